// engine.cpp -- scene-level step engine and its extern "C" entry points.
//
// The step is the stage order of updateUniformGridDEM (RigidBodies/UniformGridUpdate.cpp:939-999)
// and of the LBVH update (RigidBodies/BVHUpdate.cpp:606-759, Particles/updateFunctions.cpp:2183-2330):
//   body integrate -> particle pose + grain integrate + broadphase key -> radix sort ->
//   (grid) cell bounds + reorder -> 27-cell DEM          |
//   (LBVH) Karras build -> leaf init + refit -> traversal DEM
//   -> segmented per-body reduce + apply.
// All buffers are allocated once when the engine is built; a step enqueues kernels only, on one
// stream, with no host synchronisation, and can be replayed from a captured hipGraph.
#include "engine.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "host_common.hpp"
#include "params.hpp"
#include "scene_owned.hpp"

using arsim_host::fail;

namespace arsim {

namespace {

int ceil_log2(long long v) {
    int b = 0;
    while ((1ll << b) < v) ++b;
    return b;
}

bool same_phys(const DevPhys &a, const DevPhys &b) { return std::memcmp(&a, &b, sizeof(DevPhys)) == 0; }

// The ranges a sort needs cleared before its histograms are accumulated.
ZeroJobs sort_zero_jobs(const SortPlan &pl) {
    ZeroJobs z{};
    z.p[0] = pl.hist;
    z.count[0] = pl.histWords;
    z.p[1] = pl.counters;
    z.count[1] = (size_t)pl.passes;
    z.p[2] = pl.statusA;
    z.count[2] = pl.statusWords;
    return z;
}

// Lets the key kernel accumulate the sort's digit histograms.
void attach_hist(KeyOut &ko, const SortPlan &pl, const SortTuning &stt, int keyBits, int beginBit = 0) {
    ko.hist = pl.hist;
    ko.histPasses = pl.passes;
    ko.histRbits = stt.radixBits;
    ko.histEndBit = keyBits;
    ko.histBase = beginBit;
}

}  // namespace

// Makes *p hold at least count elements.  A buffer that is large enough is kept (its contents are
// not preserved); a smaller one is replaced by one 25 % larger than asked, so that an engine that is
// rebuilt for a growing scene set reallocates rarely.
template <class T> int Engine::alloc(T **p, size_t count) {
    const size_t bytes = std::max<size_t>(count * sizeof(T), 16);
    for (auto &a : allocs) {
        if (a.first != (void **)p) continue;
        if (a.second >= bytes) return ARSIM_OK;
        hipFree(*p);
        *p = nullptr;
        devBytes -= a.second;
        const size_t grown = bytes + bytes / 4;
        void *d = nullptr;
        hipError_t e = hipMalloc(&d, grown);
        if (e != hipSuccess) {
            a.second = 0;
            return fail(e == hipErrorOutOfMemory ? ARSIM_ERR_OUT_OF_MEMORY : ARSIM_ERR_HIP,
                        std::string("engine allocation: ") + hipGetErrorString(e));
        }
        a.second = grown;
        devBytes += grown;
        *p = (T *)d;
        return ARSIM_OK;
    }
    void *d = nullptr;
    hipError_t e = hipMalloc(&d, bytes);
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? ARSIM_ERR_OUT_OF_MEMORY : ARSIM_ERR_HIP,
                    std::string("engine allocation: ") + hipGetErrorString(e));
    allocs.push_back({(void **)p, bytes});
    devBytes += bytes;
    *p = (T *)d;
    return ARSIM_OK;
}

Engine::~Engine() {
    free_static();
    if (gexec) hipGraphExecDestroy(gexec);
    for (auto &a : allocs)
        if (*a.first) hipFree(*a.first);
    if (st && ownStream) hipStreamDestroy(st);
}

void Engine::free_static() {
    if (stPos) hipFree(stPos);
    if (stNorm) hipFree(stNorm);
    if (stTable) hipFree(stTable);
    stPos = stNorm = nullptr;
    stTable = nullptr;
    numStatic = 0;
}

// Indexes a static point set once: cell keys, a stable sort by key, sorted copies of the points
// and normals, and the exact cell table over the scene's grid.
int Engine::set_static(const float *pos4, const float *norm4, int n) {
    if (S != 1) return fail(ARSIM_ERR_UNSUPPORTED, "engine_set_static: one-scene engines only");
    if (n < 0 || (n > 0 && (!pos4 || !norm4))) return fail(ARSIM_ERR_INVALID_ARG, "engine_set_static: bad arguments");
    ARSIM_HIP(hipSetDevice(device));
    ARSIM_HIP(hipStreamSynchronize(st));
    invalidate_graph();
    primed = false;   // the step form may change with the static set (integrate_ahead)
    free_static();
    if (n == 0) return ARSIM_OK;
    if ((flags & ARSIM_CONTACT_SIS) && sisCap < 2 * N) {
        const int rc0 = alloc_sis(2 * N);   // room for one static and one body contact per particle
        if (rc0) return rc0;
    }
    const SceneGeom &g = geoms[0];
    const unsigned cells = (g.gridMask[0] + 1u) * (g.gridMask[1] + 1u) * (g.gridMask[2] + 1u);
    const size_t wsBytes = sort_workspace_bytes(n, 4);
    void *tmp = nullptr;
    const size_t v4 = (size_t)n * 16, v1 = (size_t)n * 4;
    const unsigned longCapS = fill_long_capacity((unsigned long long)cells + 1ull);
    const size_t tmpBytes = 2 * v4 + 4 * v1 + wsBytes + 256 + 16 * (size_t)longCapS + 256 + 16;
    ARSIM_HIP(hipMalloc(&tmp, tmpBytes));
    char *c = (char *)tmp;
    float4 *dPos = (float4 *)c; c += v4;
    float4 *dNorm = (float4 *)c; c += v4;
    unsigned *kA = (unsigned *)c; c += v1;
    unsigned *kB = (unsigned *)c; c += v1;
    unsigned *vA = (unsigned *)c; c += v1;
    unsigned *vB = (unsigned *)c; c += v1;
    void *ws = (void *)(((uintptr_t)c + 255) & ~(uintptr_t)255);
    // long empty runs of the static table (its tail, gaps between occupied z slabs): list + count
    uint4 *longListS = (uint4 *)(((uintptr_t)ws + wsBytes + 255) & ~(uintptr_t)255);
    unsigned *longCntS = (unsigned *)(longListS + longCapS);
    int rc = ARSIM_OK;
    hipError_t e = hipMalloc(&stPos, v4);
    if (e == hipSuccess) e = hipMalloc(&stNorm, v4);
    if (e == hipSuccess) e = hipMalloc(&stTable, ((size_t)cells + 1) * 4);
    if (e == hipSuccess) e = hipMemcpy(dPos, pos4, v4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dNorm, norm4, v4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_fill_u32(longCntS, 0u, 1, st);
        launch_static_hash(dPos, n, g, kA, st);
        int bits = 1;
        while ((1ull << bits) < (unsigned long long)cells) ++bits;
        SortTuning stt{};
        if (sort_pairs<unsigned>(&kA, &vA, &kB, &vB, n, 0, bits, true, ws, wsBytes, stt, st))
            rc = fail(ARSIM_ERR_INVALID_ARG, "engine_set_static: sort rejected");
        else
            launch_static_reorder(kA, vA, dPos, dNorm, n, cells, stPos, stNorm, stTable,
                                  LongFill{longCntS, longListS, longCapS}, st);
        e = hipStreamSynchronize(st);
        if (e == hipSuccess) e = hipGetLastError();
    }
    hipFree(tmp);
    if (e != hipSuccess) rc = fail(ARSIM_ERR_HIP, std::string("engine_set_static: ") + hipGetErrorString(e));
    if (rc) {
        free_static();
        return rc;
    }
    numStatic = n;
    return ARSIM_OK;
}

void Engine::invalidate_graph() {
    if (gexec) hipGraphExecDestroy(gexec);
    gexec = nullptr;
    graphReady = false;
}

SceneView Engine::scene_view() const {
    SceneView sv{};
    sv.count = S;
    sv.table = geomD;
    if (S == 1) sv.single = geoms[0];
    else sv.sceneStart = sceneStartD;
    return sv;
}

StateBufs Engine::state_bufs() const {
    StateBufs b{};
    b.pos = pos; b.vel = vel; b.bodyFrame = bodyFrame; b.rb = rb; b.scene = scene;
    b.rbPos = rbPos; b.rbVel = rbVel; b.rbAng = rbAng; b.qinc = qinc; b.qcum = qcum;
    b.Ib = Ib; b.Iw = Iw; b.mass = mass; b.model = model; b.pstart = pstartD; b.pcount = pcountD;
    return b;
}

BodyBufs Engine::ahead_bufs() const {
    BodyBufs b = body_bufs();
    b.pos = aPos; b.vel = aVel; b.ang = aAng; b.qinc = aQinc; b.qcum = aQcum; b.Iw = aIw;
    b.model = nullptr;
    return b;
}

int Engine::prime() {
    ZeroJobs z{};
    z.p[0] = bucketHist;
    z.count[0] = (size_t)numBuckets;
    const BodyBufs a = ahead_bufs();
    launch_body_integrate(body_bufs(), B, ph, &z, st, &a);
    ARSIM_HIP(hipGetLastError());
    primed = true;
    return ARSIM_OK;
}

BodyBufs Engine::body_bufs() const {
    BodyBufs b{};
    b.pos = rbPos; b.vel = rbVel; b.ang = rbAng; b.qinc = qinc; b.qcum = qcum;
    b.Ib = Ib; b.Iw = Iw; b.model = model; b.mass = mass; b.pstart = pstartD; b.pcount = pcountD;
    return b;
}

PartBufs Engine::part_bufs() const {
    PartBufs p{};
    p.pos = pos; p.vel = vel; p.relPos = relPos; p.bodyFrame = bodyFrame; p.rb = rb;
    p.scene = S > 1 ? scene : nullptr;
    p.rec = gatherRec;
    return p;
}

const char *const *Engine::stage_names(int method, unsigned flags, bool bucket, bool fusedDem, bool ahead,
                                       int *count) {
    const char *const *names = stage_names_all(method, flags, bucket, fusedDem, count);
    if (ahead) {   // every list starts with the body-integration stage
        --*count;
        return names + 1;
    }
    return names;
}

const char *const *Engine::stage_names_all(int method, unsigned flags, bool bucket, bool fusedDem, int *count) {
    // with the bucket sort the pose stage also ranks each tile by bucket (the sort's first pass)
    // and the bucket stage finishes the sort with the reorder
    static const char *const sis[] = {"body_integrate", "pose_key", "sort", "reorder_cellstart", "sis_find",
                                      "sis_gather", "sis_solve", "pose"};
    static const char *const sisB[] = {"body_integrate", "pose_key", "bucket_sort_reorder", "sis_find",
                                       "sis_gather", "sis_solve", "pose"};
    if (flags & ARSIM_CONTACT_SIS) {
        *count = bucket ? 7 : 8;
        return bucket ? sisB : sis;
    }
    static const char *const grid[] = {"body_integrate", "pose_key", "sort", "reorder_cellstart", "dem_grid",
                                       "reduce_apply"};
    static const char *const gridB[] = {"body_integrate", "pose_key", "bucket_sort_reorder", "dem_grid",
                                        "reduce_apply"};
    static const char *const lbvh[] = {"body_integrate", "pose_key", "sort", "lbvh_build", "lbvh_refit", "dem_lbvh",
                                       "reduce_apply"};
    // Morton codes ranked per tile by bucket in the pose stage and sorted per bucket in the bucket stage
    static const char *const lbvhB[] = {"body_integrate", "pose_key", "bucket_sort", "lbvh_build", "lbvh_refit",
                                        "dem_lbvh", "reduce_apply"};
    static const char *const gridF[] = {"body_integrate", "pose_key", "sort", "reorder_cellstart", "dem_reduce"};
    static const char *const gridBF[] = {"body_integrate", "pose_key", "bucket_sort_reorder", "dem_reduce"};
    if (method == ARSIM_METHOD_GRID) {
        if (fusedDem) {
            *count = bucket ? 4 : 5;
            return bucket ? gridBF : gridF;
        }
        *count = bucket ? 5 : 6;
        return bucket ? gridB : grid;
    }
    *count = 7;
    return bucket ? lbvhB : lbvh;
}

size_t Engine::scene_bytes_estimate() const { return devBytes; }

int Engine::state_checksum(int s, uint64_t *out) {
    if (s < 0 || s >= S || !out) return fail(ARSIM_ERR_INVALID_ARG, "state_checksum: bad scene");
    ARSIM_HIP(hipSetDevice(device));
    ARSIM_HIP(hipStreamSynchronize(st));
    const size_t po = pStart[s], n = pCount[s], bo = bStart[s], b = bCount[s];
    std::vector<float4> buf(std::max(n, b));
    uint64_t h = 0xcbf29ce484222325ull;
    auto mix = [&](const void *src, size_t count) -> int {
        ARSIM_HIP(hipMemcpy(buf.data(), src, count * 16, hipMemcpyDeviceToHost));
        const unsigned char *c = reinterpret_cast<const unsigned char *>(buf.data());
        for (size_t k = 0; k < count * 16; ++k) {
            h ^= c[k];
            h *= 0x100000001b3ull;
        }
        return ARSIM_OK;
    };
    int rc = mix(pos + po, n);
    if (!rc) rc = mix(vel + po, n);
    if (!rc && b) rc = mix(rbPos + bo, b);
    if (!rc && b) rc = mix(rbVel + bo, b);
    if (!rc && b) rc = mix(rbAng + bo, b);
    if (!rc && b) rc = mix(qcum + bo, b);
    if (rc) return rc;
    *out = h;
    return ARSIM_OK;
}

namespace {

// The fullest bucket of a scene's grid keys for every bucket width 2^b cells (b = 0..16, UINT_MAX
// where there would be more than 1024 buckets), from host positions (float4 per particle).
std::vector<unsigned> bucket_peaks(const float *pos4, int n, const SceneGeom &g) {
    std::vector<unsigned> keys(n);
    for (int i = 0; i < n; ++i) {
        unsigned c[3];
        for (int a = 0; a < 3; ++a) {
            const double q = std::floor(((double)pos4[4 * (size_t)i + a] - g.origin[a]) / g.cellSize[a]);
            const long long v = q > -1e18 && q < 1e18 ? (long long)q : 0;   // NaN and huge values: cell 0
            c[a] = (unsigned)v & g.gridMask[a];
        }
        keys[i] = c[2] * g.gxy + c[1] * g.gx + c[0];
    }
    const unsigned long long cells =
        (unsigned long long)(g.gridMask[0] + 1u) * (g.gridMask[1] + 1u) * (g.gridMask[2] + 1u);
    std::vector<unsigned> peak(17, 0xffffffffu);
    std::vector<unsigned> hist(1024);
    for (int b = 16; b >= 0; --b) {
        if (((cells + (1ull << b) - 1) >> b) > 1024) break;
        std::fill(hist.begin(), hist.end(), 0u);
        for (unsigned k : keys) ++hist[k >> b];
        peak[b] = *std::max_element(hist.begin(), hist.end());
    }
    return peak;
}

// The fullest bucket of a scene's 30-bit Morton codes by their top 10 bits (1,024 buckets), at index 20 of
// the plan's per-shift peaks (other widths: UINT_MAX), from host positions normalised as the device does
// (for planning only).
std::vector<unsigned> morton_peaks(const float *pos4, int n, const SceneGeom &g) {
    auto spread = [](unsigned v) {   // bit k of v to bit 3k
        unsigned r = 0u;
        for (int k = 0; k < 10; ++k) r |= ((v >> k) & 1u) << (3 * k);
        return r;
    };
    std::vector<unsigned> peak(21, 0xffffffffu), hist(1024, 0u);
    for (int i = 0; i < n; ++i) {
        unsigned q[3];
        for (int a = 0; a < 3; ++a) {
            const double u = g.mExt[a] == 0.0f ? 0.0 : ((double)pos4[4 * (size_t)i + a] - g.mLo[a]) / g.mExt[a];
            const double v = u * 1024.0;
            q[a] = v > 0.0 ? (v < 1023.0 ? (unsigned)v : 1023u) : 0u;   // NaN: 0
        }
        const unsigned code = spread(q[0]) * 4u + spread(q[1]) * 2u + spread(q[2]);
        ++hist[code >> 20];
    }
    peak[20] = *std::max_element(hist.begin(), hist.end());
    return peak;
}

}  // namespace

int Engine::build(const std::vector<SceneSource> &src, const ArSimTuning *tuning, int dev, hipStream_t stream) {
    if (src.empty()) return fail(ARSIM_ERR_INVALID_ARG, "engine: no scenes");
    if (src.size() > 65535) return fail(ARSIM_ERR_INVALID_ARG, "engine: at most 65535 scenes");
    device = dev;
    ARSIM_HIP(hipSetDevice(device));
    if (stream) {
        if (st && ownStream) hipStreamDestroy(st);
        st = stream;
        ownStream = false;
    } else if (!st) {
        ARSIM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        ownStream = true;
    }
    // a rebuild starts from a clean step state; buffers are kept where they are large enough
    invalidate_graph();
    free_static();
    if (tuning) tu = *tuning;
    if (tu.demBlock <= 0 || tu.demBlock > 1024 || (tu.demBlock & 63)) return fail(ARSIM_ERR_INVALID_ARG, "bad demBlock");
    if (!sort_tuning_supported(tu.sortItemsPerThread, tu.sortRadixBits))
        return fail(ARSIM_ERR_INVALID_ARG, "unsupported sort tuning");
    if (tu.bucketSplit < 0 || tu.bucketSplit > 2) return fail(ARSIM_ERR_INVALID_ARG, "bad bucketSplit");

    // ---- layout
    S = (int)src.size();
    bucketSort = false;
    mortonBucket = false;
    primed = false;
    struct Desc { int n, b; ArSimParams p; ArSimStepConfig c; };
    std::vector<Desc> d(S);
    for (int s = 0; s < S; ++s) {
        if (src[s].host) {
            const ArSimSceneHost *h = src[s].host;
            if (h->numParticles <= 0 || h->numBodies < 0)
                return fail(ARSIM_ERR_INVALID_ARG, "engine: scene needs particles");
            if (!h->pos || !h->vel || !h->rbIndices || !h->bodyFramePos)
                return fail(ARSIM_ERR_INVALID_ARG, "engine: null particle array");
            if (h->numBodies > 0 && (!h->rbPositions || !h->rbVelocities || !h->rbAngularVelocity || !h->rbQuaternion ||
                                     !h->rbInertia || !h->rbMass || !h->particlesPerObject))
                return fail(ARSIM_ERR_INVALID_ARG, "engine: null body array");
            d[s] = Desc{h->numParticles, h->numBodies, h->params, h->cfg};
            if (h->cfg.flags & ARSIM_CONTACT_SIS) {
                long long inBodies = 0;
                for (int k = 0; k < h->numBodies; ++k) inBodies += h->particlesPerObject[k];
                if (inBodies != h->numParticles)
                    return fail(ARSIM_ERR_UNSUPPORTED, "engine: the sequential-impulse model takes bodies only (no grains)");
            }
        } else if (src[s].from) {
            const Engine *o = src[s].from;
            if (o == this) return fail(ARSIM_ERR_INVALID_ARG, "engine: a rebuild cannot read its own buffers");
            int k = src[s].fromScene;
            if (k < 0 || k >= o->S) return fail(ARSIM_ERR_INVALID_ARG, "engine: bad source scene");
            d[s] = Desc{o->pCount[k], o->bCount[k], o->p0, o->c0};
            // per-scene geometry of the source
            const SceneGeom &g = o->geoms[k];
            d[s].p.gridSize = {g.gridMask[0] + 1u, g.gridMask[1] + 1u, g.gridMask[2] + 1u};
            for (int a = 0; a < 3; ++a) { d[s].c.boxMin[a] = g.boxMin[a]; d[s].c.boxMax[a] = g.boxMax[a]; }
            d[s].c.wallMask = g.wallMask;
        } else {
            return fail(ARSIM_ERR_INVALID_ARG, "engine: empty scene source");
        }
    }
    p0 = d[0].p;
    c0 = d[0].c;
    method = c0.method;
    flags = c0.flags;
    ph = make_phys(p0, c0);
    if (method < 0 || method > 2) return fail(ARSIM_ERR_INVALID_ARG, "engine: bad method");
    if (method == ARSIM_METHOD_LBVH63 && S > 1)
        return fail(ARSIM_ERR_UNSUPPORTED, "engine: 63-bit Morton codes leave no key bits for scene ids");
    if ((flags & ARSIM_DYNAMIC_AABB) && S > 1)
        return fail(ARSIM_ERR_UNSUPPORTED, "engine: a dynamic Morton box is per scene only with one scene");
    if ((flags & ARSIM_CONTACT_SIS) && method != ARSIM_METHOD_GRID)
        return fail(ARSIM_ERR_UNSUPPORTED, "engine: the sequential-impulse model runs on the grid method");
    geoms.resize(S);
    pStart.resize(S); pCount.resize(S); bStart.resize(S); bCount.resize(S);
    std::vector<int> newBodyParticles(S, 0);   // assigned to bodyParticles once the sources are read
    long long nTot = 0, bTot = 0;
    int maxCellBits = 0;
    long long totalCells = 0;
    std::vector<long long> sceneCells(S, 0);
    for (int s = 0; s < S; ++s) {
        if (d[s].c.method != method) return fail(ARSIM_ERR_INVALID_ARG, "engine: scenes mix methods");
        if (!same_phys(make_phys(d[s].p, d[s].c), ph))
            return fail(ARSIM_ERR_INVALID_ARG, "engine: scenes must share the physics parameters");
        pStart[s] = (int)nTot; pCount[s] = d[s].n;
        bStart[s] = (int)bTot; bCount[s] = d[s].b;
        nTot += d[s].n; bTot += d[s].b;
        int rc = make_geom(d[s].p, d[s].c, pStart[s], pCount[s], geoms[s]);
        if (rc) return rc;
        if (method == ARSIM_METHOD_GRID) {
            long long cells = (long long)d[s].p.gridSize.x * d[s].p.gridSize.y * d[s].p.gridSize.z;
            if (cells <= 0) return fail(ARSIM_ERR_INVALID_ARG, "engine: grid has no cells");
            maxCellBits = std::max(maxCellBits, ceil_log2(cells));
            sceneCells[s] = cells;
            totalCells += cells;
        }
    }
    if (nTot >= (1ll << 30)) return fail(ARSIM_ERR_INVALID_ARG, "engine: too many particles");
    N = (int)nTot;
    B = (int)bTot;
    sceneBits = S > 1 ? ceil_log2(S) : 0;
    if (method == ARSIM_METHOD_GRID) {
        // scenes' cell ranges are packed back to back: key = cellBase[scene] + cell hash
        cellBits = maxCellBits;
        if (totalCells > (1ll << 32)) return fail(ARSIM_ERR_OVERFLOW, "engine: total cells of all scenes exceed 2^32");
        keyBits = std::max(1, ceil_log2(totalCells));
        keyBytes = 4;
        numCells = (size_t)totalCells;
        // one entry per cell (the DEM then reads no keys) up to 2^30 cells = 4 GiB of table, rewritten
        // every step; beyond that, segments of 2^segBits cells keep the table at that size
        segBits = keyBits > kExactTableBits ? keyBits - kExactTableBits : 0;
        numSegs = (long long)((numCells - 1) >> segBits) + 1;
        if (segBits == 0) {
            // bucket widths from the particles of host scenes, kept for scenes carried over from an
            // engine that planned them
            std::vector<BucketSceneShape> shapes(S);
            for (int s = 0; s < S; ++s) {
                shapes[s].particles = pCount[s];
                for (int a = 0; a < 3; ++a) {
                    shapes[s].grid[a] = geoms[s].gridMask[a] + 1u;
                    shapes[s].extent[a] = ((double)geoms[s].boxMax[a] - (double)geoms[s].boxMin[a]) /
                                          (double)geoms[s].cellSize[a];
                }
                if (src[s].host) shapes[s].peak = bucket_peaks(src[s].host->pos, pCount[s], geoms[s]);
                else if (src[s].from && src[s].from->bucketSort)
                    shapes[s].shift = src[s].from->bkPlan.scenes[src[s].fromScene].shift;
            }
            bucketSort = plan_bucket_sort(shapes, &bkPlan);
        }
        if ((flags & ARSIM_CONTACT_SIS) && segBits != 0)
            return fail(ARSIM_ERR_UNSUPPORTED, "engine: the sequential-impulse model needs at most 2^30 cells");
        long long base = 0;
        for (int s = 0; s < S; ++s) {
            geoms[s].cellBase = (unsigned)base;
            base += sceneCells[s];
        }
    } else if ((method == ARSIM_METHOD_LBVH30 || method == ARSIM_METHOD_LBVH63) && S == 1) {
        const bool m63 = method == ARSIM_METHOD_LBVH63;
        keyBits = m63 ? 63 : 30; keyBytes = m63 ? 8 : 4;
        // the codes sorted per bucket of their top 10 bits (tiles of 4096, fixed Morton box): 1,024 buckets
        // of 2^20 codes, a bucket's positions in 12 bits of an LDS word.  63-bit codes: the same sort of
        // their top 30 bits (the 30-bit code), then the ties among equal top bits ranked by the full code
        if (!(flags & ARSIM_DYNAMIC_AABB) && N > 65536) {
            std::vector<BucketSceneShape> shapes(1);
            shapes[0].particles = N;
            shapes[0].grid[0] = shapes[0].grid[1] = shapes[0].grid[2] = 1024u;
            if (src[0].host) shapes[0].peak = morton_peaks(src[0].host->pos, N, geoms[0]);
            else if (src[0].from && src[0].from->mortonBucket) shapes[0].shift = 20;
            mortonBucket = plan_bucket_sort(shapes, &bkPlan, 20, 4096) && bkPlan.scenes[0].shift == 20 &&
                           bkPlan.tileKeys == kTileKeys;
        }
    } else if (method == ARSIM_METHOD_LBVH63) {
        keyBits = 63; keyBytes = 8;   // a forest of 63-bit codes
    } else {
        keyBits = 30 + sceneBits; keyBytes = 8;
    }

    // ---- allocation (everything a step touches, once)
    int rc = 0;
#define A(ptr, cnt) if ((rc = alloc(&(ptr), (cnt))) != 0) return rc
    A(pos, N); A(vel, N); A(relPos, N); A(bodyFrame, N); A(pForce, N);
    A(rb, N); A(scene, N); A(gatherRec, 2 * (size_t)N);
    size_t nb = std::max(B, 1);
    A(rbPos, nb); A(rbVel, nb); A(rbAng, nb); A(qinc, nb); A(qcum, nb);
    A(Ib, nb * 9); A(Iw, nb * 9); A(mass, nb); A(pstartD, nb); A(pcountD, nb); A(model, nb * 16);
    A(geomD, (size_t)S); A(sceneStartD, (size_t)S); A(sceneCountD, (size_t)S); A(carryTab, (size_t)S);
    // (alloc keys its records by the member's address: only members are allocated through it)
    A(keyBufA, (size_t)N * keyBytes); A(keyBufB, (size_t)N * keyBytes);
    keysA = keyBufA; keysB = keyBufB;
    A(valsA, N); A(valsB, N);
    sortWsBytes = sort_workspace_bytes(N, keyBytes);
    A(sortWsBuf, sortWsBytes);
    sortWs = sortWsBuf;
    A(overflowD, 4); A(flagD, 4);
    if (method == ARSIM_METHOD_GRID) {
        if (tu.fuseReduce) { A(grainIdx, N); A(grainRanges, (size_t)S); }
        A(segStart, (size_t)numSegs + 1);
        A(sortedPos, N); A(sortedVel, N);
        longCap = fill_long_capacity((unsigned long long)numSegs + 1ull);
        A(longCnt, 2); A(longList, (size_t)longCap);
    } else {
        size_t nodes = (size_t)2 * N;
        A(tree.left, nodes); A(tree.right, nodes); A(tree.parent, nodes);
        // range refit: node records, per-leaf workgroup prefix / suffix boxes, table of workgroup boxes
        A(tree.meta, (size_t)N); A(tree.pre, 2 * (size_t)N); A(tree.suf, 2 * (size_t)N);
        tree.blockTabLevels = lbvh_refit_levels(N);
        A(tree.blockTab, 2 * (size_t)tree.blockTabLevels * (size_t)((N + lbvh_refit_block() - 1) / lbvh_refit_block()));
        A(tree.packed, 4 * (size_t)N);
        // node boxes: the step keeps them in the traversal records only (tree.bounds stays NULL);
        // treeBounds receives them when a debug view asks
        A(treeBounds, 2 * nodes); A(tree.cms, N); A(tree.sortedVel, N); A(tree.counter, N);
        A(tree.crossList, (size_t)N); A(tree.crossCount, (size_t)N / 1024 + 1);
        if (mortonBucket && method == ARSIM_METHOD_LBVH63) A(code63, (size_t)N);
        if (S > 1) A(sceneRootBuf, S);
        tree.sceneRoot = S > 1 ? sceneRootBuf : nullptr;
        A(aabbPart, (size_t)2 * aabb_max_partials()); A(aabbBox, 2);
    }
    if (bucketSort || mortonBucket) {
        numTiles = (int)bkPlan.tileScene.size();
        numBuckets = (int)bkPlan.bucketScene.size();
        tileKeys = bkPlan.tileKeys;
        A(tileInfo, std::max<size_t>(bkPlan.infoWords, 1)); A(bkScenesD, (size_t)S);
        A(tileSceneD, std::max(numTiles, 1)); A(bucketSceneD, std::max(numBuckets, 1));
        A(bucketHist, std::max(numBuckets, 1));
        if (!(flags & ARSIM_CONTACT_SIS)) {
            A(aPos, nb); A(aVel, nb); A(aAng, nb); A(aQinc, nb); A(aQcum, nb); A(aIw, nb * 9);
        }
    }
    if (flags & ARSIM_CONTACT_SIS) {
        A(sisDistA, N); A(sisPartA, N); A(sisDistB, N); A(sisBodyB, N); A(sisPartB, N); A(sceneBodyD, (size_t)S);
        if ((rc = alloc_sis(N)) != 0) return rc;
    }
#undef A

    // ---- fill
    std::vector<int> hStart(S), hCount(S);
    for (int s = 0; s < S; ++s) { hStart[s] = pStart[s]; hCount[s] = pCount[s]; }
    // Scenes carried over from one other engine that form a prefix of the new layout move in one
    // launch per array group (a rebuild for a new scene set); others are copied scene by scene.
    int bulk = 0;
    const Engine *bulkFrom = S > 0 ? src[0].from : nullptr;
    while (bulk < S && bulkFrom && src[bulk].from == bulkFrom) ++bulk;
    if (bulk > 0) {
        std::vector<CarryRange> tab(bulk);
        for (int s = 0; s < bulk; ++s) {
            const int k = src[s].fromScene;
            tab[s] = CarryRange{pStart[s], bulkFrom->pStart[k], pCount[s], bStart[s], bulkFrom->bStart[k], bCount[s], s};
            newBodyParticles[s] = bulkFrom->bodyParticles[k];
        }
        ARSIM_HIP(hipMemcpyAsync(carryTab, tab.data(), sizeof(CarryRange) * bulk, hipMemcpyHostToDevice, st));
        const int carriedN = pStart[bulk - 1] + pCount[bulk - 1], carriedB = bStart[bulk - 1] + bCount[bulk - 1];
        launch_carry(carryTab, bulk, carriedN, carriedB, bulkFrom->state_bufs(), state_bufs(), st);
        ARSIM_HIP(hipStreamSynchronize(st));   // tab is a host temporary
    }
    for (int s = bulk; s < S; ++s) {
        const size_t po = pStart[s], bo = bStart[s], n = pCount[s], b = bCount[s];
        if (src[s].host) {
            const ArSimSceneHost *h = src[s].host;
            ARSIM_HIP(hipMemcpyAsync(pos + po, h->pos, n * 16, hipMemcpyHostToDevice, st));
            ARSIM_HIP(hipMemcpyAsync(vel + po, h->vel, n * 16, hipMemcpyHostToDevice, st));
            ARSIM_HIP(hipMemcpyAsync(bodyFrame + po, h->bodyFramePos, n * 16, hipMemcpyHostToDevice, st));
            ARSIM_HIP(hipMemsetAsync(relPos + po, 0, n * 16, st));
            ARSIM_HIP(hipMemsetAsync(pForce + po, 0, n * 16, st));
            std::vector<int> rbg(n);
            for (size_t i = 0; i < n; ++i) {
                int v = h->rbIndices[i];
                if (v >= (int)b) return fail(ARSIM_ERR_INVALID_ARG, "engine: rbIndex out of range");
                rbg[i] = v >= 0 ? v + (int)bo : -1;
            }
            ARSIM_HIP(hipMemcpyAsync(rb + po, rbg.data(), n * 4, hipMemcpyHostToDevice, st));
            ARSIM_HIP(hipStreamSynchronize(st));   // rbg is a host temporary
            if (b > 0) {
                ARSIM_HIP(hipMemcpyAsync(rbPos + bo, h->rbPositions, b * 16, hipMemcpyHostToDevice, st));
                ARSIM_HIP(hipMemcpyAsync(rbVel + bo, h->rbVelocities, b * 16, hipMemcpyHostToDevice, st));
                ARSIM_HIP(hipMemcpyAsync(rbAng + bo, h->rbAngularVelocity, b * 16, hipMemcpyHostToDevice, st));
                ARSIM_HIP(hipMemcpyAsync(qcum + bo, h->rbQuaternion, b * 16, hipMemcpyHostToDevice, st));
                ARSIM_HIP(hipMemcpyAsync(Ib + bo * 9, h->rbInertia, b * 36, hipMemcpyHostToDevice, st));
                ARSIM_HIP(hipMemcpyAsync(Iw + bo * 9, h->rbInertia, b * 36, hipMemcpyHostToDevice, st));
                ARSIM_HIP(hipMemcpyAsync(mass + bo, h->rbMass, b * 4, hipMemcpyHostToDevice, st));
                std::vector<float4> ident(b, make_float4(0.f, 0.f, 0.f, 1.f));
                std::vector<int> ps(b), pc(b);
                int acc = (int)po;
                for (size_t k = 0; k < b; ++k) {
                    ps[k] = acc;
                    pc[k] = h->particlesPerObject[k];
                    acc += pc[k];
                    for (int j = ps[k]; j < acc; ++j)
                        if (j - (int)po >= (int)n || h->rbIndices[j - po] != (int)k)
                            return fail(ARSIM_ERR_INVALID_ARG, "engine: body particles must be contiguous, in body order");
                }
                // grains follow the bodies (the reference layout)
                for (int j = acc; j < (int)(po + n); ++j)
                    if (h->rbIndices[j - po] != -1)
                        return fail(ARSIM_ERR_INVALID_ARG, "engine: grains must follow the body particles");
                newBodyParticles[s] = acc - (int)po;
                ARSIM_HIP(hipMemcpyAsync(qinc + bo, ident.data(), b * 16, hipMemcpyHostToDevice, st));
                ARSIM_HIP(hipMemcpyAsync(pstartD + bo, ps.data(), b * 4, hipMemcpyHostToDevice, st));
                ARSIM_HIP(hipMemcpyAsync(pcountD + bo, pc.data(), b * 4, hipMemcpyHostToDevice, st));
                ARSIM_HIP(hipStreamSynchronize(st));
            } else {
                for (size_t j = 0; j < n; ++j)
                    if (h->rbIndices[j] != -1) return fail(ARSIM_ERR_INVALID_ARG, "engine: rbIndex without a body");
            }
        } else {
            const Engine *o = src[s].from;
            const int k = src[s].fromScene;
            const size_t opo = o->pStart[k], obo = o->bStart[k];
            newBodyParticles[s] = o->bodyParticles[k];
            ARSIM_HIP(hipMemcpyAsync(pos + po, o->pos + opo, n * 16, hipMemcpyDeviceToDevice, st));
            ARSIM_HIP(hipMemcpyAsync(vel + po, o->vel + opo, n * 16, hipMemcpyDeviceToDevice, st));
            ARSIM_HIP(hipMemcpyAsync(bodyFrame + po, o->bodyFrame + opo, n * 16, hipMemcpyDeviceToDevice, st));
            ARSIM_HIP(hipMemcpyAsync(relPos + po, o->relPos + opo, n * 16, hipMemcpyDeviceToDevice, st));
            ARSIM_HIP(hipMemcpyAsync(pForce + po, o->pForce + opo, n * 16, hipMemcpyDeviceToDevice, st));
            launch_copy_offset(rb + po, o->rb + opo, (int)n, (int)bo - (int)obo, st);
            if (b > 0) {
                ARSIM_HIP(hipMemcpyAsync(rbPos + bo, o->rbPos + obo, b * 16, hipMemcpyDeviceToDevice, st));
                ARSIM_HIP(hipMemcpyAsync(rbVel + bo, o->rbVel + obo, b * 16, hipMemcpyDeviceToDevice, st));
                ARSIM_HIP(hipMemcpyAsync(rbAng + bo, o->rbAng + obo, b * 16, hipMemcpyDeviceToDevice, st));
                ARSIM_HIP(hipMemcpyAsync(qinc + bo, o->qinc + obo, b * 16, hipMemcpyDeviceToDevice, st));
                ARSIM_HIP(hipMemcpyAsync(qcum + bo, o->qcum + obo, b * 16, hipMemcpyDeviceToDevice, st));
                ARSIM_HIP(hipMemcpyAsync(Ib + bo * 9, o->Ib + obo * 9, b * 36, hipMemcpyDeviceToDevice, st));
                ARSIM_HIP(hipMemcpyAsync(Iw + bo * 9, o->Iw + obo * 9, b * 36, hipMemcpyDeviceToDevice, st));
                ARSIM_HIP(hipMemcpyAsync(mass + bo, o->mass + obo, b * 4, hipMemcpyDeviceToDevice, st));
                launch_copy_offset(pstartD + bo, o->pstartD + obo, (int)b, (int)po - (int)opo, st);
                ARSIM_HIP(hipMemcpyAsync(pcountD + bo, o->pcountD + obo, b * 4, hipMemcpyDeviceToDevice, st));
            }
        }
        launch_fill_u16(scene + po, (uint16_t)s, n, st);
    }
    bodyParticles = newBodyParticles;
    numGrains = 0;
    if (method == ARSIM_METHOD_GRID && tu.fuseReduce) {
        // the list of grain indices the fused contact kernel walks: each scene's grains follow its bodies
        std::vector<int4> gr(S);
        for (int s = 0; s < S; ++s) {
            const int cnt = pCount[s] - bodyParticles[s];
            gr[s] = make_int4(numGrains, pStart[s] + bodyParticles[s], cnt, 0);
            numGrains += cnt;
        }
        ARSIM_HIP(hipMemcpyAsync(grainRanges, gr.data(), sizeof(int4) * S, hipMemcpyHostToDevice, st));
        launch_iota_ranges(grainRanges, S, grainIdx, st);
        ARSIM_HIP(hipStreamSynchronize(st));   // gr is a host temporary
    }
    ARSIM_HIP(hipMemcpyAsync(geomD, geoms.data(), sizeof(SceneGeom) * S, hipMemcpyHostToDevice, st));
    ARSIM_HIP(hipMemcpyAsync(sceneStartD, hStart.data(), 4 * (size_t)S, hipMemcpyHostToDevice, st));
    ARSIM_HIP(hipMemcpyAsync(sceneCountD, hCount.data(), 4 * (size_t)S, hipMemcpyHostToDevice, st));
    if (tree.sceneRoot) launch_fill_u32((unsigned *)tree.sceneRoot, 0xffffffffu, (size_t)S, st);
    launch_fill_u32((unsigned *)overflowD, 0u, 1, st);
    if (longCnt) launch_fill_u32(longCnt, 0u, 1, st);
    maxSceneBodies = 0;
    for (int s = 0; s < S; ++s) maxSceneBodies = std::max(maxSceneBodies, bCount[s]);
    if (bucketSort || mortonBucket) {
        ARSIM_HIP(hipMemcpyAsync(bkScenesD, bkPlan.scenes.data(), sizeof(BucketScene) * S, hipMemcpyHostToDevice, st));
        ARSIM_HIP(hipMemcpyAsync(tileSceneD, bkPlan.tileScene.data(), sizeof(int) * numTiles, hipMemcpyHostToDevice,
                                 st));
        ARSIM_HIP(hipMemcpyAsync(bucketSceneD, bkPlan.bucketScene.data(), sizeof(int) * numBuckets,
                                 hipMemcpyHostToDevice, st));
        ARSIM_HIP(hipStreamSynchronize(st));   // the plan's vectors may be replaced by the next build
    }
    if (sceneBodyD) {
        std::vector<int2> sb(S);
        for (int s = 0; s < S; ++s) sb[s] = make_int2(bStart[s], bCount[s]);
        ARSIM_HIP(hipMemcpyAsync(sceneBodyD, sb.data(), sizeof(int2) * S, hipMemcpyHostToDevice, st));
        ARSIM_HIP(hipStreamSynchronize(st));   // sb is a host temporary
    }
    ARSIM_HIP(hipGetLastError());
    ARSIM_HIP(hipStreamSynchronize(st));
    first = true;
    prevSorted = nullptr;
    return ARSIM_OK;
}

// Contact list of `cap` entries and the sweeps' scratch (one body contact per particle, plus one
// static contact per particle once a static set is attached).
int Engine::alloc_sis(int cap) {
    int rc = 0;
    if ((rc = alloc(&cNormal, (size_t)cap)) || (rc = alloc(&cPoint, (size_t)cap)) || (rc = alloc(&cImpulse, (size_t)cap)) ||
        (rc = alloc(&cBodies, (size_t)cap)) || (rc = alloc(&cBias, (size_t)cap)))
        return rc;
    const size_t ws = sis_workspace_bytes(cap, S, N);
    if ((rc = alloc(&sisWsBuf, ws))) return rc;
    sisW = sis_work_layout(sisWsBuf, cap, S, N);
    sisCap = cap;
    return ARSIM_OK;
}

int Engine::reset_grid_tables() {
    invalidate_graph();
    first = true;
    prevSorted = nullptr;
    ARSIM_HIP(hipGetLastError());
    return ARSIM_OK;
}

// Body integration, pose + keys, the sort and the reorder / cell table of the grid methods (the DEM
// and the SIS steps share them).  Records the stage boundaries through rec; on return *kAOut / *vAOut
// hold the sorted keys and particle indices.
template <class Rec>
int Engine::enqueue_grid_broadphase(const BodyBufs &bb, const PartBufs &pb, const SceneView &sv, const SortTuning &tuIn,
                                    bool grainVelDead, bool ahead, unsigned **kAOut, unsigned **vAOut, Rec &rec) {
    if (bucketSort) {
        // the pose kernel ranks each tile of particles by bucket (the sort's first pass); the bucket
        // kernel gathers each bucket's runs, finishes the sort there, and writes the sorted copies and
        // the cell table.  ahead: the bodies' integration is already in the ahead buffers
        if (!ahead) {
            ZeroJobs z{};
            z.p[0] = bucketHist;
            z.count[0] = (size_t)numBuckets;
            launch_body_integrate(bb, B, ph, &z, st);
            rec();
        }
        const TileKeyOut to{(unsigned *)keysB, valsB, tileInfo, bucketHist, bkScenesD, tileSceneD, bucketSceneD,
                            numTiles, numBuckets, tileKeys, bkPlan.bucketBits};
        launch_pose_key_tiles(pb, ahead ? ahead_bufs() : bb, N, ph, sv, to, grainVelDead, st);
        rec();
        const int split = tu.bucketSplit == 1 || tu.bucketSplit == 2 ? tu.bucketSplit : (numBuckets < 512 ? 2 : 1);
        launch_bucket_sort(to, (unsigned *)keysA, valsA, gatherRec, sortedPos, sortedVel, segStart, (unsigned)numCells,
                           split, st);
        rec();
        *kAOut = (unsigned *)keysA;
        *vAOut = valsA;
        return ARSIM_OK;
    }
    const SortTuning stt = tuIn;
    const SortPlan pl = sort_plan(sortWs, N, keyBytes, 0, keyBits, stt);
    ZeroJobs z = sort_zero_jobs(pl);
    z.p[3] = longCnt;   // the cell table's long-run count
    z.count[3] = 1;
    launch_body_integrate(bb, B, ph, &z, st);
    rec();
    KeyOut ko{keysA, nullptr, nullptr};
    attach_hist(ko, pl, stt, keyBits);
    ko.grainVelDead = grainVelDead;
    launch_pose_key(pb, bb, N, ph, sv, kKeyGrid, ko, true, true, st);
    rec();
    unsigned *kA = (unsigned *)keysA, *kB = (unsigned *)keysB, *vA = valsA, *vB = valsB;
    if (sort_pairs<unsigned>(&kA, &vA, &kB, &vB, N, 0, keyBits, true, sortWs, sortWsBytes, stt, st, true))
        return fail(ARSIM_ERR_INVALID_ARG, "engine: sort rejected");
    rec();
    launch_reorder_seg(kA, vA, pos, vel, rb, S > 1 ? scene : nullptr, gatherRec, sortedPos, sortedVel, segStart, segBits,
                       numSegs, N, long_fill(), st);
    rec();
    *kAOut = kA;
    *vAOut = vA;
    return ARSIM_OK;
}

int Engine::enqueue_step(hipEvent_t *ev, int *numStages) {
    const BodyBufs bb = body_bufs();
    const PartBufs pb = part_bufs();
    const SceneView sv = scene_view();
    SortTuning stt;
    stt.ipt = tu.sortItemsPerThread;
    stt.radixBits = tu.sortRadixBits;
    int k = 0;
    auto rec = [&]() {
        if (ev) hipEventRecord(ev[k], st);
        ++k;
    };
    rec();
    const bool ahead = integrate_ahead();
    ZeroJobs zHist{};   // with ahead: the bucket counts, zeroed by the reduction for the next step
    zHist.p[0] = bucketHist;
    zHist.count[0] = (size_t)numBuckets;
    if (flags & ARSIM_CONTACT_SIS) {
        // updateUniformGridSIS (RigidBodies/UniformGridUpdate.cpp:734-797): integrate and pose, the grid
        // broadphase, deepest contacts (static set first), the contact list, the sweeps, the pose again
        unsigned *kA = nullptr, *vA = nullptr;
        if (int rc = enqueue_grid_broadphase(bb, pb, sv, stt, false, false, &kA, &vA, rec)) return rc;
        if (numStatic)
            launch_sis_find_static(sortedPos, vA, N, stPos, nullptr, stTable, nullptr, geoms[0], ph.r, false, sisDistA,
                                   sisPartA, nullptr, st);
        launch_sis_find(sortedPos, sortedVel, vA, segStart, sv, S > 1, N, ph.r, sisDistB, sisBodyB, sisPartB, st);
        rec();
        const SisGatherIn gi{numStatic ? sisDistA : nullptr, sisPartA, stNorm, sisDistB, sisBodyB, sisPartB, rb,
                             S > 1 ? scene : nullptr, rbPos, rbVel, rbAng, relPos, ph.r, p0.ARrestitution,
                             p0.RBrestitution};
        const SisContacts ct{cNormal, cPoint, cBodies, cBias, cImpulse};
        launch_sis_gather(gi, N, ct, sisW, S > 1, st);
        rec();
        const SisBodies bd{rbPos, rbVel, rbAng, Iw, mass};
        const SisSolvePhys sp{p0.ARfriction, p0.gravity.y};
        if (launch_sis_solve(ct, bd, sp, sisCap, S, B, S > 1 ? sceneBodyD : nullptr, maxSceneBodies, kSisIterations,
                             sisW, stt, st))
            return fail(ARSIM_ERR_INVALID_ARG, "engine: contact sort rejected");
        rec();
        KeyOut none{};
        launch_pose_key(pb, bb, N, ph, sv, kKeyNone, none, true, false, st);
        rec();
        lastKeys = kA;
        lastIdx = vA;
    } else if (method == ARSIM_METHOD_GRID) {
        // the sort's workspace is cleared by the body launch and its histograms are built while the
        // keys are generated, so the sort itself is only its digit passes
        unsigned *kA = nullptr, *vA = nullptr;
        // the DEM below writes every grain's velocity, so the pose kernel's copy is not stored
        if (int rc = enqueue_grid_broadphase(bb, pb, sv, stt, true, ahead, &kA, &vA, rec)) return rc;
        if (fused_dem()) {
            const FusedDemIn in{gatherRec, relPos, grainIdx, numGrains, vel};
            launch_dem_fused(sortedPos, sortedVel, vA, segStart, in, bb, B, ph, sv, S > 1, st);
            rec();
        } else {
            DemOut o{};
            o.pForce = pForce; o.vel = vel; o.relPos = relPos;
            launch_dem_grid_seg(sortedPos, sortedVel, kA, vA, segStart, segBits, N, ph, sv, S > 1, o, tu.demBlock,
                                st);
            if (numStatic)
                launch_dem_static(sortedPos, sortedVel, vA, N, stPos, stNorm, nullptr, stTable, nullptr, ph, geoms[0],
                                  pForce, nullptr, relPos, false, st);
            rec();
            if (ahead) launch_reduce_apply_ahead(pForce, relPos, bb, ahead_bufs(), B, ph, &zHist, st);
            else launch_reduce_apply(pForce, nullptr, relPos, bb, B, ph, st);
            rec();
        }
        prevSorted = kA;
        lastKeys = kA;
        lastIdx = vA;
    } else if (mortonBucket) {
        // Morton codes ranked per tile by their top 10 bits in the pose kernel, then sorted per bucket
        if (!ahead) {
            launch_body_integrate(bb, B, ph, &zHist, st);
            rec();
        }
        TileKeyOut to{(unsigned *)keysB, valsB, tileInfo, bucketHist, bkScenesD, tileSceneD, bucketSceneD,
                      numTiles, numBuckets, tileKeys, bkPlan.bucketBits};
        const bool m63 = method == ARSIM_METHOD_LBVH63;
        to.morton = m63 ? 2 : 1;
        to.code63 = code63;
        launch_pose_key_tiles(pb, ahead ? ahead_bufs() : bb, N, ph, sv, to, false, st);
        rec();
        // (the tree's leaf arrays are the second buffer of a bucket sorted in global memory; the refit
        // writes them afterwards)
        launch_bucket_sort_morton(to, (unsigned *)keysA, valsA, tree.cms, tree.sortedVel, st);
        // 63-bit codes: sorted codes and indices into (keysB, valsB), the tile arrays being free now
        void *kS = keysA;
        unsigned *vS = valsA;
        if (m63) {
            launch_morton63_ties((const unsigned *)keysA, valsA, code63, N, (uint64_t *)keysB, valsB, st);
            kS = keysB;
            vS = valsB;
        }
        rec();
        launch_lbvh_build(kS, keyBytes, N, 0, sceneStartD, sceneCountD, S, tree, st);
        rec();
        launch_lbvh_refit((const int *)vS, pos, vel, rb, nullptr, gatherRec, ph.r, N, tree, st);
        rec();
        DemOut o{};
        o.pForce = pForce; o.vel = vel; o.relPos = relPos;
        launch_dem_lbvh((const int *)vS, tree, N, ph, sv, false, o, tu.demBlock, st, overflowD);
        if (numStatic)
            launch_dem_static(tree.cms, tree.sortedVel, vS, N, stPos, stNorm, nullptr, stTable, nullptr, ph,
                              geoms[0], pForce, nullptr, relPos, false, st);
        rec();
        if (ahead) launch_reduce_apply_ahead(pForce, relPos, bb, ahead_bufs(), B, ph, &zHist, st);
        else launch_reduce_apply(pForce, nullptr, relPos, bb, B, ph, st);
        rec();
        lastKeys = kS;
        lastIdx = vS;
    } else {
        const SortPlan pl = sort_plan(sortWs, N, keyBytes, 0, keyBits, stt);
        const bool fused = pl.passes <= 8;   // the key kernel holds at most 8 digit histograms
        const ZeroJobs z = sort_zero_jobs(pl);
        launch_body_integrate(bb, B, ph, fused ? &z : nullptr, st);
        rec();
        const int mode = S > 1 ? kKeyForest30 : (method == ARSIM_METHOD_LBVH30 ? kKeyMorton30 : kKeyMorton63);
        if (flags & ARSIM_DYNAMIC_AABB) {
            KeyOut none{};
            launch_pose_key(pb, bb, N, ph, sv, kKeyNone, none, true, true, st);
            launch_aabb(pos, N, aabbBox, aabbBox + 1, aabbPart, aabb_max_partials(), st);
            KeyOut ko{keysA, nullptr, aabbBox};
            if (fused) attach_hist(ko, pl, stt, keyBits);
            launch_pose_key(pb, bb, N, ph, sv, mode, ko, false, false, st);
        } else {
            KeyOut ko{keysA, nullptr, nullptr};
            if (fused) attach_hist(ko, pl, stt, keyBits);
            launch_pose_key(pb, bb, N, ph, sv, mode, ko, true, true, st);
        }
        rec();
        void *sortedKeys;
        unsigned *vA = valsA, *vB = valsB;
        if (keyBytes == 4) {
            unsigned *kA = (unsigned *)keysA, *kB = (unsigned *)keysB;
            if (sort_pairs<unsigned>(&kA, &vA, &kB, &vB, N, 0, keyBits, true, sortWs, sortWsBytes, stt, st, fused))
                return fail(ARSIM_ERR_INVALID_ARG, "engine: sort rejected");
            sortedKeys = kA;
        } else {
            uint64_t *kA = (uint64_t *)keysA, *kB = (uint64_t *)keysB;
            if (sort_pairs<uint64_t>(&kA, &vA, &kB, &vB, N, 0, keyBits, true, sortWs, sortWsBytes, stt, st, fused))
                return fail(ARSIM_ERR_INVALID_ARG, "engine: sort rejected");
            sortedKeys = kA;
        }
        rec();
        launch_lbvh_build(sortedKeys, keyBytes, N, S > 1 ? 30 : 0, sceneStartD, sceneCountD, S, tree, st);
        rec();
        launch_lbvh_refit((const int *)vA, pos, vel, rb, S > 1 ? scene : nullptr, gatherRec, ph.r, N, tree, st);
        rec();
        DemOut o{};
        o.pForce = pForce; o.vel = vel; o.relPos = relPos;
        launch_dem_lbvh((const int *)vA, tree, N, ph, sv, S > 1, o, tu.demBlock, st, overflowD);
        if (numStatic)
            launch_dem_static(tree.cms, tree.sortedVel, vA, N, stPos, stNorm, nullptr, stTable, nullptr, ph,
                              geoms[0], pForce, nullptr, relPos, false, st);
        rec();
        launch_reduce_apply(pForce, nullptr, relPos, bb, B, ph, st);
        rec();
        lastKeys = sortedKeys;
        lastIdx = vA;
    }
    first = false;
    if (numStages) *numStages = k - 1;
    return ARSIM_OK;
}

int Engine::step(int n) {
    if (n > 0 && integrate_ahead() && !primed)
        if (int rc = prime()) return rc;
    for (int i = 0; i < n; ++i) {
        if (tu.useGraph && !first) {
            if (!graphReady) {
                hipGraph_t g = nullptr;
                ARSIM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                int rc = enqueue_step(nullptr, nullptr);
                hipError_t e = hipStreamEndCapture(st, &g);
                if (rc) { if (g) hipGraphDestroy(g); return rc; }
                if (e != hipSuccess) return fail(ARSIM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
                e = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
                hipGraphDestroy(g);
                if (e != hipSuccess) return fail(ARSIM_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
                graphReady = true;
            }
            ARSIM_HIP(hipGraphLaunch(gexec, st));
        } else {
            int rc = enqueue_step(nullptr, nullptr);
            if (rc) return rc;
        }
    }
    ARSIM_HIP(hipGetLastError());
    return ARSIM_OK;
}

// Broadphase of the current state (no integration) and contact-list emission.
int Engine::enqueue_broadphase(int *pairsD, int *countsD, int maxPer) {
    invalidate_graph();
    const BodyBufs bb = body_bufs();
    const PartBufs pb = part_bufs();
    const SceneView sv = scene_view();
    SortTuning stt;
    stt.ipt = tu.sortItemsPerThread;
    stt.radixBits = tu.sortRadixBits;
    DemOut o{};
    o.pairs = pairsD; o.counts = countsD; o.maxPer = maxPer;
    if (method == ARSIM_METHOD_GRID) {
        KeyOut ko{keysA, nullptr, nullptr};
        launch_pose_key(pb, bb, N, ph, sv, kKeyGrid, ko, false, false, st);
        unsigned *kA = (unsigned *)keysA, *kB = (unsigned *)keysB, *vA = valsA, *vB = valsB;
        if (sort_pairs<unsigned>(&kA, &vA, &kB, &vB, N, 0, keyBits, true, sortWs, sortWsBytes, stt, st))
            return fail(ARSIM_ERR_INVALID_ARG, "engine: sort rejected");
        launch_fill_u32(longCnt, 0u, 1, st);
        launch_reorder_seg(kA, vA, pos, vel, rb, S > 1 ? scene : nullptr, gatherRec, sortedPos, sortedVel, segStart,
                           segBits, numSegs, N, long_fill(), st);
        launch_dem_grid_seg(sortedPos, sortedVel, kA, vA, segStart, segBits, N, ph, sv, S > 1, o, tu.demBlock, st);
        prevSorted = kA;
        lastKeys = kA;
        lastIdx = vA;
        first = false;
    } else {
        const int mode = S > 1 ? kKeyForest30 : (method == ARSIM_METHOD_LBVH30 ? kKeyMorton30 : kKeyMorton63);
        if (flags & ARSIM_DYNAMIC_AABB) launch_aabb(pos, N, aabbBox, aabbBox + 1, aabbPart, aabb_max_partials(), st);
        KeyOut ko{keysA, nullptr, (flags & ARSIM_DYNAMIC_AABB) ? aabbBox : nullptr};
        launch_pose_key(pb, bb, N, ph, sv, mode, ko, false, false, st);
        void *sortedKeys;
        unsigned *vA = valsA, *vB = valsB;
        if (keyBytes == 4) {
            unsigned *kA = (unsigned *)keysA, *kB = (unsigned *)keysB;
            if (sort_pairs<unsigned>(&kA, &vA, &kB, &vB, N, 0, keyBits, true, sortWs, sortWsBytes, stt, st))
                return fail(ARSIM_ERR_INVALID_ARG, "engine: sort rejected");
            sortedKeys = kA;
        } else {
            uint64_t *kA = (uint64_t *)keysA, *kB = (uint64_t *)keysB;
            if (sort_pairs<uint64_t>(&kA, &vA, &kB, &vB, N, 0, keyBits, true, sortWs, sortWsBytes, stt, st))
                return fail(ARSIM_ERR_INVALID_ARG, "engine: sort rejected");
            sortedKeys = kA;
        }
        launch_lbvh_build(sortedKeys, keyBytes, N, S > 1 ? 30 : 0, sceneStartD, sceneCountD, S, tree, st);
        launch_lbvh_refit((const int *)vA, pos, vel, rb, S > 1 ? scene : nullptr, gatherRec, ph.r, N, tree, st);
        launch_dem_lbvh((const int *)vA, tree, N, ph, sv, S > 1, o, tu.demBlock, st, overflowD);
        lastKeys = sortedKeys;
        lastIdx = vA;
    }
    ARSIM_HIP(hipGetLastError());
    return ARSIM_OK;
}

}  // namespace arsim

using arsim::Engine;

extern "C" {

int arsim_engine_create(const ArSimSceneHost *scenes, int numScenes, const ArSimTuning *tuning, int device,
                        arsim_engine **out) {
    if (!scenes || numScenes <= 0 || !out) return fail(ARSIM_ERR_INVALID_ARG, "engine_create: bad arguments");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(ARSIM_ERR_NO_DEVICE, "no HIP device");
    std::vector<arsim::SceneSource> src(numScenes);
    for (int s = 0; s < numScenes; ++s) src[s].host = &scenes[s];
    arsim_engine *e = new arsim_engine();
    int rc = e->e.build(src, tuning, device, nullptr);
    if (rc) { delete e; return rc; }
    *out = e;
    return ARSIM_OK;
}

int arsim_engine_rebuild(arsim_engine *e, arsim_engine *from, const int *keepScenes, int numKeep,
                         const ArSimSceneHost *added, int numAdded) {
    if (!e || numKeep < 0 || numAdded < 0 || numKeep + numAdded <= 0 || (numKeep > 0 && (!from || !keepScenes)) ||
        (numAdded > 0 && !added) || from == e)
        return fail(ARSIM_ERR_INVALID_ARG, "engine_rebuild: bad arguments");
    std::vector<arsim::SceneSource> src(numKeep + numAdded);
    for (int k = 0; k < numKeep; ++k) {
        if (keepScenes[k] < 0 || keepScenes[k] >= from->e.S)
            return fail(ARSIM_ERR_INVALID_ARG, "engine_rebuild: bad scene index");
        src[k].from = &from->e;
        src[k].fromScene = keepScenes[k];
    }
    for (int k = 0; k < numAdded; ++k) src[numKeep + k].host = &added[k];
    if (from) {
        ARSIM_HIP(hipSetDevice(from->e.device));
        ARSIM_HIP(hipStreamSynchronize(from->e.st));
    }
    const ArSimTuning tu = e->e.tu;
    return e->e.build(src, &tu, e->e.device, nullptr);
}

int arsim_engine_destroy(arsim_engine *e) {
    if (!e) return ARSIM_OK;
    hipSetDevice(e->e.device);
    if (e->e.st) hipStreamSynchronize(e->e.st);
    delete e;
    return ARSIM_OK;
}

int arsim_engine_step(arsim_engine *e, int numSteps) {
    if (!e || numSteps < 0) return fail(ARSIM_ERR_INVALID_ARG, "engine_step: bad arguments");
    return e->e.step(numSteps);
}

int arsim_engine_synchronize(arsim_engine *e) {
    if (!e) return fail(ARSIM_ERR_INVALID_ARG, "null engine");
    ARSIM_HIP(hipStreamSynchronize(e->e.st));
    int ov = 0;
    if (e->e.overflowD) {
        ARSIM_HIP(hipMemcpy(&ov, e->e.overflowD, 4, hipMemcpyDeviceToHost));
        if (ov) return fail(ARSIM_ERR_OVERFLOW, "LBVH traversal stack overflow");
    }
    return ARSIM_OK;
}

int arsim_engine_set_dt(arsim_engine *e, float dt) {
    if (!e || !(dt > 0.f)) return fail(ARSIM_ERR_INVALID_ARG, "engine_set_dt: bad arguments");
    if (e->e.ph.dt != dt) {
        e->e.ph.dt = dt;
        e->e.c0.dt = dt;
        e->e.invalidate_graph();   // dt is a kernel argument of the captured launches
        e->e.primed = false;       // the integration held ahead used the old dt
    }
    return ARSIM_OK;
}

int arsim_engine_stream(arsim_engine *e, arsim_stream_t *stream) {
    if (!e || !stream) return fail(ARSIM_ERR_INVALID_ARG, "null argument");
    *stream = (arsim_stream_t)e->e.st;
    return ARSIM_OK;
}

int arsim_engine_num_particles(arsim_engine *e, int sceneIdx) {
    if (!e) return -1;
    if (sceneIdx < 0) return e->e.N;
    if (sceneIdx >= e->e.S) return -1;
    return e->e.pCount[sceneIdx];
}

size_t arsim_engine_device_bytes(arsim_engine *e) { return e ? e->e.devBytes : 0; }

int arsim_engine_download(arsim_engine *e, int s, float *pos, float *vel, float *relativePos, float *pForce,
                          float *pTorque, float *rbPositions, float *rbVelocities, float *rbAngularVelocity,
                          float *rbQuaternion, float *rbCurrentInertia) {
    if (!e || s < 0 || s >= e->e.S) return fail(ARSIM_ERR_INVALID_ARG, "engine_download: bad scene");
    Engine &E = e->e;
    ARSIM_HIP(hipStreamSynchronize(E.st));
    const size_t po = E.pStart[s], n = E.pCount[s], bo = E.bStart[s], b = E.bCount[s];
    if (pos) ARSIM_HIP(hipMemcpy(pos, E.pos + po, n * 16, hipMemcpyDeviceToHost));
    if (vel) ARSIM_HIP(hipMemcpy(vel, E.vel + po, n * 16, hipMemcpyDeviceToHost));
    if (relativePos) ARSIM_HIP(hipMemcpy(relativePos, E.relPos + po, n * 16, hipMemcpyDeviceToHost));
    std::vector<float> f, r;
    std::vector<int> rbh;
    if (pForce || pTorque) {
        f.resize(4 * n);
        r.resize(4 * n);
        rbh.resize(n);
        ARSIM_HIP(hipMemcpy(f.data(), E.pForce + po, n * 16, hipMemcpyDeviceToHost));
        ARSIM_HIP(hipMemcpy(r.data(), E.relPos + po, n * 16, hipMemcpyDeviceToHost));
        ARSIM_HIP(hipMemcpy(rbh.data(), E.rb + po, n * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            if (rbh[i] < 0) { for (int c = 0; c < 4; ++c) f[4 * i + c] = 0.f; }
        }
        if (pForce) std::memcpy(pForce, f.data(), n * 16);
        if (pTorque) {
            // the contact kernel's per-particle torque, relativePos × pForce
            for (size_t i = 0; i < n; ++i) {
                float rx = r[4 * i], ry = r[4 * i + 1], rz = r[4 * i + 2];
                float fx = f[4 * i], fy = f[4 * i + 1], fz = f[4 * i + 2];
                bool body = rbh[i] >= 0;
                pTorque[4 * i] = body ? ry * fz - rz * fy : 0.f;
                pTorque[4 * i + 1] = body ? rz * fx - rx * fz : 0.f;
                pTorque[4 * i + 2] = body ? rx * fy - ry * fx : 0.f;
                pTorque[4 * i + 3] = 0.f;
            }
        }
    }
    if (b > 0) {
        if (rbPositions) ARSIM_HIP(hipMemcpy(rbPositions, E.rbPos + bo, b * 16, hipMemcpyDeviceToHost));
        if (rbVelocities) ARSIM_HIP(hipMemcpy(rbVelocities, E.rbVel + bo, b * 16, hipMemcpyDeviceToHost));
        if (rbAngularVelocity) ARSIM_HIP(hipMemcpy(rbAngularVelocity, E.rbAng + bo, b * 16, hipMemcpyDeviceToHost));
        if (rbQuaternion) ARSIM_HIP(hipMemcpy(rbQuaternion, E.qcum + bo, b * 16, hipMemcpyDeviceToHost));
        if (rbCurrentInertia) ARSIM_HIP(hipMemcpy(rbCurrentInertia, E.Iw + bo * 9, b * 36, hipMemcpyDeviceToHost));
    }
    return ARSIM_OK;
}

int arsim_engine_state_checksum(arsim_engine *e, int s, uint64_t *out) {
    if (!e) return fail(ARSIM_ERR_INVALID_ARG, "null engine");
    return e->e.state_checksum(s, out);
}

int arsim_engine_set_static(arsim_engine *e, const float *pos4, const float *norm4, int n) {
    if (!e) return fail(ARSIM_ERR_INVALID_ARG, "null engine");
    return e->e.set_static(pos4, norm4, n);
}

// Render state of scene s after the last step: per-body model matrices (column-major mat4, the
// rotation of the cumulative quaternion and the position, rigidBodyKernels.cu:322-332) and the
// cumulative quaternions.  Either pointer may be NULL.
int arsim_engine_render_state(arsim_engine *e, int s, float *modelMatrices, float *cumulativeQuaternions) {
    if (!e || s < 0 || s >= e->e.S) return fail(ARSIM_ERR_INVALID_ARG, "engine_render_state: bad scene");
    Engine &E = e->e;
    ARSIM_HIP(hipSetDevice(E.device));
    ARSIM_HIP(hipStreamSynchronize(E.st));
    const size_t bo = E.bStart[s], b = E.bCount[s];
    if (b == 0) return ARSIM_OK;
    if (modelMatrices) ARSIM_HIP(hipMemcpy(modelMatrices, E.model + bo * 16, b * 64, hipMemcpyDeviceToHost));
    if (cumulativeQuaternions) ARSIM_HIP(hipMemcpy(cumulativeQuaternions, E.qcum + bo, b * 16, hipMemcpyDeviceToHost));
    return ARSIM_OK;
}

// A host copy of scene s that rebuilds the same state: the per-step state (particle pos / vel, body
// position, velocity, angular velocity, cumulative orientation) plus the scene's fixed data (body
// frames, inverse body inertia, masses, particle counts, grid, walls, the current dt).  Quantities a
// step recomputes before reading them (relativePos, the world inertia, the incremental quaternion,
// forces) are not part of it.
int arsim_engine_snapshot(arsim_engine *e, int s, ArSimSceneOwned **out) {
    if (!e || !out || s < 0 || s >= e->e.S) return fail(ARSIM_ERR_INVALID_ARG, "engine_snapshot: bad arguments");
    *out = nullptr;
    Engine &E = e->e;
    ARSIM_HIP(hipSetDevice(E.device));
    ARSIM_HIP(hipStreamSynchronize(E.st));
    const size_t po = E.pStart[s], n = E.pCount[s], bo = E.bStart[s], b = E.bCount[s];
    ArSimSceneOwned *o = new (std::nothrow) ArSimSceneOwned();
    if (!o) return fail(ARSIM_ERR_OUT_OF_MEMORY, "engine_snapshot");
    o->N = (int)n;
    o->B = (int)b;
    o->pos.resize(4 * n); o->vel.resize(4 * n); o->bodyFrame.resize(4 * n); o->rb.resize(n);
    o->rbPos.resize(4 * b); o->rbVel.resize(4 * b); o->rbAng.resize(4 * b); o->rbQuat.resize(4 * b);
    o->rbInertia.resize(9 * b); o->rbMass.resize(b); o->ppo.resize(b);
    auto get = [&](void *dst, const void *src, size_t bytes) {
        return bytes ? hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) : hipSuccess;
    };
    hipError_t err = hipSuccess;
    const struct { void *dst; const void *src; size_t bytes; } copies[] = {
        {o->pos.data(), E.pos + po, n * 16}, {o->vel.data(), E.vel + po, n * 16},
        {o->bodyFrame.data(), E.bodyFrame + po, n * 16}, {o->rb.data(), E.rb + po, n * 4},
        {o->rbPos.data(), E.rbPos + bo, b * 16}, {o->rbVel.data(), E.rbVel + bo, b * 16},
        {o->rbAng.data(), E.rbAng + bo, b * 16}, {o->rbQuat.data(), E.qcum + bo, b * 16},
        {o->rbInertia.data(), E.Ib + bo * 9, b * 36}, {o->rbMass.data(), E.mass + bo, b * 4},
        {o->ppo.data(), E.pcountD + bo, b * 4}};
    for (const auto &c : copies)
        if (err == hipSuccess) err = get(c.dst, c.src, c.bytes);
    if (err != hipSuccess) {
        delete o;
        return fail(ARSIM_ERR_HIP, std::string("engine_snapshot: ") + hipGetErrorString(err));
    }
    for (int &v : o->rb)
        if (v >= 0) v -= (int)bo;
    const arsim::SceneGeom &g = E.geoms[s];
    o->params = E.p0;
    o->params.gridSize = {g.gridMask[0] + 1u, g.gridMask[1] + 1u, g.gridMask[2] + 1u};
    o->params.numCells = o->params.gridSize.x * o->params.gridSize.y * o->params.gridSize.z;
    o->params.worldOrigin = {g.origin[0], g.origin[1], g.origin[2]};
    o->params.cellSize = {g.cellSize[0], g.cellSize[1], g.cellSize[2]};
    o->params.numBodies = (unsigned)b;
    o->cfg = E.c0;
    o->cfg.dt = E.ph.dt;
    for (int a = 0; a < 3; ++a) {
        o->cfg.boxMin[a] = g.boxMin[a];
        o->cfg.boxMax[a] = g.boxMax[a];
    }
    o->cfg.wallMask = g.wallMask;
    *out = o;
    return ARSIM_OK;
}

int arsim_engine_profile_step(arsim_engine *e, float *ms, const char **names, int maxStages) {
    if (!e) return fail(ARSIM_ERR_INVALID_ARG, "null engine");
    Engine &E = e->e;
    int count = 0;
    const char *const *nm =
        Engine::stage_names(E.method, E.flags, E.bucketSort || E.mortonBucket, E.fused_dem(), E.integrate_ahead(),
                            &count);
    std::vector<hipEvent_t> ev(count + 1);
    for (auto &x : ev) ARSIM_HIP(hipEventCreate(&x));
    if (E.integrate_ahead() && !E.primed)
        if (int r = E.prime()) return r;
    int ns = 0;
    int rc = E.enqueue_step(ev.data(), &ns);
    if (rc == 0) {
        ARSIM_HIP(hipEventSynchronize(ev[count]));
        for (int k = 0; k < count && k < maxStages; ++k) {
            float t = 0.f;
            hipEventElapsedTime(&t, ev[k], ev[k + 1]);
            if (ms) ms[k] = t;
            if (names) names[k] = nm[k];
        }
    }
    for (auto &x : ev) hipEventDestroy(x);
    E.invalidate_graph();   // captured graphs stay valid; re-capture keeps the state machine simple
    return rc ? rc : count;
}

int arsim_engine_debug_buffers(arsim_engine *e, void **sortedKeys, void **sortedIndex, void **cellStart,
                               void **cellEnd, void **sortedPos, void **sortedVel, void **lbvhLeft, void **lbvhRight,
                               void **lbvhBounds, void **lbvhCMs) {
    if (!e) return fail(ARSIM_ERR_INVALID_ARG, "null engine");
    Engine &E = e->e;
    if (sortedKeys) *sortedKeys = E.lastKeys;
    if (sortedIndex) *sortedIndex = E.lastIdx;
    if (cellStart) *cellStart = E.segStart;   // the engine's cell index: CSR offsets over cell segments
    if (cellEnd) *cellEnd = nullptr;
    if (sortedPos) *sortedPos = E.sortedPos;
    if (sortedVel) *sortedVel = E.method == ARSIM_METHOD_GRID ? (void *)E.sortedVel : (void *)E.tree.sortedVel;
    if (lbvhLeft) *lbvhLeft = E.tree.left;
    if (lbvhRight) *lbvhRight = E.tree.right;
    if (lbvhBounds) {
        *lbvhBounds = E.treeBounds;
        if (E.treeBounds) {   // the last step's records (uninitialised before the first step)
            ARSIM_HIP(hipSetDevice(E.device));
            arsim::Lbvh t = E.tree;
            t.bounds = E.treeBounds;
            launch_lbvh_bounds(t, E.N, E.ph.r, E.st);
            ARSIM_HIP(hipStreamSynchronize(E.st));
        }
    }
    if (lbvhCMs) *lbvhCMs = E.tree.cms;
    return ARSIM_OK;
}

int arsim_engine_contact_pairs(arsim_engine *e, int *pairsHost, int *countsHost, int maxPer) {
    if (!e || !pairsHost || !countsHost || maxPer <= 0) return fail(ARSIM_ERR_INVALID_ARG, "contact_pairs: bad arguments");
    Engine &E = e->e;
    int *pd = nullptr, *cd = nullptr;
    ARSIM_HIP(hipSetDevice(E.device));
    ARSIM_HIP(hipMalloc(&pd, (size_t)E.N * maxPer * 4));
    if (hipMalloc(&cd, (size_t)E.N * 4) != hipSuccess) { hipFree(pd); return fail(ARSIM_ERR_OUT_OF_MEMORY, "contact_pairs"); }
    int rc = E.enqueue_broadphase(pd, cd, maxPer);
    if (rc == 0) {
        hipStreamSynchronize(E.st);
        hipMemcpy(pairsHost, pd, (size_t)E.N * maxPer * 4, hipMemcpyDeviceToHost);
        hipMemcpy(countsHost, cd, (size_t)E.N * 4, hipMemcpyDeviceToHost);
        for (int i = 0; i < E.N; ++i)
            if (countsHost[i] > maxPer) rc = fail(ARSIM_ERR_OVERFLOW, "contact list longer than maxPerParticle");
    }
    hipFree(pd);
    hipFree(cd);
    return rc;
}

int arsim_engine_check_finite(arsim_engine *e, int *nonFinite) {
    if (!e || !nonFinite) return fail(ARSIM_ERR_INVALID_ARG, "null argument");
    Engine &E = e->e;
    arsim::launch_fill_u32((unsigned *)E.flagD, 0u, 1, E.st);
    arsim::launch_check_finite(E.pos, E.N, E.flagD, E.st);
    arsim::launch_check_finite(E.vel, E.N, E.flagD, E.st);
    arsim::launch_check_finite(E.rbPos, E.B, E.flagD, E.st);
    arsim::launch_check_finite(E.rbVel, E.B, E.flagD, E.st);
    arsim::launch_check_finite(E.rbAng, E.B, E.flagD, E.st);
    ARSIM_HIP(hipStreamSynchronize(E.st));
    int f = 0;
    ARSIM_HIP(hipMemcpy(&f, E.flagD, 4, hipMemcpyDeviceToHost));
    *nonFinite = f;
    return f ? fail(ARSIM_ERR_NAN, "non-finite state") : ARSIM_OK;
}

namespace {
// Autotune results per shape bucket (method, flags, bucket sort or not, key bits, log2 of the particle
// and scene counts), kept for the life of the process: an engine of a shape already tuned takes the
// result without trials.
struct TuneKey {
    int method, flags, bucket, keyBits, logN, logS;
    bool operator<(const TuneKey &o) const {
        const int a[6] = {method, flags, bucket, keyBits, logN, logS}, b[6] = {o.method, o.flags, o.bucket, o.keyBits,
                                                                               o.logN, o.logS};
        for (int k = 0; k < 6; ++k)
            if (a[k] != b[k]) return a[k] < b[k];
        return false;
    }
};
std::mutex g_tuneMutex;
std::map<TuneKey, std::pair<ArSimTuning, float>> g_tuneCache;
int log2_bucket(long long v) {
    int l = 0;
    while ((1ll << l) < v) ++l;
    return l;
}
}  // namespace

int arsim_engine_autotune(arsim_engine *e, int stepsPerTrial, ArSimTuning *best, float *bestMs) {
    if (!e || stepsPerTrial <= 0) return fail(ARSIM_ERR_INVALID_ARG, "autotune: bad arguments");
    Engine &E = e->e;
    const TuneKey key{E.method, (int)E.flags, (E.bucketSort || E.mortonBucket) ? 1 : 0, E.keyBits, log2_bucket(E.N),
                      log2_bucket(E.S)};
    {
        std::lock_guard<std::mutex> lock(g_tuneMutex);
        auto it = g_tuneCache.find(key);
        if (it != g_tuneCache.end()) {
            ArSimTuning t = it->second.first;
            t.useGraph = E.tu.useGraph;
            t.fuseReduce = E.tu.fuseReduce;
            E.tu = t;
            E.reset_grid_tables();
            if (best) *best = t;
            if (bestMs) *bestMs = it->second.second;
            return ARSIM_OK;
        }
    }
    // with the bucket sort (and no contact-list sort) the onesweep tuning is unused: only the block size
    // and the workgroups per bucket
    const bool sortUnused = (E.method == ARSIM_METHOD_GRID && E.bucketSort && !(E.flags & ARSIM_CONTACT_SIS)) ||
                            E.mortonBucket;
    // the trials advance a scratch copy of the state: snapshot everything a step writes
    struct Snap { void *dst; const void *src; size_t bytes; };
    std::vector<Snap> snaps = {
        {nullptr, E.pos, (size_t)E.N * 16},  {nullptr, E.vel, (size_t)E.N * 16},
        {nullptr, E.relPos, (size_t)E.N * 16}, {nullptr, E.pForce, (size_t)E.N * 16},
        {nullptr, E.rbPos, (size_t)E.B * 16}, {nullptr, E.rbVel, (size_t)E.B * 16},
        {nullptr, E.rbAng, (size_t)E.B * 16}, {nullptr, E.qinc, (size_t)E.B * 16},
        {nullptr, E.qcum, (size_t)E.B * 16}, {nullptr, E.Iw, (size_t)E.B * 36}};
    int rc = ARSIM_OK;
    for (auto &s : snaps) {
        if (s.bytes == 0) continue;
        if (hipMalloc(&s.dst, s.bytes) != hipSuccess) { rc = fail(ARSIM_ERR_OUT_OF_MEMORY, "autotune snapshot"); break; }
        hipMemcpyAsync(s.dst, s.src, s.bytes, hipMemcpyDeviceToDevice, E.st);
    }
    const ArSimTuning orig = E.tu;
    ArSimTuning bestT = orig;
    float bestT_ms = 1e30f;
    const int blocks[] = {64, 128, 256};
    const int ipts[] = {4, 8, 12, 16};
    const int rbits[] = {7, 8};
    hipEvent_t a, b;
    hipEventCreate(&a);
    hipEventCreate(&b);
    // one trial: restore the state, two steps to settle, then time stepsPerTrial steps
    auto trial = [&](const ArSimTuning &t) -> int {
        for (auto &s : snaps)
            if (s.dst) hipMemcpyAsync((void *)s.src, s.dst, s.bytes, hipMemcpyDeviceToDevice, E.st);
        E.primed = false;   // the state was restored: the integration ahead is recomputed from it
        E.tu = t;
        E.reset_grid_tables();
        if (int r = E.step(2)) return r;
        hipEventRecord(a, E.st);
        if (int r = E.step(stepsPerTrial)) return r;
        hipEventRecord(b, E.st);
        hipEventSynchronize(b);
        float ms = 0.f;
        hipEventElapsedTime(&ms, a, b);
        ms /= stepsPerTrial;
        if (ms < bestT_ms) { bestT_ms = ms; bestT = t; }
        return ARSIM_OK;
    };
    // the bucket sort's workgroups per bucket are tried in place of the onesweep tuning
    const int splits[] = {1, 2};
    for (int bl : blocks)
        for (int ip : ipts)
            for (int rb : rbits) {
                if (rc) break;
                if (!arsim::sort_tuning_supported(ip, rb)) continue;
                if (sortUnused && (ip != orig.sortItemsPerThread || rb != orig.sortRadixBits)) continue;
                if (!E.bucketSort || E.method != ARSIM_METHOD_GRID) {
                    rc = trial(ArSimTuning{bl, ip, rb, orig.useGraph, orig.fuseReduce, orig.bucketSplit});
                    continue;
                }
                for (int sp : splits)
                    if (!rc) rc = trial(ArSimTuning{bl, ip, rb, orig.useGraph, orig.fuseReduce, sp});
            }
    hipEventDestroy(a);
    hipEventDestroy(b);
    for (auto &s : snaps)
        if (s.dst) hipMemcpyAsync((void *)s.src, s.dst, s.bytes, hipMemcpyDeviceToDevice, E.st);
    E.primed = false;
    hipStreamSynchronize(E.st);
    for (auto &s : snaps)
        if (s.dst) hipFree(s.dst);
    E.tu = rc ? orig : bestT;
    E.reset_grid_tables();
    hipStreamSynchronize(E.st);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lock(g_tuneMutex);
        g_tuneCache[key] = {bestT, bestT_ms};
    }
    if (best) *best = bestT;
    if (bestMs) *bestMs = bestT_ms;
    return ARSIM_OK;
}

}  // extern "C"
