// body.hip -- rigid-body integration, particle pose / grain integration with key generation,
// the segmented per-body reduction, and small utility kernels.
#include "kernels.hpp"
#include "reduce_apply.hpp"

namespace arsim {

namespace {

constexpr int kMaxHistWords = 8 * 256;   // up to 8 digit passes of up to 8 bits

// Zeroes the ranges of z with blocks [0, blocks) of the caller's extra blocks.
__device__ __forceinline__ void zero_jobs(const ZeroJobs &z, int block, int blocks) {
    const size_t t = (size_t)block * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)blocks * blockDim.x;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        for (size_t e = t; e < z.count[k]; e += stride) z.p[k][e] = 0u;
}

// GPUintegratorKernel (Kernels/rigidBodyKernels.cu:262-348), with the orientation carried as a
// cumulative quaternion (q_cum = normalize(q_inc * q_cum)) and I_w^-1 = R_cum I_b^-1 R_cum^T.
// Blocks past bodyBlocks zero the given ranges (the next sort's histograms, tile counters and
// first look-back array, the cell table's long-run count), which saves those a launch of their own.
__global__ void k_body_integrate(BodyBufs b, BodyBufs o, int B, DevPhys ph, ZeroJobs z, int bodyBlocks) {
    if ((int)blockIdx.x >= bodyBlocks) {
        zero_jobs(z, (int)blockIdx.x - bodyBlocks, (int)gridDim.x - bodyBlocks);
        return;
    }
    int bi = blockIdx.x * blockDim.x + threadIdx.x;
    if (bi >= B) return;
    const float4 X = b.pos[bi], V = b.vel[bi], W = b.ang[bi];
    const BodyStep bs = body_step(X, V, W, b.qcum[bi], ph);
    float R[9], Iw[9];
    body_world_inertia(bs.qc, b.Ib + (size_t)bi * 9, R, Iw);
#pragma unroll
    for (int k = 0; k < 9; ++k) o.Iw[(size_t)bi * 9 + k] = Iw[k];
    const f3 x = bs.x;
    o.pos[bi] = f4(x, X.w);
    o.vel[bi] = f4(bs.v, V.w);
    o.ang[bi] = f4(bs.w, W.w);
    o.qinc[bi] = make_float4(bs.qi.x, bs.qi.y, bs.qi.z, bs.qi.w);
    o.qcum[bi] = make_float4(bs.qc.x, bs.qc.y, bs.qc.z, bs.qc.w);
    write_model(o, bi, R, x);
}

// computeGlobalAttributes (rigidBodyKernels.cu:386-421) for body particles, integrate_functor
// (particleKernels.cu:50-129) for grains, then the broadphase key of the updated position:
// calcHashD (particleKernels.cu:170-189) or generateMortonCodes (BVHkernels.cu:670-716).
// Where pose_key_one takes the scene geometry from: each particle's scene (per lane), the table entry
// of scene `scene` for all of the caller's particles (scalar loads), or the one scene passed by value.
// The loads a particle's pose starts from: a body particle's cumulative quaternion, body frame,
// body centre, velocity and angular velocity (a .. e), or a grain's position and velocity (a, b).
struct PoseIn {
    float4 a, b, c, d, e;
};
__device__ __forceinline__ PoseIn pose_load(int i, int rbi, const PartBufs &p, const BodyBufs &b, int doBodies,
                                            int doGrains) {
    PoseIn in;
    if (rbi >= 0) {
        if (doBodies) {
            in.a = b.qcum[rbi];
            in.b = p.bodyFrame[i];
            in.c = b.pos[rbi];
            in.d = b.vel[rbi];
            in.e = b.ang[rbi];
        } else {
            in.a = p.pos[i];
        }
    } else {
        in.a = p.pos[i];
        if (doGrains) in.b = p.vel[i];
    }
    return in;
}
// Every load of a pose precedes its first store (pose_load, then pose_key_in): the compiler may not
// move a load above a store to a buffer it cannot tell apart, so a load after a store would cost a
// round trip of its own.
enum { kPoseGeomLane = 0, kPoseGeomScene = 1, kPoseGeomSingle = 2 };
template <int MODE, int GSRC = kPoseGeomLane>
__device__ __forceinline__ uint64_t pose_key_one(int i, PartBufs p, BodyBufs b, int N, DevPhys ph, SceneView sv,
                                                 KeyOut ko, int doBodies, int doGrains, unsigned *shist,
                                                 int scene = 0);
template <int MODE>
__device__ __forceinline__ uint64_t pose_key_in(int i, int rbi, int s, const SceneGeom &g, const PoseIn &in,
                                                PartBufs p, DevPhys ph, KeyOut ko, int doBodies, int doGrains,
                                                unsigned *shist);

// Adds 1 to an LDS histogram bin for every active lane.  Spatially ordered keys share digits across
// a wave, so lanes are grouped by digit (8 ballots find the lanes with the same 8-bit digit) and
// the lowest lane of each group adds the group's size: at most one LDS atomic per distinct digit.
__device__ __forceinline__ void hist_add(unsigned *h, unsigned d) {
    uint64_t peers = __ballot(1);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t m = __ballot(bit);
        peers &= bit ? m : ~m;
    }
    const unsigned lane = threadIdx.x & 63;
    if ((unsigned)(__ffsll((long long)peers) - 1) == lane) atomicAdd(&h[d], (unsigned)__popcll(peers));
}

template <int MODE, int GSRC>
__global__ __launch_bounds__(256) void k_pose_key(PartBufs p, BodyBufs b, int N, DevPhys ph, SceneView sv, KeyOut ko,
                                                  int doBodies, int doGrains) {
    // digit histograms of this block's keys for every sort pass (when the sort asked for them)
    __shared__ unsigned shist[kMaxHistWords];
    const bool hist = MODE != kKeyNone && ko.hist != nullptr;
    if (hist) {
        for (int k = threadIdx.x; k < ko.histPasses * 256; k += blockDim.x) shist[k] = 0u;
        __syncthreads();
    }
    // a block handles several 256-particle chunks (grid-stride), so that fewer block histograms
    // are flushed into the global ones: the flush is a global atomic per non-empty bin
    const int stride = (int)(gridDim.x * blockDim.x);
    if constexpr (GSRC == kPoseGeomSingle || GSRC == kPoseGeomLane) {
        // particles in pairs (i, i + stride): both particles' loads are issued before either's stores, so
        // a pair costs the round trips of one particle (a load after a store also waits for the store)
        for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < N; i += 2 * stride) {
            const int j = i + stride;
            const bool hasB = j < N;
            const int rbA = p.rb ? p.rb[i] : -1;
            const int rbB = hasB && p.rb ? p.rb[j] : -1;
            const int sA = GSRC == kPoseGeomLane && p.scene ? (int)p.scene[i] : 0;
            const int sB = GSRC == kPoseGeomLane && p.scene && hasB ? (int)p.scene[j] : 0;
            SceneGeom gA, gB;
            if (GSRC == kPoseGeomSingle) {
                gA = sv.single;
                gB = sv.single;
            } else {
                load_geom(sv, sA, gA);
                load_geom(sv, sB, gB);
            }
            const PoseIn inA = pose_load(i, rbA, p, b, doBodies, doGrains);
            PoseIn inB{};
            if (hasB) inB = pose_load(j, rbB, p, b, doBodies, doGrains);
            pose_key_in<MODE>(i, rbA, sA, gA, inA, p, ph, ko, doBodies, doGrains, hist ? shist : nullptr);
            if (hasB) pose_key_in<MODE>(j, rbB, sB, gB, inB, p, ph, ko, doBodies, doGrains, hist ? shist : nullptr);
        }
    } else {
        for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < N; i += stride)
            pose_key_one<MODE, GSRC>(i, p, b, N, ph, sv, ko, doBodies, doGrains, hist ? shist : nullptr);
    }
    if (hist) {
        __syncthreads();
        for (int k = threadIdx.x; k < ko.histPasses * 256; k += blockDim.x)
            if (shist[k]) atomicAdd(&ko.hist[k], shist[k]);
    }
}

// Returns the key (0 without one); a grid key is stored only when ko.keys is set.  GSRC: where the
// scene geometry comes from (kPoseGeomScene: the caller's particles all belong to scene `scene`, read
// from the scene table, which holds every scene of an engine, also a single one).
// The pose and key of particle i (rbIndex rbi, scene s with geometry g) from its loaded inputs.
template <int MODE>
__device__ __forceinline__ uint64_t pose_key_in(int i, int rbi, int s, const SceneGeom &g, const PoseIn &in,
                                                PartBufs p, DevPhys ph, KeyOut ko, int doBodies, int doGrains,
                                                unsigned *shist) {
    float4 P, Vfin;
    bool haveV = false;
    if (rbi >= 0) {
        if (doBodies) {
            const float4 Qc = in.a, frame = in.b, CM = in.c, V = in.d, W = in.e;
            float R[9];
            mat3_cast(qf{Qc.x, Qc.y, Qc.z, Qc.w}, R);
            f3 r = mat_vec(R, xyz(frame));
            p.relPos[i] = f4(r, 0.0f);
            P = f4(add3(xyz(CM), r), CM.w);
            p.pos[i] = P;
            Vfin = f4(add3(xyz(V), cross3(xyz(W), r)), V.w);
            haveV = true;
            p.vel[i] = Vfin;
        } else {
            P = in.a;
        }
    } else {
        P = in.a;
        if (doGrains) {
            const float4 Vv = in.b;
            f3 x = xyz(P), v = xyz(Vv);
            v = add3(v, mk3(ph.gx * ph.dt, ph.gy * ph.dt, ph.gz * ph.dt));
            v = mk3(v.x * ph.globalDamping, v.y * ph.globalDamping, v.z * ph.globalDamping);
            x = add3(x, mk3(v.x * ph.dt, v.y * ph.dt, v.z * ph.dt));
            if (ph.flags & kGrainClamp) {
                // reference box clamp: 0.01 margin added in double, then compared in float
                float hx = (float)((double)g.boxMax[0] + 0.01), hy = (float)((double)g.boxMax[1] + 0.01),
                      hz = (float)((double)g.boxMax[2] + 0.01);
                float lx = (float)((double)g.boxMin[0] - 0.01), ly = (float)((double)g.boxMin[1] - 0.01),
                      lz = (float)((double)g.boxMin[2] - 0.01);
                if (x.x > hx - ph.r) { x.x = hx - ph.r; v.x *= ph.boundaryDamping; }
                if (x.x < lx + ph.r) { x.x = lx + ph.r; v.x *= ph.boundaryDamping; }
                if (x.y > hy - ph.r && v.y > 0.0f) { x.y = hy - ph.r; v.y *= ph.boundaryDamping; }
                if (x.z > hz - ph.r) { x.z = hz - ph.r; v.z *= ph.boundaryDamping; }
                if (x.z < lz + ph.r) { x.z = lz + ph.r; v.z *= ph.boundaryDamping; }
                if (x.y < ly + ph.r) { x.y = ly + ph.r; v.y *= ph.boundaryDamping; }
            }
            P = f4(x, P.w);
            p.pos[i] = P;
            Vfin = f4(v, Vv.w);
            haveV = true;
            if (!ko.grainVelDead) p.vel[i] = Vfin;
        }
    }
    if (MODE == kKeyNone) return 0;
    if (p.rec) {
        if (!haveV) Vfin = p.vel[i];
        p.rec[2 * (size_t)i] = make_float4(P.x, P.y, P.z, __int_as_float(rbi));
        p.rec[2 * (size_t)i + 1] = make_float4(Vfin.x, Vfin.y, Vfin.z, p.scene ? __int_as_float(s) : Vfin.w);
    }
    uint64_t key;
    if (MODE == kKeyGrid) {
        int gx, gy, gz;
        grid_pos(xyz(P), g, gx, gy, gz);
        key = g.cellBase + grid_hash(gx, gy, gz, g);
        if (ko.keys) ((unsigned *)ko.keys)[i] = (unsigned)key;
    } else {
        float lo0 = g.mLo[0], lo1 = g.mLo[1], lo2 = g.mLo[2];
        float e0 = g.mExt[0], e1 = g.mExt[1], e2 = g.mExt[2];
        if (ko.box) {
            float4 mn = ko.box[0], mx = ko.box[1];
            lo0 = mn.x; lo1 = mn.y; lo2 = mn.z;
            e0 = mx.x - mn.x; e1 = mx.y - mn.y; e2 = mx.z - mn.z;
        }
        float u = mnorm(P.x, lo0, e0), v = mnorm(P.y, lo1, e1), w = mnorm(P.z, lo2, e2);
        if (MODE == kKeyMorton30) {
            key = morton30(u, v, w);
            if (ko.keys) ((unsigned *)ko.keys)[i] = (unsigned)key;
        } else if (MODE == kKeyMorton63) {
            key = morton63(u, v, w);
            if (ko.keys) ((uint64_t *)ko.keys)[i] = key;
        } else {
            key = ((uint64_t)(unsigned)s << 30) | (uint64_t)morton30(u, v, w);
            ((uint64_t *)ko.keys)[i] = key;
        }
    }
    if (ko.index) ko.index[i] = (unsigned)i;
    if (shist) {
        for (int pp = 0; pp < ko.histPasses; ++pp) {
            const int sh0 = ko.histBase + pp * ko.histRbits;
            const int bits = ko.histEndBit - sh0 < ko.histRbits ? ko.histEndBit - sh0 : ko.histRbits;
            const unsigned d = (unsigned)(key >> sh0) & ((1u << bits) - 1u);
            hist_add(shist + pp * 256, d);
        }
    }
    return key;
}

template <int MODE, int GSRC>
__device__ __forceinline__ uint64_t pose_key_one(int i, PartBufs p, BodyBufs b, int N, DevPhys ph, SceneView sv,
                                                 KeyOut ko, int doBodies, int doGrains, unsigned *shist, int scene) {
    if (i >= N) return 0;
    const int rbi = p.rb ? p.rb[i] : -1;
    const PoseIn in = pose_load(i, rbi, p, b, doBodies, doGrains);
    const int s = GSRC == kPoseGeomScene ? scene : GSRC == kPoseGeomSingle ? 0 : (p.scene ? (int)p.scene[i] : 0);
    SceneGeom g;
    if (GSRC == kPoseGeomScene) g = sv.table[scene];
    else if (GSRC == kPoseGeomSingle) g = sv.single;
    else load_geom(sv, s, g);
    return pose_key_in<MODE>(i, rbi, s, g, in, p, ph, ko, doBodies, doGrains, shist);
}

// Pose and grid key of one tile of up to THREADS * 4 consecutive particles of one scene, the tile then
// ranked by the key's bucket (its scene's cell range cut into 2^shift-cell ranges) in LDS and written
// in bucket order (stable: particle order within a bucket), with the tile's per-bucket count and first
// slot in info[infoBase + j * tiles + k] = count << 16 | first.  This stands in for the first pass of
// the grid sort: the bucket kernel gathers each bucket's keys from its scene's tiles' runs, in tile
// order.  Wave w owns particles [tile + 256 w, tile + 256 w + 256), four per lane, so (wave, item,
// lane) order is particle order.
constexpr int kTileItems = 4;   // particles per thread

template <int BITS>
__device__ __forceinline__ uint64_t match_bits(unsigned d, bool ok) {
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
        const unsigned t = 0u - ((d >> b) & 1u);
        const uint64_t m = __ballot(t != 0u);
        const unsigned lo = ~((unsigned)m ^ t), hi = ~((unsigned)(m >> 32) ^ t);
        peers &= ((uint64_t)hi << 32) | lo;
    }
    return peers;
}

template <int MODE, int THREADS, int BITS>
__global__ __launch_bounds__(THREADS) void k_pose_key_tiles(PartBufs p, BodyBufs b, int N, DevPhys ph, SceneView sv,
                                                            TileKeyOut to, int grainVelDead) {
    constexpr int WAVES = THREADS / 64;
    constexpr int TILE = THREADS * kTileItems;
    constexpr int RADIX = 1 << BITS;   // buckets a scene may have
    static_assert(RADIX <= THREADS, "a thread per bucket");
    __shared__ unsigned sWaveHist[WAVES][RADIX];
    __shared__ unsigned sDigitStart[RADIX];
    __shared__ unsigned sScan[WAVES];
    const int t = (int)threadIdx.x, lane = t & 63, wave = t >> 6;
    for (int k = t; k < WAVES * RADIX; k += THREADS) (&sWaveHist[0][0])[k] = 0u;
    const int scene = to.tileScene[blockIdx.x];
    const BucketScene bs = to.scenes[scene];
    const int kt = (int)blockIdx.x - bs.tileBase;   // the tile's place among its scene's tiles
    const long long tileBase = (long long)bs.pStart + (long long)kt * TILE;
    const long long sceneEnd = (long long)bs.pStart + bs.n;
    const long long tileEnd = tileBase + TILE < sceneEnd ? tileBase + TILE : sceneEnd;
    const long long waveBase = tileBase + (long long)wave * (kTileItems * 64);
    KeyOut ko{};
    ko.grainVelDead = grainVelDead != 0;
    if (MODE == kKeyMorton63) ko.keys = to.code63;   // the full codes, in particle order
    // the scene's geometry and every item's rbIndex requested first: the items' poses depend on them,
    // and the stores of one item's pose keep the compiler from moving the next item's loads above them
    const SceneGeom geom = sv.table[scene];
    int rbv[kTileItems];
#pragma unroll
    for (int k = 0; k < kTileItems; ++k) {
        const long long i = waveBase + k * 64 + lane;
        rbv[k] = i < tileEnd && p.rb ? p.rb[i] : -1;
    }
    // wait for them here, on every path: otherwise the wait for a later item's rbIndex, placed after
    // the previous item's stores, waits for those stores as well (vector memory counter 0, the rest
    // of the counters left at their maxima in the gfx9 encoding)
    __builtin_amdgcn_s_waitcnt(0x0f70);
    // items in pairs: both items' loads are issued before either item's stores, so a pair costs one
    // round trip (a load after a store also waits for the store)
    unsigned keys[kTileItems];
    static_assert(kTileItems % 2 == 0, "items in pairs");
#pragma unroll
    for (int k = 0; k < kTileItems; k += 2) {
        const long long iA = waveBase + k * 64 + lane, iB = iA + 64;
        PoseIn inA{}, inB{};
        if (iA < tileEnd) inA = pose_load((int)iA, rbv[k], p, b, 1, 1);
        if (iB < tileEnd) inB = pose_load((int)iB, rbv[k + 1], p, b, 1, 1);
        const uint64_t keyA =
            iA < tileEnd ? pose_key_in<MODE>((int)iA, rbv[k], scene, geom, inA, p, ph, ko, 1, 1, nullptr) : 0ull;
        const uint64_t keyB =
            iB < tileEnd ? pose_key_in<MODE>((int)iB, rbv[k + 1], scene, geom, inB, p, ph, ko, 1, 1, nullptr) : 0ull;
        // a 63-bit code's top 30 bits are the 30-bit code of the same position (both scale the
        // normalised coordinate by a power of two and truncate)
        keys[k] = MODE == kKeyMorton63 ? (unsigned)(keyA >> 33) : (unsigned)keyA;
        keys[k + 1] = MODE == kKeyMorton63 ? (unsigned)(keyB >> 33) : (unsigned)keyB;
    }
    __syncthreads();
    // ranks within the wave, in particle order
    const uint64_t lower = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    unsigned rank[kTileItems];
#pragma unroll
    for (int k = 0; k < kTileItems; ++k) {
        const bool valid = waveBase + k * 64 + lane < tileEnd;
        const unsigned d = valid ? (keys[k] - bs.cellBase) >> bs.shift : 0u;
        const uint64_t peers = match_bits<BITS>(d, valid);
        const unsigned prior = (unsigned)__popcll(peers & lower);
        const unsigned old = sWaveHist[wave][d];
        rank[k] = old + prior;
        if (valid && prior == 0u) sWaveHist[wave][d] = old + (unsigned)__popcll(peers);
    }
    __syncthreads();
    // per digit: exclusive over waves and the tile's count, then the digits' first slots
    unsigned cnt = 0u;
    if (t < RADIX) {
        unsigned run = 0u;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
            const unsigned c = sWaveHist[w][t];
            sWaveHist[w][t] = run;
            run += c;
        }
        cnt = run;
    }
    unsigned x = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) sScan[wave] = x;
    __syncthreads();
    unsigned before = 0u;
    for (int w = 0; w < wave; ++w) before += sScan[w];
    const unsigned first = before + x - cnt;
    if (t < RADIX) {
        sDigitStart[t] = first;
        if (t < bs.buckets) {
            to.info[(size_t)bs.infoBase + (size_t)t * bs.tiles + kt] = (cnt << 16) | first;
            if (cnt) atomicAdd(&to.hist[bs.bucketBase + t], cnt);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kTileItems; ++k) {
        const long long i = waveBase + k * 64 + lane;
        if (i < tileEnd) {
            const unsigned d = (keys[k] - bs.cellBase) >> bs.shift;
            const unsigned pos = sDigitStart[d] + sWaveHist[wave][d] + rank[k];
            to.keys[tileBase + pos] = keys[k];
            to.vals[tileBase + pos] = (unsigned)i;
        }
    }
}

// Segmented per-body reduction + apply (the intended sum of UniformGridDEM.cu:421-460).
// One wave per body: lane l sums particles l, l+64, ... of the body's contiguous range, then a
// butterfly over lane distance 32..1; lane 0 applies v += J dt / m and ω += I_w^-1 τ dt.
// Torque per particle is pTorque when given, else relativePos × pForce (the same expression
// the contact kernel would have stored).
__global__ __launch_bounds__(256) void k_reduce_apply(const float4 *__restrict__ pF, const float4 *__restrict__ pT,
                                                      const float4 *__restrict__ rel, BodyBufs b, int B, DevPhys ph) {
    const int body = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (body >= B) return;
    const int st = b.pstart[body], cnt = b.pcount[body];
    asm volatile("" : : "v"(st), "v"(cnt));   // both loads issued together (k_reduce_apply_ahead)
    f3 aF = mk3(0.f, 0.f, 0.f), aT = mk3(0.f, 0.f, 0.f);
    for (int j = reduce_first(body, ph) + lane; j < cnt; j += 64) {
        size_t o = (size_t)st + j;
        f3 F = xyz(pF[o]);
        aF = add3(aF, F);
        f3 T = pT ? xyz(pT[o]) : cross3(xyz(rel[o]), F);
        aT = add3(aT, T);
    }
    wave_reduce_apply(aF, aT, lane, body, b, ph);
}

// The reduction of a step whose bodies' integration the previous step stored ahead (in `a`): one wave
// per body sums its particles' forces and torques as k_reduce_apply does (lane 0's butterfly value),
// then one thread per body applies them to the integrated state, stores the state after the step in
// `b` (what the integration kernel and k_reduce_apply would have left there) and the integration of
// that state for the next step in `a`.
constexpr int kAheadBodies = 4;   // bodies per workgroup of 256 threads
__global__ __launch_bounds__(256) void k_reduce_apply_ahead(const float4 *__restrict__ pF,
                                                            const float4 *__restrict__ rel, BodyBufs b, BodyBufs a,
                                                            int B, DevPhys ph, ZeroJobs z, int bodyBlocks) {
    __shared__ float sSum[kAheadBodies][6];
    if ((int)blockIdx.x >= bodyBlocks) {
        zero_jobs(z, (int)blockIdx.x - bodyBlocks, (int)gridDim.x - bodyBlocks);
        return;
    }
    const int wave = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    const int body = (int)blockIdx.x * kAheadBodies + wave;
    f3 aF = mk3(0.f, 0.f, 0.f), aT = mk3(0.f, 0.f, 0.f);
    if (body < B) {
        const int st = b.pstart[body], cnt = b.pcount[body];
        // both values needed here, so that their loads are issued together (the compiler would
        // otherwise load the start only once the count has let the loop begin: two round trips)
        asm volatile("" : : "v"(st), "v"(cnt));
        for (int j = reduce_first(body, ph) + lane; j < cnt; j += 64) {
            const size_t o = (size_t)st + j;
            const f3 F = xyz(pF[o]);
            aF = add3(aF, F);
            aT = add3(aT, cross3(xyz(rel[o]), F));
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        aF.x = aF.x + __shfl_xor(aF.x, off, 64);
        aF.y = aF.y + __shfl_xor(aF.y, off, 64);
        aF.z = aF.z + __shfl_xor(aF.z, off, 64);
        aT.x = aT.x + __shfl_xor(aT.x, off, 64);
        aT.y = aT.y + __shfl_xor(aT.y, off, 64);
        aT.z = aT.z + __shfl_xor(aT.z, off, 64);
    }
    if (lane == 0) {
        sSum[wave][0] = aF.x; sSum[wave][1] = aF.y; sSum[wave][2] = aF.z;
        sSum[wave][3] = aT.x; sSum[wave][4] = aT.y; sSum[wave][5] = aT.z;
    }
    __syncthreads();
    const int k = (int)threadIdx.x;
    const int bd = (int)blockIdx.x * kAheadBodies + k;
    if (k >= kAheadBodies || bd >= B) return;
    const f3 sF = mk3(sSum[k][0], sSum[k][1], sSum[k][2]), sT = mk3(sSum[k][3], sSum[k][4], sSum[k][5]);
    // every load before the first store: the compiler may not move a load above a store to a buffer
    // it cannot tell apart, so a load after a store would cost a round trip of its own
    const float m = b.mass[bd];
    float Iw[9], Ib[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) Iw[q] = a.Iw[(size_t)bd * 9 + q];
#pragma unroll
    for (int q = 0; q < 9; ++q) Ib[q] = b.Ib[(size_t)bd * 9 + q];
    float4 V = a.vel[bd], W = a.ang[bd];
    const float4 X = a.pos[bd], Qc = a.qcum[bd], Qi = a.qinc[bd];
    // the apply of wave_reduce_apply, on the integrated state
    const float sc = (ph.flags & kRefUnits) ? 1.0f : ph.dt;
    const f3 dv = mk3((sF.x * sc) / m, (sF.y * sc) / m, (sF.z * sc) / m);
    const f3 dw = mat_vec(Iw, mk3(sT.x * sc, sT.y * sc, sT.z * sc));
    V.x += dv.x; V.y += dv.y; V.z += dv.z;
    W.x += dw.x; W.y += dw.y; W.z += dw.z;
    // the state after the step
    b.pos[bd] = X;
    b.vel[bd] = V;
    b.ang[bd] = W;
    b.qinc[bd] = Qi;
    b.qcum[bd] = Qc;
#pragma unroll
    for (int q = 0; q < 9; ++q) b.Iw[(size_t)bd * 9 + q] = Iw[q];
    if (b.model) {
        float R[9];
        mat3_cast(qf{Qc.x, Qc.y, Qc.z, Qc.w}, R);
        write_model(b, bd, R, xyz(X));
    }
    // the next step's integration of it
    const BodyStep bs = body_step(X, V, W, Qc, ph);
    float R2[9], Iw2[9];
    body_world_inertia(bs.qc, Ib, R2, Iw2);
#pragma unroll
    for (int q = 0; q < 9; ++q) a.Iw[(size_t)bd * 9 + q] = Iw2[q];
    a.pos[bd] = f4(bs.x, X.w);
    a.vel[bd] = f4(bs.v, V.w);
    a.ang[bd] = f4(bs.w, W.w);
    a.qinc[bd] = make_float4(bs.qi.x, bs.qi.y, bs.qi.z, bs.qi.w);
    a.qcum[bd] = make_float4(bs.qc.x, bs.qc.y, bs.qc.z, bs.qc.w);
}

__global__ void k_fill_u32(unsigned *p, unsigned v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

constexpr int kAabbBlocks = 512;

__device__ __forceinline__ float4 min4(float4 a, float4 b) {
    return make_float4(b.x < a.x ? b.x : a.x, b.y < a.y ? b.y : a.y, b.z < a.z ? b.z : a.z, b.w < a.w ? b.w : a.w);
}
__device__ __forceinline__ float4 max4(float4 a, float4 b) {
    return make_float4(b.x > a.x ? b.x : a.x, b.y > a.y ? b.y : a.y, b.z > a.z ? b.z : a.z, b.w > a.w ? b.w : a.w);
}

// findAABBCub (BVHSource/createMortonCodes.cu:37-107): exact min / max, two stages.
__global__ __launch_bounds__(256) void k_aabb_partial(const float4 *__restrict__ pos, int n, float4 *partials) {
    __shared__ float4 smin[256], smax[256];
    const float inf = __int_as_float(0x7f800000);
    float4 mn = make_float4(inf, inf, inf, inf), mx = make_float4(-inf, -inf, -inf, -inf);
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        float4 p = pos[i];
        mn = min4(mn, p);
        mx = max4(mx, p);
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smin[threadIdx.x] = min4(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = max4(smax[threadIdx.x], smax[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = smin[0];
        partials[2 * blockIdx.x + 1] = smax[0];
    }
}
__global__ __launch_bounds__(256) void k_aabb_final(const float4 *partials, int np, float4 *outMin, float4 *outMax) {
    __shared__ float4 smin[256], smax[256];
    const float inf = __int_as_float(0x7f800000);
    float4 mn = make_float4(inf, inf, inf, inf), mx = make_float4(-inf, -inf, -inf, -inf);
    for (int i = threadIdx.x; i < np; i += 256) {
        mn = min4(mn, partials[2 * i]);
        mx = max4(mx, partials[2 * i + 1]);
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smin[threadIdx.x] = min4(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = max4(smax[threadIdx.x], smax[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *outMin = smin[0];
        *outMax = smax[0];
    }
}

__global__ void k_check_finite(const float4 *__restrict__ a, int n, int *flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 v = a[i];
    bool bad = !(isfinite(v.x) && isfinite(v.y) && isfinite(v.z));
    if (bad) atomicOr(flag, 1);
}

__global__ void k_copy_offset(int *dst, const int *src, int n, int delta) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int v = src[i];
    dst[i] = v >= 0 ? v + delta : v;
}

// dst[r.x + k] = r.y + k for k < r.z, one range r per workgroup (grid-stride over the ranges).
__global__ void k_iota_ranges(const int4 *__restrict__ ranges, int numRanges, int *__restrict__ dst) {
    for (int r = blockIdx.x; r < numRanges; r += gridDim.x) {
        const int4 q = ranges[r];
        for (int k = threadIdx.x; k < q.z; k += blockDim.x) dst[q.x + k] = q.y + k;
    }
}

__global__ void k_fill_u16(uint16_t *p, uint16_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

inline unsigned blocks_for(long long n, int block) { return (unsigned)((n + block - 1) / block); }

// Last range whose start is <= i (ranges sorted by start, tiling the index space).
template <class F>
__device__ __forceinline__ int range_of(const CarryRange *__restrict__ tab, int R, int i, F start) {
    int lo = 0, hi = R;   // invariant: start(lo) <= i < start(hi)
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (start(tab[mid]) <= i) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ void k_carry_particles(const CarryRange *__restrict__ tab, int R, int carriedN, StateBufs src,
                                  StateBufs dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= carriedN) return;
    const CarryRange r = tab[range_of(tab, R, i, [](const CarryRange &c) { return c.dstP; })];
    const int j = r.srcP + (i - r.dstP);
    dst.pos[i] = src.pos[j];
    dst.vel[i] = src.vel[j];
    dst.bodyFrame[i] = src.bodyFrame[j];
    const int v = src.rb[j];
    dst.rb[i] = v >= 0 ? v - r.srcB + r.dstB : v;
    dst.scene[i] = (uint16_t)r.scene;
}

__global__ void k_carry_bodies(const CarryRange *__restrict__ tab, int R, int carriedB, StateBufs src,
                               StateBufs dst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= carriedB) return;
    const CarryRange r = tab[range_of(tab, R, k, [](const CarryRange &c) { return c.dstB; })];
    const int j = r.srcB + (k - r.dstB);
    dst.rbPos[k] = src.rbPos[j];
    dst.rbVel[k] = src.rbVel[j];
    dst.rbAng[k] = src.rbAng[j];
    dst.qinc[k] = src.qinc[j];
    dst.qcum[k] = src.qcum[j];
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        dst.Ib[9 * (size_t)k + c] = src.Ib[9 * (size_t)j + c];
        dst.Iw[9 * (size_t)k + c] = src.Iw[9 * (size_t)j + c];
    }
#pragma unroll
    for (int c = 0; c < 16; ++c) dst.model[16 * (size_t)k + c] = src.model[16 * (size_t)j + c];
    dst.mass[k] = src.mass[j];
    dst.pstart[k] = src.pstart[j] - r.srcP + r.dstP;
    dst.pcount[k] = src.pcount[j];
}

}  // namespace

void launch_copy_offset(int *dst, const int *src, int n, int delta, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_copy_offset, dim3(blocks_for(n, 256)), dim3(256), 0, st, dst, src, n, delta);
}

void launch_iota_ranges(const int4 *ranges, int numRanges, int *dst, hipStream_t st) {
    if (numRanges <= 0) return;
    hipLaunchKernelGGL(k_iota_ranges, dim3((unsigned)(numRanges < 4096 ? numRanges : 4096)), dim3(256), 0, st, ranges, numRanges,
                       dst);
}

void launch_fill_u16(uint16_t *p, uint16_t value, size_t count, hipStream_t st) {
    if (count == 0) return;
    unsigned blocks = (unsigned)((count + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u16, dim3(blocks), dim3(256), 0, st, p, value, count);
}

void launch_body_integrate(const BodyBufs &b, int numBodies, const DevPhys &ph, const ZeroJobs *zero,
                           hipStream_t st, const BodyBufs *out) {
    const int block = 128;
    int bodyBlocks = (int)blocks_for(numBodies, block);
    ZeroJobs z{};
    size_t zmax = 0;
    if (zero) {
        z = *zero;
        for (int k = 0; k < 4; ++k) zmax = z.count[k] > zmax ? z.count[k] : zmax;
    }
    int zeroBlocks = zmax ? (int)blocks_for((long long)zmax, block) : 0;
    if (zeroBlocks > 512) zeroBlocks = 512;
    if (bodyBlocks + zeroBlocks == 0) return;
    hipLaunchKernelGGL(k_body_integrate, dim3(bodyBlocks + zeroBlocks), dim3(block), 0, st, b, out ? *out : b,
                       numBodies, ph, z, bodyBlocks);
}

void launch_pose_key(const PartBufs &p, const BodyBufs &b, int numParticles, const DevPhys &ph, const SceneView &sv,
                     int mode, const KeyOut &ko, bool doBodies, bool doGrains, hipStream_t st) {
    if (numParticles <= 0) return;
    const int block = 256;
    // With digit histograms the blocks are capped: each flushes its LDS histogram once with one global
    // atomic per non-empty digit, and those atomics cost more than the occupancy that more blocks
    // would add.  Measured: 2^20 keys 512 blocks 28.8 us, 1024 29.4, 2048 32.2, 384 31.2; 4.2 M keys
    // 1024 blocks 98 us, 512 129, 2048 109; 8.9 M keys 1024 blocks 250 us, 512 350, 2048 257.
    const long long need = blocks_for(numParticles, block);
    const long long cap = need <= 4096 ? 512 : 1024;
    dim3 grid((unsigned)(ko.hist && mode != kKeyNone && need > cap ? cap : need));
    int db = doBodies ? 1 : 0, dg = doGrains ? 1 : 0;
    // one scene: its geometry passed by value (kernel arguments, scalar registers); several: per lane
    const bool single = sv.count <= 1;
    auto run = [&](auto one, auto many) {
        if (single) hipLaunchKernelGGL(one, grid, dim3(block), 0, st, p, b, numParticles, ph, sv, ko, db, dg);
        else hipLaunchKernelGGL(many, grid, dim3(block), 0, st, p, b, numParticles, ph, sv, ko, db, dg);
    };
    switch (mode) {
        case kKeyGrid:
            run(k_pose_key<kKeyGrid, kPoseGeomSingle>, k_pose_key<kKeyGrid, kPoseGeomLane>);
            break;
        case kKeyMorton30:
            run(k_pose_key<kKeyMorton30, kPoseGeomSingle>, k_pose_key<kKeyMorton30, kPoseGeomLane>);
            break;
        case kKeyMorton63:
            run(k_pose_key<kKeyMorton63, kPoseGeomSingle>, k_pose_key<kKeyMorton63, kPoseGeomLane>);
            break;
        case kKeyForest30:
            run(k_pose_key<kKeyForest30, kPoseGeomSingle>, k_pose_key<kKeyForest30, kPoseGeomLane>);
            break;
        default:
            run(k_pose_key<kKeyNone, kPoseGeomSingle>, k_pose_key<kKeyNone, kPoseGeomLane>);
            break;
    }
}

void launch_pose_key_tiles(const PartBufs &p, const BodyBufs &b, int numParticles, const DevPhys &ph,
                           const SceneView &sv, const TileKeyOut &to, bool grainVelDead, hipStream_t st) {
    if (numParticles <= 0) return;
    const int g = grainVelDead ? 1 : 0;
    const dim3 grid(to.numTiles);
    if (to.morton == 2) {   // 63-bit Morton codes: tiles ranked by the top 10 bits
        hipLaunchKernelGGL((k_pose_key_tiles<kKeyMorton63, 1024, 10>), grid, dim3(1024), 0, st, p, b, numParticles, ph,
                           sv, to, g);
        return;
    }
    if (to.morton) {   // 30-bit Morton keys, buckets of their top 10 bits
        hipLaunchKernelGGL((k_pose_key_tiles<kKeyMorton30, 1024, 10>), grid, dim3(1024), 0, st, p, b, numParticles, ph,
                           sv, to, g);
        return;
    }
    if (to.tileKeys == 1024)
        hipLaunchKernelGGL((k_pose_key_tiles<kKeyGrid, 256, 8>), grid, dim3(256), 0, st, p, b, numParticles, ph, sv, to,
                           g);
    else if (to.bucketBits <= 8)
        hipLaunchKernelGGL((k_pose_key_tiles<kKeyGrid, 1024, 8>), grid, dim3(1024), 0, st, p, b, numParticles, ph, sv,
                           to, g);
    else if (to.bucketBits == 9)
        hipLaunchKernelGGL((k_pose_key_tiles<kKeyGrid, 1024, 9>), grid, dim3(1024), 0, st, p, b, numParticles, ph, sv,
                           to, g);
    else
        hipLaunchKernelGGL((k_pose_key_tiles<kKeyGrid, 1024, 10>), grid, dim3(1024), 0, st, p, b, numParticles, ph, sv,
                           to, g);
}

void launch_reduce_apply(const float4 *pForce, const float4 *pTorque, const float4 *relPos, const BodyBufs &b,
                         int numBodies, const DevPhys &ph, hipStream_t st) {
    if (numBodies <= 0) return;
    const int block = 256;
    dim3 grid(blocks_for((long long)numBodies * 64, block));
    hipLaunchKernelGGL(k_reduce_apply, grid, dim3(block), 0, st, pForce, pTorque, relPos, b, numBodies, ph);
}

void launch_reduce_apply_ahead(const float4 *pForce, const float4 *relPos, const BodyBufs &b, const BodyBufs &ahead,
                               int numBodies, const DevPhys &ph, const ZeroJobs *zero, hipStream_t st) {
    const int block = 256;
    const int bodyBlocks = numBodies > 0 ? (numBodies + kAheadBodies - 1) / kAheadBodies : 0;
    ZeroJobs z{};
    size_t zmax = 0;
    if (zero) {
        z = *zero;
        for (int k = 0; k < 4; ++k) zmax = z.count[k] > zmax ? z.count[k] : zmax;
    }
    int zeroBlocks = zmax ? (int)blocks_for((long long)zmax, block) : 0;
    if (zeroBlocks > 512) zeroBlocks = 512;
    if (bodyBlocks + zeroBlocks == 0) return;
    hipLaunchKernelGGL(k_reduce_apply_ahead, dim3(bodyBlocks + zeroBlocks), dim3(block), 0, st, pForce, relPos, b, ahead,
                       numBodies, ph, z, bodyBlocks);
}

void launch_fill_u32(unsigned *p, unsigned value, size_t count, hipStream_t st) {
    if (count == 0) return;
    unsigned blocks = (unsigned)((count + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, st, p, value, count);
}

int aabb_max_partials() { return kAabbBlocks; }

void launch_aabb(const float4 *pos, int n, float4 *outMin, float4 *outMax, float4 *partials, int maxPartials,
                 hipStream_t st) {
    int blocks = (int)blocks_for(n, 256 * 8);
    if (blocks < 1) blocks = 1;
    if (blocks > maxPartials) blocks = maxPartials;
    hipLaunchKernelGGL(k_aabb_partial, dim3(blocks), dim3(256), 0, st, pos, n, partials);
    hipLaunchKernelGGL(k_aabb_final, dim3(1), dim3(256), 0, st, partials, blocks, outMin, outMax);
}

void launch_check_finite(const float4 *a, int n, int *flag, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_check_finite, dim3(blocks_for(n, 256)), dim3(256), 0, st, a, n, flag);
}

}  // namespace arsim

namespace arsim {
void launch_carry(const CarryRange *tab, int numRanges, int carriedN, int carriedB, const StateBufs &src,
                  const StateBufs &dst, hipStream_t st) {
    if (numRanges <= 0) return;
    if (carriedN > 0)
        hipLaunchKernelGGL(k_carry_particles, dim3(blocks_for(carriedN, 256)), dim3(256), 0, st, tab, numRanges,
                           carriedN, src, dst);
    if (carriedB > 0)
        hipLaunchKernelGGL(k_carry_bodies, dim3(blocks_for(carriedB, 256)), dim3(256), 0, st, tab, numRanges,
                           carriedB, src, dst);
}
}  // namespace arsim
