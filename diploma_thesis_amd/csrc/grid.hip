// grid.hip -- uniform-grid broadphase: cell bounds + reorder, and the 27-cell DEM kernel.
#include "csr_fill.hpp"
#include "kernels.hpp"
#include "reduce_apply.hpp"

namespace arsim {

namespace {

// reorderDataAndFindCellStartD (Kernels/particleKernels.cu:193-258).  The previous key is read
// straight from global memory (it is in the same or the previous cache line).  The sorted copies
// carry the rbIndex in sortedPos.w and, when scene ids are given, the scene id in sortedVel.w.
__global__ void k_reorder(const unsigned *__restrict__ keys, const unsigned *__restrict__ idx,
                          const float4 *__restrict__ pos, const float4 *__restrict__ vel, const int *__restrict__ rb,
                          const uint16_t *__restrict__ scene, float4 *__restrict__ sPos, float4 *__restrict__ sVel,
                          unsigned *__restrict__ cs, unsigned *__restrict__ ce, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned h = keys[i];
    if (i == 0 || h != keys[i - 1]) {
        cs[h] = (unsigned)i;
        if (i > 0) ce[keys[i - 1]] = (unsigned)i;
    }
    if (i == n - 1) ce[h] = (unsigned)n;
    unsigned o = idx[i];
    float4 p = pos[o];
    p.w = __int_as_float(rb[o]);
    sPos[i] = p;
    float4 v = vel[o];
    if (scene) v.w = __int_as_float((int)scene[o]);
    sVel[i] = v;
}

// FindAndHandleRigidBodyCollisionsUniformGridKernel (RigidBodies/UniformGridDEM.cu:133-204) and
// collideD (particleKernels.cu:302-406): 27 cells in z, y, x order, per-cell partial sums, the
// contact predicate j != i && (rbA != rbB || rbA == -1), then the wall term.  EMIT writes the
// contact list instead of forces.
template <bool EMIT>
__global__ void k_dem_grid(const float4 *__restrict__ sPos, const float4 *__restrict__ sVel,
                           const unsigned *__restrict__ sIdx, const unsigned *__restrict__ cs,
                           const unsigned *__restrict__ ce, int n, DevPhys ph, SceneView sv, int sceneInVelW,
                           DemOut o) {
    const int lb = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int i = lb * (int)blockDim.x + (int)threadIdx.x;
    if (i >= n) return;
    const float4 P = sPos[i];
    const float4 Vq = sVel[i];
    const int ra = __float_as_int(P.w);
    const int s = sceneInVelW ? __float_as_int(Vq.w) : 0;
    SceneGeom g;
    load_geom(sv, s, g);
    const f3 pa = xyz(P), va = xyz(Vq);
    int gx, gy, gz;
    grid_pos(pa, g, gx, gy, gz);
    const unsigned orig = sIdx[i];
    f3 F = mk3(0.f, 0.f, 0.f);
    int cnt = 0;
    for (int z = -1; z <= 1; ++z)
        for (int y = -1; y <= 1; ++y)
            for (int x = -1; x <= 1; ++x) {
                unsigned h = g.cellBase + grid_hash(gx + x, gy + y, gz + z, g);
                unsigned st = cs[h];
                if (st == kEmpty) continue;
                unsigned en = ce[h];
                f3 csum = mk3(0.f, 0.f, 0.f);
                for (unsigned j = st; j < en; ++j) {
                    float4 Q = sPos[j];
                    if (!candidate(Q, pa, ra, (int)j == i, ph)) continue;
                    if (EMIT) {
                        if (in_contact(pa, xyz(Q), ph)) {
                            if (cnt < o.maxPer) o.pairs[(size_t)orig * o.maxPer + cnt] = (int)sIdx[j];
                            ++cnt;
                        }
                    } else {
                        f3 fj;
                        if (pair_force(pa, xyz(Q), va, xyz(sVel[j]), ph, fj)) csum = add3(csum, fj);
                    }
                }
                F = add3(F, csum);
            }
    if (EMIT) {
        o.counts[orig] = cnt;
        return;
    }
    if (!((ph.flags & kGrainClamp) && ra < 0)) F = add3(F, wall_force(pa, va, ph, g));
    if (ra >= 0) {
        o.pForce[orig] = f4(F, 0.0f);
        if (o.pTorque) o.pTorque[orig] = f4(cross3(xyz(o.relPos[orig]), F), 0.0f);
    } else {
        const float fs = (ph.flags & kRefUnits) ? 1.0f : ph.dt;
        o.vel[orig] = make_float4(va.x + F.x * fs, va.y + F.y * fs, va.z + F.z * fs, 0.0f);
    }
}

// Engine form of the reorder: the same gather of the sorted copies, but instead of the two
// cell-sized tables it writes a dense CSR offset array over segments of 2^segBits cells:
// segStart[s] = first sorted index whose key >= s << segBits (segStart[numSegs] = n).  Every entry
// is rewritten each step, so nothing has to be cleared.  With segBits = 0 it is cellStart with
// cellEnd[c] = segStart[c + 1], one table instead of two.  numSegs < 2^32.
// The long runs handed over by wave_fill, one chunk of at most kLongChunk entries per workgroup
// (workgroup g takes chunks g, g + G, ...).  With nothing listed the kernel only reads the count.
constexpr unsigned kFillBlocks = 512u;
__global__ __launch_bounds__(256) void k_fill_long(unsigned *__restrict__ table, unsigned long long size, LongFill lf) {
    const unsigned cnt = min(*lf.count, lf.cap);
    for (unsigned r = blockIdx.x; r < cnt; r += gridDim.x) {
        const uint4 e = lf.list[r];
        // (the capacity bound means every listed entry was written; the clamp keeps stores in the table)
        const unsigned a = e.x, b = (unsigned long long)e.y < size ? e.y : (unsigned)size;
        if (a >= b) continue;
        // scalar head and tail, 16-byte stores for the aligned middle
        const unsigned a4 = (a + 3u) & ~3u, b4 = b & ~3u;
        if (a4 >= b4) {
            for (unsigned k = a + threadIdx.x; k < b; k += 256u) table[k] = e.z;
            continue;
        }
        if (threadIdx.x < a4 - a) table[a + threadIdx.x] = e.z;
        if (threadIdx.x < b - b4) table[b4 + threadIdx.x] = e.z;
        const uint4 v4 = make_uint4(e.z, e.z, e.z, e.z);
        for (unsigned k = a4 + 4u * threadIdx.x; k < b4; k += 1024u) *reinterpret_cast<uint4 *>(table + k) = v4;
    }
}

__global__ void k_reorder_seg(const unsigned *__restrict__ keys, const unsigned *__restrict__ idx,
                              const float4 *__restrict__ pos, const float4 *__restrict__ vel,
                              const int *__restrict__ rb, const uint16_t *__restrict__ scene,
                              const float4 *__restrict__ rec, float4 *__restrict__ sPos, float4 *__restrict__ sVel,
                              unsigned *__restrict__ segStart, int segBits, long long numSegs, int n,
                              LongFill lf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const unsigned k = valid ? keys[i] : 0u;
    const unsigned s = k >> segBits;
    const unsigned sp1 = (valid && i > 0) ? (keys[i - 1] >> segBits) + 1u : 0u;   // first segment not yet started
    wave_fill(segStart, valid && s + 1u > sp1, sp1, s + 1u, (unsigned)i, lf);
    wave_fill(segStart, i == n - 1, s + 1u, (unsigned)numSegs + 1u, (unsigned)n, lf);
    if (!valid) return;
    const unsigned o = idx[i];
    if (rec) {
        sPos[i] = rec[2 * (size_t)o];
        sVel[i] = rec[2 * (size_t)o + 1];
        return;
    }
    float4 p = pos[o];
    p.w = __int_as_float(rb[o]);
    sPos[i] = p;
    float4 v = vel[o];
    if (scene) v.w = __int_as_float((int)scene[o]);
    sVel[i] = v;
}

// Contacts of one sorted particle within the sorted-key run [j, ...) of cells h0..h1 (one row of
// up to three consecutive cells), cell by cell in ascending order, each cell's contacts summed
// into a partial that is added to F when the cell ends -- the same grouping and order as the
// 27-cell loop over cellStart / cellEnd.
// Forward scan of the sorted keys from j (a segment start at or before h0): skips keys below h0,
// stops past h1.  (Reading several candidates per round trip was measured slower: the extra
// registers cost more occupancy than the latency it hid.)
template <bool EMIT>
__device__ __forceinline__ void scan_cells(const unsigned *__restrict__ keys, const float4 *__restrict__ sPos,
                                           const float4 *__restrict__ sVel, const unsigned *__restrict__ sIdx,
                                           unsigned j, int n, unsigned h0, unsigned h1, int i, int ra, f3 pa, f3 va,
                                           const DevPhys &ph, f3 &F, int &cnt, unsigned orig, const DemOut &o) {
    unsigned cur = 0xffffffffu;
    f3 csum = mk3(0.f, 0.f, 0.f);
    for (; (int)j < n; ++j) {
        const unsigned kj = keys[j];
        if (kj > h1) break;
        if (kj < h0) continue;
        if (kj != cur) {
            if (cur != 0xffffffffu) F = add3(F, csum);
            csum = mk3(0.f, 0.f, 0.f);
            cur = kj;
        }
        const float4 Q = sPos[j];
        if (!candidate(Q, pa, ra, (int)j == i, ph)) continue;
        if (EMIT) {
            if (in_contact(pa, xyz(Q), ph)) {
                if (cnt < o.maxPer) o.pairs[(size_t)orig * o.maxPer + cnt] = (int)sIdx[j];
                ++cnt;
            }
        } else {
            f3 fj;
            if (pair_force(pa, xyz(Q), va, xyz(sVel[j]), ph, fj)) csum = add3(csum, fj);
        }
    }
    if (cur != 0xffffffffu) F = add3(F, csum);
}

// Contacts of one particle with the sorted run [j, e) that holds up to three consecutive cells, the
// second starting at b1 and the third at b2 (exact cell offsets: no key is read).  A cell's partial
// is added to F when a non-empty cell ends, as in the 27-cell loop.
// Candidates are loaded two at a time (the second index clamped into the run), so that a wave has
// two loads in flight per lane: the loop is bound by load latency, not by the force arithmetic.
// Three or more per round trip cost more occupancy (registers) than they hide.
// The particle itself is excluded by its sorted index i (SELF_BY_ORIG false), or, when the caller
// walks particles in their original order, by its original index: only a candidate at distance 0
// from a grain can be the grain itself, so only then is the candidate's original index read (a body
// particle's own copy is excluded by the same-body rule anyway).
template <bool EMIT, bool SELF_BY_ORIG = false>
__device__ __forceinline__ void scan_run(const float4 *__restrict__ sPos, const float4 *__restrict__ sVel,
                                         const unsigned *__restrict__ sIdx, unsigned j, unsigned b1, unsigned b2,
                                         unsigned e, int i, int ra, f3 pa, f3 va, const DevPhys &ph, f3 &F, int &cnt,
                                         unsigned orig, const DemOut &o) {
    const unsigned j0 = j;
    f3 csum = mk3(0.f, 0.f, 0.f);
    auto visit = [&](unsigned jj, const float4 &Q) {
        if (jj != j0 && (jj == b1 || jj == b2)) {
            F = add3(F, csum);
            csum = mk3(0.f, 0.f, 0.f);
        }
        if (SELF_BY_ORIG) {
            const f3 rel = sub3(mk3(Q.x, Q.y, Q.z), pa);
            const float d2 = dot3(rel, rel);
            bool ok = !(d2 > ph.cd2Reject) & pair_allowed(ra, __float_as_int(Q.w));
            if (ok && ra < 0 && d2 == 0.0f) ok = sIdx[jj] != orig;
            if (!ok) return;
        } else {
            if (!candidate(Q, pa, ra, (int)jj == i, ph)) return;
        }
        if (EMIT) {
            if (in_contact(pa, xyz(Q), ph)) {
                if (cnt < o.maxPer) o.pairs[(size_t)orig * o.maxPer + cnt] = (int)sIdx[jj];
                ++cnt;
            }
        } else {
            f3 fj;
            if (pair_force(pa, xyz(Q), va, xyz(sVel[jj]), ph, fj)) csum = add3(csum, fj);
        }
    };
    for (; j < e; j += 2u) {
        const float4 Q0 = sPos[j];
        const float4 Q1 = sPos[j + 1u < e ? j + 1u : j];
        visit(j, Q0);
        if (j + 1u < e) visit(j + 1u, Q1);
    }
    if (e > j0) F = add3(F, csum);
}

// Σ of the pair forces on one particle over its 27 neighbour cells with the exact cell table
// (segStart = cellStart, numCells + 1 entries): for each of the 9 (dz, dy) rows in z, y order the
// three x-cells are one contiguous run [segStart[h0], segStart[h0 + 3]) unless x wraps, visited in
// the order x-1, x, x+1.  The walls are not included.
template <bool EMIT, bool SELF_BY_ORIG>
__device__ __forceinline__ f3 dem_rows_exact(const float4 *__restrict__ sPos, const float4 *__restrict__ sVel,
                                             const unsigned *__restrict__ sIdx, const unsigned *__restrict__ segStart,
                                             const SceneGeom &g, int i, int ra, f3 pa, f3 va, const DevPhys &ph,
                                             int &cnt, unsigned orig, const DemOut &o) {
    int gx, gy, gz;
    grid_pos(pa, g, gx, gy, gz);
    const unsigned x0 = (unsigned)(gx - 1) & g.gridMask[0], x1 = (unsigned)gx & g.gridMask[0],
                   x2 = (unsigned)(gx + 1) & g.gridMask[0];
    const bool contiguous = (x0 + 1u == x1) && (x1 + 1u == x2);
    f3 F = mk3(0.f, 0.f, 0.f);
    auto row_of = [&](int r) {
        const unsigned uz = (unsigned)(gz + (r / 3) - 1) & g.gridMask[2];
        const unsigned uy = (unsigned)(gy + (r % 3) - 1) & g.gridMask[1];
        return g.cellBase + uz * g.gxy + uy * g.gx;
    };
    for (int r = 0; r < 9; ++r) {
        const unsigned row = row_of(r);
        if (contiguous) {
            // the row's four offsets in one 16-byte load (4-byte aligned): the kernel issues fewer
            // vector memory instructions, which bound it more than bytes do
            uint4 bnd;
            __builtin_memcpy(&bnd, segStart + row + x0, sizeof(bnd));
            scan_run<EMIT, SELF_BY_ORIG>(sPos, sVel, sIdx, bnd.x, bnd.y, bnd.z, bnd.w, i, ra, pa, va, ph, F, cnt,
                                         orig, o);
        } else {
            const unsigned xs[3] = {x0, x1, x2};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const unsigned h = row + xs[k];
                const unsigned e = segStart[h + 1];
                scan_run<EMIT, SELF_BY_ORIG>(sPos, sVel, sIdx, segStart[h], e, e, e, i, ra, pa, va, ph, F, cnt,
                                             orig, o);
            }
        }
    }
    return F;
}

// Engine DEM over the segment index: for each of the 9 (dz, dy) rows the three x-cells are
// consecutive keys unless x wraps, so one segment lookup and one forward scan of the sorted
// keys visit them in the order x-1, x, x+1.  EXACT (segments of one cell, segStart = cellStart
// with numCells + 1 entries): a row is the run [segStart[h0], segStart[h0 + 3]) with its cell
// boundaries read from the table, and the keys are not read at all.
// Where the kernel takes a scene's geometry from: the one scene's, passed by value (kGeomSingle);
// for several scenes (scene id in sortedVel.w), the table entry of the block's scene when all of the
// block's particles belong to it (kGeomBlock: the block's first and last sorted particle tell, and
// the geometry is then read through scalar loads), or each lane's own entry (kGeomLane).  A batch runs
// the kGeomBlock launch and then the kGeomLane launch, each doing the blocks the other skips.
enum { kGeomSingle = 0, kGeomBlock = 1, kGeomLane = 2 };

template <bool EMIT, bool EXACT, int GEOM>
__global__ void k_dem_grid_seg(const float4 *__restrict__ sPos, const float4 *__restrict__ sVel,
                               const unsigned *__restrict__ keys, const unsigned *__restrict__ sIdx,
                               const unsigned *__restrict__ segStart, int segBits, int n, DevPhys ph, SceneView sv,
                               DemOut o) {
    int lb;
    if (GEOM == kGeomLane && sv.sceneStart) {
        // one workgroup per boundary between scenes k - 1 and k: the block that holds it, if the boundary
        // lies strictly inside a block and no earlier boundary lies strictly inside the same one
        const int k = (int)blockIdx.x + 1, bs = (int)blockDim.x;
        const int b = sv.sceneStart[k], pb = sv.sceneStart[k - 1];
        lb = b / bs;
        if (b % bs == 0 || b >= n) return;
        if (k > 1 && pb / bs == lb && pb % bs != 0) return;
    } else {
        lb = xcd_block((int)blockIdx.x, (int)gridDim.x);
    }
    const int first = lb * (int)blockDim.x;
    SceneGeom g;
    if (GEOM == kGeomSingle) {
        g = sv.single;
    } else {
        const int last = min(n, first + (int)blockDim.x) - 1;
        const int sf = __float_as_int(sVel[first].w), sl = __float_as_int(sVel[last].w);
        if ((sf == sl) != (GEOM == kGeomBlock)) return;
        if (GEOM == kGeomBlock) g = sv.table[sf];
    }
    const int i = first + (int)threadIdx.x;
    if (i >= n) return;
    const float4 P = sPos[i];
    const float4 Vq = sVel[i];
    const int ra = __float_as_int(P.w);
    if (GEOM == kGeomLane) g = sv.table[__float_as_int(Vq.w)];
    const f3 pa = xyz(P), va = xyz(Vq);
    int gx, gy, gz;
    grid_pos(pa, g, gx, gy, gz);
    const unsigned orig = sIdx[i];
    const unsigned x0 = (unsigned)(gx - 1) & g.gridMask[0], x1 = (unsigned)gx & g.gridMask[0],
                   x2 = (unsigned)(gx + 1) & g.gridMask[0];
    const bool contiguous = (x0 + 1u == x1) && (x1 + 1u == x2);
    f3 F = mk3(0.f, 0.f, 0.f);
    int cnt = 0;
    // the nine row starts are independent: issue all their loads before scanning any row
    auto row_of = [&](int r) {
        const unsigned uz = (unsigned)(gz + (r / 3) - 1) & g.gridMask[2];
        const unsigned uy = (unsigned)(gy + (r % 3) - 1) & g.gridMask[1];
        return g.cellBase + uz * g.gxy + uy * g.gx;
    };
    if (EXACT) {
        for (int r = 0; r < 9; ++r) {
            const unsigned row = row_of(r);
            if (contiguous) {
                const unsigned h0 = row + x0;
                // the row's four offsets in one 16-byte load (4-byte aligned): the kernel issues fewer
                // vector memory instructions, which bound it more than bytes do
                uint4 bnd;
                __builtin_memcpy(&bnd, segStart + h0, sizeof(bnd));
                scan_run<EMIT>(sPos, sVel, sIdx, bnd.x, bnd.y, bnd.z, bnd.w, i, ra, pa, va, ph, F, cnt, orig, o);
            } else {
                const unsigned xs[3] = {x0, x1, x2};
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const unsigned h = row + xs[k];
                    const unsigned e = segStart[h + 1];
                    scan_run<EMIT>(sPos, sVel, sIdx, segStart[h], e, e, e, i, ra, pa, va, ph, F, cnt, orig, o);
                }
            }
        }
    } else {
        unsigned rowJ[9];
        if (contiguous) {
#pragma unroll
            for (int r = 0; r < 9; ++r) rowJ[r] = segStart[(row_of(r) + x0) >> segBits];
        }
        for (int r = 0; r < 9; ++r) {
            const unsigned row = row_of(r);
            if (contiguous) {
                const unsigned h0 = row + x0;
                scan_cells<EMIT>(keys, sPos, sVel, sIdx, rowJ[r], n, h0, h0 + 2u, i, ra, pa, va, ph, F, cnt, orig, o);
            } else {
                const unsigned xs[3] = {x0, x1, x2};
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const unsigned h = row + xs[k];
                    scan_cells<EMIT>(keys, sPos, sVel, sIdx, segStart[h >> segBits], n, h, h, i, ra, pa, va, ph, F,
                                     cnt, orig, o);
                }
            }
        }
    }
    if (EMIT) {
        o.counts[orig] = cnt;
        return;
    }
    if (!((ph.flags & kGrainClamp) && ra < 0)) F = add3(F, wall_force(pa, va, ph, g));
    if (ra >= 0) {
        o.pForce[orig] = f4(F, 0.0f);
        if (o.pTorque) o.pTorque[orig] = f4(cross3(xyz(o.relPos[orig]), F), 0.0f);
    } else {
        const float fs = (ph.flags & kRefUnits) ? 1.0f : ph.dt;
        o.vel[orig] = make_float4(va.x + F.x * fs, va.y + F.y * fs, va.z + F.z * fs, 0.0f);
    }
}

// Grid DEM fused with the per-body reduction and apply (engine, exact cell table, no static set):
// one launch in place of k_dem_grid_seg + k_reduce_apply, with no per-particle force round trip
// through memory.  Waves [0, B) take one body each: lane l walks the body's particles l, l+64, ... in
// index order, computes each particle's force exactly as k_dem_grid_seg does (same cells, same order,
// same per-cell partials, walls last) and accumulates force and torque relPos × F in the order of
// k_reduce_apply; the butterfly and the apply follow.  The remaining waves take 64 grains each (grain
// indices in original order) and update their velocities.  A wave's lanes are then particles of one
// kind, close in space (one body, or grains in index order), so their candidate loops have similar
// lengths, where a wave of consecutive sorted particles mixes body particles and grains.
__global__ __launch_bounds__(256) void k_dem_fused(const float4 *__restrict__ sPos, const float4 *__restrict__ sVel,
                                                   const unsigned *__restrict__ sIdx,
                                                   const unsigned *__restrict__ segStart, FusedDemIn in, BodyBufs b,
                                                   int B, DevPhys ph, SceneView sv, int sceneInVelW) {
    const int lb = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int wave = lb * 4 + (int)(threadIdx.x >> 6);
    const int lane = (int)(threadIdx.x & 63);
    const DemOut none{};
    int cnt = 0;
    if (wave < B) {
        const int body = wave;
        const int st = b.pstart[body], count = b.pcount[body];
        f3 aF = mk3(0.f, 0.f, 0.f), aT = mk3(0.f, 0.f, 0.f);
        for (int j = reduce_first(body, ph) + lane; j < count; j += 64) {
            const size_t o = (size_t)st + j;
            const float4 P = in.rec[2 * o], V = in.rec[2 * o + 1];
            const int ra = __float_as_int(P.w);
            SceneGeom g;
            load_geom(sv, sceneInVelW ? __float_as_int(V.w) : 0, g);
            const f3 pa = xyz(P), va = xyz(V);
            f3 F = dem_rows_exact<false, true>(sPos, sVel, sIdx, segStart, g, -1, ra, pa, va, ph, cnt, (unsigned)o,
                                               none);
            F = add3(F, wall_force(pa, va, ph, g));
            aF = add3(aF, F);
            aT = add3(aT, cross3(xyz(in.relPos[o]), F));
        }
        wave_reduce_apply(aF, aT, lane, body, b, ph);
        return;
    }
    const long long k = (long long)(wave - B) * 64 + lane;
    if (k >= in.numGrains) return;
    const unsigned o = (unsigned)in.grainIdx[k];
    const float4 P = in.rec[2 * (size_t)o], V = in.rec[2 * (size_t)o + 1];
    const int ra = __float_as_int(P.w);
    SceneGeom g;
    load_geom(sv, sceneInVelW ? __float_as_int(V.w) : 0, g);
    const f3 pa = xyz(P), va = xyz(V);
    f3 F = dem_rows_exact<false, true>(sPos, sVel, sIdx, segStart, g, -1, ra, pa, va, ph, cnt, o, none);
    if (!(ph.flags & kGrainClamp)) F = add3(F, wall_force(pa, va, ph, g));
    const float fs = (ph.flags & kRefUnits) ? 1.0f : ph.dt;
    in.vel[o] = make_float4(va.x + F.x * fs, va.y + F.y * fs, va.z + F.z * fs, 0.0f);
}

}  // namespace

void launch_reorder_seg(const unsigned *sortedKeys, const unsigned *sortedIdx, const float4 *pos, const float4 *vel,
                        const int *rb, const uint16_t *scene, const float4 *rec, float4 *sortedPos,
                        float4 *sortedVel, unsigned *segStart, int segBits, long long numSegs, int n,
                        const LongFill &lf, hipStream_t st) {
    if (n <= 0) return;
    const int block = 256;
    hipLaunchKernelGGL(k_reorder_seg, dim3((n + block - 1) / block), dim3(block), 0, st, sortedKeys, sortedIdx, pos,
                       vel, rb, scene, rec, sortedPos, sortedVel, segStart, segBits, numSegs, n, lf);
    if (lf.list) launch_fill_long(segStart, (unsigned long long)numSegs + 1ull, lf, st);
}

// Runs are disjoint and longer than kLongRange, so there are fewer than size / kLongRange of them, and
// each has at most len / kLongChunk + 1 chunks.
unsigned fill_long_capacity(unsigned long long tableSize) {
    return (unsigned)(tableSize / kLongRange + tableSize / kLongChunk + 4ull);
}

void launch_fill_long(unsigned *table, unsigned long long tableSize, const LongFill &lf, hipStream_t st) {
    unsigned long long blocks = (tableSize + kLongChunk - 1) / kLongChunk;
    if (blocks > kFillBlocks) blocks = kFillBlocks;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_fill_long, dim3((unsigned)blocks), dim3(256), 0, st, table, tableSize, lf);
}

void launch_dem_grid_seg(const float4 *sortedPos, const float4 *sortedVel, const unsigned *sortedKeys,
                         const unsigned *sortedIdx, const unsigned *segStart, int segBits, int n, const DevPhys &ph,
                         const SceneView &sv, bool sceneInVelW, const DemOut &o, int block, hipStream_t st) {
    if (n <= 0) return;
    if (block <= 0) block = 128;
    dim3 grid((n + block - 1) / block);
    auto run = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, grid, dim3(block), 0, st, sortedPos, sortedVel, sortedKeys, sortedIdx, segStart,
                           segBits, n, ph, sv, o);
    };
    auto geoms = [&](auto single, auto perBlock, auto perLane) {
        if (!sceneInVelW) {
            run(single);
        } else {
            run(perBlock);
            // the blocks that straddle two scenes: with the scenes' sorted ranges known, one workgroup
            // per boundary instead of a full grid whose other blocks return at once
            if (!sv.sceneStart) run(perLane);
            else if (sv.count > 1)
                hipLaunchKernelGGL(perLane, dim3((unsigned)(sv.count - 1)), dim3(block), 0, st, sortedPos, sortedVel,
                                   sortedKeys, sortedIdx, segStart, segBits, n, ph, sv, o);
        }
    };
    if (o.pairs) {
        if (segBits == 0)
            geoms(k_dem_grid_seg<true, true, kGeomSingle>, k_dem_grid_seg<true, true, kGeomBlock>,
                  k_dem_grid_seg<true, true, kGeomLane>);
        else
            geoms(k_dem_grid_seg<true, false, kGeomSingle>, k_dem_grid_seg<true, false, kGeomBlock>,
                  k_dem_grid_seg<true, false, kGeomLane>);
    } else {
        if (segBits == 0)
            geoms(k_dem_grid_seg<false, true, kGeomSingle>, k_dem_grid_seg<false, true, kGeomBlock>,
                  k_dem_grid_seg<false, true, kGeomLane>);
        else
            geoms(k_dem_grid_seg<false, false, kGeomSingle>, k_dem_grid_seg<false, false, kGeomBlock>,
                  k_dem_grid_seg<false, false, kGeomLane>);
    }
}

void launch_reorder(const unsigned *sortedKeys, const unsigned *sortedIdx, const float4 *pos, const float4 *vel,
                    const int *rb, const uint16_t *scene, float4 *sortedPos, float4 *sortedVel, unsigned *cellStart,
                    unsigned *cellEnd, int n, hipStream_t st) {
    if (n <= 0) return;
    const int block = 256;
    hipLaunchKernelGGL(k_reorder, dim3((n + block - 1) / block), dim3(block), 0, st, sortedKeys, sortedIdx, pos, vel,
                       rb, scene, sortedPos, sortedVel, cellStart, cellEnd, n);
}

void launch_dem_grid(const float4 *sortedPos, const float4 *sortedVel, const unsigned *sortedIdx,
                     const unsigned *cellStart, const unsigned *cellEnd, int n, const DevPhys &ph,
                     const SceneView &sv, bool sceneInVelW, const DemOut &o, int block, hipStream_t st) {
    if (n <= 0) return;
    if (block <= 0) block = 128;
    dim3 grid((n + block - 1) / block);
    if (o.pairs)
        hipLaunchKernelGGL(k_dem_grid<true>, grid, dim3(block), 0, st, sortedPos, sortedVel, sortedIdx, cellStart,
                           cellEnd, n, ph, sv, sceneInVelW ? 1 : 0, o);
    else
        hipLaunchKernelGGL(k_dem_grid<false>, grid, dim3(block), 0, st, sortedPos, sortedVel, sortedIdx, cellStart,
                           cellEnd, n, ph, sv, sceneInVelW ? 1 : 0, o);
}

void launch_dem_fused(const float4 *sortedPos, const float4 *sortedVel, const unsigned *sortedIdx,
                      const unsigned *segStart, const FusedDemIn &in, const BodyBufs &b, int numBodies,
                      const DevPhys &ph, const SceneView &sv, bool sceneInVelW, hipStream_t st) {
    const long long waves = (long long)numBodies + (in.numGrains + 63) / 64;
    if (waves <= 0) return;
    const long long blocks = (waves + 3) / 4;
    hipLaunchKernelGGL(k_dem_fused, dim3((unsigned)blocks), dim3(256), 0, st, sortedPos, sortedVel, sortedIdx, segStart,
                       in, b, numBodies, ph, sv, sceneInVelW ? 1 : 0);
}

}  // namespace arsim
