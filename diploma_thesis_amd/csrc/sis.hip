// sis.hip -- the sequential-impulse contact model on the device: the deepest contact of every
// particle, the contact list in the reference's order, and the Gauss-Seidel impulse sweeps.
//
// The sweeps keep the reference's sequential semantics exactly: contact c depends on the last
// contact before it that touches one of its bodies (its predecessors); level(c) = 1 + the larger
// of its predecessors' levels.  Contacts of one level touch disjoint bodies and everything they
// read was written by lower levels, so running a sweep level by level, with a barrier between
// levels, reads and writes every body value exactly as the serial loop does.
#include <climits>

#include "kernels.hpp"

namespace arsim {

namespace {

constexpr int kSisBlock = 1024;   // one workgroup per scene for levels / buckets / sweeps
constexpr int kSisWaves = kSisBlock / 64;
constexpr int kSisLdsBodies = 2048;   // 136 KiB of (vel, ang) and inverse inertia per scene in LDS

__device__ __forceinline__ float dot4(float4 a, float4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }

// FindRigidBodyCollisionsUniformGridKernel (Kernels/RB_UniformGridKernels.cu:730-845): per sorted
// particle, the 27 cells in z, y, x order, each cell's particles in sorted order; a candidate of
// another body whose penetration 2r - |p - q| exceeds the best so far becomes the contact.
//  REF: reference layout -- cellStart / cellEnd (0xffffffff = empty), bodies read through
//       rbIndices[gridParticleIndex[j]], the best starts at contactDistance[orig] and the outputs are
//       written only when a contact is found (color set to (1, 0, 0, 0) then, if given);
//  engine: the exact per-cell offset table (cell c = [cs[c], cs[c + 1])), the body index in the bits
//       of sortedPos.w, the scene id in sortedVel.w when there are several scenes; the best starts at
//       0 and all three outputs are written (-1 partners when there is no contact).
template <bool REF>
__global__ void k_sis_find(const float4 *__restrict__ sPos, const float4 *__restrict__ sVel,
                           const unsigned *__restrict__ sIdx, const int *__restrict__ rbIdx,
                           const unsigned *__restrict__ cs, const unsigned *__restrict__ ce, SceneView sv,
                           int sceneInVelW, int n, float thr, float *__restrict__ dist, int *__restrict__ pBody,
                           int *__restrict__ pPart, float4 *__restrict__ color) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 P = sPos[i];
    const unsigned orig = sIdx[i];
    const int ra = REF ? rbIdx[orig] : __float_as_int(P.w);
    const int s = (!REF && sceneInVelW) ? __float_as_int(sVel[i].w) : 0;
    SceneGeom g;
    load_geom(sv, s, g);
    const f3 pa = xyz(P);
    int gx, gy, gz;
    grid_pos(pa, g, gx, gy, gz);
    float best = REF ? dist[orig] : 0.0f;
    int bj = -1;
    for (int z = -1; z <= 1; ++z)
        for (int y = -1; y <= 1; ++y)
            for (int x = -1; x <= 1; ++x) {
                const unsigned h = g.cellBase + grid_hash(gx + x, gy + y, gz + z, g);
                unsigned j = cs[h], e;
                if (REF) {
                    if (j == kEmpty) continue;
                    e = ce[h];
                } else {
                    e = cs[h + 1];
                }
                for (; j < e; ++j) {
                    const float4 Q = sPos[j];
                    const int rq = REF ? rbIdx[sIdx[j]] : __float_as_int(Q.w);
                    if ((int)j == i || ra == rq) continue;
                    const f3 d = sub3(pa, xyz(Q));
                    const float len = sqrtf(dot3(d, d));
                    if (thr - len > best) {
                        best = thr - len;
                        bj = (int)j;
                    }
                }
            }
    if (bj >= 0) {
        const unsigned o2 = sIdx[bj];
        dist[orig] = best;
        pPart[orig] = (int)o2;
        pBody[orig] = REF ? rbIdx[o2] : __float_as_int(sPos[bj].w);
        if (REF && color) color[orig] = make_float4(1.f, 0.f, 0.f, 0.f);
    } else if (!REF) {
        dist[orig] = 0.0f;
        pPart[orig] = -1;
        pBody[orig] = -1;
    }
}

// FindAugmentedRealityCollisionsUniformGridKernel (RB_UniformGridKernels.cu:879-1022): the same search
// over the static points of one scene.
//  REF: static cellStart / cellEnd, the partner is the static point's original index
//       (staticSortedIndex[j]);
//  engine: the static points' exact cell table; the partner is the sorted index j (the engine keeps
//       the static normals in the same sorted order).
template <bool REF>
__global__ void k_sis_find_static(const float4 *__restrict__ sPos, const unsigned *__restrict__ sIdx, int n,
                                  const float4 *__restrict__ stPos, const unsigned *__restrict__ stIdx,
                                  const unsigned *__restrict__ cs, const unsigned *__restrict__ ce, SceneGeom g,
                                  float thr, float *__restrict__ dist, int *__restrict__ partner,
                                  float4 *__restrict__ color) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const f3 pa = xyz(sPos[i]);
    const unsigned orig = sIdx[i];
    int gx, gy, gz;
    grid_pos(pa, g, gx, gy, gz);
    float best = REF ? dist[orig] : 0.0f;
    int bj = -1;
    for (int z = -1; z <= 1; ++z)
        for (int y = -1; y <= 1; ++y)
            for (int x = -1; x <= 1; ++x) {
                const unsigned h = grid_hash(gx + x, gy + y, gz + z, g);
                unsigned j = cs[h], e;
                if (REF) {
                    if (j == kEmpty) continue;
                    e = ce[h];
                } else {
                    e = cs[h + 1];
                }
                for (; j < e; ++j) {
                    const f3 d = sub3(pa, xyz(stPos[j]));
                    const float len = sqrtf(dot3(d, d));
                    if (thr - len > best) {
                        best = thr - len;
                        bj = (int)j;
                    }
                }
            }
    if (bj >= 0) {
        dist[orig] = best;
        partner[orig] = REF ? (int)stIdx[bj] : bj;
        if (REF && color) color[orig] = make_float4(1.f, 0.f, 0.f, 0.f);
    } else if (!REF) {
        dist[orig] = 0.0f;
        partner[orig] = -1;
    }
}

__device__ __forceinline__ bool flag_static(const SisGatherIn &in, int i) {
    return in.distA != nullptr && in.distA[i] > 0.0f;
}
__device__ __forceinline__ bool flag_body(const SisGatherIn &in, int i) {
    return in.distB[i] > 0.0f && in.bodyB[i] >= 0;
}

// Block-wide exclusive rank of a 0/1 flag (kSisBlock threads) and the block's total.
__device__ __forceinline__ unsigned block_rank(bool f, unsigned *sWave, unsigned &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t m = __ballot(f);
    const uint64_t below = lane == 0 ? 0ull : (m & (~0ull >> (64 - lane)));
    const unsigned inWave = (unsigned)__popcll(below);
    if (lane == 0) sWave[wave] = (unsigned)__popcll(m);
    __syncthreads();
    unsigned before = 0u, tot = 0u;
    for (int w = 0; w < kSisWaves; ++w) {
        const unsigned c = sWave[w];
        before += w < wave ? c : 0u;
        tot += c;
    }
    __syncthreads();
    total = tot;
    return before + inWave;
}

// Contact counts of each block of kSisBlock body particles: static contacts and body contacts.
__global__ __launch_bounds__(kSisBlock) void k_sis_count(SisGatherIn in, int nb, unsigned *__restrict__ blockCnt) {
    __shared__ unsigned sWave[kSisWaves];
    const int i = blockIdx.x * kSisBlock + threadIdx.x;
    const bool ok = i < nb;
    unsigned ta, tb;
    block_rank(ok && flag_static(in, i), sWave, ta);
    block_rank(ok && flag_body(in, i), sWave, tb);
    if (threadIdx.x == 0) {
        blockCnt[2 * blockIdx.x] = ta;
        blockCnt[2 * blockIdx.x + 1] = tb;
    }
}

// Exclusive offsets of the blocks' contacts: static contacts first (all blocks), then body contacts;
// the total; and, for one scene, its contact range.  One workgroup.
__global__ __launch_bounds__(kSisBlock) void k_sis_scan(unsigned *__restrict__ blockCnt, int numBlocks,
                                                        int *__restrict__ numContacts, int2 *__restrict__ sceneRange,
                                                        int singleScene) {
    __shared__ unsigned sWave[kSisWaves];
    __shared__ unsigned sCarry[2];
    if (threadIdx.x == 0) sCarry[0] = sCarry[1] = 0u;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int part = 0; part < 2; ++part) {
        for (int base = 0; base < numBlocks; base += kSisBlock) {
            const int b = base + (int)threadIdx.x;
            const unsigned v = b < numBlocks ? blockCnt[2 * b + part] : 0u;
            // inclusive wave scan
            unsigned x = v;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const unsigned y = __shfl_up(x, d, 64);
                if (lane >= d) x += y;
            }
            if (lane == 63) sWave[wave] = x;
            __syncthreads();
            unsigned before = 0u, tot = 0u;
            for (int w = 0; w < kSisWaves; ++w) {
                before += w < wave ? sWave[w] : 0u;
                tot += sWave[w];
            }
            const unsigned carry = sCarry[part];
            if (b < numBlocks) blockCnt[2 * b + part] = carry + before + x - v;
            __syncthreads();
            if (threadIdx.x == 0) sCarry[part] = carry + tot;
            __syncthreads();
        }
    }
    // body contacts follow every static contact
    const unsigned ta = sCarry[0];
    for (int b = threadIdx.x; b < numBlocks; b += kSisBlock) blockCnt[2 * b + 1] += ta;
    if (threadIdx.x == 0) {
        numContacts[0] = (int)(ta + sCarry[1]);
        if (singleScene) sceneRange[0] = make_int2(0, (int)(ta + sCarry[1]));
    }
}

// GatherAugmentedRealityCollisions (Particles/updateFunctions.cpp:1521-1578) and
// GatherRigidBodyCollisions (:1431-1519) with findExactContactPoint (Particles/CPU_Engine.cpp:461-473).
//  static contact: normal = the static point's normal, point = relativePos + CM,
//    bias = ARrestitution * (v + w x r) . n;
//  body contact: p1 = CM1 + r1, p2 = CM2 + r2 (float4), point = p1 + (p2 - p1) * (r / (r + r)),
//    normal = normalize(p2 - p1) * -1 with normalize(v) = v * (1 / sqrt(v . v)) over four components,
//    bias = RBrestitution * ((v1 + w1 x r1) . n - (v2 + w2 x r2) . n).
// With several scenes (no static set), the first and last body particle of each scene record the
// scene's contact range.
__global__ __launch_bounds__(kSisBlock) void k_sis_emit(SisGatherIn in, int nb, const unsigned *__restrict__ blockCnt,
                                                        SisContacts out, int2 *__restrict__ sceneRange) {
    __shared__ unsigned sWave[kSisWaves];
    const int i = blockIdx.x * kSisBlock + threadIdx.x;
    const bool ok = i < nb;
    const bool fa = ok && flag_static(in, i);
    const bool fb = ok && flag_body(in, i);
    unsigned ta, tb;
    const unsigned ra = block_rank(fa, sWave, ta);
    const unsigned rb = block_rank(fb, sWave, tb);
    if (!ok) return;
    const unsigned kb = blockCnt[2 * blockIdx.x + 1] + rb;
    if (in.scene) {
        const int s = in.scene[i];
        if (i == 0 || in.scene[i - 1] != s) sceneRange[s].x = (int)kb;
        if (i == nb - 1 || in.scene[i + 1] != s) sceneRange[s].y = (int)kb + (fb ? 1 : 0);
    }
    if (!fa && !fb) return;
    const int b = in.rb[i];
    const float4 cm = in.rbPos[b], r4 = in.relPos[i];
    const f3 v1 = xyz(in.rbVel[b]), w1 = xyz(in.rbAng[b]), r1 = xyz(r4);
    if (fa) {
        const unsigned k = blockCnt[2 * blockIdx.x] + ra;
        const float4 n = in.staticNorm[in.partA[i]];
        const float4 pt = make_float4(r4.x + cm.x, r4.y + cm.y, r4.z + cm.z, r4.w + cm.w);
        const float vrel = dot3(add3(v1, cross3(w1, r1)), xyz(n));
        out.normal[k] = n;
        out.point[k] = pt;
        out.bodies[k] = make_int2(b, -1);
        out.bias[k] = in.arRest * vrel;
        out.impulse[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    if (fb) {
        const int b2 = in.bodyB[i], j = in.partB[i];
        const float4 cm2 = in.rbPos[b2], r24 = in.relPos[j];
        const float4 p1 = make_float4(cm.x + r4.x, cm.y + r4.y, cm.z + r4.z, cm.w + r4.w);
        const float4 p2 = make_float4(cm2.x + r24.x, cm2.y + r24.y, cm2.z + r24.z, cm2.w + r24.w);
        const float4 d = make_float4(p2.x - p1.x, p2.y - p1.y, p2.z - p1.z, p2.w - p1.w);
        const float inv = 1.0f / sqrtf(dot4(d, d));
        const float t = in.radius / (in.radius + in.radius);
        const float4 pt = make_float4(p1.x + d.x * t, p1.y + d.y * t, p1.z + d.z * t, p1.w + d.w * t);
        const float4 n = make_float4((d.x * inv) * -1.0f, (d.y * inv) * -1.0f, (d.z * inv) * -1.0f,
                                     (d.w * inv) * -1.0f);
        const f3 n3 = xyz(n);
        const float a = dot3(add3(v1, cross3(w1, r1)), n3);
        const float c = dot3(add3(xyz(in.rbVel[b2]), cross3(xyz(in.rbAng[b2]), xyz(r24))), n3);
        out.normal[kb] = n;
        out.point[kb] = pt;
        out.bodies[kb] = make_int2(b, b2);
        out.bias[kb] = in.rbRest * (a - c);
        out.impulse[kb] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

__device__ __forceinline__ int2 scene_range(const int2 *sceneRange, const int *numContacts, int s) {
    return sceneRange ? sceneRange[s] : make_int2(0, numContacts[0]);
}

// Entries of the predecessor search: entry e = 2c + k names body k of contact c (k = 1 only for
// body contacts); the key is the body, or numBodies for an unused entry.  Sorted stably by key,
// the entries of one body come in contact order.
__global__ void k_sis_entries(const int2 *__restrict__ bodies, const int *__restrict__ numContacts, int cap,
                              unsigned invalid, unsigned *__restrict__ keys, int *__restrict__ pred) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 2 * cap) return;
    const int c = e >> 1, k = e & 1;
    unsigned key = invalid;
    if (c < numContacts[0]) {
        const int2 b = bodies[c];
        const int body = k ? b.y : b.x;
        if (body >= 0) key = (unsigned)body;
    }
    keys[e] = key;
    pred[e] = -1;
}

// pred[e] = the contact of the previous entry with the same body (-1 for the first).
__global__ void k_sis_pred(const unsigned *__restrict__ keys, const unsigned *__restrict__ vals, int n,
                           unsigned invalid, int *__restrict__ pred) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned key = keys[i];
    if (key == invalid) return;
    if (i > 0 && keys[i - 1] == key) pred[vals[i]] = (int)(vals[i - 1] >> 1);
}

__device__ __forceinline__ int block_max(int v, int *sRed) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = max(v, __shfl_xor(v, d, 64));
    if (lane == 0) sRed[wave] = v;
    __syncthreads();
    int m = sRed[0];
    for (int w = 1; w < kSisWaves; ++w) m = max(m, sRed[w]);
    __syncthreads();
    return m;
}

// Block-wide inclusive max-scan of x over the threads in segments that start where head is set.
__device__ __forceinline__ int block_seg_max_scan(int x, bool head, int *sAgg, int *sHead) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    bool f = head;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        const int g = __shfl_up((int)f, d, 64);
        if (lane >= d) {
            if (!f) x = max(x, y);
            f = f || g;
        }
    }
    if (lane == 63) {
        sAgg[wave] = x;
        sHead[wave] = f ? 1 : 0;
    }
    __syncthreads();
    // carry into this wave: the running maximum of the segment open at the end of the wave before
    int carry = INT_MIN;
    for (int w = 0; w < wave; ++w) carry = sHead[w] ? sAgg[w] : max(carry, sAgg[w]);
    __syncthreads();
    return f ? x : max(x, carry);
}

// Levels of one scene's contacts (one workgroup per scene), chunk by chunk in contact order:
// level(c) = 1 + max(level of its predecessors), 0 without any.  A predecessor before the chunk has
// its final level already.  Inside the chunk, a contact whose predecessor is the contact just before
// it continues that contact's run (typically the contacts of one body): along a run
// L(c) = max(B(c), L(c - 1) + 1), with B(c) = 1 + the levels of its other predecessors, so
// L(c) = c + max over the run up to c of (B(k) - k), a segmented max-scan.  The scan is repeated with
// the updated B until nothing changes, which takes as many passes as the longest path has jumps
// between runs.
__global__ __launch_bounds__(kSisBlock) void k_sis_levels(const int *__restrict__ pred,
                                                          const int2 *__restrict__ sceneRange,
                                                          const int *__restrict__ numContacts, int *__restrict__ lvl,
                                                          int *__restrict__ maxLvl) {
    __shared__ int sL[kSisBlock];
    __shared__ int sAgg[kSisWaves], sHead[kSisWaves];
    __shared__ int sRed[kSisWaves];
    const int2 rg = scene_range(sceneRange, numContacts, blockIdx.x);
    const int t = (int)threadIdx.x;
    int mx = -1;
    for (int c0 = rg.x; c0 < rg.y; c0 += kSisBlock) {
        const int c = c0 + t;
        const bool ok = c < rg.y;
        int p1 = -1, p2 = -1, outer = 0;
        if (ok) {
            p1 = pred[2 * c];
            p2 = pred[2 * c + 1];
            if (p1 >= 0 && p1 < c0) outer = max(outer, lvl[p1] + 1);
            if (p2 >= 0 && p2 < c0) outer = max(outer, lvl[p2] + 1);
        }
        // the run link (the previous contact is a predecessor) and the other in-chunk predecessors
        const bool link = ok && t > 0 && (p1 == c - 1 || p2 == c - 1);
        const bool side1 = p1 >= c0 && p1 != c - 1, side2 = p2 >= c0 && p2 != c - 1;
        sL[t] = 0;
        __syncthreads();
        while (true) {
            int b = outer;
            if (side1) b = max(b, sL[p1 - c0] + 1);
            if (side2) b = max(b, sL[p2 - c0] + 1);
            const int L = t + block_seg_max_scan(b - t, !link, sAgg, sHead);
            const bool changed = ok && L != sL[t];
            if (changed) sL[t] = L;
            if (!__syncthreads_or(changed)) break;
        }
        if (ok) {
            lvl[c] = sL[t];
            mx = max(mx, sL[t]);
        }
        __syncthreads();
    }
    mx = block_max(mx, sRed);
    if (threadIdx.x == 0) maxLvl[blockIdx.x] = mx;
}

// Contacts of one scene grouped by level: off[L] = first position of level L in order[], for
// L = 0 .. maxLvl + 1; the scene's offsets start at off + range.x + scene.  The per-level counters
// live in LDS when the levels fit, in global memory (cur) otherwise.
constexpr int kBucketLds = 12288;
__global__ __launch_bounds__(kSisBlock) void k_sis_bucket(const int *__restrict__ lvl,
                                                          const int2 *__restrict__ sceneRange,
                                                          const int *__restrict__ numContacts,
                                                          const int *__restrict__ maxLvl, int *__restrict__ offAll,
                                                          int *__restrict__ curAll, int *__restrict__ order) {
    __shared__ int sCnt[kBucketLds];
    __shared__ unsigned sWave[kSisWaves];
    __shared__ int sCarry;
    const int s = blockIdx.x;
    const int2 rg = scene_range(sceneRange, numContacts, s);
    const int levels = maxLvl[s] + 1;
    const bool lds = levels + 1 <= kBucketLds;
    int *off = offAll + rg.x + s;
    int *cnt = lds ? sCnt : curAll + rg.x + s;
    for (int L = threadIdx.x; L <= levels; L += kSisBlock) cnt[L] = 0;
    __syncthreads();
    for (int c = rg.x + (int)threadIdx.x; c < rg.y; c += kSisBlock) atomicAdd(&cnt[lvl[c] + 1], 1);
    __syncthreads();
    // inclusive scan of cnt[0 .. levels] (cnt[0] = 0): the first position of each level
    if (threadIdx.x == 0) sCarry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int base = 0; base <= levels; base += kSisBlock) {
        const int L = base + (int)threadIdx.x;
        const unsigned v = L <= levels ? (unsigned)cnt[L] : 0u;
        unsigned x = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const unsigned y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        if (lane == 63) sWave[wave] = x;
        __syncthreads();
        unsigned before = 0u, tot = 0u;
        for (int w = 0; w < kSisWaves; ++w) {
            before += w < wave ? sWave[w] : 0u;
            tot += sWave[w];
        }
        const int carry = sCarry;
        if (L <= levels) {
            off[L] = carry + (int)(before + x);
            cnt[L] = carry + (int)(before + x);
        }
        __syncthreads();
        if (threadIdx.x == 0) sCarry = carry + (int)tot;
        __syncthreads();
    }
    for (int c = rg.x + (int)threadIdx.x; c < rg.y; c += kSisBlock) order[rg.x + atomicAdd(&cnt[lvl[c]], 1)] = c;
}

__device__ __forceinline__ void ld_mat(const float *__restrict__ Iw, int b, float m[9]) {
#pragma unroll
    for (int k = 0; k < 9; ++k) m[k] = Iw[9 * b + k];
}
// (I^-1 (p x d)) x p . d
__device__ __forceinline__ float sis_k(const float I[9], f3 p, f3 d) {
    return dot3(cross3(mat_vec(I, cross3(p, d)), p), d);
}
// glm::normalize(vec3) of GLM 0.9.7: v * (1 / sqrt(v . v)).
__device__ __forceinline__ f3 gnormalize(f3 v) {
    const float inv = 1.0f / sqrtf(dot3(v, v));
    return f3{v.x * inv, v.y * inv, v.z * inv};
}
__device__ __forceinline__ f3 divs3(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }

// Where the sweeps keep the bodies' linear / angular velocities and read their inverse world-frame
// inertia (fixed over the sweeps): global memory, or a copy of the scene's bodies in LDS (index b - b0).
struct VelGlobal {
    float4 *vel, *ang;
    const float *Iw;
    __device__ float4 v(int b) const { return vel[b]; }
    __device__ float4 w(int b) const { return ang[b]; }
    __device__ void set(int b, f3 lv, f3 av) const {
        vel[b] = f4(lv, 0.0f);
        ang[b] = f4(av, 0.0f);
    }
    __device__ void inertia(int b, float I[9]) const { ld_mat(Iw, b, I); }
};
struct VelLds {
    float4 *s;   // (vel, ang) pairs
    const float *I;   // 9 per body
    int b0;
    __device__ float4 v(int b) const { return s[2 * (b - b0)]; }
    __device__ float4 w(int b) const { return s[2 * (b - b0) + 1]; }
    __device__ void set(int b, f3 lv, f3 av) const {
        s[2 * (b - b0)] = f4(lv, 0.0f);
        s[2 * (b - b0) + 1] = f4(av, 0.0f);
    }
    __device__ void inertia(int b, float Im[9]) const { ld_mat(I, b - b0, Im); }
};

// Sweep records in level order.  k_sis_prep writes, for position q of order[] (contact c =
// order[q]), the contact's data and the terms that stay fixed over the sweeps of a step, formed with
// the expressions the serial loop evaluates in every sweep (so they are the same bits), into planes
// of kSisPlanes float4 (plane k at rec + k * cap):
//   0: (n, bias)   1: (p1, mc)   2: static (t, kt), body (p2, 0)   3: static (t2, kt2)
//   4: (b1 bits, b2 bits, m1, static: fb / body: m2)
//   5: the accumulated impulses (normal, friction, friction 2, 0)
// mc already has the |mc| < 1e-5 -> 1 substitution applied.  A sweep then reads its level's records
// with coalesced loads (a level's positions are consecutive) instead of gathering by contact; the
// bodies' inverse inertia, the same over the sweeps, comes from the bodies (in LDS when they fit).
constexpr int kSisPlanes = 6;
__global__ void k_sis_prep(SisContacts ct, SisBodies bd, SisSolvePhys ph, const int *__restrict__ numContacts,
                           const int *__restrict__ order, int cap, float4 *__restrict__ rec) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= numContacts[0]) return;
    const int c = order[q];
    const size_t P = (size_t)cap;
    const float4 n4 = ct.normal[c], pt = ct.point[c];
    const int2 bb = ct.bodies[c];
    const f3 n = xyz(n4);
    float I1[9], I2[9];
    ld_mat(bd.Iw, bb.x, I1);
    const float m1 = bd.mass[bb.x];
    const float4 c1 = bd.pos[bb.x];
    const f3 p1 = f3{pt.x - c1.x, pt.y - c1.y, pt.z - c1.z};
    rec[q] = make_float4(n.x, n.y, n.z, ct.bias[c]);
    if (bb.y < 0) {
        float mc = 1.0f / m1 + sis_k(I1, p1, n);
        if (fabs((double)mc) < 0.00001) mc = 1.0f;
        const float fb = ph.arFric * mc * fabsf(ph.gy);
        f3 t = f3{1.0f, 0.0f, 0.0f};
        if (fabs((double)(t.x - n.x)) < 0.0001) t = f3{0.0f, 1.0f, 0.0f};
        else if (fabs((double)(t.x + n.x)) < 0.0001) t = f3{0.0f, 1.0f, 0.0f};
        t = sub3(t, scl3(dot3(t, n), n));
        t = gnormalize(t);
        const float kt = 1.0f / m1 + sis_k(I1, p1, t);
        const f3 t2 = gnormalize(cross3(t, n));
        const float kt2 = 1.0f / m1 + sis_k(I1, p1, t2);
        rec[P + q] = f4(p1, mc);
        rec[2 * P + q] = f4(t, kt);
        rec[3 * P + q] = f4(t2, kt2);
        rec[4 * P + q] = make_float4(__int_as_float(bb.x), __int_as_float(-1), m1, fb);
        for (int k = 0; k < 9; ++k) I2[k] = 0.0f;
    } else {
        ld_mat(bd.Iw, bb.y, I2);
        const float m2 = bd.mass[bb.y];
        const float4 c2 = bd.pos[bb.y];
        const f3 p2 = f3{pt.x - c2.x, pt.y - c2.y, pt.z - c2.z};
        float mc = 1.0f / m1 + 1.0f / m2 + sis_k(I1, p1, n) + sis_k(I2, p2, n);
        if (fabs((double)mc) < 0.00001) mc = 1.0f;
        rec[P + q] = f4(p1, mc);
        rec[2 * P + q] = f4(p2, 0.0f);
        rec[4 * P + q] = make_float4(__int_as_float(bb.x), __int_as_float(bb.y), m1, m2);
    }
    rec[5 * P + q] = ct.impulse[c];
}

// One contact of one sweep, record q: SequentialImpulseSolver (Particles/updateFunctions.cpp:1580-1874).
//  static contact (:1603-1713): normal impulse accumulated and clamped to [0, 100]; friction along
//    t and t2, each accumulated and clamped to +-fb; v += J / m, w += I^-1 (p x J),
//    J = jn n + jt t + jt2 t2;
//  body contact (:1714-1860): the normal impulse with both bodies' terms, applied with opposite
//    signs (its friction impulses are not applied there, so they are not formed here).
// Velocities are stored with w = 0.
template <class VA>
__device__ void sis_contact(size_t q, float4 *__restrict__ rec, size_t P, const VA &va) {
    const float4 r0 = rec[q], r1 = rec[P + q], r2 = rec[2 * P + q], r4 = rec[4 * P + q];
    float4 acc = rec[5 * P + q];
    const f3 n = xyz(r0);
    const float bias = r0.w, mc = r1.w;
    const int b1 = __float_as_int(r4.x), b2 = __float_as_int(r4.y);
    float I1[9];
    va.inertia(b1, I1);
    const float upper = 100.0f;
    if (b2 < 0) {
        const float4 r3 = rec[3 * P + q];
        const f3 p = xyz(r1), t = xyz(r2), t2 = xyz(r3);
        const float kt = r2.w, kt2 = r3.w, m = r4.z, fb = r4.w;
        f3 v = xyz(va.v(b1)), w = xyz(va.w(b1));
        const float vrel = dot3(add3(v, cross3(w, p)), n);
        float ci = -(vrel + bias) / mc;
        float tmp = acc.x + ci;
        if (tmp < 0.0f) tmp = 0.0f;
        else if (tmp > upper) tmp = upper;
        ci = tmp - acc.x;
        acc.x = tmp;
        const f3 vel = add3(v, cross3(w, p));
        float cf = -dot3(vel, t) / kt;
        float tf = acc.y + cf;
        if (tf < -fb) tf = -fb;
        else if (tf > fb) tf = fb;
        cf = tf - acc.y;
        acc.y = tf;
        float cf2 = -dot3(vel, t2) / kt2;
        float tf2 = acc.z + cf2;
        if (tf2 < -fb) tf2 = -fb;
        else if (tf2 > fb) tf2 = fb;
        cf2 = tf2 - acc.z;
        acc.z = tf2;
        const f3 J = add3(add3(scl3(ci, n), scl3(cf, t)), scl3(cf2, t2));
        v = add3(v, divs3(J, m));
        w = add3(w, mat_vec(I1, cross3(p, J)));
        va.set(b1, v, w);
    } else {
        float I2[9];
        va.inertia(b2, I2);
        const f3 p1 = xyz(r1), p2 = xyz(r2);
        const float m1 = r4.z, m2 = r4.w;
        f3 v1 = xyz(va.v(b1)), w1 = xyz(va.w(b1)), v2 = xyz(va.v(b2)), w2 = xyz(va.w(b2));
        const float vrel = dot3(add3(v1, cross3(w1, p1)), n) - dot3(add3(v2, cross3(w2, p2)), n);
        float ci = -(vrel + bias) / mc;
        float tmp = acc.x + ci;
        if (tmp < 0.0f) tmp = 0.0f;
        else if (tmp > upper) tmp = upper;
        ci = tmp - acc.x;
        acc.x = tmp;
        const f3 J = scl3(ci, n);
        v1 = add3(v1, divs3(J, m1));
        w1 = add3(w1, mat_vec(I1, cross3(p1, J)));
        va.set(b1, v1, w1);
        v2 = sub3(v2, divs3(J, m2));
        w2 = sub3(w2, mat_vec(I2, cross3(p2, J)));
        va.set(b2, v2, w2);
    }
    rec[5 * P + q] = acc;
}

// The sweeps of one scene (one workgroup per scene): level by level, a barrier between levels.
// LDS: the scene's bodies' velocities and inverse inertia live in LDS for the whole kernel (copied in,
// velocities copied out), so a level's round trip is an LDS access instead of a global one.
template <bool LDS>
__global__ __launch_bounds__(kSisBlock) void k_sis_solve(float4 *__restrict__ rec, int cap, float4 *__restrict__ impulse,
                                                         SisBodies bd, const int2 *__restrict__ sceneRange,
                                                         const int *__restrict__ numContacts,
                                                         const int2 *__restrict__ sceneBodies, int numBodies,
                                                         const int *__restrict__ maxLvl,
                                                         const int *__restrict__ offAll,
                                                         const int *__restrict__ order, int iterations) {
    extern __shared__ float4 sVel[];
    const int s = blockIdx.x;
    const int2 rg = scene_range(sceneRange, numContacts, s);
    const int levels = maxLvl[s] + 1;
    const int *off = offAll + rg.x + s;
    const size_t P = (size_t)cap;
    const int2 br = sceneBodies ? sceneBodies[s] : make_int2(0, numBodies);
    float *sI = reinterpret_cast<float *>(sVel + 2 * (size_t)br.y);   // after the (vel, ang) pairs
    if (LDS) {
        for (int k = threadIdx.x; k < br.y; k += kSisBlock) {
            sVel[2 * k] = bd.vel[br.x + k];
            sVel[2 * k + 1] = bd.ang[br.x + k];
        }
        for (int k = threadIdx.x; k < 9 * br.y; k += kSisBlock) sI[k] = bd.Iw[9 * (size_t)br.x + k];
        __syncthreads();
    }
    const VelLds vl{sVel, sI, br.x};
    const VelGlobal vg{bd.vel, bd.ang, bd.Iw};
    for (int it = 0; it < iterations; ++it)
        for (int L = 0; L < levels; ++L) {
            const int a = off[L], e = off[L + 1];
            for (int k = a + (int)threadIdx.x; k < e; k += kSisBlock) {
                if (LDS) sis_contact((size_t)(rg.x + k), rec, P, vl);
                else sis_contact((size_t)(rg.x + k), rec, P, vg);
            }
            __syncthreads();
        }
    if (LDS)
        for (int k = threadIdx.x; k < br.y; k += kSisBlock) {
            bd.vel[br.x + k] = sVel[2 * k];
            bd.ang[br.x + k] = sVel[2 * k + 1];
        }
    for (int k = rg.x + (int)threadIdx.x; k < rg.y; k += kSisBlock) impulse[order[k]] = rec[5 * P + k];
}

}  // namespace

void launch_sis_find(const float4 *sPos, const float4 *sVel, const unsigned *sIdx, const unsigned *segStart,
                     const SceneView &sv, bool sceneInVelW, int n, float radius, float *dist, int *pBody, int *pPart,
                     hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_sis_find<false>, dim3((n + 127) / 128), dim3(128), 0, st, sPos, sVel, sIdx, nullptr,
                       segStart, nullptr, sv, sceneInVelW ? 1 : 0, n, 2.0f * radius, dist, pBody, pPart, nullptr);
}

void launch_sis_find_ref(const float4 *sPos, const unsigned *sIdx, const int *rbIdx, const unsigned *cellStart,
                         const unsigned *cellEnd, const SceneGeom &g, int n, float radius, float *dist, int *pBody,
                         int *pPart, float4 *color, hipStream_t st) {
    if (n <= 0) return;
    SceneView sv{};
    sv.count = 1;
    sv.single = g;
    hipLaunchKernelGGL(k_sis_find<true>, dim3((n + 127) / 128), dim3(128), 0, st, sPos, nullptr, sIdx, rbIdx,
                       cellStart, cellEnd, sv, 0, n, 2.0f * radius, dist, pBody, pPart, color);
}

void launch_sis_find_static(const float4 *sPos, const unsigned *sIdx, int n, const float4 *stPos,
                            const unsigned *stIdx, const unsigned *cellStart, const unsigned *cellEnd,
                            const SceneGeom &g, float radius, bool refLayout, float *dist, int *partner,
                            float4 *color, hipStream_t st) {
    if (n <= 0) return;
    const dim3 grid((n + 127) / 128);
    if (refLayout)
        hipLaunchKernelGGL(k_sis_find_static<true>, grid, dim3(128), 0, st, sPos, sIdx, n, stPos, stIdx, cellStart,
                           cellEnd, g, 2.0f * radius, dist, partner, color);
    else
        hipLaunchKernelGGL(k_sis_find_static<false>, grid, dim3(128), 0, st, sPos, sIdx, n, stPos, stIdx, cellStart,
                           cellEnd, g, 2.0f * radius, dist, partner, color);
}

size_t sis_workspace_bytes(int cap, int numScenes, int numBodyParticles) {
    const SisWork w = sis_work_layout(nullptr, cap, numScenes, numBodyParticles);
    return (size_t)w.end;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

SisWork sis_work_layout(void *base, int cap, int numScenes, int numBodyParticles) {
    SisWork w{};
    char *p = (char *)base;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align256(o + bytes);
        return p ? (void *)(p + at) : nullptr;
    };
    const size_t ne = 2 * (size_t)cap;
    const int nblk = (numBodyParticles + kSisBlock - 1) / kSisBlock;
    w.keysA = (unsigned *)take(ne * 4);
    w.keysB = (unsigned *)take(ne * 4);
    w.valsA = (unsigned *)take(ne * 4);
    w.valsB = (unsigned *)take(ne * 4);
    w.sortWsBytes = sort_workspace_bytes((int)ne, 4);
    w.sortWs = take(w.sortWsBytes);
    w.pred = (int *)take(ne * 4);
    w.lvl = (int *)take((size_t)cap * 4);
    w.order = (int *)take((size_t)cap * 4);
    w.off = (int *)take(((size_t)cap + numScenes + 1) * 4);
    w.cur = (int *)take(((size_t)cap + numScenes + 1) * 4);
    w.maxLvl = (int *)take((size_t)numScenes * 4);
    w.sceneRange = (int2 *)take((size_t)numScenes * 8);
    w.prep = (float4 *)take((size_t)kSisPlanes * cap * 16);
    w.blockCnt = (unsigned *)take(2 * (size_t)(nblk + 1) * 4);
    w.numContacts = (int *)take(4);
    w.end = o;
    return w;
}

void launch_sis_gather(const SisGatherIn &in, int nb, const SisContacts &out, const SisWork &w, bool perScene,
                       hipStream_t st) {
    const int nblk = (nb + kSisBlock - 1) / kSisBlock;
    if (nblk > 0) hipLaunchKernelGGL(k_sis_count, dim3(nblk), dim3(kSisBlock), 0, st, in, nb, w.blockCnt);
    hipLaunchKernelGGL(k_sis_scan, dim3(1), dim3(kSisBlock), 0, st, w.blockCnt, nblk, w.numContacts, w.sceneRange,
                       perScene ? 0 : 1);
    if (nblk > 0)
        hipLaunchKernelGGL(k_sis_emit, dim3(nblk), dim3(kSisBlock), 0, st, in, nb, w.blockCnt, out,
                           perScene ? w.sceneRange : nullptr);
}

int launch_sis_solve(const SisContacts &ct, const SisBodies &bd, const SisSolvePhys &ph, int cap, int numScenes,
                     int numBodies, const int2 *sceneBodies, int maxSceneBodies, int iterations, const SisWork &w,
                     const SortTuning &tu, hipStream_t st) {
    if (cap <= 0 || numScenes <= 0) return 0;
    const int ne = 2 * cap;
    const unsigned invalid = (unsigned)numBodies;
    int bits = 1;
    while (bits < 32 && (invalid >> bits) != 0u) ++bits;
    hipLaunchKernelGGL(k_sis_entries, dim3((ne + 255) / 256), dim3(256), 0, st, ct.bodies, w.numContacts, cap,
                       invalid, w.keysA, w.pred);
    unsigned *kA = w.keysA, *kB = w.keysB, *vA = w.valsA, *vB = w.valsB;
    if (sort_pairs<unsigned>(&kA, &vA, &kB, &vB, ne, 0, bits, true, w.sortWs, w.sortWsBytes, tu, st)) return 1;
    hipLaunchKernelGGL(k_sis_pred, dim3((ne + 255) / 256), dim3(256), 0, st, kA, vA, ne, invalid, w.pred);
    const int2 *rg = numScenes > 1 ? w.sceneRange : nullptr;
    hipLaunchKernelGGL(k_sis_levels, dim3(numScenes), dim3(kSisBlock), 0, st, w.pred, rg, w.numContacts, w.lvl,
                       w.maxLvl);
    hipLaunchKernelGGL(k_sis_bucket, dim3(numScenes), dim3(kSisBlock), 0, st, w.lvl, rg, w.numContacts, w.maxLvl,
                       w.off, w.cur, w.order);
    if (iterations > 0) {
        hipLaunchKernelGGL(k_sis_prep, dim3((cap + 255) / 256), dim3(256), 0, st, ct, bd, ph, w.numContacts, w.order,
                           cap, w.prep);
        if (maxSceneBodies <= kSisLdsBodies)
            hipLaunchKernelGGL(k_sis_solve<true>, dim3(numScenes), dim3(kSisBlock),
                               (size_t)maxSceneBodies * (2 * sizeof(float4) + 9 * sizeof(float)), st, w.prep, cap,
                               ct.impulse, bd, rg,
                               w.numContacts, sceneBodies, numBodies, w.maxLvl, w.off, w.order, iterations);
        else
            hipLaunchKernelGGL(k_sis_solve<false>, dim3(numScenes), dim3(kSisBlock), 0, st, w.prep, cap, ct.impulse,
                               bd, rg, w.numContacts, sceneBodies, numBodies, w.maxLvl, w.off, w.order, iterations);
    }
    return 0;
}

}  // namespace arsim
