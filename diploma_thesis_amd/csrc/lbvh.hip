// lbvh.hip -- Karras LBVH: radix-tree build, leaf init + atomic bottom-up refit, and the
// stack traversal with the DEM contact force.
#include <algorithm>

#include "kernels.hpp"

namespace arsim {

namespace {

constexpr int kStackDepth = 64;   // per-thread LDS stack entries (the reference's int stack[64], RB_BVHKernels.cu:129)
constexpr int kWaveStackDepth = 64;   // per-wave LDS stack entries of the packet traversal

#ifdef ARSIM_TRAVERSAL_STATS
// Counting build only: totals of the packet traversal.
//   [0] waves  [1] internal nodes visited, summed over packets  [2] valid lanes  [3] lane-leaf tests
//   [4] lanes whose box overlaps an entered node, summed over entered nodes  [5] largest stack depth
//   [6] lanes that completed  [7] contacts found  [8] packets  [9] iterations of the waves' node loops
constexpr int kTravStats = 10;
__device__ unsigned long long g_travStats[kTravStats];
#endif

__device__ __forceinline__ int clz_key(unsigned v) { return __clz((int)v); }
__device__ __forceinline__ int clz_key(uint64_t v) { return __clzll((long long)v); }

// Leaf block of the engine's range refit: a node whose leaf range spans two blocks "crosses".
constexpr int kBlk = 1024;
__device__ __forceinline__ bool range_crosses(int lo, int hi) { return lo / kBlk != hi / kBlk; }

// The codes a workgroup's searches read, [i0 - kKarrasHalo, i0 + kKarrasBlock + kKarrasHalo), are
// staged in LDS: a node whose range is at most kKarrasHalo / 2 long (most of them) searches LDS only;
// wider ranges read the rest from global memory.
constexpr int kKarrasBlock = 1024;
constexpr int kKarrasHalo = 1024;
constexpr int kKarrasWin = kKarrasBlock + 2 * kKarrasHalo;

// longestCommonPrefix (BVHSource/BVHAuxiliary.cuh:31-43).  Equal keys always belong to one
// scene (the scene id is in the high key bits), so the tie-break uses scene-local indices
// li = i - base, lj = j - base and a batched subtree matches the standalone tree.
template <class K>
__device__ __forceinline__ int lcp(const K *__restrict__ codes, const K *win, int w0, int n, K ki, int i, int j,
                                   int base) {
    if (j < 0 || j >= n) return -1;
    const unsigned q = (unsigned)(j - w0);
    K kj = q < (unsigned)kKarrasWin ? win[q] : codes[j];
    if (ki != kj) return clz_key((K)(ki ^ kj));
    return (int)sizeof(K) * 8 + __clz((int)((unsigned)(i - base) ^ (unsigned)(j - base)));
}

// Workgroup g of the build lists its nodes whose range crosses a refit block in
// crossList[g * kKarrasBlock + j], j < crossCount[g] (the order within a workgroup is immaterial).
// Every thread of the workgroup calls this once (it holds barriers).
__device__ __forceinline__ void karras_list_crossing(const Lbvh &t, bool crosses, int i, unsigned &crossings) {
    const unsigned j = crosses ? __hip_atomic_fetch_add(&crossings, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)
                               : 0u;
    if (crosses) t.crossList[(size_t)blockIdx.x * kKarrasBlock + j] = i;
    __syncthreads();
    if (threadIdx.x == 0) t.crossCount[blockIdx.x] = crossings;
}

// kernelConstructRadixTreeSoA (BVHSource/BVHSoAKernels.cu:7-99).  t = ceil(l / 2^k) is computed
// with integer shifts instead of the reference's double division (identical for l < 2^31).
template <class K>
__global__ __launch_bounds__(kKarrasBlock) void k_karras(const K *__restrict__ codes, int n, int sceneShift,
                                                         const int *__restrict__ sceneStart,
                                                         const int *__restrict__ sceneCount, Lbvh t) {
    __shared__ K win[kKarrasWin];
    const int i0 = (int)blockIdx.x * kKarrasBlock, w0 = i0 - kKarrasHalo;
    for (int q = (int)threadIdx.x; q < kKarrasWin; q += kKarrasBlock) {
        const int j = w0 + q;
        if (j >= 0 && j < n) win[q] = codes[j];
    }
    __shared__ unsigned crossings;
    if (threadIdx.x == 0) crossings = 0;
    __syncthreads();
    const int i = i0 + (int)threadIdx.x;
    if (i == 0) t.parent[n] = -1;
    bool crosses = false;
    if (i < n - 1) {
        if (!t.meta) t.counter[i] = 0;   // the bottom-up refit's arrival counters (the range refit has none)
        auto lcp = [&](K ki_, int i_, int j_, int base_) {
            return arsim::lcp(codes, win, w0, n, ki_, i_, j_, base_);
        };
        const K ki = win[i - w0];
        const int sci = sceneShift > 0 ? (int)(ki >> sceneShift) : 0;
        const int base = sceneShift > 0 ? sceneStart[sci] : 0;
        // loaded here rather than after the node's stores, where it would cost a round trip of its own
        const int sceneCnt = t.sceneRoot && sceneShift > 0 ? sceneCount[sci] : 0;
        const int dPrev = lcp(ki, i, i - 1, base);
        const int dNext = lcp(ki, i, i + 1, base);
        const int d = (dNext - dPrev) > 0 ? 1 : -1;
        const int sigMin = lcp(ki, i, i - d, base);
        long long lmax = 2;
        while (lcp(ki, i, (int)(i + lmax * d), base) > sigMin) lmax *= 2;
        long long l = 0;
        for (long long s = lmax / 2; s >= 1; s /= 2) {
            if (lcp(ki, i, (int)(i + (l + s) * d), base) > sigMin) l += s;
            if (s == 1) break;
        }
        const int j = (int)(i + l * d);
        const int sigNode = lcp(ki, i, j, base);
        long long s = 0;
        for (int k = 1;; ++k) {
            long long tt = (l + (1ll << k) - 1) >> k;
            if (lcp(ki, i, (int)(i + (s + tt) * d), base) > sigNode) s += tt;
            if (tt == 1) break;
        }
        const int gamma = (int)(i + s * d + (d < 0 ? d : 0));
        const int node = i + n;
        const int lo = i < j ? i : j, hi = i < j ? j : i;
        const int L = (lo == gamma) ? gamma : gamma + n;
        const int R = (hi == gamma + 1) ? gamma + 1 : gamma + 1 + n;
        t.left[node] = L;
        t.right[node] = R;
        t.parent[L] = node;
        t.parent[R] = node;
        if (t.minR) { t.minR[node] = lo; t.maxR[node] = hi; }
        if (t.meta) t.meta[i] = make_int4(L, R, lo, hi);
        if (t.isLeaf) t.isLeaf[node] = 0;
        // the root of leaf i's scene: a range that starts in another scene starts before this one's
        // first leaf and spans two scenes, so testing against leaf i's scene alone decides
        if (t.sceneRoot && sceneShift > 0 && lo == base && hi == base + sceneCnt - 1) t.sceneRoot[sci] = node;
        crosses = range_crosses(lo, hi);
    }
    if (t.crossList) karras_list_crossing(t, crosses, i, crossings);   // every thread: it holds a barrier
}

// Node bounds that other workgroups read during the refit travel as agent-scope atomic stores and
// loads (coherent across the XCDs' L2s), so the counter add needs no L2 writeback / invalidate: the
// writer waits for its stores to complete before its add, and the merging thread loads only after
// its own add has returned the sibling's increment.
__device__ __forceinline__ uint64_t pack2(float lo, float hi) {
    return ((uint64_t)__float_as_uint(hi) << 32) | (uint64_t)__float_as_uint(lo);
}
__device__ __forceinline__ void store_bounds(float4 *bounds, size_t node, float4 mn, float4 mx) {
    uint64_t *p = reinterpret_cast<uint64_t *>(bounds + 2 * node);
    __hip_atomic_store(p + 0, pack2(mn.x, mn.y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 1, pack2(mn.z, mn.w), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 2, pack2(mx.x, mx.y), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 3, pack2(mx.z, mx.w), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void load_bounds(float4 *bounds, size_t node, float4 &mn, float4 &mx) {
    uint64_t *p = reinterpret_cast<uint64_t *>(bounds + 2 * node);
    const uint64_t a = __hip_atomic_load(p + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t b = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t c = __hip_atomic_load(p + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t d = __hip_atomic_load(p + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    mn = make_float4(__uint_as_float((unsigned)a), __uint_as_float((unsigned)(a >> 32)), __uint_as_float((unsigned)b),
                     __uint_as_float((unsigned)(b >> 32)));
    mx = make_float4(__uint_as_float((unsigned)c), __uint_as_float((unsigned)(c >> 32)), __uint_as_float((unsigned)d),
                     __uint_as_float((unsigned)(d >> 32)));
}
// Waits until this thread's stores have completed (the compiler may not move memory ops across it).
__device__ __forceinline__ void stores_complete() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// kernelConstructLeafNodesSoA (BVHSoAKernels.cu:108-138) + kernelConstructInternalNodesSoA
// (:148-176) in one launch: each leaf writes its centre (w = rbIndex bits), velocity copy and
// AABB = c ± r, then walks up; the first arrival at a node stops, the second merges the children.
// With the node ranges (minR / maxR), nodes whose leaves all belong to this workgroup's 256 leaves
// are merged through LDS (counters and bounds in LDS, workgroup-scope ordering); only nodes whose
// range leaves the workgroup use the device-wide counters and coherent bound loads.  A node's bounds
// are the exact min / max of its leaves' either way.
__device__ __forceinline__ float4 fmin4(float4 a, float4 b) {
    return make_float4(a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z, 0.0f);
}
__device__ __forceinline__ float4 fmax4(float4 a, float4 b) {
    return make_float4(a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z, 0.0f);
}
__device__ __forceinline__ void write_packed(const Lbvh &t, int n, int node, int L, int R, float4 aMin, float4 aMax,
                                             float4 bMin, float4 bMax) {
    if (!t.packed) return;
    float4 *q = t.packed + 4 * (size_t)(node - n);
    q[0] = make_float4(aMin.x, aMin.y, aMin.z, __int_as_float(L));
    q[1] = make_float4(aMax.x, aMax.y, aMax.z, __int_as_float(R));
    q[2] = make_float4(bMin.x, bMin.y, bMin.z, 0.0f);
    q[3] = make_float4(bMax.x, bMax.y, bMax.z, 0.0f);
}

// Leaf initialisation shared by both refit forms: centre (w = rbIndex bits), velocity copy, AABB.
// o = sIdx[i], loaded by the caller.  Every load precedes the first store: the compiler may not move
// a load above a store to a buffer it cannot tell apart, so a load after a store would cost a round
// trip of its own.
__device__ __forceinline__ void leaf_init(int i, int o, const float4 *__restrict__ pos,
                                          const float4 *__restrict__ vel, const int *__restrict__ rb,
                                          const uint16_t *__restrict__ scene, const float4 *__restrict__ rec, float r,
                                          const Lbvh &t, float4 &lmin, float4 &lmax, float4 &centre) {
    float4 p, v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (rec) {   // the key kernel's gather record: (pos, rb bits), (vel, scene bits)
        p = rec[2 * (size_t)o];
        if (t.sortedVel) v = rec[2 * (size_t)o + 1];
    } else {
        p = pos[o];
        if (rb) p.w = __int_as_float(rb[o]);
        if (t.sortedVel) {
            v = vel[o];
            if (scene) v.w = __int_as_float((int)scene[o]);
        }
    }
    t.cms[i] = p;
    if (t.sortedVel) t.sortedVel[i] = v;
    centre = p;
    lmin = make_float4(p.x - r, p.y - r, p.z - r, 0.0f);
    lmax = make_float4(p.x + r, p.y + r, p.z + r, 0.0f);
    if (t.isLeaf) t.isLeaf[i] = 1;
    if (t.minR) { t.minR[i] = i; t.maxR[i] = i; }
}

// Without node ranges: every node by the device-wide protocol (coherent bounds, counters).
__global__ __launch_bounds__(256) void k_leaf_refit(const int *__restrict__ sIdx, const float4 *__restrict__ pos,
                                                    const float4 *__restrict__ vel, const int *__restrict__ rb,
                                                    const uint16_t *__restrict__ scene,
                                                    const float4 *__restrict__ rec, float r, int n, Lbvh t) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    float4 lmin, lmax, centre;
    leaf_init(i, sIdx[i], pos, vel, rb, scene, rec, r, t, lmin, lmax, centre);
    store_bounds(t.bounds, (size_t)i, lmin, lmax);
    if (n < 2) return;
    int node = t.parent[i];
    stores_complete();
    int prev = __hip_atomic_fetch_add(&t.counter[node - n], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (prev == 1) {
        const int L = t.left[node], R = t.right[node];
        float4 aMin, aMax, bMin, bMax;
        load_bounds(t.bounds, (size_t)L, aMin, aMax);
        load_bounds(t.bounds, (size_t)R, bMin, bMax);
        const float4 mn = fmin4(aMin, bMin), mx = fmax4(aMax, bMax);
        store_bounds(t.bounds, (size_t)node, mn, mx);
        write_packed(t, n, node, L, R, aMin, aMax, bMin, bMax);
        const int up = t.parent[node];
        if (up < 0) break;
        node = up;
        stores_complete();
        prev = __hip_atomic_fetch_add(&t.counter[node - n], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Range refit (engine), three launches and no device-wide synchronisation.  A node's AABB is the
// min / max over the leaves of its range [minR, maxR], whatever the order of the merges, so:
//  1. k_refit_local, one workgroup per kBlk consecutive leaves: leaf initialisation; inclusive
//     prefix (pre) and suffix (suf) min / max of the leaf boxes within the workgroup; the
//     workgroup's box (level 0 of the table); and the bottom-up merge, through LDS counters and
//     bounds, of every node whose range lies inside the workgroup ("local").
//  2. k_refit_table, one workgroup: a sparse table of the workgroup boxes (level l holds the box of
//     2^l consecutive workgroups).
//  3. k_refit_cross, one thread per internal node: a node whose range crosses a workgroup boundary
//     takes suf[lo] ∪ pre[hi] ∪ (the full workgroups between, two table lookups), and the same for
//     its crossing children, for its packed traversal record.
__device__ __forceinline__ void store_plain(float4 *bounds, int node, float4 mn, float4 mx) {
    bounds[2 * (size_t)node] = mn;
    bounds[2 * (size_t)node + 1] = mx;
}


__device__ __forceinline__ void shfl_minmax(float4 &mn, float4 &mx, int d, bool up) {
    float v[6] = {mn.x, mn.y, mn.z, mx.x, mx.y, mx.z};
#pragma unroll
    for (int c = 0; c < 6; ++c) v[c] = up ? __shfl_up(v[c], d) : __shfl_down(v[c], d);
    mn = make_float4(v[0], v[1], v[2], 0.0f);
    mx = make_float4(v[3], v[4], v[5], 0.0f);
}

__global__ __launch_bounds__(kBlk) void k_refit_local(const int *__restrict__ sIdx, const float4 *__restrict__ pos,
                                                      const float4 *__restrict__ vel, const int *__restrict__ rb,
                                                      const uint16_t *__restrict__ scene,
                                                      const float4 *__restrict__ rec, float r, int n, Lbvh t) {
    __shared__ float4 sCtr[kBlk];        // leaf centres of this workgroup (a leaf's box is centre -/+ r)
    __shared__ float4 sNode[2 * kBlk];   // boxes of its local internal nodes
    __shared__ int4 sTop[kBlk];          // internal node b0 + slot: left, right, parent, other end of its range
    __shared__ int sCnt[kBlk];
    __shared__ unsigned char sNeedSuf[kBlk];   // a crossing range starts at this leaf: its suf is read
    __shared__ unsigned char sNeedPre[kBlk];   // a crossing range ends at this leaf: its pre is read
    __shared__ float4 sWave[4 * (kBlk / 64)];   // per wave: prefix total min / max, suffix total min / max
    const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int b0 = (int)blockIdx.x * kBlk;
    const int i = b0 + tid;
    const bool valid = i < n;
    sCnt[tid] = 0;
    sNeedSuf[tid] = 0;
    sNeedPre[tid] = 0;
    // The topology of internal nodes [b0, b0 + kBlk), staged so that the walk up reads LDS only: an
    // internal node's index is one end of its leaf range, so every node whose range lies inside this
    // workgroup's leaves is one of them.
    // (the topology's loads and the leaf's sorted index are issued together, the leaf's own loads
    // next, and the staging waits for them after that)
    int4 m = make_int4(0, 0, 0, 0);
    int upTop = -1;
    if (i < n - 1) {
        m = t.meta[i];
        upTop = t.parent[i + n];
    }
    int node = valid && n >= 2 ? t.parent[i] : -1;
    const int o = valid ? sIdx[i] : 0;
    const float inf = __int_as_float(0x7f800000);
    float4 lmin = make_float4(inf, inf, inf, 0.0f), lmax = make_float4(-inf, -inf, -inf, 0.0f);
    if (valid) {
        float4 c;
        leaf_init(i, o, pos, vel, rb, scene, rec, r, t, lmin, lmax, c);
        if (t.bounds) store_plain(t.bounds, i, lmin, lmax);
        sCtr[tid] = c;
    }
    if (i < n - 1) sTop[tid] = make_int4(m.x, m.y, upTop, m.z == i ? m.w : m.z);
    // inclusive prefix and suffix min / max within the wave
    float4 pmn = lmin, pmx = lmax, smn = lmin, smx = lmax;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        float4 a = pmn, b = pmx;
        shfl_minmax(a, b, d, true);
        if (lane >= d) { pmn = fmin4(pmn, a); pmx = fmax4(pmx, b); }
        float4 c = smn, e = smx;
        shfl_minmax(c, e, d, false);
        if (lane + d < 64) { smn = fmin4(smn, c); smx = fmax4(smx, e); }
    }
    if (lane == 63) { sWave[4 * wave] = pmn; sWave[4 * wave + 1] = pmx; }
    if (lane == 0) { sWave[4 * wave + 2] = smn; sWave[4 * wave + 3] = smx; }
    __syncthreads();
    const int nw = kBlk / 64;
    for (int w = 0; w < nw; ++w) {   // the waves before this one for the prefix, after it for the suffix
        if (w < wave) { pmn = fmin4(pmn, sWave[4 * w]); pmx = fmax4(pmx, sWave[4 * w + 1]); }
        if (w > wave) { smn = fmin4(smn, sWave[4 * w + 2]); smx = fmax4(smx, sWave[4 * w + 3]); }
    }
    if (tid == kBlk - 1) {   // the whole workgroup's box: the last thread's prefix
        t.blockTab[2 * (size_t)blockIdx.x] = pmn;
        t.blockTab[2 * (size_t)blockIdx.x + 1] = pmx;
    }
    auto box = [&](int c, float4 &mn, float4 &mx) {
        if (c < n) {
            const float4 p = sCtr[c - b0];
            mn = make_float4(p.x - r, p.y - r, p.z - r, 0.0f);
            mx = make_float4(p.x + r, p.y + r, p.z + r, 0.0f);
        } else {
            mn = sNode[2 * (c - n - b0)];
            mx = sNode[2 * (c - n - b0) + 1];
        }
    };
    // Walk up while the nodes are local; the second arrival at a node merges it.  A local node (or
    // leaf) whose parent's range crosses the workgroup's end marks its range's ends: every crossing
    // range starts where the left-most such node of its left spine starts and ends where the
    // right-most of its right spine ends, so the crossing kernel reads suf / pre only there.
    int lo = i, hi = i;
    while (valid && node >= 0) {
        const int slot = node - n - b0;
        int4 m = make_int4(0, 0, 0, 0);
        bool crosses = (unsigned)slot >= (unsigned)kBlk;   // not one of ours: its range leaves the workgroup
        if (!crosses) {
            m = sTop[slot];
            const int k = node - n;
            crosses = range_crosses(min(k, m.w), max(k, m.w));
        }
        if (crosses) {
            sNeedSuf[lo - b0] = 1;
            sNeedPre[hi - b0] = 1;
            break;
        }
        const int prev = __hip_atomic_fetch_add(&sCnt[slot], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (prev == 0) break;
        float4 aMin, aMax, bMin, bMax;
        box(m.x, aMin, aMax);
        box(m.y, bMin, bMax);
        const float4 mn = fmin4(aMin, bMin), mx = fmax4(aMax, bMax);
        sNode[2 * slot] = mn;
        sNode[2 * slot + 1] = mx;
        if (t.bounds) store_plain(t.bounds, node, mn, mx);
        write_packed(t, n, node, m.x, m.y, aMin, aMax, bMin, bMax);
        const int k = node - n;
        lo = min(k, m.w);
        hi = max(k, m.w);
        node = m.z;
    }
    if ((int)gridDim.x == 1) return;   // one workgroup: no range crosses
    __syncthreads();
    if (valid) {
        if (sNeedSuf[tid]) { t.suf[2 * (size_t)i] = smn; t.suf[2 * (size_t)i + 1] = smx; }
        if (sNeedPre[tid]) { t.pre[2 * (size_t)i] = pmn; t.pre[2 * (size_t)i + 1] = pmx; }
    }
}

// Level l, entry b: box of workgroups [b, b + 2^l).  Up to kTabLds workgroup boxes one workgroup builds
// the levels in order, each from the previous one held in LDS (as six planes of floats); larger tables
// take a launch per level (k_refit_table_level).
constexpr int kTabLds = 4096;
constexpr int kTabPer = kTabLds / 1024;   // entries per thread
__global__ __launch_bounds__(1024) void k_refit_table(int numBlocks, int levels, float4 *__restrict__ tab) {
    __shared__ float sBox[6][kTabLds];
    const int tid = (int)threadIdx.x;
    if (numBlocks > kTabLds) return;   // larger tables: k_refit_table_level
    {
        for (int b = tid; b < numBlocks; b += 1024) {
            const float4 mn = tab[2 * b], mx = tab[2 * b + 1];
            sBox[0][b] = mn.x; sBox[1][b] = mn.y; sBox[2][b] = mn.z;
            sBox[3][b] = mx.x; sBox[4][b] = mx.y; sBox[5][b] = mx.z;
        }
        __syncthreads();
        for (int l = 1; l < levels; ++l) {
            float4 *dst = tab + 2 * (size_t)l * numBlocks;
            const int half = 1 << (l - 1), cnt = numBlocks - (1 << l) + 1;
            float4 mn[kTabPer], mx[kTabPer];
#pragma unroll
            for (int q = 0; q < kTabPer; ++q) {
                const int b = tid + 1024 * q;
                if (b >= cnt) break;
                const float4 aMin = make_float4(sBox[0][b], sBox[1][b], sBox[2][b], 0.0f);
                const float4 aMax = make_float4(sBox[3][b], sBox[4][b], sBox[5][b], 0.0f);
                const int c = b + half;
                const float4 bMin = make_float4(sBox[0][c], sBox[1][c], sBox[2][c], 0.0f);
                const float4 bMax = make_float4(sBox[3][c], sBox[4][c], sBox[5][c], 0.0f);
                mn[q] = fmin4(aMin, bMin);
                mx[q] = fmax4(aMax, bMax);
                dst[2 * b] = mn[q];
                dst[2 * b + 1] = mx[q];
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < kTabPer; ++q) {
                const int b = tid + 1024 * q;
                if (b >= cnt) break;
                sBox[0][b] = mn[q].x; sBox[1][b] = mn[q].y; sBox[2][b] = mn[q].z;
                sBox[3][b] = mx[q].x; sBox[4][b] = mx[q].y; sBox[5][b] = mx[q].z;
            }
            __syncthreads();
        }
    }
}

// Level l of the table from level l - 1, one thread per entry (tables of more than kTabLds blocks).
__global__ void k_refit_table_level(int numBlocks, int l, float4 *__restrict__ tab) {
    const int b = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int half = 1 << (l - 1), cnt = numBlocks - (1 << l) + 1;
    if (b >= cnt) return;
    const float4 *src = tab + 2 * (size_t)(l - 1) * numBlocks;
    float4 *dst = tab + 2 * (size_t)l * numBlocks;
    dst[2 * b] = fmin4(src[2 * b], src[2 * (b + half)]);
    dst[2 * b + 1] = fmax4(src[2 * b + 1], src[2 * (b + half) + 1]);
}

// Box of the leaves [lo, hi] that span at least two workgroups.
__device__ __forceinline__ void range_box(const Lbvh &t, int numBlocks, int lo, int hi, float4 &mn, float4 &mx) {
    mn = fmin4(t.suf[2 * (size_t)lo], t.pre[2 * (size_t)hi]);
    mx = fmax4(t.suf[2 * (size_t)lo + 1], t.pre[2 * (size_t)hi + 1]);
    const int bl = lo / kBlk + 1, bh = hi / kBlk - 1;
    if (bl <= bh) {
        const int len = bh - bl + 1;
        const int l = 31 - __clz(len);
        const float4 *lev = t.blockTab + 2 * (size_t)l * numBlocks;
        const int b2 = bh - (1 << l) + 1;
        mn = fmin4(mn, fmin4(lev[2 * bl], lev[2 * b2]));
        mx = fmax4(mx, fmax4(lev[2 * bl + 1], lev[2 * b2 + 1]));
    }
}

// The box of a node that k_refit_local merged (or of a leaf), from what the step keeps: a leaf's
// centre -/+ r, an internal node's two child boxes in its traversal record (the same min / max as
// its merge took).
__device__ __forceinline__ void kept_box(const Lbvh &t, int n, float r, int c, float4 &mn, float4 &mx) {
    if (c < n) {
        const float4 p = t.cms[c];
        mn = make_float4(p.x - r, p.y - r, p.z - r, 0.0f);
        mx = make_float4(p.x + r, p.y + r, p.z + r, 0.0f);
    } else {
        const float4 *q = t.packed + 4 * (size_t)(c - n);
        mn = fmin4(q[0], q[2]);
        mx = fmax4(q[1], q[3]);
    }
}

__device__ __forceinline__ void child_box(const Lbvh &t, int n, int numBlocks, float r, int c, float4 &mn,
                                          float4 &mx) {
    if (c >= n) {
        const int4 m = t.meta[c - n];
        if (range_crosses(m.z, m.w)) {
            range_box(t, numBlocks, m.z, m.w, mn, mx);
            return;
        }
    }
    if (t.bounds) {
        mn = t.bounds[2 * (size_t)c];
        mx = t.bounds[2 * (size_t)c + 1];
    } else {
        kept_box(t, n, r, c, mn, mx);
    }
}

__device__ __forceinline__ void refit_crossing(int n, int numBlocks, float r, const Lbvh &t, int k, int4 m) {
    float4 mn, mx, aMin, aMax, bMin, bMax;
    range_box(t, numBlocks, m.z, m.w, mn, mx);
    child_box(t, n, numBlocks, r, m.x, aMin, aMax);
    child_box(t, n, numBlocks, r, m.y, bMin, bMax);
    if (t.bounds) store_plain(t.bounds, k + n, mn, mx);
    write_packed(t, n, k + n, m.x, m.y, aMin, aMax, bMin, bMax);
}

// Every internal node: those whose range crosses a block boundary are refitted.
__global__ void k_refit_cross(int n, int numBlocks, float r, Lbvh t) {
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= n - 1) return;
    const int4 m = t.meta[k];
    if (!range_crosses(m.z, m.w)) return;
    refit_crossing(n, numBlocks, r, t, k, m);
}

// The crossing nodes that k_karras listed: workgroup g takes build workgroup g's list.
__global__ void k_refit_cross_list(int n, int numBlocks, float r, Lbvh t) {
    const unsigned cnt = t.crossCount[blockIdx.x];
    const int *list = t.crossList + (size_t)blockIdx.x * kKarrasBlock;
    for (unsigned e = threadIdx.x; e < cnt; e += blockDim.x) {
        const int k = list[e];
        refit_crossing(n, numBlocks, r, t, k, t.meta[k]);
    }
}

// Node boxes for callers that read them (the engine's step keeps none): the boxes k_refit_local and
// k_refit_cross merged, restated from the leaf centres and the traversal records.
__global__ void k_bounds_from_records(int n, float r, Lbvh t) {
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= 2 * n - 1) return;
    float4 mn, mx;
    kept_box(t, n, r, c, mn, mx);
    store_plain(t.bounds, c, mn, mx);
}

__device__ __forceinline__ bool overlap(float4 aMin, float4 aMax, float4 bMin, float4 bMax) {
    // checkOverlap(AABB, AABB) (BVHAuxiliary.cuh:139-152); evaluated without short-circuiting, so
    // that a node's boxes are needed (and loaded) together
    return (aMin.x <= bMax.x) & (aMax.x >= bMin.x) & (aMin.y <= bMax.y) & (aMax.y >= bMin.y) &
           (aMin.z <= bMax.z) & (aMax.z >= bMin.z);
}

// collideBVHSoARigidBody (Kernels/RB_BVHKernels.cu:100-244): per sorted leaf, traversal from the
// root (or the leaf's scene root), left child before right, the right child pushed when both
// are entered.  Leaves are the nodes < n.  The stack lives in LDS, one column per thread.
template <bool EMIT>
__global__ void k_dem_lbvh(const int *__restrict__ sIdx, Lbvh t, int n, DevPhys ph, SceneView sv, int sceneInVelW,
                           DemOut o, int *overflow) {
    extern __shared__ int sStack[];
    const int lb = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int i = lb * (int)blockDim.x + (int)threadIdx.x;
    if (i >= n) return;
    const int bd = (int)blockDim.x, tid = (int)threadIdx.x;
    const float4 C = t.cms[i];
    const float4 Vq = t.sortedVel[i];
    const int ra = __float_as_int(C.w);
    const int s = sceneInVelW ? __float_as_int(Vq.w) : 0;
    const float4 qMin = t.bounds[2 * (size_t)i], qMax = t.bounds[2 * (size_t)i + 1];
    const f3 pa = xyz(C), va = xyz(Vq);
    const int orig = sIdx[i];
    int node = t.sceneRoot ? t.sceneRoot[s] : (n >= 2 ? n : -1);
    f3 F = mk3(0.f, 0.f, 0.f);
    int cnt = 0, sp = 0;
    bool ovf = false;   // stack overflow: the traversal is incomplete and nothing is written
    while (node >= n) {
        int L, R;
        bool oL, oR;
        if (t.packed) {
            // one 64-byte node: both children's boxes and ids
            const float4 *q = t.packed + 4 * (size_t)(node - n);
            const float4 a0 = q[0], a1 = q[1], b0 = q[2], b1 = q[3];
            L = __float_as_int(a0.w);
            R = __float_as_int(a1.w);
            oL = L != i && overlap(qMin, qMax, a0, a1);
            oR = R != i && overlap(qMin, qMax, b0, b1);
        } else {
            L = t.left[node];
            R = t.right[node];
            oL = L != i && overlap(qMin, qMax, t.bounds[2 * (size_t)L], t.bounds[2 * (size_t)L + 1]);
            oR = R != i && overlap(qMin, qMax, t.bounds[2 * (size_t)R], t.bounds[2 * (size_t)R + 1]);
        }
        const bool lL = L < n, lR = R < n;
        if (oL && lL) {
            float4 CL = t.cms[L];
            if (candidate(CL, pa, ra, false, ph)) {
                if (EMIT) {
                    if (in_contact(pa, xyz(CL), ph)) {
                        if (cnt < o.maxPer) o.pairs[(size_t)orig * o.maxPer + cnt] = sIdx[L];
                        ++cnt;
                    }
                } else {
                    f3 fj;
                    if (pair_force(pa, xyz(CL), va, xyz(t.sortedVel[L]), ph, fj)) F = add3(F, fj);
                }
            }
        }
        if (oR && lR) {
            float4 CR = t.cms[R];
            if (candidate(CR, pa, ra, false, ph)) {
                if (EMIT) {
                    if (in_contact(pa, xyz(CR), ph)) {
                        if (cnt < o.maxPer) o.pairs[(size_t)orig * o.maxPer + cnt] = sIdx[R];
                        ++cnt;
                    }
                } else {
                    f3 fj;
                    if (pair_force(pa, xyz(CR), va, xyz(t.sortedVel[R]), ph, fj)) F = add3(F, fj);
                }
            }
        }
        const bool tL = oL && !lL, tR = oR && !lR;
        if (!tL && !tR) {
            if (sp == 0) break;
            --sp;
            node = sStack[sp * bd + tid];
        } else {
            node = tL ? L : R;
            if (tL && tR) {
                if (sp == kStackDepth) {
                    atomicOr(overflow, 1);
                    ovf = true;
                    break;
                }
                sStack[sp * bd + tid] = R;
                ++sp;
            }
        }
    }
    if (ovf) return;
    if (EMIT) {
        o.counts[orig] = cnt;
        return;
    }
    SceneGeom g;
    load_geom(sv, s, g);
    if (!((ph.flags & kGrainClamp) && ra < 0)) F = add3(F, wall_force(pa, va, ph, g));
    if (ra >= 0) {
        o.pForce[orig] = f4(F, 0.0f);
        if (o.pTorque) o.pTorque[orig] = f4(cross3(xyz(o.relPos[orig]), F), 0.0f);
    } else {
        const float fs = (ph.flags & kRefUnits) ? 1.0f : ph.dt;
        o.vel[orig] = make_float4(va.x + F.x * fs, va.y + F.y * fs, va.z + F.z * fs, 0.0f);
    }
}

// Packet form of the same traversal: 16 consecutive leaves in Morton order (four packets per wave)
// walk the tree together, with the packet's own node, LDS stack (kWaveStackDepth entries) and depth.
// The packet visits a node when any of its lanes' boxes overlaps it.  A lane handles a leaf only when
// its own box overlaps the leaf's, and every ancestor's box contains the leaf's box, so each lane meets
// exactly the leaves of its own traversal, in the same depth-first order (left before right): the
// per-lane force sums are bit-identical.  A smaller packet's lanes are closer together, so it enters
// fewer nodes that none of its lanes overlaps (at the 2^20-particle scene a 64-lane packet enters 157
// nodes, a 16-lane packet 78, and the wave, which runs as long as its slowest packet, 95).  A visit
// loads the packet's 64-byte node record as one dword per lane and passes the words round the packet
// by DPP row broadcasts.
constexpr int kPacketLanes = 16;
// Word K of each 16-lane row (DPP row_newbcast: every lane of a row receives lane K's value), as a float.
template <int K>
__device__ __forceinline__ float row_word(unsigned w) {
    return __uint_as_float((unsigned)__builtin_amdgcn_update_dpp(0, (int)w, 0x150 + K, 0xf, 0xf, false));
}
// FOREST: scenes' trees under their own roots (scene id in sortedVel.w); a packet whose lanes start at
// different roots walks them one root after another, the lanes of the current root taking part.
template <bool EMIT, bool FOREST>
__global__ __attribute__((amdgpu_waves_per_eu(8))) void k_dem_lbvh_packets(const int *__restrict__ sIdx,
                                                                           const float4 *__restrict__ packed,
                                   const float4 *__restrict__ cms, const float4 *__restrict__ sVel,
                                   const int *__restrict__ sceneRoot, int n,
                                   DevPhys ph, SceneView sv, DemOut o, int *overflow) {
    extern __shared__ int pStack[];
    const int lb = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int i = lb * (int)blockDim.x + (int)threadIdx.x;
    const bool valid = i < n;
    const int ii = valid ? i : n - 1;
    const int lane = (int)(threadIdx.x & 63);
    const int pl = lane & (kPacketLanes - 1);
    const unsigned long long pmask = 0xffffull << (lane & ~(kPacketLanes - 1));
    int *stack = pStack + ((int)threadIdx.x / kPacketLanes) * kWaveStackDepth;
    const float4 C = cms[ii];
    const float4 Vq = sVel[ii];
    const int ra = __float_as_int(C.w);
    // the leaf's box, centre -/+ r as the refit forms it
    const float4 qMin = make_float4(C.x - ph.r, C.y - ph.r, C.z - ph.r, 0.0f);
    const float4 qMax = make_float4(C.x + ph.r, C.y + ph.r, C.z + ph.r, 0.0f);
    const f3 pa = xyz(C), va = xyz(Vq);
    const int orig = sIdx[ii];
    f3 F = mk3(0.f, 0.f, 0.f);
    int cnt = 0;
    bool ovf = false;   // this lane's packet hit the stack bound: nothing is written for it
    const int s = FOREST ? __float_as_int(Vq.w) : 0;
    const int myRoot = FOREST ? sceneRoot[s] : (n >= 2 ? n : -1);
    // the packet's lanes whose root has not been walked yet; node < 0: no walk under way
    unsigned long long pending = __ballot(valid && myRoot >= n) & pmask;
    int root = -1, node = -1, sp = 0;
    if (!FOREST && pending != 0ull) {
        root = node = n;
        pending = 0ull;
    }
    const unsigned *words = reinterpret_cast<const unsigned *>(packed);
    // The next node's record is requested as soon as the next node is known (from the overlap tests and
    // the stack), before this node's leaf children are tested, so that its load overlaps theirs.
    unsigned wn = node >= 0 ? words[16 * (size_t)(node - n) + pl] : 0u;
#ifdef ARSIM_TRAVERSAL_STATS
    const bool packetUsed = pending != 0ull || node >= 0;
    unsigned long long stVisits = 0, stLeafTests = 0, stInterest = 0, stContacts = 0, stIters = 0;
    int stMaxSp = 0;
#endif
    while (true) {
        bool fresh = false;   // a walk that starts here: its root's record is not requested yet
        if (FOREST) {
            // a packet without a walk starts the next pending root (all lanes take part in the exchange)
            const bool needRoot = node < 0 && pending != 0ull;
            const int first = needRoot ? __builtin_ctzll(pending) : lane;
            const int cand = __shfl(myRoot, first, 64);
            if (needRoot) {
                root = cand;
                node = cand;
                sp = 0;
                fresh = true;
            }
            pending &= ~(__ballot(needRoot && valid && myRoot == cand) & pmask);
        }
        if (__ballot(node >= 0) == 0ull) break;
        const bool live = node >= 0;
#ifdef ARSIM_TRAVERSAL_STATS
        ++stIters;
        if (live) ++stVisits;
#endif
        const unsigned w = fresh ? words[16 * (size_t)(node - n) + pl] : wn;
        const float4 a0 = make_float4(row_word<0>(w), row_word<1>(w), row_word<2>(w), 0.f);
        const float4 a1 = make_float4(row_word<4>(w), row_word<5>(w), row_word<6>(w), 0.f);
        const float4 b0 = make_float4(row_word<8>(w), row_word<9>(w), row_word<10>(w), 0.f);
        const float4 b1 = make_float4(row_word<12>(w), row_word<13>(w), row_word<14>(w), 0.f);
        const int L = __float_as_int(row_word<3>(w)), R = __float_as_int(row_word<7>(w));
        const bool act = live && valid && myRoot == root;
        const bool oL = act & (L != i) & overlap(qMin, qMax, a0, a1);
        const bool oR = act & (R != i) & overlap(qMin, qMax, b0, b1);
        const bool tL = L >= n && (__ballot(oL) & pmask) != 0ull;
        const bool tR = R >= n && (__ballot(oR) & pmask) != 0ull;
#ifdef ARSIM_TRAVERSAL_STATS
        stLeafTests += (L < n && oL ? 1 : 0) + (R < n && oR ? 1 : 0);
        if (L < n && oL && in_contact(pa, xyz(cms[L]), ph)) ++stContacts;
        if (R < n && oR && in_contact(pa, xyz(cms[R]), ph)) ++stContacts;
        if (L >= n) stInterest += (unsigned long long)__popcll(__ballot(oL) & pmask);
        if (R >= n) stInterest += (unsigned long long)__popcll(__ballot(oR) & pmask);
        if (live && tL && tR) stMaxSp = sp + 1 > stMaxSp ? sp + 1 : stMaxSp;
#endif
        // next node: the left child entered (the right one pushed when both are), else the right one,
        // else the top of the stack; selects rather than branches (every value is uniform in a packet)
        const bool both = tL && tR, full = both && sp == kWaveStackDepth;
        const int top = stack[sp > 0 ? sp - 1 : 0];
        if (live && full && act) {
            atomicOr(overflow, 1);
            ovf = true;
        }
        if (live && both && !full && pl == 0) stack[sp] = R;
        const int next = full ? -1 : (tL ? L : (tR ? R : (sp > 0 ? top : -1)));
        const int dsp = (both && !full) ? 1 : ((!tL && !tR && sp > 0) ? -1 : 0);
        node = live ? next : node;
        sp = live ? sp + dsp : sp;
        wn = node >= 0 ? words[16 * (size_t)(node - n) + pl] : 0u;
        // both leaf children's centres requested before either is tested (one round trip, not two)
        const bool hL = L < n && oL, hR = R < n && oR;
        float4 CL, CR;
        if (hL) CL = cms[L];
        if (hR) CR = cms[R];
        if (hL) {
            if (candidate(CL, pa, ra, false, ph)) {
                if (EMIT) {
                    if (in_contact(pa, xyz(CL), ph)) {
                        if (cnt < o.maxPer) o.pairs[(size_t)orig * o.maxPer + cnt] = sIdx[L];
                        ++cnt;
                    }
                } else {
                    f3 fj;
                    if (pair_force(pa, xyz(CL), va, xyz(sVel[L]), ph, fj)) F = add3(F, fj);
                }
            }
        }
        if (hR) {
            if (candidate(CR, pa, ra, false, ph)) {
                if (EMIT) {
                    if (in_contact(pa, xyz(CR), ph)) {
                        if (cnt < o.maxPer) o.pairs[(size_t)orig * o.maxPer + cnt] = sIdx[R];
                        ++cnt;
                    }
                } else {
                    f3 fj;
                    if (pair_force(pa, xyz(CR), va, xyz(sVel[R]), ph, fj)) F = add3(F, fj);
                }
            }
        }
    }
#ifdef ARSIM_TRAVERSAL_STATS
    {
        // per-lane leaf tests and contacts summed over the wave; the packet counts from each packet's
        // first lane, the wave counts from lane 0 (stInterest and stVisits are packet-uniform)
        unsigned long long lt = stLeafTests, ct = stContacts;
        for (int off = 32; off >= 1; off >>= 1) {
            lt += __shfl_xor(lt, off, 64);
            ct += __shfl_xor(ct, off, 64);
        }
        const unsigned long long lanes = (unsigned long long)__popcll(__ballot(valid));
        const unsigned long long done = (unsigned long long)__popcll(__ballot(valid && !ovf));
        if (pl == 0 && packetUsed) {
            atomicAdd(&g_travStats[1], stVisits);
            atomicAdd(&g_travStats[4], stInterest);
            atomicMax(&g_travStats[5], (unsigned long long)stMaxSp);
            atomicAdd(&g_travStats[8], 1ull);
        }
        if (lane == 0) {
            atomicAdd(&g_travStats[0], 1ull);
            atomicAdd(&g_travStats[2], lanes);
            atomicAdd(&g_travStats[3], lt);
            atomicAdd(&g_travStats[6], done);
            atomicAdd(&g_travStats[7], ct);
            atomicAdd(&g_travStats[9], stIters);
        }
    }
#endif
    if (!valid || ovf) return;
    if (EMIT) {
        o.counts[orig] = cnt;
        return;
    }
    SceneGeom g;
    if (FOREST) g = sv.table[s];
    else g = sv.single;
    if (!((ph.flags & kGrainClamp) && ra < 0)) F = add3(F, wall_force(pa, va, ph, g));
    if (ra >= 0) {
        o.pForce[orig] = f4(F, 0.0f);
        if (o.pTorque) o.pTorque[orig] = f4(cross3(xyz(o.relPos[orig]), F), 0.0f);
    } else {
        const float fs = (ph.flags & kRefUnits) ? 1.0f : ph.dt;
        o.vel[orig] = make_float4(va.x + F.x * fs, va.y + F.y * fs, va.z + F.z * fs, 0.0f);
    }
}

}  // namespace

#ifdef ARSIM_TRAVERSAL_STATS
// Counting build only: copies the packet traversal's totals (10 counters, see g_travStats) and, when
// reset is non-zero, clears them.
extern "C" int arsim_debug_traversal_stats(unsigned long long *out, int reset) {
    if (out && hipMemcpyFromSymbol(out, HIP_SYMBOL(g_travStats), sizeof(g_travStats)) != hipSuccess) return 3;
    if (reset) {
        const unsigned long long zero[kTravStats] = {};
        if (hipMemcpyToSymbol(HIP_SYMBOL(g_travStats), zero, sizeof(zero)) != hipSuccess) return 3;
    }
    return hipDeviceSynchronize() == hipSuccess ? 0 : 3;
}
#endif

int lbvh_refit_block() { return kBlk; }
int lbvh_refit_levels(int n) {
    const int numBlocks = (n + kBlk - 1) / kBlk;
    int l = 1;
    while ((1 << l) <= numBlocks) ++l;
    return l;
}

void launch_lbvh_build(const void *sortedKeys, int keyBytes, int n, int sceneShift, const int *sceneStart,
                       const int *sceneCount, int numScenes, const Lbvh &t, hipStream_t st) {
    (void)numScenes;
    if (n <= 0) return;
    const int block = kKarrasBlock;
    int threads = n - 1 > 1 ? n - 1 : 1;
    dim3 grid((threads + block - 1) / block);
    if (keyBytes == 8)
        hipLaunchKernelGGL(k_karras<uint64_t>, grid, dim3(block), 0, st, (const uint64_t *)sortedKeys, n, sceneShift,
                           sceneStart, sceneCount, t);
    else
        hipLaunchKernelGGL(k_karras<unsigned>, grid, dim3(block), 0, st, (const unsigned *)sortedKeys, n, sceneShift,
                           sceneStart, sceneCount, t);
}

void launch_lbvh_refit(const int *sortedIdx, const float4 *pos, const float4 *vel, const int *rb,
                       const uint16_t *scene, const float4 *rec, float radius, int n, const Lbvh &t, hipStream_t st) {
    if (n <= 0) return;
    if (t.meta && t.pre && t.suf && t.blockTab) {
        const int numBlocks = (n + kBlk - 1) / kBlk;
        hipLaunchKernelGGL(k_refit_local, dim3(numBlocks), dim3(kBlk), 0, st, sortedIdx, pos, vel, rb, scene, rec,
                           radius, n, t);
        if (numBlocks > 1) {
            const int levels = lbvh_refit_levels(n);
            if (numBlocks <= kTabLds) {
                hipLaunchKernelGGL(k_refit_table, dim3(1), dim3(1024), 0, st, numBlocks, levels, t.blockTab);
            } else {
                // one launch per level: a single workgroup walking thousands of entries per level waits a
                // global-memory round trip per handful of them
                for (int l = 1; l < levels; ++l) {
                    const int cnt = numBlocks - (1 << l) + 1;
                    hipLaunchKernelGGL(k_refit_table_level, dim3((cnt + 255) / 256), dim3(256), 0, st, numBlocks, l,
                                       t.blockTab);
                }
            }
            if (t.crossList) {
                const int blocks = (n - 1 + kKarrasBlock - 1) / kKarrasBlock;   // the build's workgroups
                hipLaunchKernelGGL(k_refit_cross_list, dim3(blocks), dim3(64), 0, st, n, numBlocks, radius, t);
            } else {
                hipLaunchKernelGGL(k_refit_cross, dim3((n - 1 + 255) / 256), dim3(256), 0, st, n, numBlocks, radius,
                                   t);
            }
        }
        return;
    }
    const int block = 256;
    hipLaunchKernelGGL(k_leaf_refit, dim3((n + block - 1) / block), dim3(block), 0, st, sortedIdx, pos, vel, rb,
                       scene, rec, radius, n, t);
}

void launch_lbvh_bounds(const Lbvh &t, int n, float radius, hipStream_t st) {
    if (n <= 0 || !t.bounds || !t.packed) return;
    const int nodes = 2 * n - 1;
    hipLaunchKernelGGL(k_bounds_from_records, dim3((nodes + 255) / 256), dim3(256), 0, st, n, radius, t);
}

void launch_dem_lbvh(const int *sortedIdx, const Lbvh &t, int n, const DevPhys &ph, const SceneView &sv,
                     bool sceneInVelW, const DemOut &o, int block, hipStream_t st, int *overflow) {
    if (n <= 0) return;
    if (block <= 0) block = 128;
    dim3 grid((n + block - 1) / block);
    if (t.packed) {
        // packets of 16 lanes, one scene or a forest (whose scene ids ride in sortedVel.w)
        block = (block + 63) & ~63;
        grid = dim3((n + block - 1) / block);
        const size_t pl = (size_t)kWaveStackDepth * (block / kPacketLanes) * sizeof(int);
        auto run = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, grid, dim3(block), pl, st, sortedIdx, t.packed, t.cms, t.sortedVel, t.sceneRoot, n,
                               ph, sv, o, overflow);
        };
        if (t.sceneRoot) {
            if (o.pairs) run(k_dem_lbvh_packets<true, true>);
            else run(k_dem_lbvh_packets<false, true>);
        } else {
            if (o.pairs) run(k_dem_lbvh_packets<true, false>);
            else run(k_dem_lbvh_packets<false, false>);
        }
        return;
    }
    size_t lds = (size_t)kStackDepth * block * sizeof(int);
    if (o.pairs)
        hipLaunchKernelGGL(k_dem_lbvh<true>, grid, dim3(block), lds, st, sortedIdx, t, n, ph, sv, sceneInVelW ? 1 : 0,
                           o, overflow);
    else
        hipLaunchKernelGGL(k_dem_lbvh<false>, grid, dim3(block), lds, st, sortedIdx, t, n, ph, sv,
                           sceneInVelW ? 1 : 0, o, overflow);
}

}  // namespace arsim
