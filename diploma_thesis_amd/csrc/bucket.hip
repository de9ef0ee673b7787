// bucket.hip -- the grid key's sort finished per bucket, fused with the reorder and the cell table.
//
// The pose kernel leaves every tile of particles ranked by bucket (tile-local, stable): a bucket is
// the range [c0, c0 + 2^shift) of one scene's cells.  The keys of a bucket are then the runs of that
// bucket in its scene's tiles taken in tile order, and its place in the sorted order is its scene's
// first particle plus the counts of the scene's earlier buckets (scenes' cell ranges are packed in
// scene order).  One or two workgroups per bucket then
//   1. sort the bucket by the key's offset from c0, stably: a bucket that fits in LDS is sorted there
//      as packed words (offset << 16 | position in the bucket), so one 32-bit word carries key and
//      order; a larger bucket is sorted in global memory by the same stable passes (slow, but correct);
//   2. write the sorted keys / indices and the gathered sorted copies (the reorder), a share each;
//   3. write their part of the cell table: table[c] = first sorted index with key >= c.
// The result is the same as the stable LSD sort followed by the reorder, for any distribution.
#include <algorithm>
#include <cmath>

#include "csr_fill.hpp"
#include "kernels.hpp"

namespace arsim {

namespace {

constexpr int kBkThreads = 1024;
constexpr int kBkWaves = kBkThreads / 64;
// Bucket size sorted in LDS: two buffers of packed words plus the 16 KiB cursor table, the wave sums
// and the tile map stay within 80 KiB, half of a gfx950 CU's 160 KiB, so two workgroups share a CU.
constexpr int kBkCap = 7776;
constexpr int kBkMaxTiles = 384;  // tiles of one scene, which a bucket gathers from (2^20 keys: 256)
constexpr int kBkLdsBytes = 2 * kBkCap * 4 + kBkWaves * 256 * 4 + kBkWaves * 4 + 2 * kBkMaxTiles * 4;
static_assert(kBkLdsBytes <= 80 * 1024, "two bucket workgroups must fit in one CU's LDS");

// Lanes of the wave whose digit (bits <= 8 wide) equals this lane's, among lanes with ok set.  Per
// bit: the ballot of lanes with the bit set, kept as is by lanes with the bit set and complemented
// by the others (xnor with the lane's bit spread to a mask).
__device__ __forceinline__ uint64_t match_digit(unsigned d, int bits, bool ok) {
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        if (b >= bits) break;
        const unsigned t = 0u - ((d >> b) & 1u);
        const uint64_t m = __ballot(t != 0u);
        const unsigned lo = ~((unsigned)m ^ t), hi = ~((unsigned)(m >> 32) ^ t);
        peers &= ((uint64_t)hi << 32) | lo;
    }
    return peers;
}

__device__ __forceinline__ unsigned block_excl_scan(unsigned v, unsigned *sWave) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) sWave[wave] = x;
    __syncthreads();
    unsigned before = 0u;
    for (int w = 0; w < wave; ++w) before += sWave[w];
    __syncthreads();
    return before + x - v;
}

__device__ __forceinline__ unsigned block_sum(unsigned v, unsigned *sWave) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) sWave[wave] = v;
    __syncthreads();
    unsigned s = 0u;
    for (int w = 0; w < kBkWaves; ++w) s += sWave[w];
    __syncthreads();
    return s;
}

// Item accessors of the stable pass: packed words in LDS, or (key, value) pairs in global memory.
struct LdsItems {
    unsigned *a;
    __device__ unsigned digit(int i, int shift, unsigned mask) const { return (a[i] >> shift) & mask; }
    __device__ void move_to(const LdsItems &dst, int pos, int i) const { dst.a[pos] = a[i]; }
};
struct GlobalItems {
    unsigned *k, *v;
    __device__ unsigned digit(int i, int shift, unsigned mask) const { return (k[i] >> shift) & mask; }
    __device__ void move_to(const GlobalItems &dst, int pos, int i) const {
        dst.k[pos] = k[i];
        dst.v[pos] = v[i];
    }
};

// One stable counting-sort pass of cnt items by the digit at `shift` (bits wide), src -> dst.
// Wave w owns a contiguous segment of the items: it counts its digits with LDS atomics, then walks
// the segment in order 64 at a time, a lane's rank among equal digits of its group coming from
// ballots, so the order of equal digits is kept.  (The ballots are the pass's VALU cost: a wave64
// instruction takes four cycles, so the count phase does without them.)
template <class IT>
__device__ void stable_pass(const IT &src, const IT &dst, int cnt, int shift, int bits, unsigned (*sCur)[256],
                            unsigned *sWave) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned mask = (1u << bits) - 1u;
    const int radix = 1 << bits;
    const int seg = ((cnt + kBkThreads - 1) / kBkThreads) * 64;
    const int s0 = min(wave * seg, cnt), s1 = min(s0 + seg, cnt);
    const uint64_t lower = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    for (int k = threadIdx.x; k < kBkWaves * 256; k += kBkThreads) (&sCur[0][0])[k] = 0u;
    __syncthreads();
    // the counts need no order: an LDS atomic per item
    for (int i = s0 + lane; i < s1; i += 64) atomicAdd(&sCur[wave][src.digit(i, shift, mask)], 1u);
    __syncthreads();
    // cursors: digits in order, and within a digit the waves in order
    {
        const int t = threadIdx.x;
        unsigned c[kBkWaves], tot = 0u;
        for (int w = 0; w < kBkWaves; ++w) {
            c[w] = t < radix ? sCur[w][t] : 0u;
            tot += c[w];
        }
        const unsigned excl = block_excl_scan(tot, sWave);
        if (t < radix) {
            unsigned run = excl;
            for (int w = 0; w < kBkWaves; ++w) {
                sCur[w][t] = run;
                run += c[w];
            }
        }
    }
    __syncthreads();
    for (int g = s0; g < s1; g += 64) {
        const int i = g + lane;
        const bool ok = i < s1;
        const unsigned d = ok ? src.digit(i, shift, mask) : 0u;
        const uint64_t peers = match_digit(d, bits, ok);
        const unsigned prior = (unsigned)__popcll(peers & lower);
        const unsigned base = ok ? sCur[wave][d] : 0u;
        if (ok) src.move_to(dst, (int)(base + prior), i);
        if (ok && prior == 0u) sCur[wave][d] = base + (unsigned)__popcll(peers);
    }
    __syncthreads();
}

// Fills table[a, b) with v using the whole workgroup (16-byte stores for the aligned middle).
__device__ __forceinline__ void block_fill(unsigned *table, unsigned long long a, unsigned long long b, unsigned v) {
    if (a >= b) return;
    const unsigned long long a4 = (a + 3ull) & ~3ull, b4 = b & ~3ull;
    if (a4 >= b4) {
        for (unsigned long long k = a + threadIdx.x; k < b; k += kBkThreads) table[k] = v;
        return;
    }
    if (threadIdx.x < a4 - a) table[a + threadIdx.x] = v;
    if (threadIdx.x < b - b4) table[b4 + threadIdx.x] = v;
    const uint4 v4 = make_uint4(v, v, v, v);
    for (unsigned long long k = a4 + 4ull * threadIdx.x; k < b4; k += 4ull * kBkThreads)
        *reinterpret_cast<uint4 *>(table + k) = v4;
}

// POSBITS: bits of a bucket position in the packed LDS words (the offset from c0 takes the other 32 -
// POSBITS; a bucket sorted in LDS holds at most 2^POSBITS keys).  REORDER: also the gathered sorted
// copies and the cell table (grid keys); without it only the sorted keys and indices (Morton keys,
// whose sorted-copy arrays are then only the global fallback's second buffer).
// Eight waves per SIMD (two workgroups per CU): registers are held to that occupancy.
template <int POSBITS, bool REORDER>
__global__ __launch_bounds__(kBkThreads) __attribute__((amdgpu_waves_per_eu(8))) void k_bucket_sort(
    TileKeyOut to, int split, unsigned *__restrict__ kOut, unsigned *__restrict__ vOut,
    const float4 *__restrict__ rec, float4 *__restrict__ sPos, float4 *__restrict__ sVel,
    unsigned *__restrict__ table, unsigned numCells) {
    __shared__ unsigned sA[kBkCap], sB[kBkCap];
    __shared__ unsigned sCur[kBkWaves][256];
    __shared__ unsigned sWave[kBkWaves];
    __shared__ unsigned sOff[kBkMaxTiles], sSrc[kBkMaxTiles];   // tile map: bucket offset and source slot
    const unsigned *__restrict__ tKeys = to.keys;
    const unsigned *__restrict__ tVals = to.vals;
    const int t = threadIdx.x;
    // consecutive buckets (neighbouring cells, and one scene's particles) run on one XCD
    const int blk = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int g = blk / split, part = blk % split;
    const BucketScene bs = to.scenes[to.bucketScene[g]];
    const int j = g - bs.bucketBase;
    const int lowBits = bs.shift;
    const unsigned *hist = to.hist + bs.bucketBase;
    // the bucket counts before this one and the tile map's words are independent: both loads in flight
    const int numTiles = bs.tiles;
    const unsigned hPrev = t < j ? hist[t] : 0u;
    const unsigned tileWord = t < numTiles ? to.info[(size_t)bs.infoBase + (size_t)j * numTiles + t] : 0u;
    const unsigned start = bs.pStart + block_sum(hPrev, sWave);
    const int cnt = (int)hist[j];
    constexpr int kCap = kBkCap < (1 << POSBITS) ? kBkCap : (1 << POSBITS);
    constexpr unsigned kPosMask = (1u << POSBITS) - 1u;
    const bool inLds = cnt <= kCap;
    // a bucket too large for LDS is sorted in global memory by its first workgroup alone
    if (!inLds && part != 0) return;
    const int nparts = inLds ? split : 1;
    const int i0 = (int)((long long)cnt * part / nparts), i1 = (int)((long long)cnt * (part + 1) / nparts);
    // the bucket's cells [c0, c1): table[c] = start + (number of bucket keys below c); the last bucket
    // also writes the table's closing entry, table[numCells]
    const unsigned long long c0 = (unsigned long long)bs.cellBase + ((unsigned long long)j << lowBits);
    const unsigned long long sceneEnd = (unsigned long long)bs.cellBase + bs.cells;
    const unsigned long long c1 = g == to.numBuckets - 1 ? sceneEnd + 1ull
                                  : c0 + (1ull << lowBits) < sceneEnd ? c0 + (1ull << lowBits) : sceneEnd;
    if (cnt == 0) {
        const unsigned long long n = c1 - c0;
        if (REORDER) block_fill(table, c0 + n * part / split, c0 + n * (part + 1) / split, start);
        return;
    }
    // The bucket's keys are the runs of bucket j in the scene's tiles, in tile order: tile k's run
    // starts at bucket position sOff[k] and at slot sSrc[k] of the tile-ordered arrays.
    {
        const unsigned c = tileWord >> 16, f = tileWord & 0xffffu;
        const unsigned off = block_excl_scan(c, sWave);
        if (t < numTiles) {
            sOff[t] = off;
            sSrc[t] = bs.pStart + (unsigned)t * (unsigned)to.tileKeys + f;
        }
        __syncthreads();
    }
    // slot of bucket position i in the tile-ordered arrays: the last tile whose run starts at or before i
    auto slot = [&](unsigned i) -> unsigned {
        int lo = 0, hi = numTiles;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (sOff[mid] <= i) lo = mid;
            else hi = mid;
        }
        return sSrc[lo] + (i - sOff[lo]);
    };
    // low(i) = offset from c0 of the i-th sorted key of the bucket
    auto fill_table = [&](auto low) {
        // head: cells up to the first key; tail: cells past the last key
        const unsigned kFirst = low(0), kLast = low(cnt - 1);
        if (part == 0) block_fill(table, c0, c0 + kFirst + 1ull < c1 ? c0 + kFirst + 1ull : c1, start);
        if (part == nparts - 1) block_fill(table, c0 + kLast + 1ull, c1, start + (unsigned)cnt);
        // between keys: the cells (k[i-1], k[i]] start at sorted position i (a lane per key, waves help
        // with long runs); this workgroup's share of the positions
        const int rounds = (i1 - i0 + kBkThreads - 1) / kBkThreads;
        for (int r = 0; r < rounds; ++r) {
            const int i = i0 + r * kBkThreads + t;
            const bool ok = i > 0 && i < i1;
            const unsigned kp = ok ? low(i - 1) : 0u, ki = ok ? low(i) : 0u;
            const bool has = ok && ki != kp;
            const unsigned a = (unsigned)c0 + kp + 1u, b = (unsigned)c0 + ki + 1u;
            wave_fill(table, has, a, b, start + (unsigned)i);
        }
    };
    const unsigned base = (unsigned)c0;
    if (inLds) {
        // every slot found first (LDS only), then every key and value load in flight at once; the
        // values wait in registers while the keys are sorted, then go to the buffer the sort left free,
        // so that the gathers below find a sorted key's value in LDS rather than in global memory
        constexpr int kPer = (kCap + kBkThreads - 1) / kBkThreads;
        unsigned sl[kPer], kv[kPer], vv[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int i = t + q * kBkThreads;
            sl[q] = i < cnt ? slot((unsigned)i) : 0u;
        }
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int i = t + q * kBkThreads;
            kv[q] = i < cnt ? tKeys[sl[q]] : 0u;
            vv[q] = i < cnt ? tVals[sl[q]] : 0u;
        }
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int i = t + q * kBkThreads;
            if (i < cnt) sA[i] = ((kv[q] - base) << POSBITS) | (unsigned)i;
        }
        __syncthreads();
        unsigned *src = sA, *dst = sB;
        for (int sh = 0; sh < lowBits; sh += 8) {
            stable_pass(LdsItems{src}, LdsItems{dst}, cnt, POSBITS + sh, min(8, lowBits - sh), sCur, sWave);
            unsigned *x = src;
            src = dst;
            dst = x;
        }
        const unsigned *sorted = src;   // packed words in sorted order
        unsigned *valAt = dst;          // valAt[i]: the value of the bucket's i-th key in gather order
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int i = t + q * kBkThreads;
            if (i < cnt) valAt[i] = vv[q];
        }
        __syncthreads();
        // the gathers of the first two rounds of positions are issued before the table is filled, so
        // that their latency overlaps the table's stores; a lane past the part's end loads sorted
        // position 0 (the bucket is not empty) and stores nothing, so the loads need no branch
        const int ia = i0 + t, ib = ia + kBkThreads;
        const bool ha = ia < i1, hb = ib < i1;
        const unsigned wa = sorted[ha ? ia : 0], wb = sorted[hb ? ib : 0];
        const unsigned va = valAt[wa & kPosMask], vb = valAt[wb & kPosMask];
        float4 pa, qa, pb, qb;
        if (REORDER) {
            pa = rec[2 * (size_t)va];
            qa = rec[2 * (size_t)va + 1];
            pb = rec[2 * (size_t)vb];
            qb = rec[2 * (size_t)vb + 1];
            fill_table([&](int i) -> unsigned { return sorted[i] >> POSBITS; });
        }
        if (ha) {
            kOut[start + ia] = base + (wa >> POSBITS);
            vOut[start + ia] = va;
            if (REORDER) {
                sPos[start + ia] = pa;
                sVel[start + ia] = qa;
            }
        }
        if (hb) {
            kOut[start + ib] = base + (wb >> POSBITS);
            vOut[start + ib] = vb;
            if (REORDER) {
                sPos[start + ib] = pb;
                sVel[start + ib] = qb;
            }
        }
        // further rounds in pairs: both positions' gathers are issued before either's stores (a load
        // after a store also waits for the store); a position past the end reloads the first one
        for (int i = ib + kBkThreads; i < i1; i += 2 * kBkThreads) {
            const int i2 = i + kBkThreads;
            const bool h2 = i2 < i1;
            const unsigned w = sorted[i], w2 = sorted[h2 ? i2 : i];
            const unsigned val = valAt[w & kPosMask], val2 = valAt[w2 & kPosMask];
            float4 P, Q, P2, Q2;
            if (REORDER) {
                P = rec[2 * (size_t)val];
                Q = rec[2 * (size_t)val + 1];
                P2 = rec[2 * (size_t)val2];
                Q2 = rec[2 * (size_t)val2 + 1];
            }
            kOut[start + i] = base + (w >> POSBITS);
            vOut[start + i] = val;
            if (REORDER) {
                sPos[start + i] = P;
                sVel[start + i] = Q;
            }
            if (h2) {
                kOut[start + i2] = base + (w2 >> POSBITS);
                vOut[start + i2] = val2;
                if (REORDER) {
                    sPos[start + i2] = P2;
                    sVel[start + i2] = Q2;
                }
            }
        }
    } else {
        // a bucket too large for LDS, sorted in global memory by its first workgroup alone: its runs are
        // gathered into (kOut, vOut) at [start, start + cnt) as offsets from c0, and the sorted-copy
        // arrays' range (16 B per key) serves as the second buffer of the passes until the copies are
        // written; the keys get c0 back at the end
        for (int i = t; i < cnt; i += kBkThreads) {
            const unsigned src = slot((unsigned)i);
            kOut[start + i] = tKeys[src] - base;
            vOut[start + i] = tVals[src];
        }
        __syncthreads();
        GlobalItems a{kOut + start, vOut + start}, b{reinterpret_cast<unsigned *>(sPos + start),
                                                     reinterpret_cast<unsigned *>(sVel + start)};
        int passes = 0;
        for (int sh = 0; sh < lowBits; sh += 8, ++passes) {
            if ((passes & 1) == 0) stable_pass(a, b, cnt, sh, min(8, lowBits - sh), sCur, sWave);
            else stable_pass(b, a, cnt, sh, min(8, lowBits - sh), sCur, sWave);
        }
        if (passes & 1) {
            for (int i = t; i < cnt; i += kBkThreads) {
                a.k[i] = b.k[i];
                a.v[i] = b.v[i];
            }
            __syncthreads();
        }
        if (REORDER) {
            for (int i = t; i < cnt; i += kBkThreads) {
                const unsigned val = vOut[start + i];
                sPos[start + i] = rec[2 * (size_t)val];
                sVel[start + i] = rec[2 * (size_t)val + 1];
            }
            __syncthreads();
            fill_table([&](int i) -> unsigned { return kOut[start + i]; });
            __syncthreads();
        }
        for (int i = t; i < cnt; i += kBkThreads) kOut[start + i] += base;
    }
}

// One thread per sorted position.  A position alone in its run of equal top bits keeps its place;
// the members of a longer run rank themselves among the run (O(run length) each).
__global__ void k_morton63_ties(const unsigned *__restrict__ k30, const unsigned *__restrict__ vIn,
                                const uint64_t *__restrict__ code63, int n, uint64_t *__restrict__ kOut,
                                unsigned *__restrict__ vOut) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const unsigned c = k30[i];
    const unsigned v = vIn[i];
    const uint64_t key = code63[v];
    const bool prevSame = i > 0 && k30[i - 1] == c, nextSame = i + 1 < n && k30[i + 1] == c;
    if (!prevSame && !nextSame) {
        kOut[i] = key;
        vOut[i] = v;
        return;
    }
    int a = i;
    while (a > 0 && k30[a - 1] == c) --a;
    int e = i + 1;
    while (e < n && k30[e] == c) ++e;
    int rank = 0;
    for (int j = a; j < e; ++j) {
        if (j == i) continue;
        const uint64_t kj = code63[vIn[j]];
        rank += (kj < key || (kj == key && j < i)) ? 1 : 0;
    }
    kOut[a + rank] = key;
    vOut[a + rank] = v;
}

}  // namespace

void launch_morton63_ties(const unsigned *k30, const unsigned *vIn, const uint64_t *code63, int n, uint64_t *kOut,
                          unsigned *vOut, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_morton63_ties, dim3((n + 255) / 256), dim3(256), 0, st, k30, vIn, code63, n, kOut, vOut);
}

bool plan_bucket_sort(const std::vector<BucketSceneShape> &shapes, BucketPlan *out, int maxShift, int capKeys) {
    const size_t S = shapes.size();
    long long nTot = 0, tiles4k = 0;
    for (const auto &sh : shapes) {
        nTot += sh.particles;
        tiles4k += (sh.particles + kTileKeys - 1) / kTileKeys;
    }
    // tiles of 4096 particles, or of 1024 for at most 65,536 particles (too few tiles otherwise) and
    // where more than a fifth of the larger tiles' slots would lie past scene ends
    const int tk = nTot <= 65536 || (double)tiles4k * kTileKeys > 1.25 * (double)nTot ? 1024 : kTileKeys;
    // keys per bucket on average: a 512th of all keys (512 buckets fill the GPU with one workgroup
    // each; 256 buckets with two workgroups each, which sort every bucket twice, and 1,024 buckets
    // were slower at 2^20 keys), at least 64 and at most 3/4 of what a workgroup sorts in LDS
    const double target = std::min(0.75 * capKeys, std::max(64.0, (double)nTot / 512.0));
    // buckets per scene: the tile rank's digit, up to 10 bits with tiles of 4096 (1024 threads)
    const int maxBuckets = tk == kTileKeys ? 1024 : 256;
    BucketPlan p;
    p.tileKeys = tk;
    int maxSceneBuckets = 1;
    p.scenes.resize(S);
    long long pBase = 0, cBase = 0, info = 0;
    int bBase = 0, tBase = 0;
    for (size_t s = 0; s < S; ++s) {
        const BucketSceneShape &sc = shapes[s];
        const long long n = sc.particles;
        const long long c = (long long)sc.grid[0] * sc.grid[1] * sc.grid[2];
        if (c <= 0 || n < 0) return false;
        const long long tiles = (n + tk - 1) / tk;
        if (tiles > kBkMaxTiles) return false;
        double uneven[3];
        for (int a = 0; a < 3; ++a) {
            const double r = sc.extent[a] / (double)sc.grid[a];
            uneven[a] = r > 0.0 ? std::ceil(r) / r : 1.0;
        }
        const bool measured = (int)sc.peak.size() > maxShift;
        // the widest shift (fewest buckets, at most maxBuckets, offsets of at most maxShift bits) whose
        // buckets average at most `target` keys and whose fullest bucket holds at most 4/5 of the LDS capacity
        // (measured, leaving room for the particles to move) or 3/4 (estimated); else the narrowest
        // one if its buckets still average within that
        int shift = -1, narrowest = -1;
        for (int b = maxShift; b >= 0; --b) {
            const long long nb = (c + (1ll << b) - 1) >> b;
            if (nb > maxBuckets) break;
            const long long span = 1ll << b;
            const double avg = (double)n * (double)span / (double)c;
            double peak = avg * uneven[2];
            if (span < (long long)sc.grid[0] * sc.grid[1]) peak *= uneven[1];
            if (span < (long long)sc.grid[0]) peak *= uneven[0];
            const bool fits = measured ? sc.peak[b] <= 0.8 * capKeys : peak <= 0.75 * capKeys;
            if (avg <= 0.75 * capKeys) narrowest = b;
            if (avg <= target && fits) {
                shift = b;
                break;
            }
        }
        if (shift < 0) shift = narrowest;
        if (sc.shift >= 0 && sc.shift <= maxShift && ((c + (1ll << sc.shift) - 1) >> sc.shift) <= maxBuckets)
            shift = sc.shift;
        if (shift < 0) return false;
        BucketScene &b = p.scenes[s];
        b.pStart = (unsigned)pBase;
        b.n = (unsigned)n;
        b.cellBase = (unsigned)cBase;
        b.cells = (unsigned)c;
        b.shift = shift;
        b.buckets = (int)((c + (1ll << shift) - 1) >> shift);
        maxSceneBuckets = std::max(maxSceneBuckets, b.buckets);
        b.bucketBase = bBase;
        b.tileBase = tBase;
        b.tiles = (int)tiles;
        b.infoBase = (int)info;
        for (int k = 0; k < b.tiles; ++k) p.tileScene.push_back((int)s);
        for (int k = 0; k < b.buckets; ++k) p.bucketScene.push_back((int)s);
        pBase += n;
        cBase += c;
        bBase += b.buckets;
        tBase += b.tiles;
        info += (long long)b.buckets * tiles;
        if (cBase > 0xffffffffll || info > 0x7fffffffll) return false;
    }
    p.infoWords = (size_t)info;
    p.bucketBits = maxSceneBuckets <= 256 ? 8 : maxSceneBuckets <= 512 ? 9 : 10;
    *out = std::move(p);
    return true;
}

void launch_bucket_sort(const TileKeyOut &to, unsigned *kOut, unsigned *vOut, const float4 *rec, float4 *sortedPos,
                        float4 *sortedVel, unsigned *table, unsigned numCells, int split, hipStream_t st) {
    if (to.numBuckets <= 0) return;
    hipLaunchKernelGGL((k_bucket_sort<16, true>), dim3(to.numBuckets * split), dim3(kBkThreads), 0, st, to, split, kOut,
                       vOut, rec, sortedPos, sortedVel, table, numCells);
}

void launch_bucket_sort_morton(const TileKeyOut &to, unsigned *kOut, unsigned *vOut, float4 *scratchA,
                               float4 *scratchB, hipStream_t st) {
    if (to.numBuckets <= 0) return;
    hipLaunchKernelGGL((k_bucket_sort<12, false>), dim3(to.numBuckets), dim3(kBkThreads), 0, st, to, 1, kOut, vOut,
                       nullptr, scratchA, scratchB, nullptr, 0u);
}

}  // namespace arsim
