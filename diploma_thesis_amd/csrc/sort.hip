// sort.hip -- stable LSD radix sort of (key, u32 value) pairs for gfx950 (wave64).
//
// Replaces cub::DeviceRadixSort::SortPairs as used by sortParticlesPreallocated
// (Kernels/particleKernelWrappers.cu:419-534) and createMortonCodes (BVHSource/createMortonCodes.cu:196-202).
//
// Structure (onesweep):
//   k_sort_init   zero the digit histograms, the tile counters and the first look-back array;
//   k_sort_hist   one read of the keys builds the digit histogram of every pass at once;
//   k_onesweep    one launch per digit pass.  A workgroup takes the next tile id from an atomic
//                 counter (so every earlier tile is already resident), ranks its keys per wave with
//                 RBITS ballots (peers = lanes with the same digit), builds per-wave LDS digit
//                 histograms, publishes its per-digit count and resolves its global digit offset by
//                 decoupled look-back over the tiles before it, scatters the tile into LDS in digit
//                 order and writes it out in runs of equal digits.
// Look-back words pack a 2-bit flag (1 = aggregate, 2 = inclusive prefix) over a 30-bit count, so a
// relaxed agent-scope atomic load sees flag and value together.  Pass p uses one of two look-back
// arrays and clears the other for pass p + 1.
#include <mutex>

#include "../../include/arsim.h"
#include "kernels.hpp"

namespace arsim {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr unsigned kFlagAgg = 1u << 30;
constexpr unsigned kFlagInc = 2u << 30;
constexpr unsigned kValMask = (1u << 30) - 1u;

// Exclusive scan of one value per thread over the block (threads >= count contribute 0).
__device__ __forceinline__ unsigned block_exclusive_scan(unsigned v, unsigned *sWave) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    unsigned x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        unsigned y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) sWave[wave] = x;
    __syncthreads();
    unsigned waveOff = 0;
    for (int w = 0; w < wave; ++w) waveOff += sWave[w];
    __syncthreads();
    return waveOff + x - v;
}

__global__ void k_sort_init(unsigned *hist, int histWords, unsigned *counters, int numCounters,
                            unsigned *status0, int statusWords) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int stride = gridDim.x * blockDim.x;
    for (int k = i; k < histWords; k += stride) hist[k] = 0u;
    for (int k = i; k < numCounters; k += stride) counters[k] = 0u;
    for (int k = i; k < statusWords; k += stride) status0[k] = 0u;
}

template <class K, int RBITS>
__global__ __launch_bounds__(kBlock) void k_sort_hist(const K *__restrict__ keys, int n, int beginBit,
                                                       int endBit, int passes, unsigned *__restrict__ hist) {
    constexpr int RADIX = 1 << RBITS;
    constexpr int MAXP = (int)(sizeof(K) * 8 + RBITS - 1) / RBITS;
    __shared__ unsigned sh[MAXP * RADIX];
    for (int k = threadIdx.x; k < passes * RADIX; k += kBlock) sh[k] = 0u;
    __syncthreads();
    int stride = gridDim.x * kBlock;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        K key = keys[i];
        for (int p = 0; p < passes; ++p) {
            int sh0 = beginBit + p * RBITS;
            int bits = min(RBITS, endBit - sh0);
            unsigned d = (unsigned)((key >> sh0) & (K)((1u << bits) - 1u));
            // keys in spatial order often share a digit across the wave: one add for all of them
            const uint64_t active = __ballot(1);
            const int first = __ffsll((long long)active) - 1;
            const unsigned d0 = __shfl(d, first, 64);
            if (__ballot(d == d0) == active) {
                if ((int)(threadIdx.x & 63) == first) atomicAdd(&sh[p * RADIX + d], (unsigned)__popcll(active));
            } else {
                atomicAdd(&sh[p * RADIX + d], 1u);
            }
        }
    }
    __syncthreads();
    // global layout: pass p's histogram starts at p * 256 whatever the digit width
    for (int k = threadIdx.x; k < passes * RADIX; k += kBlock)
        if (sh[k]) atomicAdd(&hist[(k / RADIX) * 256 + (k % RADIX)], sh[k]);
}

template <class K, int IPT, int RBITS>
__global__ __launch_bounds__(kBlock) void k_onesweep(const K *__restrict__ kin, const unsigned *__restrict__ vin,
                                                      K *__restrict__ kout, unsigned *__restrict__ vout, int n,
                                                      int shift, int bits, int valsIota,
                                                      const unsigned *__restrict__ histPass,
                                                      unsigned *tileCounter, unsigned *status,
                                                      unsigned *statusNext) {
    constexpr int RADIX = 1 << RBITS;
    constexpr int TILE = kBlock * IPT;
    __shared__ K sKeys[TILE];
    __shared__ unsigned sVals[TILE];
    __shared__ unsigned sWaveHist[kWaves * RADIX];
    __shared__ unsigned sDigitStart[RADIX];
    __shared__ unsigned sGlobalBase[RADIX];
    __shared__ unsigned sScan[kWaves];
    __shared__ unsigned sTile;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const unsigned dmask = (1u << bits) - 1u;
    if (tid == 0) sTile = atomicAdd(tileCounter, 1u);
    for (int k = tid; k < kWaves * RADIX; k += kBlock) sWaveHist[k] = 0u;
    __syncthreads();
    const unsigned tile = sTile;
    const long long base = (long long)tile * TILE;
    if (statusNext)
        for (int d = tid; d < RADIX; d += kBlock) statusNext[(size_t)tile * RADIX + d] = 0u;

    // ---- load (wave-striped: item k of lane l at waveBase + k*64 + l keeps the input order)
    K keys[IPT];
    unsigned vals[IPT];
    unsigned rank[IPT];
    const long long waveBase = base + (long long)wave * IPT * 64;
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        long long idx = waveBase + k * 64 + lane;
        bool valid = idx < n;
        keys[k] = valid ? kin[idx] : (K)0;
        vals[k] = valid ? (valsIota ? (unsigned)idx : vin[idx]) : 0u;
    }
    // ---- rank within the wave: peers = valid lanes with the same digit
    const uint64_t lowerMask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        bool valid = (waveBase + k * 64 + lane) < n;
        unsigned d = (unsigned)(keys[k] >> shift) & dmask;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < RBITS; ++b) {
            bool bit = (d >> b) & 1u;
            uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        unsigned prior = (unsigned)__popcll(peers & lowerMask);
        unsigned cnt = (unsigned)__popcll(peers);
        unsigned old = sWaveHist[wave * RADIX + d];
        rank[k] = old + prior;
        if (valid && prior == 0u) sWaveHist[wave * RADIX + d] = old + cnt;
    }
    __syncthreads();

    // ---- per digit: counts across waves (exclusive over waves) and the tile total
    unsigned tileCnt = 0u;
    if (tid < RADIX) {
        unsigned run = 0u;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            unsigned c = sWaveHist[w * RADIX + tid];
            sWaveHist[w * RADIX + tid] = run;
            run += c;
        }
        tileCnt = run;
        // publish this tile's aggregate as early as possible
        if (tile > 0)
            __hip_atomic_store(&status[(size_t)tile * RADIX + tid], kFlagAgg | tileCnt, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
    unsigned digitStart = block_exclusive_scan(tid < RADIX ? tileCnt : 0u, sScan);
    unsigned globalExcl = 0u;
    if (tile == 0) globalExcl = block_exclusive_scan(tid < RADIX ? histPass[tid < RADIX ? tid : 0] : 0u, sScan);
    if (tid < RADIX) {
        sDigitStart[tid] = digitStart;
        unsigned excl;
        if (tile == 0) {
            excl = globalExcl;
        } else {
            // decoupled look-back over the tiles before this one
            excl = 0u;
            long long pt = (long long)tile - 1;
            while (true) {
                unsigned s = __hip_atomic_load(&status[(size_t)pt * RADIX + tid], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                unsigned f = s & ~kValMask;
                if (f == 0u) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                excl += s & kValMask;
                if (f == kFlagInc) break;
                --pt;
            }
        }
        __hip_atomic_store(&status[(size_t)tile * RADIX + tid], kFlagInc | ((excl + tileCnt) & kValMask),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        sGlobalBase[tid] = excl;
    }
    __syncthreads();

    // ---- scatter into LDS in digit order
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        bool valid = (waveBase + k * 64 + lane) < n;
        if (valid) {
            unsigned d = (unsigned)(keys[k] >> shift) & dmask;
            unsigned pos = sDigitStart[d] + sWaveHist[wave * RADIX + d] + rank[k];
            sKeys[pos] = keys[k];
            sVals[pos] = vals[k];
        }
    }
    __syncthreads();

    // ---- write out: consecutive LDS slots of one digit go to consecutive global slots
    long long rem = (long long)n - base;
    int tileCount = rem < TILE ? (int)rem : TILE;
    for (int e = tid; e < tileCount; e += kBlock) {
        K key = sKeys[e];
        unsigned d = (unsigned)(key >> shift) & dmask;
        unsigned o = sGlobalBase[d] + ((unsigned)e - sDigitStart[d]);
        kout[o] = key;
        vout[o] = sVals[e];
    }
}

inline int num_passes(int beginBit, int endBit, int rbits) { return (endBit - beginBit + rbits - 1) / rbits; }

template <class K, int IPT, int RBITS>
void launch_passes(K **keys, unsigned **vals, K **alt, unsigned **altVals, int n, int beginBit, int endBit,
                   bool valsIota, unsigned *hist, unsigned *counters, unsigned *statusA, unsigned *statusB,
                   int numTiles, bool doHist, hipStream_t st) {
    const int passes = num_passes(beginBit, endBit, RBITS);
    if (doHist) {
        int histBlocks = (n + kBlock * 8 - 1) / (kBlock * 8);
        histBlocks = histBlocks < 1 ? 1 : (histBlocks > 4096 ? 4096 : histBlocks);
        hipLaunchKernelGGL((k_sort_hist<K, RBITS>), dim3(histBlocks), dim3(kBlock), 0, st, *keys, n, beginBit, endBit,
                           passes, hist);
    }
    for (int p = 0; p < passes; ++p) {
        int shift = beginBit + p * RBITS;
        int bits = endBit - shift < RBITS ? endBit - shift : RBITS;
        unsigned *cur = (p & 1) ? statusB : statusA;
        unsigned *nxt = (p + 1 < passes) ? ((p & 1) ? statusA : statusB) : nullptr;
        hipLaunchKernelGGL((k_onesweep<K, IPT, RBITS>), dim3(numTiles), dim3(kBlock), 0, st, *keys, *vals, *alt,
                           *altVals, n, shift, bits, (valsIota && p == 0) ? 1 : 0, hist + (size_t)p * 256,
                           counters + p, cur, nxt);
        K *tk = *keys; *keys = *alt; *alt = tk;
        unsigned *tv = *vals; *vals = *altVals; *altVals = tv;
    }
}

}  // namespace

int sort_max_passes(int keyBytes) { return (keyBytes * 8 + 5) / 6; }

namespace {
// sized for the smallest tile (IPT 4) and the narrowest digit (6 bits) any tuning may use
constexpr size_t workspace_bytes(long long n, int keyBytes) {
    const long long minTile = kBlock * 4;
    long long tiles = (n + minTile - 1) / minTile;
    if (tiles < 1) tiles = 1;
    const long long passes = (keyBytes * 8 + 5) / 6;
    const long long words = passes * 256 + 64 + 2 * tiles * 256;
    return (size_t)words * sizeof(unsigned) + 256;
}
}  // namespace

size_t sort_workspace_bytes(int n, int keyBytes) { return workspace_bytes(n, keyBytes); }

SortPlan sort_plan(void *ws, int n, int keyBytes, int beginBit, int endBit, const SortTuning &tu) {
    SortPlan p{};
    const int tile = kBlock * tu.ipt;
    p.numTiles = n > 0 ? (n + tile - 1) / tile : 1;
    p.passes = num_passes(beginBit, endBit, tu.radixBits);
    p.radix = 1 << tu.radixBits;
    unsigned *w = (unsigned *)ws;
    p.hist = w;
    p.counters = p.hist + (size_t)sort_max_passes(keyBytes) * 256;
    p.statusA = p.counters + 64;
    p.statusB = p.statusA + (size_t)p.numTiles * p.radix;
    // the per-pass histograms are laid out with a 256-word stride whatever the digit width
    p.histWords = (size_t)p.passes * 256;
    p.statusWords = (size_t)p.numTiles * p.radix;
    return p;
}

template <class K>
int sort_pairs(K **keys, unsigned **vals, K **alt, unsigned **altVals, int n, int beginBit, int endBit,
               bool valsIota, void *ws, size_t wsBytes, const SortTuning &tu, hipStream_t st, bool prepared) {
    if (n <= 0) return 0;
    if (n >= (1 << 30)) return -1;
    const int keyBits = (int)sizeof(K) * 8;
    if (beginBit < 0 || endBit > keyBits || beginBit >= endBit) return -1;
    if (wsBytes < sort_workspace_bytes(n, (int)sizeof(K))) return -2;
    const int ipt = tu.ipt, rbits = tu.radixBits;
    const SortPlan pl = sort_plan(ws, n, (int)sizeof(K), beginBit, endBit, tu);
    const int numTiles = pl.numTiles;
    unsigned *hist = pl.hist, *counters = pl.counters, *statusA = pl.statusA, *statusB = pl.statusB;
    if (!prepared) {
        int initWords = (int)pl.statusWords;
        int initBlocks = (initWords + kBlock - 1) / kBlock;
        initBlocks = initBlocks < 1 ? 1 : (initBlocks > 1024 ? 1024 : initBlocks);
        hipLaunchKernelGGL(k_sort_init, dim3(initBlocks), dim3(kBlock), 0, st, hist, (int)pl.histWords, counters,
                           pl.passes, statusA, initWords);
    }
#define ARSIM_SORT_CASE(I, R)                                                                                 \
    if (ipt == I && rbits == R) {                                                                             \
        launch_passes<K, I, R>(keys, vals, alt, altVals, n, beginBit, endBit, valsIota, hist, counters,       \
                               statusA, statusB, numTiles, !prepared, st);                                    \
        return 0;                                                                                             \
    }
    ARSIM_SORT_CASE(8, 8)
    ARSIM_SORT_CASE(12, 8)
    ARSIM_SORT_CASE(16, 8)
    ARSIM_SORT_CASE(8, 7)
    ARSIM_SORT_CASE(12, 7)
    ARSIM_SORT_CASE(16, 7)
    ARSIM_SORT_CASE(8, 6)
    ARSIM_SORT_CASE(16, 6)
    ARSIM_SORT_CASE(4, 8)
#undef ARSIM_SORT_CASE
    return -3;
}

template int sort_pairs<unsigned>(unsigned **, unsigned **, unsigned **, unsigned **, int, int, int, bool, void *,
                                  size_t, const SortTuning &, hipStream_t, bool);
template int sort_pairs<uint64_t>(uint64_t **, unsigned **, uint64_t **, unsigned **, int, int, int, bool, void *,
                                  size_t, const SortTuning &, hipStream_t, bool);

// ---------------------------------------------------------------- sort without a caller workspace
// The workspace of a u32-key sort of up to ARSIM_SORT_NOALLOC_MAX_ITEMS items lives in this
// file-scope device array (8.0 MiB; every device the library runs on holds its own copy until the
// process ends).  Host code must never pass the array's name to a kernel: on the host it is only a
// stand-in for the symbol.  The device address comes from hipGetSymbolAddress, taken once per device
// with that device current, and is kept per device.  Every sort on one device shares the array, so
// sorts that overlap on one device (two streams, two threads) race; the sort's first kernel clears
// the histograms, tile counters and look-back words it relies on, so each call starts clean.
static_assert(ARSIM_SORT_NOALLOC_MAX_ITEMS < (1 << 30), "item limit must fit the sort's index range");
constexpr size_t kStaticSortScratchBytes = workspace_bytes(ARSIM_SORT_NOALLOC_MAX_ITEMS, 4);
static __device__ __align__(256) unsigned char scratch[kStaticSortScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
void *g_scratchPtr[kMaxDevices] = {};
}  // namespace

hipError_t sort_static_scratch(void **ptr, size_t *bytes) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (!g_scratchPtr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratchPtr[dev] = p;
    }
    *ptr = g_scratchPtr[dev];
    *bytes = kStaticSortScratchBytes;
    return hipSuccess;
}

bool sort_tuning_supported(int ipt, int rbits) {
    return (rbits == 8 && (ipt == 4 || ipt == 8 || ipt == 12 || ipt == 16)) ||
           (rbits == 7 && (ipt == 8 || ipt == 12 || ipt == 16)) || (rbits == 6 && (ipt == 8 || ipt == 16));
}

}  // namespace arsim
