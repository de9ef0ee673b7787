// csr_fill.hpp -- device helpers that write the CSR offset tables (cell / segment starts).
#pragma once

#include <hip/hip_runtime.h>

namespace arsim {

// Writes v to p[a..b) with 16-byte stores for the aligned middle of long runs.
__device__ __forceinline__ void fill_range(unsigned *p, long long a, long long b, unsigned v) {
    long long k = a;
    while (k < b && (k & 3)) p[k++] = v;
    const uint4 v4 = make_uint4(v, v, v, v);
    for (; k + 4 <= b; k += 4) *reinterpret_cast<uint4 *>(p + k) = v4;
    for (; k < b; ++k) p[k] = v;
}

// Ranges longer than kLongRange entries that a kernel hands to k_fill_long (launch_fill_long), cut
// into chunks of at most kLongChunk entries: list[k] = (a, b, v, 0) for k < *count.  One wave filling
// a run of millions of empty cells (a scene that occupies a small part of its grid) would take
// milliseconds; k_fill_long gives every chunk to a workgroup.  *count must be zero before the kernel
// that appends runs.
struct LongFill {
    unsigned *count;
    uint4 *list;
    unsigned cap;
};
constexpr unsigned kLongRange = 4096u;
constexpr unsigned kLongChunk = 65536u;

// Fills p[a, b) with v for every lane that has a range: short ranges by the lane itself, long
// ones (empty-cell runs, the tail of the table) by the whole wave with coalesced stores, and runs
// longer than kLongRange through lf when lf.list is set.  Every lane of the wave must call it.
__device__ __forceinline__ void wave_fill(unsigned *p, bool has, unsigned a, unsigned b, unsigned v,
                                          const LongFill &lf = LongFill{nullptr, nullptr, 0u}) {
    if (lf.list && has && b - a > kLongRange) {
        const unsigned nch = (b - a + kLongChunk - 1u) / kLongChunk;
        const unsigned slot = atomicAdd(lf.count, nch);
        if (slot + nch <= lf.cap) {
            for (unsigned q = 0; q < nch; ++q) {
                const unsigned c0 = a + q * kLongChunk;
                lf.list[slot + q] = make_uint4(c0, b - c0 > kLongChunk ? c0 + kLongChunk : b, v, 0u);
            }
            has = false;
        }
    }
    const bool lng = has && b - a > 32u;
    if (has && !lng) fill_range(p, a, b, v);
    unsigned long long m = __ballot(lng);
    const unsigned lane = threadIdx.x & 63u;
    while (m) {
        const int l = __builtin_ctzll(m);
        m &= m - 1;
        const unsigned la = __shfl(a, l), lb = __shfl(b, l), lv = __shfl(v, l);
        for (unsigned k = la + lane; k < lb; k += 64u) p[k] = lv;
    }
}

}  // namespace arsim
