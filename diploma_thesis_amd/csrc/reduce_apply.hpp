// reduce_apply.hpp -- the per-body sum over a wave and the velocity update, shared by the separate
// reduction kernel and the fused contact + reduction kernel so that both produce the same bits.
#pragma once

#include "kernels.hpp"

namespace arsim {

// The intended sum of UniformGridDEM.cu:421-460: every lane holds the partial sums (aF, aT) of the
// body's particles l, l+64, ... in index order; a butterfly over lane distance 32..1 completes them,
// and lane 0 applies v += J dt / m and ω += I_w^-1 τ dt (J and τ without dt in the reference units).
// The body's render model matrix (column-major 4x4: R and the centre), when kept.
__device__ __forceinline__ void write_model(const BodyBufs &b, int body, const float R[9], f3 x) {
    if (!b.model) return;
    float *m = b.model + (size_t)body * 16;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        m[c * 4 + 0] = R[c * 3 + 0];
        m[c * 4 + 1] = R[c * 3 + 1];
        m[c * 4 + 2] = R[c * 3 + 2];
        m[c * 4 + 3] = 0.0f;
    }
    m[12] = x.x; m[13] = x.y; m[14] = x.z; m[15] = 1.0f;
}

__device__ __forceinline__ void wave_reduce_apply(f3 aF, f3 aT, int lane, int body, const BodyBufs &b,
                                                  const DevPhys &ph) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        aF.x = aF.x + __shfl_xor(aF.x, off, 64);
        aF.y = aF.y + __shfl_xor(aF.y, off, 64);
        aF.z = aF.z + __shfl_xor(aF.z, off, 64);
        aT.x = aT.x + __shfl_xor(aT.x, off, 64);
        aT.y = aT.y + __shfl_xor(aT.y, off, 64);
        aT.z = aT.z + __shfl_xor(aT.z, off, 64);
    }
    if (lane != 0) return;
    const float sc = (ph.flags & kRefUnits) ? 1.0f : ph.dt;
    const float m = b.mass[body];
    f3 dv = mk3((aF.x * sc) / m, (aF.y * sc) / m, (aF.z * sc) / m);
    float Iw[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Iw[k] = b.Iw[(size_t)body * 9 + k];
    f3 dw = mat_vec(Iw, mk3(aT.x * sc, aT.y * sc, aT.z * sc));
    float4 V = b.vel[body], W = b.ang[body];
    V.x += dv.x; V.y += dv.y; V.z += dv.z;
    W.x += dw.x; W.y += dw.y; W.z += dw.z;
    b.vel[body] = V;
    b.ang[body] = W;
}

// First particle of a body that enters its sum: the reference's scan difference drops it for every
// body but the first (compatibility flag).
__device__ __forceinline__ int reduce_first(int body, const DevPhys &ph) {
    return ((ph.flags & kDropFirst) && body > 0) ? 1 : 0;
}

}  // namespace arsim
