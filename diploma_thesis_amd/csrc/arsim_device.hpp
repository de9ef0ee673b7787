// arsim_device.hpp -- device-side math and shared structs of the gfx950 DEM pipeline.
//
// Every FP32 expression below is written in a fixed evaluation order and the library is built
// with -ffp-contract=off and correctly rounded division / sqrt, so a host restatement that
// evaluates the same expressions in the same order reproduces the device results bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace arsim {

// ----------------------------------------------------------------------------- shared structs
// Geometry of one scene (grid hashing, walls, Morton box, particle range).
struct SceneGeom {
    float origin[3];
    float cellSize[3];
    unsigned gridMask[3];    // gridSize - 1 per axis (sizes are powers of two)
    unsigned gx, gxy;        // gridSize.x, gridSize.x * gridSize.y
    unsigned cellBase;       // first global cell of this scene (scenes' cell ranges are packed)
    unsigned wallMask;
    float boxMin[3], boxMax[3];
    float mLo[3], mExt[3];   // Morton normalisation: lo and (hi - lo), computed in FP32 on the host
    int pStart, pCount;
};

// One scene passed by value, or a device table of scenes indexed by scene id.
struct SceneView {
    const SceneGeom *table;
    int count;
    SceneGeom single;
    // several scenes: scene s's particles are the sorted positions [sceneStart[s], sceneStart[s + 1])
    // (NULL where the caller has no such table)
    const int *sceneStart;
};

struct DevPhys {
    float k, c, s, a, r, dt;
    float gx, gy, gz;
    float globalDamping, Wdamping, boundaryDamping;
    unsigned flags;
    // a float >= (2r)^2 in exact arithmetic: d2 > cd2Reject implies sqrtf(d2) >= 2r, because sqrtf is
    // correctly rounded and monotone and 2r is a float; the contact test skips the square root then
    float cd2Reject;
};

constexpr unsigned kEmpty = 0xffffffffu;
constexpr unsigned kRefUnits = 0x01u, kDropFirst = 0x02u, kGrainClamp = 0x04u, kDynamicAabb = 0x08u;

// Maps hardware block b of nb to a logical block so that the blocks one XCD runs (dispatch is
// round-robin over the 8 XCDs) cover one contiguous slice of the work; neighbouring particles then
// share that XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int X = 8;
    int q = nb / X, r = nb % X;
    int x = b % X, k = b / X;
    return x * q + (x < r ? x : r) + k;
}

// ----------------------------------------------------------------------------- float3 helpers
struct f3 { float x, y, z; };
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }
__device__ __forceinline__ f3 add3(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 scl3(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 div3(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float4 f4(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

__device__ __forceinline__ void load_geom(const SceneView &sv, int s, SceneGeom &g) {
    if (sv.count <= 1) g = sv.single; else g = sv.table[s];
}

// ----------------------------------------------------------------------------- quaternions
struct qf { float x, y, z, w; };
__device__ __forceinline__ qf qmul(qf p, qf q) {    // glm tquat::operator*=
    qf r;
    r.w = p.w * q.w - p.x * q.x - p.y * q.y - p.z * q.z;
    r.x = p.w * q.x + p.x * q.w + p.y * q.z - p.z * q.y;
    r.y = p.w * q.y + p.y * q.w + p.z * q.x - p.x * q.z;
    r.z = p.w * q.z + p.z * q.w + p.x * q.y - p.y * q.x;
    return r;
}
__device__ __forceinline__ qf qnorm(qf q) {          // glm normalize(quat)
    float d = (q.x * q.x + q.y * q.y) + (q.z * q.z + q.w * q.w);
    float len = sqrtf(d);
    if (len <= 0.0f) return qf{0.f, 0.f, 0.f, 1.f};
    float inv = 1.0f / len;
    return qf{q.x * inv, q.y * inv, q.z * inv, q.w * inv};
}
__device__ __forceinline__ void mat3_cast(qf q, float m[9]) {   // glm mat3_cast, column-major
    float qxx = q.x * q.x, qyy = q.y * q.y, qzz = q.z * q.z;
    float qxz = q.x * q.z, qxy = q.x * q.y, qyz = q.y * q.z;
    float qwx = q.w * q.x, qwy = q.w * q.y, qwz = q.w * q.z;
    m[0] = 1.0f - 2.0f * (qyy + qzz); m[1] = 2.0f * (qxy + qwz);        m[2] = 2.0f * (qxz - qwy);
    m[3] = 2.0f * (qxy - qwz);        m[4] = 1.0f - 2.0f * (qxx + qzz); m[5] = 2.0f * (qyz + qwx);
    m[6] = 2.0f * (qxz + qwy);        m[7] = 2.0f * (qyz - qwx);        m[8] = 1.0f - 2.0f * (qxx + qyy);
}
__device__ __forceinline__ f3 mat_vec(const float m[9], f3 v) {
    return f3{m[0] * v.x + m[3] * v.y + m[6] * v.z,
              m[1] * v.x + m[4] * v.y + m[7] * v.z,
              m[2] * v.x + m[5] * v.y + m[8] * v.z};
}
__device__ __forceinline__ void mat_mul(const float A[9], const float B[9], float C[9]) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r)
            C[c * 3 + r] = A[r] * B[c * 3 + 0] + A[3 + r] * B[c * 3 + 1] + A[6 + r] * B[c * 3 + 2];
}

// Deterministic sin/cos of a float: double-precision reduction by the nearest multiple of
// pi/2, Taylor polynomials of degree 15 / 16 on |r| <= pi/4, rounded to float.
__device__ __forceinline__ void sincos_det(float x, float *s_out, float *c_out) {
    const double two_over_pi = 0.63661977236758134308;
    const double pio2 = 1.57079632679489661923;
    double xd = (double)x;
    double kd = rint(xd * two_over_pi);
    double r = xd - kd * pio2;
    double r2 = r * r;
    double s = r * (1.0 + r2 * (-1.0 / 6.0 + r2 * (1.0 / 120.0 + r2 * (-1.0 / 5040.0 + r2 * (1.0 / 362880.0 +
               r2 * (-1.0 / 39916800.0 + r2 * (1.0 / 6227020800.0 + r2 * (-1.0 / 1307674368000.0))))))));
    double c = 1.0 + r2 * (-1.0 / 2.0 + r2 * (1.0 / 24.0 + r2 * (-1.0 / 720.0 + r2 * (1.0 / 40320.0 +
               r2 * (-1.0 / 3628800.0 + r2 * (1.0 / 479001600.0 + r2 * (-1.0 / 87178291200.0 +
               r2 * (1.0 / 20922789888000.0))))))));
    long long k = (long long)kd;
    int qd = (int)(k & 3);
    double sv = s, cv = c;
    if (qd == 1) { sv = c; cv = -s; }
    else if (qd == 2) { sv = -s; cv = -c; }
    else if (qd == 3) { sv = -c; cv = s; }
    *s_out = (float)sv;
    *c_out = (float)cv;
}

// One step of the rigid-body integrator (GPUintegratorKernel, Kernels/rigidBodyKernels.cu:262-348): position by
// the old velocity, the incremental rotation from the angular velocity (theta = |w| dt), the
// cumulative orientation qc = normalize(qi * qc), gravity and angular damping.
struct BodyStep {
    f3 x, v, w;
    qf qi, qc;
};
__device__ __forceinline__ BodyStep body_step(float4 X, float4 V, float4 W, float4 Qc, const DevPhys &ph) {
    BodyStep r;
    const f3 v = xyz(V), w = xyz(W);
    r.x = add3(xyz(X), mk3(v.x * ph.dt, v.y * ph.dt, v.z * ph.dt));
    const float w2 = dot3(w, w);
    const float theta = sqrtf(w2) * ph.dt;
    qf qi{0.f, 0.f, 0.f, 1.f};
    if (theta > 1e-5f) {
        const float inv = 1.0f / sqrtf(w2);
        const f3 n = mk3(w.x * inv, w.y * inv, w.z * inv);
        float sh, ch;
        sincos_det(theta / 2.0f, &sh, &ch);
        qi = qf{sh * n.x, sh * n.y, sh * n.z, ch};
    }
    r.qi = qnorm(qi);
    r.qc = qnorm(qmul(r.qi, qf{Qc.x, Qc.y, Qc.z, Qc.w}));
    const bool ref = (ph.flags & kRefUnits) != 0;
    const f3 gs = ref ? mk3(ph.gx, ph.gy, ph.gz) : mk3(ph.gx * ph.dt, ph.gy * ph.dt, ph.gz * ph.dt);
    r.v = add3(v, gs);
    r.w = mk3(w.x * ph.Wdamping, w.y * ph.Wdamping, w.z * ph.Wdamping);
    return r;
}
// R = mat3_cast(qc) and the world inverse inertia I_w^-1 = R I_b^-1 R^T (Ib: the body's 9 floats).
__device__ __forceinline__ void body_world_inertia(qf qc, const float *Ib, float R[9], float Iw[9]) {
    float RT[9], B[9], tmp[9];
    mat3_cast(qc, R);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) RT[c * 3 + r] = R[r * 3 + c];
#pragma unroll
    for (int k = 0; k < 9; ++k) B[k] = Ib[k];
    mat_mul(R, B, tmp);
    mat_mul(tmp, RT, Iw);
}

// ----------------------------------------------------------------------------- DEM
// Spring-dashpot-shear(-attraction) pair force of collideSpheres (particleKernels.cu:261-299).
__device__ __forceinline__ bool pair_force(f3 pa, f3 pb, f3 va, f3 vb, const DevPhys &ph, f3 &out) {
    f3 rel = sub3(pb, pa);
    const float d2 = dot3(rel, rel);
    if (d2 > ph.cd2Reject) return false;   // certainly no contact; same outcome as the test below
    float dist = sqrtf(d2);
    float cd = 2.0f * ph.r;
    if (!(dist < cd)) return false;
    f3 n = div3(rel, dist);
    f3 rv = sub3(vb, va);
    float vn = dot3(rv, n);
    f3 tv = sub3(rv, scl3(vn, n));
    float sp = -ph.k * (cd - dist);
    f3 F = scl3(sp, n);
    F = add3(F, scl3(ph.c, rv));
    F = add3(F, scl3(ph.s, tv));
    F = add3(F, scl3(ph.a, rel));
    out = F;
    return true;
}
// Force on a particle from a static scene point (FindAndHandleAugmentedRealityCollisionsUniformGridCell,
// RigidBodies/UniformGridDEM.cu:243-301): the pair formula with the normal taken as minus the scene
// point's normal and a partner at rest (relVel = 0 - va).
__device__ __forceinline__ bool static_force(f3 pa, f3 ps, f3 ns, f3 va, const DevPhys &ph, f3 &out) {
    f3 rel = sub3(ps, pa);
    const float d2 = dot3(rel, rel);
    if (d2 > ph.cd2Reject) return false;
    float dist = sqrtf(d2);
    float cd = 2.0f * ph.r;
    if (!(dist < cd)) return false;
    f3 n = mk3(-ns.x, -ns.y, -ns.z);
    f3 rv = sub3(mk3(0.f, 0.f, 0.f), va);
    float vn = dot3(rv, n);
    f3 tv = sub3(rv, scl3(vn, n));
    float sp = -ph.k * (cd - dist);
    f3 F = scl3(sp, n);
    F = add3(F, scl3(ph.c, rv));
    F = add3(F, scl3(ph.s, tv));
    F = add3(F, scl3(ph.a, rel));
    out = F;
    return true;
}
// Contact predicate only (same arithmetic as the distance test above).
__device__ __forceinline__ bool in_contact(f3 pa, f3 pb, const DevPhys &ph) {
    f3 rel = sub3(pb, pa);
    const float d2 = dot3(rel, rel);
    if (d2 > ph.cd2Reject) return false;
    float dist = sqrtf(d2);
    return dist < 2.0f * ph.r;
}
__device__ __forceinline__ bool pair_allowed(int ra, int rb) { return ra != rb || ra == -1; }
// True when the sorted copy Q (xyz position, w = rbIndex bits) may be in contact with particle a:
// not a itself, an allowed pair, and not certainly out of reach.  The three tests are combined
// without short-circuiting so that all of Q is needed at once (one 16-byte load, not two).
__device__ __forceinline__ bool candidate(float4 Q, f3 pa, int ra, bool self, const DevPhys &ph) {
    const f3 rel = sub3(mk3(Q.x, Q.y, Q.z), pa);
    const bool near = !(dot3(rel, rel) > ph.cd2Reject);
    return near & !self & pair_allowed(ra, __float_as_int(Q.w));
}

// Static-plane walls: the static-contact formula with a plane partner, walls in bit order.
__device__ __forceinline__ f3 wall_force(f3 p, f3 v, const DevPhys &ph, const SceneGeom &g) {
    f3 tot = mk3(0.f, 0.f, 0.f);
#pragma unroll
    for (int w = 0; w < 6; ++w) {
        if (!(g.wallMask & (1u << w))) continue;
        int axis = w >> 1;
        bool isMax = (w & 1) != 0;
        float pc = axis == 0 ? p.x : (axis == 1 ? p.y : p.z);
        float d = isMax ? (g.boxMax[axis] - pc) : (pc - g.boxMin[axis]);
        if (!(d < ph.r)) continue;
        float sgn = isMax ? 1.0f : -1.0f;
        f3 n = mk3(axis == 0 ? sgn : 0.0f, axis == 1 ? sgn : 0.0f, axis == 2 ? sgn : 0.0f);
        f3 rv = mk3(-v.x, -v.y, -v.z);
        float vn = dot3(rv, n);
        f3 tv = sub3(rv, scl3(vn, n));
        float sp = -ph.k * (ph.r - d);
        f3 F = scl3(sp, n);
        F = add3(F, scl3(ph.c, rv));
        F = add3(F, scl3(ph.s, tv));
        F = add3(F, scl3(ph.a, scl3(d, n)));
        tot = add3(tot, F);
    }
    return tot;
}

// ----------------------------------------------------------------------------- grid hashing
// calcGridPos / calcGridHash (particleKernels.cu:151-167)
__device__ __forceinline__ void grid_pos(f3 p, const SceneGeom &g, int &x, int &y, int &z) {
    x = (int)floorf((p.x - g.origin[0]) / g.cellSize[0]);
    y = (int)floorf((p.y - g.origin[1]) / g.cellSize[1]);
    z = (int)floorf((p.z - g.origin[2]) / g.cellSize[2]);
}
__device__ __forceinline__ unsigned grid_hash(int x, int y, int z, const SceneGeom &g) {
    unsigned ux = (unsigned)x & g.gridMask[0];
    unsigned uy = (unsigned)y & g.gridMask[1];
    unsigned uz = (unsigned)z & g.gridMask[2];
    return uz * g.gxy + uy * g.gx + ux;
}

// ----------------------------------------------------------------------------- Morton codes
// expandBits / morton3D (BVHkernels.cu:646-668) and the 21-bit variant.
__device__ __forceinline__ unsigned expand10(unsigned v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
__device__ __forceinline__ uint64_t expand21(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}
__device__ __forceinline__ float mnorm(float p, float lo, float ext) {
    if (ext == 0.0f) return 0.0f;
    return (p - lo) / ext;
}
__device__ __forceinline__ unsigned morton30(float x, float y, float z) {
    x = fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f);
    y = fminf(fmaxf(y * 1024.0f, 0.0f), 1023.0f);
    z = fminf(fmaxf(z * 1024.0f, 0.0f), 1023.0f);
    return expand10((unsigned)x) * 4u + expand10((unsigned)y) * 2u + expand10((unsigned)z);
}
__device__ __forceinline__ uint64_t morton63(float x, float y, float z) {
    x = fminf(fmaxf(x * 2097152.0f, 0.0f), 2097151.0f);
    y = fminf(fmaxf(y * 2097152.0f, 0.0f), 2097151.0f);
    z = fminf(fmaxf(z * 2097152.0f, 0.0f), 2097151.0f);
    return (expand21((uint64_t)x) << 2) | (expand21((uint64_t)y) << 1) | expand21((uint64_t)z);
}

}  // namespace arsim
