/*
 * arsim_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * A clean-room, single-threaded host restatement of AR_Sim's per-step DEM pipeline, used as
 * the checker for the gfx950 kernels.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * Two instantiations of every stage:
 *   T = float  : the FP32 "mirror" -- same operation order as the device kernels, built with
 *                -ffp-contract=off and IEEE division / sqrtf, so integer stages are bit-exact
 *                and FP32 stages are expected to match the device bit for bit;
 *   T = double : the FP64 reference-semantics model (forces, reduction, integration, drift).
 *
 * Each function cites the reference source it restates.  How each is pinned (known answers,
 * brute force, invariants) is stated in tests/test_oracle_*.py; functions without such a
 * check would be marked "parity unpinned" here -- at present every exported function has one.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/arsim.h"

namespace {

template <class T> struct V3 { T x, y, z; };
template <class T> inline V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> add(V3<T> a, V3<T> b) { return mk<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> inline V3<T> sub(V3<T> a, V3<T> b) { return mk<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> inline V3<T> scl(T s, V3<T> a) { return mk<T>(s * a.x, s * a.y, s * a.z); }
template <class T> inline V3<T> divs(V3<T> a, T s) { return mk<T>(a.x / s, a.y / s, a.z / s); }
// helper_math.h semantics: dot = x*x' + y*y' + z*z' (left to right), length = sqrt(dot).
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return mk<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class T> inline T vsqrt(T v) { return std::sqrt(v); }
template <class T> inline V3<T> ld3(const T *p4, long i) { return mk<T>(p4[4 * i], p4[4 * i + 1], p4[4 * i + 2]); }
template <class T> inline void st4(T *p4, long i, V3<T> v, T w) {
    p4[4 * i] = v.x; p4[4 * i + 1] = v.y; p4[4 * i + 2] = v.z; p4[4 * i + 3] = w;
}

// Quaternion (x,y,z,w) as glm 0.9.7 stores it (Libs/glm/glm/gtc/quaternion.hpp:77).
template <class T> struct Q { T x, y, z, w; };
// glm tquat::operator*= (gtc/quaternion.inl:263-272)
template <class T> inline Q<T> qmul(Q<T> p, Q<T> q) {
    Q<T> r;
    r.w = p.w * q.w - p.x * q.x - p.y * q.y - p.z * q.z;
    r.x = p.w * q.x + p.x * q.w + p.y * q.z - p.z * q.y;
    r.y = p.w * q.y + p.y * q.w + p.z * q.x - p.x * q.z;
    r.z = p.w * q.z + p.z * q.w + p.x * q.y - p.y * q.x;
    return r;
}
// glm normalize(quat) (gtc/quaternion.inl:390-397); the dot is summed as (xx+yy)+(zz+ww).
template <class T> inline Q<T> qnorm(Q<T> q) {
    T d = (q.x * q.x + q.y * q.y) + (q.z * q.z + q.w * q.w);
    T len = vsqrt(d);
    if (len <= T(0)) return Q<T>{T(0), T(0), T(0), T(1)};
    T inv = T(1) / len;
    return Q<T>{q.x * inv, q.y * inv, q.z * inv, q.w * inv};
}
// glm mat3_cast (gtc/quaternion.inl:598-624); m[c*3 + r] = column c, row r.
template <class T> inline void mat3_cast(Q<T> q, T m[9]) {
    T qxx = q.x * q.x, qyy = q.y * q.y, qzz = q.z * q.z;
    T qxz = q.x * q.z, qxy = q.x * q.y, qyz = q.y * q.z;
    T qwx = q.w * q.x, qwy = q.w * q.y, qwz = q.w * q.z;
    m[0] = T(1) - T(2) * (qyy + qzz); m[1] = T(2) * (qxy + qwz);        m[2] = T(2) * (qxz - qwy);
    m[3] = T(2) * (qxy - qwz);        m[4] = T(1) - T(2) * (qxx + qzz); m[5] = T(2) * (qyz + qwx);
    m[6] = T(2) * (qxz + qwy);        m[7] = T(2) * (qyz - qwx);        m[8] = T(1) - T(2) * (qxx + qyy);
}
// glm mat3 * vec3: row r = (m[0][r]*v.x + m[1][r]*v.y) + m[2][r]*v.z
template <class T> inline V3<T> mv(const T m[9], V3<T> v) {
    return mk<T>(m[0] * v.x + m[3] * v.y + m[6] * v.z,
                 m[1] * v.x + m[4] * v.y + m[7] * v.z,
                 m[2] * v.x + m[5] * v.y + m[8] * v.z);
}
// C = A * B (column-major), C[c][r] = (A[0][r]*B[c][0] + A[1][r]*B[c][1]) + A[2][r]*B[c][2]
template <class T> inline void mm(const T A[9], const T B[9], T C[9]) {
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            C[c * 3 + r] = A[0 * 3 + r] * B[c * 3 + 0] + A[1 * 3 + r] * B[c * 3 + 1] + A[2 * 3 + r] * B[c * 3 + 2];
}
template <class T> inline void mt(const T A[9], T B[9]) {
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) B[c * 3 + r] = A[r * 3 + c];
}

// Deterministic sin/cos: argument in double, reduced by k*pi/2 (k = nearest integer),
// then Taylor series to degree 15 (sin) / 16 (cos) on |r| <= pi/4, rounded to float.
// Written from the Taylor expansion; restated by the device code in the same order.
inline void sincos_det(float x, float *s_out, float *c_out) {
    const double two_over_pi = 0.63661977236758134308;
    const double pio2 = 1.57079632679489661923;
    double xd = (double)x;
    double kd = std::nearbyint(xd * two_over_pi);
    double r = xd - kd * pio2;
    double r2 = r * r;
    double s = r * (1.0 + r2 * (-1.0 / 6.0 + r2 * (1.0 / 120.0 + r2 * (-1.0 / 5040.0 + r2 * (1.0 / 362880.0 +
               r2 * (-1.0 / 39916800.0 + r2 * (1.0 / 6227020800.0 + r2 * (-1.0 / 1307674368000.0))))))));
    double c = 1.0 + r2 * (-1.0 / 2.0 + r2 * (1.0 / 24.0 + r2 * (-1.0 / 720.0 + r2 * (1.0 / 40320.0 +
               r2 * (-1.0 / 3628800.0 + r2 * (1.0 / 479001600.0 + r2 * (-1.0 / 87178291200.0 +
               r2 * (1.0 / 20922789888000.0))))))));
    long long k = (long long)kd;
    int qd = (int)(k & 3);
    double sv, cv;
    switch (qd) {
        case 0: sv = s; cv = c; break;
        case 1: sv = c; cv = -s; break;
        case 2: sv = -s; cv = -c; break;
        default: sv = -c; cv = s; break;
    }
    *s_out = (float)sv;
    *c_out = (float)cv;
}
inline void sincos_t(float x, float *s, float *c) { sincos_det(x, s, c); }
inline void sincos_t(double x, double *s, double *c) { *s = std::sin(x); *c = std::cos(x); }

template <class T> struct Phys {
    T k, c, s, a, r, dt;
    V3<T> g;
    T globalDamping, Wdamping, boundaryDamping;
    unsigned flags, wallMask;
    V3<T> bmin, bmax;
};
template <class T> Phys<T> mkphys(const ArSimParams *p, const ArSimStepConfig *cfg) {
    Phys<T> f;
    f.k = (T)p->spring; f.c = (T)p->damping; f.s = (T)p->shear; f.a = (T)p->attraction;
    f.r = (T)p->particleRadius; f.dt = (T)cfg->dt;
    f.g = mk<T>((T)p->gravity.x, (T)p->gravity.y, (T)p->gravity.z);
    f.globalDamping = (T)p->globalDamping; f.Wdamping = (T)p->Wdamping;
    f.boundaryDamping = (T)p->boundaryDamping;
    f.flags = cfg->flags; f.wallMask = cfg->wallMask;
    f.bmin = mk<T>((T)cfg->boxMin[0], (T)cfg->boxMin[1], (T)cfg->boxMin[2]);
    f.bmax = mk<T>((T)cfg->boxMax[0], (T)cfg->boxMax[1], (T)cfg->boxMax[2]);
    return f;
}

// DEM pair force: collideSpheres (Kernels/particleKernels.cu:261-299), identical to
// UniformGridDEM.cu:96-123 and BVHAuxiliary.cuh:96-130.  Returns true on contact.
template <class T>
inline bool pair_force(V3<T> pa, V3<T> pb, V3<T> va, V3<T> vb, const Phys<T> &f, V3<T> *out) {
    V3<T> rel = sub(pb, pa);
    T dist = vsqrt(dot(rel, rel));
    T cd = T(2) * f.r;
    if (!(dist < cd)) return false;
    V3<T> n = divs(rel, dist);
    V3<T> rv = sub(vb, va);
    T vn = dot(rv, n);
    V3<T> tv = sub(rv, scl(vn, n));
    T sp = -f.k * (cd - dist);
    V3<T> F = scl(sp, n);
    F = add(F, scl(f.c, rv));
    F = add(F, scl(f.s, tv));
    F = add(F, scl(f.a, rel));
    *out = F;
    return true;
}

// DEM force against a static scene point (FindAndHandleAugmentedRealityCollisionsUniformGridCell,
// RigidBodies/UniformGridDEM.cu:243-301): contact when |ps - pa| < 2r; the normal is minus the
// scene point's normal, the partner velocity is zero (relVel = 0 - va), the rest is the pair formula.
template <class T>
inline bool static_force(V3<T> pa, V3<T> ps, V3<T> ns, V3<T> va, const Phys<T> &f, V3<T> *out) {
    V3<T> rel = sub(ps, pa);
    T dist = vsqrt(dot(rel, rel));
    T cd = T(2) * f.r;
    if (!(dist < cd)) return false;
    V3<T> n = mk<T>(-ns.x, -ns.y, -ns.z);
    V3<T> rv = sub(mk<T>(0, 0, 0), va);
    T vn = dot(rv, n);
    V3<T> tv = sub(rv, scl(vn, n));
    T sp = -f.k * (cd - dist);
    V3<T> F = scl(sp, n);
    F = add(F, scl(f.c, rv));
    F = add(F, scl(f.s, tv));
    F = add(F, scl(f.a, rel));
    *out = F;
    return true;
}

// Static-plane wall term: the static-contact formula (UniformGridDEM.cu:281-300,
// particleKernels.cu:408-445) with the wall as the static partner: relVel = -v, contact
// distance r, overlap r - d, normal = -(inward wall normal).  Walls in bit order.
template <class T> inline V3<T> wall_force(V3<T> p, V3<T> v, const Phys<T> &f) {
    V3<T> tot = mk<T>(0, 0, 0);
    for (int w = 0; w < 6; ++w) {
        if (!(f.wallMask & (1u << w))) continue;
        int axis = w >> 1;
        bool isMax = (w & 1) != 0;
        T pc = axis == 0 ? p.x : (axis == 1 ? p.y : p.z);
        T lo = axis == 0 ? f.bmin.x : (axis == 1 ? f.bmin.y : f.bmin.z);
        T hi = axis == 0 ? f.bmax.x : (axis == 1 ? f.bmax.y : f.bmax.z);
        T d = isMax ? (hi - pc) : (pc - lo);
        if (!(d < f.r)) continue;
        T sgn = isMax ? T(1) : T(-1);     // direction from the particle into the wall
        V3<T> n = mk<T>(axis == 0 ? sgn : T(0), axis == 1 ? sgn : T(0), axis == 2 ? sgn : T(0));
        V3<T> rv = mk<T>(-v.x, -v.y, -v.z);
        T vn = dot(rv, n);
        V3<T> tv = sub(rv, scl(vn, n));
        T sp = -f.k * (f.r - d);
        V3<T> F = scl(sp, n);
        F = add(F, scl(f.c, rv));
        F = add(F, scl(f.s, tv));
        F = add(F, scl(f.a, scl(d, n)));
        tot = add(tot, F);
    }
    return tot;
}

// calcGridPos / calcGridHash (Kernels/particleKernels.cu:151-167)
inline uint32_t grid_hash(float px, float py, float pz, const ArSimParams *p) {
    int gx = (int)std::floor((px - p->worldOrigin.x) / p->cellSize.x);
    int gy = (int)std::floor((py - p->worldOrigin.y) / p->cellSize.y);
    int gz = (int)std::floor((pz - p->worldOrigin.z) / p->cellSize.z);
    uint32_t ux = (uint32_t)gx & (p->gridSize.x - 1u);
    uint32_t uy = (uint32_t)gy & (p->gridSize.y - 1u);
    uint32_t uz = (uint32_t)gz & (p->gridSize.z - 1u);
    return uz * p->gridSize.y * p->gridSize.x + uy * p->gridSize.x + ux;
}
inline uint32_t grid_hash_cell(int gx, int gy, int gz, const ArSimParams *p) {
    uint32_t ux = (uint32_t)gx & (p->gridSize.x - 1u);
    uint32_t uy = (uint32_t)gy & (p->gridSize.y - 1u);
    uint32_t uz = (uint32_t)gz & (p->gridSize.z - 1u);
    return uz * p->gridSize.y * p->gridSize.x + uy * p->gridSize.x + ux;
}

// expandBits / morton3D (BVHSource/BVHkernels.cu:646-668)
inline uint32_t expand10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
// 21-bit spread for the 63-bit code (build extension; BVHkernels.cu:4-5 reserve 21 bits).
inline uint64_t expand21(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}
// normalised coordinate (generateMortonCodes, BVHkernels.cu:670-716) with a zero-extent guard.
inline float mnorm(float p, float lo, float hi) {
    float ext = hi - lo;
    if (ext == 0.0f) return 0.0f;
    return (p - lo) / ext;
}
inline uint32_t morton30(float x, float y, float z) {
    x = std::fmin(std::fmax(x * 1024.0f, 0.0f), 1023.0f);
    y = std::fmin(std::fmax(y * 1024.0f, 0.0f), 1023.0f);
    z = std::fmin(std::fmax(z * 1024.0f, 0.0f), 1023.0f);
    return expand10((uint32_t)x) * 4u + expand10((uint32_t)y) * 2u + expand10((uint32_t)z);
}
inline uint64_t morton63(float x, float y, float z) {
    x = std::fmin(std::fmax(x * 2097152.0f, 0.0f), 2097151.0f);
    y = std::fmin(std::fmax(y * 2097152.0f, 0.0f), 2097151.0f);
    z = std::fmin(std::fmax(z * 2097152.0f, 0.0f), 2097151.0f);
    return (expand21((uint64_t)x) << 2) | (expand21((uint64_t)y) << 1) | expand21((uint64_t)z);
}

inline int clz32(uint32_t v) { return v ? __builtin_clz(v) : 32; }
inline int clz64(uint64_t v) { return v ? __builtin_clzll(v) : 64; }

// longestCommonPrefix (BVHSource/BVHAuxiliary.cuh:31-43); equal keys fall back to the
// leaf index (local to the leaf's scene when local != nullptr: base[i] is the scene start).
template <class K>
inline int lcp(const K *codes, int n, int i, int j, const int *base) {
    if (j < 0 || j >= n) return -1;
    K a = codes[i], b = codes[j];
    const int kb = (int)sizeof(K) * 8;
    if (a != b) return sizeof(K) == 8 ? clz64((uint64_t)(a ^ b)) : clz32((uint32_t)(a ^ b));
    uint32_t li = (uint32_t)(base ? i - base[i] : i), lj = (uint32_t)(base ? j - base[j] : j);
    return kb + clz32(li ^ lj);
}

// kernelConstructRadixTreeSoA (BVHSource/BVHSoAKernels.cu:7-99).  The split search's
// t = ceil(l / 2^k) is computed in integers (the reference goes through doubles; l < 2^31 so
// l / 2^k is exact in double and both round up to the same integer).
template <class K>
void karras(const K *codes, int n, const int *base, int *left, int *right, int *parent,
            int *minR, int *maxR, unsigned char *isLeaf) {
    for (int i = 0; i < n; ++i) { isLeaf[i] = 1; minR[i] = i; maxR[i] = i; }
    if (n < 2) { if (n == 1) parent[0] = -1; return; }
    parent[n] = -1;
    for (int i = 0; i < n - 1; ++i) {
        int dPrev = lcp(codes, n, i, i - 1, base);
        int dNext = lcp(codes, n, i, i + 1, base);
        int d = (dNext - dPrev) > 0 ? 1 : -1;
        int sigMin = lcp(codes, n, i, i - d, base);
        long long lmax = 2;
        while (lcp(codes, n, i, (int)(i + lmax * d), base) > sigMin) lmax *= 2;
        long long l = 0;
        for (long long t = lmax / 2; t >= 1; t /= 2) {
            if (lcp(codes, n, i, (int)(i + (l + t) * d), base) > sigMin) l += t;
            if (t == 1) break;
        }
        int j = (int)(i + l * d);
        int sigNode = lcp(codes, n, i, j, base);
        long long s = 0;
        for (int k = 1;; ++k) {
            long long t = (l + (1ll << k) - 1) >> k;
            if (lcp(codes, n, i, (int)(i + (s + t) * d), base) > sigNode) s += t;
            if (t == 1) break;
        }
        int gamma = (int)(i + s * d + std::min(d, 0));
        int node = i + n;
        isLeaf[node] = 0;
        int lo = std::min(i, j), hi = std::max(i, j);
        int L = (lo == gamma) ? gamma : gamma + n;
        int R = (hi == gamma + 1) ? gamma + 1 : gamma + 1 + n;
        left[node] = L; right[node] = R;
        parent[L] = node; parent[R] = node;
        minR[node] = lo; maxR[node] = hi;
    }
}

// Leaves (kernelConstructLeafNodesSoA, BVHSoAKernels.cu:108-138, initBound BVHAuxiliary.cuh:53-68)
// and internal bounds as the union of the children (mergeBounds, BVHAuxiliary.cuh:83-94).  The
// reference merges bottom-up with an atomic counter; min/max are exact, so a recursive union
// yields the same bits.
void refit(const int *sortedIdx, const float *pos4, float r, int n, const int *left, const int *right,
           float *cms4, float *bounds8) {
    for (int i = 0; i < n; ++i) {
        long o = sortedIdx[i];
        float x = pos4[4 * o], y = pos4[4 * o + 1], z = pos4[4 * o + 2];
        cms4[4 * i] = x; cms4[4 * i + 1] = y; cms4[4 * i + 2] = z; cms4[4 * i + 3] = pos4[4 * o + 3];
        float *b = bounds8 + 8 * (long)i;
        b[0] = x - r; b[1] = y - r; b[2] = z - r; b[3] = 0.f;
        b[4] = x + r; b[5] = y + r; b[6] = z + r; b[7] = 0.f;
    }
    if (n < 2) return;
    // post-order over internal nodes with an explicit stack
    std::vector<std::pair<int, int>> st;
    st.push_back({n, 0});
    while (!st.empty()) {
        auto &top = st.back();
        int node = top.first;
        if (top.second == 0) {
            top.second = 1;
            if (left[node] >= n) st.push_back({left[node], 0});
            continue;
        }
        if (top.second == 1) {
            top.second = 2;
            if (right[node] >= n) st.push_back({right[node], 0});
            continue;
        }
        const float *a = bounds8 + 8 * (long)left[node];
        const float *b = bounds8 + 8 * (long)right[node];
        float *t = bounds8 + 8 * (long)node;
        for (int c = 0; c < 3; ++c) {
            t[c] = a[c] < b[c] ? a[c] : b[c];
            t[4 + c] = a[4 + c] > b[4 + c] ? a[4 + c] : b[4 + c];
        }
        t[3] = 0.f; t[7] = 0.f;
        st.pop_back();
    }
}

inline bool overlap(const float *a, const float *b) {
    // checkOverlap(AABB, AABB) (BVHAuxiliary.cuh:139-152)
    return (a[0] <= b[4] && a[4] >= b[0]) && (a[1] <= b[5] && a[5] >= b[1]) && (a[2] <= b[6] && a[6] >= b[2]);
}

inline bool pair_allowed(int ra, int rb) { return ra != rb || ra == -1; }   // particleKernels.cu:330

}  // namespace

extern "C" {

int orc_abi_version(void) { return 1; }

/* ---- F5 calcHashD (particleKernels.cu:170-189) */
void orc_calc_hash(const float *pos4, int n, const ArSimParams *p, uint32_t *hash, uint32_t *index) {
    for (int i = 0; i < n; ++i) {
        hash[i] = grid_hash(pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2], p);
        index[i] = (uint32_t)i;
    }
}

/* ---- F12 morton codes (BVHkernels.cu:670-716) */
void orc_morton30(const float *pos4, int n, const float *bmin, const float *bmax, uint32_t *codes) {
    for (int i = 0; i < n; ++i)
        codes[i] = morton30(mnorm(pos4[4 * i], bmin[0], bmax[0]), mnorm(pos4[4 * i + 1], bmin[1], bmax[1]),
                            mnorm(pos4[4 * i + 2], bmin[2], bmax[2]));
}
void orc_morton63(const float *pos4, int n, const float *bmin, const float *bmax, uint64_t *codes) {
    for (int i = 0; i < n; ++i)
        codes[i] = morton63(mnorm(pos4[4 * i], bmin[0], bmax[0]), mnorm(pos4[4 * i + 1], bmin[1], bmax[1]),
                            mnorm(pos4[4 * i + 2], bmin[2], bmax[2]));
}
void orc_morton30_unit(const float *xyz, int n, uint32_t *codes) {
    for (int i = 0; i < n; ++i) codes[i] = morton30(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}
void orc_morton63_unit(const float *xyz, int n, uint64_t *codes) {
    for (int i = 0; i < n; ++i) codes[i] = morton63(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}
void orc_expand10(const uint32_t *v, int n, uint32_t *out) { for (int i = 0; i < n; ++i) out[i] = expand10(v[i]); }
void orc_expand21(const uint64_t *v, int n, uint64_t *out) { for (int i = 0; i < n; ++i) out[i] = expand21(v[i]); }

/* ---- F6/F13 stable LSD sort semantics: stable by the key bits [beginBit, endBit) */
void orc_sort_pairs_u32(const uint32_t *keys, const uint32_t *vals, int n, int beginBit, int endBit,
                        uint32_t *okeys, uint32_t *ovals) {
    uint32_t mask = (endBit - beginBit >= 32) ? 0xffffffffu : ((1u << (endBit - beginBit)) - 1u);
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
        return ((keys[a] >> beginBit) & mask) < ((keys[b] >> beginBit) & mask);
    });
    for (int i = 0; i < n; ++i) { okeys[i] = keys[idx[i]]; ovals[i] = vals[idx[i]]; }
}
void orc_sort_pairs_u64(const uint64_t *keys, const uint32_t *vals, int n, int beginBit, int endBit,
                        uint64_t *okeys, uint32_t *ovals) {
    uint64_t mask = (endBit - beginBit >= 64) ? ~0ull : ((1ull << (endBit - beginBit)) - 1ull);
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
        return ((keys[a] >> beginBit) & mask) < ((keys[b] >> beginBit) & mask);
    });
    for (int i = 0; i < n; ++i) { okeys[i] = keys[idx[i]]; ovals[i] = vals[idx[i]]; }
}

/* ---- F7 reorderDataAndFindCellStartD (particleKernels.cu:193-258) */
void orc_cell_bounds(const uint32_t *sortedHash, int n, uint32_t numCells, uint32_t *cellStart, uint32_t *cellEnd) {
    for (uint32_t c = 0; c < numCells; ++c) cellStart[c] = 0xffffffffu;
    for (int i = 0; i < n; ++i) {
        uint32_t h = sortedHash[i];
        if (i == 0 || h != sortedHash[i - 1]) {
            cellStart[h] = (uint32_t)i;
            if (i > 0) cellEnd[sortedHash[i - 1]] = (uint32_t)i;
        }
        if (i == n - 1) cellEnd[h] = (uint32_t)n;
    }
}

/* ---- F14 Karras (BVHSoAKernels.cu:7-99); base: per-leaf scene start or NULL */
void orc_karras_u32(const uint32_t *codes, int n, const int *base, int *left, int *right, int *parent,
                    int *minR, int *maxR, unsigned char *isLeaf) {
    karras<uint32_t>(codes, n, base, left, right, parent, minR, maxR, isLeaf);
}
void orc_karras_u64(const uint64_t *codes, int n, const int *base, int *left, int *right, int *parent,
                    int *minR, int *maxR, unsigned char *isLeaf) {
    karras<uint64_t>(codes, n, base, left, right, parent, minR, maxR, isLeaf);
}

/* ---- F15/F16 leaves + bounds */
void orc_refit(const int *sortedIdx, const float *pos4, float r, int n, const int *left, const int *right,
               float *cms4, float *bounds8) {
    refit(sortedIdx, pos4, r, n, left, right, cms4, bounds8);
}

/* ---- contact predicate by brute force over all pairs (BVHUpdate.cpp:315-414 pattern).
 * Writes pairs (i<j, original indices) and returns their number (or -1 if > cap). */
long orc_contacts_brute(const float *pos4, const int *rb, int n, float r, int *pairs, long cap) {
    long cnt = 0;
    float cd = 2.0f * r;
    for (int i = 0; i < n; ++i) {
        V3<float> pi = ld3(pos4, i);
        for (int j = i + 1; j < n; ++j) {
            if (!pair_allowed(rb[i], rb[j])) continue;
            V3<float> rel = sub(ld3(pos4, j), pi);
            float d = std::sqrt(dot(rel, rel));
            if (d < cd) {
                if (cnt >= cap) return -1;
                pairs[2 * cnt] = i; pairs[2 * cnt + 1] = j;
                ++cnt;
            }
        }
    }
    return cnt;
}

}  // extern "C"

/* ===================================================================== full step */
namespace {

template <class T> struct State {
    int N, B;
    T *pos, *vel, *relPos, *bodyFrame, *pForce, *pTorque;  // float4-shaped
    const int *rb;
    T *rbPos, *rbVel, *rbAng, *qinc, *qcum;                // float4-shaped
    const T *Ib;                                           // 9 per body
    T *Iw;                                                 // 9 per body
    const T *mass;
    const int *pstart, *pcount;
    T *model;                                              // optional: 16 per body (column-major mat4)
    const float *staticPos, *staticNorm;                   // optional static scene points (float4 each)
    int numStatic;
};

template <class T> inline void integrate_bodies(State<T> &s, const Phys<T> &f) {
    // GPUintegratorKernel (rigidBodyKernels.cu:262-348) with the orientation kept as a
    // cumulative quaternion and I_w^-1 = R_cum I_b^-1 R_cum^T.
    bool ref = (f.flags & ARSIM_COMPAT_REF_UNITS) != 0;
    for (int b = 0; b < s.B; ++b) {
        V3<T> x = ld3(s.rbPos, b), v = ld3(s.rbVel, b), w = ld3(s.rbAng, b);
        x = add(x, mk<T>(v.x * f.dt, v.y * f.dt, v.z * f.dt));
        T w2 = dot(w, w);
        T theta = vsqrt(w2) * f.dt;
        Q<T> qi{T(0), T(0), T(0), T(1)};
        if (theta > T(1e-5f)) {
            T inv = T(1) / vsqrt(w2);
            V3<T> nrm = mk<T>(w.x * inv, w.y * inv, w.z * inv);
            T sh, ch;
            sincos_t(theta / T(2), &sh, &ch);
            qi = Q<T>{sh * nrm.x, sh * nrm.y, sh * nrm.z, ch};
        }
        qi = qnorm(qi);
        Q<T> qc{s.qcum[4 * b], s.qcum[4 * b + 1], s.qcum[4 * b + 2], s.qcum[4 * b + 3]};
        qc = qnorm(qmul(qi, qc));
        T R[9], RT[9], tmp[9], Iw[9];
        mat3_cast(qc, R);
        mt(R, RT);
        mm(R, s.Ib + 9 * (long)b, tmp);
        mm(tmp, RT, Iw);
        for (int k = 0; k < 9; ++k) s.Iw[9 * (long)b + k] = Iw[k];
        if (s.model) {
            // modelMatrix of rigidBodyKernels.cu:324-331: rotation of the cumulative quaternion, then the
            // updated position in the last column
            T *m = s.model + 16 * (long)b;
            for (int c = 0; c < 3; ++c) {
                for (int r = 0; r < 3; ++r) m[4 * c + r] = R[3 * c + r];
                m[4 * c + 3] = T(0);
            }
            m[12] = x.x; m[13] = x.y; m[14] = x.z; m[15] = T(1);
        }
        V3<T> gs = ref ? f.g : mk<T>(f.g.x * f.dt, f.g.y * f.dt, f.g.z * f.dt);
        v = add(v, gs);
        w = mk<T>(w.x * f.Wdamping, w.y * f.Wdamping, w.z * f.Wdamping);
        s.rbPos[4 * b] = x.x; s.rbPos[4 * b + 1] = x.y; s.rbPos[4 * b + 2] = x.z;
        s.rbVel[4 * b] = v.x; s.rbVel[4 * b + 1] = v.y; s.rbVel[4 * b + 2] = v.z;
        s.rbAng[4 * b] = w.x; s.rbAng[4 * b + 1] = w.y; s.rbAng[4 * b + 2] = w.z;
        s.qinc[4 * b] = qi.x; s.qinc[4 * b + 1] = qi.y; s.qinc[4 * b + 2] = qi.z; s.qinc[4 * b + 3] = qi.w;
        s.qcum[4 * b] = qc.x; s.qcum[4 * b + 1] = qc.y; s.qcum[4 * b + 2] = qc.z; s.qcum[4 * b + 3] = qc.w;
    }
}

template <class T> inline void pose_and_grains(State<T> &s, const Phys<T> &f) {
    bool clamp = (f.flags & ARSIM_COMPAT_GRAIN_CLAMP) != 0;
    for (int i = 0; i < s.N; ++i) {
        int b = s.rb[i];
        if (b >= 0) {
            // computeGlobalAttributes (rigidBodyKernels.cu:386-421), r = R_cum * r_body
            Q<T> qc{s.qcum[4 * b], s.qcum[4 * b + 1], s.qcum[4 * b + 2], s.qcum[4 * b + 3]};
            T R[9];
            mat3_cast(qc, R);
            V3<T> r = mv(R, ld3(s.bodyFrame, i));
            st4(s.relPos, i, r, T(0));
            V3<T> cm = ld3(s.rbPos, b);
            st4(s.pos, i, add(cm, r), s.rbPos[4 * b + 3]);
            V3<T> V = ld3(s.rbVel, b), W = ld3(s.rbAng, b);
            st4(s.vel, i, add(V, cross(W, r)), s.rbVel[4 * b + 3]);
        } else {
            // integrate_functor (particleKernels.cu:50-129)
            V3<T> p = ld3(s.pos, i), v = ld3(s.vel, i);
            v = add(v, mk<T>(f.g.x * f.dt, f.g.y * f.dt, f.g.z * f.dt));
            v = mk<T>(v.x * f.globalDamping, v.y * f.globalDamping, v.z * f.globalDamping);
            p = add(p, mk<T>(v.x * f.dt, v.y * f.dt, v.z * f.dt));
            if (clamp) {
                // 0.01 margin added in double, as the reference's double literal does
                V3<T> hi = mk<T>((T)((double)f.bmax.x + 0.01), (T)((double)f.bmax.y + 0.01), (T)((double)f.bmax.z + 0.01));
                V3<T> lo = mk<T>((T)((double)f.bmin.x - 0.01), (T)((double)f.bmin.y - 0.01), (T)((double)f.bmin.z - 0.01));
                if (p.x > hi.x - f.r) { p.x = hi.x - f.r; v.x *= f.boundaryDamping; }
                if (p.x < lo.x + f.r) { p.x = lo.x + f.r; v.x *= f.boundaryDamping; }
                if (p.y > hi.y - f.r && v.y > T(0)) { p.y = hi.y - f.r; v.y *= f.boundaryDamping; }
                if (p.z > hi.z - f.r) { p.z = hi.z - f.r; v.z *= f.boundaryDamping; }
                if (p.z < lo.z + f.r) { p.z = lo.z + f.r; v.z *= f.boundaryDamping; }
                if (p.y < lo.y + f.r) { p.y = lo.y + f.r; v.y *= f.boundaryDamping; }
            }
            s.pos[4 * i] = p.x; s.pos[4 * i + 1] = p.y; s.pos[4 * i + 2] = p.z;
            s.vel[4 * i] = v.x; s.vel[4 * i + 1] = v.y; s.vel[4 * i + 2] = v.z;
        }
    }
}

// Per-particle DEM result for one query, shared by grid and LBVH drivers.
template <class T> inline void dem_write(State<T> &s, const Phys<T> &f, int orig, V3<T> F, V3<T> vq) {
    bool ref = (f.flags & ARSIM_COMPAT_REF_UNITS) != 0;
    if (s.rb[orig] >= 0) {
        st4(s.pForce, orig, F, T(0));
        st4(s.pTorque, orig, cross(ld3(s.relPos, orig), F), T(0));
    } else {
        T fs = ref ? T(1) : f.dt;
        st4(s.vel, orig, mk<T>(vq.x + F.x * fs, vq.y + F.y * fs, vq.z + F.z * fs), T(0));
    }
}

// Grid DEM (UniformGridDEM.cu:57-204 + collideD particleKernels.cu:302-406): cells in z, y, x
// order from -1 to 1, each cell's contacts summed into a cell partial that is then added to
// the total; then the wall term.
template <class T>
void dem_grid(State<T> &s, const Phys<T> &f, const ArSimParams *p, const uint32_t *sIdx,
              const uint32_t *cellStart, const uint32_t *cellEnd, const T *sPos, const T *sVel,
              const int *sRb) {
    bool clamp = (f.flags & ARSIM_COMPAT_GRAIN_CLAMP) != 0;
    for (int i = 0; i < s.N; ++i) {
        V3<T> pa = ld3(sPos, i), va = ld3(sVel, i);
        int ra = sRb[i];
        int gx = (int)std::floor(((float)pa.x - p->worldOrigin.x) / p->cellSize.x);
        int gy = (int)std::floor(((float)pa.y - p->worldOrigin.y) / p->cellSize.y);
        int gz = (int)std::floor(((float)pa.z - p->worldOrigin.z) / p->cellSize.z);
        V3<T> F = mk<T>(0, 0, 0);
        for (int z = -1; z <= 1; ++z)
            for (int y = -1; y <= 1; ++y)
                for (int x = -1; x <= 1; ++x) {
                    uint32_t h = grid_hash_cell(gx + x, gy + y, gz + z, p);
                    uint32_t st = cellStart[h];
                    if (st == 0xffffffffu) continue;
                    uint32_t en = cellEnd[h];
                    V3<T> cs = mk<T>(0, 0, 0);
                    for (uint32_t j = st; j < en; ++j) {
                        if ((int)j == i || !pair_allowed(ra, sRb[j])) continue;
                        V3<T> fj;
                        if (pair_force(pa, ld3(sPos, j), va, ld3(sVel, j), f, &fj)) cs = add(cs, fj);
                    }
                    F = add(F, cs);
                }
        if (!(clamp && ra < 0)) F = add(F, wall_force(pa, va, f));
        dem_write(s, f, (int)sIdx[i], F, va);
    }
}

// Static scene DEM (FindAndHandleAugmentedRealityCollisionsUniformGridKernel, UniformGridDEM.cu:306-378)
// for the body particles: the 27 cells of the static points' grid in z, y, x order, each cell's
// contacts summed into a partial that is added to the total for every cell (empty ones too, as the
// reference does), then pForce += total and the particle torque becomes relativePos x pForce.
// Static points come sorted by cell with their normals (ssPos / ssNorm) and the static cellStart /
// cellEnd of the same grid.
// ref = true: the stage as the reference wrapper runs it: every particle, pForce += F and
// pTorque += relativePos x F.
template <class T>
void dem_static(State<T> &s, const Phys<T> &f, const ArSimParams *p, const uint32_t *cellStart,
                const uint32_t *cellEnd, const float *ssPos, const float *ssNorm, bool ref = false) {
    for (int i = 0; i < s.N; ++i) {
        if (!ref && s.rb[i] < 0) continue;
        V3<T> pa = ld3(s.pos, i), va = ld3(s.vel, i);
        int gx = (int)std::floor(((float)pa.x - p->worldOrigin.x) / p->cellSize.x);
        int gy = (int)std::floor(((float)pa.y - p->worldOrigin.y) / p->cellSize.y);
        int gz = (int)std::floor(((float)pa.z - p->worldOrigin.z) / p->cellSize.z);
        V3<T> F = mk<T>(0, 0, 0);
        for (int z = -1; z <= 1; ++z)
            for (int y = -1; y <= 1; ++y)
                for (int x = -1; x <= 1; ++x) {
                    uint32_t h = grid_hash_cell(gx + x, gy + y, gz + z, p);
                    V3<T> cs = mk<T>(0, 0, 0);
                    uint32_t st = cellStart[h];
                    if (st != 0xffffffffu) {
                        for (uint32_t j = st; j < cellEnd[h]; ++j) {
                            V3<T> ps = mk<T>((T)ssPos[4 * j], (T)ssPos[4 * j + 1], (T)ssPos[4 * j + 2]);
                            V3<T> ns = mk<T>((T)ssNorm[4 * j], (T)ssNorm[4 * j + 1], (T)ssNorm[4 * j + 2]);
                            V3<T> fj;
                            if (static_force(pa, ps, ns, va, f, &fj)) cs = add(cs, fj);
                        }
                    }
                    F = add(F, cs);
                }
        V3<T> tot = add(ld3(s.pForce, i), F);
        st4(s.pForce, i, tot, T(0));
        if (ref) st4(s.pTorque, i, add(ld3(s.pTorque, i), cross(ld3(s.relPos, i), F)), T(0));
        else st4(s.pTorque, i, cross(ld3(s.relPos, i), tot), T(0));
    }
}

// LBVH traversal DEM (collideBVHSoARigidBody, RB_BVHKernels.cu:100-244): stack traversal from the
// root, left child tested before right, the right child deferred when both are entered.
template <class T>
void dem_lbvh(State<T> &s, const Phys<T> &f, int n, const int *sIdx, const int *left, const int *right,
              const float *bounds8, const T *cms, const T *sVel, const int *sRb) {
    bool clamp = (f.flags & ARSIM_COMPAT_GRAIN_CLAMP) != 0;
    std::vector<int> stack;
    for (int i = 0; i < n; ++i) {
        V3<T> pa = ld3(cms, i), va = ld3(sVel, i);
        int ra = sRb[i];
        const float *qb = bounds8 + 8 * (long)i;
        V3<T> F = mk<T>(0, 0, 0);
        if (n > 1) {
            stack.clear();
            int node = n;
            while (true) {
                int L = left[node], R = right[node];
                bool oL = overlap(qb, bounds8 + 8 * (long)L) && L != i;
                bool oR = overlap(qb, bounds8 + 8 * (long)R) && R != i;
                bool lL = L < n, lR = R < n;
                V3<T> fj;
                if (oL && lL && pair_allowed(ra, sRb[L]) && pair_force(pa, ld3(cms, L), va, ld3(sVel, L), f, &fj))
                    F = add(F, fj);
                if (oR && lR && pair_allowed(ra, sRb[R]) && pair_force(pa, ld3(cms, R), va, ld3(sVel, R), f, &fj))
                    F = add(F, fj);
                bool tL = oL && !lL, tR = oR && !lR;
                if (!tL && !tR) {
                    if (stack.empty()) break;
                    node = stack.back();
                    stack.pop_back();
                } else {
                    node = tL ? L : R;
                    if (tL && tR) stack.push_back(R);
                }
            }
        }
        if (!(clamp && ra < 0)) F = add(F, wall_force(pa, va, f));
        dem_write(s, f, sIdx[i], F, va);
    }
}

// Segmented per-body reduction + apply (intended sum of UniformGridDEM.cu:421-460).  The FP32
// mirror follows the device order: 64 lane partials (lane l takes particles l, l+64, ...),
// then a butterfly over lane distance 32, 16, 8, 4, 2, 1; lane 0 holds the total.
template <class T> void reduce_apply(State<T> &s, const Phys<T> &f) {
    bool ref = (f.flags & ARSIM_COMPAT_REF_UNITS) != 0;
    bool dropFirst = (f.flags & ARSIM_COMPAT_DROP_FIRST) != 0;
    for (int b = 0; b < s.B; ++b) {
        V3<T> accF[64], accT[64];
        for (int l = 0; l < 64; ++l) { accF[l] = mk<T>(0, 0, 0); accT[l] = mk<T>(0, 0, 0); }
        int st = s.pstart[b], cnt = s.pcount[b];
        int j0 = (dropFirst && b > 0) ? 1 : 0;
        for (int j = j0; j < cnt; ++j) {
            int l = (j - j0) & 63;
            long o = st + j;
            accF[l] = add(accF[l], ld3(s.pForce, o));
            accT[l] = add(accT[l], ld3(s.pTorque, o));
        }
        for (int off = 32; off >= 1; off >>= 1) {
            V3<T> nF[64], nT[64];
            for (int l = 0; l < 64; ++l) { nF[l] = add(accF[l], accF[l ^ off]); nT[l] = add(accT[l], accT[l ^ off]); }
            for (int l = 0; l < 64; ++l) { accF[l] = nF[l]; accT[l] = nT[l]; }
        }
        V3<T> J = accF[0], tau = accT[0];
        T sc = ref ? T(1) : f.dt;
        T m = s.mass[b];
        V3<T> dv = mk<T>((J.x * sc) / m, (J.y * sc) / m, (J.z * sc) / m);
        V3<T> dw = mv(s.Iw + 9 * (long)b, mk<T>(tau.x * sc, tau.y * sc, tau.z * sc));
        s.rbVel[4 * b] += dv.x; s.rbVel[4 * b + 1] += dv.y; s.rbVel[4 * b + 2] += dv.z;
        s.rbAng[4 * b] += dw.x; s.rbAng[4 * b + 1] += dw.y; s.rbAng[4 * b + 2] += dw.z;
    }
}

/* ===================================================================== sequential impulses (SIS)
 * Contact model of updateUniformGridSIS (RigidBodies/UniformGridUpdate.cpp:734-816): after the body
 * integration, the deepest contact of every particle against the static scene points (if any) and
 * against other bodies' particles, one contact per particle, gathered in body / particle order
 * (static contacts first), then 18 Gauss-Seidel sweeps of sequential impulses over the list. */

// One contact of the list (ContactNormal / ContactPoint / ContactRigidBody_1/2 / ContactBias /
// ContactAccumulated* of Particles/particleSystem.h:606-613).
constexpr int kSisIterations = 18;   // iterations, updateFunctions.cpp:1598

template <class T> struct SisContact {
    T n[4], pt[4];
    int b1, b2;                 // b2 = -1: static contact
    T bias;
    T acc, accF1, accF2;
};

// FindRigidBodyCollisionsUniformGridKernel + ...Cell (Kernels/RB_UniformGridKernels.cu:730-845):
// per sorted particle, the 27 cells in z, y, x order, each cell's particles in sorted order; a
// candidate of another body (grains all share -1) whose penetration 2r - |p - q| exceeds the best
// so far (starting at dist[orig], cleared to 0 by the caller) becomes the contact.
template <class T>
void sis_find_deepest(int N, const ArSimParams *p, const uint32_t *sIdx, const uint32_t *cellStart,
                      const uint32_t *cellEnd, const T *sPos, const int *rb, T *dist, int *partnerBody,
                      int *partnerParticle) {
    const T thr = T(2 * p->particleRadius);
    for (int i = 0; i < N; ++i) {
        V3<T> pa = ld3(sPos, i);
        const uint32_t orig = sIdx[i];
        const int ra = rb[orig];
        int gx = (int)std::floor(((float)pa.x - p->worldOrigin.x) / p->cellSize.x);
        int gy = (int)std::floor(((float)pa.y - p->worldOrigin.y) / p->cellSize.y);
        int gz = (int)std::floor(((float)pa.z - p->worldOrigin.z) / p->cellSize.z);
        for (int z = -1; z <= 1; ++z)
            for (int y = -1; y <= 1; ++y)
                for (int x = -1; x <= 1; ++x) {
                    uint32_t h = grid_hash_cell(gx + x, gy + y, gz + z, p);
                    uint32_t st = cellStart[h];
                    if (st == 0xffffffffu) continue;
                    T best = dist[orig];
                    for (uint32_t j = st; j < cellEnd[h]; ++j) {
                        const int o2 = (int)sIdx[j];
                        const int r2 = rb[o2];
                        if ((int)j == i || ra == r2) continue;
                        const V3<T> d = sub(pa, ld3(sPos, j));
                        const T len = vsqrt(dot(d, d));
                        if (thr - len > best) {
                            best = thr - len;
                            dist[orig] = best;
                            partnerParticle[orig] = o2;
                            partnerBody[orig] = r2;
                        }
                    }
                }
    }
}

// FindAugmentedRealityCollisionsUniformGridKernel + ...Cell (RB_UniformGridKernels.cu:879-1022):
// the same search over the static points (sorted by cell, cellStart / cellEnd of the static grid);
// the contact records the static point's original index.
template <class T>
void sis_find_static(int N, const ArSimParams *p, const uint32_t *sIdx, const T *sPos, const uint32_t *stIdx,
                     const uint32_t *cellStart, const uint32_t *cellEnd, const float *ssPos, T *dist,
                     int *partner) {
    const T thr = T(2 * p->particleRadius);
    for (int i = 0; i < N; ++i) {
        V3<T> pa = ld3(sPos, i);
        const uint32_t orig = sIdx[i];
        int gx = (int)std::floor(((float)pa.x - p->worldOrigin.x) / p->cellSize.x);
        int gy = (int)std::floor(((float)pa.y - p->worldOrigin.y) / p->cellSize.y);
        int gz = (int)std::floor(((float)pa.z - p->worldOrigin.z) / p->cellSize.z);
        for (int z = -1; z <= 1; ++z)
            for (int y = -1; y <= 1; ++y)
                for (int x = -1; x <= 1; ++x) {
                    uint32_t h = grid_hash_cell(gx + x, gy + y, gz + z, p);
                    uint32_t st = cellStart[h];
                    if (st == 0xffffffffu) continue;
                    T best = dist[orig];
                    for (uint32_t j = st; j < cellEnd[h]; ++j) {
                        const V3<T> q = mk<T>((T)ssPos[4 * j], (T)ssPos[4 * j + 1], (T)ssPos[4 * j + 2]);
                        const V3<T> d = sub(pa, q);
                        const T len = vsqrt(dot(d, d));
                        if (thr - len > best) {
                            best = thr - len;
                            dist[orig] = best;
                            partner[orig] = (int)stIdx[j];
                        }
                    }
                }
    }
}

template <class T> inline T dot4(const T *a, const T *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3]; }

// GatherAugmentedRealityCollisions (Particles/updateFunctions.cpp:1521-1578): for every body
// particle (bodies in order, their particles in order) with a static contact: normal = the static
// point's normal, point = relativePos + CM, bias = ARrestitution * (v + w x r) . n.
template <class T>
void sis_gather_static(const State<T> &s, const ArSimParams *p, const T *dist, const int *partner,
                       const float *staticNorm, std::vector<SisContact<T>> &out) {
    const T eps = (T)p->ARrestitution;
    for (int b = 0; b < s.B; ++b)
        for (int k = 0; k < s.pcount[b]; ++k) {
            const long i = (long)s.pstart[b] + k;
            if (!(dist[i] > T(0))) continue;
            SisContact<T> c{};
            const long q = partner[i];
            for (int a = 0; a < 4; ++a) {
                c.n[a] = (T)staticNorm[4 * q + a];
                c.pt[a] = s.relPos[4 * i + a] + s.rbPos[4 * b + a];
            }
            c.b1 = b; c.b2 = -1;
            const V3<T> r = ld3(s.relPos, i), n = mk<T>(c.n[0], c.n[1], c.n[2]);
            const T vrel = dot(add(ld3(s.rbVel, b), cross(ld3(s.rbAng, b), r)), n);
            c.bias = eps * vrel;
            out.push_back(c);
        }
}

// GatherRigidBodyCollisions (updateFunctions.cpp:1431-1519) with findExactContactPoint
// (Particles/CPU_Engine.cpp:461-473): for every body particle with a contact, p1 = CM1 + r1,
// p2 = CM2 + r2 (float4, w included), point = p1 + (p2 - p1) * (r / (r + r)),
// normal = normalize(p2 - p1) * -1 (normalize = v * (1 / sqrt(v . v)) over four components),
// bias = RBrestitution * ((v1 + w1 x r1) . n - (v2 + w2 x r2) . n).  A partner that is a grain
// (body -1) has no body state; such contacts are left out.
template <class T>
void sis_gather_bodies(const State<T> &s, const ArSimParams *p, const T *dist, const int *pBody, const int *pPart,
                       std::vector<SisContact<T>> &out) {
    const T eps = (T)p->RBrestitution;
    const T r = (T)p->particleRadius;
    const T t = r / (r + r);
    for (int b = 0; b < s.B; ++b)
        for (int k = 0; k < s.pcount[b]; ++k) {
            const long i = (long)s.pstart[b] + k;
            if (!(dist[i] > T(0))) continue;
            const int b2 = pBody[i];
            if (b2 < 0) continue;
            const long j = pPart[i];
            SisContact<T> c{};
            T p1[4], p2[4], d[4];
            for (int a = 0; a < 4; ++a) {
                p1[a] = s.rbPos[4 * b + a] + s.relPos[4 * i + a];
                p2[a] = s.rbPos[4 * (long)b2 + a] + s.relPos[4 * j + a];
                d[a] = p2[a] - p1[a];
            }
            const T inv = T(1) / vsqrt(dot4(d, d));
            for (int a = 0; a < 4; ++a) {
                c.pt[a] = p1[a] + d[a] * t;
                c.n[a] = (d[a] * inv) * T(-1);
            }
            c.b1 = b; c.b2 = b2;
            const V3<T> n = mk<T>(c.n[0], c.n[1], c.n[2]);
            const T v1 = dot(add(ld3(s.rbVel, b), cross(ld3(s.rbAng, b), ld3(s.relPos, i))), n);
            const T v2 = dot(add(ld3(s.rbVel, b2), cross(ld3(s.rbAng, b2), ld3(s.relPos, j))), n);
            c.bias = eps * (v1 - v2);
            out.push_back(c);
        }
}

// glm::normalize(vec3) of GLM 0.9.7: v * inversesqrt(v . v), inversesqrt(x) = 1 / sqrt(x).
template <class T> inline V3<T> gnormalize(V3<T> v) {
    const T inv = T(1) / vsqrt(dot(v, v));
    return mk<T>(v.x * inv, v.y * inv, v.z * inv);
}
// Effective mass term (I^-1 (p x d)) x p . d of SequentialImpulseSolver.
template <class T> inline T sis_k(const T *Iinv, V3<T> p, V3<T> d) { return dot(cross(mv(Iinv, cross(p, d)), p), d); }

// One contact of one sweep: the two branches of SequentialImpulseSolver
// (Particles/updateFunctions.cpp:1580-1874).
//  static contact (b2 == -1, :1603-1713): normal impulse clamped to [0, 100] accumulated, two
//    friction impulses along t = normalize((1,0,0) - ((1,0,0).n) n) (or (0,1,0) when n.x is within
//    1e-4 of +-1) and t2 = normalize(t x n), each clamped to +-ARfriction * mc * |g.y| accumulated;
//    v += J / m, w += I^-1 (p x J) with J = jn n + jt t + jt2 t2;
//  body contact (:1714-1860): the same normal impulse with both bodies' terms, applied with
//    opposite signs; its friction impulses are computed but not applied (impulse_vector =
//    normal_impulse, :1784), so they have no effect and are not formed here.
// Velocities are stored with w = 0, as make_float4(v.x, v.y, v.z, 0) does.
template <class T>
void sis_contact(SisContact<T> &c, const T *rbPos, T *rbVel, T *rbAng, const T *Iw, const T *mass,
                 const ArSimParams *p) {
    const V3<T> n = mk<T>(c.n[0], c.n[1], c.n[2]);
    const T upper = T(100);
    if (c.b2 < 0) {
        const long b = c.b1;
        const V3<T> pp = mk<T>(c.pt[0] - rbPos[4 * b], c.pt[1] - rbPos[4 * b + 1], c.pt[2] - rbPos[4 * b + 2]);
        const T *I = Iw + 9 * b;
        const T m = mass[b];
        V3<T> v = ld3(rbVel, b), w = ld3(rbAng, b);
        T mc = T(1) / m + sis_k(I, pp, n);
        if (std::fabs((double)mc) < 0.00001) mc = T(1);
        const T vrel = dot(add(v, cross(w, pp)), n);
        T ci = -(vrel + c.bias) / mc;
        T tmp = c.acc + ci;
        if (tmp < T(0)) tmp = T(0);
        else if (tmp > upper) tmp = upper;
        ci = tmp - c.acc;
        c.acc = tmp;
        const T fb = (T)p->ARfriction * mc * std::fabs((T)p->gravity.y);
        const V3<T> vel = add(v, cross(w, pp));
        V3<T> t = mk<T>(T(1), T(0), T(0));
        if (std::fabs((double)(t.x - n.x)) < 0.0001) t = mk<T>(T(0), T(1), T(0));
        else if (std::fabs((double)(t.x + n.x)) < 0.0001) t = mk<T>(T(0), T(1), T(0));
        t = sub(t, scl(dot(t, n), n));
        t = gnormalize(t);
        const T kt = T(1) / m + sis_k(I, pp, t);
        T cf = -dot(vel, t) / kt;
        T tf = c.accF1 + cf;
        if (tf < -fb) tf = -fb;
        else if (tf > fb) tf = fb;
        cf = tf - c.accF1;
        c.accF1 = tf;
        const V3<T> t2 = gnormalize(cross(t, n));
        const T kt2 = T(1) / m + sis_k(I, pp, t2);
        T cf2 = -dot(vel, t2) / kt2;
        T tf2 = c.accF2 + cf2;
        if (tf2 < -fb) tf2 = -fb;
        else if (tf2 > fb) tf2 = fb;
        cf2 = tf2 - c.accF2;
        c.accF2 = tf2;
        const V3<T> J = add(add(scl(ci, n), scl(cf, t)), scl(cf2, t2));
        v = add(v, divs(J, m));
        w = add(w, mv(I, cross(pp, J)));
        st4(rbVel, b, v, T(0));
        st4(rbAng, b, w, T(0));
    } else {
        const long b1 = c.b1, b2 = c.b2;
        const V3<T> p1 = mk<T>(c.pt[0] - rbPos[4 * b1], c.pt[1] - rbPos[4 * b1 + 1], c.pt[2] - rbPos[4 * b1 + 2]);
        const V3<T> p2 = mk<T>(c.pt[0] - rbPos[4 * b2], c.pt[1] - rbPos[4 * b2 + 1], c.pt[2] - rbPos[4 * b2 + 2]);
        const T *I1 = Iw + 9 * b1, *I2 = Iw + 9 * b2;
        const T m1 = mass[b1], m2 = mass[b2];
        V3<T> v1 = ld3(rbVel, b1), w1 = ld3(rbAng, b1), v2 = ld3(rbVel, b2), w2 = ld3(rbAng, b2);
        T mc = T(1) / m1 + T(1) / m2 + sis_k(I1, p1, n) + sis_k(I2, p2, n);
        if (std::fabs((double)mc) < 0.00001) mc = T(1);
        const T vrel = dot(add(v1, cross(w1, p1)), n) - dot(add(v2, cross(w2, p2)), n);
        T ci = -(vrel + c.bias) / mc;
        T tmp = c.acc + ci;
        if (tmp < T(0)) tmp = T(0);
        else if (tmp > upper) tmp = upper;
        ci = tmp - c.acc;
        c.acc = tmp;
        const V3<T> J = scl(ci, n);
        v1 = add(v1, divs(J, m1));
        w1 = add(w1, mv(I1, cross(p1, J)));
        st4(rbVel, b1, v1, T(0));
        st4(rbAng, b1, w1, T(0));
        v2 = sub(v2, divs(J, m2));
        w2 = sub(w2, mv(I2, cross(p2, J)));
        st4(rbVel, b2, v2, T(0));
        st4(rbAng, b2, w2, T(0));
    }
}

template <class T>
void sis_solve(std::vector<SisContact<T>> &cs, const T *rbPos, T *rbVel, T *rbAng, const T *Iw, const T *mass,
               const ArSimParams *p, int iterations) {
    for (int k = 0; k < iterations; ++k)
        for (auto &c : cs) sis_contact(c, rbPos, rbVel, rbAng, Iw, mass, p);
}

// One SIS step (updateUniformGridSIS, RigidBodies/UniformGridUpdate.cpp:734-797): integrate the
// bodies and pose their particles (Integrate_Rigid_Body_System_GPU), static contacts then body
// contacts, the solver, and the particles' pose from the solved velocities (the
// computeGlobalAttributesWrapper call at the end of SequentialImpulseSolver).
template <class T> struct SisScratch {
    std::vector<T> distA, distB;
    std::vector<int> partA, bodyB, partB;
    std::vector<SisContact<T>> contacts;
};

template <class T> struct Scratch {
    std::vector<uint32_t> hash, idx, sh, si, cs, ce;
    std::vector<uint64_t> k64, sk64;
    std::vector<uint32_t> k32, sk32;
    std::vector<T> sPos, sVel;
    std::vector<int> sRb, sIdxI, left, right, parent, minR, maxR;
    std::vector<unsigned char> isLeaf;
    std::vector<float> bounds, cmsF, posF;
    std::vector<T> cms;
    std::vector<uint32_t> stH, stI, stSH, stSI, stCs, stCe;
    std::vector<float> stPos, stNorm;
};

template <class T>
int step_impl(State<T> &s, const ArSimParams *p, const ArSimStepConfig *cfg, int nsteps) {
    Phys<T> f = mkphys<T>(p, cfg);
    Scratch<T> w;
    SisScratch<T> sw;
    const int N = s.N;
    const bool sis = (cfg->flags & ARSIM_CONTACT_SIS) != 0;
    if (sis) {
        // the SIS model is defined for bodies on the grid broadphase
        if (cfg->method != ARSIM_METHOD_GRID) return -1;
        for (int i = 0; i < N; ++i)
            if (s.rb[i] < 0) return -1;
    }
    for (int it = 0; it < nsteps; ++it) {
        // stage order of updateUniformGridDEM (RigidBodies/UniformGridUpdate.cpp:939-999)
        integrate_bodies(s, f);
        pose_and_grains(s, f);
        for (long i = 0; i < 4l * N; ++i) { s.pForce[i] = T(0); s.pTorque[i] = T(0); }
        // the broadphase keys are computed from the FP32 rounding of the positions
        w.posF.resize(4 * (size_t)N);
        for (long i = 0; i < 4l * N; ++i) w.posF[i] = (float)s.pos[i];
        if (cfg->method == ARSIM_METHOD_GRID) {
            w.hash.resize(N); w.idx.resize(N); w.sh.resize(N); w.si.resize(N);
            orc_calc_hash(w.posF.data(), N, p, w.hash.data(), w.idx.data());
            orc_sort_pairs_u32(w.hash.data(), w.idx.data(), N, 0, 32, w.sh.data(), w.si.data());
            w.cs.resize(p->numCells); w.ce.resize(p->numCells);
            orc_cell_bounds(w.sh.data(), N, p->numCells, w.cs.data(), w.ce.data());
            w.sPos.resize(4 * (size_t)N); w.sVel.resize(4 * (size_t)N); w.sRb.resize(N);
            for (int i = 0; i < N; ++i) {
                long o = w.si[i];
                for (int c = 0; c < 4; ++c) { w.sPos[4 * i + c] = s.pos[4 * o + c]; w.sVel[4 * i + c] = s.vel[4 * o + c]; }
                w.sRb[i] = s.rb[o];
            }
            if (!sis)
                dem_grid(s, f, p, w.si.data(), w.cs.data(), w.ce.data(), w.sPos.data(), w.sVel.data(), w.sRb.data());
        } else {
            float bmin[3] = {cfg->boxMin[0], cfg->boxMin[1], cfg->boxMin[2]};
            float bmax[3] = {cfg->boxMax[0], cfg->boxMax[1], cfg->boxMax[2]};
            if (cfg->flags & ARSIM_DYNAMIC_AABB) {
                for (int c = 0; c < 3; ++c) { bmin[c] = INFINITY; bmax[c] = -INFINITY; }
                for (int i = 0; i < N; ++i)
                    for (int c = 0; c < 3; ++c) {
                        float v = w.posF[4 * i + c];
                        bmin[c] = v < bmin[c] ? v : bmin[c];
                        bmax[c] = v > bmax[c] ? v : bmax[c];
                    }
            }
            w.idx.resize(N); w.si.resize(N);
            for (int i = 0; i < N; ++i) w.idx[i] = (uint32_t)i;
            w.left.assign(2 * (size_t)N, -1); w.right.assign(2 * (size_t)N, -1); w.parent.assign(2 * (size_t)N, -1);
            w.minR.assign(2 * (size_t)N, 0); w.maxR.assign(2 * (size_t)N, 0); w.isLeaf.assign(2 * (size_t)N, 0);
            if (cfg->method == ARSIM_METHOD_LBVH30) {
                w.k32.resize(N); w.sk32.resize(N);
                orc_morton30(w.posF.data(), N, bmin, bmax, w.k32.data());
                orc_sort_pairs_u32(w.k32.data(), w.idx.data(), N, 0, 32, w.sk32.data(), w.si.data());
                karras<uint32_t>(w.sk32.data(), N, nullptr, w.left.data(), w.right.data(), w.parent.data(),
                                 w.minR.data(), w.maxR.data(), w.isLeaf.data());
            } else {
                w.k64.resize(N); w.sk64.resize(N);
                orc_morton63(w.posF.data(), N, bmin, bmax, w.k64.data());
                orc_sort_pairs_u64(w.k64.data(), w.idx.data(), N, 0, 64, w.sk64.data(), w.si.data());
                karras<uint64_t>(w.sk64.data(), N, nullptr, w.left.data(), w.right.data(), w.parent.data(),
                                 w.minR.data(), w.maxR.data(), w.isLeaf.data());
            }
            w.sIdxI.resize(N);
            for (int i = 0; i < N; ++i) w.sIdxI[i] = (int)w.si[i];
            w.bounds.assign(8 * (size_t)(2 * N), 0.f); w.cmsF.resize(4 * (size_t)N);
            refit(w.sIdxI.data(), w.posF.data(), p->particleRadius, N, w.left.data(), w.right.data(),
                  w.cmsF.data(), w.bounds.data());
            w.cms.resize(4 * (size_t)N); w.sVel.resize(4 * (size_t)N); w.sRb.resize(N);
            for (int i = 0; i < N; ++i) {
                long o = w.si[i];
                for (int c = 0; c < 4; ++c) { w.cms[4 * i + c] = s.pos[4 * o + c]; w.sVel[4 * i + c] = s.vel[4 * o + c]; }
                w.sRb[i] = s.rb[o];
            }
            dem_lbvh(s, f, N, w.sIdxI.data(), w.left.data(), w.right.data(), w.bounds.data(), w.cms.data(),
                     w.sVel.data(), w.sRb.data());
        }
        if (s.numStatic > 0) {
            // the static points' grid index (rebuilt every step, as UniformGridUpdate.cpp:880-912 does)
            const int M = s.numStatic;
            w.stH.resize(M); w.stI.resize(M); w.stSH.resize(M); w.stSI.resize(M);
            orc_calc_hash(s.staticPos, M, p, w.stH.data(), w.stI.data());
            orc_sort_pairs_u32(w.stH.data(), w.stI.data(), M, 0, 32, w.stSH.data(), w.stSI.data());
            w.stCs.resize(p->numCells); w.stCe.resize(p->numCells);
            orc_cell_bounds(w.stSH.data(), M, p->numCells, w.stCs.data(), w.stCe.data());
            w.stPos.resize(4 * (size_t)M); w.stNorm.resize(4 * (size_t)M);
            for (int j = 0; j < M; ++j)
                for (int c = 0; c < 4; ++c) {
                    w.stPos[4 * j + c] = s.staticPos[4 * (size_t)w.stSI[j] + c];
                    w.stNorm[4 * j + c] = s.staticNorm[4 * (size_t)w.stSI[j] + c];
                }
            if (!sis) dem_static(s, f, p, w.stCs.data(), w.stCe.data(), w.stPos.data(), w.stNorm.data());
        }
        if (!sis) {
            reduce_apply(s, f);
            continue;
        }
        sw.contacts.clear();
        if (s.numStatic > 0) {
            sw.distA.assign(N, T(0));
            sw.partA.assign(N, -1);
            sis_find_static(N, p, w.si.data(), w.sPos.data(), w.stSI.data(), w.stCs.data(), w.stCe.data(),
                            w.stPos.data(), sw.distA.data(), sw.partA.data());
            sis_gather_static(s, p, sw.distA.data(), sw.partA.data(), s.staticNorm, sw.contacts);
        }
        sw.distB.assign(N, T(0));
        sw.bodyB.assign(N, -1);
        sw.partB.assign(N, -1);
        sis_find_deepest(N, p, w.si.data(), w.cs.data(), w.ce.data(), w.sPos.data(), s.rb, sw.distB.data(),
                         sw.bodyB.data(), sw.partB.data());
        sis_gather_bodies(s, p, sw.distB.data(), sw.bodyB.data(), sw.partB.data(), sw.contacts);
        sis_solve(sw.contacts, s.rbPos, s.rbVel, s.rbAng, s.Iw, s.mass, p, kSisIterations);
        pose_and_grains(s, f);
    }
    return 0;
}

}  // namespace

extern "C" {

typedef struct OrcStateF {
    int N, B;
    float *pos, *vel, *relPos, *bodyFrame, *pForce, *pTorque;
    const int *rb;
    float *rbPos, *rbVel, *rbAng, *qinc, *qcum;
    const float *Ib;
    float *Iw;
    const float *mass;
    const int *pstart, *pcount;
    float *model;
    const float *staticPos, *staticNorm;
    int numStatic;
} OrcStateF;
typedef struct OrcStateD {
    int N, B;
    double *pos, *vel, *relPos, *bodyFrame, *pForce, *pTorque;
    const int *rb;
    double *rbPos, *rbVel, *rbAng, *qinc, *qcum;
    const double *Ib;
    double *Iw;
    const double *mass;
    const int *pstart, *pcount;
    double *model;
    const float *staticPos, *staticNorm;
    int numStatic;
} OrcStateD;

}  // extern "C"

template <class T, class S> static State<T> to_state(S *o) {
    State<T> s;
    s.N = o->N; s.B = o->B;
    s.pos = o->pos; s.vel = o->vel; s.relPos = o->relPos; s.bodyFrame = o->bodyFrame;
    s.pForce = o->pForce; s.pTorque = o->pTorque; s.rb = o->rb;
    s.rbPos = o->rbPos; s.rbVel = o->rbVel; s.rbAng = o->rbAng; s.qinc = o->qinc; s.qcum = o->qcum;
    s.Ib = o->Ib; s.Iw = o->Iw; s.mass = o->mass; s.pstart = o->pstart; s.pcount = o->pcount;
    s.model = o->model;
    s.staticPos = o->staticPos; s.staticNorm = o->staticNorm; s.numStatic = o->numStatic;
    return s;
}

/* `iterations` sweeps over the given contacts; acc4 holds (normal, friction, friction 2, 0)
 * accumulators in and out. */
template <class T, class S>
static int sis_solve_export(S *o, const ArSimParams *p, const float *n4, const float *pt4, const int *b12,
                            const float *bias, float *acc4, long count, int iterations) {
    State<T> s = to_state<T>(o);
    std::vector<SisContact<T>> cs((size_t)count);
    for (long c = 0; c < count; ++c) {
        for (int a = 0; a < 4; ++a) { cs[c].n[a] = (T)n4[4 * c + a]; cs[c].pt[a] = (T)pt4[4 * c + a]; }
        cs[c].b1 = b12[2 * c]; cs[c].b2 = b12[2 * c + 1];
        cs[c].bias = (T)bias[c];
        cs[c].acc = (T)acc4[4 * c]; cs[c].accF1 = (T)acc4[4 * c + 1]; cs[c].accF2 = (T)acc4[4 * c + 2];
    }
    sis_solve(cs, s.rbPos, s.rbVel, s.rbAng, s.Iw, s.mass, p, iterations);
    for (long c = 0; c < count; ++c) {
        acc4[4 * c] = (float)cs[c].acc; acc4[4 * c + 1] = (float)cs[c].accF1; acc4[4 * c + 2] = (float)cs[c].accF2;
        acc4[4 * c + 3] = 0.f;
    }
    return 0;
}

extern "C" {


/* Full step(s): integrate bodies, pose + grains, broadphase (cfg->method), DEM, reduce. */
int orc_step_f32(OrcStateF *o, const ArSimParams *p, const ArSimStepConfig *cfg, int nsteps) {
    State<float> s = to_state<float>(o);
    return step_impl<float>(s, p, cfg, nsteps);
}
int orc_step_f64(OrcStateD *o, const ArSimParams *p, const ArSimStepConfig *cfg, int nsteps) {
    State<double> s = to_state<double>(o);
    return step_impl<double>(s, p, cfg, nsteps);
}

/* Individual stages on a state (for stage-level parity from identical inputs). */
int orc_integrate_bodies_f32(OrcStateF *o, const ArSimParams *p, const ArSimStepConfig *cfg) {
    State<float> s = to_state<float>(o); Phys<float> f = mkphys<float>(p, cfg); integrate_bodies(s, f); return 0;
}
int orc_integrate_bodies_f64(OrcStateD *o, const ArSimParams *p, const ArSimStepConfig *cfg) {
    State<double> s = to_state<double>(o); Phys<double> f = mkphys<double>(p, cfg); integrate_bodies(s, f); return 0;
}
int orc_pose_grains_f32(OrcStateF *o, const ArSimParams *p, const ArSimStepConfig *cfg) {
    State<float> s = to_state<float>(o); Phys<float> f = mkphys<float>(p, cfg); pose_and_grains(s, f); return 0;
}
int orc_pose_grains_f64(OrcStateD *o, const ArSimParams *p, const ArSimStepConfig *cfg) {
    State<double> s = to_state<double>(o); Phys<double> f = mkphys<double>(p, cfg); pose_and_grains(s, f); return 0;
}
int orc_reduce_apply_f32(OrcStateF *o, const ArSimParams *p, const ArSimStepConfig *cfg) {
    State<float> s = to_state<float>(o); Phys<float> f = mkphys<float>(p, cfg); reduce_apply(s, f); return 0;
}
int orc_reduce_apply_f64(OrcStateD *o, const ArSimParams *p, const ArSimStepConfig *cfg) {
    State<double> s = to_state<double>(o); Phys<double> f = mkphys<double>(p, cfg); reduce_apply(s, f); return 0;
}
/* DEM from given sorted grid arrays (sPos w ignored; sRb explicit). */
int orc_dem_grid_f32(OrcStateF *o, const ArSimParams *p, const ArSimStepConfig *cfg, const uint32_t *sIdx,
                     const uint32_t *cellStart, const uint32_t *cellEnd, const float *sPos, const float *sVel,
                     const int *sRb) {
    State<float> s = to_state<float>(o); Phys<float> f = mkphys<float>(p, cfg);
    dem_grid(s, f, p, sIdx, cellStart, cellEnd, sPos, sVel, sRb); return 0;
}
int orc_dem_grid_f64(OrcStateD *o, const ArSimParams *p, const ArSimStepConfig *cfg, const uint32_t *sIdx,
                     const uint32_t *cellStart, const uint32_t *cellEnd, const double *sPos, const double *sVel,
                     const int *sRb) {
    State<double> s = to_state<double>(o); Phys<double> f = mkphys<double>(p, cfg);
    dem_grid(s, f, p, sIdx, cellStart, cellEnd, sPos, sVel, sRb); return 0;
}
int orc_dem_lbvh_f32(OrcStateF *o, const ArSimParams *p, const ArSimStepConfig *cfg, int n, const int *sIdx,
                     const int *left, const int *right, const float *bounds8, const float *cms, const float *sVel,
                     const int *sRb) {
    State<float> s = to_state<float>(o); Phys<float> f = mkphys<float>(p, cfg);
    dem_lbvh(s, f, n, sIdx, left, right, bounds8, cms, sVel, sRb); return 0;
}
int orc_dem_lbvh_f64(OrcStateD *o, const ArSimParams *p, const ArSimStepConfig *cfg, int n, const int *sIdx,
                     const int *left, const int *right, const float *bounds8, const double *cms, const double *sVel,
                     const int *sRb) {
    State<double> s = to_state<double>(o); Phys<double> f = mkphys<double>(p, cfg);
    dem_lbvh(s, f, n, sIdx, left, right, bounds8, cms, sVel, sRb); return 0;
}
/* Single pair / wall force, for closed-form checks. */
int orc_pair_force_f32(const float *pa, const float *pb, const float *va, const float *vb,
                       const ArSimParams *p, const ArSimStepConfig *cfg, float *out) {
    Phys<float> f = mkphys<float>(p, cfg);
    V3<float> F{0, 0, 0};
    bool hit = pair_force(mk(pa[0], pa[1], pa[2]), mk(pb[0], pb[1], pb[2]), mk(va[0], va[1], va[2]),
                          mk(vb[0], vb[1], vb[2]), f, &F);
    out[0] = F.x; out[1] = F.y; out[2] = F.z;
    return hit ? 1 : 0;
}
void orc_wall_force_f64(const double *p3, const double *v3, const ArSimParams *p, const ArSimStepConfig *cfg,
                        double *out) {
    Phys<double> f = mkphys<double>(p, cfg);
    V3<double> F = wall_force(mk(p3[0], p3[1], p3[2]), mk(v3[0], v3[1], v3[2]), f);
    out[0] = F.x; out[1] = F.y; out[2] = F.z;
}
void orc_sincos(const float *x, int n, float *s, float *c) {
    for (int i = 0; i < n; ++i) sincos_det(x[i], s + i, c + i);
}

/* Range image -> static scene points (RGBDloader, Kernels/customRangeImageLoader.cu:14-109) and
 * normals (computeNormals, :118-162), FP32 in the device kernel's operation order.
 *  - depth d = image / 5000 (d < 0.01 -> -1000), X = d (x - cx) / fx, Y = d (y - cy) / fy with the
 *    hard-coded intrinsics fx 517.3, fy 516.5, cx 318.6, cy 255.3;
 *  - p = transpose(cam) * (X, Y, d, 0), then the fixed gravity rotation G (columns (1,0,0),
 *    -g, (0, g.z, -g.y)) with g the normalised accelerometer vector (0.478957, -8.010560,
 *    4.166928); a 4x4 matrix-vector product sums (c0 v0 + c1 v1) + (c2 v2 + c3 v3) by columns, as
 *    GLM 0.9.7 does;
 *  - the point is (p.x, -p.y, -p.z); pixels with y*W + x > (W-1)(H-1) are not written (the
 *    reference's bound);
 *  - normal of an interior pixel = normalize(cross(P[y+1][x] - P[y-1][x], P[y][x+1] - P[y][x-1])),
 *    the border is not written; normalize = v * (1 / sqrt(v.v)).
 * cam is column-major.  vao (5 floats per pixel: position, x / W, y / H) and radii may be NULL.
 * The reference is built with nvcc's default multiply-add contraction, so its values can differ
 * from these (and from the device kernel's) in the last bits; g and the normalisation use a
 * correctly rounded square root where the reference uses an approximate reciprocal square root. */
void orc_range_image(const uint16_t *img, int W, int H, const float *cam, float radius, float *pos, float *norm,
                     float *radii, float *vao) {
    const float fx = 517.3f, fy = 516.5f, cx = 318.6f, cy = 255.3f;
    const double g0 = 0.478957, g1 = -8.010560, g2 = 4.166928;
    const double gl = std::sqrt(g0 * g0 + g1 * g1 + g2 * g2);
    const float gx = (float)(g0 / gl), gy = (float)(g1 / gl), gz = (float)(g2 / gl);
    const long bound = (long)(W - 1) * (long)(H - 1);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const long idx = (long)y * W + x;
            if (idx > bound) continue;
            float d = (float)img[idx];
            d = d / 5000.f;
            if (radii) radii[idx] = radius;
            if ((double)d < 0.01) d = -1000.f;   // compared in double, as the reference's literal is
            const float X = d * ((float)x - cx) / fx;
            const float Y = d * ((float)y - cy) / fy;
            const float v[4] = {X, Y, d, 0.f};
            float t[4];
            for (int r = 0; r < 4; ++r)   // transpose(cam): element (r, c) = cam[4 r + c]
                t[r] = (cam[4 * r + 0] * v[0] + cam[4 * r + 1] * v[1]) + (cam[4 * r + 2] * v[2] + cam[4 * r + 3] * v[3]);
            const float G[16] = {1.f, 0.f, 0.f, 0.f, -gx, -gy, -gz, 0.f, 0.f, gz, -gy, 0.f, 0.f, 0.f, 0.f, 1.f};
            float u[4];
            for (int r = 0; r < 4; ++r)
                u[r] = (G[r] * t[0] + G[4 + r] * t[1]) + (G[8 + r] * t[2] + G[12 + r] * t[3]);
            const float px = u[0], py = -u[1], pz = -u[2];
            if (vao) {
                vao[5 * idx] = px; vao[5 * idx + 1] = py; vao[5 * idx + 2] = pz;
                vao[5 * idx + 3] = (float)x / (float)W;
                vao[5 * idx + 4] = (float)y / (float)H;
            }
            pos[4 * idx] = px; pos[4 * idx + 1] = py; pos[4 * idx + 2] = pz; pos[4 * idx + 3] = 0.f;
        }
    for (int y = 1; y <= H - 2; ++y)
        for (int x = 1; x <= W - 2; ++x) {
            const long idx = (long)y * W + x;
            auto P = [&](int yy, int xx, int c) { return pos[4 * ((long)yy * W + xx) + c]; };
            const float dx[3] = {P(y, x + 1, 0) - P(y, x - 1, 0), P(y, x + 1, 1) - P(y, x - 1, 1), P(y, x + 1, 2) - P(y, x - 1, 2)};
            const float dy[3] = {P(y + 1, x, 0) - P(y - 1, x, 0), P(y + 1, x, 1) - P(y - 1, x, 1), P(y + 1, x, 2) - P(y - 1, x, 2)};
            const float n0 = dy[1] * dx[2] - dy[2] * dx[1];
            const float n1 = dy[2] * dx[0] - dy[0] * dx[2];
            const float n2 = dy[0] * dx[1] - dy[1] * dx[0];
            const float inv = 1.0f / std::sqrt(n0 * n0 + n1 * n1 + n2 * n2);
            norm[4 * idx] = n0 * inv; norm[4 * idx + 1] = n1 * inv; norm[4 * idx + 2] = n2 * inv; norm[4 * idx + 3] = 0.f;
        }
}

/* Sequential-impulse stages (deepest contacts, gather, solve) for stage-level checks. */
void orc_sis_find_deepest_f32(int N, const ArSimParams *p, const uint32_t *sIdx, const uint32_t *cellStart,
                              const uint32_t *cellEnd, const float *sPos, const int *rb, float *dist, int *pBody,
                              int *pPart) {
    sis_find_deepest<float>(N, p, sIdx, cellStart, cellEnd, sPos, rb, dist, pBody, pPart);
}
void orc_sis_find_static_f32(int N, const ArSimParams *p, const uint32_t *sIdx, const float *sPos,
                             const uint32_t *stIdx, const uint32_t *cellStart, const uint32_t *cellEnd,
                             const float *ssPos, float *dist, int *partner) {
    sis_find_static<float>(N, p, sIdx, sPos, stIdx, cellStart, cellEnd, ssPos, dist, partner);
}
/* Contact list (static contacts first, then body contacts) into n4 / pt4 (float4 each), b12 (two ints),
 * bias; returns the count, or -1 past cap.  distA / partA / staticNorm may be NULL (no static set). */
long orc_sis_gather_f32(OrcStateF *o, const ArSimParams *p, const float *distA, const int *partA,
                        const float *staticNorm, const float *distB, const int *bodyB, const int *partB, float *n4,
                        float *pt4, int *b12, float *bias, long cap) {
    State<float> s = to_state<float>(o);
    std::vector<SisContact<float>> cs;
    if (distA) sis_gather_static(s, p, distA, partA, staticNorm, cs);
    sis_gather_bodies(s, p, distB, bodyB, partB, cs);
    if ((long)cs.size() > cap) return -1;
    for (size_t c = 0; c < cs.size(); ++c) {
        for (int a = 0; a < 4; ++a) { n4[4 * c + a] = cs[c].n[a]; pt4[4 * c + a] = cs[c].pt[a]; }
        b12[2 * c] = cs[c].b1; b12[2 * c + 1] = cs[c].b2;
        bias[c] = cs[c].bias;
    }
    return (long)cs.size();
}
int orc_sis_solve_f32(OrcStateF *o, const ArSimParams *p, const float *n4, const float *pt4, const int *b12,
                      const float *bias, float *acc4, long count, int iterations) {
    return sis_solve_export<float>(o, p, n4, pt4, b12, bias, acc4, count, iterations);
}
int orc_sis_solve_f64(OrcStateD *o, const ArSimParams *p, const float *n4, const float *pt4, const int *b12,
                      const float *bias, float *acc4, long count, int iterations) {
    return sis_solve_export<double>(o, p, n4, pt4, b12, bias, acc4, count, iterations);
}

/* Static scene DEM as a stage on a state (sorted static points, their normals, static cell bounds). */
int orc_dem_static_f32(OrcStateF *o, const ArSimParams *p, const ArSimStepConfig *cfg, const uint32_t *cs,
                       const uint32_t *ce, const float *ssPos, const float *ssNorm, int ref) {
    State<float> s = to_state<float>(o); Phys<float> f = mkphys<float>(p, cfg);
    dem_static(s, f, p, cs, ce, ssPos, ssNorm, ref != 0); return 0;
}
int orc_dem_static_f64(OrcStateD *o, const ArSimParams *p, const ArSimStepConfig *cfg, const uint32_t *cs,
                       const uint32_t *ce, const float *ssPos, const float *ssNorm, int ref) {
    State<double> s = to_state<double>(o); Phys<double> f = mkphys<double>(p, cfg);
    dem_static(s, f, p, cs, ce, ssPos, ssNorm, ref != 0); return 0;
}

}  // extern "C"
