// scenegen.cpp -- seeded synthetic scenes in the reference's particle / body layout.
//
// After addRigidBody / initObject (RigidBodies/RigidBodyAddition.cpp:20-228, 1486-1903) and
// addNewSphere (Particles/sphereAddition.cpp:151-301): a body is a set of equal spheres with
// offsets from its centre of mass, its inverse inertia is the inverse of the point-mass sum
// Σ (|r|² I - r rᵀ) over its particles (RigidBodyAddition.cpp:1633-1643), body particles come
// first in body order, free grains (rbIndex -1) follow.  Body mass = particle count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <unordered_map>
#include <vector>

#include "../../include/arsim.h"
#include "host_common.hpp"
#include "scene_owned.hpp"

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {   // SplitMix64
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {   // Box-Muller
        double u1 = uni(), u2 = uni();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct BodyShape {
    std::vector<double> off;   // xyz per particle, centred
    double circum = 0.0;       // max |offset| + r
};

BodyShape cube_shape(int edge, double r) {
    BodyShape s;
    double h = (edge - 1) * 0.5;
    for (int i = 0; i < edge; ++i)
        for (int j = 0; j < edge; ++j)
            for (int k = 0; k < edge; ++k) {
                s.off.push_back((i - h) * 2.0 * r);
                s.off.push_back((j - h) * 2.0 * r);
                s.off.push_back((k - h) * 2.0 * r);
            }
    s.circum = std::sqrt(3.0) * h * 2.0 * r + r;
    return s;
}

// Voxelised ellipsoid: lattice points of spacing sp inside (x/a)² + (y/b)² + (z/c)² <= 1.
BodyShape ellipsoid_shape(double a, double b, double c, double r, double sp) {
    BodyShape s;
    int na = (int)std::floor(a / sp), nb = (int)std::floor(b / sp), nc = (int)std::floor(c / sp);
    double cx = 0, cy = 0, cz = 0;
    for (int i = -na; i <= na; ++i)
        for (int j = -nb; j <= nb; ++j)
            for (int k = -nc; k <= nc; ++k) {
                double x = i * sp, y = j * sp, z = k * sp;
                if ((x * x) / (a * a) + (y * y) / (b * b) + (z * z) / (c * c) <= 1.0) {
                    s.off.push_back(x); s.off.push_back(y); s.off.push_back(z);
                    cx += x; cy += y; cz += z;
                }
            }
    if (s.off.empty()) { s.off = {0.0, 0.0, 0.0}; }
    size_t n = s.off.size() / 3;
    cx /= n; cy /= n; cz /= n;
    double m = 0;
    for (size_t p = 0; p < n; ++p) {
        s.off[3 * p] -= cx; s.off[3 * p + 1] -= cy; s.off[3 * p + 2] -= cz;
        double d = std::sqrt(s.off[3 * p] * s.off[3 * p] + s.off[3 * p + 1] * s.off[3 * p + 1] +
                             s.off[3 * p + 2] * s.off[3 * p + 2]);
        m = std::max(m, d);
    }
    s.circum = m + r;
    return s;
}

// 3x3 inverse (column-major), zero matrix when singular.
void inv3(const double m[9], double out[9]) {
    double a = m[0], b = m[3], c = m[6], d = m[1], e = m[4], f = m[7], g = m[2], h = m[5], i = m[8];
    double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    double det = a * A + b * B + c * C;
    if (std::fabs(det) < 1e-300) { for (int k = 0; k < 9; ++k) out[k] = 0.0; return; }
    double inv[9];
    inv[0] = A / det; inv[3] = -(b * i - c * h) / det; inv[6] = (b * f - c * e) / det;
    inv[1] = B / det; inv[4] = (a * i - c * g) / det;  inv[7] = -(a * f - c * d) / det;
    inv[2] = C / det; inv[5] = -(a * h - b * g) / det; inv[8] = (a * e - b * d) / det;
    for (int k = 0; k < 9; ++k) out[k] = inv[k];
}

void quat_to_mat(const double q[4], double R[9]) {   // q = (x,y,z,w), column-major
    double x = q[0], y = q[1], z = q[2], w = q[3];
    R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y + w * z);     R[2] = 2 * (x * z - w * y);
    R[3] = 2 * (x * y - w * z);     R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z + w * x);
    R[6] = 2 * (x * z + w * y);     R[7] = 2 * (y * z - w * x);     R[8] = 1 - 2 * (x * x + y * y);
}

struct CellKey {
    size_t operator()(int64_t k) const { return (size_t)(k * 0x9e3779b97f4a7c15ull); }
};

// 3 x 21-bit Morton code of a point in [0, box] (ordering only).
uint64_t morton_code(double x, double y, double z, const double box[3]) {
    auto q = [](double v, double L) {
        double u = L > 0 ? v / L : 0.0;
        u = u < 0 ? 0 : (u > 1 ? 1 : u);
        uint64_t k = (uint64_t)(u * 2097151.0);
        uint64_t r = 0;
        for (int b = 0; b < 21; ++b) r |= ((k >> b) & 1ull) << (3 * b);
        return r;
    };
    return (q(x, box[0]) << 2) | (q(y, box[1]) << 1) | q(z, box[2]);
}

}  // namespace

extern "C" int arsim_generate_scene(const ArSimGenSpec *spec, ArSimSceneOwned **out) {
    if (!spec || !out) return arsim_host::fail(ARSIM_ERR_INVALID_ARG, "null argument");
    if (spec->numCubes < 0 || spec->numGrains < 0 || spec->numEllipsoids < 0 || spec->radius <= 0.f)
        return arsim_host::fail(ARSIM_ERR_INVALID_ARG, "bad generator counts or radius");
    for (int a = 0; a < 3; ++a)
        if (spec->gridLog2[a] < 0 || spec->gridLog2[a] > 12)
            return arsim_host::fail(ARSIM_ERR_INVALID_ARG, "grid log2 out of range");
    ArSimSceneOwned *s = new (std::nothrow) ArSimSceneOwned();
    if (!s) return arsim_host::fail(ARSIM_ERR_OUT_OF_MEMORY, "host allocation failed");
    Rng rng(spec->seed);
    const double r = spec->radius;
    const double box[3] = {spec->box[0], spec->box[1], spec->box[2]};

    // ---- body shapes
    std::vector<BodyShape> shapes;
    int cubeEdge = spec->cubeEdge > 0 ? spec->cubeEdge : 4;
    for (int b = 0; b < spec->numCubes; ++b) shapes.push_back(cube_shape(cubeEdge, r));
    for (int b = 0; b < spec->numEllipsoids; ++b) {
        double lo = spec->ellipsoidMinSemi, hi = spec->ellipsoidMaxSemi;
        double a = lo + (hi - lo) * rng.uni(), bb = lo + (hi - lo) * rng.uni(), c = lo + (hi - lo) * rng.uni();
        shapes.push_back(ellipsoid_shape(a, bb, c, r, 2.0 * r));
    }
    const int B = (int)shapes.size();

    // ---- jittered lattice placement of bodies (no initial overlap)
    double maxCirc = 0;
    for (auto &sh : shapes) maxCirc = std::max(maxCirc, sh.circum);
    std::vector<double> centres(3 * (size_t)B);
    if (B > 0) {
        int n = (int)std::ceil(std::cbrt((double)B) - 1e-9);
        while ((long long)n * n * n < B) ++n;
        double pitch[3];
        int cells[3];
        double hmax = spec->dropHeightMax > 0.f ? spec->dropHeightMax : box[1];
        for (int a = 0; a < 3; ++a) {
            double ext = (a == 1) ? std::min(box[1], hmax) : box[a];
            cells[a] = n;
            pitch[a] = ext / n;
        }
        // drop the vertical lattice to as few layers as the horizontal ones allow
        while ((long long)cells[0] * cells[2] * (cells[1] - 1) >= B && cells[1] > 1) --cells[1];
        pitch[1] = ((spec->dropHeightMax > 0.f) ? std::min(box[1], (double)spec->dropHeightMax) : box[1]) / cells[1];
        double minPitch = std::min(pitch[0], std::min(pitch[1], pitch[2]));
        if (minPitch < 2.0 * maxCirc) {
            // dense pile: horizontal layers at the smallest non-overlapping pitch, stacked upwards
            // from the floor (the container is open at the top)
            const double need = 2.0 * maxCirc * 1.05;
            cells[0] = (int)std::floor(box[0] / need);
            cells[2] = (int)std::floor(box[2] / need);
            if (cells[0] < 1 || cells[2] < 1) {
                delete s;
                return arsim_host::fail(ARSIM_ERR_INVALID_ARG, "box too small for the bodies");
            }
            cells[1] = (int)((B + (long long)cells[0] * cells[2] - 1) / ((long long)cells[0] * cells[2]));
            pitch[0] = box[0] / cells[0];
            pitch[2] = box[2] / cells[2];
            pitch[1] = need;
        }
        long long total = (long long)cells[0] * cells[1] * cells[2];
        std::vector<long long> order(total);
        for (long long k = 0; k < total; ++k) order[k] = k;
        for (long long k = total - 1; k > 0; --k) {   // Fisher-Yates
            long long j = (long long)(rng.uni() * (double)(k + 1));
            if (j > k) j = k;
            std::swap(order[k], order[j]);
        }
        for (int b = 0; b < B; ++b) {
            long long c = order[b];
            int cx = (int)(c % cells[0]), cy = (int)((c / cells[0]) % cells[1]), cz = (int)(c / ((long long)cells[0] * cells[1]));
            int ci[3] = {cx, cy, cz};
            for (int a = 0; a < 3; ++a) {
                double slack = 0.5 * pitch[a] - shapes[b].circum;
                double jit = slack > 0 ? (2.0 * rng.uni() - 1.0) * slack * 0.9 : 0.0;
                centres[3 * b + a] = (ci[a] + 0.5) * pitch[a] + jit;
            }
        }
    }

    // ---- bodies in Morton order of their centres: particles that are close in space get close
    // indices, so the broadphase's gathers and scatters by particle index stay coherent
    if (B > 1) {
        std::vector<int> order(B);
        std::vector<uint64_t> code(B);
        for (int b = 0; b < B; ++b) {
            order[b] = b;
            code[b] = morton_code(centres[3 * b], centres[3 * b + 1], centres[3 * b + 2], box);
        }
        std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return code[a] < code[c]; });
        std::vector<double> c2(3 * (size_t)B);
        std::vector<BodyShape> s2(B);
        for (int k = 0; k < B; ++k) {
            for (int a = 0; a < 3; ++a) c2[3 * k + a] = centres[3 * order[k] + a];
            s2[k] = std::move(shapes[order[k]]);
        }
        centres.swap(c2);
        shapes.swap(s2);
    }

    // ---- particles of bodies
    std::vector<double> P;   // world positions (double) for overlap checks
    for (int b = 0; b < B; ++b) {
        // Shoemake uniform random rotation, (x,y,z,w)
        double u1 = rng.uni(), u2 = rng.uni(), u3 = rng.uni();
        double q[4] = {std::sqrt(1 - u1) * std::sin(6.283185307179586 * u2), std::sqrt(1 - u1) * std::cos(6.283185307179586 * u2),
                       std::sqrt(u1) * std::sin(6.283185307179586 * u3), std::sqrt(u1) * std::cos(6.283185307179586 * u3)};
        double R[9];
        quat_to_mat(q, R);
        double v[3] = {rng.normal() * spec->bodyVelSigma, rng.normal() * spec->bodyVelSigma, rng.normal() * spec->bodyVelSigma};
        const BodyShape &sh = shapes[b];
        size_t np = sh.off.size() / 3;
        double I[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (size_t p = 0; p < np; ++p) {
            float fx = (float)sh.off[3 * p], fy = (float)sh.off[3 * p + 1], fz = (float)sh.off[3 * p + 2];
            double x = fx, y = fy, z = fz;
            I[0] += y * y + z * z; I[3] -= x * y; I[6] -= x * z;
            I[1] -= x * y; I[4] += x * x + z * z; I[7] -= y * z;
            I[2] -= x * z; I[5] -= y * z; I[8] += x * x + y * y;
            double w[3] = {R[0] * x + R[3] * y + R[6] * z, R[1] * x + R[4] * y + R[7] * z, R[2] * x + R[5] * y + R[8] * z};
            double wp[3] = {centres[3 * b] + w[0], centres[3 * b + 1] + w[1], centres[3 * b + 2] + w[2]};
            s->pos.insert(s->pos.end(), {(float)wp[0], (float)wp[1], (float)wp[2], 0.f});
            s->vel.insert(s->vel.end(), {(float)v[0], (float)v[1], (float)v[2], 0.f});
            s->bodyFrame.insert(s->bodyFrame.end(), {fx, fy, fz, 0.f});
            s->rb.push_back(b);
            P.insert(P.end(), {wp[0], wp[1], wp[2]});
        }
        double Iinv[9];
        inv3(I, Iinv);
        for (int k = 0; k < 9; ++k) s->rbInertia.push_back((float)Iinv[k]);
        s->rbPos.insert(s->rbPos.end(), {(float)centres[3 * b], (float)centres[3 * b + 1], (float)centres[3 * b + 2], 0.f});
        s->rbVel.insert(s->rbVel.end(), {(float)v[0], (float)v[1], (float)v[2], 0.f});
        s->rbAng.insert(s->rbAng.end(), {0.f, 0.f, 0.f, 0.f});
        s->rbQuat.insert(s->rbQuat.end(), {(float)q[0], (float)q[1], (float)q[2], (float)q[3]});
        s->rbMass.push_back((float)np);
        s->ppo.push_back((int)np);
    }

    // ---- free grains: uniform in the box, rejected if closer than 2r to any particle
    if (spec->numGrains > 0) {
        const double cell = 2.0 * r;
        std::unordered_map<int64_t, std::vector<int>, CellKey> grid;
        grid.reserve((P.size() / 3 + spec->numGrains) * 2);
        auto key = [&](int x, int y, int z) { return ((int64_t)(x + 1048576) << 42) | ((int64_t)(y + 1048576) << 21) | (int64_t)(z + 1048576); };
        auto insert = [&](size_t idx) {
            int cx = (int)std::floor(P[3 * idx] / cell), cy = (int)std::floor(P[3 * idx + 1] / cell), cz = (int)std::floor(P[3 * idx + 2] / cell);
            grid[key(cx, cy, cz)].push_back((int)idx);
        };
        for (size_t k = 0; k < P.size() / 3; ++k) insert(k);
        struct GrainRec { uint64_t code; double v[6]; };
        std::vector<GrainRec> grains;
        grains.reserve(spec->numGrains);
        int placed = 0;
        long long attempts = 0, maxAttempts = (long long)spec->numGrains * 200 + 1000;
        while (placed < spec->numGrains && attempts < maxAttempts) {
            ++attempts;
            double x = r + (box[0] - 2 * r) * rng.uni();
            double y = r + (box[1] - 2 * r) * rng.uni();
            double z = r + (box[2] - 2 * r) * rng.uni();
            int cx = (int)std::floor(x / cell), cy = (int)std::floor(y / cell), cz = (int)std::floor(z / cell);
            bool ok = true;
            for (int dz = -1; dz <= 1 && ok; ++dz)
                for (int dy = -1; dy <= 1 && ok; ++dy)
                    for (int dx = -1; dx <= 1 && ok; ++dx) {
                        auto it = grid.find(key(cx + dx, cy + dy, cz + dz));
                        if (it == grid.end()) continue;
                        for (int j : it->second) {
                            double ddx = P[3 * (size_t)j] - x, ddy = P[3 * (size_t)j + 1] - y, ddz = P[3 * (size_t)j + 2] - z;
                            if (ddx * ddx + ddy * ddy + ddz * ddz < 4.0 * r * r * 1.0001) { ok = false; break; }
                        }
                    }
            if (!ok) continue;
            P.insert(P.end(), {x, y, z});
            insert(P.size() / 3 - 1);
            double vs = spec->grainVelSigma;
            double v0 = rng.normal() * vs, v1 = rng.normal() * vs, v2 = rng.normal() * vs;
            grains.push_back({morton_code(x, y, z, box), {x, y, z, v0, v1, v2}});
            ++placed;
        }
        if (placed < spec->numGrains) {
            delete s;
            return arsim_host::fail(ARSIM_ERR_INVALID_ARG, "could not place all grains without overlap");
        }
        // grains in Morton order, after the bodies
        std::stable_sort(grains.begin(), grains.end(),
                         [](const GrainRec &a, const GrainRec &b) { return a.code < b.code; });
        for (const GrainRec &g : grains) {
            s->pos.insert(s->pos.end(), {(float)g.v[0], (float)g.v[1], (float)g.v[2], 0.f});
            s->vel.insert(s->vel.end(), {(float)g.v[3], (float)g.v[4], (float)g.v[5], 0.f});
            s->bodyFrame.insert(s->bodyFrame.end(), {0.f, 0.f, 0.f, 0.f});
            s->rb.push_back(-1);
        }
    }

    s->N = (int)s->rb.size();
    s->B = B;
    ArSimParams &p = s->params;
    std::memset(&p, 0, sizeof(p));
    p.gravity = {0.f, spec->gravityY, 0.f};
    p.globalDamping = 1.0f;
    p.particleRadius = spec->radius;
    p.gridSize = {1u << spec->gridLog2[0], 1u << spec->gridLog2[1], 1u << spec->gridLog2[2]};
    p.numCells = p.gridSize.x * p.gridSize.y * p.gridSize.z;
    p.worldOrigin = {0.f, 0.f, 0.f};
    p.cellSize = {2.0f * spec->radius, 2.0f * spec->radius, 2.0f * spec->radius};
    p.numBodies = (unsigned)B;
    p.spring = spec->spring;
    p.damping = spec->damping;
    p.shear = spec->shear;
    p.attraction = 0.0f;
    p.boundaryDamping = -0.5f;
    p.Wdamping = 1.0f;
    p.Vdamping = 1.0f;
    ArSimStepConfig &c = s->cfg;
    std::memset(&c, 0, sizeof(c));
    c.dt = spec->dt;
    for (int a = 0; a < 3; ++a) { c.boxMin[a] = 0.f; c.boxMax[a] = spec->box[a]; }
    c.wallMask = ARSIM_WALLS_OPEN_TOP;
    c.method = spec->method;
    c.flags = 0;
    *out = s;
    return ARSIM_OK;
}

namespace {

// Points on a uniform grid of cells of edge h over [0, n*h)^3 (coordinates clamped into it), with
// linked lists per cell: "is any point closer than d to (x, y, z)?" for d <= h.
struct PointGrid {
    double h;
    int n[3];
    std::vector<int> head, next;
    std::vector<double> P;
    PointGrid(double cell, const double ext[3]) : h(cell) {
        for (int a = 0; a < 3; ++a) n[a] = std::max(1, (int)std::ceil(ext[a] / cell) + 1);
        head.assign((size_t)n[0] * n[1] * n[2], -1);
    }
    int c(double v, int a) const {
        int k = (int)std::floor(v / h);
        return k < 0 ? 0 : (k >= n[a] ? n[a] - 1 : k);
    }
    void add(double x, double y, double z) {
        size_t cell = ((size_t)c(z, 2) * n[1] + c(y, 1)) * n[0] + c(x, 0);
        next.push_back(head[cell]);
        head[cell] = (int)(P.size() / 3);
        P.insert(P.end(), {x, y, z});
    }
    bool near(double x, double y, double z, double d) const {
        const int cx = c(x, 0), cy = c(y, 1), cz = c(z, 2);
        const double d2 = d * d;
        for (int k = std::max(cz - 1, 0); k <= std::min(cz + 1, n[2] - 1); ++k)
            for (int j = std::max(cy - 1, 0); j <= std::min(cy + 1, n[1] - 1); ++j)
                for (int i = std::max(cx - 1, 0); i <= std::min(cx + 1, n[0] - 1); ++i)
                    for (int q = head[((size_t)k * n[1] + j) * n[0] + i]; q >= 0; q = next[q]) {
                        const double dx = P[3 * (size_t)q] - x, dy = P[3 * (size_t)q + 1] - y, dz = P[3 * (size_t)q + 2] - z;
                        if (dx * dx + dy * dy + dz * dz < d2) return true;
                    }
        return false;
    }
};

}  // namespace

// Dense settled pile: ellipsoids voxelised at spacing r, placed at random positions and orientations
// (largest first) so that no two bodies' particles are closer than the lattice distance, then the
// space around them filled from the floor up with grains on a face-centred cubic lattice.
extern "C" int arsim_generate_packed_pile(const ArSimPileSpec *spec, ArSimSceneOwned **out) {
    if (!spec || !out) return arsim_host::fail(ARSIM_ERR_INVALID_ARG, "null argument");
    if (spec->numEllipsoids < 0 || spec->targetParticles < 0 || !(spec->edge > 0.f) || !(spec->radius > 0.f) ||
        !(spec->latticeGap > 0.f && spec->latticeGap <= 1.f) || !(spec->surfaceGap > 0.f && spec->surfaceGap <= 1.f) || spec->gridLog2 < 1 || spec->gridLog2 > 12)
        return arsim_host::fail(ARSIM_ERR_INVALID_ARG, "packed pile: bad counts, sizes or grid");
    const double r = spec->radius, edge = spec->edge;
    const double d0 = 2.0 * r * spec->latticeGap;   // nearest-neighbour distance of the lattice
    const double dS = 2.0 * r * spec->surfaceGap;   // closest a grain may sit to a body particle
    Rng rng(spec->seed);

    // ---- shapes: semi-axes 2r * 4^(u^2), u ~ U[0,1] per axis, on [2r, 8r] with a heavy tail in size
    std::vector<BodyShape> shapes;
    shapes.reserve(spec->numEllipsoids);
    double bodyParticles = 0, maxCirc = 0;
    for (int b = 0; b < spec->numEllipsoids; ++b) {
        double ax[3];
        for (int a = 0; a < 3; ++a) {
            const double u = rng.uni();
            ax[a] = 2.0 * r * std::pow(4.0, u * u);
        }
        shapes.push_back(ellipsoid_shape(ax[0], ax[1], ax[2], r, r));
        bodyParticles += (double)(shapes.back().off.size() / 3);
        maxCirc = std::max(maxCirc, shapes.back().circum);
    }
    if (bodyParticles > spec->targetParticles)
        return arsim_host::fail(ARSIM_ERR_INVALID_ARG, "packed pile: the bodies alone exceed the particle target");
    if (edge < 2.0 * maxCirc + 2.0 * r)
        return arsim_host::fail(ARSIM_ERR_INVALID_ARG, "packed pile: container narrower than the largest body");
    const double siteVolume = d0 * d0 * d0 / std::sqrt(2.0);
    const double grainsWanted = spec->targetParticles - bodyParticles;
    // pile height holding the bodies' voxels and the grains' lattice sites, with 10 % slack
    double H = 1.1 * (bodyParticles * r * r * r + grainsWanted * siteVolume) / (edge * edge);
    H = std::max(H, 2.0 * maxCirc + 2.0 * r);

    // ---- random sequential placement of the bodies, largest first
    std::vector<int> order(shapes.size());
    for (size_t b = 0; b < shapes.size(); ++b) order[b] = (int)b;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return shapes[a].off.size() > shapes[b].off.size(); });
    const double ext[3] = {edge, 4.0 * H + 8.0 * maxCirc, edge};
    PointGrid bodyGrid(2.0 * r, ext);
    const int B = (int)shapes.size();
    std::vector<double> centre(3 * (size_t)B), quat(4 * (size_t)B);
    std::vector<double> world;
    double yTop = H;
    for (int k = 0; k < B; ++k) {
        const int b = order[k];
        const BodyShape &sh = shapes[b];
        const size_t np = sh.off.size() / 3;
        bool placed = false;
        for (int attempt = 0; !placed; ++attempt) {
            if (attempt > 0 && attempt % 400 == 0) yTop *= 1.05;   // crowded: let the pile grow
            if (yTop > ext[1] - 2.0 * sh.circum)
                return arsim_host::fail(ARSIM_ERR_INVALID_ARG, "packed pile: bodies do not fit the container");
            const double c[3] = {sh.circum + (edge - 2.0 * sh.circum) * rng.uni(),
                                 sh.circum + std::max(0.0, yTop - 2.0 * sh.circum) * rng.uni(),
                                 sh.circum + (edge - 2.0 * sh.circum) * rng.uni()};
            const double u1 = rng.uni(), u2 = rng.uni(), u3 = rng.uni();
            const double q[4] = {std::sqrt(1 - u1) * std::sin(6.283185307179586 * u2),
                                 std::sqrt(1 - u1) * std::cos(6.283185307179586 * u2),
                                 std::sqrt(u1) * std::sin(6.283185307179586 * u3),
                                 std::sqrt(u1) * std::cos(6.283185307179586 * u3)};
            double R[9];
            quat_to_mat(q, R);
            world.resize(3 * np);
            bool ok = true;
            for (size_t p = 0; p < np && ok; ++p) {
                // the same FP32 body-frame offsets the scene stores
                const double x = (float)sh.off[3 * p], y = (float)sh.off[3 * p + 1], z = (float)sh.off[3 * p + 2];
                world[3 * p] = c[0] + R[0] * x + R[3] * y + R[6] * z;
                world[3 * p + 1] = c[1] + R[1] * x + R[4] * y + R[7] * z;
                world[3 * p + 2] = c[2] + R[2] * x + R[5] * y + R[8] * z;
                ok = !bodyGrid.near(world[3 * p], world[3 * p + 1], world[3 * p + 2], d0);
            }
            if (!ok) continue;
            for (size_t p = 0; p < np; ++p) bodyGrid.add(world[3 * p], world[3 * p + 1], world[3 * p + 2]);
            for (int a = 0; a < 3; ++a) centre[3 * (size_t)b + a] = c[a];
            for (int a = 0; a < 4; ++a) quat[4 * (size_t)b + a] = q[a];
            placed = true;
        }
    }

    // ---- grains: FCC half-layers from the floor up, sites closer than d0 to a body particle skipped.
    // Half-layer L at height r + L a/2 holds the points (r + m a/2, r + n a/2) with m + n + L even,
    // a = d0 sqrt(2) the cubic cell: every site has 12 neighbours at distance d0.
    const double a = d0 * std::sqrt(2.0);
    const int steps = (int)std::floor((edge - 2.0 * r) / (0.5 * a)) + 1;
    const long long want = (long long)std::llround(grainsWanted);
    std::vector<double> G;
    G.reserve(3 * (size_t)std::max(want, 0LL));
    for (int layer = 0; (long long)(G.size() / 3) < want; ++layer) {
        const double y = r + layer * 0.5 * a;
        if (y > ext[1] - r) return arsim_host::fail(ARSIM_ERR_INVALID_ARG, "packed pile: grains do not fit");
        for (int n = 0; n < steps && (long long)(G.size() / 3) < want; ++n)
            for (int m = 0; m < steps && (long long)(G.size() / 3) < want; ++m) {
                if ((m + n + layer) & 1) continue;
                const double x = r + m * 0.5 * a, z = r + n * 0.5 * a;
                if (bodyGrid.near(x, y, z, dS)) continue;
                G.insert(G.end(), {x, y, z});
            }
    }

    // ---- output: bodies in Morton order of their centres, then grains in Morton order
    ArSimSceneOwned *s = new (std::nothrow) ArSimSceneOwned();
    if (!s) return arsim_host::fail(ARSIM_ERR_OUT_OF_MEMORY, "host allocation failed");
    double top = 0;
    for (size_t p = 1; p < bodyGrid.P.size(); p += 3) top = std::max(top, bodyGrid.P[p]);
    for (size_t p = 1; p < G.size(); p += 3) top = std::max(top, G[p]);
    const double box[3] = {edge, std::max(edge, top + 2.0 * r), edge};
    std::vector<int> bo(B);
    std::vector<uint64_t> code(B);
    for (int b = 0; b < B; ++b) {
        bo[b] = b;
        code[b] = morton_code(centre[3 * (size_t)b], centre[3 * (size_t)b + 1], centre[3 * (size_t)b + 2], box);
    }
    std::stable_sort(bo.begin(), bo.end(), [&](int x, int y) { return code[x] < code[y]; });
    for (int k = 0; k < B; ++k) {
        const int b = bo[k];
        const BodyShape &sh = shapes[b];
        const double *q = &quat[4 * (size_t)b], *c = &centre[3 * (size_t)b];
        double R[9];
        quat_to_mat(q, R);
        const size_t np = sh.off.size() / 3;
        double I[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (size_t p = 0; p < np; ++p) {
            const float fx = (float)sh.off[3 * p], fy = (float)sh.off[3 * p + 1], fz = (float)sh.off[3 * p + 2];
            const double x = fx, y = fy, z = fz;
            I[0] += y * y + z * z; I[3] -= x * y; I[6] -= x * z;
            I[1] -= x * y; I[4] += x * x + z * z; I[7] -= y * z;
            I[2] -= x * z; I[5] -= y * z; I[8] += x * x + y * y;
            const double w[3] = {c[0] + R[0] * x + R[3] * y + R[6] * z, c[1] + R[1] * x + R[4] * y + R[7] * z,
                                 c[2] + R[2] * x + R[5] * y + R[8] * z};
            s->pos.insert(s->pos.end(), {(float)w[0], (float)w[1], (float)w[2], 0.f});
            s->vel.insert(s->vel.end(), {0.f, 0.f, 0.f, 0.f});
            s->bodyFrame.insert(s->bodyFrame.end(), {fx, fy, fz, 0.f});
            s->rb.push_back(k);
        }
        double Iinv[9];
        inv3(I, Iinv);
        for (int m = 0; m < 9; ++m) s->rbInertia.push_back((float)Iinv[m]);
        s->rbPos.insert(s->rbPos.end(), {(float)c[0], (float)c[1], (float)c[2], 0.f});
        s->rbVel.insert(s->rbVel.end(), {0.f, 0.f, 0.f, 0.f});
        s->rbAng.insert(s->rbAng.end(), {0.f, 0.f, 0.f, 0.f});
        s->rbQuat.insert(s->rbQuat.end(), {(float)q[0], (float)q[1], (float)q[2], (float)q[3]});
        s->rbMass.push_back((float)np);
        s->ppo.push_back((int)np);
    }
    const size_t ng = G.size() / 3;
    std::vector<uint64_t> gcode(ng);
    std::vector<int> go(ng);
    for (size_t g = 0; g < ng; ++g) {
        go[g] = (int)g;
        gcode[g] = morton_code(G[3 * g], G[3 * g + 1], G[3 * g + 2], box);
    }
    std::stable_sort(go.begin(), go.end(), [&](int x, int y) { return gcode[x] < gcode[y]; });
    for (size_t k = 0; k < ng; ++k) {
        const size_t g = (size_t)go[k];
        s->pos.insert(s->pos.end(), {(float)G[3 * g], (float)G[3 * g + 1], (float)G[3 * g + 2], 0.f});
        s->vel.insert(s->vel.end(), {0.f, 0.f, 0.f, 0.f});
        s->bodyFrame.insert(s->bodyFrame.end(), {0.f, 0.f, 0.f, 0.f});
        s->rb.push_back(-1);
    }
    s->N = (int)s->rb.size();
    s->B = B;
    ArSimParams &p = s->params;
    std::memset(&p, 0, sizeof(p));
    p.gravity = {0.f, -9.81f, 0.f};
    p.globalDamping = 1.0f;
    p.particleRadius = spec->radius;
    const unsigned g = 1u << spec->gridLog2;
    p.gridSize = {g, g, g};
    p.numCells = g * g * g;
    p.worldOrigin = {0.f, 0.f, 0.f};
    p.cellSize = {2.0f * spec->radius, 2.0f * spec->radius, 2.0f * spec->radius};
    p.numBodies = (unsigned)B;
    p.spring = 5e4f;
    p.damping = 50.f;
    p.shear = 10.f;
    p.attraction = 0.0f;
    p.boundaryDamping = -0.5f;
    p.Wdamping = 1.0f;
    p.Vdamping = 1.0f;
    ArSimStepConfig &cf = s->cfg;
    std::memset(&cf, 0, sizeof(cf));
    cf.dt = 1e-4f;
    for (int m = 0; m < 3; ++m) { cf.boxMin[m] = 0.f; cf.boxMax[m] = (float)box[m]; }
    cf.wallMask = ARSIM_WALLS_OPEN_TOP;
    cf.method = spec->method;
    cf.flags = 0;
    *out = s;
    return ARSIM_OK;
}

extern "C" int arsim_scene_view(const ArSimSceneOwned *s, ArSimSceneHost *v) {
    if (!s || !v) return arsim_host::fail(ARSIM_ERR_INVALID_ARG, "null argument");
    v->numParticles = s->N;
    v->numBodies = s->B;
    v->pos = s->pos.data();
    v->vel = s->vel.data();
    v->rbIndices = s->rb.data();
    v->bodyFramePos = s->bodyFrame.data();
    v->rbPositions = s->rbPos.data();
    v->rbVelocities = s->rbVel.data();
    v->rbAngularVelocity = s->rbAng.data();
    v->rbQuaternion = s->rbQuat.data();
    v->rbInertia = s->rbInertia.data();
    v->rbMass = s->rbMass.data();
    v->particlesPerObject = s->ppo.data();
    v->params = s->params;
    v->cfg = s->cfg;
    return ARSIM_OK;
}

extern "C" int arsim_scene_free(ArSimSceneOwned *s) {
    delete s;
    return ARSIM_OK;
}
