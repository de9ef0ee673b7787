// Host-code driver for an AddressSanitizer / UndefinedBehaviorSanitizer build of the pure host C++:
// the scene generators (scenegen.cpp), the template loader and scene builder (sceneio.cpp) and the
// oracle (oracle/arsim_oracle.cpp).  Each part runs on small inputs; any sanitizer report aborts
// the process with a non-zero status.  argv[1] = the OBJparticles template directory.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/arsim.h"

namespace arsim_host {
int fail(int code, const std::string &msg) {
    std::fprintf(stderr, "library error %d: %s\n", code, msg.c_str());
    return code;
}
}  // namespace arsim_host

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
int orc_step_f32(OrcStateF *o, const ArSimParams *p, const ArSimStepConfig *cfg, int nsteps);
int orc_step_f64(OrcStateD *o, const ArSimParams *p, const ArSimStepConfig *cfg, int nsteps);
void orc_range_image(const uint16_t *img, int W, int H, const float *cam, float radius, float *pos, float *norm,
                     float *radii, float *vao);
}

namespace {

int failures = 0;
void check(bool ok, const char *what) {
    if (!ok) {
        std::fprintf(stderr, "FAILED: %s\n", what);
        ++failures;
    }
}

// Arrays of one oracle state in precision T, initialised from a scene view as the tests do.
template <class T> struct Arrays {
    std::vector<T> pos, vel, rel, frame, force, torque, rbPos, rbVel, rbAng, qinc, qcum, Ib, Iw, mass, model;
    std::vector<int> rb, pstart, pcount;
    explicit Arrays(const ArSimSceneHost &v) {
        const size_t N = v.numParticles, B = v.numBodies;
        auto cp = [](const float *src, size_t n) { return std::vector<T>(src, src + n); };
        pos = cp(v.pos, 4 * N); vel = cp(v.vel, 4 * N); frame = cp(v.bodyFramePos, 4 * N);
        rel.assign(4 * N, T(0)); force.assign(4 * N, T(0)); torque.assign(4 * N, T(0));
        rb.assign(v.rbIndices, v.rbIndices + N);
        rbPos = cp(v.rbPositions, 4 * B); rbVel = cp(v.rbVelocities, 4 * B); rbAng = cp(v.rbAngularVelocity, 4 * B);
        qcum = cp(v.rbQuaternion, 4 * B);
        qinc.assign(4 * B, T(0));
        for (size_t b = 0; b < B; ++b) qinc[4 * b + 3] = T(1);
        Ib = cp(v.rbInertia, 9 * B); Iw = Ib; mass = cp(v.rbMass, B);
        pcount.assign(v.particlesPerObject, v.particlesPerObject + B);
        pstart.resize(B);
        int acc = 0;
        for (size_t b = 0; b < B; ++b) { pstart[b] = acc; acc += pcount[b]; }
        model.assign(16 * B + 16, T(0));
    }
    template <class S> void fill(S &s, int N, int B) {
        s.N = N; s.B = B;
        s.pos = pos.data(); s.vel = vel.data(); s.relPos = rel.data(); s.bodyFrame = frame.data();
        s.pForce = force.data(); s.pTorque = torque.data(); s.rb = rb.data();
        s.rbPos = rbPos.data(); s.rbVel = rbVel.data(); s.rbAng = rbAng.data(); s.qinc = qinc.data();
        s.qcum = qcum.data(); s.Ib = Ib.data(); s.Iw = Iw.data(); s.mass = mass.data();
        s.pstart = pstart.data(); s.pcount = pcount.data(); s.model = model.data();
        s.staticPos = nullptr; s.staticNorm = nullptr; s.numStatic = 0;
    }
    bool finite() const {
        for (T x : pos) if (!std::isfinite((double)x)) return false;
        for (T x : vel) if (!std::isfinite((double)x)) return false;
        return true;
    }
};

void step_both(const ArSimSceneOwned *sc, int method, int steps, const char *what) {
    ArSimSceneHost v;
    check(arsim_scene_view(sc, &v) == ARSIM_OK, "scene_view");
    ArSimStepConfig cfg = v.cfg;
    cfg.method = method;
    Arrays<float> f(v);
    OrcStateF sf;
    f.fill(sf, v.numParticles, v.numBodies);
    check(orc_step_f32(&sf, &v.params, &cfg, steps) == 0 && f.finite(), what);
    Arrays<double> d(v);
    OrcStateD sd;
    d.fill(sd, v.numParticles, v.numBodies);
    check(orc_step_f64(&sd, &v.params, &cfg, steps) == 0 && d.finite(), what);
}

}  // namespace

int main(int argc, char **argv) {
    // cubes, ellipsoids and grains from the lattice generator, every broadphase
    ArSimGenSpec g{};
    g.seed = 5; g.numCubes = 24; g.cubeEdge = 4; g.numEllipsoids = 6; g.ellipsoidMinSemi = 1.f;
    g.ellipsoidMaxSemi = 3.f; g.numGrains = 500;
    for (int a = 0; a < 3; ++a) { g.box[a] = 32.f; g.gridLog2[a] = 5; }
    g.radius = 0.5f; g.bodyVelSigma = 0.5f; g.grainVelSigma = 0.5f; g.method = ARSIM_METHOD_GRID;
    g.dt = 1e-4f; g.spring = 5e4f; g.damping = 50.f; g.shear = 10.f; g.gravityY = -9.81f;
    ArSimSceneOwned *sc = nullptr;
    check(arsim_generate_scene(&g, &sc) == ARSIM_OK, "generate_scene");
    if (sc) {
        step_both(sc, ARSIM_METHOD_GRID, 3, "lattice scene, grid");
        step_both(sc, ARSIM_METHOD_LBVH30, 2, "lattice scene, lbvh30");
        step_both(sc, ARSIM_METHOD_LBVH63, 2, "lattice scene, lbvh63");
        arsim_scene_free(sc);
    }
    // the packed pile at about 1 % of its full size
    ArSimPileSpec ps{};
    ps.seed = 3; ps.numEllipsoids = 60; ps.targetParticles = 30000; ps.edge = 16.f; ps.radius = 0.5f;
    ps.latticeGap = 0.98f; ps.surfaceGap = 0.9f; ps.gridLog2 = 7; ps.method = ARSIM_METHOD_GRID;
    sc = nullptr;
    check(arsim_generate_packed_pile(&ps, &sc) == ARSIM_OK, "generate_packed_pile");
    if (sc) {
        step_both(sc, ARSIM_METHOD_GRID, 1, "packed pile, grid");
        arsim_scene_free(sc);
    }
    // bad generator inputs are refused
    ArSimPileSpec bad = ps;
    bad.edge = 4.f;
    sc = nullptr;
    check(arsim_generate_packed_pile(&bad, &sc) != ARSIM_OK && !sc, "packed pile: narrow container refused");
    // templates in the reference's OBJparticles format and the scene builder
    if (argc > 1) {
        ArSimSceneOwned *built = nullptr;
        ArSimParams p{};
        p.gravity = {0.f, -9.81f, 0.f}; p.globalDamping = 1.f; p.particleRadius = 0.5f;
        p.gridSize = {32, 32, 32}; p.numCells = 32 * 32 * 32; p.cellSize = {1.f, 1.f, 1.f};
        p.spring = 5e4f; p.damping = 50.f; p.shear = 10.f; p.boundaryDamping = -0.5f; p.Wdamping = 1.f;
        p.Vdamping = 1.f;
        ArSimStepConfig c{};
        c.dt = 1e-4f;
        for (int a = 0; a < 3; ++a) { c.boxMin[a] = 0.f; c.boxMax[a] = 32.f; }
        c.wallMask = ARSIM_WALLS_OPEN_TOP;
        c.method = ARSIM_METHOD_GRID;
        check(arsim_scene_create(&p, &c, &built) == ARSIM_OK, "scene_create");
        const char *models[] = {"bunny", "teapot", "banana", "cube"};
        const float scales[] = {1.0f, 1.5f, 2.0f, 2.5f};
        int k = 0;
        for (const char *m : models)
            for (float s : scales) {
                ArSimTemplate *t = nullptr;
                if (arsim_template_load(argv[1], m, s, 0.5f, &t) != ARSIM_OK) {
                    check(false, "template_load");
                    continue;
                }
                int n = 0;
                check(arsim_template_info(t, &n, nullptr, nullptr, nullptr, nullptr) == ARSIM_OK && n > 0,
                      "template_info");
                const float pos[3] = {4.f + 6.f * (k % 4), 8.f + 6.f * (k / 4), 16.f};
                const float vel[3] = {0.f, 0.f, 0.f}, ang[3] = {0.f, 0.f, 0.f};
                check(arsim_scene_add_body(built, t, pos, vel, ang) == ARSIM_OK, "scene_add_body");
                arsim_template_free(t);
                ++k;
            }
        std::vector<float> gp(4 * 64, 0.f), gv(4 * 64, 0.f);
        for (int i = 0; i < 64; ++i) { gp[4 * i] = 2.f + (i % 8); gp[4 * i + 1] = 1.f; gp[4 * i + 2] = 2.f + (i / 8); }
        check(arsim_scene_add_grains(built, 64, gp.data(), gv.data()) == ARSIM_OK, "scene_add_grains");
        step_both(built, ARSIM_METHOD_GRID, 2, "template scene, grid");
        step_both(built, ARSIM_METHOD_LBVH30, 2, "template scene, lbvh30");
        arsim_scene_free(built);
    }
    // range image to static points
    const int W = 32, H = 24;
    std::vector<uint16_t> img(W * H, 1200);
    for (int i = 0; i < W; ++i) img[i] = 0;   // missing depth
    const float cam[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    std::vector<float> pos(4 * W * H, 0.f), norm(4 * W * H, 0.f), radii(W * H, 0.f), vao(5 * W * H, 0.f);
    orc_range_image(img.data(), W, H, cam, 0.5f, pos.data(), norm.data(), radii.data(), vao.data());
    check(std::isfinite(pos[4 * (W + 1)]) && radii[W + 1] == 0.5f, "range image");
    if (failures) return 1;
    std::puts("host code clean under the sanitizers");
    return 0;
}
