// sceneio.cpp -- particle templates in the reference's OBJparticles format and a host-side scene
// builder (bodies placed from templates, free grains), the input side of the step engine.
//
// Template loading follows ParticleSystem::initObject (RigidBodies/RigidBodyAddition.cpp:1486-1756):
//   * file  <dataDir>/<model>/<model>_<suffix>.txt, a particle count on the first line and one
//     "x y z" line per particle (:1505-1593); the suffix is chosen from the scale by the reference's
//     comparison chain (:1494-1503), including its quirk (see arsim_template_suffix);
//   * the centre of mass is the FP32 mean of the points (:1575-1595), and each point is moved by
//     -(cm + per-model offset), the offsets being bunny (-r, r, -r), banana (0, 3r, 0), teapot
//     (0, 2r, 0) (:1599-1625);
//   * the inertia tensor is the identity plus the FP32 point-mass sum Σ (|p|² I - p pᵀ) over the
//     moved points (:1577, :1633-1643; the identity is glm::mat3's default value), replaced for some
//     models and scales by the diagonal tensors the reference hard-codes (:1693-1729) and inverted;
//     for cubes the inverse itself is hard-coded (:1730-1752);
//   * the body "radius" is the largest coordinate of any moved point (:1646-1649: the maximum of x,
//     y and z, not of |p|), the mass is 1 (:1858), the orientation the identity (:1828).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <string>
#include <vector>

#include "../../include/arsim.h"
#include "host_common.hpp"
#include "scene_owned.hpp"

using arsim_host::fail;

struct ArSimTemplate {
    std::vector<float> rel;   // float4 per particle, w = 0
    float invInertia[9];      // column-major
    float radius = 0.f;
    float mass = 1.f;
    std::string file;
};

namespace {

// Inverse of a column-major 3x3 matrix through the adjugate, in double, rounded to float once (the
// reference inverts in FP32 with glm::inverse; the two agree to a few ulp).
bool inverse3(const float *m, float *out) {
    const double a = m[0], b = m[3], c = m[6], d = m[1], e = m[4], f = m[7], g = m[2], h = m[5], k = m[8];
    const double A = e * k - f * h, B = -(d * k - f * g), C = d * h - e * g;
    const double det = a * A + b * B + c * C;
    if (det == 0.0 || !std::isfinite(det)) return false;
    const double r[9] = {A, B, C,                                       // column 0
                         -(b * k - c * h), a * k - c * g, -(a * h - b * g),   // column 1
                         b * f - c * e, -(a * f - c * d), a * e - b * d};     // column 2
    for (int i = 0; i < 9; ++i) out[i] = (float)(r[i] / det);
    return true;
}

bool near_scale(float scale, float s) { return std::fabs(scale - s) < 0.0001f; }

void set_diag(float *m, float x, float y, float z) {
    for (int i = 0; i < 9; ++i) m[i] = 0.f;
    m[0] = x; m[4] = y; m[8] = z;
}

}  // namespace

extern "C" const char *arsim_template_suffix(float scale) {
    // The reference writes abs(scale - 1.0 < 0.01): the comparison is taken before abs, so each test
    // is "scale < threshold" and the first one that holds wins.
    const double s = scale;
    if (s - 1.0 < 0.01) return "1_0";
    if (s - 1.5 < 0.01) return "1_5";
    if (s - 2.0 < 0.01) return "2_0";
    if (s - 2.5 < 0.01) return "2_5";
    return "1_5";
}

extern "C" int arsim_template_load(const char *dataDir, const char *model, float scale, float particleRadius,
                                   ArSimTemplate **out) {
    if (!dataDir || !model || !out) return fail(ARSIM_ERR_INVALID_ARG, "template_load: null argument");
    *out = nullptr;
    std::string path = std::string(dataDir) + "/" + model + "/" + model + "_" + arsim_template_suffix(scale) + ".txt";
    std::ifstream f(path);
    if (!f.is_open()) return fail(ARSIM_ERR_INVALID_ARG, "template_load: cannot open " + path);
    std::string line;
    if (!std::getline(f, line)) return fail(ARSIM_ERR_INVALID_ARG, "template_load: empty file " + path);
    char *end = nullptr;
    const long count = std::strtol(line.c_str(), &end, 10);
    if (end == line.c_str() || count <= 0 || count > (1l << 24))
        return fail(ARSIM_ERR_INVALID_ARG, "template_load: bad particle count in " + path);
    ArSimTemplate *t = new (std::nothrow) ArSimTemplate();
    if (!t) return fail(ARSIM_ERR_OUT_OF_MEMORY, "template_load");
    t->file = path;
    std::vector<float> p(3 * (size_t)count);
    float cx = 0.f, cy = 0.f, cz = 0.f;
    for (long i = 0; i < count; ++i) {
        if (!std::getline(f, line)) {
            delete t;
            return fail(ARSIM_ERR_INVALID_ARG, "template_load: fewer points than announced in " + path);
        }
        const char *q = line.c_str();
        for (int c = 0; c < 3; ++c) {
            char *e = nullptr;
            const float v = std::strtof(q, &e);
            if (e == q) {
                delete t;
                return fail(ARSIM_ERR_INVALID_ARG, "template_load: bad point line in " + path);
            }
            p[3 * i + c] = v;
            q = e;
        }
        cx += p[3 * i];
        cy += p[3 * i + 1];
        cz += p[3 * i + 2];
    }
    const float n = (float)count;
    cx = cx / n;
    cy = cy / n;
    cz = cz / n;
    const float r = particleRadius;
    float ox = 0.f, oy = 0.f, oz = 0.f;
    if (!std::strcmp(model, "bunny")) { ox = -r; oy = r; oz = -r; }
    if (!std::strcmp(model, "banana")) oy = 3.f * r;
    else if (!std::strcmp(model, "teapot")) oy = 2.f * r;
    const float sx = cx + ox, sy = cy + oy, sz = cz + oz;
    // column-major, symmetric; the reference's tensor is a default-constructed glm::mat3, which GLM
    // 0.9.7 initialises to the identity, so the point-mass sum is added to I
    float I[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
    float maxc = -100000000.f;
    t->rel.resize(4 * (size_t)count);
    for (long i = 0; i < count; ++i) {
        const float x = p[3 * i] - sx, y = p[3 * i + 1] - sy, z = p[3 * i + 2] - sz;
        t->rel[4 * i] = x; t->rel[4 * i + 1] = y; t->rel[4 * i + 2] = z; t->rel[4 * i + 3] = 0.f;
        I[0] += y * y + z * z; I[3] -= x * y;         I[6] -= x * z;
        I[1] -= x * y;         I[4] += x * x + z * z; I[7] -= y * z;
        I[2] -= x * z;         I[5] -= y * z;         I[8] += x * x + y * y;
        maxc = maxc > x ? maxc : x;
        maxc = maxc > y ? maxc : y;
        maxc = maxc > z ? maxc : z;
    }
    // hard-coded tensors of the reference, then the inverse
    if (!std::strcmp(model, "banana") && near_scale(scale, 2.f)) set_diag(I, 0.0140314f, 0.0218576f, 0.0309458f);
    else if (!std::strcmp(model, "banana") && near_scale(scale, 1.5f)) set_diag(I, 0.00596544f, 0.00895867f, 0.0125508f);
    else if (!std::strcmp(model, "banana") && near_scale(scale, 1.f)) set_diag(I, 0.00321249f, 0.00464048f, 0.00641495f);
    if (!std::strcmp(model, "bunny") && near_scale(scale, 1.5f)) set_diag(I, 0.00989096f, 0.00979847f, 0.0119784f);
    if (!std::strcmp(model, "teapot") && near_scale(scale, 2.f)) set_diag(I, 0.0122465f, 0.0232885f, 0.0208828f);
    else if (!std::strcmp(model, "teapot") && near_scale(scale, 1.5f)) set_diag(I, 0.0075797f, 0.0140341f, 0.012614f);
    if (!inverse3(I, t->invInertia)) {
        delete t;
        return fail(ARSIM_ERR_INVALID_ARG, "template_load: singular inertia tensor in " + path);
    }
    // cubes: the inverse tensor itself is fixed, 6 / (edge)^2 on the diagonal (edge 0.1 * scale)
    if (!std::strcmp(model, "cube")) {
        const double edges[3] = {0.1, 0.2, 0.15};
        const float scales[3] = {1.f, 2.f, 1.5f};
        for (int k = 0; k < 3; ++k)
            if (near_scale(scale, scales[k])) {
                const float v = (float)(6.0 / (edges[k] * edges[k]));
                set_diag(t->invInertia, v, v, v);
            }
    }
    t->radius = maxc;
    t->mass = 1.f;
    *out = t;
    return ARSIM_OK;
}

extern "C" int arsim_template_info(const ArSimTemplate *t, int *numParticles, float *relPos, float *invInertia,
                                   float *radius, float *mass) {
    if (!t) return fail(ARSIM_ERR_INVALID_ARG, "template_info: null template");
    if (numParticles) *numParticles = (int)(t->rel.size() / 4);
    if (relPos) std::memcpy(relPos, t->rel.data(), t->rel.size() * sizeof(float));
    if (invInertia) std::memcpy(invInertia, t->invInertia, sizeof(t->invInertia));
    if (radius) *radius = t->radius;
    if (mass) *mass = t->mass;
    return ARSIM_OK;
}

extern "C" int arsim_template_free(ArSimTemplate *t) {
    delete t;
    return ARSIM_OK;
}

extern "C" int arsim_scene_create(const ArSimParams *params, const ArSimStepConfig *cfg, ArSimSceneOwned **out) {
    if (!params || !cfg || !out) return fail(ARSIM_ERR_INVALID_ARG, "scene_create: null argument");
    ArSimSceneOwned *s = new (std::nothrow) ArSimSceneOwned();
    if (!s) return fail(ARSIM_ERR_OUT_OF_MEMORY, "scene_create");
    s->params = *params;
    s->params.numBodies = 0;
    s->cfg = *cfg;
    *out = s;
    return ARSIM_OK;
}

// addRigidBody for a template instance (RigidBodyAddition.cpp:1760-1903): the body's particles go
// after the existing bodies' particles and before the free grains; world position = pos + rel
// (identity orientation), particle velocity = the body velocity.
extern "C" int arsim_scene_add_body(ArSimSceneOwned *s, const ArSimTemplate *t, const float *pos, const float *vel,
                                    const float *angVel) {
    if (!s || !t || !pos || !vel || !angVel) return fail(ARSIM_ERR_INVALID_ARG, "scene_add_body: null argument");
    const size_t n = t->rel.size() / 4;
    size_t at = 0;   // first grain = end of the body block
    for (int k : s->ppo) at += (size_t)k;
    std::vector<float> p(4 * n), v(4 * n);
    for (size_t i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c) {
            p[4 * i + c] = pos[c] + t->rel[4 * i + c];
            v[4 * i + c] = vel[c];
        }
        p[4 * i + 3] = 0.f;
        v[4 * i + 3] = 0.f;
    }
    s->pos.insert(s->pos.begin() + 4 * at, p.begin(), p.end());
    s->vel.insert(s->vel.begin() + 4 * at, v.begin(), v.end());
    s->bodyFrame.insert(s->bodyFrame.begin() + 4 * at, t->rel.begin(), t->rel.end());
    s->rb.insert(s->rb.begin() + at, n, s->B);
    s->rbPos.insert(s->rbPos.end(), {pos[0], pos[1], pos[2], 0.f});
    s->rbVel.insert(s->rbVel.end(), {vel[0], vel[1], vel[2], 0.f});
    s->rbAng.insert(s->rbAng.end(), {angVel[0], angVel[1], angVel[2], 0.f});
    s->rbQuat.insert(s->rbQuat.end(), {0.f, 0.f, 0.f, 1.f});
    s->rbInertia.insert(s->rbInertia.end(), t->invInertia, t->invInertia + 9);
    s->rbMass.push_back(t->mass);
    s->ppo.push_back((int)n);
    s->N += (int)n;
    s->B += 1;
    s->params.numBodies = (unsigned)s->B;
    return ARSIM_OK;
}

// Free grains (rbIndex -1) appended after everything else; pos4 / vel4 are float4 per grain.
extern "C" int arsim_scene_add_grains(ArSimSceneOwned *s, int n, const float *pos4, const float *vel4) {
    if (!s || n < 0 || (n > 0 && (!pos4 || !vel4))) return fail(ARSIM_ERR_INVALID_ARG, "scene_add_grains: bad arguments");
    s->pos.insert(s->pos.end(), pos4, pos4 + 4 * (size_t)n);
    s->vel.insert(s->vel.end(), vel4, vel4 + 4 * (size_t)n);
    s->bodyFrame.insert(s->bodyFrame.end(), 4 * (size_t)n, 0.f);
    s->rb.insert(s->rb.end(), (size_t)n, -1);
    s->N += n;
    return ARSIM_OK;
}
