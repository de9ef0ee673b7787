// capi.cpp -- extern "C" stage functions and device-memory helpers of libarsim.
//
// Each stage function validates its arguments, converts the reference's SimParams into the
// kernels' parameter structs and enqueues the kernels on the caller's stream.  None of them
// allocates device memory or waits for the device.
#include <cmath>
#include <cstring>
#include <string>

#include "host_common.hpp"
#include "kernels.hpp"
#include "params.hpp"

namespace arsim_host {

static thread_local std::string g_lastError;

int fail(int code, const std::string &msg) {
    g_lastError = msg;
    return code;
}

}  // namespace arsim_host

using namespace arsim;
using arsim_host::fail;

namespace {

inline hipStream_t S(arsim_stream_t s) { return (hipStream_t)s; }

// Launch errors surface through hipGetLastError right after the enqueue.
int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ARSIM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return ARSIM_OK;
}

template <class K>
int sort_common(K **keys, unsigned **vals, K **altKeys, unsigned **altVals, int n, int beginBit, int endBit,
                       void *ws, size_t wsBytes, arsim_stream_t st) {
    if (!keys || !vals || !altKeys || !altVals || !*keys || !*vals || !*altKeys || !*altVals)
        return fail(ARSIM_ERR_INVALID_ARG, "sort_pairs: null buffer");
    if (n < 0 || n >= (1 << 30)) return fail(ARSIM_ERR_INVALID_ARG, "sort_pairs: item count out of range");
    if (beginBit < 0 || endBit > (int)sizeof(K) * 8 || beginBit >= endBit)
        return fail(ARSIM_ERR_INVALID_ARG, "sort_pairs: bad bit range");
    if (n == 0) return ARSIM_OK;
    if (!ws || wsBytes < sort_workspace_bytes(n, (int)sizeof(K)))
        return fail(ARSIM_ERR_WORKSPACE, "sort_pairs: workspace too small");
    SortTuning tu;
    int rc = sort_pairs<K>(keys, vals, altKeys, altVals, n, beginBit, endBit, false, ws, wsBytes, tu, S(st));
    if (rc) return fail(ARSIM_ERR_INVALID_ARG, "sort_pairs: rejected arguments");
    return check_launch("sort_pairs");
}
}  // namespace

extern "C" {

int arsim_abi_version(void) { return 2; }

const char *arsim_last_error(void) { return arsim_host::g_lastError.c_str(); }

int arsim_device_count(int *count) {
    if (!count) return fail(ARSIM_ERR_INVALID_ARG, "null count");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) { *count = 0; return fail(ARSIM_ERR_NO_DEVICE, hipGetErrorString(e)); }
    *count = c;
    return ARSIM_OK;
}
int arsim_set_device(int device) { ARSIM_HIP(hipSetDevice(device)); return ARSIM_OK; }
int arsim_malloc(void **dptr, size_t bytes) {
    if (!dptr) return fail(ARSIM_ERR_INVALID_ARG, "null dptr");
    ARSIM_HIP(hipMalloc(dptr, bytes ? bytes : 4));
    return ARSIM_OK;
}
int arsim_free(void *dptr) { ARSIM_HIP(hipFree(dptr)); return ARSIM_OK; }
int arsim_memcpy_htod(void *dst, const void *src, size_t bytes, arsim_stream_t st) {
    ARSIM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(st)));
    ARSIM_HIP(hipStreamSynchronize(S(st)));
    return ARSIM_OK;
}
int arsim_memcpy_dtoh(void *dst, const void *src, size_t bytes, arsim_stream_t st) {
    ARSIM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(st)));
    ARSIM_HIP(hipStreamSynchronize(S(st)));
    return ARSIM_OK;
}
int arsim_memcpy_dtod(void *dst, const void *src, size_t bytes, arsim_stream_t st) {
    ARSIM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(st)));
    return ARSIM_OK;
}
int arsim_memset(void *dptr, int value, size_t bytes, arsim_stream_t st) {
    ARSIM_HIP(hipMemsetAsync(dptr, value, bytes, S(st)));
    return ARSIM_OK;
}
int arsim_stream_create(arsim_stream_t *st) {
    if (!st) return fail(ARSIM_ERR_INVALID_ARG, "null stream out");
    hipStream_t s;
    ARSIM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *st = (arsim_stream_t)s;
    return ARSIM_OK;
}
int arsim_stream_destroy(arsim_stream_t st) { ARSIM_HIP(hipStreamDestroy(S(st))); return ARSIM_OK; }
int arsim_stream_synchronize(arsim_stream_t st) { ARSIM_HIP(hipStreamSynchronize(S(st))); return ARSIM_OK; }

/* ---------------------------------------------------------------- F2 */
int arsim_integrate_bodies(float *rbPositions, float *rbVelocities, float *rbAngularVelocity, float *rbQuaternion,
                           float *cumulativeQuaternion, const float *rbInertia, float *rbCurrentInertia,
                           float *modelMatrix, int numBodies, const ArSimParams *params, const ArSimStepConfig *cfg,
                           arsim_stream_t st) {
    if (numBodies < 0 || !params || !cfg) return fail(ARSIM_ERR_INVALID_ARG, "integrate_bodies: bad arguments");
    if (numBodies == 0) return ARSIM_OK;
    if (!rbPositions || !rbVelocities || !rbAngularVelocity || !rbQuaternion || !cumulativeQuaternion || !rbInertia ||
        !rbCurrentInertia)
        return fail(ARSIM_ERR_INVALID_ARG, "integrate_bodies: null buffer");
    BodyBufs b{};
    b.pos = (float4 *)rbPositions; b.vel = (float4 *)rbVelocities; b.ang = (float4 *)rbAngularVelocity;
    b.qinc = (float4 *)rbQuaternion; b.qcum = (float4 *)cumulativeQuaternion;
    b.Ib = rbInertia; b.Iw = rbCurrentInertia; b.model = modelMatrix;
    launch_body_integrate(b, numBodies, make_phys(*params, *cfg), nullptr, S(st));
    return check_launch("integrate_bodies");
}

/* ---------------------------------------------------------------- F3 */
int arsim_particle_pose(const float *rbPositions, const float *rbVelocities, const float *rbAngularVelocity,
                        const float *cumulativeQuaternion, const float *bodyFramePos, const int *rbIndices,
                        float *relativePos, float *pos, float *vel, int numParticles, arsim_stream_t st) {
    if (numParticles < 0) return fail(ARSIM_ERR_INVALID_ARG, "particle_pose: negative count");
    if (numParticles == 0) return ARSIM_OK;
    if (!rbPositions || !rbVelocities || !rbAngularVelocity || !cumulativeQuaternion || !bodyFramePos || !rbIndices ||
        !relativePos || !pos || !vel)
        return fail(ARSIM_ERR_INVALID_ARG, "particle_pose: null buffer");
    BodyBufs b{};
    b.pos = (float4 *)rbPositions; b.vel = (float4 *)rbVelocities; b.ang = (float4 *)rbAngularVelocity;
    b.qcum = (float4 *)cumulativeQuaternion;
    PartBufs p{};
    p.pos = (float4 *)pos; p.vel = (float4 *)vel; p.relPos = (float4 *)relativePos;
    p.bodyFrame = (const float4 *)bodyFramePos; p.rb = rbIndices;
    SceneView sv{};
    sv.count = 1;
    DevPhys ph{};
    KeyOut ko{};
    launch_pose_key(p, b, numParticles, ph, sv, kKeyNone, ko, true, false, S(st));
    return check_launch("particle_pose");
}

/* ---------------------------------------------------------------- F4 */
int arsim_integrate_grains(float *pos, float *vel, const int *rbIndices, int numParticles, const ArSimParams *params,
                           const ArSimStepConfig *cfg, arsim_stream_t st) {
    if (numParticles < 0 || !params || !cfg) return fail(ARSIM_ERR_INVALID_ARG, "integrate_grains: bad arguments");
    if (numParticles == 0) return ARSIM_OK;
    if (!pos || !vel || !rbIndices) return fail(ARSIM_ERR_INVALID_ARG, "integrate_grains: null buffer");
    PartBufs p{};
    p.pos = (float4 *)pos; p.vel = (float4 *)vel; p.rb = rbIndices;
    BodyBufs b{};
    SceneView sv{};
    sv.count = 1;
    int rc = make_geom(*params, *cfg, 0, numParticles, sv.single);
    if (rc) return rc;
    KeyOut ko{};
    launch_pose_key(p, b, numParticles, make_phys(*params, *cfg), sv, kKeyNone, ko, false, true, S(st));
    return check_launch("integrate_grains");
}

/* ---------------------------------------------------------------- F5 */
int arsim_calc_hash(unsigned int *hash, unsigned int *index, const float *pos, int numParticles,
                    const ArSimParams *params, arsim_stream_t st) {
    if (numParticles < 0 || !params) return fail(ARSIM_ERR_INVALID_ARG, "calc_hash: bad arguments");
    if (numParticles == 0) return ARSIM_OK;
    if (!hash || !index || !pos) return fail(ARSIM_ERR_INVALID_ARG, "calc_hash: null buffer");
    ArSimStepConfig cfg{};
    SceneView sv{};
    sv.count = 1;
    int rc = make_geom(*params, cfg, 0, numParticles, sv.single);
    if (rc) return rc;
    PartBufs p{};
    p.pos = (float4 *)pos;
    BodyBufs b{};
    KeyOut ko{hash, index, nullptr};
    DevPhys ph{};
    launch_pose_key(p, b, numParticles, ph, sv, kKeyGrid, ko, false, false, S(st));
    return check_launch("calc_hash");
}

/* ---------------------------------------------------------------- F6 / F13 */
size_t arsim_sort_workspace_bytes(int numItems, int keyBytes) {
    if (numItems < 0 || (keyBytes != 4 && keyBytes != 8)) return 0;
    return sort_workspace_bytes(numItems, keyBytes);
}

int arsim_sort_pairs_u32(unsigned int **keys, unsigned int **values, unsigned int **altKeys, unsigned int **altValues,
                         int numItems, int beginBit, int endBit, void *workspace, size_t workspaceBytes,
                         arsim_stream_t st) {
    return sort_common<unsigned>(keys, values, altKeys, altValues, numItems, beginBit, endBit, workspace,
                                 workspaceBytes, st);
}
int arsim_sort_pairs_u64(uint64_t **keys, unsigned int **values, uint64_t **altKeys, unsigned int **altValues,
                         int numItems, int beginBit, int endBit, void *workspace, size_t workspaceBytes,
                         arsim_stream_t st) {
    return sort_common<uint64_t>(keys, values, altKeys, altValues, numItems, beginBit, endBit, workspace,
                                 workspaceBytes, st);
}

int arsim_sort_pairs_u32_noalloc(unsigned int **keys, unsigned int **values, unsigned int **altKeys,
                                 unsigned int **altValues, int numItems, int beginBit, int endBit,
                                 arsim_stream_t st) {
    if (numItems > ARSIM_SORT_NOALLOC_MAX_ITEMS)
        return fail(ARSIM_ERR_WORKSPACE, "sort_pairs_u32_noalloc: " + std::to_string(numItems) +
                                             " items exceed the static scratch's limit of " +
                                             std::to_string(ARSIM_SORT_NOALLOC_MAX_ITEMS));
    if (numItems <= 0)
        return sort_common<unsigned>(keys, values, altKeys, altValues, numItems, beginBit, endBit, nullptr, 0, st);
    void *ws = nullptr;
    size_t wsBytes = 0;
    hipError_t e = sort_static_scratch(&ws, &wsBytes);
    if (e != hipSuccess)
        return fail(ARSIM_ERR_HIP, std::string("sort_pairs_u32_noalloc: static scratch lookup: ") + hipGetErrorString(e));
    return sort_common<unsigned>(keys, values, altKeys, altValues, numItems, beginBit, endBit, ws, wsBytes, st);
}

/* ---------------------------------------------------------------- F7 */
int arsim_reorder_find_cell_start(const int *rbIndices, unsigned int *cellStart, unsigned int *cellEnd,
                                  float *sortedPos, float *sortedVel, const unsigned int *hash,
                                  const unsigned int *index, const float *pos, const float *vel, int numParticles,
                                  unsigned int numCells, arsim_stream_t st) {
    if (numParticles < 0) return fail(ARSIM_ERR_INVALID_ARG, "reorder: negative count");
    if (!rbIndices || !cellStart || !cellEnd || !sortedPos || !sortedVel || !hash || !index || !pos || !vel)
        return fail(ARSIM_ERR_INVALID_ARG, "reorder: null buffer");
    launch_fill_u32(cellStart, kEmpty, numCells, S(st));
    launch_reorder(hash, index, (const float4 *)pos, (const float4 *)vel, rbIndices, nullptr, (float4 *)sortedPos,
                   (float4 *)sortedVel, cellStart, cellEnd, numParticles, S(st));
    return check_launch("reorder");
}

/* ---------------------------------------------------------------- F8 / F9 / F18 */
int arsim_dem_grid(const int *rbIndices, float *pForce, float *pTorque, float *vel, const float *sortedPos,
                   const float *sortedVel, const float *relativePos, const unsigned int *index,
                   const unsigned int *cellStart, const unsigned int *cellEnd, int numParticles,
                   const ArSimParams *params, const ArSimStepConfig *cfg, arsim_stream_t st) {
    (void)rbIndices;   // the rbIndex of every sorted particle travels in sortedPos.w
    if (numParticles < 0 || !params || !cfg) return fail(ARSIM_ERR_INVALID_ARG, "dem_grid: bad arguments");
    if (numParticles == 0) return ARSIM_OK;
    if (!pForce || !vel || !sortedPos || !sortedVel || !index || !cellStart || !cellEnd)
        return fail(ARSIM_ERR_INVALID_ARG, "dem_grid: null buffer");
    if (pTorque && !relativePos) return fail(ARSIM_ERR_INVALID_ARG, "dem_grid: pTorque needs relativePos");
    SceneView sv{};
    sv.count = 1;
    int rc = make_geom(*params, *cfg, 0, numParticles, sv.single);
    if (rc) return rc;
    DemOut o{};
    o.pForce = (float4 *)pForce; o.pTorque = (float4 *)pTorque; o.vel = (float4 *)vel;
    o.relPos = (const float4 *)relativePos;
    launch_dem_grid((const float4 *)sortedPos, (const float4 *)sortedVel, index, cellStart, cellEnd, numParticles,
                    make_phys(*params, *cfg), sv, false, o, 128, S(st));
    return check_launch("dem_grid");
}

/* ---------------------------------------------------------------- F10 */
int arsim_segmented_reduce_apply(const float *pForce, const float *pTorque, const float *relativePos,
                                 float *rbVelocities, float *rbAngularVelocity, const float *rbMass,
                                 const float *rbCurrentInertia, const int *particleStart,
                                 const int *particlesPerObject, int numBodies, const ArSimParams *params,
                                 const ArSimStepConfig *cfg, arsim_stream_t st) {
    if (numBodies < 0 || !params || !cfg) return fail(ARSIM_ERR_INVALID_ARG, "reduce: bad arguments");
    if (numBodies == 0) return ARSIM_OK;
    if (!pForce || (!pTorque && !relativePos) || !rbVelocities || !rbAngularVelocity || !rbMass || !rbCurrentInertia ||
        !particleStart || !particlesPerObject)
        return fail(ARSIM_ERR_INVALID_ARG, "reduce: null buffer");
    BodyBufs b{};
    b.vel = (float4 *)rbVelocities; b.ang = (float4 *)rbAngularVelocity; b.mass = rbMass;
    b.Iw = (float *)rbCurrentInertia; b.pstart = particleStart; b.pcount = particlesPerObject;
    launch_reduce_apply((const float4 *)pForce, (const float4 *)pTorque, (const float4 *)relativePos, b, numBodies,
                        make_phys(*params, *cfg), S(st));
    return check_launch("reduce");
}

/* ---------------------------------------------------------------- F11 */
size_t arsim_aabb_workspace_bytes(int numItems) {
    (void)numItems;
    return (size_t)aabb_max_partials() * 2 * sizeof(float4);
}
int arsim_aabb(const float *pos, int numItems, float *outMin4, float *outMax4, void *ws, size_t wsBytes,
               arsim_stream_t st) {
    if (numItems <= 0 || !pos || !outMin4 || !outMax4) return fail(ARSIM_ERR_INVALID_ARG, "aabb: bad arguments");
    if (!ws || wsBytes < arsim_aabb_workspace_bytes(numItems)) return fail(ARSIM_ERR_WORKSPACE, "aabb: workspace");
    launch_aabb((const float4 *)pos, numItems, (float4 *)outMin4, (float4 *)outMax4, (float4 *)ws,
                aabb_max_partials(), S(st));
    return check_launch("aabb");
}

/* ---------------------------------------------------------------- F12 */
static int morton_common(const float *pos, void *codes, int *indices, int n, const float boxMin[3],
                         const float boxMax[3], const float *box4, int mode, arsim_stream_t st) {
    if (n < 0 || !pos || !codes || !indices) return fail(ARSIM_ERR_INVALID_ARG, "morton: bad arguments");
    if (!box4 && (!boxMin || !boxMax)) return fail(ARSIM_ERR_INVALID_ARG, "morton: no box");
    if (n == 0) return ARSIM_OK;
    SceneView sv{};
    sv.count = 1;
    if (!box4)
        for (int a = 0; a < 3; ++a) {
            sv.single.mLo[a] = boxMin[a];
            sv.single.mExt[a] = boxMax[a] - boxMin[a];
        }
    PartBufs p{};
    p.pos = (float4 *)pos;
    BodyBufs b{};
    KeyOut ko{codes, (unsigned *)indices, (const float4 *)box4};
    DevPhys ph{};
    launch_pose_key(p, b, n, ph, sv, mode, ko, false, false, S(st));
    return check_launch("morton");
}
int arsim_morton30(const float *pos, unsigned int *codes, int *indices, int n, const float boxMin[3],
                   const float boxMax[3], const float *box4, arsim_stream_t st) {
    return morton_common(pos, codes, indices, n, boxMin, boxMax, box4, kKeyMorton30, st);
}
int arsim_morton63(const float *pos, uint64_t *codes, int *indices, int n, const float boxMin[3],
                   const float boxMax[3], const float *box4, arsim_stream_t st) {
    return morton_common(pos, codes, indices, n, boxMin, boxMax, box4, kKeyMorton63, st);
}

/* ---------------------------------------------------------------- F14 */
int arsim_lbvh_build(const void *sortedCodes, int keyBytes, int n, unsigned char *isLeaf, int *left, int *right,
                     int *parent, int *minRange, int *maxRange, int *nodeCounter, arsim_stream_t st) {
    if (n < 0 || (keyBytes != 4 && keyBytes != 8)) return fail(ARSIM_ERR_INVALID_ARG, "lbvh_build: bad arguments");
    if (n == 0) return ARSIM_OK;
    if (!sortedCodes || !left || !right || !parent || !nodeCounter)
        return fail(ARSIM_ERR_INVALID_ARG, "lbvh_build: null buffer");
    Lbvh t{};
    t.left = left; t.right = right; t.parent = parent; t.minR = minRange; t.maxR = maxRange; t.isLeaf = isLeaf;
    t.counter = nodeCounter;
    if ((minRange == nullptr) != (maxRange == nullptr))
        return fail(ARSIM_ERR_INVALID_ARG, "lbvh_build: minRange and maxRange go together");
    launch_lbvh_build(sortedCodes, keyBytes, n, 0, nullptr, nullptr, 1, t, S(st));
    return check_launch("lbvh_build");
}

/* ---------------------------------------------------------------- F15 / F16 */
int arsim_lbvh_refit(const int *sortedIndices, const float *pos, const float *vel, const int *rbIndices, float radius,
                     int n, unsigned char *isLeaf, int *minRange, int *maxRange, const int *left, const int *right,
                     const int *parent, float *CMs, float *sortedVel, float *bounds, int *nodeCounter,
                     arsim_stream_t st) {
    if (n < 0) return fail(ARSIM_ERR_INVALID_ARG, "lbvh_refit: negative count");
    if (n == 0) return ARSIM_OK;
    if (!sortedIndices || !pos || !left || !right || !parent || !CMs || !bounds || !nodeCounter)
        return fail(ARSIM_ERR_INVALID_ARG, "lbvh_refit: null buffer");
    if (sortedVel && !vel) return fail(ARSIM_ERR_INVALID_ARG, "lbvh_refit: sortedVel needs vel");
    if ((minRange == nullptr) != (maxRange == nullptr))
        return fail(ARSIM_ERR_INVALID_ARG, "lbvh_refit: minRange and maxRange go together");
    Lbvh t{};
    t.left = (int *)left; t.right = (int *)right; t.parent = (int *)parent; t.minR = minRange; t.maxR = maxRange;
    t.isLeaf = isLeaf; t.bounds = (float4 *)bounds; t.cms = (float4 *)CMs; t.sortedVel = (float4 *)sortedVel;
    t.counter = nodeCounter;
    launch_lbvh_refit(sortedIndices, (const float4 *)pos, (const float4 *)vel, rbIndices, nullptr, nullptr, radius, n, t,
                      S(st));
    return check_launch("lbvh_refit");
}

/* ---------------------------------------------------------------- F17 */
int arsim_dem_lbvh(const int *rbIndices, float *pForce, float *pTorque, float *vel, const float *sortedVel,
                   const float *relativePos, const int *sortedIndices, const int *left, const int *right,
                   const float *CMs, const float *bounds, int n, const ArSimParams *params,
                   const ArSimStepConfig *cfg, int *overflowFlag, arsim_stream_t st) {
    (void)rbIndices;   // the rbIndex of every leaf travels in CMs.w
    if (n < 0 || !params || !cfg) return fail(ARSIM_ERR_INVALID_ARG, "dem_lbvh: bad arguments");
    if (n == 0) return ARSIM_OK;
    if (!pForce || !vel || !sortedVel || !sortedIndices || !left || !right || !CMs || !bounds || !overflowFlag)
        return fail(ARSIM_ERR_INVALID_ARG, "dem_lbvh: null buffer");
    if (pTorque && !relativePos) return fail(ARSIM_ERR_INVALID_ARG, "dem_lbvh: pTorque needs relativePos");
    SceneView sv{};
    sv.count = 1;
    int rc = make_geom(*params, *cfg, 0, n, sv.single);
    if (rc) return rc;
    Lbvh t{};
    t.left = (int *)left; t.right = (int *)right; t.cms = (float4 *)CMs; t.bounds = (float4 *)bounds;
    t.sortedVel = (float4 *)sortedVel;
    DemOut o{};
    o.pForce = (float4 *)pForce; o.pTorque = (float4 *)pTorque; o.vel = (float4 *)vel;
    o.relPos = (const float4 *)relativePos;
    launch_dem_lbvh(sortedIndices, t, n, make_phys(*params, *cfg), sv, false, o, 128, S(st), overflowFlag);
    return check_launch("dem_lbvh");
}

int arsim_rebase_indices(int *dst, const int *src, int n, int delta, arsim_stream_t st) {
    if (n < 0 || (n > 0 && (!dst || !src))) return fail(ARSIM_ERR_INVALID_ARG, "rebase_indices: bad arguments");
    launch_copy_offset(dst, src, n, delta, S(st));
    return check_launch("rebase_indices");
}

/* ---------------------------------------------------------------- contact pairs (test mode) */
int arsim_contact_pairs_grid(const int *rbIndices, const float *sortedPos, const unsigned int *index,
                             const unsigned int *cellStart, const unsigned int *cellEnd, int n,
                             const ArSimParams *params, int *pairs, int *counts, int maxPer, arsim_stream_t st) {
    (void)rbIndices;
    if (n < 0 || !params || maxPer <= 0) return fail(ARSIM_ERR_INVALID_ARG, "contact_pairs_grid: bad arguments");
    if (n == 0) return ARSIM_OK;
    if (!sortedPos || !index || !cellStart || !cellEnd || !pairs || !counts)
        return fail(ARSIM_ERR_INVALID_ARG, "contact_pairs_grid: null buffer");
    ArSimStepConfig cfg{};
    SceneView sv{};
    sv.count = 1;
    int rc = make_geom(*params, cfg, 0, n, sv.single);
    if (rc) return rc;
    DemOut o{};
    o.pairs = pairs; o.counts = counts; o.maxPer = maxPer;
    // sortedVel is not read in pair mode; sortedPos stands in for it
    launch_dem_grid((const float4 *)sortedPos, (const float4 *)sortedPos, index, cellStart, cellEnd, n,
                    make_phys(*params, cfg), sv, false, o, 128, S(st));
    return check_launch("contact_pairs_grid");
}
int arsim_contact_pairs_lbvh(const int *rbIndices, const int *sortedIndices, const int *left, const int *right,
                             const float *CMs, const float *bounds, int n, const ArSimParams *params,
                             int *overflowFlag, int *pairs, int *counts, int maxPer, arsim_stream_t st) {
    (void)rbIndices;
    if (n < 0 || !params || maxPer <= 0) return fail(ARSIM_ERR_INVALID_ARG, "contact_pairs_lbvh: bad arguments");
    if (n == 0) return ARSIM_OK;
    if (!sortedIndices || !left || !right || !CMs || !bounds || !pairs || !counts || !overflowFlag)
        return fail(ARSIM_ERR_INVALID_ARG, "contact_pairs_lbvh: null buffer");
    ArSimStepConfig cfg{};
    SceneView sv{};
    sv.count = 1;
    int rc = make_geom(*params, cfg, 0, n, sv.single);
    if (rc) return rc;
    Lbvh t{};
    t.left = (int *)left; t.right = (int *)right; t.cms = (float4 *)CMs; t.bounds = (float4 *)bounds;
    t.sortedVel = (float4 *)CMs;   // not read in pair mode beyond w
    DemOut o{};
    o.pairs = pairs; o.counts = counts; o.maxPer = maxPer;
    launch_dem_lbvh(sortedIndices, t, n, make_phys(*params, cfg), sv, false, o, 128, S(st), overflowFlag);
    return check_launch("contact_pairs_lbvh");
}

/* ---------------------------------------------------------------- static collider set */
int arsim_range_image_load(const uint16_t *depth, int width, int height, const float *cameraTransformation,
                           float particleRadius, float *staticPos, float *staticNorm, float *staticRadii,
                           float *vaoData, arsim_stream_t st) {
    if (width <= 2 || height <= 2 || !depth || !cameraTransformation || !staticPos || !staticNorm)
        return fail(ARSIM_ERR_INVALID_ARG, "range_image_load: bad arguments");
    RangeCamera cam{};
    std::memcpy(cam.m, cameraTransformation, sizeof(cam.m));
    const double g0 = 0.478957, g1 = -8.010560, g2 = 4.166928;   // customRangeImageLoader.cu:76-78
    const double gl = std::sqrt(g0 * g0 + g1 * g1 + g2 * g2);
    cam.g[0] = (float)(g0 / gl); cam.g[1] = (float)(g1 / gl); cam.g[2] = (float)(g2 / gl);
    launch_range_image(depth, width, height, cam, particleRadius, (float4 *)staticPos, (float4 *)staticNorm,
                       staticRadii, vaoData, S(st));
    return check_launch("range_image_load");
}

int arsim_dem_static_grid(float *pForce, float *pTorque, const float *sortedPos, const float *sortedVel,
                          const float *relativePos, const float *staticSortedPos, const float *staticNorm,
                          const unsigned int *sortedIndex, const unsigned int *staticSortedIndex,
                          const unsigned int *staticCellStart, const unsigned int *staticCellEnd, int numParticles,
                          const ArSimParams *params, const ArSimStepConfig *cfg, arsim_stream_t st) {
    if (numParticles < 0 || !params || !cfg) return fail(ARSIM_ERR_INVALID_ARG, "dem_static_grid: bad arguments");
    if (numParticles == 0) return ARSIM_OK;
    if (!pForce || !pTorque || !sortedPos || !sortedVel || !relativePos || !staticSortedPos || !staticNorm ||
        !sortedIndex || !staticSortedIndex || !staticCellStart || !staticCellEnd)
        return fail(ARSIM_ERR_INVALID_ARG, "dem_static_grid: null buffer");
    SceneGeom g{};
    int rc = make_geom(*params, *cfg, 0, numParticles, g);
    if (rc) return rc;
    launch_dem_static((const float4 *)sortedPos, (const float4 *)sortedVel, sortedIndex, numParticles,
                      (const float4 *)staticSortedPos, (const float4 *)staticNorm, staticSortedIndex, staticCellStart,
                      staticCellEnd, make_phys(*params, *cfg), g, (float4 *)pForce, (float4 *)pTorque,
                      (const float4 *)relativePos, true, S(st));
    return check_launch("dem_static_grid");
}

int arsim_find_deepest_contacts_grid(const int *rbIndices, int *collidingRigidBodyIndex, int *collidingParticleIndex,
                                     float *contactDistance, float *color, const float *sortedPos,
                                     const unsigned int *gridParticleIndex, const unsigned int *cellStart,
                                     const unsigned int *cellEnd, int numParticles, const ArSimParams *params,
                                     arsim_stream_t st) {
    if (numParticles < 0 || !params) return fail(ARSIM_ERR_INVALID_ARG, "find_deepest_contacts_grid: bad arguments");
    if (numParticles == 0) return ARSIM_OK;
    if (!rbIndices || !collidingRigidBodyIndex || !collidingParticleIndex || !contactDistance || !sortedPos ||
        !gridParticleIndex || !cellStart || !cellEnd)
        return fail(ARSIM_ERR_INVALID_ARG, "find_deepest_contacts_grid: null buffer");
    ArSimStepConfig cfg{};
    SceneGeom g{};
    int rc = make_geom(*params, cfg, 0, numParticles, g);
    if (rc) return rc;
    launch_sis_find_ref((const float4 *)sortedPos, gridParticleIndex, rbIndices, cellStart, cellEnd, g, numParticles,
                        params->particleRadius, contactDistance, collidingRigidBodyIndex, collidingParticleIndex,
                        (float4 *)color, S(st));
    return check_launch("find_deepest_contacts_grid");
}

int arsim_find_static_contacts_grid(int *collidingParticleIndex, float *contactDistance, float *color,
                                    const float *sortedPos, const float *staticSortedPos,
                                    const unsigned int *gridParticleIndex, const unsigned int *staticSortedIndex,
                                    const unsigned int *staticCellStart, const unsigned int *staticCellEnd,
                                    int numParticles, const ArSimParams *params, arsim_stream_t st) {
    if (numParticles < 0 || !params) return fail(ARSIM_ERR_INVALID_ARG, "find_static_contacts_grid: bad arguments");
    if (numParticles == 0) return ARSIM_OK;
    if (!collidingParticleIndex || !contactDistance || !sortedPos || !staticSortedPos || !gridParticleIndex ||
        !staticSortedIndex || !staticCellStart || !staticCellEnd)
        return fail(ARSIM_ERR_INVALID_ARG, "find_static_contacts_grid: null buffer");
    ArSimStepConfig cfg{};
    SceneGeom g{};
    int rc = make_geom(*params, cfg, 0, numParticles, g);
    if (rc) return rc;
    launch_sis_find_static((const float4 *)sortedPos, gridParticleIndex, numParticles, (const float4 *)staticSortedPos,
                           staticSortedIndex, staticCellStart, staticCellEnd, g, params->particleRadius, true,
                           contactDistance, collidingParticleIndex, (float4 *)color, S(st));
    return check_launch("find_static_contacts_grid");
}

size_t arsim_sis_workspace_bytes(int maxContacts, int numBodyParticles) {
    if (maxContacts < 1 || numBodyParticles < 0) return 0;
    return sis_workspace_bytes(maxContacts, 1, numBodyParticles);
}

int arsim_sis_gather(const float *rbPositions, const float *rbVelocities, const float *rbAngularVelocity,
                     const float *relativePos, const int *rbIndices, int numBodyParticles,
                     const float *staticContactDistance, const int *staticContactIndex, const float *staticNorm,
                     const float *contactDistance, const int *collidingRigidBodyIndex,
                     const int *collidingParticleIndex, float *contactNormal, float *contactPoint,
                     int *contactBodies, float *contactBias, float *contactImpulse, int *numContacts,
                     int maxContacts, const ArSimParams *params, void *workspace, size_t workspaceBytes,
                     arsim_stream_t st) {
    if (numBodyParticles < 0 || maxContacts < 1 || !params)
        return fail(ARSIM_ERR_INVALID_ARG, "sis_gather: bad arguments");
    if (!rbPositions || !rbVelocities || !rbAngularVelocity || !relativePos || !rbIndices || !contactDistance ||
        !collidingRigidBodyIndex || !collidingParticleIndex || !contactNormal || !contactPoint || !contactBodies ||
        !contactBias || !contactImpulse || !numContacts || !workspace)
        return fail(ARSIM_ERR_INVALID_ARG, "sis_gather: null buffer");
    if (staticContactDistance && (!staticContactIndex || !staticNorm))
        return fail(ARSIM_ERR_INVALID_ARG, "sis_gather: static contacts need their indices and normals");
    const long long need = (long long)numBodyParticles * (staticContactDistance ? 2 : 1);
    if (need > maxContacts) return fail(ARSIM_ERR_OVERFLOW, "sis_gather: maxContacts below the worst case");
    if (workspaceBytes < sis_workspace_bytes(maxContacts, 1, numBodyParticles))
        return fail(ARSIM_ERR_WORKSPACE, "sis_gather: workspace too small");
    const SisWork w = sis_work_layout(workspace, maxContacts, 1, numBodyParticles);
    const SisGatherIn gi{staticContactDistance, staticContactIndex, (const float4 *)staticNorm, contactDistance,
                         collidingRigidBodyIndex, collidingParticleIndex, rbIndices, nullptr,
                         (const float4 *)rbPositions, (const float4 *)rbVelocities, (const float4 *)rbAngularVelocity,
                         (const float4 *)relativePos, params->particleRadius, params->ARrestitution,
                         params->RBrestitution};
    const SisContacts ct{(float4 *)contactNormal, (float4 *)contactPoint, (int2 *)contactBodies, contactBias,
                         (float4 *)contactImpulse};
    launch_sis_gather(gi, numBodyParticles, ct, w, false, S(st));
    ARSIM_HIP(hipMemcpyAsync(numContacts, w.numContacts, sizeof(int), hipMemcpyDeviceToDevice, S(st)));
    return check_launch("sis_gather");
}

int arsim_sis_solve(float *contactNormal, float *contactPoint, int *contactBodies, float *contactBias,
                    float *contactImpulse, const int *numContacts, int maxContacts, const float *rbPositions,
                    float *rbVelocities, float *rbAngularVelocity, const float *rbCurrentInertia, const float *rbMass,
                    int numBodies, int iterations, const ArSimParams *params, void *workspace,
                    size_t workspaceBytes, arsim_stream_t st) {
    if (maxContacts < 1 || numBodies < 0 || iterations < 0 || !params)
        return fail(ARSIM_ERR_INVALID_ARG, "sis_solve: bad arguments");
    if (!contactNormal || !contactPoint || !contactBodies || !contactBias || !contactImpulse || !numContacts ||
        !rbPositions || !rbVelocities || !rbAngularVelocity || !rbCurrentInertia || !rbMass || !workspace)
        return fail(ARSIM_ERR_INVALID_ARG, "sis_solve: null buffer");
    if (workspaceBytes < sis_workspace_bytes(maxContacts, 1, 0))
        return fail(ARSIM_ERR_WORKSPACE, "sis_solve: workspace too small");
    const SisWork w = sis_work_layout(workspace, maxContacts, 1, 0);
    ARSIM_HIP(hipMemcpyAsync(w.numContacts, numContacts, sizeof(int), hipMemcpyDeviceToDevice, S(st)));
    const SisContacts ct{(float4 *)contactNormal, (float4 *)contactPoint, (int2 *)contactBodies, contactBias,
                         (float4 *)contactImpulse};
    const SisBodies bd{(const float4 *)rbPositions, (float4 *)rbVelocities, (float4 *)rbAngularVelocity,
                       rbCurrentInertia, rbMass};
    const SisSolvePhys sp{params->ARfriction, params->gravity.y};
    SortTuning stt{};
    if (launch_sis_solve(ct, bd, sp, maxContacts, 1, numBodies, nullptr, numBodies, iterations, w, stt, S(st)))
        return fail(ARSIM_ERR_INVALID_ARG, "sis_solve: contact sort rejected");
    return check_launch("sis_solve");
}

}  // extern "C"
