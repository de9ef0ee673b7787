// kernels.hpp -- host-side launch interface of the gfx950 kernels (internal to libarsim).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "arsim_device.hpp"
#include "csr_fill.hpp"

namespace arsim {

struct SortTuning {
    int ipt = 16;        // items per thread
    int radixBits = 8;   // digit width
};

// ---------------------------------------------------------------- sort (sort.hip)
int sort_max_passes(int keyBytes);
size_t sort_workspace_bytes(int n, int keyBytes);
bool sort_tuning_supported(int ipt, int rbits);
// Returns 0 on success; swaps *keys/*alt and *vals/*altVals once per pass so that on return
// *keys / *vals hold the sorted data.  valsIota: the input values are 0..n-1 (not read).
template <class K>
int sort_pairs(K **keys, unsigned **vals, K **alt, unsigned **altVals, int n, int beginBit, int endBit,
               bool valsIota, void *ws, size_t wsBytes, const SortTuning &tu, hipStream_t st, bool prepared = false);
// Where the sort keeps its digit histograms, tile counters and look-back words inside the workspace.
// A caller may zero {hist, counters, statusA} and accumulate the histograms itself (for instance while
// it generates the keys) and then call sort_pairs(..., prepared = true), which skips both steps; the
// histogram of pass p is that of the digit at beginBit + p * radixBits.
struct SortPlan {
    unsigned *hist, *counters, *statusA, *statusB;
    int numTiles, passes, radix;
    size_t histWords, statusWords;
};
SortPlan sort_plan(void *ws, int n, int keyBytes, int beginBit, int endBit, const SortTuning &tu);
// Device address and size of the sort's static scratch array on the current device, looked up once
// per device.  It serves as the workspace of u32 sorts of up to ARSIM_SORT_NOALLOC_MAX_ITEMS items.
hipError_t sort_static_scratch(void **ptr, size_t *bytes);

// ---------------------------------------------------------------- buffers
struct BodyBufs {
    float4 *pos, *vel, *ang;      // rbPositions, rbVelocities, rbAngularVelocity
    float4 *qinc, *qcum;          // rbQuaternion (incremental), cumulative quaternion
    const float *Ib;              // inverse body-frame inertia, 9 per body, column-major
    float *Iw;                    // inverse world-frame inertia (rbCurrentInertia)
    float *model;                 // optional column-major mat4 per body
    const float *mass;
    const int *pstart, *pcount;
};

struct PartBufs {
    float4 *pos, *vel, *relPos;
    const float4 *bodyFrame;
    const int *rb;
    const uint16_t *scene;        // per-particle scene id (NULL when there is one scene)
    // optional gather record, written by the key kernel in the original order: rec[2i] = (pos.xyz,
    // rbIndex bits), rec[2i+1] = (vel.xyz, scene id bits or vel.w); the reorder / refit then gather
    // one 32-byte record per particle instead of pos, vel, rbIndex and scene separately
    float4 *rec;
};

enum KeyMode { kKeyNone = 0, kKeyGrid = 1, kKeyMorton30 = 2, kKeyMorton63 = 3, kKeyForest30 = 4 };

struct KeyOut {
    void *keys;                   // u32 (grid / morton30) or u64 (morton63 / forest30)
    unsigned *index;              // may be NULL when the sort generates the identity
    const float4 *box;            // dynamic Morton box (min, max) on the device, or NULL
    unsigned *hist = nullptr;     // optional: the sort's digit histograms, accumulated here
    int histPasses = 0, histRbits = 8, histEndBit = 32;
    int histBase = 0;             // bit of the first digit
    bool grainVelDead = false;    // the step's DEM rewrites every grain's velocity: skip writing it here
};

// Up to four device ranges zeroed by extra blocks of the body-integration launch.
struct ZeroJobs {
    unsigned *p[4] = {nullptr, nullptr, nullptr, nullptr};
    size_t count[4] = {0, 0, 0, 0};
};

// ---------------------------------------------------------------- particle / body stages (body.hip)
// out: where the integrated state goes (NULL: in place, into b)
void launch_body_integrate(const BodyBufs &b, int numBodies, const DevPhys &ph, const ZeroJobs *zero,
                           hipStream_t st, const BodyBufs *out = nullptr);
// doBodies / doGrains select which particles are updated; mode selects the key written.
void launch_pose_key(const PartBufs &p, const BodyBufs &b, int numParticles, const DevPhys &ph,
                     const SceneView &sv, int mode, const KeyOut &ko, bool doBodies, bool doGrains,
                     hipStream_t st);
void launch_reduce_apply(const float4 *pForce, const float4 *pTorque, const float4 *relPos, const BodyBufs &b,
                         int numBodies, const DevPhys &ph, hipStream_t st);
// The reduction of a step whose bodies were integrated ahead, into `ahead` (the step's pose read them
// there): the impulses are applied to the integrated state, the state after the step is stored in
// `b`, and the next step's integration of it in `ahead` (the integration kernel's arithmetic).  zero:
// ranges zeroed by extra blocks, for the next step.
void launch_reduce_apply_ahead(const float4 *pForce, const float4 *relPos, const BodyBufs &b, const BodyBufs &ahead,
                               int numBodies, const DevPhys &ph, const ZeroJobs *zero, hipStream_t st);
void launch_fill_u32(unsigned *p, unsigned value, size_t count, hipStream_t st);
void launch_aabb(const float4 *pos, int n, float4 *outMin, float4 *outMax, float4 *partials, int maxPartials,
                 hipStream_t st);
int  aabb_max_partials();
void launch_check_finite(const float4 *a, int n, int *flag, hipStream_t st);
// dst[i] = src[i] + delta for src[i] >= 0, else src[i] (rebases body indices when scenes move).
void launch_copy_offset(int *dst, const int *src, int n, int delta, hipStream_t st);
void launch_fill_u16(uint16_t *p, uint16_t value, size_t count, hipStream_t st);
// dst[r.x + k] = r.y + k for every range r = (first destination, first value, count).
void launch_iota_ranges(const int4 *ranges, int numRanges, int *dst, hipStream_t st);

// Scene carry-over between two engines' buffers (a rebuild for a new scene set): destination
// scenes [dstP, dstP + n) / [dstB, dstB + b) come from [srcP, ...) / [srcB, ...), ranges in
// ascending dstP / dstB order and tiling [0, carriedN) / [0, carriedB).
struct CarryRange {
    int dstP, srcP, n, dstB, srcB, b, scene;
};
struct StateBufs {
    float4 *pos, *vel, *bodyFrame;
    int *rb;
    uint16_t *scene;
    float4 *rbPos, *rbVel, *rbAng, *qinc, *qcum;
    float *Ib, *Iw, *mass, *model;
    int *pstart, *pcount;
};
void launch_carry(const CarryRange *tab, int numRanges, int carriedN, int carriedB, const StateBufs &src,
                  const StateBufs &dst, hipStream_t st);

// ---------------------------------------------------------------- grid path (grid.hip)
// sortedPos.w carries the particle's rbIndex bits; with sceneInVelW sortedVel.w carries its scene id.
void launch_reorder(const unsigned *sortedKeys, const unsigned *sortedIdx, const float4 *pos, const float4 *vel,
                    const int *rb, const uint16_t *scene, float4 *sortedPos, float4 *sortedVel,
                    unsigned *cellStart, unsigned *cellEnd, int n, hipStream_t st);
struct DemOut {
    float4 *pForce;               // body particles: Σ F (original index)
    float4 *pTorque;              // optional: relativePos × Σ F
    float4 *vel;                  // grains: velocity after the contact impulse
    const float4 *relPos;
    int *pairs;                   // pair-emission mode when non-NULL
    int *counts;
    int maxPer;
};
void launch_dem_grid(const float4 *sortedPos, const float4 *sortedVel, const unsigned *sortedIdx,
                     const unsigned *cellStart, const unsigned *cellEnd, int n, const DevPhys &ph,
                     const SceneView &sv, bool sceneInVelW, const DemOut &o, int block, hipStream_t st);

// Engine form of the grid broadphase: a dense CSR offset array over segments of 2^segBits cells
// (segStart has numSegs + 1 entries) replaces cellStart / cellEnd; the DEM scans the sorted keys.
// rec: the key kernel's gather record (pos / vel / rb / scene are then not read), or NULL
void launch_reorder_seg(const unsigned *sortedKeys, const unsigned *sortedIdx, const float4 *pos, const float4 *vel,
                        const int *rb, const uint16_t *scene, const float4 *rec, float4 *sortedPos,
                        float4 *sortedVel, unsigned *segStart, int segBits, long long numSegs, int n,
                        const LongFill &lf, hipStream_t st);
// The bucket sort's plan (bucket.hip).  Scene s holds particles [pStart, pStart + n) and cells
// [cellBase, cellBase + cells); the scenes' cell ranges are packed in scene order, so the keys of
// scene s sort below those of scene s + 1 and the scene's sorted range is its particle range.  A
// scene's cells are cut into `buckets` (at most 1,024; 256 with tiles of 1024) ranges of 2^shift cells
// (shift <= 16), the scene's bucket j being global bucket bucketBase + j; its particles into `tiles`
// tiles of tileKeys particles, global tiles tileBase .. tileBase + tiles - 1.  One scene of 2^20
// particles over 2^24 cells is cut into 512 buckets of its top 9 key bits.
struct BucketScene {
    unsigned pStart, n, cellBase, cells;
    int shift, buckets, bucketBase, tileBase, tiles, infoBase;
};
// Grid keys ranked per tile of particles inside the pose kernel (the first pass of the bucket sort,
// body.hip): the tile's keys and particle indices in bucket order (stable), info[infoBase + j * tiles
// + k] = count << 16 | first slot of bucket j in the scene's tile k, and every bucket's count
// accumulated into hist[bucketBase + j] (zeroed beforehand).  Tiles hold tileKeys particles: 4096
// (kTileKeys), or 1024 where tiles of 4096 would be too few to fill the GPU or scenes are small.
constexpr int kTileKeys = 4096;
struct TileKeyOut {
    unsigned *keys, *vals, *info, *hist;
    const BucketScene *scenes;
    const int *tileScene, *bucketScene;   // scene of each tile and of each bucket
    int numTiles, numBuckets, tileKeys;
    int bucketBits;                       // a scene has at most 2^bucketBits buckets (8, 9 or 10)
    // Morton keys (one scene, 1,024 buckets of 2^20 codes): 1 = 30-bit codes; 2 = 63-bit codes, whose
    // top 30 bits (the 30-bit code) are the tile keys while the full codes go to code63 in particle order
    int morton = 0;
    uint64_t *code63 = nullptr;
};
void launch_pose_key_tiles(const PartBufs &p, const BodyBufs &b, int numParticles, const DevPhys &ph,
                           const SceneView &sv, const TileKeyOut &to, bool grainVelDead, hipStream_t st);
// The host side of the plan: tiles of `tileKeys` particles, and per scene the widest shift whose
// buckets fit in LDS.  A scene is described by its particle count, its grid (cells per axis), and
// either the fullest bucket measured on its particles for every shift (peak[b], b = 0..16), or the
// extent of its container in cells per axis (0 when unknown): where the container is not a whole
// number of grid periods long the hashed cells are unevenly filled, and the plan then expects the
// fullest bucket to hold ceil(r) / r times the average along each axis a bucket does not span,
// r = extent / grid size.  A scene that kept its shift from an earlier plan (a rebuild) passes it.
// False when a scene needs more buckets than its tiles can rank or more than 384 tiles.
struct BucketSceneShape {
    long long particles = 0;
    unsigned grid[3] = {1, 1, 1};
    double extent[3] = {0.0, 0.0, 0.0};
    std::vector<unsigned> peak;   // empty, or one entry per shift 0 .. maxShift
    int shift = -1;
};
struct BucketPlan {
    std::vector<BucketScene> scenes;
    std::vector<int> tileScene, bucketScene;
    size_t infoWords = 0;
    int tileKeys = kTileKeys;
    int bucketBits = 8;   // the most buckets of one scene, rounded up to a power of two (at least 2^8)
};
// maxShift / capKeys: the widest bucket (2^maxShift cells) and the most keys a bucket sorted in LDS holds
// (grid keys: 2^16 and 7,776; Morton keys: 2^20 and 4,096, bucket positions taking 12 bits of a word).
bool plan_bucket_sort(const std::vector<BucketSceneShape> &shapes, BucketPlan *out, int maxShift = 16,
                      int capKeys = 7776);
// The grid sort finished per bucket (bucket.hip): each bucket gathers its keys from its scene's
// tiles' runs (tile order keeps the sort stable), `split` workgroups (1 or 2) sort it by the low bits,
// each writing its share of the sorted keys / indices to (kOut, vOut), of the gathered sorted copies
// and of the bucket's range of the exact cell table (whose last entry, numCells, the last bucket writes).
void launch_bucket_sort(const TileKeyOut &to, unsigned *kOut, unsigned *vOut, const float4 *rec, float4 *sortedPos,
                        float4 *sortedVel, unsigned *table, unsigned numCells, int split, hipStream_t st);
// The same for 30-bit Morton keys ranked by their top 10 bits (TileKeyOut.morton): each bucket of up to
// 4,096 keys sorted by its low 20 bits in LDS, written as sorted keys and indices only; a larger bucket is
// sorted in global memory with scratchA / scratchB (16 B per key, N entries) as its second buffer.
// The 63-bit order from the stable sort by the codes' top 30 bits: code63 is in particle order, k30 /
// vIn the sorted top 30 bits and particle indices.  Writes the sorted 63-bit codes and indices; within
// a run of equal top bits the members are ranked by their full codes, equal codes keeping the run's
// (particle) order, so the result is the stable sort by the 63-bit code.
void launch_morton63_ties(const unsigned *k30, const unsigned *vIn, const uint64_t *code63, int n, uint64_t *kOut,
                          unsigned *vOut, hipStream_t st);
void launch_bucket_sort_morton(const TileKeyOut &to, unsigned *kOut, unsigned *vOut, float4 *scratchA,
                               float4 *scratchB, hipStream_t st);
// Capacity of the long-run list of a table of tableSize entries (runs are disjoint and longer than
// kLongRange), and the launch that fills the listed runs.
unsigned fill_long_capacity(unsigned long long tableSize);
void launch_fill_long(unsigned *table, unsigned long long tableSize, const LongFill &lf, hipStream_t st);
void launch_dem_grid_seg(const float4 *sortedPos, const float4 *sortedVel, const unsigned *sortedKeys,
                         const unsigned *sortedIdx, const unsigned *segStart, int segBits, int n, const DevPhys &ph,
                         const SceneView &sv, bool sceneInVelW, const DemOut &o, int block, hipStream_t st);
// Grid DEM fused with the per-body reduction and apply (exact cell table, no static collider set):
// bodies take one wave each, grains 64 per wave.  rec is the pose kernel's gather record in the
// original order (pos.xyz + rbIndex bits, vel.xyz + scene id); grainIdx lists every grain's index.
struct FusedDemIn {
    const float4 *rec;
    const float4 *relPos;
    const int *grainIdx;
    int numGrains;
    float4 *vel;                  // grains: velocity after the contact impulse
};
void launch_dem_fused(const float4 *sortedPos, const float4 *sortedVel, const unsigned *sortedIdx,
                      const unsigned *segStart, const FusedDemIn &in, const BodyBufs &b, int numBodies,
                      const DevPhys &ph, const SceneView &sv, bool sceneInVelW, hipStream_t st);

// ---------------------------------------------------------------- LBVH path (lbvh.hip)
struct Lbvh {
    int *left, *right, *parent, *minR, *maxR;
    unsigned char *isLeaf;        // optional
    float4 *bounds;               // 2 float4 per node: (min, 0), (max, 0); NULL: not kept (range refit)
    float4 *cms;                  // leaf centres, w = rbIndex bits
    float4 *sortedVel;            // leaf velocities, w = scene id bits when batched
    int *counter;                 // N - 1 refit counters
    int *sceneRoot;               // per scene (batched) or NULL
    // optional traversal layout: internal node k = node - N as 4 float4
    // {Lmin.xyz, L}, {Lmax.xyz, R}, {Rmin.xyz, 0}, {Rmax.xyz, 0} (child ids as int bits)
    float4 *packed;
    // optional (engine): internal node k as (left, right, minR, maxR), and the buffers of the range
    // refit: per leaf the AABB of the leaves from its refit workgroup's start (pre) and to its end
    // (suf), 2 float4 each, written only at the leaves where a range that crosses workgroups ends or
    // starts, and a sparse table of workgroup AABBs (levels x numBlocks x 2 float4)
    int4 *meta;
    float4 *pre, *suf, *blockTab;
    int blockTabLevels;
    // optional (engine, with meta): the build lists the internal nodes whose range crosses a refit
    // block: build workgroup g (1024 nodes) writes its count to crossCount[g] (ceil(n / 1024) entries)
    // and its nodes to crossList[1024 g ...] (n entries in all)
    int *crossList;
    unsigned *crossCount;
};
// Workgroup (leaf block) size of the range refit and the sizes of its buffers for n leaves.
int lbvh_refit_block();
int lbvh_refit_levels(int n);
// keyBytes 4 or 8; sceneShift > 0: the scene id is key >> sceneShift and equal-key ties use the
// scene-local leaf index (sceneStart[scene] is subtracted).
void launch_lbvh_build(const void *sortedKeys, int keyBytes, int n, int sceneShift, const int *sceneStart,
                       const int *sceneCount, int numScenes, const Lbvh &t, hipStream_t st);
// With the range-refit buffers (meta, pre, suf, blockTab, packed) t.bounds may be NULL: the step then
// keeps node boxes only in the traversal records, and launch_lbvh_bounds restates them on request.
void launch_lbvh_refit(const int *sortedIdx, const float4 *pos, const float4 *vel, const int *rb,
                       const uint16_t *scene, const float4 *rec, float radius, int n, const Lbvh &t, hipStream_t st);
// Fills t.bounds (2 float4 per node) from the leaf centres (box = centre -/+ radius) and the packed
// traversal records of a range refit that ran without bounds.
void launch_lbvh_bounds(const Lbvh &t, int n, float radius, hipStream_t st);
// overflow: device int set to 1 when a traversal stack would exceed its LDS depth.
void launch_dem_lbvh(const int *sortedIdx, const Lbvh &t, int n, const DevPhys &ph, const SceneView &sv,
                     bool sceneInVelW, const DemOut &o, int block, hipStream_t st, int *overflow);

// ---------------------------------------------------------------- static collider set (static.hip)
struct RangeCamera {
    float m[16];                  // camera transformation, column-major
    float g[3];                   // normalised accelerometer direction
};
// RGBDloader + computeNormals: pos / norm hold a float4 per pixel; radii and vao may be NULL.
void launch_range_image(const uint16_t *img, int W, int H, const RangeCamera &cam, float radius, float4 *pos,
                        float4 *norm, float *radii, float *vao, hipStream_t st);
void launch_static_hash(const float4 *pos, int m, const SceneGeom &g, unsigned *keys, hipStream_t st);
void launch_static_reorder(const unsigned *keys, const unsigned *idx, const float4 *pos, const float4 *norm, int m,
                           unsigned numCells, float4 *sPos, float4 *sNorm, unsigned *table, const LongFill &lf,
                           hipStream_t st);
// refLayout: cs / ce are the reference's static cellStart / cellEnd, normals are read through stIdx and
// every particle gets pForce += F, pTorque += relPos x F; otherwise cs is the exact static cell table
// (ce unused), normals are sorted and only body particles get pForce += F.
void launch_dem_static(const float4 *sPos, const float4 *sVel, const unsigned *sIdx, int n, const float4 *stPos,
                       const float4 *stNorm, const unsigned *stIdx, const unsigned *cs, const unsigned *ce,
                       const DevPhys &ph, const SceneGeom &g, float4 *pForce, float4 *pTorque,
                       const float4 *relPos, bool refLayout, hipStream_t st);


// ---------------------------------------------------------------- sequential impulses (sis.hip)
constexpr int kSisIterations = 18;   // SequentialImpulseSolver sweeps (Particles/updateFunctions.cpp:1598)
// Inputs of the contact gather, per body particle in the original order.
struct SisGatherIn {
    const float *distA;            // static contacts: penetration (NULL: no static set)
    const int *partA;              //   the static point (index into staticNorm)
    const float4 *staticNorm;
    const float *distB;            // body contacts: penetration, partner body and partner particle
    const int *bodyB, *partB;
    const int *rb;
    const uint16_t *scene;         // per-particle scene id (NULL: one scene)
    const float4 *rbPos, *rbVel, *rbAng, *relPos;
    float radius, arRest, rbRest;
};
// The contact list (ContactNormal / ContactPoint / ContactRigidBody_1/2 / ContactBias /
// ContactAccumulated* of Particles/particleSystem.h:606-613); impulse = (normal, friction,
// friction 2, 0) accumulators.
struct SisContacts {
    float4 *normal, *point;
    int2 *bodies;                  // (b1, b2), b2 = -1 for a static contact
    float *bias;
    float4 *impulse;
};
struct SisBodies {
    const float4 *pos;
    float4 *vel, *ang;
    const float *Iw;               // inverse world-frame inertia, 9 per body, column-major
    const float *mass;
};
struct SisSolvePhys {
    float arFric, gy;
};
// Scratch of the gather and the level-scheduled sweeps, carved from one device buffer.
struct SisWork {
    unsigned *keysA, *keysB, *valsA, *valsB;
    void *sortWs;
    size_t sortWsBytes;
    int *pred, *lvl, *order, *off, *cur, *maxLvl;
    int2 *sceneRange;
    float4 *prep;                  // the sweeps' records in level order (planes of cap float4)
    unsigned *blockCnt;
    int *numContacts;
    size_t end;
};
// Deepest contact per sorted particle against other bodies' particles (engine layout: exact cell
// table, body index in sortedPos.w, scene id in sortedVel.w when sceneInVelW).
void launch_sis_find(const float4 *sPos, const float4 *sVel, const unsigned *sIdx, const unsigned *segStart,
                     const SceneView &sv, bool sceneInVelW, int n, float radius, float *dist, int *pBody, int *pPart,
                     hipStream_t st);
// The same in the reference layout (cellStart / cellEnd, rbIndices), writing only found contacts.
void launch_sis_find_ref(const float4 *sPos, const unsigned *sIdx, const int *rbIdx, const unsigned *cellStart,
                         const unsigned *cellEnd, const SceneGeom &g, int n, float radius, float *dist, int *pBody,
                         int *pPart, float4 *color, hipStream_t st);
// Deepest contact against the static points (refLayout: static cellStart / cellEnd and original
// indices; otherwise the exact static table and sorted indices).
void launch_sis_find_static(const float4 *sPos, const unsigned *sIdx, int n, const float4 *stPos,
                            const unsigned *stIdx, const unsigned *cellStart, const unsigned *cellEnd,
                            const SceneGeom &g, float radius, bool refLayout, float *dist, int *partner,
                            float4 *color, hipStream_t st);
size_t sis_workspace_bytes(int cap, int numScenes, int numBodyParticles);
SisWork sis_work_layout(void *base, int cap, int numScenes, int numBodyParticles);
// Contact list of nb body particles into out (static contacts first); count in w.numContacts.
// perScene: the particles' scene ids give each scene's contact range (no static set then).
void launch_sis_gather(const SisGatherIn &in, int nb, const SisContacts &out, const SisWork &w, bool perScene,
                       hipStream_t st);
// Level schedule and `iterations` sweeps per scene; cap = capacity of the contact list; sceneBodies
// (start, count) per scene, NULL for one scene of numBodies bodies; maxSceneBodies the largest count.
int launch_sis_solve(const SisContacts &ct, const SisBodies &bd, const SisSolvePhys &ph, int cap, int numScenes,
                     int numBodies, const int2 *sceneBodies, int maxSceneBodies, int iterations, const SisWork &w,
                     const SortTuning &tu, hipStream_t st);

}  // namespace arsim
