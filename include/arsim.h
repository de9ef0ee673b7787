/*
 * arsim.h -- C ABI of the MI355X-native (gfx950) per-step DEM pipeline of AR_Sim.
 *
 * The reference exposes this path as free C++-linkage host functions that take raw
 * device pointers, a SimParams POD by value and a block size, and that allocate,
 * synchronise and exit() internally.  This ABI exports the same stages with the same
 * buffer layouts, but:
 *   - every entry point is extern "C", returns an int status (ARSIM_OK == 0) and never
 *     exits the process;
 *   - every launch goes to an explicit stream (a hipStream_t, passed as void*);
 *   - no stage allocates device memory or synchronises the host; scratch memory comes
 *     from a caller-provided workspace whose size is queried first;
 *   - there is no global __constant__ state: parameters travel by pointer.
 *
 * Buffer layouts (reference names, particleSystem.h:211-242, 336-360):
 *   particles  float4 pos[N] (w = 0 for body particles), float4 vel[N],
 *              float4 relativePos[N] (world-frame offset from the body's centre of mass),
 *              int rbIndices[N] (-1 = free grain), float4 pForce[N], float4 pTorque[N].
 *              A body's particles are contiguous and in body order.
 *   bodies     float4 rbPositions[B], rbVelocities[B], rbAngularVelocity[B],
 *              quaternion (x,y,z,w) rbQuaternion[B], column-major mat3 rbInertia[B]
 *              (inverse, body frame) and rbCurrentInertia[B] (inverse, world frame),
 *              float rbMass[B], int particlesPerObject[B].
 *   LBVH       leaves [0,N), internal nodes [N,2N-1), root N; isLeaf u8[2N-1],
 *              left/right/parent/minRange/maxRange int[2N-1], bounds AABB[2N-1]
 *              (float3 min, float3 max, 16-byte aligned = 8 floats), CMs float4[N].
 *
 * Device pointers are plain pointers into HIP device memory.  Floats are FP32.
 */
#ifndef ARSIM_H
#define ARSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status codes */
#define ARSIM_OK                 0
#define ARSIM_ERR_INVALID_ARG    1   /* null pointer, negative size, bad enum */
#define ARSIM_ERR_NOT_POW2       2   /* grid size not a power of two (calcGridHash wraps with &) */
#define ARSIM_ERR_HIP            3   /* a HIP runtime call failed; see arsim_last_error() */
#define ARSIM_ERR_OUT_OF_MEMORY  4
#define ARSIM_ERR_WORKSPACE      5   /* caller workspace smaller than *_workspace_bytes() */
#define ARSIM_ERR_NAN            6   /* NaN guard tripped (state or force) */
#define ARSIM_ERR_OVERFLOW       7   /* fixed-capacity output (contact list, key bits) overflowed */
#define ARSIM_ERR_BUDGET         8   /* serving admission refused: HBM budget exhausted */
#define ARSIM_ERR_NO_DEVICE      9
#define ARSIM_ERR_UNSUPPORTED   10

typedef void *arsim_stream_t;        /* a hipStream_t; NULL = the null stream */

/* ------------------------------------------------------------------ SimParams
 * Byte-identical to the reference's SimParams (Kernels/particleSystem.cuh:27-55):
 * float3 / uint3 are 12-byte, 4-byte-aligned; sizeof == 128. */
typedef struct { float x, y, z; } arsim_float3;
typedef struct { unsigned int x, y, z; } arsim_uint3;

typedef struct ArSimParams {
    arsim_float3 colliderPos;
    float        colliderRadius;
    arsim_float3 gravity;
    float        globalDamping;
    float        particleRadius;
    arsim_uint3  gridSize;
    unsigned int numCells;
    arsim_float3 worldOrigin;
    arsim_float3 cellSize;
    unsigned int numBodies;
    unsigned int maxParticlesPerCell;
    float spring;
    float damping;
    float shear;
    float attraction;
    float boundaryDamping;
    float ARrestitution;
    float RBrestitution;
    float ARfriction;
    float RBfriction;
    float Wdamping;
    float Vdamping;
} ArSimParams;

/* ------------------------------------------------------------------ step configuration
 * What the reference keeps in ParticleSystem members or hard-codes, made explicit. */
#define ARSIM_METHOD_GRID    0       /* hashed uniform grid + 27-cell search */
#define ARSIM_METHOD_LBVH30  1       /* 30-bit Morton LBVH */
#define ARSIM_METHOD_LBVH63  2       /* 63-bit Morton LBVH (21 bits per axis) */

/* wall bits: one static plane per bit, inward normal in brackets */
#define ARSIM_WALL_XMIN 0x01u        /* x = boxMin.x (+x) */
#define ARSIM_WALL_XMAX 0x02u        /* x = boxMax.x (-x) */
#define ARSIM_WALL_YMIN 0x04u        /* y = boxMin.y (+y), the floor */
#define ARSIM_WALL_YMAX 0x08u        /* y = boxMax.y (-y) */
#define ARSIM_WALL_ZMIN 0x10u
#define ARSIM_WALL_ZMAX 0x20u
#define ARSIM_WALLS_OPEN_TOP 0x37u   /* the five walls of an open-top container */

/* flags: the default (0) is the documented resolution of each reference quirk;
 * the COMPAT bits reproduce the reference behaviour for comparison only. */
#define ARSIM_COMPAT_REF_UNITS    0x01u /* Δv = F (no dt, no mass) and body gravity per step, UniformGridDEM.cu:446-459, rigidBodyKernels.cu:334 */
#define ARSIM_COMPAT_DROP_FIRST   0x02u /* per-body sum omits the body's first particle (bodies > 0), UniformGridDEM.cu:435-442 */
#define ARSIM_COMPAT_GRAIN_CLAMP  0x04u /* grains: position clamp + boundaryDamping reflect instead of wall DEM, particleKernels.cu:76-122 */
#define ARSIM_DYNAMIC_AABB        0x08u /* Morton box = per-step particle AABB instead of the container box */
#define ARSIM_CONTACT_SIS         0x10u /* contact model of updateUniformGridSIS (RigidBodies/UniformGridUpdate.cpp:734-816):
                                          deepest contact per particle + sequential impulses, bodies only, grid method */

typedef struct ArSimStepConfig {
    float        dt;
    float        boxMin[3];          /* container / fixed Morton normalisation box */
    float        boxMax[3];
    unsigned int wallMask;           /* ARSIM_WALL_* bits */
    int          method;             /* ARSIM_METHOD_* */
    unsigned int flags;              /* ARSIM_COMPAT_* | ARSIM_DYNAMIC_AABB | ARSIM_CONTACT_SIS */
} ArSimStepConfig;

/* ------------------------------------------------------------------ device-memory helpers
 * Thin wrappers over the HIP runtime so that hosts without HIP headers (ctypes, cgo)
 * can manage the buffers the stage functions take. */
int  arsim_device_count(int *count);
int  arsim_set_device(int device);
int  arsim_malloc(void **dptr, size_t bytes);
int  arsim_free(void *dptr);
int  arsim_memcpy_htod(void *dst, const void *src, size_t bytes, arsim_stream_t stream);
int  arsim_memcpy_dtoh(void *dst, const void *src, size_t bytes, arsim_stream_t stream);
int  arsim_memcpy_dtod(void *dst, const void *src, size_t bytes, arsim_stream_t stream);
int  arsim_memset(void *dptr, int value, size_t bytes, arsim_stream_t stream);
int  arsim_stream_create(arsim_stream_t *stream);
int  arsim_stream_destroy(arsim_stream_t stream);
int  arsim_stream_synchronize(arsim_stream_t stream);
const char *arsim_last_error(void);
int  arsim_abi_version(void);

/* ------------------------------------------------------------------ stage functions (F2-F17)
 * Each replaces one reference wrapper; the citation names the function it replaces. */

/* F2  GPUintegratorWrapper (RigidBodies/rigidBodyKernelWrappers.h:395-407, Kernels/rigidBodyKernels.cu:262-348).
 * x += v dt; incremental q from ω dt; cumulative q = normalize(q_inc * q_cum);
 * rbCurrentInertia = R_cum I_b^-1 R_cum^T; v += g dt (g per step under REF_UNITS); ω *= Wdamping.
 * rbQuaternion receives the incremental rotation, cumulativeQuaternion the orientation,
 * modelMatrix (nullable, column-major mat4) the render pose. */
int arsim_integrate_bodies(float *rbPositions, float *rbVelocities, float *rbAngularVelocity,
                           float *rbQuaternion, float *cumulativeQuaternion,
                           const float *rbInertia, float *rbCurrentInertia, float *modelMatrix,
                           int numBodies, const ArSimParams *params, const ArSimStepConfig *cfg,
                           arsim_stream_t stream);

/* F3  computeGlobalAttributesWrapper (rigidBodyKernelWrappers.h:38-47, rigidBodyKernels.cu:386-421).
 * For body particles: relativePos = R(cumulativeQuaternion) * bodyFramePos; pos = CM + relativePos;
 * vel = V + ω × relativePos.  Grains (rbIndices < 0) are left untouched. */
int arsim_particle_pose(const float *rbPositions, const float *rbVelocities,
                        const float *rbAngularVelocity, const float *cumulativeQuaternion,
                        const float *bodyFramePos, const int *rbIndices,
                        float *relativePos, float *pos, float *vel, int numParticles,
                        arsim_stream_t stream);

/* F4  integrateSystem (Particles/ParticleAuxiliaryFunctions.h:14-20, particleKernels.cu:50-148).
 * Grains only: v += g dt; v *= globalDamping; x += v dt; with ARSIM_COMPAT_GRAIN_CLAMP also the
 * reference's box clamp (±0.01 margin, reflect with boundaryDamping, top wall only if v.y > 0). */
int arsim_integrate_grains(float *pos, float *vel, const int *rbIndices, int numParticles,
                           const ArSimParams *params, const ArSimStepConfig *cfg,
                           arsim_stream_t stream);

/* F5  calcHash (ParticleAuxiliaryFunctions.h:21-24, particleKernels.cu:151-189).
 * hash = wrap(floor((p - origin) / cellSize)) flattened z-y-x; index = i. */
int arsim_calc_hash(unsigned int *gridParticleHash, unsigned int *gridParticleIndex,
                    const float *pos, int numParticles, const ArSimParams *params,
                    arsim_stream_t stream);

/* F6 / F13  sortParticlesPreallocated (ParticleAuxiliaryFunctions.h:52-57,
 * particleKernelWrappers.cu:419-534) and the sort inside createMortonCodesPreallocated
 * (BVHSource/createMortonCodes.cu:253-347).  Stable ascending LSD radix sort of (key, value)
 * pairs over key bits [beginBit, endBit).  Like the reference's cub::DoubleBuffer use it
 * exchanges the caller's pointers: on return *keys / *values point at the sorted data and
 * *altKeys / *altValues at the scratch buffers. */
size_t arsim_sort_workspace_bytes(int numItems, int keyBytes);
int arsim_sort_pairs_u32(unsigned int **keys, unsigned int **values,
                         unsigned int **altKeys, unsigned int **altValues,
                         int numItems, int beginBit, int endBit,
                         void *workspace, size_t workspaceBytes, arsim_stream_t stream);
int arsim_sort_pairs_u64(uint64_t **keys, unsigned int **values,
                         uint64_t **altKeys, unsigned int **altValues,
                         int numItems, int beginBit, int endBit,
                         void *workspace, size_t workspaceBytes, arsim_stream_t stream);

/* No counterpart in the reference.  arsim_sort_pairs_u32 without a workspace, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, results and pointer exchange as arsim_sort_pairs_u32; the
 * workspace is a static device array of the library (8.0 MiB per device, held until the process
 * ends), so the call allocates nothing.  numItems above ARSIM_SORT_NOALLOC_MAX_ITEMS is refused
 * with ARSIM_ERR_WORKSPACE.  All calls on one device share that array: calls that overlap on one
 * device (different streams or threads) race and must be serialised by the caller. */
#define ARSIM_SORT_NOALLOC_MAX_ITEMS (1 << 22)
int arsim_sort_pairs_u32_noalloc(unsigned int **keys, unsigned int **values,
                                 unsigned int **altKeys, unsigned int **altValues,
                                 int numItems, int beginBit, int endBit, arsim_stream_t stream);

/* F7  reorderDataAndFindCellStart (ParticleAuxiliaryFunctions.h:25-35, particleKernels.cu:193-258,
 * particleKernelWrappers.cu:213-254).  cellStart is set to 0xffffffff for all numCells, then
 * cellStart[h] / cellEnd[h] bracket the sorted run of hash h; sortedPos/Vel[i] = pos/vel[index[i]],
 * except that sortedPos[i].w carries the bit pattern of rbIndices[index[i]] (the reference never
 * reads w of its sorted copies), which lets the contact kernel skip the rbIndices[index[j]] gather. */
int arsim_reorder_find_cell_start(const int *rbIndices, unsigned int *cellStart, unsigned int *cellEnd,
                                  float *sortedPos, float *sortedVel,
                                  const unsigned int *gridParticleHash,
                                  const unsigned int *gridParticleIndex,
                                  const float *pos, const float *vel,
                                  int numParticles, unsigned int numCells,
                                  arsim_stream_t stream);

/* F8 / F9 / F18  FindAndHandleRigidBodyCollisionsUniformGridWrapper (rigidBodyKernelWrappers.h:361-376,
 * UniformGridDEM.cu:57-241) fused with the grain branch of collideD (particleKernels.cu:302-406)
 * and the static-plane wall term.  For each sorted particle: Σ spring-dashpot-shear force over
 * the 27 neighbour cells (self and same-body pairs excluded, grain-grain included), plus walls.
 * Body particles: pForce[orig] = F, pTorque[orig] = relativePos[orig] × F.
 * Grains: vel[orig] = sortedVel + F * dt (F without dt under REF_UNITS). */
int arsim_dem_grid(const int *rbIndices, float *pForce, float *pTorque, float *vel,
                   const float *sortedPos, const float *sortedVel, const float *relativePos,
                   const unsigned int *gridParticleIndex, const unsigned int *cellStart,
                   const unsigned int *cellEnd, int numParticles,
                   const ArSimParams *params, const ArSimStepConfig *cfg, arsim_stream_t stream);

/* F10  ReducePerParticleImpulses (rigidBodyKernelWrappers.h:348-359, UniformGridDEM.cu:421-666).
 * Per body, in a fixed order: J = Σ pForce, τ = Σ pTorque over its contiguous particles
 * (particleStart[b], particlesPerObject[b]); v += J dt / m; ω += I_w^-1 τ dt where
 * I_w^-1 = rbCurrentInertia.  No scan over N, no host prefix, no allocation.  pTorque may be NULL:
 * the per-particle torque is then relativePos × pForce, the expression the contact kernel stores. */
int arsim_segmented_reduce_apply(const float *pForce, const float *pTorque, const float *relativePos,
                                 float *rbVelocities, float *rbAngularVelocity,
                                 const float *rbMass, const float *rbCurrentInertia,
                                 const int *particleStart, const int *particlesPerObject,
                                 int numBodies, const ArSimParams *params,
                                 const ArSimStepConfig *cfg, arsim_stream_t stream);

/* F11  findAABBCub (BVHSource/createMortonCodes.cu:37-107).  Exact component-wise min / max
 * over float4[N] (w included); results stay on the device (no D2H). */
size_t arsim_aabb_workspace_bytes(int numItems);
int arsim_aabb(const float *pos, int numItems, float *outMin4, float *outMax4,
               void *workspace, size_t workspaceBytes, arsim_stream_t stream);

/* F12  generateMortonCodes (BVHSource/BVHkernels.cu:646-716).  u = (p - min) / (max - min),
 * quantised to 10 bits (30-bit code) or 21 bits (63-bit code) per axis and interleaved x,y,z.
 * box4 == NULL uses boxMin/boxMax; otherwise two device float4 (min, max), e.g. from arsim_aabb. */
int arsim_morton30(const float *pos, unsigned int *codes, int *indices, int numItems,
                   const float boxMin[3], const float boxMax[3], const float *box4,
                   arsim_stream_t stream);
int arsim_morton63(const float *pos, uint64_t *codes, int *indices, int numItems,
                   const float boxMin[3], const float boxMax[3], const float *box4,
                   arsim_stream_t stream);

/* F14  wrapperConstructRadixTreeSoA (BVHSource/BVHcreation.h:111-120, BVHSoAKernels.cu:7-99).
 * Karras radix tree over sorted codes (keyBytes 4 or 8).  Equal keys are ordered by the
 * leaf index (δ = keyBits + clz(i ^ j)).  Also clears nodeCounter[N-1] for the refit. */
int arsim_lbvh_build(const void *sortedCodes, int keyBytes, int numLeaves,
                     unsigned char *isLeaf, int *leftIndices, int *rightIndices,
                     int *parentIndices, int *minRange, int *maxRange, int *nodeCounter,
                     arsim_stream_t stream);

/* F15 / F16  wrapperConstructLeafNodesSoA + wrapperConstructInternalNodesSoA (BVHcreation.h:122-145,
 * BVHSoAKernels.cu:108-176).  Leaf i: CMs[i] = pos[sortedIndices[i]], bounds = CM ± radius;
 * then each leaf walks up with an acquire-release counter; the second arrival merges.
 * CMs[i].w carries the bit pattern of rbIndices[sortedIndices[i]]; sortedVel[i] = vel[sortedIndices[i]]. */
int arsim_lbvh_refit(const int *sortedIndices, const float *pos, const float *vel, const int *rbIndices,
                     float radius, int numLeaves,
                     unsigned char *isLeaf, int *minRange, int *maxRange,
                     const int *leftIndices, const int *rightIndices, const int *parentIndices,
                     float *CMs, float *sortedVel, float *bounds, int *nodeCounter, arsim_stream_t stream);

/* F17  collideBVHSoARigidBodyWrapper (rigidBodyKernelWrappers.h:49-68, RB_BVHKernels.cu:100-244).
 * Stack traversal from the root per sorted leaf; same contact predicate, force, wall term and
 * outputs as arsim_dem_grid.  Neighbour velocities are read from a sorted copy (sortedVel), so
 * no thread reads a velocity another thread is writing.  The per-thread stack holds 64 entries in
 * LDS, as the reference's int stack[64]; a traversal that would exceed it sets the device int
 * *overflowFlag to 1 and writes nothing for that particle (its pForce / pTorque / vel keep their
 * previous contents), so a truncated sum is never returned as a result. */
int arsim_dem_lbvh(const int *rbIndices, float *pForce, float *pTorque, float *vel,
                   const float *sortedVel, const float *relativePos, const int *sortedIndices,
                   const int *leftIndices, const int *rightIndices, const float *CMs,
                   const float *bounds, int numLeaves,
                   const ArSimParams *params, const ArSimStepConfig *cfg, int *overflowFlag,
                   arsim_stream_t stream);

/* Contact-pair emission (test mode of F9 / F17, after FindRigidBodyCollisionsBVHKernel,
 * RB_BVHKernels.cu:760-876).  For each particle (original index i) the list of original
 * indices j it touches (same predicate as the force kernels), at most maxPerParticle each:
 * pairs[i * maxPerParticle + k], counts[i].  Returns ARSIM_ERR_OVERFLOW via *overflow != 0. */
int arsim_contact_pairs_grid(const int *rbIndices, const float *sortedPos,
                             const unsigned int *gridParticleIndex,
                             const unsigned int *cellStart, const unsigned int *cellEnd,
                             int numParticles, const ArSimParams *params,
                             int *pairs, int *counts, int maxPerParticle, arsim_stream_t stream);
int arsim_contact_pairs_lbvh(const int *rbIndices, const int *sortedIndices,
                             const int *leftIndices, const int *rightIndices,
                             const float *CMs, const float *bounds, int numLeaves,
                             const ArSimParams *params, int *overflowFlag,
                             int *pairs, int *counts, int maxPerParticle, arsim_stream_t stream);

/* Static (augmented-reality) collider set.
 * loadRangeImage (Kernels/customRangeImageLoader.cu:165-216 = RGBDloader :14-109 + computeNormals
 * :118-162): a W x H depth image (uint16, device) -> one scene point per pixel (float4, w = 0) with its
 * normal (float4), through the reference's hard-coded intrinsics and accelerometer rotation and the
 * host-side camera transformation (16 floats, column-major, host memory).  Pixels past (W-1)(H-1)
 * and the border's normals are not written, as in the reference.  staticRadii (float per pixel) and
 * vaoData (5 floats per pixel: position, x / W, y / H) may be NULL. */
int arsim_range_image_load(const uint16_t *depth, int width, int height, const float *cameraTransformation,
                           float particleRadius, float *staticPos, float *staticNorm, float *staticRadii,
                           float *vaoData, arsim_stream_t stream);
/* FindAndHandleAugmentedRealityCollisionsUniformGridWrapper (RigidBodies/UniformGridDEM.cu:382-417):
 * contact forces of the sorted dynamic particles against the static points (sorted by cell, with
 * their cellStart / cellEnd from arsim_calc_hash + sort + arsim_reorder_find_cell_start on the static
 * set, and their unsorted normals read through staticSortedIndex), ADDED to pForce, with
 * relativePos x F added to pTorque, for every particle. */
int arsim_dem_static_grid(float *pForce, float *pTorque, const float *sortedPos, const float *sortedVel,
                          const float *relativePos, const float *staticSortedPos, const float *staticNorm,
                          const unsigned int *sortedIndex, const unsigned int *staticSortedIndex,
                          const unsigned int *staticCellStart, const unsigned int *staticCellEnd,
                          int numParticles, const ArSimParams *params, const ArSimStepConfig *cfg,
                          arsim_stream_t stream);

/* Sequential-impulse contact model (the step of updateUniformGridSIS, RigidBodies/UniformGridUpdate.cpp:734-816).
 * FindRigidBodyCollisionsUniformGridWrapper (RigidBodies/rigidBodyKernelWrappers.h:220-232,
 * Kernels/RB_UniformGridKernels.cu:730-876): per sorted particle, over the 27 neighbour cells in z, y, x
 * order and each cell's particles in sorted order, the candidate of another body with the largest
 * penetration 2r - |p - q| above contactDistance[orig] (cleared to 0 by the caller, as
 * UniformGridUpdate.cpp:412 does).  Writes contactDistance / collidingParticleIndex (original index) /
 * collidingRigidBodyIndex only where a contact is found, and color[orig] = (1,0,0,0) then (color may
 * be NULL). */
int arsim_find_deepest_contacts_grid(const int *rbIndices, int *collidingRigidBodyIndex, int *collidingParticleIndex,
                                     float *contactDistance, float *color, const float *sortedPos,
                                     const unsigned int *gridParticleIndex, const unsigned int *cellStart,
                                     const unsigned int *cellEnd, int numParticles, const ArSimParams *params,
                                     arsim_stream_t stream);
/* FindAugmentedRealityCollisionsUniformGridWrapper (RB_UniformGridKernels.cu:879-1108): the same search
 * over the static points (sorted by cell, static cellStart / cellEnd); collidingParticleIndex receives
 * the static point's original index (staticSortedIndex[j]). */
int arsim_find_static_contacts_grid(int *collidingParticleIndex, float *contactDistance, float *color,
                                    const float *sortedPos, const float *staticSortedPos,
                                    const unsigned int *gridParticleIndex, const unsigned int *staticSortedIndex,
                                    const unsigned int *staticCellStart, const unsigned int *staticCellEnd,
                                    int numParticles, const ArSimParams *params, arsim_stream_t stream);
/* Device scratch of arsim_sis_gather and arsim_sis_solve for maxContacts contacts. */
size_t arsim_sis_workspace_bytes(int maxContacts, int numBodyParticles);
/* GatherAugmentedRealityCollisions + GatherRigidBodyCollisions (Particles/updateFunctions.cpp:1431-1578)
 * on the device: for every body particle i < numBodyParticles (bodies' particles come first, in body
 * order) with a static contact (staticContactDistance[i] > 0; the three static arrays may be NULL) and
 * then for every one with a body contact (contactDistance[i] > 0 and a partner body >= 0), one contact:
 * contactNormal / contactPoint (float4), contactBodies (b1, b2; b2 = -1 static), contactBias, and
 * contactImpulse (4 floats, cleared: normal, friction, friction 2 accumulators).  Static contacts come
 * first.  The count goes to the device int *numContacts; maxContacts must cover numBodyParticles
 * (twice that with static contacts).  A partner that is a grain has no body state and is skipped. */
int arsim_sis_gather(const float *rbPositions, const float *rbVelocities, const float *rbAngularVelocity,
                     const float *relativePos, const int *rbIndices, int numBodyParticles,
                     const float *staticContactDistance, const int *staticContactIndex, const float *staticNorm,
                     const float *contactDistance, const int *collidingRigidBodyIndex,
                     const int *collidingParticleIndex, float *contactNormal, float *contactPoint,
                     int *contactBodies, float *contactBias, float *contactImpulse, int *numContacts,
                     int maxContacts, const ArSimParams *params, void *workspace, size_t workspaceBytes,
                     arsim_stream_t stream);
/* SequentialImpulseSolver (Particles/updateFunctions.cpp:1580-1874) on the device: `iterations` (the
 * reference uses 18) Gauss-Seidel sweeps over the contacts in list order, updating rbVelocities /
 * rbAngularVelocity (w = 0) and the accumulators, with the same results as the serial loop: contacts
 * are run level by level, a level holding contacts whose bodies no earlier pending contact touches.
 * Call arsim_particle_pose afterwards for the particles (computeGlobalAttributesWrapper, :1862-1871). */
int arsim_sis_solve(float *contactNormal, float *contactPoint, int *contactBodies, float *contactBias,
                    float *contactImpulse, const int *numContacts, int maxContacts, const float *rbPositions,
                    float *rbVelocities, float *rbAngularVelocity, const float *rbCurrentInertia, const float *rbMass,
                    int numBodies, int iterations, const ArSimParams *params, void *workspace,
                    size_t workspaceBytes, arsim_stream_t stream);

/* Body-index rebasing copy: dst[i] = src[i] + delta where src[i] >= 0, else src[i].  The reference
 * offsets a new body's rbIndices by the current body count when it appends it
 * (RigidBodies/RigidBodyAddition.cpp:111-139); the engine does the same when it moves scenes
 * between batch layouts.  With delta = 0 it is a plain 4-byte copy (used to calibrate counters). */
int arsim_rebase_indices(int *dst, const int *src, int n, int delta, arsim_stream_t stream);

/* ------------------------------------------------------------------ engine (scene-level API)
 * ParticleSystem::update(dt) (Particles/particleSystem.h:288) for one or many scenes.
 * One engine holds S >= 1 independent scenes resident in HBM; one step advances all of them
 * with one launch per stage (S == 1 is the plain path).  The host arrays below are copied in. */
typedef struct ArSimSceneHost {
    int numParticles;
    int numBodies;
    const float *pos;                 /* float4[N] */
    const float *vel;                 /* float4[N] */
    const int   *rbIndices;           /* int[N], scene-local body index or -1 */
    const float *bodyFramePos;        /* float4[N], offset from the CM in the body frame */
    const float *rbPositions;         /* float4[B] */
    const float *rbVelocities;        /* float4[B] */
    const float *rbAngularVelocity;   /* float4[B] */
    const float *rbQuaternion;        /* float4[B] cumulative orientation (x,y,z,w) */
    const float *rbInertia;           /* float[9*B] inverse body-frame inertia, column-major */
    const float *rbMass;              /* float[B] */
    const int   *particlesPerObject;  /* int[B] */
    ArSimParams     params;
    ArSimStepConfig cfg;
} ArSimSceneHost;

typedef struct ArSimTuning {
    int demBlock;                     /* threads per block of the DEM kernels (64..1024) */
    int sortItemsPerThread;           /* onesweep items per thread (4..16) */
    int sortRadixBits;                /* onesweep digit width (4..8) */
    int useGraph;                     /* capture the step in a hipGraph and replay it */
    int fuseReduce;                   /* grid method: one kernel for the contact forces and the
                                         per-body reduction (bodies one wave each, grains 64 per
                                         wave, original order); same results, measured slower than
                                         the sorted-order contact kernel plus the reduction */
    int bucketSplit;                  /* grid method, bucket sort: workgroups per bucket (1 or 2;
                                         0 = two while there are fewer than 512 buckets, else one) */
} ArSimTuning;

typedef struct arsim_engine arsim_engine;

int arsim_engine_create(const ArSimSceneHost *scenes, int numScenes, const ArSimTuning *tuning,
                        int device, arsim_engine **out);
int arsim_engine_destroy(arsim_engine *e);
int arsim_engine_step(arsim_engine *e, int numSteps);            /* enqueue only */
int arsim_engine_synchronize(arsim_engine *e);
/* ParticleSystem::update(dt) takes dt per call (updateFunctions.cpp:2803); set it before stepping. */
int arsim_engine_set_dt(arsim_engine *e, float dt);
int arsim_engine_stream(arsim_engine *e, arsim_stream_t *stream);
int arsim_engine_num_particles(arsim_engine *e, int sceneIdx);   /* sceneIdx < 0: total */
size_t arsim_engine_device_bytes(arsim_engine *e);
/* Copies one scene's state back (any pointer may be NULL). */
int arsim_engine_download(arsim_engine *e, int sceneIdx,
                          float *pos, float *vel, float *relativePos, float *pForce, float *pTorque,
                          float *rbPositions, float *rbVelocities, float *rbAngularVelocity,
                          float *rbQuaternion, float *rbCurrentInertia);
/* Per-stage GPU time (ms) of the last profiled step: measured with HIP events on the engine
 * stream.  names[k] is a static string.  Returns the number of stages. */
int arsim_engine_profile_step(arsim_engine *e, float *msPerStage, const char **names, int maxStages);
/* Debug views of the last step's broadphase (device pointers, valid until the next step).  The LBVH
 * node boxes are not kept by the step (the traversal reads the packed child boxes): asking for
 * lbvhBounds restates them from the leaf centres and the packed records and waits for that. */
int arsim_engine_debug_buffers(arsim_engine *e, void **sortedKeys, void **sortedIndex,
                               void **cellStart, void **cellEnd, void **sortedPos, void **sortedVel,
                               void **lbvhLeft, void **lbvhRight, void **lbvhBounds, void **lbvhCMs);
/* Contact pairs of the current state (original particle indices), computed with the engine's
 * broadphase: runs the broadphase stages of a step without integrating. */
int arsim_engine_contact_pairs(arsim_engine *e, int *pairsHost, int *countsHost, int maxPerParticle);
int arsim_engine_check_finite(arsim_engine *e, int *nonFinite);   /* NaN guard (device reduce) */
/* Times candidate tunings on this engine's scenes and keeps the fastest. */
int arsim_engine_autotune(arsim_engine *e, int stepsPerTrial, ArSimTuning *best, float *bestMsPerStep);

/* ------------------------------------------------------------------ seeded scene generator
 * Synthetic scenes of rigid cubes (edge^3 particles, spacing 2r), voxelised ellipsoids and
 * free grains, after addRigidBody / initObject / addNewSphere (RigidBodies/RigidBodyAddition.cpp:
 * 20-228, 1486-1903; Particles/sphereAddition.cpp:151-301).  Body masses are the particle
 * counts, inertia the point-mass sum over the body's particles, inverted. */
typedef struct ArSimGenSpec {
    uint64_t seed;
    int   numCubes;
    int   cubeEdge;                   /* particles per cube edge (4) */
    int   numEllipsoids;
    float ellipsoidMinSemi, ellipsoidMaxSemi;  /* semi-axes, uniform per axis */
    int   numGrains;
    float box[3];                     /* container [0, box] */
    float radius;
    float bodyVelSigma;               /* N(0, sigma) per component */
    float grainVelSigma;
    float dropHeightMax;              /* > 0: bodies dropped from y ~ U[0, dropHeightMax] */
    int   gridLog2[3];                /* grid size per axis = 1 << gridLog2 */
    int   method;
    float dt, spring, damping, shear, gravityY;
} ArSimGenSpec;

typedef struct ArSimSceneOwned ArSimSceneOwned;   /* host-side scene arrays owned by the library */
int arsim_generate_scene(const ArSimGenSpec *spec, ArSimSceneOwned **out);
/* The small-scene generator scaled to n particles: 4x4x4 cubes holding ~75 % of them, the rest
 * free grains, box edge 64 * (n / 16384)^(1/3), r = 0.5, k = 5e4, c = 50, shear = 10, dt = 1e-4.
 * Used for the batched and served mixes. */
int arsim_generate_mix_scene(int n, uint64_t seed, int method, ArSimSceneOwned **out);
/* Dense settled pile (the 4 M-particle workload).  numEllipsoids ellipsoids with semi-axes
 * 2r * 4^(u^2), u ~ U[0,1] per axis (on [2r, 8r]), voxelised at spacing r (33 to 2,145 particles,
 * heavy-tailed), are placed at random positions and orientations, largest first, so that no two
 * bodies' particles are closer than 2r * latticeGap.  The space around them is then filled from the
 * floor up with grains on a face-centred cubic lattice of nearest-neighbour distance 2r * latticeGap
 * (12 neighbours in contact for latticeGap < 1), skipping sites closer than 2r * surfaceGap to a body
 * particle, until the scene holds targetParticles.  Container [0, edge] in x and z with an open top, v = 0, grid
 * 2^gridLog2 cells of 2r per axis, k = 5e4, c = 50, shear 10, dt = 1e-4, g = -9.81. */
typedef struct ArSimPileSpec {
    uint64_t seed;
    int   numEllipsoids;
    int   targetParticles;
    float edge;
    float radius;
    float latticeGap;
    float surfaceGap;
    int   gridLog2;
    int   method;
} ArSimPileSpec;
int arsim_generate_packed_pile(const ArSimPileSpec *spec, ArSimSceneOwned **out);
int arsim_scene_view(const ArSimSceneOwned *s, ArSimSceneHost *view);
int arsim_scene_free(ArSimSceneOwned *s);

/* ------------------------------------------------------------------ scene I/O
 * Particle templates in the reference's OBJparticles format, loaded as ParticleSystem::initObject
 * does (RigidBodies/RigidBodyAddition.cpp:1486-1756): <dataDir>/<model>/<model>_<suffix>.txt, a
 * particle count line then one "x y z" line per particle; points centred on their FP32 mean and
 * moved by the per-model offset; inverse inertia from the identity plus the point-mass sum, or the
 * reference's hard-coded tensors for banana / bunny / teapot / cube scales; radius = the largest
 * coordinate; mass 1. */
typedef struct ArSimTemplate ArSimTemplate;
/* The file suffix initObject picks for a scale ("1_0", "1_5", "2_0" or "2_5"; RigidBodyAddition.cpp:1494-1503). */
const char *arsim_template_suffix(float scale);
int arsim_template_load(const char *dataDir, const char *modelName, float scale, float particleRadius,
                        ArSimTemplate **out);
/* relPos (float4 per particle) and invInertia (9, column-major) may be NULL. */
int arsim_template_info(const ArSimTemplate *t, int *numParticles, float *relPos, float *invInertia,
                        float *radius, float *mass);
int arsim_template_free(ArSimTemplate *t);
/* An empty scene to be filled with template bodies and grains (ParticleSystem's add paths). */
int arsim_scene_create(const ArSimParams *params, const ArSimStepConfig *cfg, ArSimSceneOwned **out);
/* addRigidBody of a template instance (RigidBodyAddition.cpp:1760-1903): identity orientation,
 * particles at pos + rel with the body velocity; inserted before the scene's free grains. */
int arsim_scene_add_body(ArSimSceneOwned *s, const ArSimTemplate *t, const float pos[3], const float vel[3],
                         const float angularVelocity[3]);
/* Free grains (rbIndex -1): pos4 / vel4 hold a float4 per grain. */
int arsim_scene_add_grains(ArSimSceneOwned *s, int n, const float *pos4, const float *vel4);
/* A host scene that rebuilds scene s of an engine in its current state (for eviction / restore and
 * fixtures): particle pos / vel, body position, velocity, angular velocity and orientation, plus the
 * scene's fixed data and current dt.  arsim_engine_create on it continues the run bit for bit. */
int arsim_engine_snapshot(arsim_engine *e, int scene, ArSimSceneOwned **out);
/* Rebuilds engine e for a new scene set: scenes keepScenes[0..numKeep) of engine `from` (moved on
 * the device, state and all, in one launch per array group) followed by numAdded new host scenes.
 * e keeps its allocations where they are large enough, so a serving loop that alternates two engines
 * changes its resident set without new allocations; e != from.  Every kept scene continues bit for
 * bit as if it had never moved.  No counterpart in the reference (one scene per process there). */
int arsim_engine_rebuild(arsim_engine *e, arsim_engine *from, const int *keepScenes, int numKeep,
                         const ArSimSceneHost *added, int numAdded);
/* Headless render state after the last step (the per-body modelMatrixGPU / cumulativeQuaternionGPU
 * outputs of GPUintegratorKernel, Kernels/rigidBodyKernels.cu:322-332): 16 floats (column-major mat4)
 * and 4 floats (x,y,z,w) per body of the scene.  Either pointer may be NULL. */
int arsim_engine_render_state(arsim_engine *e, int scene, float *modelMatrices, float *cumulativeQuaternions);
/* Static collider set of a one-scene engine (host float4 arrays of points and normals, e.g. from
 * arsim_range_image_load): indexed once on the device, then every step adds the static contact
 * forces of the body particles to their particle forces after the DEM (the reference's
 * Find_And_Handle_Augmented_Reality_Collisions_Uniform_Grid_DEM, UniformGridUpdate.cpp:880-935).
 * n = 0 removes the set.  Not available for batched engines. */
int arsim_engine_set_static(arsim_engine *e, const float *pos4, const float *norm4, int n);

/* ------------------------------------------------------------------ serving
 * Continuous batching over a Poisson request stream: requests are admitted into the running
 * batch as slots and the HBM budget allow, their state stays on the device, every engine
 * iteration steps all resident scenes in the batched launches, and finished scenes retire. */
typedef struct ArSimServeConfig {
    uint64_t seed;
    int    numRequests;
    double arrivalRate;               /* requests per second (Poisson) */
    int    minParticles, maxParticles;/* log-uniform */
    int    minSteps, maxSteps;        /* uniform */
    double hbmBudgetBytes;
    int    maxResident;               /* slot limit */
    int    oneAtATime;                /* 1: serve FIFO, one scene at a time, plain path */
    double timeScale;                 /* >1 compresses arrival times (for short runs) */
    int    method;
    /* Optional (NULL = off), numRequests entries: request i's result, the FNV-1a checksum of its
     * final state as arsim_engine_state_checksum computes it (0 for a request that was not served).
     * Costs a device-to-host copy of each finished scene. */
    uint64_t *resultChecksums;
} ArSimServeConfig;

typedef struct ArSimServeReport {
    int    completed;
    double wallSeconds;
    double particleUpdates;           /* Σ N_s * steps_s */
    double updatesPerSecond;
    double stepLatencyP50Ms, stepLatencyP99Ms;       /* per-scene step latency */
    double completionLatencyP50Ms, completionLatencyP99Ms;
    double queueWaitP50Ms, queueWaitP99Ms;
    int    maxResidentSeen;
    int    rejectedForBudget;         /* requests larger than the whole HBM budget: refused, never served */
    double peakDeviceBytes;
    int    deferredForBudget;         /* requests that waited at least once for budget to free up */
} ArSimServeReport;

/* Serves the stream.  Batched mode (oneAtATime = 0) needs maxResident > 0 and hbmBudgetBytes > 0
 * (else ARSIM_ERR_INVALID_ARG).  A request whose memory estimate exceeds the budget on its own is
 * refused and counted in rejectedForBudget; the run still returns ARSIM_OK once every other
 * request has completed, so completed + rejectedForBudget == numRequests. */
int arsim_serve_run(const ArSimServeConfig *cfg, ArSimServeReport *report);
/* Request i of the stream cfg describes (the same deterministic draw arsim_serve_run makes): its
 * particle count, step count, scene seed (arsim_generate_mix_scene(numParticles, sceneSeed, method))
 * and arrival time in seconds (after timeScale).  Any output pointer may be NULL. */
int arsim_serve_request(const ArSimServeConfig *cfg, int i, int *numParticles, int *numSteps, uint64_t *sceneSeed,
                        double *arrivalSeconds);
/* FNV-1a (64-bit) over scene s's state bytes: particle pos and vel (float4 each), then the body
 * positions, velocities, angular velocities and cumulative quaternions.  Synchronises the engine.
 * No counterpart in the reference; it lets a served request's result be compared with the same
 * scene stepped alone. */
int arsim_engine_state_checksum(arsim_engine *e, int scene, uint64_t *out);

#ifdef __cplusplus
}
#endif

#endif /* ARSIM_H */
