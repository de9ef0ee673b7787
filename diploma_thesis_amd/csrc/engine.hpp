// engine.hpp -- the scene-level step engine: S >= 1 independent scenes resident in HBM, one
// launch per stage per step for all of them (S == 1 is the plain path).
#pragma once

#include <string>
#include <vector>

#include "../../include/arsim.h"
#include "kernels.hpp"

namespace arsim {

// Where a scene's initial state comes from when an engine is (re)built.
struct SceneSource {
    const ArSimSceneHost *host = nullptr;   // upload from host arrays, or
    const struct Engine *from = nullptr;    // copy scene `fromScene` of another engine on the device
    int fromScene = -1;
};

// Widest grid key (all scenes' cells) with one cell-table entry per cell.
constexpr int kExactTableBits = 30;

struct Engine {
    int device = 0;
    hipStream_t st = nullptr;
    bool ownStream = true;
    int S = 0, N = 0, B = 0;
    int method = ARSIM_METHOD_GRID;
    unsigned flags = 0;
    ArSimParams p0{};
    ArSimStepConfig c0{};
    DevPhys ph{};
    ArSimTuning tu{128, 16, 8, 1, 0, 0};
    std::vector<SceneGeom> geoms;
    std::vector<int> pStart, pCount, bStart, bCount;
    std::vector<int> bodyParticles;         // per scene: its body particles (they precede its grains)
    int numGrains = 0;
    int cellBits = 0, sceneBits = 0, keyBits = 0, keyBytes = 4;
    size_t numCells = 0;

    // device buffers: each member pointer keeps its allocation (and capacity) across builds, so an
    // engine object can be rebuilt for a new scene set without new allocations when it fits
    std::vector<std::pair<void **, size_t>> allocs;
    size_t devBytes = 0;
    int *sceneRootBuf = nullptr;            // tree.sceneRoot points here when S > 1
    CarryRange *carryTab = nullptr;         // scene ranges of a bulk carry-over
    float4 *pos = nullptr, *vel = nullptr, *relPos = nullptr, *bodyFrame = nullptr, *pForce = nullptr;
    int *rb = nullptr;
    uint16_t *scene = nullptr;
    float4 *gatherRec = nullptr;   // the key kernel's 32-byte gather record per particle
    float4 *rbPos = nullptr, *rbVel = nullptr, *rbAng = nullptr, *qinc = nullptr, *qcum = nullptr;
    float *Ib = nullptr, *Iw = nullptr, *mass = nullptr;
    float *model = nullptr;       // per-body column-major mat4 (render state), written each step
    int *pstartD = nullptr, *pcountD = nullptr;
    int *grainIdx = nullptr;                // grid: every grain's particle index, in order
    int4 *grainRanges = nullptr;            // grid: per scene (first list slot, first grain, grain count)
    SceneGeom *geomD = nullptr;
    int *sceneStartD = nullptr, *sceneCountD = nullptr;
    void *keysA = nullptr, *keysB = nullptr;
    unsigned char *keyBufA = nullptr, *keyBufB = nullptr, *sortWsBuf = nullptr;
    unsigned *valsA = nullptr, *valsB = nullptr;
    unsigned *segStart = nullptr;           // grid: CSR offsets over segments of 2^segBits cells
    unsigned *tileInfo = nullptr;           // grid, bucket sort: per-tile bucket runs (count << 16 | first)
    BucketScene *bkScenesD = nullptr;       // grid, bucket sort: the plan (bucket.hip)
    int *tileSceneD = nullptr, *bucketSceneD = nullptr;
    unsigned *bucketHist = nullptr;         // keys per bucket (zeroed each step)
    int numTiles = 0, numBuckets = 0, tileKeys = 0;
    BucketPlan bkPlan;
    unsigned *longCnt = nullptr;            // grid: run count of the long-run table fill (zeroed each step)
    uint4 *longList = nullptr;
    unsigned longCap = 0;
    int segBits = 3;
    bool bucketSort = false;                // grid: sort finished per bucket of cells (bucket.hip)
    bool mortonBucket = false;              // LBVH30, one scene: Morton codes sorted per bucket of their top 10 bits
    long long numSegs = 0;
    float4 *sortedPos = nullptr, *sortedVel = nullptr;
    void *sortWs = nullptr;
    size_t sortWsBytes = 0;
    Lbvh tree{};
    float4 *treeBounds = nullptr;   // node boxes for debug views (the step keeps tree.bounds NULL)
    // the bodies' integration one step ahead (integrate_ahead): position, velocities, quaternions,
    // inverse world inertia
    float4 *aPos = nullptr, *aVel = nullptr, *aAng = nullptr, *aQinc = nullptr, *aQcum = nullptr;
    float *aIw = nullptr;
    uint64_t *code63 = nullptr;     // LBVH63 bucket path: the 63-bit codes in particle order
    int *overflowD = nullptr, *flagD = nullptr;
    float4 *aabbPart = nullptr, *aabbBox = nullptr;
    // static collider set (one-scene engines): points and normals sorted by cell, exact cell table
    int numStatic = 0;
    float4 *stPos = nullptr, *stNorm = nullptr;
    unsigned *stTable = nullptr;
    // sequential-impulse contact model: deepest contacts per particle, the contact list, scratch
    int sisCap = 0;
    float *sisDistA = nullptr, *sisDistB = nullptr;
    int *sisPartA = nullptr, *sisBodyB = nullptr, *sisPartB = nullptr;
    float4 *cNormal = nullptr, *cPoint = nullptr, *cImpulse = nullptr;
    int2 *cBodies = nullptr;
    float *cBias = nullptr;
    unsigned char *sisWsBuf = nullptr;
    int2 *sceneBodyD = nullptr;             // (first body, body count) per scene
    int maxSceneBodies = 0;
    SisWork sisW{};

    // step state
    bool first = true;
    void *prevSorted = nullptr;
    void *lastKeys = nullptr;
    unsigned *lastIdx = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool graphReady = false;

    ~Engine();
    int build(const std::vector<SceneSource> &src, const ArSimTuning *tuning, int dev, hipStream_t stream);
    int enqueue_step(hipEvent_t *ev, int *numStages);
    int step(int n);
    int enqueue_broadphase(int *pairsD, int *countsD, int maxPer);
    void invalidate_graph();
    int set_static(const float *pos4, const float *norm4, int n);
    void free_static();
    int reset_grid_tables();
    static const char *const *stage_names_all(int method, unsigned flags, bool bucket, bool fusedDem, int *count);
    // ahead: the step has no body-integration stage of its own (integrate_ahead)
    static const char *const *stage_names(int method, unsigned flags, bool bucket, bool fusedDem, bool ahead,
                                          int *count);
    // Bucket-sorted steps with the plain DEM keep the bodies' integration one step ahead: the previous
    // step's reduction stored it in the ahead buffers, the pose reads it there, and the reduction
    // stores the state after the step (in the state buffers) and the next integration (ahead).  The
    // step then has no integration launch.  The ahead buffers are filled from the state (prime) before
    // a step whenever the state or dt may have changed outside the steps (build, a trial's restore, a
    // new dt or static set).
    bool integrate_ahead() const { return (bucketSort || mortonBucket) && !(flags & ARSIM_CONTACT_SIS) && !fused_dem(); }
    int prime();
    BodyBufs ahead_bufs() const;
    bool primed = false;
    // the grid DEM and the per-body reduction run as one kernel (on request; exact cell table, no static set)
    bool fused_dem() const { return tu.fuseReduce && method == ARSIM_METHOD_GRID && !(flags & ARSIM_CONTACT_SIS) &&
                                    segBits == 0 && numStatic == 0; }
    SceneView scene_view() const;
    BodyBufs body_bufs() const;
    StateBufs state_bufs() const;
    PartBufs part_bufs() const;
    LongFill long_fill() const { return LongFill{longCnt, longList, longCap}; }
    size_t scene_bytes_estimate() const;
    // FNV-1a (64-bit) over scene s's state in the reference layout: particle pos and vel, then the
    // body positions, velocities, angular velocities and cumulative quaternions (synchronises)
    int state_checksum(int s, uint64_t *out);

  private:
    template <class T> int alloc(T **p, size_t count);
    int alloc_sis(int cap);
    template <class Rec>
    int enqueue_grid_broadphase(const BodyBufs &bb, const PartBufs &pb, const SceneView &sv, const SortTuning &stt,
                                bool grainVelDead, bool ahead, unsigned **kAOut, unsigned **vAOut, Rec &rec);
};

}  // namespace arsim

struct arsim_engine {
    arsim::Engine e;
};
