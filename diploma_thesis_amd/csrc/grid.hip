// grid.hip -- uniform-grid broadphase: cell bounds + reorder, and the 27-cell DEM kernel.
#include "kernels.hpp"

namespace arsim {

namespace {

// Maps hardware block b to a logical block so that the blocks one XCD runs (dispatch is
// round-robin over 8 XCDs) cover one contiguous slice of the sorted particles; neighbouring
// particles then share that XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int X = 8;
    int q = nb / X, r = nb % X;
    int x = b % X, k = b / X;
    return x * q + (x < r ? x : r) + k;
}

// reorderDataAndFindCellStartD (Kernels/particleKernels.cu:193-258).  The previous key is read
// straight from global memory (it is in the same or the previous cache line).  The sorted copies
// carry the rbIndex in sortedPos.w and, when scene ids are given, the scene id in sortedVel.w.
__global__ void k_reorder(const unsigned *__restrict__ keys, const unsigned *__restrict__ idx,
                          const float4 *__restrict__ pos, const float4 *__restrict__ vel, const int *__restrict__ rb,
                          const uint16_t *__restrict__ scene, float4 *__restrict__ sPos, float4 *__restrict__ sVel,
                          unsigned *__restrict__ cs, unsigned *__restrict__ ce, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned h = keys[i];
    if (i == 0 || h != keys[i - 1]) {
        cs[h] = (unsigned)i;
        if (i > 0) ce[keys[i - 1]] = (unsigned)i;
    }
    if (i == n - 1) ce[h] = (unsigned)n;
    unsigned o = idx[i];
    float4 p = pos[o];
    p.w = __int_as_float(rb[o]);
    sPos[i] = p;
    float4 v = vel[o];
    if (scene) v.w = __int_as_float((int)scene[o]);
    sVel[i] = v;
}

// FindAndHandleRigidBodyCollisionsUniformGridKernel (RigidBodies/UniformGridDEM.cu:133-204) and
// collideD (particleKernels.cu:302-406): 27 cells in z, y, x order, per-cell partial sums, the
// contact predicate j != i && (rbA != rbB || rbA == -1), then the wall term.  EMIT writes the
// contact list instead of forces.
template <bool EMIT>
__global__ void k_dem_grid(const float4 *__restrict__ sPos, const float4 *__restrict__ sVel,
                           const unsigned *__restrict__ sIdx, const unsigned *__restrict__ cs,
                           const unsigned *__restrict__ ce, int n, DevPhys ph, SceneView sv, int sceneInVelW,
                           DemOut o) {
    const int lb = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int i = lb * (int)blockDim.x + (int)threadIdx.x;
    if (i >= n) return;
    const float4 P = sPos[i];
    const float4 Vq = sVel[i];
    const int ra = __float_as_int(P.w);
    const int s = sceneInVelW ? __float_as_int(Vq.w) : 0;
    SceneGeom g;
    load_geom(sv, s, g);
    const f3 pa = xyz(P), va = xyz(Vq);
    int gx, gy, gz;
    grid_pos(pa, g, gx, gy, gz);
    const unsigned orig = sIdx[i];
    f3 F = mk3(0.f, 0.f, 0.f);
    int cnt = 0;
    for (int z = -1; z <= 1; ++z)
        for (int y = -1; y <= 1; ++y)
            for (int x = -1; x <= 1; ++x) {
                unsigned h = g.cellBase + grid_hash(gx + x, gy + y, gz + z, g);
                unsigned st = cs[h];
                if (st == kEmpty) continue;
                unsigned en = ce[h];
                f3 csum = mk3(0.f, 0.f, 0.f);
                for (unsigned j = st; j < en; ++j) {
                    float4 Q = sPos[j];
                    if ((int)j == i || !pair_allowed(ra, __float_as_int(Q.w))) continue;
                    if (EMIT) {
                        if (in_contact(pa, xyz(Q), ph.r)) {
                            if (cnt < o.maxPer) o.pairs[(size_t)orig * o.maxPer + cnt] = (int)sIdx[j];
                            ++cnt;
                        }
                    } else {
                        f3 fj;
                        if (pair_force(pa, xyz(Q), va, xyz(sVel[j]), ph, fj)) csum = add3(csum, fj);
                    }
                }
                F = add3(F, csum);
            }
    if (EMIT) {
        o.counts[orig] = cnt;
        return;
    }
    if (!((ph.flags & kGrainClamp) && ra < 0)) F = add3(F, wall_force(pa, va, ph, g));
    if (ra >= 0) {
        o.pForce[orig] = f4(F, 0.0f);
        if (o.pTorque) o.pTorque[orig] = f4(cross3(xyz(o.relPos[orig]), F), 0.0f);
    } else {
        const float fs = (ph.flags & kRefUnits) ? 1.0f : ph.dt;
        o.vel[orig] = make_float4(va.x + F.x * fs, va.y + F.y * fs, va.z + F.z * fs, 0.0f);
    }
}

}  // namespace

void launch_reorder(const unsigned *sortedKeys, const unsigned *sortedIdx, const float4 *pos, const float4 *vel,
                    const int *rb, const uint16_t *scene, float4 *sortedPos, float4 *sortedVel, unsigned *cellStart,
                    unsigned *cellEnd, int n, hipStream_t st) {
    if (n <= 0) return;
    const int block = 256;
    hipLaunchKernelGGL(k_reorder, dim3((n + block - 1) / block), dim3(block), 0, st, sortedKeys, sortedIdx, pos, vel,
                       rb, scene, sortedPos, sortedVel, cellStart, cellEnd, n);
}

void launch_dem_grid(const float4 *sortedPos, const float4 *sortedVel, const unsigned *sortedIdx,
                     const unsigned *cellStart, const unsigned *cellEnd, int n, const DevPhys &ph,
                     const SceneView &sv, bool sceneInVelW, const DemOut &o, int block, hipStream_t st) {
    if (n <= 0) return;
    if (block <= 0) block = 128;
    dim3 grid((n + block - 1) / block);
    if (o.pairs)
        hipLaunchKernelGGL(k_dem_grid<true>, grid, dim3(block), 0, st, sortedPos, sortedVel, sortedIdx, cellStart,
                           cellEnd, n, ph, sv, sceneInVelW ? 1 : 0, o);
    else
        hipLaunchKernelGGL(k_dem_grid<false>, grid, dim3(block), 0, st, sortedPos, sortedVel, sortedIdx, cellStart,
                           cellEnd, n, ph, sv, sceneInVelW ? 1 : 0, o);
}

}  // namespace arsim
