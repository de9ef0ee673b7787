// static.hip -- the static (augmented-reality) collider set: range image -> scene points and
// normals, the static points' grid index, and the DEM of the dynamic particles against them.
#include "csr_fill.hpp"
#include "kernels.hpp"

namespace arsim {

namespace {

// RGBDloader (Kernels/customRangeImageLoader.cu:14-109), one thread per pixel.  Intrinsics and the
// accelerometer rotation are the reference's constants; g is the normalised accelerometer vector.
// A 4x4 matrix-vector product is summed by columns as (c0 v0 + c1 v1) + (c2 v2 + c3 v3).
__global__ void k_range_image(const uint16_t *__restrict__ img, int W, int H, RangeCamera cam, float radius,
                              float4 *__restrict__ pos, float *__restrict__ radii, float *__restrict__ vao) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)W * H) return;
    if (idx > (long)(W - 1) * (long)(H - 1)) return;   // the reference's bound
    const int y = (int)(idx / W), x = (int)(idx - (long)y * W);
    float d = (float)img[idx];
    d = d / 5000.f;
    if (radii) radii[idx] = radius;
    if ((double)d < 0.01) d = -1000.f;
    const float X = d * ((float)x - 318.6f) / 517.3f;
    const float Y = d * ((float)y - 255.3f) / 516.5f;
    const float v0 = X, v1 = Y, v2 = d, v3 = 0.f;
    float t[4];
#pragma unroll
    for (int r = 0; r < 4; ++r)   // transpose(cam): element (r, c) = cam[4 r + c]
        t[r] = (cam.m[4 * r + 0] * v0 + cam.m[4 * r + 1] * v1) + (cam.m[4 * r + 2] * v2 + cam.m[4 * r + 3] * v3);
    const float gx = cam.g[0], gy = cam.g[1], gz = cam.g[2];
    const float G[16] = {1.f, 0.f, 0.f, 0.f, -gx, -gy, -gz, 0.f, 0.f, gz, -gy, 0.f, 0.f, 0.f, 0.f, 1.f};
    float u[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) u[r] = (G[r] * t[0] + G[4 + r] * t[1]) + (G[8 + r] * t[2] + G[12 + r] * t[3]);
    const float px = u[0], py = -u[1], pz = -u[2];
    if (vao) {
        vao[5 * idx] = px;
        vao[5 * idx + 1] = py;
        vao[5 * idx + 2] = pz;
        vao[5 * idx + 3] = (float)x / (float)W;
        vao[5 * idx + 4] = (float)y / (float)H;
    }
    pos[idx] = make_float4(px, py, pz, 0.f);
}

// computeNormals (customRangeImageLoader.cu:118-162): interior pixels only,
// n = normalize(cross(P[y+1][x] - P[y-1][x], P[y][x+1] - P[y][x-1])), normalize = v * (1 / |v|).
__global__ void k_range_normals(const float4 *__restrict__ pos, int W, int H, float4 *__restrict__ norm) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)W * H) return;
    const int y = (int)(idx / W), x = (int)(idx - (long)y * W);
    if (x > W - 2 || x == 0 || y == 0 || y > H - 2) return;
    const float4 r = pos[idx + 1], l = pos[idx - 1], b = pos[idx + W], t = pos[idx - W];
    const float dx0 = r.x - l.x, dx1 = r.y - l.y, dx2 = r.z - l.z;
    const float dy0 = b.x - t.x, dy1 = b.y - t.y, dy2 = b.z - t.z;
    const float n0 = dy1 * dx2 - dy2 * dx1;
    const float n1 = dy2 * dx0 - dy0 * dx2;
    const float n2 = dy0 * dx1 - dy1 * dx0;
    const float inv = 1.0f / sqrtf(n0 * n0 + n1 * n1 + n2 * n2);
    norm[idx] = make_float4(n0 * inv, n1 * inv, n2 * inv, 0.f);
}

// Cell key of every static point (calcHashD for the static set) and the identity index.
__global__ void k_static_hash(const float4 *__restrict__ pos, int m, SceneGeom g, unsigned *__restrict__ keys) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    int gx, gy, gz;
    grid_pos(xyz(pos[j]), g, gx, gy, gz);
    keys[j] = grid_hash(gx, gy, gz, g);
}

// Sorted copies of the static points and normals plus the exact cell table:
// table[c] = first sorted index with key >= c, for c = 0 .. numCells.
__global__ void k_static_reorder(const unsigned *__restrict__ keys, const unsigned *__restrict__ idx,
                                 const float4 *__restrict__ pos, const float4 *__restrict__ norm, int m,
                                 unsigned numCells, float4 *__restrict__ sPos, float4 *__restrict__ sNorm,
                                 unsigned *__restrict__ table, LongFill lf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < m;
    const unsigned k = valid ? keys[i] : 0u;
    const unsigned kp1 = (valid && i > 0) ? keys[i - 1] + 1u : 0u;
    wave_fill(table, valid && k + 1u > kp1, kp1, k + 1u, (unsigned)i, lf);
    wave_fill(table, i == m - 1, k + 1u, numCells + 1u, (unsigned)m, lf);
    if (!valid) return;
    const unsigned o = idx[i];
    sPos[i] = pos[o];
    sNorm[i] = norm[o];
}

// FindAndHandleAugmentedRealityCollisionsUniformGridKernel (RigidBodies/UniformGridDEM.cu:306-378):
// per sorted dynamic particle, the 27 cells of the static grid in z, y, x order, each cell's
// contacts summed into a partial that is added to the total for every cell.
//  REF (reference layout): static cellStart / cellEnd with 0xffffffff for an empty cell, normals read
//    through the static sorted index, every particle: pForce += F, pTorque += relativePos x F.
//  engine: the exact static cell table and sorted normals, body particles only: pForce += F (the
//    torque is relativePos x pForce, formed by the reduction).
template <bool REF>
__global__ void k_dem_static(const float4 *__restrict__ sPos, const float4 *__restrict__ sVel,
                             const unsigned *__restrict__ sIdx, int n, const float4 *__restrict__ stPos,
                             const float4 *__restrict__ stNorm, const unsigned *__restrict__ stIdx,
                             const unsigned *__restrict__ cs, const unsigned *__restrict__ ce, DevPhys ph,
                             SceneGeom g, float4 *__restrict__ pForce, float4 *__restrict__ pTorque,
                             const float4 *__restrict__ relPos) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 P = sPos[i];
    if (!REF && __float_as_int(P.w) < 0) return;
    const f3 pa = xyz(P), va = xyz(sVel[i]);
    int gx, gy, gz;
    grid_pos(pa, g, gx, gy, gz);
    f3 F = mk3(0.f, 0.f, 0.f);
    for (int z = -1; z <= 1; ++z)
        for (int y = -1; y <= 1; ++y)
            for (int x = -1; x <= 1; ++x) {
                const unsigned h = grid_hash(gx + x, gy + y, gz + z, g);
                f3 csum = mk3(0.f, 0.f, 0.f);
                unsigned j, e;
                if (REF) {
                    j = cs[h];
                    e = j == kEmpty ? j : ce[h];
                } else {
                    j = cs[h];
                    e = cs[h + 1];
                }
                for (; j < e; ++j) {
                    const f3 ps = xyz(stPos[j]);
                    const f3 ns = xyz(stNorm[REF ? stIdx[j] : j]);
                    f3 fj;
                    if (static_force(pa, ps, ns, va, ph, fj)) csum = add3(csum, fj);
                }
                F = add3(F, csum);
            }
    const unsigned o = sIdx[i];
    const float4 old = pForce[o];
    pForce[o] = make_float4(old.x + F.x, old.y + F.y, old.z + F.z, old.w + 0.0f);
    if (REF) {
        const float4 t = pTorque[o];
        const f3 tq = cross3(xyz(relPos[o]), F);
        pTorque[o] = make_float4(t.x + tq.x, t.y + tq.y, t.z + tq.z, t.w + 0.0f);
    }
}

}  // namespace

void launch_range_image(const uint16_t *img, int W, int H, const RangeCamera &cam, float radius, float4 *pos,
                        float4 *norm, float *radii, float *vao, hipStream_t st) {
    const long total = (long)W * H;
    if (total <= 0) return;
    const int block = 256;
    const dim3 grid((unsigned)((total + block - 1) / block));
    hipLaunchKernelGGL(k_range_image, grid, dim3(block), 0, st, img, W, H, cam, radius, pos, radii, vao);
    hipLaunchKernelGGL(k_range_normals, grid, dim3(block), 0, st, pos, W, H, norm);
}

void launch_static_hash(const float4 *pos, int m, const SceneGeom &g, unsigned *keys, hipStream_t st) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_static_hash, dim3((m + 255) / 256), dim3(256), 0, st, pos, m, g, keys);
}

void launch_static_reorder(const unsigned *keys, const unsigned *idx, const float4 *pos, const float4 *norm, int m,
                           unsigned numCells, float4 *sPos, float4 *sNorm, unsigned *table, const LongFill &lf,
                           hipStream_t st) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_static_reorder, dim3((m + 255) / 256), dim3(256), 0, st, keys, idx, pos, norm, m, numCells,
                       sPos, sNorm, table, lf);
    if (lf.list) launch_fill_long(table, (unsigned long long)numCells + 1ull, lf, st);
}

void launch_dem_static(const float4 *sPos, const float4 *sVel, const unsigned *sIdx, int n, const float4 *stPos,
                       const float4 *stNorm, const unsigned *stIdx, const unsigned *cs, const unsigned *ce,
                       const DevPhys &ph, const SceneGeom &g, float4 *pForce, float4 *pTorque,
                       const float4 *relPos, bool refLayout, hipStream_t st) {
    if (n <= 0) return;
    const dim3 grid((n + 127) / 128);
    if (refLayout)
        hipLaunchKernelGGL(k_dem_static<true>, grid, dim3(128), 0, st, sPos, sVel, sIdx, n, stPos, stNorm, stIdx, cs,
                           ce, ph, g, pForce, pTorque, relPos);
    else
        hipLaunchKernelGGL(k_dem_static<false>, grid, dim3(128), 0, st, sPos, sVel, sIdx, n, stPos, stNorm, stIdx,
                           cs, ce, ph, g, pForce, pTorque, relPos);
}

}  // namespace arsim
