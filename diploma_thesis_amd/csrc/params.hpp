// params.hpp -- conversion of the reference's SimParams + step configuration into the kernels'
// parameter structs (host side).
#pragma once

#include <cmath>
#include <string>

#include "../../include/arsim.h"
#include "host_common.hpp"
#include "kernels.hpp"

namespace arsim {

inline bool is_pow2(unsigned v) { return v != 0u && (v & (v - 1u)) == 0u; }

inline DevPhys make_phys(const ArSimParams &p, const ArSimStepConfig &c) {
    DevPhys ph{};
    ph.k = p.spring; ph.c = p.damping; ph.s = p.shear; ph.a = p.attraction;
    ph.r = p.particleRadius; ph.dt = c.dt;
    ph.gx = p.gravity.x; ph.gy = p.gravity.y; ph.gz = p.gravity.z;
    ph.globalDamping = p.globalDamping; ph.Wdamping = p.Wdamping; ph.boundaryDamping = p.boundaryDamping;
    ph.flags = c.flags;
    // smallest float not below (2r)^2 in exact arithmetic (see DevPhys::cd2Reject)
    const float cd = 2.0f * p.particleRadius;
    const double cd2 = (double)cd * (double)cd;
    float t = (float)cd2;
    if ((double)t < cd2) t = std::nextafter(t, INFINITY);
    ph.cd2Reject = t;
    return ph;
}

// Grid sizes must be powers of two: calcGridHash wraps cells with & (size - 1)
// (Kernels/particleKernels.cu:161-167).
inline int make_geom(const ArSimParams &p, const ArSimStepConfig &c, int pStart, int pCount, SceneGeom &g) {
    g = SceneGeom{};
    const unsigned gs[3] = {p.gridSize.x, p.gridSize.y, p.gridSize.z};
    for (int a = 0; a < 3; ++a)
        if (gs[a] != 0u && !is_pow2(gs[a]))
            return arsim_host::fail(ARSIM_ERR_NOT_POW2, "grid size must be a power of two per axis");
    g.origin[0] = p.worldOrigin.x; g.origin[1] = p.worldOrigin.y; g.origin[2] = p.worldOrigin.z;
    g.cellSize[0] = p.cellSize.x; g.cellSize[1] = p.cellSize.y; g.cellSize[2] = p.cellSize.z;
    for (int a = 0; a < 3; ++a) g.gridMask[a] = gs[a] ? gs[a] - 1u : 0u;
    g.gx = gs[0];
    g.gxy = gs[0] * gs[1];
    g.cellBase = 0u;
    g.wallMask = c.wallMask;
    for (int a = 0; a < 3; ++a) {
        g.boxMin[a] = c.boxMin[a];
        g.boxMax[a] = c.boxMax[a];
        g.mLo[a] = c.boxMin[a];
        g.mExt[a] = c.boxMax[a] - c.boxMin[a];
    }
    g.pStart = pStart;
    g.pCount = pCount;
    return ARSIM_OK;
}

}  // namespace arsim
