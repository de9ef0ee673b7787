// scene_owned.hpp -- host-side scene arrays owned by the library (ArSimSceneOwned), shared by the
// generator, the template loader / scene builder and the engine snapshot.
#pragma once

#include <vector>

#include "../../include/arsim.h"

// Layout of ArSimSceneHost: body particles first, contiguous and in body order, then free grains.
struct ArSimSceneOwned {
    int N = 0, B = 0;
    std::vector<float> pos, vel, bodyFrame, rbPos, rbVel, rbAng, rbQuat, rbInertia, rbMass;
    std::vector<int> rb, ppo;
    ArSimParams params{};
    ArSimStepConfig cfg{};
};
