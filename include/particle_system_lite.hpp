/*
 * particle_system_lite.hpp -- the simulator's C++ step interface over the C ABI of arsim.h.
 *
 * ParticleSystemLite keeps the shape of the reference's ParticleSystem (Particles/particleSystem.h):
 * update(deltaTime) (:288, updateFunctions.cpp:2803) advances the scene by one DEM step,
 * getArray(POSITION / VELOCITY) (:92) returns the particle state as float4 host arrays,
 * getNumParticles() (:95), getNumberOfObjects() (:643) and getModelMatrixArray() (:642) the counts and
 * the per-body render state.  Unlike the reference it owns no GL buffers, never exits the process
 * (errors throw std::runtime_error with arsim_last_error()), and keeps the whole state on the device
 * between updates: getArray copies it back on demand.
 *
 * Header-only; link with -larsim.
 */
#ifndef ARSIM_PARTICLE_SYSTEM_LITE_HPP
#define ARSIM_PARTICLE_SYSTEM_LITE_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "arsim.h"

namespace arsim_cpp {

class ParticleSystemLite {
  public:
    enum ParticleArray { POSITION, VELOCITY };

    // The scene is copied to the device; the host arrays behind `scene` may be released afterwards.
    explicit ParticleSystemLite(const ArSimSceneHost &scene, int device = 0, bool useGraph = true)
        : numParticles_(scene.numParticles), numBodies_(scene.numBodies), dt_(scene.cfg.dt) {
        ArSimTuning tuning = {128, 16, 8, useGraph ? 1 : 0, 0, 0};
        check(arsim_engine_create(&scene, 1, &tuning, device, &engine_), "engine_create");
    }
    ~ParticleSystemLite() {
        if (engine_) arsim_engine_destroy(engine_);
    }
    ParticleSystemLite(const ParticleSystemLite &) = delete;
    ParticleSystemLite &operator=(const ParticleSystemLite &) = delete;

    // One step of deltaTime (enqueued on the engine's stream; the host does not wait).
    void update(float deltaTime) {
        if (deltaTime != dt_) {
            check(arsim_engine_set_dt(engine_, deltaTime), "engine_set_dt");
            dt_ = deltaTime;
        }
        check(arsim_engine_step(engine_, 1), "engine_step");
    }

    // Waits for the enqueued steps; reports a traversal-stack overflow or a device error.
    void synchronize() { check(arsim_engine_synchronize(engine_), "engine_synchronize"); }

    // float4 per particle (x, y, z, w), copied from the device after the enqueued steps.
    const float *getArray(ParticleArray array) {
        synchronize();
        std::vector<float> &dst = array == POSITION ? hostPos_ : hostVel_;
        dst.resize(4 * (size_t)numParticles_);
        check(arsim_engine_download(engine_, 0, array == POSITION ? dst.data() : nullptr,
                                    array == VELOCITY ? dst.data() : nullptr, nullptr, nullptr, nullptr, nullptr,
                                    nullptr, nullptr, nullptr, nullptr),
              "engine_download");
        return dst.data();
    }

    // Column-major mat4 per body: rotation of the cumulative quaternion and the centre of mass.
    const float *getModelMatrixArray() {
        synchronize();
        hostModel_.resize(16 * (size_t)numBodies_);
        check(arsim_engine_render_state(engine_, 0, hostModel_.data(), nullptr), "engine_render_state");
        return hostModel_.data();
    }

    int getNumParticles() const { return numParticles_; }
    const int &getNumberOfObjects() const { return numBodies_; }

    // FNV-1a checksum of the state in the reference layout (see arsim_engine_state_checksum).
    uint64_t stateChecksum() {
        uint64_t h = 0;
        check(arsim_engine_state_checksum(engine_, 0, &h), "engine_state_checksum");
        return h;
    }

    arsim_engine *engine() { return engine_; }

  private:
    static void check(int rc, const char *what) {
        if (rc != ARSIM_OK) throw std::runtime_error(std::string(what) + ": " + arsim_last_error());
    }

    arsim_engine *engine_ = nullptr;
    int numParticles_ = 0, numBodies_ = 0;
    float dt_ = 0.f;
    std::vector<float> hostPos_, hostVel_, hostModel_;
};

}  // namespace arsim_cpp

#endif  // ARSIM_PARTICLE_SYSTEM_LITE_HPP
