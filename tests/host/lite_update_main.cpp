// A C++ caller of the library through include/arsim.h and include/particle_system_lite.hpp only:
// builds the small-scene generator's scene of `n` particles (seed, grid method), calls
// ParticleSystemLite::update(dt) `steps` times, and writes the particle positions, velocities and
// body model matrices as raw float32 (in that order) to `out`.
//   lite_update <out> <n> <seed> <steps>
#include <cstdio>
#include <cstdlib>
#include <exception>

#include "arsim.h"
#include "particle_system_lite.hpp"

int main(int argc, char **argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s <out> <n> <seed> <steps>\n", argv[0]);
        return 2;
    }
    const int n = std::atoi(argv[2]), steps = std::atoi(argv[4]);
    const uint64_t seed = std::strtoull(argv[3], nullptr, 10);
    ArSimSceneOwned *owned = nullptr;
    if (arsim_generate_mix_scene(n, seed, ARSIM_METHOD_GRID, &owned) != ARSIM_OK) {
        std::fprintf(stderr, "generate: %s\n", arsim_last_error());
        return 1;
    }
    ArSimSceneHost view;
    arsim_scene_view(owned, &view);
    try {
        arsim_cpp::ParticleSystemLite ps(view);
        for (int k = 0; k < steps; ++k) ps.update(view.cfg.dt);
        const int N = ps.getNumParticles(), B = ps.getNumberOfObjects();
        FILE *f = std::fopen(argv[1], "wb");
        if (!f) return 1;
        std::fwrite(ps.getArray(arsim_cpp::ParticleSystemLite::POSITION), sizeof(float), 4 * (size_t)N, f);
        std::fwrite(ps.getArray(arsim_cpp::ParticleSystemLite::VELOCITY), sizeof(float), 4 * (size_t)N, f);
        std::fwrite(ps.getModelMatrixArray(), sizeof(float), 16 * (size_t)B, f);
        std::fclose(f);
        std::printf("particles %d bodies %d steps %d checksum %016llx\n", N, B, steps,
                    (unsigned long long)ps.stateChecksum());
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        arsim_scene_free(owned);
        return 1;
    }
    arsim_scene_free(owned);
    return 0;
}
