// serve.cpp -- continuous batching of independent scenes over a Poisson request stream.
//
// Requests arrive as scenes (config-0 generator: 4x4x4 cubes with 75 % of the particles plus
// 25 % free grains, box scaled to the particle count).  Each iteration of the loop:
//   1. retires scenes that finished their step count (their slots and bytes free up);
//   2. admits queued arrivals while the slot limit and the HBM budget allow;
//   3. if the resident set changed, rebuilds the batched engine by copying the surviving scenes'
//      device state into the new layout (device-to-device) and uploading the new ones;
//   4. advances every resident scene by one step in the batched launches.
// Per-scene results do not depend on the batch they run in: every stage is scene-local and
// order-preserving, which tests/ check by running a scene alone and in a mix.
// The one-at-a-time mode serves the same stream FIFO on the plain path (one engine per scene).
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "host_common.hpp"

using arsim_host::fail;

namespace {

struct Req {
    double arrival = 0;   // seconds (already divided by the time scale)
    int n = 0, steps = 0;
    uint64_t seed = 0;
    int done = 0;
    double admitted = -1, finished = -1;
    size_t bytes = 0;
    ArSimSceneOwned *scene = nullptr;
};

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

double pct(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    double idx = q * (v.size() - 1);
    size_t lo = (size_t)std::floor(idx), hi = (size_t)std::ceil(idx);
    return v[lo] + (v[hi] - v[lo]) * (idx - lo);
}

// Bytes a scene of n particles and b bodies occupies in a batched grid engine.
size_t scene_bytes(int n, int b, int method, int cellBits) {
    size_t per = 16 * 5 + 4 + 2 + 8 * 2 + 4 * 2 + 16 * 2 + 64;   // particle state, keys, values, sorted copies, sort ws
    if (method != ARSIM_METHOD_GRID) per += 2 * 32 * 2 + 16 * 2 + 4 * 6;
    size_t body = 16 * 5 + 36 * 2 + 12;
    size_t cells = method == ARSIM_METHOD_GRID ? ((size_t)8 << cellBits) : 0;
    return (size_t)n * per + (size_t)b * body + cells;
}

bool config_ok(const ArSimServeConfig *cfg) {
    return cfg && cfg->numRequests > 0 && cfg->arrivalRate > 0 && cfg->minParticles > 0 &&
           cfg->maxParticles >= cfg->minParticles && cfg->minSteps > 0 && cfg->maxSteps >= cfg->minSteps;
}

// The request stream of cfg: Poisson arrivals, log-uniform particle counts, uniform step counts,
// one scene seed per request.  Deterministic in cfg->seed.
std::vector<Req> draw_requests(const ArSimServeConfig *cfg) {
    const double scale = cfg->timeScale > 0 ? cfg->timeScale : 1.0;
    Rng rng(cfg->seed);
    std::vector<Req> reqs(cfg->numRequests);
    double t = 0;
    const double lmin = std::log((double)cfg->minParticles), lmax = std::log((double)cfg->maxParticles);
    for (int i = 0; i < cfg->numRequests; ++i) {
        double u = rng.uni();
        t += -std::log(1.0 - u) / cfg->arrivalRate;
        reqs[i].arrival = t / scale;
        reqs[i].n = (int)std::lround(std::exp(lmin + (lmax - lmin) * rng.uni()));
        reqs[i].steps = cfg->minSteps + (int)(rng.uni() * (cfg->maxSteps - cfg->minSteps + 1));
        if (reqs[i].steps > cfg->maxSteps) reqs[i].steps = cfg->maxSteps;
        reqs[i].seed = rng.next();
        reqs[i].bytes = scene_bytes(reqs[i].n, (int)(0.75 * reqs[i].n / 64), cfg->method, 18);
    }
    return reqs;
}

}  // namespace

extern "C" int arsim_serve_request(const ArSimServeConfig *cfg, int i, int *numParticles, int *numSteps,
                                   uint64_t *sceneSeed, double *arrivalSeconds) {
    if (!config_ok(cfg) || i < 0 || i >= cfg->numRequests)
        return fail(ARSIM_ERR_INVALID_ARG, "serve_request: bad configuration or index");
    const std::vector<Req> reqs = draw_requests(cfg);
    if (numParticles) *numParticles = reqs[i].n;
    if (numSteps) *numSteps = reqs[i].steps;
    if (sceneSeed) *sceneSeed = reqs[i].seed;
    if (arrivalSeconds) *arrivalSeconds = reqs[i].arrival;
    return ARSIM_OK;
}

// The config-0 generator scaled to n particles (used by serving and by the batched bench).
extern "C" int arsim_generate_mix_scene(int n, uint64_t seed, int method, ArSimSceneOwned **out) {
    ArSimGenSpec g{};
    g.seed = seed;
    g.cubeEdge = 4;
    g.numCubes = (int)std::lround(0.75 * n / 64.0);
    g.numGrains = n - 64 * g.numCubes;
    if (g.numGrains < 0) { g.numCubes = n / 64; g.numGrains = n - 64 * g.numCubes; }
    double L = 64.0 * std::cbrt((double)n / 16384.0);
    L = std::max(L, 16.0);
    for (int a = 0; a < 3; ++a) g.box[a] = (float)L;
    g.radius = 0.5f;
    g.bodyVelSigma = 0.5f;
    g.grainVelSigma = 0.5f;
    int lg = (int)std::ceil(std::log2(L));
    lg = std::min(std::max(lg, 2), 6);
    for (int a = 0; a < 3; ++a) g.gridLog2[a] = lg;
    g.method = method;
    g.dt = 1e-4f;
    g.spring = 5e4f;
    g.damping = 50.f;
    g.shear = 10.f;
    g.gravityY = -9.81f;
    return arsim_generate_scene(&g, out);
}

extern "C" int arsim_serve_run(const ArSimServeConfig *cfg, ArSimServeReport *rep) {
    using clk = std::chrono::steady_clock;
    if (!config_ok(cfg) || !rep) return fail(ARSIM_ERR_INVALID_ARG, "serve: bad configuration");
    // the batched loop admits while a slot and the budget allow: with no slot or no budget it could
    // never admit anything
    if (!cfg->oneAtATime && (cfg->maxResident <= 0 || !(cfg->hbmBudgetBytes > 0)))
        return fail(ARSIM_ERR_INVALID_ARG, "serve: batched serving needs maxResident > 0 and hbmBudgetBytes > 0");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(ARSIM_ERR_NO_DEVICE, "no HIP device");
    std::memset(rep, 0, sizeof(*rep));
    const int method = cfg->method;
    if (cfg->resultChecksums)
        for (int i = 0; i < cfg->numRequests; ++i) cfg->resultChecksums[i] = 0;

    // ---- request stream
    std::vector<Req> reqs = draw_requests(cfg);

    // ---- clients: request inputs are prepared ahead of their arrival by producer threads (a request
    // arrives with its scene), at most kLookahead requests ahead of the server.  Eight threads keep
    // up with arrivals 10x the specified rate; four left the server waiting 5.6 s for scenes.
    const int kLookahead = 1024;
    std::vector<std::atomic<int>> ready(cfg->numRequests);
    for (auto &x : ready) x.store(0);
    std::atomic<int> nextGen{0}, served{0}, genError{0};
    auto producer = [&]() {
        for (;;) {
            if (genError.load()) return;
            const int i = nextGen.fetch_add(1);
            if (i >= cfg->numRequests) return;
            while (i - served.load() > kLookahead && !genError.load())
                std::this_thread::sleep_for(std::chrono::microseconds(200));
            ArSimSceneOwned *sc = nullptr;
            if (arsim_generate_mix_scene(reqs[i].n, reqs[i].seed, method, &sc) != ARSIM_OK) sc = nullptr;
            reqs[i].scene = sc;
            ready[i].store(sc ? 1 : -1);
            if (!sc) return;
        }
    };
    std::vector<std::thread> producers;
    for (int k = 0; k < 8; ++k) producers.emplace_back(producer);
    struct Joiner {   // on every exit: stop the producers and free scenes nobody took
        std::vector<std::thread> &t;
        std::atomic<int> &stop;
        std::vector<Req> &reqs;
        ~Joiner() {
            stop.store(1);
            for (auto &x : t) x.join();
            for (auto &r : reqs)
                if (r.scene) { arsim_scene_free(r.scene); r.scene = nullptr; }
        }
    } joiner{producers, genError, reqs};
    // the request's scene, waiting for its producer if it is not ready yet (NULL on failure)
    auto take = [&](int i) -> ArSimSceneOwned * {
        while (ready[i].load() == 0) std::this_thread::sleep_for(std::chrono::microseconds(50));
        ArSimSceneOwned *sc = reqs[i].scene;
        reqs[i].scene = nullptr;
        served.store(i + 1);
        return ready[i].load() > 0 ? sc : nullptr;
    };

    std::vector<double> stepLat, complLat, waitLat;
    double updates = 0;
    int completed = 0, maxResident = 0, rejected = 0, deferred = 0;
    double peakBytes = 0;
    const auto t0 = clk::now();
    auto now = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
    ArSimTuning tune{128, 16, 8, 0, 0, 0};
    double lastReport = 0.0;
    auto progress = [&](int done) {   // a line on stderr every 10 s of wall time
        const double t = now();
        if (t - lastReport >= 10.0) {
            lastReport = t;
            std::fprintf(stderr, "[serve] %.0f s: %d / %d requests completed\n", t, done, cfg->numRequests);
            std::fflush(stderr);
        }
    };

    if (cfg->oneAtATime) {
        arsim::Engine eng;   // rebuilt for every request, keeping its allocations
        for (auto &r : reqs) {
            while (now() < r.arrival) std::this_thread::sleep_for(std::chrono::microseconds(200));
            ArSimSceneOwned *sc = take((int)(&r - &reqs[0]));
            if (!sc) return fail(ARSIM_ERR_INVALID_ARG, "serve: request generation failed");
            int rc = 0;
            ArSimSceneHost v;
            arsim_scene_view(sc, &v);
            r.admitted = now();
            std::vector<arsim::SceneSource> src(1);
            src[0].host = &v;
            ArSimTuning tg = tune;
            tg.useGraph = 1;
            rc = eng.build(src, &tg, 0, nullptr);
            arsim_scene_free(sc);
            if (rc) return rc;
            double s0 = now();
            if ((rc = eng.step(r.steps)) != 0) return rc;
            hipStreamSynchronize(eng.st);
            double s1 = now();
            r.finished = s1;
            if (cfg->resultChecksums && (rc = eng.state_checksum(0, &cfg->resultChecksums[&r - &reqs[0]])) != 0)
                return rc;
            for (int k = 0; k < r.steps; ++k) stepLat.push_back((s1 - s0) / r.steps * 1e3);
            complLat.push_back((r.finished - r.arrival) * 1e3);
            waitLat.push_back((r.admitted - r.arrival) * 1e3);
            updates += (double)v.numParticles * r.steps;
            peakBytes = std::max(peakBytes, (double)eng.devBytes);
            ++completed;
            maxResident = 1;
            progress(completed);
        }
    } else {
        std::deque<int> pending;
        for (int i = 0; i < cfg->numRequests; ++i) pending.push_back(i);
        std::vector<int> resident;            // request ids, in engine scene order
        // two engines alternate: a change of the resident set rebuilds the idle one from the
        // active one (surviving scenes move on the device in one launch per array group) and
        // keeps its allocations from earlier rebuilds
        std::unique_ptr<arsim::Engine> engs[2];
        int cur = 0;
        arsim::Engine *eng = nullptr;
        size_t residentBytes = 0;
        std::vector<char> countedDefer(cfg->numRequests, 0);
        while (completed + rejected < cfg->numRequests) {
            bool changed = false;
            // retire
            std::vector<int> keep;
            for (size_t k = 0; k < resident.size(); ++k) {
                const int id = resident[k];
                Req &r = reqs[id];
                if (r.done >= r.steps) {
                    if (cfg->resultChecksums) {
                        const int rc = eng->state_checksum((int)k, &cfg->resultChecksums[id]);
                        if (rc) return rc;
                    }
                    complLat.push_back((r.finished - r.arrival) * 1e3);
                    residentBytes -= r.bytes;
                    ++completed;
                    changed = true;
                } else {
                    keep.push_back(id);
                }
            }
            // admit
            std::vector<int> admitted;
            double tn = now();
            while (!pending.empty() && reqs[pending.front()].arrival <= tn &&
                   (int)(keep.size() + admitted.size()) < cfg->maxResident) {
                Req &r = reqs[pending.front()];
                if ((double)r.bytes > cfg->hbmBudgetBytes) {
                    // larger than the whole budget: it can never be admitted, so it is refused
                    ArSimSceneOwned *sc = take(pending.front());
                    if (sc) arsim_scene_free(sc);
                    ++rejected;
                    pending.pop_front();
                    continue;
                }
                if ((double)(residentBytes + r.bytes) > cfg->hbmBudgetBytes) {
                    if (!countedDefer[pending.front()]) { ++deferred; countedDefer[pending.front()] = 1; }
                    break;
                }
                r.scene = take(pending.front());
                if (!r.scene) return fail(ARSIM_ERR_INVALID_ARG, "serve: request generation failed");
                r.admitted = now();
                waitLat.push_back((r.admitted - r.arrival) * 1e3);
                residentBytes += r.bytes;
                admitted.push_back(pending.front());
                pending.pop_front();
                changed = true;
            }
            if (keep.empty() && admitted.empty()) {
                eng = nullptr;
                resident.clear();
                if (pending.empty()) break;
                double wait = reqs[pending.front()].arrival - now();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::microseconds((long long)(wait * 1e6)));
                continue;
            }
            if (changed) {
                std::vector<arsim::SceneSource> src;
                std::vector<ArSimSceneHost> views(admitted.size());
                for (size_t k = 0; k < keep.size(); ++k) {
                    arsim::SceneSource s;
                    s.from = eng;
                    s.fromScene = (int)(std::find(resident.begin(), resident.end(), keep[k]) - resident.begin());
                    src.push_back(s);
                }
                for (size_t k = 0; k < admitted.size(); ++k) {
                    arsim_scene_view(reqs[admitted[k]].scene, &views[k]);
                    arsim::SceneSource s;
                    s.host = &views[k];
                    src.push_back(s);
                }
                const int nxt = eng ? 1 - cur : cur;
                if (!engs[nxt]) engs[nxt].reset(new arsim::Engine());
                if (eng) hipStreamSynchronize(eng->st);
                int rc = engs[nxt]->build(src, &tune, 0, nullptr);
                for (int id : admitted) { arsim_scene_free(reqs[id].scene); reqs[id].scene = nullptr; }
                if (rc) return rc;
                cur = nxt;
                eng = engs[cur].get();
                resident = keep;
                resident.insert(resident.end(), admitted.begin(), admitted.end());
                maxResident = std::max(maxResident, (int)resident.size());
                peakBytes = std::max(peakBytes, (double)(engs[0] ? engs[0]->devBytes : 0) +
                                                    (double)(engs[1] ? engs[1]->devBytes : 0));
            }
            progress(completed);
            double s0 = now();
            int rc = eng->step(1);
            if (rc) return rc;
            hipStreamSynchronize(eng->st);
            double s1 = now();
            for (size_t k = 0; k < resident.size(); ++k) {
                Req &r = reqs[resident[k]];
                ++r.done;
                updates += (double)eng->pCount[k];
                stepLat.push_back((s1 - s0) * 1e3);
                if (r.done >= r.steps) r.finished = s1;
            }
        }
    }
    rep->completed = completed;
    rep->wallSeconds = now();
    rep->particleUpdates = updates;
    rep->updatesPerSecond = updates / std::max(rep->wallSeconds, 1e-9);
    rep->stepLatencyP50Ms = pct(stepLat, 0.50);
    rep->stepLatencyP99Ms = pct(stepLat, 0.99);
    rep->completionLatencyP50Ms = pct(complLat, 0.50);
    rep->completionLatencyP99Ms = pct(complLat, 0.99);
    rep->queueWaitP50Ms = pct(waitLat, 0.50);
    rep->queueWaitP99Ms = pct(waitLat, 0.99);
    rep->maxResidentSeen = maxResident;
    rep->rejectedForBudget = rejected;
    rep->deferredForBudget = deferred;
    rep->peakDeviceBytes = peakBytes;
    return ARSIM_OK;
}
