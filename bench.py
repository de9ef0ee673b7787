"""Benchmark of the gfx950 per-step DEM pipeline.

Headline line (value): particle-updates/s of the plain one-scene path on the large single scene
(2^20 particles: 8,192 rigid 4x4x4 cubes + 524,288 grains in [0,512]^3, hashed 256^3 grid), with
inputs resident in HBM.  Beside it on the same JSON line:
  - p50 / p99 per-step latency of that scene (HIP events, one step each);
  - roofline of the dominant kernel (algorithmic bytes / live HIP-event duration vs 8 TB/s HBM);
  - cpu_baseline: the oracle's serial FP32 restatement on a bounded sample of the same scene;
  - batched: 512 scenes log-uniform in [1024, 65536] particles stepped in one launch per stage,
    autotuned, vs the plain path looping over the same scenes.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic HBM bytes per launch of each stage: compulsory traffic in the reference layout with
# neighbour re-reads assumed cached, as (per particle, per grid cell, per body particle, per body).
# A fused kernel counts the sum of its stages.
ALGO_BYTES = {
    "dem_grid": (120.0, 8.0, 0.0, 0.0),     # sorted pos+vel 32, index 4, rbIndex 4, relativePos 16, pForce/pTorque
                                            # RMW 64; cellStart + cellEnd 8 per cell
    "dem_lbvh": (120.0, 0.0, 0.0, 0.0),
    "reorder_cellstart": (72.0, 4.0, 0.0, 0.0),   # hash 4, index 4, gathered pos+vel 32, sorted pos+vel 32; cellStart
                                                  # clear 4 per cell (the 8 B per occupied cell are left out)
    "pose_key": (68.0 + 24.0, 0.0, 0.0, 0.0),     # body integrate + pose 68, hash 24
    "sort": (4.0 + 16.0 * 3, 0.0, 0.0, 0.0),      # onesweep, u32 key + u32 value, 3 digit passes over the 24-bit key
    "reduce_apply": (0.0, 0.0, 32.0, 240.0),      # pForce + relativePos of each body particle; body state
    # bucket sort: the top digit ranked per tile inside pose_key, the two remaining passes fused with the
    # reorder and the cell table
    "bucket_sort_reorder": (72.0 + 2 * 16.0, 4.0, 0.0, 0.0),
    # the grid DEM fused with the per-body reduction and apply
    "dem_reduce": (120.0, 8.0, 32.0, 240.0),
    # LBVH 30-bit codes: the sort within each bucket of the top 10 bits (tile-ranked keys and indices in,
    # sorted keys and indices out)
    "bucket_sort": (16.0, 0.0, 0.0, 0.0),
}
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--method", choices=["grid", "lbvh30", "lbvh63"], default="grid")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-scenes", type=int, default=512)
    ap.add_argument("--batched-steps", type=int, default=100)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-rate", type=float, default=500.0,
                    help="Poisson arrival rate of the serving leg (scenes/s); 10x the specified 50/s by default, so "
                         "that the leg measures serving capacity rather than the arrival rate")
    ap.add_argument("--quick", action="store_true",
                    help="timed steps only (no latency / stage loops, no CPU baseline, no batched run): for profilers")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a plain 256 MiB device copy kernel (3 launches) to calibrate HBM byte counters, then exit")
    ap.add_argument("--batched-only", action="store_true",
                    help="only step the 512-scene batch (default tuning, no autotune): for profilers")
    ap.add_argument("--traversal-stats", action="store_true",
                    help="LBVH packet traversal statistics of the 2^20-particle scene (nodes visited per wave, leaf "
                         "tests per particle, stack depth) from the counting build libarsim_traversal_stats.so")
    ap.add_argument("--scene", choices=["small", "pile", "sis"], default=None,
                    help="time another plain scene instead of the headline one: 'small' = 256 cubes (16,384 "
                         "particles, 64^3 grid), 'pile' = the dense ellipsoid pile (~4.2 M particles), 'sis' = the "
                         "sequential-impulse model on two layers of 32 x 32 cubes resting on a static floor "
                         "(131,072 particles)")
    return ap.parse_args()


def other_scene(args):
    """Plain-path throughput of the small scene or the dense pile (stage times included)."""
    from diploma_thesis_amd import Engine, capi, scenes
    method = {"grid": capi.METHOD_GRID, "lbvh30": capi.METHOD_LBVH30, "lbvh63": capi.METHOD_LBVH63}[args.method]
    floor = None
    if args.scene == "sis":
        sc, fp, fn = scenes.stacked_cubes(32, 32, 2)
        floor = (fp, fn)
    else:
        sc = scenes.config0(method) if args.scene == "small" else scenes.config2(method)
    cpu = None
    if floor is not None and not args.no_cpu_baseline:
        cpu = cpu_leg(sc, 5, 10.0, "the stacked-cubes scene", static=floor)
    with Engine(sc, use_graph=True) as e:
        if floor is not None:
            e.set_static(*floor)
        e.step(args.warmup)
        e.sync()
        t0 = time.perf_counter()
        e.step(args.steps)
        e.sync()
        sec = time.perf_counter() - t0
        stage_ms = {}
        for _ in range(3):
            for name, ms in e.profile():
                stage_ms[name] = stage_ms.get(name, 0.0) + ms / 3
        finite = e.check_finite()
    print(json.dumps({"scene": args.scene, "method": args.method, "particles": sc.n, "bodies": sc.b,
                      "steps": args.steps, "updates_per_s": sc.n * args.steps / sec,
                      "ms_per_step": sec * 1e3 / args.steps, "stage_ms": stage_ms, "finite": finite,
                      "cpu_baseline": cpu}))


def traversal_stats(args):
    """Counters of the LBVH packet traversal over `args.steps` steps of the 2^20-particle scene."""
    os.environ["ARSIM_LIB"] = os.path.join(ROOT, "diploma_thesis_amd", "libarsim_traversal_stats.so")
    from diploma_thesis_amd import Engine, capi, scenes
    fn = capi.lib.arsim_debug_traversal_stats
    fn.restype, fn.argtypes = C.c_int, [C.c_void_p, C.c_int]
    method = {"grid": capi.METHOD_LBVH30, "lbvh30": capi.METHOD_LBVH30, "lbvh63": capi.METHOD_LBVH63}[args.method]
    sc = scenes.config1(method)
    out = np.zeros(10, np.uint64)
    with Engine(sc, use_graph=False) as e:
        e.step(1)
        e.sync()
        capi.check(fn(None, 1), "reset")
        e.step(args.steps)
        e.sync()
        capi.check(fn(capi.ptr(out), 1), "stats")
    waves, visits, lanes, leaf_tests, interest, max_sp, done, contacts, packets, iters = (int(x) for x in out)
    width = 64 * waves / max(packets, 1)   # lanes per packet
    print(json.dumps({"traversal_stats": True, "method": "lbvh30" if method == capi.METHOD_LBVH30 else "lbvh63",
                      "particles": sc.n, "steps": args.steps, "waves": waves, "packets": packets,
                      "lanes_per_packet": width,
                      "nodes_visited_per_packet": visits / max(packets, 1),
                      "node_loop_iterations_per_wave": iters / max(waves, 1),
                      "leaf_tests_per_particle": leaf_tests / max(lanes, 1),
                      "contacts_per_particle": contacts / max(lanes, 1),
                      "packet_efficiency": interest / max(visits * width, 1),
                      "max_stack_depth": max_sp, "completed_lanes": done, "valid_lanes": lanes}))


def batched_only(args):
    from diploma_thesis_amd import Engine, capi, scenes
    mix = scenes.batch_mix(args.batched_scenes, seed=0, method=capi.METHOD_GRID)
    total = sum(s.n for s in mix)
    be = Engine(mix, use_graph=True)
    be.step(args.warmup)
    be.sync()
    t0 = time.perf_counter()
    be.step(args.steps)
    be.sync()
    sec = time.perf_counter() - t0
    stage_ms = {}
    for _ in range(3):
        for name, ms in be.profile():
            stage_ms[name] = stage_ms.get(name, 0.0) + ms / 3
    be.close()
    print(json.dumps({"batched_only": True, "particles": total, "steps": args.steps,
                      "updates_per_s": total * args.steps / sec, "ms_per_step": sec * 1e3 / args.steps,
                      "stage_ms": stage_ms}))


def calibrate_copy():
    """A 4-byte-per-element copy kernel of known size: 64 Mi elements, 256 MiB read + 256 MiB written."""
    from diploma_thesis_amd import capi
    n = 64 << 20
    a, b = C.c_void_p(), C.c_void_p()
    capi.check(capi.lib.arsim_malloc(C.byref(a), 4 * n))
    capi.check(capi.lib.arsim_malloc(C.byref(b), 4 * n))
    capi.check(capi.lib.arsim_memset(a, 0, 4 * n, None))
    for _ in range(3):
        capi.check(capi.lib.arsim_rebase_indices(b, a, n, 0, None))
    capi.check(capi.lib.arsim_stream_synchronize(None))
    capi.lib.arsim_free(a)
    capi.lib.arsim_free(b)
    print(json.dumps({"calibration": "copy", "elements": n, "bytes_read": 4 * n, "bytes_written": 4 * n}))


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_info():
    """Core count, the cores this process may use and the CPU model of the host the baseline runs on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "usable_cpus": len(os.sched_getaffinity(0)), "cpu_model": model}


class one_core:
    """Pins this process to one of its usable cores for the serial CPU baseline, then restores the set."""

    def __enter__(self):
        self.prev = os.sched_getaffinity(0)
        self.core = min(self.prev)
        os.sched_setaffinity(0, {self.core})
        return self

    def __exit__(self, *exc):
        os.sched_setaffinity(0, self.prev)


def cpu_leg(sc, min_steps, target_s, what, static=None):
    """The oracle's serial FP32 step on scene sc, pinned to one core: one untimed step, then about
    target_s seconds of steps (at least min_steps)."""
    import oracle
    st = oracle.State(sc, np.float32)
    if static is not None:
        st.set_static(*static)
    with one_core() as pin:
        t0 = time.perf_counter()
        st.step(1)
        one = time.perf_counter() - t0
        k = max(min_steps, min(200, int(target_s / max(one, 1e-3))))
        t0 = time.perf_counter()
        st.step(k)
        sec = time.perf_counter() - t0
    return {"value": sc.n * k / sec, "unit": "particle-updates/s", "cores": 1, "kind": "port",
            "sample": f"{k} steps of {what} after 1 untimed step, oracle FP32 serial C++ (g++ -O2), {sec:.1f} s",
            "pinned_core": pin.core, **host_info()}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this pipeline runs on one GPU; multi-GPU runs are not part of this build "
                                   "(independent scenes shard across GPUs with no exchange)", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.batched_only:
        batched_only(args)
        return
    if args.traversal_stats:
        traversal_stats(args)
        return
    if args.scene:
        other_scene(args)
        return
    from diploma_thesis_amd import Engine, capi, scenes

    method = {"grid": capi.METHOD_GRID, "lbvh30": capi.METHOD_LBVH30, "lbvh63": capi.METHOD_LBVH63}[args.method]
    log("generating the 2^20-particle scene")
    sc = scenes.config1(method)
    n = sc.n
    eng = Engine(sc, use_graph=True)
    tuning = None
    if not args.quick:
        # block size / items per thread / digit width picked by the engine's autotuner (state is restored)
        best, best_ms = eng.autotune(steps_per_trial=3)
        eng.close()
        eng = Engine(sc, dem_block=best[0], sort_ipt=best[1], sort_radix_bits=best[2], use_graph=True,
                     bucket_split=best[3])
        tuning = {"dem_block": best[0], "sort_items_per_thread": best[1], "sort_radix_bits": best[2],
                  "bucket_split": best[3]}
    log(f"engine ready: {n} particles, {sc.b} bodies, {eng.device_bytes / 2**20:.0f} MiB on device, tuning {tuning}")
    eng.step(args.warmup)
    eng.sync()

    # ---- timed region: K steps, inputs resident in HBM
    eng.sync()
    t0 = time.perf_counter()
    eng.step(args.steps)
    eng.sync()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = n * args.steps / (t1 - t0)
    log(f"plain: {ms_per_step:.4f} ms/step, {value:.4e} particle-updates/s")
    if not eng.check_finite():
        log("WARNING: non-finite state after the timed steps")
    if args.quick:
        eng.close()
        print(json.dumps({"metric": "particle-updates/s", "value": value, "ms_per_step": ms_per_step,
                          "steps": args.steps, "quick": True}))
        return

    # ---- per-step latency (one step per event pair) and per-stage times (HIP events on the engine stream)
    lat = []
    for _ in range(50):
        s0 = time.perf_counter()
        eng.step(1)
        eng.sync()
        lat.append((time.perf_counter() - s0) * 1e3)
    stage_ms = {}
    reps = 10
    for _ in range(reps):
        for name, ms in eng.profile():
            stage_ms[name] = stage_ms.get(name, 0.0) + ms / reps
    dom = max((k for k in stage_ms if not k.startswith("sort")), key=lambda k: stage_ms[k])
    cells = int(sc.params.numCells)

    body_particles = int((sc.rb >= 0).sum())

    def algo_bytes(stage):
        per_p, per_c, per_bp, per_b = ALGO_BYTES.get(stage, (0.0, 0.0, 0.0, 0.0))
        return per_p * n + per_c * cells + per_bp * body_particles + per_b * sc.b

    achieved = algo_bytes(dom) / (stage_ms[dom] * 1e-3) / 1e9
    traffic = None
    pmc_path = os.path.join(ROOT, "profiles", f"{dom}_c1_pmc.json")
    if os.path.exists(pmc_path):
        with open(pmc_path) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "algorithmic_bytes": int(algo_bytes(dom)), "kernel_ms": round(stage_ms[dom], 5),
                "stage_frac": {k: round(algo_bytes(k) / (v * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                               for k, v in stage_ms.items() if k in ALGO_BYTES and args.method == "grid"}}
    stage_summary = {k: round(v, 5) for k, v in stage_ms.items()}
    eng.close()

    # ---- CPU baseline: the oracle's serial FP32 restatement of the same step on the same scene
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_leg(sc, 1, 15.0, "the 2^20-particle scene")
        log(f"cpu baseline: {cpu['value']:.3e} particle-updates/s ({cpu['sample']})")

    # ---- batched form vs the plain path looping over scenes
    batched = None
    if not args.no_batched:
        log(f"generating {args.batched_scenes} scenes")
        mix = scenes.batch_mix(args.batched_scenes, seed=0, method=capi.METHOD_GRID)
        total = sum(s.n for s in mix)
        be = Engine(mix, use_graph=True)
        best, best_ms = be.autotune(steps_per_trial=3)
        be.close()
        be = Engine(mix, dem_block=best[0], sort_ipt=best[1], sort_radix_bits=best[2], use_graph=True,
                    bucket_split=best[3])
        be.step(3)
        be.sync()
        t0 = time.perf_counter()
        be.step(args.batched_steps)
        be.sync()
        bsec = time.perf_counter() - t0
        blat = []
        for _ in range(20):
            s0 = time.perf_counter()
            be.step(1)
            be.sync()
            blat.append((time.perf_counter() - s0) * 1e3)
        be.close()
        plain = [Engine(s, use_graph=True) for s in mix]
        for e in plain:
            e.step(2)
        for e in plain:
            e.sync()
        t0 = time.perf_counter()
        for _ in range(args.batched_steps):
            for e in plain:
                e.step(1)
        for e in plain:
            e.sync()
        psec = time.perf_counter() - t0
        for e in plain:
            e.close()
        bv, pv = total * args.batched_steps / bsec, total * args.batched_steps / psec
        batched = {"scenes": len(mix), "particles": total, "steps": args.batched_steps,
                   "batched_updates_per_s": bv, "plain_loop_updates_per_s": pv, "speedup": bv / pv,
                   "batched_step_ms_p50": float(np.percentile(blat, 50)),
                   "batched_step_ms_p99": float(np.percentile(blat, 99)),
                   "tuning": {"dem_block": best[0], "sort_items_per_thread": best[1], "sort_radix_bits": best[2],
                              "bucket_split": best[3]}}
        log(f"batched: {bv:.3e} vs plain loop {pv:.3e} particle-updates/s ({bv / pv:.1f}x)")

    # ---- served form: continuous batching vs one scene at a time on the same request stream
    serving = None
    if not args.no_serving:
        from diploma_thesis_amd import serve_bench
        log(f"serving: {args.serve_requests} requests at {args.serve_rate}/s, batched and one at a time")
        sv = serve_bench.compare(args.serve_requests, args.serve_rate)
        b, o = sv["batched"], sv["one_at_a_time"]
        serving = {"workload": sv["workload"], "throughput_ratio": sv["throughput_ratio"],
                   "batched_updates_per_s": b["updatesPerSecond"], "one_at_a_time_updates_per_s": o["updatesPerSecond"],
                   "step_latency_ms_p50": b["stepLatencyP50Ms"], "step_latency_ms_p99": b["stepLatencyP99Ms"],
                   "completion_latency_ms_p50": b["completionLatencyP50Ms"],
                   "completion_latency_ms_p99": b["completionLatencyP99Ms"],
                   "one_at_a_time_completion_latency_ms_p50": o["completionLatencyP50Ms"],
                   "one_at_a_time_completion_latency_ms_p99": o["completionLatencyP99Ms"],
                   "queue_wait_ms_p50": b["queueWaitP50Ms"], "max_resident": b["maxResidentSeen"],
                   "completed": b["completed"], "rejected_for_budget": b["rejectedForBudget"]}
        log(f"serving: {b['updatesPerSecond']:.3e} vs {o['updatesPerSecond']:.3e} particle-updates/s "
            f"({sv['throughput_ratio']:.2f}x)")

    line = {
        "metric": "particle-updates/s batched vs plain on 1 GPU; p50/p99 per-scene step latency",
        "value": value, "unit": "particle-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain single large scene: 2^20 particles (8,192 rigid 4x4x4 cubes + 524,288 "
                               "grains) in [0,512]^3, r=0.5, k=5e4, c=50, shear=10, dt=1e-4, 5 walls",
                   "method": args.method, "grid": "256^3 hashed" if method == capi.METHOD_GRID else None,
                   "particles": n, "bodies": sc.b, "tuning": tuning},
        "step_latency_ms_p50": float(np.percentile(lat, 50)), "step_latency_ms_p99": float(np.percentile(lat, 99)),
        "stage_ms": stage_summary,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batched": batched,
        "serving": serving,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
