"""Serving benchmark: a Poisson stream of scenes served with continuous batching, and the same stream
served one scene at a time on the plain path.

Default workload: 2,000 requests at 50 scenes/s, N log-uniform in [2,048, 262,144] particles (small-scene
generator), steps uniform in [50, 2,000], HBM budget 200 GB.  Prints one JSON line with p50/p99
per-step and completion latency and aggregate particle-updates/s for both modes.

    python -m diploma_thesis_amd.serve_bench [--requests 2000] [--rate 50] [--time-scale 1]
"""
import argparse
import ctypes as C
import json
import sys

from . import _capi as capi


def run(args, one_at_a_time):
    cfg = capi.ArSimServeConfig(seed=args.seed, numRequests=args.requests, arrivalRate=args.rate,
                                minParticles=args.min_particles, maxParticles=args.max_particles,
                                minSteps=args.min_steps, maxSteps=args.max_steps, hbmBudgetBytes=args.budget,
                                maxResident=args.max_resident, oneAtATime=1 if one_at_a_time else 0,
                                timeScale=args.time_scale, method=capi.METHOD_GRID)
    rep = capi.ArSimServeReport()
    capi.check(capi.lib.arsim_serve_run(C.byref(cfg), C.byref(rep)), "serve_run")
    return {k: getattr(rep, k) for k, _ in capi.ArSimServeReport._fields_}


def compare(requests, rate, time_scale=1.0, seed=4, min_particles=2048, max_particles=262144, min_steps=50,
            max_steps=2000, budget=200e9, max_resident=4096):
    """One request stream served with continuous batching and one scene at a time; both reports and the
    throughput ratio."""
    args = argparse.Namespace(requests=requests, rate=rate, time_scale=time_scale, seed=seed,
                              min_particles=min_particles, max_particles=max_particles, min_steps=min_steps,
                              max_steps=max_steps, budget=budget, max_resident=max_resident)
    out = {"workload": f"{requests} requests, Poisson {rate}/s (time scale {time_scale}), N log-U[{min_particles}, "
                       f"{max_particles}], steps U[{min_steps}, {max_steps}], budget {budget:.3g} B"}
    out["batched"] = run(args, False)
    out["one_at_a_time"] = run(args, True)
    out["throughput_ratio"] = out["batched"]["updatesPerSecond"] / out["one_at_a_time"]["updatesPerSecond"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--rate", type=float, default=50.0)
    ap.add_argument("--min-particles", type=int, default=2048)
    ap.add_argument("--max-particles", type=int, default=262144)
    ap.add_argument("--min-steps", type=int, default=50)
    ap.add_argument("--max-steps", type=int, default=2000)
    ap.add_argument("--budget", type=float, default=200e9)
    ap.add_argument("--max-resident", type=int, default=4096)
    ap.add_argument("--time-scale", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--skip-baseline", action="store_true")
    args = ap.parse_args()
    out = {"workload": f"{args.requests} requests, Poisson {args.rate}/s (time scale {args.time_scale}), "
                       f"N log-U[{args.min_particles}, {args.max_particles}], steps U[{args.min_steps}, "
                       f"{args.max_steps}], budget {args.budget:.3g} B"}
    print("[serve] continuous batching", file=sys.stderr, flush=True)
    out["batched"] = run(args, False)
    if not args.skip_baseline:
        print("[serve] one scene at a time", file=sys.stderr, flush=True)
        out["one_at_a_time"] = run(args, True)
        out["throughput_ratio"] = out["batched"]["updatesPerSecond"] / out["one_at_a_time"]["updatesPerSecond"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
