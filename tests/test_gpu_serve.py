"""Serving (continuous batching) checked request by request.

- Every served request's final state equals the same scene stepped alone on the plain path, bit
  for bit (FNV-1a checksums of the state bytes), in the batched loop with slot-limited churn and
  in the one-at-a-time loop, for the grid and the LBVH forest.
- A request whose memory estimate exceeds the whole HBM budget is refused and counted; the rest of
  the stream is served and the run ends.
"""
import ctypes as C

import numpy as np
import pytest

from diploma_thesis_amd import Engine, capi, scenes

pytestmark = pytest.mark.gpu


def _cfg(n, seed, lo, hi, steps, budget, slots, one, method, rate=400.0):
    return capi.ArSimServeConfig(seed=seed, numRequests=n, arrivalRate=rate, minParticles=lo, maxParticles=hi,
                                 minSteps=steps[0], maxSteps=steps[1], hbmBudgetBytes=budget, maxResident=slots,
                                 oneAtATime=one, timeScale=1.0, method=method)


def _request(cfg, i):
    n, st, seed, t = C.c_int(), C.c_int(), C.c_uint64(), C.c_double()
    capi.check(capi.lib.arsim_serve_request(C.byref(cfg), i, C.byref(n), C.byref(st), C.byref(seed), C.byref(t)))
    return n.value, st.value, seed.value


def _serve(cfg):
    sums = np.zeros(cfg.numRequests, np.uint64)
    cfg.resultChecksums = capi.ptr(sums)
    rep = capi.ArSimServeReport()
    capi.check(capi.lib.arsim_serve_run(C.byref(cfg), C.byref(rep)), "serve_run")
    return rep, sums


@pytest.mark.parametrize("method", [capi.METHOD_GRID, capi.METHOD_LBVH30])
@pytest.mark.parametrize("one", [0, 1])
def test_served_results_equal_alone(method, one):
    print(f"serve: served == alone method={method} oneAtATime={one}", flush=True)
    # three slots for ten overlapping requests: scenes are admitted, carried over and retired
    # while others run
    cfg = _cfg(10, 3, 2048, 12000, (3, 25), 8e9, 3, one, method)
    rep, sums = _serve(cfg)
    assert rep.completed == 10 and rep.rejectedForBudget == 0
    if not one:
        assert rep.maxResidentSeen == 3
    for i in range(10):
        n, steps, seed = _request(cfg, i)
        with Engine(scenes.mix_scene(n, seed, method)) as e:
            e.step(steps)
            e.sync()
            alone = e.checksum()
        assert int(sums[i]) == alone, f"request {i} (n={n}, steps={steps}) differs from the scene stepped alone"


def test_serve_refuses_requests_over_budget():
    print("serve: budget refusal", flush=True)
    # the estimate is about 206 B per particle plus 2 MiB of cell table: a 6 MB budget takes scenes
    # of up to ~18,000 particles and refuses the larger ones
    cfg = _cfg(12, 5, 2048, 60000, (2, 6), 6e6, 8, 0, capi.METHOD_GRID)
    rep, sums = _serve(cfg)
    assert rep.completed + rep.rejectedForBudget == 12
    assert rep.completed > 0 and rep.rejectedForBudget > 0
    sizes = np.array([_request(cfg, i)[0] for i in range(12)])
    served = sums != 0
    assert served.sum() == rep.completed
    assert sizes[served].max() < sizes[~served].min()
    # nothing fits: every request is refused and the run still ends
    rep, sums = _serve(_cfg(4, 5, 2048, 4096, (2, 3), 1.0, 8, 0, capi.METHOD_GRID))
    assert rep.completed == 0 and rep.rejectedForBudget == 4 and not sums.any()
