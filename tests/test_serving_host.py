"""Host logic of the ensemble server: admission by HBM footprint, slot
limits and latency accounting, with a stand-in solve (no GPU)."""
import threading
import time

import numpy as np
import pytest

from tdycore_amd import serving, workloads


def test_footprint_bounds_grow_with_size():
    small = serving.footprint_bytes(2048, 6656, 2601, 1536, serving.block_rows(16, 16, 8, (8, 8, 4)))
    big = serving.footprint_bytes(442368, 1347840, 456533, 64512, serving.block_rows(96, 96, 48, (8, 8, 4)))
    assert 0 < small < big
    assert big / 442368 < 4000  # a few KB per cell


def test_request_specs_match_the_ensemble_ranges():
    specs = workloads.ensemble_request_specs(512)
    assert len(specs) == 512
    nx = np.array([s["nx"] for s in specs]); nz = np.array([s["nz"] for s in specs])
    sig = np.array([s["sigma"] for s in specs])
    assert nx.min() >= 16 and nx.max() <= 96 and nz.min() >= 8 and nz.max() <= 48
    assert sig.min() >= 0.5 and sig.max() <= 1.5
    assert len({s["seed"] for s in specs}) == 512


def test_admission_respects_budget_and_slots():
    lock = threading.Lock()
    state = {"held": 0, "peak": 0, "running": 0, "peak_running": 0}
    sizes = {}

    def fake_solve(P, req, handle):
        need = sizes[req.rid]
        with lock:
            state["held"] += need
            state["running"] += 1
            state["peak"] = max(state["peak"], state["held"])
            state["peak_running"] = max(state["peak_running"], state["running"])
        time.sleep(0.02)
        with lock:
            state["held"] -= need
            state["running"] -= 1
        return P["U0"].copy(), {"newton_its": 0}

    reqs = [serving.Request(rid=i, nx=16 + 4 * (i % 3), ny=16, nz=8, seed=i, sigma=1.0) for i in range(12)]
    for r in reqs:
        P = workloads.ensemble_box_problem(r.nx, r.ny, r.nz, r.seed, r.sigma)
        m = P["mesh"]
        sizes[r.rid] = serving.footprint_bytes(m.nc, m.nf, m.nv, m.nbnd, serving.block_rows(r.nx, r.ny, r.nz, (8, 8, 4)))
    budget = int(2.5 * max(sizes.values()))
    srv = serving.EnsembleServer(slots=4, hbm_budget=budget, build_workers=4, solve_fn=fake_solve)
    futs = [srv.submit(r) for r in reqs]
    res = [f.result(timeout=60) for f in futs]
    srv.close()
    assert sorted(r["rid"] for r in res) == list(range(12))
    assert state["peak"] <= budget          # never more than the budget on the device
    assert state["peak_running"] <= 4       # never more than the slots
    assert srv.peak_inflight >= 2           # and requests did overlap
    for r in res:
        assert r["latency_s"] >= r["solve_s"] > 0
    s = serving.latency_summary(res)
    assert s["requests"] == 12 and s["p99_latency_s"] >= s["p50_latency_s"]


def test_request_above_budget_is_rejected():
    srv = serving.EnsembleServer(slots=1, hbm_budget=1 << 20, build_workers=1,
                                 solve_fn=lambda P, r, h: (P["U0"], {}))
    f = srv.submit(serving.Request(rid=0, nx=16, ny=16, nz=8, seed=1, sigma=1.0))
    with pytest.raises(MemoryError):
        f.result(timeout=60)
    srv.close()


def test_drain_cancels_requests_not_yet_on_the_device():
    def slow(P, r, h):
        time.sleep(0.2)
        return P["U0"], {"status": 0}
    srv = serving.EnsembleServer(slots=2, hbm_budget=1 << 40, build_workers=2, solve_fn=slow)
    futs = [srv.submit(serving.Request(rid=i, nx=16, ny=16, nz=8, seed=i, sigma=1.0)) for i in range(16)]
    time.sleep(0.6)
    srv.drain()
    srv.close()
    done = [f for f in futs if f.done() and not f.cancelled()]
    assert all(f.done() for f in futs)
    assert 0 < len(done) < 16 and all(f.result()["stats"]["status"] == 0 for f in done)


def _cancellable(run_s):
    """A stand-in solve that runs run_s seconds unless its handle stops it."""
    def solve(P, r, handle):
        stop = threading.Event()
        handle.attach(stop.set)
        try:
            if stop.wait(run_s):
                raise RuntimeError("tdyhip error 6: solve cancelled")
            return P["U0"], {"status": 0}
        finally:
            handle.attach(None)
    return solve


def test_deadline_cancels_long_solves_only():
    def solve(P, r, handle):
        return _cancellable(5.0 if r.rid % 2 else 0.05)(P, r, handle)
    srv = serving.EnsembleServer(slots=4, hbm_budget=1 << 40, build_workers=2, solve_fn=solve, deadline_s=0.3)
    futs = [srv.submit(serving.Request(rid=i, nx=16, ny=16, nz=8, seed=i, sigma=1.0)) for i in range(6)]
    t0 = time.perf_counter()
    for i, f in enumerate(futs):
        if i % 2:
            with pytest.raises(serving.DeadlineExceeded) as ei:
                f.result(timeout=30)
            assert ei.value.rid == i and ei.value.latency_s >= 0.3
        else:
            assert f.result(timeout=30)["stats"]["status"] == 0
    assert time.perf_counter() - t0 < 4.0  # nobody waited for a 5 s solve
    srv.close()


def test_drain_can_cancel_running_solves():
    srv = serving.EnsembleServer(slots=2, hbm_budget=1 << 40, build_workers=2, solve_fn=_cancellable(30.0))
    futs = [srv.submit(serving.Request(rid=i, nx=16, ny=16, nz=8, seed=i, sigma=1.0)) for i in range(5)]
    time.sleep(0.5)
    t0 = time.perf_counter()
    srv.drain(cancel_running=True)
    srv.close()
    assert time.perf_counter() - t0 < 5.0
    running = [f for f in futs if not f.cancelled()]
    assert all(f.done() for f in futs) and len(running) == 2
    for f in running:
        with pytest.raises(serving.DeadlineExceeded):
            f.result()
