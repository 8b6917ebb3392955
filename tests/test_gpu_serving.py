"""Ensemble serving on the device: each result is the same alone and in the
mix (1e-8 relative), matches the oracle, and the admission footprint bounds
what the library actually holds."""
import numpy as np
import pytest

from tdycore_amd import serving, workloads

pytestmark = pytest.mark.gpu

REQS = [serving.Request(rid=i, nx=nx, ny=ny, nz=nz, seed=7000 + 17 * i, sigma=s, steps=2)
        for i, (nx, ny, nz, s) in enumerate([(16, 16, 8, 0.5), (24, 16, 8, 0.6), (16, 20, 12, 1.0),
                                             (20, 24, 10, 0.8), (17, 19, 9, 1.2)])]


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


def test_alone_equals_mix():
    alone = {}
    srv = serving.EnsembleServer(slots=1, build_workers=1)
    for r in REQS:
        alone[r.rid] = srv.submit(r).result(timeout=120)
    srv.close()
    srv = serving.EnsembleServer(slots=4, build_workers=4)
    futs = [srv.submit(r) for r in REQS]
    mix = {f.result(timeout=120)["rid"]: f.result() for f in futs}
    srv.close()
    assert srv.peak_inflight >= 2
    for r in REQS:
        a, m = alone[r.rid], mix[r.rid]
        assert a["stats"]["status"] == 0 and m["stats"]["status"] == 0
        assert _rel(m["U"], a["U"]) <= 1e-8, r
        # every kernel reduces in a fixed order: the mix is bit-for-bit the same
        assert np.array_equal(m["U"], a["U"]), r
        P = workloads.ensemble_box_problem(r.nx, r.ny, r.nz, r.seed, r.sigma)
        mesh = P["mesh"]
        est = serving.footprint_bytes(mesh.nc, mesh.nf, mesh.nv, mesh.nbnd,
                                      serving.block_rows(r.nx, r.ny, r.nz, (8, 8, 4)))
        assert a["stats"]["device_bytes"] <= est, (a["stats"]["device_bytes"], est)


def test_served_result_matches_oracle():
    from pyoracle import Oracle
    r = REQS[0]
    srv = serving.EnsembleServer(slots=2, build_workers=2)
    res = srv.submit(r).result(timeout=120)
    srv.close()
    P = workloads.ensemble_box_problem(r.nx, r.ny, r.nz, r.seed, r.sigma)
    o = Oracle(P["mesh"], P["face_bc"], P["K"], P["porosity"], P["Sr"], **P["kw"])
    U, st = o.run_tdyti(P["U0"], r.dt, 1.25, r.dt, r.steps * r.dt, lin="gmres", rtol=1e-10)
    assert st[3] == 0
    assert _rel(res["U"], U) <= 1e-8


def test_live_contexts_have_distinct_streams():
    """Every context owns its stream (the solver captures graphs on it): 40
    contexts alive together, more than torch's pool of 32 streams, never share."""
    print("[case] 40 contexts, distinct streams", flush=True)
    from tdycore_amd.mpfao import MPFAO
    P = workloads.ensemble_box_problem(4, 4, 4, 11, 1.0)
    gs = [MPFAO(P["mesh"], P["face_bc"], P["K"], P["porosity"], P["Sr"], **P["kw"]) for _ in range(40)]
    handles = {g.stream.cuda_stream for g in gs}
    for g in gs:
        g.close()
    assert len(handles) == 40


def test_many_small_requests_on_many_slots():
    """More requests than torch's stream pool, 8 slots: each one equals its
    alone result bit for bit."""
    print("[case] 40 requests on 8 slots", flush=True)
    reqs = [serving.Request(rid=i, nx=12 + (i % 5), ny=12 + (i % 3), nz=8, seed=900 + i, sigma=0.6, steps=1)
            for i in range(40)]
    srv = serving.EnsembleServer(slots=8, build_workers=8)
    futs = [srv.submit(r) for r in reqs]
    mix = {f.result(timeout=120)["rid"]: f.result() for f in futs}
    srv.close()
    assert srv.peak_inflight >= 2
    srv = serving.EnsembleServer(slots=1, build_workers=1)
    for r in reqs[:6]:
        a = srv.submit(r).result(timeout=120)
        assert np.array_equal(a["U"], mix[r.rid]["U"]), r
    srv.close()
    assert all(m["stats"]["status"] == 0 for m in mix.values())


def test_cancel_stops_a_running_solve():
    """A request cancelled from another thread stops within a GMRES cycle or a
    Newton iteration and reports TDYHIP_ERR_CANCELLED."""
    import threading
    import time
    from tdycore_amd import _lib
    print("[case] cancel a 40x40x20, 200-step solve after 0.5 s", flush=True)
    r = serving.Request(rid=99, nx=40, ny=40, nz=20, seed=4242, sigma=1.0, steps=200)
    P = workloads.ensemble_box_problem(r.nx, r.ny, r.nz, r.seed, r.sigma)
    srv = serving.EnsembleServer(slots=1, build_workers=1)
    h, box = serving.SolveHandle(), {}

    def run():
        try:
            box["res"] = srv._gpu_solve(P, r, h)
        except _lib.TdyHipError as e:
            box["err"] = e
        box["t"] = time.perf_counter()

    th = threading.Thread(target=run)
    th.start()
    time.sleep(0.5)
    tc = time.perf_counter()
    h.cancel()
    th.join(120)
    srv.close()
    assert "err" in box and box["err"].code == _lib.ERR_CANCELLED
    assert box["t"] - tc < 2.0


def test_context_solves_again_after_a_cancel():
    """A cancel request set before the solve stops it at once; cleared, the
    same context solves to the bits of a fresh context."""
    import torch
    from tdycore_amd import _lib
    from tdycore_amd.meshgen import brick_partition
    from tdycore_amd.mpfao import MPFAO
    print("[case] cancel before start, then solve on the same context", flush=True)
    r = REQS[0]
    P = workloads.ensemble_box_problem(r.nx, r.ny, r.nz, r.seed, r.sigma)

    def ctx():
        g = MPFAO(P["mesh"], P["face_bc"], P["K"], P["porosity"], P["Sr"], **P["kw"])
        g.set_pc_partition(brick_partition(P["mesh"], 8, 8, 4))
        return g

    def solve(g):
        o = g.solver_options(snes_rtol=1e-10, krylov_fp32=0, pc_block=256, pc_coarse=1)
        U = torch.tensor(P["U0"], device="cuda", dtype=torch.float64)
        st = g.run_tdyti(U, r.dt, 1.25, r.dt, r.steps * r.dt, o)
        return U.cpu().numpy(), st

    g = ctx()
    g.cancel(True)
    with pytest.raises(_lib.TdyHipError) as ei:
        solve(g)
    assert ei.value.code == _lib.ERR_CANCELLED
    g.cancel(False)
    U1, st1 = solve(g)
    g.close()
    g2 = ctx()
    U2, st2 = solve(g2)
    g2.close()
    assert st1["status"] == 0 and st2["status"] == 0
    assert np.array_equal(U1, U2)


def test_fused_server_equals_unfused():
    """The fused engine (one segmented launch per evaluation over every
    request on the device) serves the same pressures bit for bit as the
    one-context-per-slot server, with several requests in the batch at once."""
    print("[case] fused server vs slots", flush=True)
    srv = serving.EnsembleServer(slots=4, build_workers=4)
    ref = {f.result(timeout=120)["rid"]: f.result() for f in [srv.submit(r) for r in REQS]}
    srv.close()
    srv = serving.EnsembleServer(slots=8, build_workers=4, fused=True)
    futs = [srv.submit(r) for r in REQS]
    got = {f.result(timeout=120)["rid"]: f.result() for f in futs}
    srv.close()
    assert srv.peak_inflight >= 2
    for r in REQS:
        assert got[r.rid]["stats"]["status"] == 0
        assert np.array_equal(got[r.rid]["U"], ref[r.rid]["U"]), r
        for k in ("newton_its", "linear_its", "steps"):
            assert got[r.rid]["stats"][k] == ref[r.rid]["stats"][k], (k, r)


def test_fused_server_deadline_cancels_in_the_batch():
    """A request past the deadline is cancelled inside the batch (its future
    fails with DeadlineExceeded, carrying its latency); a short one beside it
    is served."""
    print("[case] fused server deadline", flush=True)
    long_r = serving.Request(rid=100, nx=64, ny=64, nz=32, seed=9100, sigma=1.4, steps=400)
    short_r = REQS[0]
    srv = serving.EnsembleServer(slots=4, build_workers=2, fused=True, deadline_s=1.0)
    f_long, f_short = srv.submit(long_r), srv.submit(short_r)
    assert f_short.result(timeout=120)["stats"]["status"] == 0
    with pytest.raises(serving.DeadlineExceeded) as ei:
        f_long.result(timeout=120)
    srv.close()
    assert ei.value.rid == 100 and ei.value.latency_s >= 1.0
