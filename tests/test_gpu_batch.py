"""The fused serving batch (tdyhip_batch_*): requests of different sizes,
seeds and solver settings advanced together, with their state, residual and
Jacobian evaluations in one segmented launch each.  Every request's final
pressure and solver statistics are the same bit for bit
  * through MPFAO.run_tdyti alone (the single-request solver),
  * alone in a batch,
  * in a mix, with a request joining after two rounds.
A cancelled request leaves the batch with ERR_CANCELLED and the others are
unaffected."""
import numpy as np
import pytest

from tdycore_amd import workloads
from tdycore_amd.meshgen import brick_partition

pytestmark = pytest.mark.gpu

HOUR = 3600.0
# (nx, ny, nz, seed, sigma, steps, krylov_fp32, pc_coarse)
SPECS = [(12, 10, 8, 101, 1.0, 3, 0, 0), (16, 16, 8, 202, 1.4, 2, 1, 1), (9, 14, 6, 303, 0.7, 3, 0, 1)]


def _requests():
    import torch
    from tdycore_amd.mpfao import MPFAO
    assert torch.cuda.is_available()
    out = []
    for nx, ny, nz, seed, sigma, steps, fp32, coarse in SPECS:
        P = workloads.ensemble_box_problem(nx, ny, nz, seed, sigma)
        g = MPFAO(P["mesh"], P["face_bc"], P["K"], P["porosity"], P["Sr"], **P["kw"])
        g.set_pc_partition(brick_partition(P["mesh"], 4, 4, 4))
        o = g.solver_options(snes_rtol=1e-10, krylov_fp32=fp32, pc_block=64, pc_coarse=coarse)
        out.append(dict(g=g, P=P, o=o, final=steps * HOUR))
    return out


def _alone(r):
    import torch
    U = torch.tensor(r["P"]["U0"], device="cuda", dtype=torch.float64)
    st = r["g"].run_tdyti(U, HOUR, 1.25, HOUR, r["final"], r["o"])
    torch.cuda.synchronize()
    return U.cpu().numpy(), st


def _batched(reqs, late=()):
    """Run reqs in one batch; the indices in `late` join after two rounds."""
    import torch
    from tdycore_amd.batch import FusedBatch
    b = FusedBatch()
    U = {}
    ids = {}

    def add(i):
        r = reqs[i]
        U[i] = torch.tensor(r["P"]["U0"], device="cuda", dtype=torch.float64)
        torch.cuda.synchronize()
        ids[b.add(r["g"], U[i], r["o"], HOUR, 1.25, HOUR, r["final"])] = i

    for i in range(len(reqs)):
        if i not in late:
            add(i)
    out, rounds = {}, 0
    while len(b) or len(out) + len(b) < len(reqs):
        if rounds == 2:
            for i in late:
                add(i)
        for rid, rc, st in b.round():
            assert rc == 0, (rid, rc)
            out[ids[rid]] = (U[ids[rid]].cpu().numpy(), st)
        rounds += 1
        assert rounds < 2000
    b.close()
    return out, rounds


@pytest.fixture(scope="module")
def reqs():
    rs = _requests()
    yield rs
    for r in rs:
        r["g"].close()


KEYS = ("newton_its", "linear_its", "steps", "cuts", "status", "krylov_fallbacks")


def test_batch_alone_equals_run_tdyti_equals_mix(reqs):
    print("[case] fused batch: alone / run_tdyti / mix", flush=True)
    ref = [_alone(r) for r in reqs]
    for i, (U, st) in enumerate(ref):
        assert st["status"] == 0 and st["steps"] >= 2, st
        assert np.max(np.abs(U - reqs[i]["P"]["U0"])) > 0.0  # the steps moved the state
    for i in range(len(reqs)):
        one, _ = _batched([reqs[i]])
        assert np.array_equal(one[0][0], ref[i][0]), f"request {i} alone in a batch"
        assert all(one[0][1][k] == ref[i][1][k] for k in KEYS), (one[0][1], ref[i][1])
    mix, rounds = _batched(reqs, late=(2,))
    print(f"  mix of {len(reqs)} in {rounds} rounds; Newton {[m[1]['newton_its'] for m in mix.values()]}", flush=True)
    for i in range(len(reqs)):
        assert np.array_equal(mix[i][0], ref[i][0]), f"request {i} in the mix"
        assert all(mix[i][1][k] == ref[i][1][k] for k in KEYS), (mix[i][1], ref[i][1])


def test_batch_cancel_leaves_others_untouched(reqs):
    import torch
    from tdycore_amd import _lib
    from tdycore_amd.batch import FusedBatch
    print("[case] fused batch: cancel one request", flush=True)
    ref = _alone(reqs[0])
    b = FusedBatch()
    U0 = torch.tensor(reqs[0]["P"]["U0"], device="cuda", dtype=torch.float64)
    U1 = torch.tensor(reqs[1]["P"]["U0"], device="cuda", dtype=torch.float64)
    torch.cuda.synchronize()
    i0 = b.add(reqs[0]["g"], U0, reqs[0]["o"], HOUR, 1.25, HOUR, reqs[0]["final"])
    i1 = b.add(reqs[1]["g"], U1, reqs[1]["o"], HOUR, 1.25, HOUR, reqs[1]["final"])
    b.round()
    reqs[1]["g"].cancel(True)
    res = {}
    while len(b):
        for rid, rc, st in b.round():
            res[rid] = (rc, st)
    reqs[1]["g"].cancel(False)
    b.close()
    assert res[i1][0] == _lib.ERR_CANCELLED
    assert res[i0][0] == 0
    assert np.array_equal(U0.cpu().numpy(), ref[0])


def test_fuser_threads_equal_alone(reqs):
    """Requests solving in their own threads (run_tdyti) with an evaluation
    fuser attached give the same pressures and statistics bit for bit as
    without it, and evaluations that were pending together shared a launch."""
    import threading
    import torch
    from tdycore_amd.batch import Fuser
    print("[case] evaluation fuser, 4 threads", flush=True)
    extra = _requests()[1]  # a second context of request 1's problem: four solves in flight
    rs = list(reqs) + [extra]
    ref = [_alone(r) for r in rs]
    fz = Fuser(window_us=2000)
    for r in rs:
        fz.attach(r["g"])
    out = [None] * len(rs)
    go = threading.Barrier(len(rs))

    def run(i):
        r = rs[i]
        with torch.cuda.stream(r["g"].stream):
            U = torch.tensor(r["P"]["U0"], device="cuda", dtype=torch.float64)
            torch.cuda.current_stream().synchronize()
            go.wait()
            st = r["g"].run_tdyti(U, HOUR, 1.25, HOUR, r["final"], r["o"])
            out[i] = (U.cpu().numpy(), st)

    th = [threading.Thread(target=run, args=(i,)) for i in range(len(rs))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for r in rs:
        fz.detach(r["g"])
    stats = fz.stats()
    fz.close()
    extra["g"].close()
    print(f"  fuser: {stats}", flush=True)
    for i in range(len(rs)):
        assert out[i] is not None
        assert np.array_equal(out[i][0], ref[i][0]), f"request {i} through the fuser"
        assert all(out[i][1][k] == ref[i][1][k] for k in KEYS), (out[i][1], ref[i][1])
    assert stats["evaluations"] >= sum(o[1]["newton_its"] for o in out)
    assert stats["largest_group"] >= 2
