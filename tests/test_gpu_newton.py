"""Newton + GMRES/block-Jacobi ILU(0) on the device against the reference's
regression golds (pressure 1e-12 relative) and against the oracle's Newton."""
import numpy as np
import pytest

from cases import GOLDS, gold_problem, cfg0_problem
from test_oracle_golds import _check

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", sorted(GOLDS["cases"]))
def test_device_newton_reproduces_reference_gold(name, tmp_path):
    print(f"[case] device Newton gold {name}", flush=True)
    import torch
    from tdycore_amd.mpfao import MPFAO
    case = GOLDS["cases"][name]
    P = gold_problem(case)
    g = MPFAO(P["mesh"], P["face_bc"], P["K"], P["porosity"], P["Sr"], **P["kw"])
    # the golds were made with direct solves (-pc_type lu, or ILU(0) of a dense
    # 8x8 Jacobian): solve the Newton systems to 1e-14 here
    opts = g.solver_options(ksp_rtol=1e-14, ksp_restart=60, snes_rtol=case.get("snes_rtol", 1e-8))
    if (P["mesh"].cell_verts < 0).any():
        # prisms: blocks of two cells keep the in-block ILU rows within their limit
        g.set_pc_partition(np.arange(P["mesh"].nc) // 2)
    U = torch.tensor(P["U0"], device="cuda")
    if case["form"] == "ts":
        st = g.run_ts_be(U, case["dt"], case["nsteps"], opts)
    elif case["form"] == "snes":
        st = g.run_snes_steps(U, case["dt"], case["nsteps"], opts)
    else:
        for i in range(1 + case.get("restart", 0)):
            if i:
                # restart through the reference's checkpoint file (fields/IC)
                from tdycore_amd import hdf5io
                path = str(tmp_path / hdf5io.checkpoint_filename(case["final_time"]))
                hdf5io.write_checkpoint(path, U.cpu().numpy(), st["steps"])
                U = torch.tensor(hdf5io.read_dataset(path, "fields/IC"), device="cuda")
            st = g.run_tdyti(U, case["dt_init"], case["growth"], case["dt_max"], case["final_time"], opts)
            assert st["status"] == 0
    assert st["status"] == 0
    rtol = case.get("rtol", 1e-12)
    _check(U.cpu().numpy(), case["expected"], rtol)
    # the same numbers through the product's regression writer (the reference's file format)
    from tdycore_amd.regression import parse_regression, regression_text
    got = parse_regression(regression_text(U, case["nsample"]))
    exp = case["expected"]
    for key in ("Max", "Min", "Mean"):
        assert abs(got[key] - exp[key]) <= rtol * abs(exp[key]), key
    assert set(got["samples"]) == {int(k) for k in exp["samples"]}


@pytest.mark.parametrize("fp32,pc_block,coarse", [(0, 256, 0), (1, 256, 0), (0, 512, 0), (1, 512, 0), (1, 64, 0),
                                                  (0, 64, 1), (1, 64, 1), (0, 256, 1)])
def test_device_newton_matches_oracle_cfg0(fp32, pc_block, coarse):
    """Plain-method configuration on a reduced mesh: device Newton (FP64 or
    FP32 Krylov) reaches the oracle's solution to 1e-8 relative in pressure."""
    print(f"[case] cfg0 Newton fp32={fp32} pc_block={pc_block} coarse={coarse}", flush=True)
    import torch
    from pyoracle import Oracle
    from tdycore_amd.mpfao import MPFAO
    P = cfg0_problem(n=10)
    o = Oracle(P["mesh"], P["face_bc"], P["K"], P["porosity"], P["Sr"], **P["kw"])
    Uo, st_o = o.run_snes_steps(P["U0"], 3600.0, 3, lin="lu", rtol=1e-10)
    assert st_o[3] == 0
    g = MPFAO(P["mesh"], P["face_bc"], P["K"], P["porosity"], P["Sr"], **P["kw"])
    U = torch.tensor(P["U0"], device="cuda")
    st = g.run_snes_steps(U, 3600.0, 3, g.solver_options(snes_rtol=1e-10, krylov_fp32=fp32, ksp_rtol=1e-8,
                                                         pc_block=pc_block, pc_coarse=coarse))
    assert st["status"] == 0
    err = np.abs(U.cpu().numpy() - Uo).max() / np.abs(Uo).max()
    print("max rel pressure err", err, st)
    assert err < 1e-8


def test_device_newton_is_deterministic():
    """Residual gather, Jacobian gather and every Krylov reduction run in a
    fixed order: two solves from the same state agree bit for bit."""
    print("[case] determinism cfg0 n=12", flush=True)
    import torch
    from tdycore_amd.mpfao import MPFAO
    P = cfg0_problem(n=12)
    g = MPFAO(P["mesh"], P["face_bc"], P["K"], P["porosity"], P["Sr"], **P["kw"])
    for fp32, coarse in ((0, 0), (1, 0), (0, 1)):
        outs = []
        # graph replays (first two runs: capture, then replay across steps and
        # Newton iterations with eager work between) against direct launches
        for graphs in (1, 1, 0):
            U = torch.tensor(P["U0"], device="cuda")
            st = g.run_snes_steps(U, 3600.0, 2, g.solver_options(snes_rtol=1e-10, krylov_fp32=fp32,
                                                                 hip_graphs=graphs, pc_block=64,
                                                                 pc_coarse=coarse))
            assert st["status"] == 0
            outs.append(U.cpu().numpy())
        assert np.array_equal(outs[0], outs[1]), fp32
        assert np.array_equal(outs[0], outs[2]), fp32


def test_newton_iteration_graph_matches_direct_launches():
    """Each Newton iteration is one replayed graph (Jacobian, preconditioner,
    GMRES cycles, the line search's full step and its test) with dt and the
    TS shift read from device memory.  Replays must equal direct launches bit
    for bit when dt changes from step to step (TDyTimeIntegratorRunToTime's
    growth) and in the TS form, also where a solve needs more cycles than the
    graph holds and where a full step is backtracked on the host."""
    print("[case] Newton iteration graphs cfg0 n=10", flush=True)
    import torch
    from tdycore_amd.mpfao import MPFAO
    P = cfg0_problem(n=10)
    g = MPFAO(P["mesh"], P["face_bc"], P["K"], P["porosity"], P["Sr"], **P["kw"])
    for fp32 in (0, 1):
        outs, stats = [], []
        for graphs in (1, 0, 1):
            U = torch.tensor(P["U0"], device="cuda")
            o = g.solver_options(snes_rtol=1e-10, krylov_fp32=fp32, hip_graphs=graphs, pc_block=64, pc_coarse=1)
            st = g.run_tdyti(U, 600.0, 1.5, 7200.0, 4 * 3600.0, o)
            assert st["status"] == 0 and st["steps"] >= 4
            U2 = U.clone()
            st2 = g.run_ts_be(U2, 1800.0, 2, o)
            assert st2["status"] == 0
            outs.append((U.cpu().numpy(), U2.cpu().numpy()))
            stats.append((st["newton_its"], st["linear_its"], st2["newton_its"], st2["linear_its"]))
        for k in (1, 2):
            assert np.array_equal(outs[0][0], outs[k][0]) and np.array_equal(outs[0][1], outs[k][1]), (fp32, k)
            assert stats[0] == stats[k]
    g.close()
