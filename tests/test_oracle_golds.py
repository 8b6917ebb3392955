"""The CPU oracle against the reference's own MPFA-O regression golds
(pressure within 1e-12 relative, as demo/*/ *.cfg [default-test-criteria])."""
import numpy as np
import pytest

from cases import GOLDS, gold_problem
from tdycore_amd import hdf5io
from pyoracle import Oracle

RTOL = 1e-12


def _check(U, exp, rtol=RTOL):
    U = np.asarray(U)
    n = U.shape[0]
    def close(a, b):
        return abs(a - b) <= rtol * abs(b)
    assert close(U.max(), exp["Max"]), (U.max(), exp["Max"])
    assert close(U.min(), exp["Min"]), (U.min(), exp["Min"])
    assert close(U.sum() / n, exp["Mean"]), (U.sum() / n, exp["Mean"])
    for k, v in exp["samples"].items():
        assert close(U[int(k)], v), (k, U[int(k)], v)


@pytest.mark.parametrize("name", sorted(GOLDS["cases"]))
def test_oracle_reproduces_reference_gold(name, tmp_path):
    case = GOLDS["cases"][name]
    P = gold_problem(case)
    o = Oracle(P["mesh"], P["face_bc"], P["K"], P["porosity"], P["Sr"], **P["kw"])
    if case["form"] == "ts":
        U, st = o.run_ts_be(P["U0"], case["dt"], case["nsteps"])
    elif case["form"] == "snes":
        U, st = o.run_snes_steps(P["U0"], case["dt"], case["nsteps"], rtol=case.get("snes_rtol", 1e-8))
    else:
        U = P["U0"]
        for i in range(1 + case.get("restart", 0)):
            if i:
                # restart through the reference's checkpoint file (fields/IC)
                path = str(tmp_path / hdf5io.checkpoint_filename(case["final_time"]))
                hdf5io.write_checkpoint(path, U, int(st[2]))
                U = hdf5io.read_dataset(path, "fields/IC")
            U, st = o.run_tdyti(U, case["dt_init"], case["growth"], case["dt_max"], case["final_time"])
            assert st[3] == 0
    assert st[3] == 0
    _check(U, case["expected"], case.get("rtol", RTOL))


def test_gold_sampling_matches_reference_format():
    from pyoracle import regression_text
    U = np.arange(27, dtype=float) + 1e5
    txt = regression_text(U, 2)
    assert "       13:" in txt and "        0:" in txt
