"""The allocation-free residual (tdyhip_ifunction_noalloc) against
tdyhip_ifunction, through the C ABI: the same kernels, so the same bits.

Two calls in a row share the library's per-GPU scratch array: a larger mesh
first, then a smaller one (whose rows must not see the first call's leftovers),
on hexahedra and on prisms (the general path, whose unwritten rows must read
zero).  A mesh beyond the array's 8,388,608 flux rows is refused before any
device work."""
import ctypes

import numpy as np
import pytest

from cases import cfg0_problem
from tdycore_amd import workloads

pytestmark = pytest.mark.gpu


def _pair_call(P, seed):
    import torch
    from tdycore_amd import _lib
    from tdycore_amd.mpfao import MPFAO
    g = MPFAO(P["mesh"], P["face_bc"], P["K"], P["porosity"], P["Sr"], **P["kw"])
    nc = P["mesh"].nc
    rng = np.random.default_rng(seed)
    U = torch.tensor(P["U0"] + 2e4 * (rng.random(nc) - 0.5), device="cuda")
    Ud = torch.tensor(1e-2 * (rng.random(nc) - 0.5), device="cuda")
    R1 = torch.full((nc,), np.nan, dtype=torch.float64, device="cuda")
    R2 = torch.full((nc,), np.nan, dtype=torch.float64, device="cuda")
    g.IFunction(0.0, U, Ud, R1)
    L = _lib.load()
    torch.cuda.synchronize()
    # the C ABI directly, on the context's own stream (NULL)
    _lib.check(L.tdyhip_ifunction_noalloc(g.handle, ctypes.c_double(0.0), ctypes.c_void_p(U.data_ptr()),
                                          ctypes.c_void_p(Ud.data_ptr()), ctypes.c_void_p(R2.data_ptr()), None))
    s = ctypes.c_void_p()
    _lib.check(L.tdyhip_get_stream(g.handle, ctypes.byref(s)))
    torch.cuda.ExternalStream(s.value).synchronize()
    a, b = R1.cpu().numpy(), R2.cpu().numpy()
    g.close()
    return a, b


@pytest.mark.parametrize("kind", ["hex", "prism"])
def test_noalloc_residual_equals_ifunction(kind):
    print(f"[case] allocation-free residual {kind}", flush=True)
    if kind == "hex":
        big, small = cfg0_problem(n=24), cfg0_problem(n=8)
    else:
        big = workloads.prism_problem(n_tri=600, nlayers=5, seed=21, length=800.0, lattice=16)
        small = workloads.prism_problem(n_tri=100, nlayers=2, seed=22, length=300.0, lattice=8)
    for P, seed in ((big, 1), (small, 2)):
        a, b = _pair_call(P, seed)
        assert np.all(np.isfinite(a))
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_noalloc_refuses_a_mesh_beyond_its_scratch():
    import torch
    from tdycore_amd import _lib
    from tdycore_amd.mpfao import MPFAO
    print("[case] allocation-free residual size limit", flush=True)
    P = workloads.layered_box_problem(128, 128, 64)  # 3,178,496 faces: 12.7M flux rows
    g = MPFAO(P["mesh"], P["face_bc"], P["K"], P["porosity"], P["Sr"], **P["kw"])
    U = torch.tensor(P["U0"], device="cuda")
    R = torch.empty_like(U)
    with pytest.raises(_lib.TdyHipError) as e:
        g.IFunctionNoAlloc(0.0, U, torch.zeros_like(U), R)
    assert e.value.code == 1 and b"flux rows" in _lib.load().tdyhip_last_error()
    g.close()
