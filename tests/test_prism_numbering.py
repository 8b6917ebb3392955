"""Cell numbering of the prism workload: the reverse Cuthill-McKee groups of
columns give the same discrete problem as the reference's layer-major
numbering (the oracle's residual and Jacobian permute exactly), every
Jacobian column offset fits int16, and the block-Jacobi blocks are groups of
whole columns in layer order."""
import numpy as np
import scipy.sparse as sp

from tdycore_amd import workloads
from pyoracle import Oracle


def _pair():
    kw = dict(n_tri=400, nlayers=4, seed=21, length=600.0, lattice=16, infiltration=True)
    return workloads.prism_problem(column_major=False, **kw), workloads.prism_problem(column_major=True, **kw)


def _map(a, b):
    """new_of_old: the cell of mesh b at each cell's centroid of mesh a."""
    ka = np.lexsort(a.cell_cen.T[::-1])
    kb = np.lexsort(b.cell_cen.T[::-1])
    assert np.array_equal(a.cell_cen[ka], b.cell_cen[kb])
    m = np.empty(a.nc, dtype=np.int64)
    m[ka] = kb
    return m


def test_renumbered_problem_is_the_same_problem():
    A, B = _pair()
    m = _map(A["mesh"], B["mesh"])
    assert np.array_equal(np.sort(m), np.arange(A["mesh"].nc))
    for key in ("K", "porosity", "Sr", "U0"):
        assert np.array_equal(B[key][m], A[key]), key
    oa = Oracle(A["mesh"], A["face_bc"], A["K"], A["porosity"], A["Sr"], **A["kw"])
    ob = Oracle(B["mesh"], B["face_bc"], B["K"], B["porosity"], B["Sr"], **B["kw"])
    rng = np.random.default_rng(3)
    U = A["U0"] + 3e3 * rng.standard_normal(A["mesh"].nc)
    Ub = np.empty_like(U)
    Ub[m] = U
    out = []
    for o, u in ((oa, U), (ob, Ub)):
        o.set_dtime(1800.0)
        o.snes_presolve(u - 50.0)
        R = o.snes_function(u)
        J = o.jac_to_csr(o.snes_jacobian(u))
        out.append((R, J))
    (Ra, Ja), (Rb, Jb) = out
    np.testing.assert_allclose(Rb[m], Ra, rtol=1e-12, atol=1e-12 * np.abs(Ra).max())
    Jb_old = Jb[m][:, m]
    assert abs(Jb_old - Ja).max() <= 1e-12 * abs(Ja).max()


def test_column_offsets_fit_int16_and_blocks_are_columns():
    P = workloads.prism_problem(n_tri=20_000, nlayers=20, seed=5, infiltration=True, block_tris=3)
    mesh, part = P["mesh"], P["pc_part"]
    nc = mesh.nc
    rows = np.repeat(np.arange(nc), 6)
    inc = sp.csr_matrix((np.ones(rows.size), (rows, mesh.cell_verts[:, :6].ravel())), shape=(nc, mesh.nv))
    A = (inc @ inc.T).tocoo()
    assert np.abs(A.row.astype(np.int64) - A.col).max() <= 32767
    # a block: 3 whole columns (60 cells), numbered layer by layer
    sizes = np.bincount(part)
    assert sizes.max() == 60 and np.all(np.diff(part) >= 0)
    b = np.nonzero(part == 7)[0]
    z = mesh.cell_cen[b, 2].reshape(20, 3)
    assert np.all(z == z[:, :1]) and np.all(np.diff(z[:, 0]) > 0)
    xy = mesh.cell_cen[b, :2].reshape(20, 3, 2)
    assert np.all(xy == xy[:1])
