"""Right-prism meshes (a triangulation extruded through layers): the mesh
generator's geometry, and pins of the oracle's MPFA-O transmissibilities on
them that do not depend on the oracle's own formulas (row sums of a
consistent flux, exact fluxes of a linear pressure field)."""
import numpy as np
import pytest

from tdycore_amd import meshgen as mg
from pyoracle import Oracle


def _prisms(seed=3, ntri=60, nl=3, length=100.0):
    rng = np.random.default_rng(seed)
    P, T = mg.delaunay_square(ntri, length, rng)
    Z = np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 5.0, size=nl))])
    return mg.prism_mesh(P, T, Z), rng, length, Z


def test_prism_geometry_is_closed_and_exact():
    m, _, L, Z = _prisms(ntri=200, nl=4, length=1000.0)
    assert np.isclose(m.cell_vol.sum(), L * L * Z[-1], rtol=1e-12)
    S = np.zeros((m.nc, 3))
    cells = np.arange(m.nc)
    for lf in range(6):
        f = m.cell_faces[:, lf]
        ok = f >= 0
        sgn = np.where(m.face_cells[f[ok], 0] == cells[ok], 1.0, -1.0)  # outward normal
        S[ok] += sgn[:, None] * m.face_normal[f[ok]] * m.face_area[f[ok]][:, None]
    assert np.abs(S).max() < 1e-9 * m.face_area.max()
    # 6 vertices, 5 faces per prism; triangles have 3 vertices, side quads 4
    assert ((m.cell_verts >= 0).sum(1) == 6).all() and ((m.cell_faces >= 0).sum(1) == 5).all()
    nvf = (m.face_verts >= 0).sum(1)
    assert set(np.unique(nvf)) == {3, 4}


@pytest.mark.parametrize("gmatrix", ["mpfao", "tpf"])
def test_prism_trans_rows_sum_to_zero(gmatrix):
    m, rng, _, _ = _prisms()
    K = mg.random_rotated_K(m.nc, rng, -12.0, 0.5) if gmatrix == "mpfao" else \
        mg.isotropic_K(m.nc, 1e-12) * (10.0 ** rng.uniform(-0.5, 0.5, size=m.nc))[:, None]
    o = Oracle(m, m.face_bc(default=mg.BC_DIRICHLET), K, np.full(m.nc, 0.3), np.full(m.nc, 0.1),
               gmatrix=gmatrix, p_bnd_fn=np.zeros(m.nbnd))
    rp, cols, vals = o.trans_csr()
    live = rp[1:] > rp[:-1]
    rs = np.add.reduceat(vals, rp[:-1][live])
    assert np.abs(rs).max() <= 1e-13 * np.abs(vals).max()
    # every flux row of a triangle face is one of its 3 sub-faces
    rows = np.nonzero(live)[0]
    tri = (m.face_verts[rows // 4] >= 0).sum(1) == 3
    assert (rows[tri] % 4 < 3).all()


def test_prism_linear_pressure_is_reproduced_exactly():
    """MPFA-O with face-centroid continuity points is exact for linear
    pressure: T p = A_f/n_v (n . K grad p) on every sub-face, any prism."""
    m, rng, _, _ = _prisms(seed=5)
    Kc = np.tile(mg.random_rotated_K(1, rng, -12.0, 0.3), (m.nc, 1))
    o = Oracle(m, m.face_bc(default=mg.BC_DIRICHLET), Kc, np.full(m.nc, 0.3), np.full(m.nc, 0.1),
               p_bnd_fn=np.zeros(m.nbnd))
    rp, cols, vals = o.trans_csr()
    g = np.array([0.3, -0.7, 1.1])
    pv = np.concatenate([m.cell_cen @ g, m.face_cen[m.bnd_face] @ g])
    live = np.nonzero(rp[1:] > rp[:-1])[0]
    TP = np.array([vals[rp[r]:rp[r + 1]] @ pv[cols[rp[r]:rp[r + 1]]] for r in live])
    f = live // 4
    nvf = (m.face_verts[f] >= 0).sum(1)
    exact = m.face_area[f] / nvf * (m.face_normal[f] @ (Kc[0].reshape(3, 3) @ g))
    assert np.abs(TP - exact).max() <= 1e-12 * np.abs(exact).max()
    # a wrong sub-face area (a quarter of a triangle) would fail
    wrong = m.face_area[f] / 4 * (m.face_normal[f] @ (Kc[0].reshape(3, 3) @ g))
    assert np.abs(TP - wrong).max() > 1e-3 * np.abs(exact).max()


def test_prism_ts_jacobian_matches_finite_differences():
    """On prisms, as on hexahedra: where the reference Jacobian is exact
    (boundary pressures from a BC function, no upwind switch under the
    perturbation) it is the derivative of the residual; the residual of a
    closed no-flow box sums to the accumulation (the fluxes cancel)."""
    rng = np.random.default_rng(17)
    P, T = mg.delaunay_square(24, 50.0, rng)
    m = mg.prism_mesh(P, T, np.array([0.0, 2.0, 3.5]))
    K = mg.random_rotated_K(m.nc, rng)
    pb = 9e4 + 2e3 * rng.random(m.nbnd)
    o = Oracle(m, m.face_bc(default=mg.BC_DIRICHLET), K, rng.uniform(0.2, 0.4, m.nc), np.full(m.nc, 0.1),
               vg_m=1 / 3, p_bnd_fn=pb)
    U = 9.5e4 + 1.5e4 * rng.random(m.nc)
    Uprev = U - 50.0 * rng.random(m.nc)
    dt = 300.0

    def R(X):
        return o.ifunction(X, (X - Uprev) / dt)
    R(U)
    J = o.jac_to_csr(o.ijacobian(U, (U - Uprev) / dt, 1.0 / dt)).toarray()
    h = 1e-7 * np.abs(U)
    Jfd = np.empty_like(J)
    for j in range(m.nc):
        e = np.zeros(m.nc)
        e[j] = h[j]
        Jfd[:, j] = (R(U + e) - R(U - e)) / (2 * h[j])
    scale = np.abs(J).max(axis=1, keepdims=True)
    assert np.all(np.abs(J - Jfd) <= 1e-5 * scale)
    # closed box: no-flow everywhere, the fluxes cancel in the sum of R
    o2 = Oracle(m, m.face_bc(default=mg.BC_NEUMANN), K, np.full(m.nc, 0.3), np.full(m.nc, 0.1), vg_m=1 / 3)
    Rs = o2.ifunction(U, np.zeros(m.nc))
    Rf = o2.ifunction(U, np.zeros(m.nc))
    assert np.allclose(Rs, Rf) and abs(Rs.sum()) <= 1e-10 * np.abs(Rs).max()
