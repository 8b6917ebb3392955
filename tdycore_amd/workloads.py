"""Synthetic Richards problems of the benchmark and the parity tests.

Each builder returns dict(mesh, face_bc, K, porosity, Sr, kw, U0), where
``kw`` holds the MPFAO keyword arguments (van Genuchten m and alpha, G-matrix
method, density model, boundary pressures) and ``U0`` the initial pressure.
Pressure is in Pa, lengths in m, permeability in m^2.
"""
import numpy as np

from . import meshgen as mg


def jittered_cube_problem(n=16, seed=20261202, jitter=0.1):
    """Unit cube of n^3 hexes, interior vertices moved by U(-jitter, jitter)
    of the spacing; per-cell rotated full-tensor K (log10 k_i ~ N(-12, 0.5));
    phi ~ U(0.2, 0.4), Sr 0.1, van Genuchten n = 1.5 (m = 1/3), alpha 1e-4;
    hydrostatic initial state with the water table at z = 0.5; Dirichlet
    top and bottom at the hydrostatic values, no-flow sides."""
    rng = np.random.default_rng(seed)
    m = mg.box_mesh(n, n, n, jitter=jitter, seed=seed + 1)
    K = mg.random_rotated_K(m.nc, rng, -12.0, 0.5)
    phi = rng.uniform(0.2, 0.4, size=m.nc)
    Sr = np.full(m.nc, 0.1)
    fbc = m.face_bc(dirichlet=["z-", "z+"], neumann=["x-", "x+", "y-", "y+"])
    pb = mg.boundary_values(m, lambda x: mg.hydrostatic_pressure(x[:, 2], 0.5))
    U0 = mg.hydrostatic_pressure(m.cell_cen[:, 2], 0.5)
    kw = dict(vg_m=1.0 / 3.0, vg_alpha=1e-4, gmatrix="mpfao", density="constant", p_bnd_fn=pb)
    return dict(mesh=m, face_bc=fbc, K=K, porosity=phi, Sr=Sr, kw=kw, U0=U0)


def layered_box_problem(nx=128, ny=128, nz=64, seed=20261203, lx=1000.0, ly=1000.0, lz=100.0, jitter=0.1):
    """nx*ny*nz hexes on lx*ly*lz m, interior vertices jittered as in the
    unit-cube problem; per layer log10 k ~ N(-12, 1) smoothed by a 5x5 x-y
    moving average (rescaled back to sigma 1), K = diag(k, k, k/10);
    phi ~ U(0.2, 0.4), Sr 0.1, van Genuchten n = 1.5 (m = 1/3), alpha 1e-4;
    hydrostatic initial state with the water table at mid-height; Dirichlet
    top and bottom at the hydrostatic values, no-flow sides.  On a jittered
    mesh the discrete hydrostatic state is not an exact equilibrium, so every
    time step drives a genuine Newton solve."""
    rng = np.random.default_rng(seed)
    m = mg.box_mesh(nx, ny, nz, upper=(lx, ly, lz), jitter=jitter, seed=seed + 1)
    K = mg.layered_K(nx, ny, nz, rng)
    phi = rng.uniform(0.2, 0.4, size=m.nc)
    Sr = np.full(m.nc, 0.1)
    fbc = m.face_bc(dirichlet=["z-", "z+"], neumann=["x-", "x+", "y-", "y+"])
    zt = 0.5 * lz
    pb = mg.boundary_values(m, lambda x: mg.hydrostatic_pressure(x[:, 2], zt))
    U0 = mg.hydrostatic_pressure(m.cell_cen[:, 2], zt)
    kw = dict(vg_m=1.0 / 3.0, vg_alpha=1e-4, gmatrix="mpfao", density="constant", p_bnd_fn=pb)
    return dict(mesh=m, face_bc=fbc, K=K, porosity=phi, Sr=Sr, kw=kw, U0=U0)


def ensemble_box_problem(nx, ny, nz, seed, sigma_log10, jitter=0.1):
    """One request of the serving ensemble: the unit cube in nx*ny*nz hexes,
    jittered as the unit-cube problem, with its own seed for the rotated
    log-normal tensor permeability (log10 k_i ~ N(-12, sigma_log10));
    phi ~ U(0.2, 0.4), Sr 0.1, van Genuchten n = 1.5, alpha 1e-4; hydrostatic
    initial state with the water table at z = 0.5, Dirichlet top and bottom
    at the hydrostatic values, no-flow sides."""
    rng = np.random.default_rng(seed)
    m = mg.box_mesh(nx, ny, nz, jitter=jitter, seed=seed + 1)
    K = mg.random_rotated_K(m.nc, rng, -12.0, sigma_log10)
    phi = rng.uniform(0.2, 0.4, size=m.nc)
    Sr = np.full(m.nc, 0.1)
    fbc = m.face_bc(dirichlet=["z-", "z+"], neumann=["x-", "x+", "y-", "y+"])
    pb = mg.boundary_values(m, lambda x: mg.hydrostatic_pressure(x[:, 2], 0.5))
    U0 = mg.hydrostatic_pressure(m.cell_cen[:, 2], 0.5)
    kw = dict(vg_m=1.0 / 3.0, vg_alpha=1e-4, gmatrix="mpfao", density="constant", p_bnd_fn=pb)
    return dict(mesh=m, face_bc=fbc, K=K, porosity=phi, Sr=Sr, kw=kw, U0=U0)


def ensemble_request_specs(n=512, seed=20261206):
    """Sizes and seeds of the serving ensemble: nx, ny ~ U{16..96},
    nz ~ U{8..48}, sigma(log10 k) ~ U(0.5, 1.5), one seed per request."""
    rng = np.random.default_rng(seed)
    nx = rng.integers(16, 97, size=n)
    ny = rng.integers(16, 97, size=n)
    nz = rng.integers(8, 49, size=n)
    sig = rng.uniform(0.5, 1.5, size=n)
    return [dict(rid=i, nx=int(nx[i]), ny=int(ny[i]), nz=int(nz[i]), seed=int(seed + 1000 * (i + 1)),
                 sigma=float(sig[i])) for i in range(n)]


def prism_problem(n_tri=200_000, nlayers=20, seed=20261205, length=10_000.0, lattice=128, gmatrix="mpfao",
                  points=None, triangles=None, infiltration=False, column_major=True, block_tris=3):
    """Right prisms: a random Delaunay triangulation of about n_tri triangles
    on a length x length square (jittered lattice points, scipy's Delaunay),
    extruded through nlayers layers of thickness U(0.5, 5) m.  Permeability as
    the layered box: per layer log10 k ~ N(-12, 1) on a lattice x lattice grid
    smoothed by a 5x5 moving average (rescaled back to sigma 1) and sampled at
    the cell centroids, K = diag(k, k, k/10); phi ~ U(0.2, 0.4), Sr 0.1,
    van Genuchten n = 1.5 (m = 1/3), alpha 1e-4; hydrostatic initial state
    with the water table at mid-height; Dirichlet top and bottom at the
    hydrostatic values, no-flow sides.  ``points``/``triangles`` replace the
    triangulation (then n_tri is unused).  On right prisms the hydrostatic
    state is an exact discrete equilibrium (flat layers, vertical walls), so
    with ``infiltration`` the top boundary is held at atmospheric pressure
    instead (ponded infiltration into the unsaturated upper half), which makes
    every time step a genuine Newton solve.

    Cell numbering: with ``column_major`` the triangles are put in reverse
    Cuthill-McKee order and cut into groups of ``block_tris`` consecutive
    ones; the cells of a group are numbered together, layer by layer inside
    the group, so that every Jacobian column offset is small (it fits int16
    at 4M prisms) and ``pc_part`` (one block-Jacobi block per group of
    columns) keeps each block's rows in layer order, the order whose ILU(0)
    converges best here.  Otherwise cell = triangle + n_tri * layer and
    ``pc_part`` is None."""
    rng = np.random.default_rng(seed)
    if points is None:
        points, triangles = mg.delaunay_square(n_tri, length, rng)
    Z = np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 5.0, size=nlayers))])
    m = mg.prism_mesh(points, triangles, Z)
    field = mg.layered_K(lattice, lattice, nlayers, rng)[:, 0].reshape(nlayers, lattice, lattice)
    lo, hi = points.min(axis=0), points.max(axis=0)
    ix = np.clip(((m.cell_cen[:, 0] - lo[0]) / (hi[0] - lo[0]) * lattice).astype(np.int64), 0, lattice - 1)
    iy = np.clip(((m.cell_cen[:, 1] - lo[1]) / (hi[1] - lo[1]) * lattice).astype(np.int64), 0, lattice - 1)
    layer = np.arange(m.nc) // triangles.shape[0]
    k = field[layer, iy, ix]
    K = np.zeros((m.nc, 9))
    K[:, 0] = K[:, 4] = k
    K[:, 8] = k / 10.0
    phi = rng.uniform(0.2, 0.4, size=m.nc)
    Sr = np.full(m.nc, 0.1)
    fbc = m.face_bc(dirichlet=["z-", "z+"], neumann=["x-"])
    zt = 0.5 * Z[-1]
    pb = mg.boundary_values(m, lambda x: mg.hydrostatic_pressure(x[:, 2], zt))
    if infiltration:
        top = m.face_cen[m.bnd_face, 2] >= Z[-1] - 1e-9
        pb[top] = 101325.0
    U0 = mg.hydrostatic_pressure(m.cell_cen[:, 2], zt)
    pc_part = None
    if column_major:
        nt, g = triangles.shape[0], int(block_tris)
        rank = np.empty(nt, dtype=np.int64)
        rank[mg.rcm_triangle_order(triangles, points.shape[0])] = np.arange(nt)
        grp, j = rank // g, rank % g
        gsize = np.minimum(g, nt - grp * g)  # the last group may be short
        c = np.arange(m.nc)
        t, k = c % nt, c // nt
        new_of_old = grp[t] * g * nlayers + k * gsize[t] + j[t]
        m = mg.renumber_cells(m, new_of_old)
        old_of_new = np.empty_like(new_of_old)
        old_of_new[new_of_old] = c
        K, phi, Sr, U0 = K[old_of_new], phi[old_of_new], Sr[old_of_new], U0[old_of_new]
        pc_part = (c // (g * nlayers)).astype(np.int32)
    kw = dict(vg_m=1.0 / 3.0, vg_alpha=1e-4, gmatrix=gmatrix, density="constant", p_bnd_fn=pb)
    return dict(mesh=m, face_bc=fbc, K=K, porosity=phi, Sr=Sr, kw=kw, U0=U0, pc_part=pc_part)


def wheel_triangulation(n_spokes=14, rings=2, radius=100.0):
    """A centre point with n_spokes triangles around it, inside rings of
    points: its prism vertices have 2*n_spokes cells and 3*n_spokes faces,
    the largest interaction regions the setup buckets handle."""
    from scipy.spatial import Delaunay
    pts = [[0.0, 0.0]]
    for r in range(1, rings + 1):
        n = n_spokes * r
        a = 2 * np.pi * (np.arange(n) + 0.5 * (r % 2)) / n
        pts += list(np.stack([r * radius * np.cos(a), r * radius * np.sin(a)], 1))
    P = np.array(pts)
    return P, Delaunay(P).simplices.astype(np.int64)


def star_hex_problem(n_quads=6, layers=2, seed=16, radius=100.0, thickness=10.0):
    """Hexahedra around one vertex of high valence: n_quads quadrilaterals
    (centre, A_i, B_i, A_i+1) around a centre point, A_i on a circle, B_i
    outside between them, extruded into ``layers`` layers of ``thickness``.
    The centre vertex of an inner layer has 2 n_quads cells and 3 n_quads
    faces (12 and 18 for n_quads = 6), more than the hexahedral kernels'
    8 cells and 12 faces, so such a mesh takes the general path.  Rotated
    log-normal tensors (sigma 0.5), phi ~ U(0.2, 0.4), Sr 0.1, van Genuchten
    n = 1.5, alpha 1e-4; Dirichlet on every boundary face at the hydrostatic
    pressure of a water table at mid-height."""
    from . import exodus
    rng = np.random.default_rng(seed)
    a = 2 * np.pi * np.arange(n_quads) / n_quads
    b = a + np.pi / n_quads
    P = np.concatenate([[[0.0, 0.0]], radius * np.column_stack([np.cos(a), np.sin(a)]),
                        1.4 * radius * np.column_stack([np.cos(b), np.sin(b)])])
    quads = [[0, 1 + i, 1 + n_quads + i, 1 + (i + 1) % n_quads] for i in range(n_quads)]
    ns = P.shape[0]
    X = np.concatenate([np.column_stack([P, np.full(ns, l * thickness)]) for l in range(layers + 1)])
    cells = [[v + l * ns for v in q] + [v + (l + 1) * ns for v in q] for l in range(layers) for q in quads]
    m = exodus.connectivity_mesh(X, cells)
    K = mg.random_rotated_K(m.nc, rng, -12.0, 0.5)
    phi = rng.uniform(0.2, 0.4, size=m.nc)
    Sr = np.full(m.nc, 0.1)
    fbc = m.face_bc(default=mg.BC_DIRICHLET)
    zt = 0.5 * layers * thickness
    pb = mg.boundary_values(m, lambda x: mg.hydrostatic_pressure(x[:, 2], zt))
    U0 = mg.hydrostatic_pressure(m.cell_cen[:, 2], zt)
    kw = dict(vg_m=1.0 / 3.0, vg_alpha=1e-4, gmatrix="mpfao", density="constant", p_bnd_fn=pb)
    return dict(mesh=m, face_bc=fbc, K=K, porosity=phi, Sr=Sr, kw=kw, U0=U0)


# constant, symmetric positive definite and fully anisotropic (m^2)
MMS_K = 1e-12 * np.array([[1.0, 0.5, 0.25], [0.5, 1.5, 0.5], [0.25, 0.5, 2.0]])
MMS_RHO, MMS_MU, MMS_PREF = 998.0, 9.94e-4, 101325.0


def mms_pressure(x, amp=1e5):
    """Manufactured pressure p = pref + amp (1 + sin(pi x) sin(pi y) sin(pi z)):
    above pref everywhere on the unit cube, so every cell is saturated
    (S = 1, kr = 1) and the Richards residual reduces to the linear flux
    operator plus the source."""
    s = np.sin(np.pi * x)
    return MMS_PREF + amp * (1.0 + s[:, 0] * s[:, 1] * s[:, 2])


def mms_div_K_grad(x, amp=1e5, K=MMS_K):
    """div(K grad p) of mms_pressure for the constant tensor K."""
    s, c = np.sin(np.pi * x), np.cos(np.pi * x)
    sx, sy, sz, cx, cy, cz = s[:, 0], s[:, 1], s[:, 2], c[:, 0], c[:, 1], c[:, 2]
    return amp * np.pi ** 2 * (-(K[0, 0] + K[1, 1] + K[2, 2]) * sx * sy * sz + 2 * K[0, 1] * cx * cy * sz
                               + 2 * K[0, 2] * cx * sy * cz + 2 * K[1, 2] * sx * cy * cz)


def mms_problem(n, warp=0.0, amp=1e5):
    """Steady saturated flow on the unit cube in n^3 hexahedra (smoothly
    warped with ``warp``, meshgen.box_mesh) with the manufactured pressure
    mms_pressure: constant full tensor MMS_K, constant density, no gravity,
    Dirichlet on every boundary face at the exact pressure of the face
    centroid.  The SNES residual is R_c = (net mass outflow of c) - Q_c, so
    the source is Q_c = -(rho / mu) V_c div(K grad p)(x_c) (midpoint rule; the
    reference's forcing term is f V_c at the centroid too, tdympfao.c:2810-2819).
    U0 is the constant pref + amp; the exact solution is mms_pressure at the
    cell centroids up to the discretisation error, second order in h on
    smooth problems."""
    m = mg.box_mesh(n, n, n, warp=warp)
    K = np.tile(MMS_K.ravel(), (m.nc, 1))
    phi = np.full(m.nc, 0.3)
    Sr = np.full(m.nc, 0.1)
    fbc = m.face_bc(default=mg.BC_DIRICHLET)
    pb = mms_pressure(m.face_cen[m.bnd_face], amp)
    Q = -(MMS_RHO / MMS_MU) * m.cell_vol * mms_div_K_grad(m.cell_cen, amp)
    U0 = np.full(m.nc, MMS_PREF + amp)
    kw = dict(vg_m=1.0 / 3.0, vg_alpha=1e-4, gmatrix="mpfao", density="constant", gravity=0.0, p_bnd_fn=pb)
    return dict(mesh=m, face_bc=fbc, K=K, porosity=phi, Sr=Sr, kw=kw, U0=U0, Q=Q,
                p_exact=mms_pressure(m.cell_cen, amp), amp=amp)


def mms_error(P, U):
    """Volume-weighted L2 error of the cell pressures against the exact
    pressure at the centroids, relative to the amplitude."""
    V = P["mesh"].cell_vol
    return float(np.sqrt(np.sum(V * (U - P["p_exact"]) ** 2) / np.sum(V)) / P["amp"])
