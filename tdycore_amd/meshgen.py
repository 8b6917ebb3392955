"""Structured hexahedral box meshes in the unstructured form the MPFA-O path consumes.

This module produces the *inputs* of the MPFA-O Richards path: the mesh
connectivity and finite-volume geometry that TDycore obtains from PETSc
(``DMPlexCreateBoxMesh`` + ``DMPlexComputeCellGeometryFVM``, consumed by
``TDyMeshCreateFromPlex``, src/fv/share/tdymeshplex.c:3685) and the per-cell
material fields that ``InitMaterials`` (src/fv/mpfao/tdympfao.c:630-763)
evaluates once at setup.

Conventions (kept identical to the reference where they matter):

* cell ``c = i + nx*(j + ny*k)`` (x fastest, z slowest), the natural order the
  reference's regression sampler indexes (src/tdyregression.c:67-89);
* vertex ``v = i + (nx+1)*(j + (ny+1)*k)``; a hex cell lists its 8 vertices
  in lexicographic corner order ``a + 2b + 4c``;
* every face stores ``face_cells = (up, dn)`` with the unit normal pointing
  from ``up`` to ``dn``; a boundary face stores ``up = -(b+1)`` (b = boundary
  index in face order) and ``dn = cell``, so boundary normals point into the
  domain (src/fv/share/tdymeshplex.c:3126-3168);
* face normals are the mean of the four unit triangle normals fanned about
  the vertex mean (``TDyNormalToQuadrilateral``, src/tdyutils.c:395-417),
  face area is the sum of those triangle areas (``TDyQuadrilateralArea``,
  src/tdyutils.c:312-333); face and cell centroids and cell volumes use the
  same fan triangulation (divergence theorem).  On planar faces (the
  reference's unit-box golden meshes) all of these are exact.
"""
from __future__ import annotations

import dataclasses

import numpy as np

BC_INTERNAL = -1
BC_DIRICHLET = 0
BC_NEUMANN = 1
BC_SEEPAGE = 2


@dataclasses.dataclass
class Mesh:
    nx: int
    ny: int
    nz: int
    vert_xyz: np.ndarray      # (nv,3) f64
    cell_verts: np.ndarray    # (nc,8) i32
    cell_faces: np.ndarray    # (nc,6) i32  (-x,+x,-y,+y,-z,+z)
    face_verts: np.ndarray    # (nf,4) i32  cyclic
    face_cells: np.ndarray    # (nf,2) i32  (up,dn); boundary up = -(b+1)
    cell_cen: np.ndarray      # (nc,3)
    cell_vol: np.ndarray      # (nc,)
    face_cen: np.ndarray      # (nf,3)
    face_area: np.ndarray     # (nf,)
    face_normal: np.ndarray   # (nf,3), from up to dn
    bnd_face: np.ndarray      # (nbnd,) face id of boundary index b
    face_dir: np.ndarray      # (nf,) 0/1/2 axis of the face family
    face_side: np.ndarray     # (nf,) for boundary faces: 0 = low side, 1 = high side; -1 internal

    @property
    def nc(self) -> int:
        return self.cell_verts.shape[0]

    @property
    def nf(self) -> int:
        return self.face_verts.shape[0]

    @property
    def nv(self) -> int:
        return self.vert_xyz.shape[0]

    @property
    def nbnd(self) -> int:
        return self.bnd_face.shape[0]

    def face_bc(self, *, dirichlet=None, neumann=None, seepage=None, default=BC_DIRICHLET) -> np.ndarray:
        """Per-face boundary type array (``-1`` for internal faces).

        Each keyword takes a list of box sides such as ``["z-", "z+"]``; sides
        not named get ``default``.  With one type on every side this reduces
        to the reference's single global ``bc_type``
        (src/fv/mpfao/tdympfao.c:588-596).
        """
        bc = np.full(self.nf, BC_INTERNAL, dtype=np.int32)
        b = self.face_cells[:, 0] < 0
        bc[b] = default
        names = {"x-": (0, 0), "x+": (0, 1), "y-": (1, 0), "y+": (1, 1), "z-": (2, 0), "z+": (2, 1)}
        for code, sides in ((BC_DIRICHLET, dirichlet), (BC_NEUMANN, neumann), (BC_SEEPAGE, seepage)):
            for s in sides or ():
                d, side = names[s]
                sel = b & (self.face_dir == d) & (self.face_side == side)
                bc[sel] = code
        return bc


def _cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], axis=-1)


def _face_geometry(X):
    """X: (nf,4,3) cyclic vertex coordinates -> normal (unnormalised mean of unit
    triangle normals), area, centroid, and the summed vector area."""
    m = X.mean(axis=1)
    area = np.zeros(X.shape[0])
    cen = np.zeros((X.shape[0], 3))
    nrm = np.zeros((X.shape[0], 3))
    svec = np.zeros((X.shape[0], 3))
    for i in range(4):
        a = X[:, i] - m
        b = X[:, (i + 1) % 4] - m
        t = _cross(a, b)
        tl = np.sqrt((t * t).sum(axis=1))
        area += 0.5 * tl
        nrm += t / tl[:, None]
        svec += 0.5 * t
        cen += (0.5 * tl)[:, None] * (m + X[:, i] + X[:, (i + 1) % 4]) / 3.0
    nrm /= 4.0
    cen /= area[:, None]
    return nrm, area, cen, svec, m


def box_mesh(nx: int, ny: int, nz: int, lower=(0.0, 0.0, 0.0), upper=(1.0, 1.0, 1.0),
             jitter: float = 0.0, seed: int = 0, warp: float = 0.0) -> Mesh:
    """Hex box mesh of nx*ny*nz cells.  ``jitter`` moves every interior vertex
    by U(-jitter, +jitter) times the local spacing per coordinate.  ``warp``
    maps every vertex by the smooth displacement warp * L * s(x) s(y) s(z) in
    each coordinate, s = sin(2 pi (x - lower) / L) per axis (L the box
    length), which keeps the box's faces in place and perturbs the cells
    smoothly and independently of the resolution (faces become non-planar)."""
    lower = np.asarray(lower, dtype=np.float64)
    upper = np.asarray(upper, dtype=np.float64)
    h = (upper - lower) / np.array([nx, ny, nz], dtype=np.float64)
    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    # vertex order: i fastest
    I = I.transpose(2, 1, 0).ravel()
    J = J.transpose(2, 1, 0).ravel()
    K = K.transpose(2, 1, 0).ravel()
    V = np.stack([lower[0] + I * h[0], lower[1] + J * h[1], lower[2] + K * h[2]], axis=1)
    # exact end points
    V[I == nx, 0] = upper[0]
    V[J == ny, 1] = upper[1]
    V[K == nz, 2] = upper[2]
    if warp != 0.0:
        L = upper - lower
        t = 2.0 * np.pi * (V - lower[None, :]) / L[None, :]
        d = warp * np.sin(t[:, 0]) * np.sin(t[:, 1]) * np.sin(t[:, 2])
        V = V + d[:, None] * L[None, :]
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        inner = (I > 0) & (I < nx) & (J > 0) & (J < ny) & (K > 0) & (K < nz)
        d = rng.uniform(-jitter, jitter, size=(int(inner.sum()), 3)) * h[None, :]
        V[inner] += d

    def vid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci = ci.transpose(2, 1, 0).ravel()
    cj = cj.transpose(2, 1, 0).ravel()
    ck = ck.transpose(2, 1, 0).ravel()
    nc = nx * ny * nz
    cell_verts = np.empty((nc, 8), dtype=np.int32)
    for c in range(2):
        for b in range(2):
            for a in range(2):
                cell_verts[:, a + 2 * b + 4 * c] = vid(ci + a, cj + b, ck + c)

    NX = (nx + 1) * ny * nz
    NY = nx * (ny + 1) * nz
    NZ = nx * ny * (nz + 1)
    nf = NX + NY + NZ

    def cid(i, j, k):
        return i + nx * (j + ny * k)

    face_verts = np.empty((nf, 4), dtype=np.int32)
    face_cells = np.empty((nf, 2), dtype=np.int32)
    face_dir = np.empty(nf, dtype=np.int32)
    face_side = np.full(nf, -1, dtype=np.int32)

    def grid(n0, n1, n2):
        a, b, c = np.meshgrid(np.arange(n0), np.arange(n1), np.arange(n2), indexing="ij")
        return a.transpose(2, 1, 0).ravel(), b.transpose(2, 1, 0).ravel(), c.transpose(2, 1, 0).ravel()

    # x-faces
    fi, fj, fk = grid(nx + 1, ny, nz)
    sl = slice(0, NX)
    face_verts[sl] = np.stack([vid(fi, fj, fk), vid(fi, fj + 1, fk), vid(fi, fj + 1, fk + 1), vid(fi, fj, fk + 1)], 1)
    lo = np.where(fi > 0, cid(np.maximum(fi - 1, 0), fj, fk), -1)
    hi = np.where(fi < nx, cid(np.minimum(fi, nx - 1), fj, fk), -1)
    face_cells[sl, 0], face_cells[sl, 1] = lo, hi
    face_dir[sl] = 0
    face_side[sl] = np.where(fi == 0, 0, np.where(fi == nx, 1, -1))
    # y-faces
    fi, fj, fk = grid(nx, ny + 1, nz)
    sl = slice(NX, NX + NY)
    face_verts[sl] = np.stack([vid(fi, fj, fk), vid(fi, fj, fk + 1), vid(fi + 1, fj, fk + 1), vid(fi + 1, fj, fk)], 1)
    lo = np.where(fj > 0, cid(fi, np.maximum(fj - 1, 0), fk), -1)
    hi = np.where(fj < ny, cid(fi, np.minimum(fj, ny - 1), fk), -1)
    face_cells[sl, 0], face_cells[sl, 1] = lo, hi
    face_dir[sl] = 1
    face_side[sl] = np.where(fj == 0, 0, np.where(fj == ny, 1, -1))
    # z-faces
    fi, fj, fk = grid(nx, ny, nz + 1)
    sl = slice(NX + NY, nf)
    face_verts[sl] = np.stack([vid(fi, fj, fk), vid(fi + 1, fj, fk), vid(fi + 1, fj + 1, fk), vid(fi, fj + 1, fk)], 1)
    lo = np.where(fk > 0, cid(fi, fj, np.maximum(fk - 1, 0)), -1)
    hi = np.where(fk < nz, cid(fi, fj, np.minimum(fk, nz - 1)), -1)
    face_cells[sl, 0], face_cells[sl, 1] = lo, hi
    face_dir[sl] = 2
    face_side[sl] = np.where(fk == 0, 0, np.where(fk == nz, 1, -1))

    # boundary faces: store (-(b+1), cell)
    bmask = (face_cells[:, 0] < 0) | (face_cells[:, 1] < 0)
    bnd_face = np.nonzero(bmask)[0].astype(np.int32)
    nb = bnd_face.shape[0]
    cell_of_b = np.where(face_cells[bnd_face, 0] >= 0, face_cells[bnd_face, 0], face_cells[bnd_face, 1])
    face_cells[bnd_face, 0] = -(np.arange(nb, dtype=np.int32) + 1)
    face_cells[bnd_face, 1] = cell_of_b

    # cell faces (-x,+x,-y,+y,-z,+z)
    cell_faces = np.empty((nc, 6), dtype=np.int32)
    cell_faces[:, 0] = ci + (nx + 1) * (cj + ny * ck)
    cell_faces[:, 1] = ci + 1 + (nx + 1) * (cj + ny * ck)
    cell_faces[:, 2] = NX + ci + nx * (cj + (ny + 1) * ck)
    cell_faces[:, 3] = NX + ci + nx * (cj + 1 + (ny + 1) * ck)
    cell_faces[:, 4] = NX + NY + ci + nx * (cj + ny * ck)
    cell_faces[:, 5] = NX + NY + ci + nx * (cj + ny * (ck + 1))

    # ---- geometry ----
    FX = V[face_verts]                       # (nf,4,3)
    fnrm, farea, fcen, fsvec, fmean = _face_geometry(FX)

    # cells: volume and centroid by fan tetrahedra about the vertex mean
    CX = V[cell_verts]                       # (nc,8,3)
    p0 = CX.mean(axis=1)
    vol = np.zeros(nc)
    cen = np.zeros((nc, 3))
    for lf in range(6):
        f = cell_faces[:, lf]
        Xf = FX[f]                           # (nc,4,3)
        mf = fmean[f]
        # outward orientation of the stored cyclic order
        sgn = np.sign(((fsvec[f]) * (mf - p0)).sum(axis=1))
        for i in range(4):
            a = mf - p0
            b = Xf[:, i] - p0
            c = Xf[:, (i + 1) % 4] - p0
            tv = sgn * (a * _cross(b, c)).sum(axis=1) / 6.0
            vol += tv
            cen += tv[:, None] * (p0 + mf + Xf[:, i] + Xf[:, (i + 1) % 4]) / 4.0
    cen /= vol[:, None]

    # orient normals from up to dn (boundary: from the outside into the cell)
    up, dn = face_cells[:, 0], face_cells[:, 1]
    ref_from = np.where(up[:, None] >= 0, cen[np.maximum(up, 0)], fcen)
    ref_to = cen[dn]
    s = np.sign(((ref_to - ref_from) * fnrm).sum(axis=1))
    s[s == 0] = 1.0
    fnrm = fnrm * s[:, None]

    return Mesh(nx=nx, ny=ny, nz=nz, vert_xyz=V, cell_verts=cell_verts, cell_faces=cell_faces,
                face_verts=face_verts, face_cells=face_cells, cell_cen=cen, cell_vol=vol,
                face_cen=fcen, face_area=farea, face_normal=fnrm, bnd_face=bnd_face,
                face_dir=face_dir, face_side=face_side)


# --------------------------------------------------------------------------
# Material fields (per-cell inputs, evaluated once at setup)
# --------------------------------------------------------------------------

def isotropic_K(nc: int, k: float) -> np.ndarray:
    K = np.zeros((nc, 9))
    K[:, 0] = K[:, 4] = K[:, 8] = k
    return K


def random_rotated_K(nc: int, rng: np.random.Generator, mean_log10: float = -12.0,
                     sigma_log10: float = 0.5) -> np.ndarray:
    """Per-cell SPD tensors K = R diag(k1,k2,k3) R^T with log10 k_i ~ N(mean, sigma)
    and R Haar-distributed (QR of a Gaussian matrix with the sign fix)."""
    k = 10.0 ** rng.normal(mean_log10, sigma_log10, size=(nc, 3))
    A = rng.normal(size=(nc, 3, 3))
    Q, R = np.linalg.qr(A)
    d = np.sign(np.diagonal(R, axis1=1, axis2=2))
    d[d == 0] = 1.0
    Q = Q * d[:, None, :]
    K = np.einsum("nij,nj,nkj->nik", Q, k, Q)
    K = 0.5 * (K + K.transpose(0, 2, 1))
    return K.reshape(nc, 9)


def layered_K(nx: int, ny: int, nz: int, rng: np.random.Generator, mean_log10: float = -12.0,
              sigma_log10: float = 1.0, window: int = 5, anisotropy: float = 10.0) -> np.ndarray:
    """Layered heterogeneous permeability: per layer, i.i.d. N(mean, sigma) log10 k
    smoothed by a window x window x-y moving average (periodic wrap), then
    rescaled per layer back to standard deviation ``sigma`` (the plain average
    would shrink it by ``window``).  K = diag(k, k, k/anisotropy)."""
    g = rng.normal(0.0, 1.0, size=(nz, ny, nx))
    s = np.zeros_like(g)
    r = window // 2
    for dy in range(-r, r + 1):
        for dx in range(-r, r + 1):
            s += np.roll(np.roll(g, dy, axis=1), dx, axis=2)
    s /= float(window * window)
    s -= s.mean(axis=(1, 2), keepdims=True)
    sd = s.std(axis=(1, 2), keepdims=True)
    sd[sd == 0] = 1.0
    s /= sd
    logk = mean_log10 + sigma_log10 * s
    k = (10.0 ** logk).ravel()            # cell order i + nx*(j + ny*k)
    K = np.zeros((k.shape[0], 9))
    K[:, 0] = k
    K[:, 4] = k
    K[:, 8] = k / anisotropy
    return K


def hydrostatic_pressure(z: np.ndarray, z_table: float, rho: float = 998.0, g: float = 9.8068,
                         pref: float = 101325.0) -> np.ndarray:
    """p = Pref + rho g (z_table - z): atmospheric at the water table."""
    return pref + rho * g * (z_table - z)


def boundary_values(mesh: Mesh, fn) -> np.ndarray:
    """Evaluate a boundary-pressure function at boundary face centroids
    (``ConditionsComputeBoundaryPressure`` at ``faces->centroid``,
    src/fv/mpfao/tdympfao_utils.c:745-750).  Returns (nbnd,)."""
    return np.asarray(fn(mesh.face_cen[mesh.bnd_face]), dtype=np.float64)


def brick_partition(mesh: Mesh, bx: int, by: int, bz: int) -> np.ndarray:
    """Block id per cell for bx*by*bz bricks of a box mesh (block-Jacobi blocks)."""
    c = np.arange(mesh.nc)
    i, j, k = c % mesh.nx, (c // mesh.nx) % mesh.ny, c // (mesh.nx * mesh.ny)
    gx, gy = -(-mesh.nx // bx), -(-mesh.ny // by)
    return ((i // bx) + gx * ((j // by) + gy * (k // bz))).astype(np.int32)


# --------------------------------------------------------------------------
# Right-prism meshes (a 2-D triangulation extruded through layers)
# --------------------------------------------------------------------------

def prism_mesh(points: np.ndarray, triangles: np.ndarray, z_levels: np.ndarray) -> Mesh:
    """Extrude a triangulation through the levels ``z_levels`` (nl+1, increasing).

    The arrays keep the hexahedral widths, padded with -1: a prism lists 6 of
    8 vertex slots (bottom triangle a, b, c then top a, b, c), 5 of 6 face
    slots (bottom, top, then the three side quads), and a triangular face 3
    of 4 vertex slots.  Flux rows stay ``face*4 + local vertex`` as in the
    reference (src/fv/mpfao/tdympfao.c:1451-1455), so a triangle leaves row
    4f+3 unused.

    * cell ``c = t + nt*k`` (triangle t, layer k), vertex ``v = p + np*k``;
    * horizontal faces first (``t + nt*k`` at level k, normal +z from the
      cell below to the cell above), then the side quads (``nh + e + ne*k``
      for 2-D edge e at layer k);
    * boundary faces store ``(-(b+1), cell)`` with the normal into the domain;
      ``face_dir`` is 2 for the bottom/top (``face_side`` 0/1) and 0 with
      ``face_side`` 0 for the lateral boundary (``face_bc(neumann=["x-"])``);
    * geometry is exact: planar faces, cell centroid at the triangle centroid
      and mid-layer height, volume = triangle area x thickness.
    """
    P = np.asarray(points, dtype=np.float64)
    T = np.asarray(triangles, dtype=np.int64).copy()
    Z = np.asarray(z_levels, dtype=np.float64)
    npt, nt, nl = P.shape[0], T.shape[0], Z.shape[0] - 1
    # counter-clockwise triangles
    a, b, c = P[T[:, 0]], P[T[:, 1]], P[T[:, 2]]
    sarea = 0.5 * ((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0]))
    cw = sarea < 0
    T[cw, 1], T[cw, 2] = T[cw, 2].copy(), T[cw, 1].copy()
    tarea = np.abs(sarea)
    # unique 2-D edges and their one or two triangles
    E = np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]])
    Et = np.concatenate([np.arange(nt)] * 3)
    Es = np.sort(E, axis=1)
    key = Es[:, 0] * npt + Es[:, 1]
    uk, inv = np.unique(key, return_inverse=True)
    ne = uk.shape[0]
    edges = np.stack([uk // npt, uk % npt], axis=1)
    et = np.full((ne, 2), -1, dtype=np.int64)
    order = np.argsort(inv, kind="stable")
    inv_s, t_s = inv[order], Et[order]
    first = np.ones(inv_s.shape[0], dtype=bool)
    first[1:] = inv_s[1:] != inv_s[:-1]
    et[inv_s[first], 0] = t_s[first]
    et[inv_s[~first], 1] = t_s[~first]
    tri_edge = inv.reshape(3, nt).T                      # (nt,3): edges (ab, bc, ca)

    nv, nc = npt * (nl + 1), nt * nl
    V = np.zeros((nv, 3))
    V[:, 0] = np.tile(P[:, 0], nl + 1)
    V[:, 1] = np.tile(P[:, 1], nl + 1)
    V[:, 2] = np.repeat(Z, npt)
    kk = np.repeat(np.arange(nl), nt)
    tt = np.tile(np.arange(nt), nl)
    cell_verts = np.full((nc, 8), -1, dtype=np.int32)
    for s in range(3):
        cell_verts[:, s] = T[tt, s] + npt * kk
        cell_verts[:, 3 + s] = T[tt, s] + npt * (kk + 1)

    nh = nt * (nl + 1)
    nf = nh + ne * nl
    face_verts = np.full((nf, 4), -1, dtype=np.int32)
    face_cells = np.empty((nf, 2), dtype=np.int32)
    face_dir = np.zeros(nf, dtype=np.int32)
    face_side = np.full(nf, -1, dtype=np.int32)
    # horizontal faces
    hk = np.repeat(np.arange(nl + 1), nt)
    ht = np.tile(np.arange(nt), nl + 1)
    for s in range(3):
        face_verts[:nh, s] = T[ht, s] + npt * hk
    face_cells[:nh, 0] = np.where(hk > 0, ht + nt * (hk - 1), -1)
    face_cells[:nh, 1] = np.where(hk < nl, ht + nt * hk, -1)
    face_dir[:nh] = 2
    face_side[:nh] = np.where(hk == 0, 0, np.where(hk == nl, 1, -1))
    # side quads
    vk = np.repeat(np.arange(nl), ne)
    ve = np.tile(np.arange(ne), nl)
    p0, p1 = edges[ve, 0], edges[ve, 1]
    face_verts[nh:] = np.stack([p0 + npt * vk, p1 + npt * vk, p1 + npt * (vk + 1), p0 + npt * (vk + 1)], 1)
    t0, t1 = et[ve, 0], et[ve, 1]
    face_cells[nh:, 0] = t0 + nt * vk
    face_cells[nh:, 1] = np.where(t1 >= 0, t1 + nt * vk, -1)
    face_dir[nh:] = 0
    face_side[nh:] = np.where(t1 >= 0, -1, 0)

    bmask = (face_cells[:, 0] < 0) | (face_cells[:, 1] < 0)
    bnd_face = np.nonzero(bmask)[0].astype(np.int32)
    nb = bnd_face.shape[0]
    cell_of_b = np.where(face_cells[bnd_face, 0] >= 0, face_cells[bnd_face, 0], face_cells[bnd_face, 1])
    face_cells[bnd_face, 0] = -(np.arange(nb, dtype=np.int32) + 1)
    face_cells[bnd_face, 1] = cell_of_b

    cell_faces = np.full((nc, 6), -1, dtype=np.int32)
    cell_faces[:, 0] = tt + nt * kk
    cell_faces[:, 1] = tt + nt * (kk + 1)
    for s in range(3):
        cell_faces[:, 2 + s] = nh + tri_edge[tt, s] + ne * kk

    # geometry
    tc = (a + b + c) / 3.0
    dz = Z[1:] - Z[:-1]
    cell_cen = np.zeros((nc, 3))
    cell_cen[:, :2] = tc[tt]
    cell_cen[:, 2] = 0.5 * (Z[kk] + Z[kk + 1])
    cell_vol = tarea[tt] * dz[kk]
    face_cen = np.zeros((nf, 3))
    face_area = np.zeros(nf)
    fnrm = np.zeros((nf, 3))
    face_cen[:nh, :2] = tc[ht]
    face_cen[:nh, 2] = Z[hk]
    face_area[:nh] = tarea[ht]
    fnrm[:nh, 2] = 1.0
    q0, q1 = P[p0], P[p1]
    face_cen[nh:, :2] = 0.5 * (q0 + q1)
    face_cen[nh:, 2] = 0.5 * (Z[vk] + Z[vk + 1])
    el = np.sqrt(((q1 - q0) ** 2).sum(axis=1))
    face_area[nh:] = el * dz[vk]
    fnrm[nh:, 0] = (q1 - q0)[:, 1] / el
    fnrm[nh:, 1] = -(q1 - q0)[:, 0] / el
    # orient normals from up to dn (boundary: from the outside into the cell)
    up, dn = face_cells[:, 0], face_cells[:, 1]
    ref_from = np.where(up[:, None] >= 0, cell_cen[np.maximum(up, 0)], face_cen)
    ref_to = cell_cen[dn]
    sgn = np.sign(((ref_to - ref_from) * fnrm).sum(axis=1))
    sgn[sgn == 0] = 1.0
    fnrm = fnrm * sgn[:, None]
    return Mesh(nx=nt, ny=1, nz=nl, vert_xyz=V, cell_verts=cell_verts, cell_faces=cell_faces,
                face_verts=face_verts, face_cells=face_cells, cell_cen=cell_cen, cell_vol=cell_vol,
                face_cen=face_cen, face_area=face_area, face_normal=fnrm, bnd_face=bnd_face,
                face_dir=face_dir, face_side=face_side)


def delaunay_square(n_tri: int, length: float, rng: np.random.Generator, jitter: float = 0.35):
    """Random Delaunay triangulation of a square with about n_tri triangles:
    an m x m lattice (2(m-1)^2 ~ n_tri) whose interior points move by
    U(-jitter, +jitter) lattice spacings, triangulated by scipy's Delaunay.
    Returns (points (np,2), triangles (nt,3))."""
    from scipy.spatial import Delaunay
    m = int(round(np.sqrt(n_tri / 2.0))) + 1
    h = length / (m - 1)
    I, J = np.meshgrid(np.arange(m), np.arange(m), indexing="xy")
    X = np.stack([I.ravel() * h, J.ravel() * h], axis=1)
    inner = (I.ravel() > 0) & (I.ravel() < m - 1) & (J.ravel() > 0) & (J.ravel() < m - 1)
    X[inner] += rng.uniform(-jitter, jitter, size=(int(inner.sum()), 2)) * h
    tri = Delaunay(X)
    return X, tri.simplices.astype(np.int64)


def column_partition(mesh: Mesh, per_block: int, grid: int = 256) -> np.ndarray:
    """Block id per cell of a prism mesh: the vertical columns (cells with the
    same x-y centroid) of per_block triangles that are neighbours in a
    grid x grid ordering of their centroids, so a block holds per_block *
    nlayers cells and its rows couple through at most a few in-block
    neighbours.  Any cell numbering."""
    xy = np.ascontiguousarray(mesh.cell_cen[:, :2])
    key = xy.view(np.dtype((np.void, 16))).ravel()
    _, first, col_of_cell = np.unique(key, return_index=True, return_inverse=True)
    cen = xy[first]
    ncol = cen.shape[0]
    lo, hi = cen.min(axis=0), cen.max(axis=0)
    ij = np.clip(((cen - lo) / np.maximum(hi - lo, 1e-300) * grid).astype(np.int64), 0, grid - 1)
    # serpentine row order keeps consecutive triangles close
    row = ij[:, 1]
    col = np.where(row % 2 == 0, ij[:, 0], grid - 1 - ij[:, 0])
    order = np.lexsort((col, row))
    blk_of_col = np.empty(ncol, dtype=np.int32)
    blk_of_col[order] = np.arange(ncol) // per_block
    return blk_of_col[col_of_cell.ravel()].astype(np.int32)


def rcm_triangle_order(triangles: np.ndarray, npoints: int) -> np.ndarray:
    """Triangles in reverse Cuthill-McKee order of their vertex-sharing graph
    (the horizontal pattern of the Jacobian): entry i is the triangle placed
    i-th.  Neighbours end up close in the numbering, which bounds the column
    offsets of the Jacobian rows."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    T = np.asarray(triangles, dtype=np.int64)
    nt = T.shape[0]
    inc = sp.csr_matrix((np.ones(3 * nt), (np.repeat(np.arange(nt), 3), T.ravel())), shape=(nt, npoints))
    adj = (inc @ inc.T).tocsr()
    return np.asarray(reverse_cuthill_mckee(adj, symmetric_mode=True), dtype=np.int64)


def renumber_cells(mesh: Mesh, new_of_old: np.ndarray) -> Mesh:
    """The same mesh with cell c renamed new_of_old[c] (faces, vertices and
    boundary indices unchanged)."""
    new_of_old = np.asarray(new_of_old, dtype=np.int64)
    old_of_new = np.empty_like(new_of_old)
    old_of_new[new_of_old] = np.arange(new_of_old.shape[0])
    fc = mesh.face_cells.copy()
    live = fc >= 0
    fc[live] = new_of_old[fc[live]]
    return Mesh(nx=mesh.nx, ny=mesh.ny, nz=mesh.nz, vert_xyz=mesh.vert_xyz,
                cell_verts=mesh.cell_verts[old_of_new], cell_faces=mesh.cell_faces[old_of_new],
                face_verts=mesh.face_verts, face_cells=fc.astype(np.int32), cell_cen=mesh.cell_cen[old_of_new],
                cell_vol=mesh.cell_vol[old_of_new], face_cen=mesh.face_cen, face_area=mesh.face_area,
                face_normal=mesh.face_normal, bnd_face=mesh.bnd_face, face_dir=mesh.face_dir,
                face_side=mesh.face_side)


def polyhedral_mesh(V: np.ndarray, cell_verts: np.ndarray, cell_faces: np.ndarray, face_verts: np.ndarray,
                    face_cells: np.ndarray) -> Mesh:
    """Geometry and boundary bookkeeping of a mesh of hexahedra and prisms
    given by connectivity (arrays padded with -1; face_cells = (c0, c1) with
    c1 = -1 on the boundary).  Faces: triangles exact, quadrilaterals by the
    four-triangle fan of ``box_mesh``; cells: volume and centroid by fan
    tetrahedra about the vertex mean (exact for planar faces); boundary faces
    stored as (-(b+1), cell), normals from up to dn."""
    V = np.asarray(V, dtype=np.float64)
    nc, nf = cell_verts.shape[0], face_verts.shape[0]
    nvf = (face_verts >= 0).sum(axis=1)
    fnrm = np.zeros((nf, 3)); farea = np.zeros(nf); fcen = np.zeros((nf, 3))
    fsvec = np.zeros((nf, 3)); fmean = np.zeros((nf, 3))
    q = nvf == 4
    if q.any():
        n4, a4, c4, s4, m4 = _face_geometry(V[face_verts[q]])
        fnrm[q], farea[q], fcen[q], fsvec[q], fmean[q] = n4, a4, c4, s4, m4
    t = nvf == 3
    if t.any():
        X = V[face_verts[t, :3]]
        cr = _cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0])
        ln = np.sqrt((cr * cr).sum(axis=1))
        fnrm[t] = cr / ln[:, None]; farea[t] = 0.5 * ln; fcen[t] = X.mean(axis=1)
        fsvec[t] = 0.5 * cr; fmean[t] = fcen[t]
    vol = np.zeros(nc); cen = np.zeros((nc, 3))
    for c in range(nc):
        vs = cell_verts[c][cell_verts[c] >= 0]
        p0 = V[vs].mean(axis=0)
        for f in cell_faces[c][cell_faces[c] >= 0]:
            fv = face_verts[f][face_verts[f] >= 0]
            mf = fmean[f]
            sgn = np.sign(fsvec[f] @ (mf - p0))
            for i in range(len(fv)):
                a, b = V[fv[i]], V[fv[(i + 1) % len(fv)]]
                tv = sgn * ((mf - p0) @ _cross(a - p0, b - p0)) / 6.0
                vol[c] += tv
                cen[c] += tv * (p0 + mf + a + b) / 4.0
        cen[c] /= vol[c]
    fc = np.array(face_cells, dtype=np.int32).copy()
    bmask = fc[:, 1] < 0
    bnd_face = np.nonzero(bmask)[0].astype(np.int32)
    fc[bnd_face, 1] = fc[bnd_face, 0]
    fc[bnd_face, 0] = -(np.arange(bnd_face.shape[0], dtype=np.int32) + 1)
    up, dn = fc[:, 0], fc[:, 1]
    ref_from = np.where(up[:, None] >= 0, cen[np.maximum(up, 0)], fcen)
    s = np.sign(((cen[dn] - ref_from) * fnrm).sum(axis=1))
    s[s == 0] = 1.0
    fnrm = fnrm * s[:, None]
    face_side = np.where(bmask, 0, -1).astype(np.int32)
    return Mesh(nx=nc, ny=1, nz=1, vert_xyz=V, cell_verts=cell_verts.astype(np.int32),
                cell_faces=cell_faces.astype(np.int32), face_verts=face_verts.astype(np.int32), face_cells=fc,
                cell_cen=cen, cell_vol=vol, face_cen=fcen, face_area=farea, face_normal=fnrm, bnd_face=bnd_face,
                face_dir=np.zeros(nf, dtype=np.int32), face_side=face_side)
