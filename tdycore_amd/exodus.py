"""Exodus II meshes (classic netCDF files) in the form the MPFA-O path consumes.

Reads the element blocks of hexahedra and wedges (``connect*``) and the
coordinates of a classic netCDF (CDF-1 / CDF-2) Exodus file, then builds
faces, orientations and geometry as ``meshgen`` does.  This is the mesh input
of the reference's ``DMPlexCreateFromFile`` path
(demo/transient/transient_snes_mpfaof90.F90:149): cells keep the file's
element order, so regression samples index the same cells.
"""
from __future__ import annotations

import struct

import numpy as np

from . import meshgen as mg

_TYPES = {1: ("b", 1), 2: ("c", 1), 3: ("h", 2), 4: ("i", 4), 5: ("f", 4), 6: ("d", 8)}


class _Reader:
    def __init__(self, data: bytes):
        self.d = data
        self.p = 0

    def u32(self):
        v = struct.unpack_from(">i", self.d, self.p)[0]
        self.p += 4
        return v

    def u64(self):
        v = struct.unpack_from(">q", self.d, self.p)[0]
        self.p += 8
        return v

    def name(self):
        n = self.u32()
        s = self.d[self.p:self.p + n].decode("latin-1")
        self.p += (n + 3) & ~3
        return s

    def skip_values(self, t, n):
        self.p += (n * _TYPES[t][1] + 3) & ~3


def read_netcdf(path: str) -> dict:
    """Non-record variables of a classic netCDF file: {name: ndarray}."""
    data = open(path, "rb").read()
    if data[:3] != b"CDF" or data[3] not in (1, 2):
        raise ValueError(f"{path}: not a classic netCDF file")
    r = _Reader(data)
    r.p = 4
    offset64 = data[3] == 2
    r.u32()  # numrecs
    dims = []
    tag, n = r.u32(), r.u32()
    for _ in range(n if tag == 0x0A else 0):
        dims.append((r.name(), r.u32()))

    def atts():
        tag, n = r.u32(), r.u32()
        for _ in range(n if tag == 0x0C else 0):
            r.name()
            t, cnt = r.u32(), r.u32()
            r.skip_values(t, cnt)

    atts()
    out = {}
    tag, n = r.u32(), r.u32()
    for _ in range(n if tag == 0x0B else 0):
        nm = r.name()
        nd = r.u32()
        ids = [r.u32() for _ in range(nd)]
        atts()
        t = r.u32()
        r.u32()  # vsize
        begin = r.u64() if offset64 else r.u32()
        shape = tuple(dims[i][1] for i in ids)
        if any(s == 0 for s in shape):
            continue  # record variables are not needed here
        code, size = _TYPES[t]
        count = int(np.prod(shape)) if shape else 1
        if t == 2:
            continue
        arr = np.frombuffer(data, dtype=">" + code, count=count, offset=begin)
        out[nm] = arr.reshape(shape).astype(np.float64 if t in (5, 6) else np.int64)
    return out


# local faces of the Exodus element types (0-based node positions), each
# listed counter-clockwise seen from outside
_HEX_FACES = [(0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (0, 4, 7, 3), (0, 3, 2, 1), (4, 5, 6, 7)]
_WEDGE_FACES = [(0, 1, 4, 3), (1, 2, 5, 4), (0, 3, 5, 2), (0, 2, 1), (3, 4, 5)]


def _coords(v: dict) -> np.ndarray:
    if "coord" in v:
        return v["coord"].T.copy()
    cols = [v["coordx"], v["coordy"]] + ([v["coordz"]] if "coordz" in v else [])
    return np.stack(cols, 1)


def _elements(v: dict) -> list:
    out = []
    for k in sorted((k for k in v if k.startswith("connect")), key=lambda k: int(k[7:])):
        out.extend(list(e) for e in v[k] - 1)
    return out


def connectivity_mesh(X: np.ndarray, cells: list) -> mg.Mesh:
    """Mesh of hexahedra (8 nodes) and wedges (6 nodes) in Exodus node order:
    faces are found by their vertex sets, numbered in order of first use, and
    the geometry follows from ``meshgen.polyhedral_mesh``."""
    nc = len(cells)
    cell_verts = np.full((nc, 8), -1, dtype=np.int32)
    faces, face_cells, cell_faces = {}, [], np.full((nc, 6), -1, dtype=np.int32)
    face_verts = []
    for c, e in enumerate(cells):
        if len(e) == 8:
            loc = _HEX_FACES
        elif len(e) == 6:
            loc = _WEDGE_FACES
        else:
            raise ValueError("only hexahedra and wedges are supported")
        cell_verts[c, :len(e)] = e
        for lf, fl in enumerate(loc):
            fv = tuple(int(e[i]) for i in fl)
            key = tuple(sorted(fv))
            if key in faces:
                f = faces[key]
                if face_cells[f][1] >= 0:
                    raise ValueError("a face shared by more than two cells")
                face_cells[f][1] = c
            else:
                f = len(face_verts)
                faces[key] = f
                face_verts.append(fv)
                face_cells.append([c, -1])
            cell_faces[c, lf] = f
    nf = len(face_verts)
    FV = np.full((nf, 4), -1, dtype=np.int32)
    for f, fv in enumerate(face_verts):
        FV[f, :len(fv)] = fv
    FC = np.array(face_cells, dtype=np.int32)
    return mg.polyhedral_mesh(np.asarray(X, dtype=np.float64), cell_verts, cell_faces, FV, FC)


def exodus_mesh(path: str) -> mg.Mesh:
    """Hexahedra and wedges of an Exodus file as a ``meshgen.Mesh`` (arrays
    padded with -1 for wedges and triangles).  Only the default boundary type
    applies to such a mesh (``face_bc(default=...)``)."""
    v = read_netcdf(path)
    X = _coords(v)
    if X.shape[1] == 2:
        raise ValueError(f"{path}: a surface mesh; use extruded_exodus_mesh")
    return connectivity_mesh(X, _elements(v))


def extruded_exodus_mesh(path: str, layers: int, height: float = 1.0) -> mg.Mesh:
    """A surface mesh of quadrilaterals and triangles extruded along +z into
    ``layers`` equal layers of total ``height`` (PETSc's DMPlexExtrude, as the
    reference drives it with -dm_plex_extrude_layers,
    demo/transient/transient_snes_mpfaof90.cfg:35).  Surface element c gives
    cells c*layers + l, l counted from the bottom, as DMPlexExtrude numbers
    them."""
    v = read_netcdf(path)
    X2 = _coords(v)[:, :2]
    ns = X2.shape[0]
    z = np.arange(layers + 1, dtype=np.float64) * (height / layers)
    X = np.concatenate([np.column_stack([X2, np.full(ns, zz)]) for zz in z])
    cells = []
    for e in _elements(v):
        if len(e) not in (3, 4):
            raise ValueError("surface elements must be triangles or quadrilaterals")
        for l in range(layers):
            cells.append([int(i) + l * ns for i in e] + [int(i) + (l + 1) * ns for i in e])
    return connectivity_mesh(X, cells)
