"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg use
this module; the shipped library never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(path)
        P = C.c_void_p
        L.ora_create.restype = P
        L.ora_create.argtypes = [C.c_int] * 4 + [P] * 16
        L.ora_destroy.argtypes = [P]
        L.ora_last_error.restype = C.c_char_p
        L.ora_sizes.argtypes = [P, P]
        L.ora_trans_csr.argtypes = [P, P, P, P]
        L.ora_gravdis.argtypes = [P, P]
        L.ora_jac_pattern.argtypes = [P, P, P]
        L.ora_update_state.argtypes = [P, P]
        L.ora_get_state.argtypes = [P, C.c_int, P]
        L.ora_boundary_pressure.argtypes = [P, P]
        L.ora_ifunction.argtypes = [P, P, P, P]
        L.ora_ijacobian.argtypes = [P, P, P, C.c_double, P]
        L.ora_set_dtime.argtypes = [P, C.c_double]
        L.ora_snes_presolve.argtypes = [P, P]
        L.ora_snes_function.argtypes = [P, P, P]
        L.ora_snes_jacobian.argtypes = [P, P, P]
        L.ora_run_ts_be.argtypes = [P, P, C.c_double, C.c_int, C.c_int, C.c_double, P]
        L.ora_run_snes_steps.argtypes = [P, P, C.c_double, C.c_int, C.c_int, C.c_double, P]
        L.ora_set_newton_maxit.argtypes = [P, C.c_int]
        L.ora_run_tdyti.argtypes = [P, P, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                    C.c_double, P]
        L.ora_newton_snes.argtypes = [P, P, C.c_int, C.c_double, P]
        L.ora_liquid_mass.argtypes = [P, P, P]
        L.ora_set_source_sink.argtypes = [P, P]
        L.ora_set_threads.argtypes = [P, C.c_int]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


_LIN = {"lu": 0, "gmres": 1, "gmres_bjacobi": 2}

STATE = {"S": 0, "dS_dP": 1, "d2S_dP2": 2, "Kr": 3, "dKr_dS": 4, "rho": 5, "drho_dP": 6,
         "d2rho_dP2": 7, "vis": 8}


class Oracle:
    """One MPFA-O Richards problem on the CPU (reference algorithm, FP64)."""

    def __init__(self, mesh, face_bc, K, porosity, Sr, *, vg_m=0.8, vg_alpha=1e-4,
                 poly=(0.995, 0.99, 1.0, 0.005), gravity=9.8068, density="constant",
                 gmatrix="mpfao", pref=101325.0, p_bnd_fn: Optional[np.ndarray] = None):
        L = lib()
        self.mesh = mesh
        self.nc, self.nf = mesh.nc, mesh.nf
        self._keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self._keep.append(a)
            return a

        params = arr([vg_m, vg_alpha, poly[0], poly[1], poly[2], poly[3], gravity,
                      0.0 if density == "constant" else 1.0,
                      0.0 if gmatrix == "mpfao" else 1.0, pref], np.float64)
        pb = None if p_bnd_fn is None else arr(p_bnd_fn, np.float64)
        h = L.ora_create(mesh.nc, mesh.nf, mesh.nv, mesh.nbnd,
                         _p(arr(mesh.cell_verts, np.int32)), _p(arr(mesh.cell_faces, np.int32)),
                         _p(arr(mesh.face_verts, np.int32)), _p(arr(mesh.face_cells, np.int32)),
                         _p(arr(mesh.vert_xyz, np.float64)), _p(arr(mesh.cell_cen, np.float64)),
                         _p(arr(mesh.cell_vol, np.float64)), _p(arr(mesh.face_cen, np.float64)),
                         _p(arr(mesh.face_area, np.float64)), _p(arr(mesh.face_normal, np.float64)),
                         _p(arr(face_bc, np.int32)), _p(arr(K, np.float64)),
                         _p(arr(porosity, np.float64)), _p(arr(Sr, np.float64)), _p(params),
                         None if pb is None else _p(pb))
        if not h:
            raise RuntimeError("oracle setup failed: " + L.ora_last_error().decode())
        self.h = C.c_void_p(h)
        sz = np.zeros(4, dtype=np.int64)
        L.ora_sizes(self.h, _p(sz))
        self.trans_nnz, self.jac_nnz, self.max_ir_cells, self.max_ir_faces = (int(x) for x in sz)
        self._jrp = np.zeros(self.nc + 1, np.int32)
        self._jcols = np.zeros(self.jac_nnz, np.int32)
        L.ora_jac_pattern(self.h, _p(self._jrp), _p(self._jcols))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().ora_destroy(self.h)
                self.h = None
        except Exception:
            pass

    # ---- setup products ----
    def trans_csr(self):
        rp = np.zeros(4 * self.nf + 1, np.int32)
        cols = np.zeros(self.trans_nnz, np.int32)
        vals = np.zeros(self.trans_nnz, np.float64)
        lib().ora_trans_csr(self.h, _p(rp), _p(cols), _p(vals))
        return rp, cols, vals

    def trans_matrix(self):
        import scipy.sparse as sp
        rp, cols, vals = self.trans_csr()
        return sp.csr_matrix((vals, cols, rp), shape=(4 * self.nf, self.nc + self.mesh.nbnd))

    def gravdis(self):
        out = np.zeros(4 * self.nf)
        lib().ora_gravdis(self.h, _p(out))
        return out

    def jac_pattern(self):
        return self._jrp.copy(), self._jcols.copy()

    def jac_to_csr(self, vals):
        import scipy.sparse as sp
        return sp.csr_matrix((vals, self._jcols, self._jrp), shape=(self.nc, self.nc))

    # ---- state ----
    def update_state(self, p):
        p = np.ascontiguousarray(p, np.float64)
        lib().ora_update_state(self.h, _p(p))

    def state(self, name):
        out = np.zeros(self.nc)
        lib().ora_get_state(self.h, STATE[name], _p(out))
        return out

    def boundary_pressure(self):
        out = np.zeros(self.mesh.nbnd)
        lib().ora_boundary_pressure(self.h, _p(out))
        return out

    # ---- TS form ----
    def ifunction(self, U, Udot):
        U = np.ascontiguousarray(U, np.float64)
        Ud = np.ascontiguousarray(Udot, np.float64)
        R = np.zeros(self.nc)
        lib().ora_ifunction(self.h, _p(U), _p(Ud), _p(R))
        return R

    def ijacobian(self, U, Udot, shift):
        U = np.ascontiguousarray(U, np.float64)
        Ud = np.ascontiguousarray(Udot, np.float64)
        J = np.zeros(self.jac_nnz)
        lib().ora_ijacobian(self.h, _p(U), _p(Ud), float(shift), _p(J))
        return J

    # ---- SNES form ----
    def set_dtime(self, dt):
        lib().ora_set_dtime(self.h, float(dt))

    def snes_presolve(self, Uprev):
        Uprev = np.ascontiguousarray(Uprev, np.float64)
        lib().ora_snes_presolve(self.h, _p(Uprev))

    def snes_function(self, U):
        U = np.ascontiguousarray(U, np.float64)
        R = np.zeros(self.nc)
        lib().ora_snes_function(self.h, _p(U), _p(R))
        return R

    def snes_jacobian(self, U):
        U = np.ascontiguousarray(U, np.float64)
        J = np.zeros(self.jac_nnz)
        lib().ora_snes_jacobian(self.h, _p(U), _p(J))
        return J

    # ---- drivers ----
    def run_ts_be(self, U, dt, nsteps, lin="lu", rtol=1e-8):
        U = np.array(U, dtype=np.float64, copy=True)
        st = np.zeros(4, np.int32)
        lib().ora_run_ts_be(self.h, _p(U), float(dt), int(nsteps), _LIN[lin], float(rtol), _p(st))
        return U, st

    def run_snes_steps(self, U, dt, nsteps, lin="lu", rtol=1e-8):
        U = np.array(U, dtype=np.float64, copy=True)
        st = np.zeros(4, np.int32)
        lib().ora_run_snes_steps(self.h, _p(U), float(dt), int(nsteps), _LIN[lin], float(rtol), _p(st))
        return U, st

    def run_tdyti(self, U, dt_init, growth, dt_max, final_time, lin="lu", rtol=1e-8):
        U = np.array(U, dtype=np.float64, copy=True)
        st = np.zeros(4, np.int32)
        lib().ora_run_tdyti(self.h, _p(U), float(dt_init), float(growth), float(dt_max), float(final_time),
                            _LIN[lin], float(rtol), _p(st))
        return U, st

    def set_threads(self, n):
        """Threads of every parallel loop (OpenMP)."""
        lib().ora_set_threads(self.h, int(n))

    def set_newton_maxit(self, maxit):
        lib().ora_set_newton_maxit(self.h, int(maxit))

    def newton_snes(self, U, lin="lu", rtol=1e-8):
        U = np.array(U, dtype=np.float64, copy=True)
        st = np.zeros(4, np.int32)
        s = lib().ora_newton_snes(self.h, _p(U), _LIN[lin], float(rtol), _p(st))
        return U, st, s

    def set_source_sink(self, Q=None):
        """Per-cell source_sink (None: zero): the TS residual subtracts Q*V, the
        SNES residual Q (tdympfao_ts.c:211, tdympfao_snes.c:122)."""
        if Q is None:
            lib().ora_set_source_sink(self.h, None)
        else:
            self._src = np.ascontiguousarray(Q, np.float64)
            lib().ora_set_source_sink(self.h, _p(self._src))

    def liquid_mass(self, U):
        U = np.ascontiguousarray(U, np.float64)
        out = np.zeros(self.nc)
        lib().ora_liquid_mass(self.h, _p(U), _p(out))
        return out


def regression_text(U, n_sample):
    """Regression sample in the reference format (src/tdyregression.c:239-262):
    sampled natural ids are floor(n/N)*i (src/tdyregression.c:67-89)."""
    U = np.asarray(U)
    n = U.shape[0]
    lines = ["-- PRESSURE: Liquid Pressure --",
             "      Max: %21.13e" % U.max(),
             "      Min: %21.13e" % U.min(),
             "     Mean: %21.13e" % (U.sum() / n)]
    step = n // n_sample
    for i in range(n_sample):
        lines.append("%9d: %21.13e" % (step * i, U[step * i]))
    return "\n".join(lines) + "\n"


def parse_gold(path):
    """Parse a reference .regression.gold file -> dict."""
    out = {"samples": {}}
    with open(path) as fh:
        for line in fh:
            line = line.strip()
            if not line or line.startswith("--"):
                continue
            k, v = line.split(":")
            k = k.strip()
            if k in ("Max", "Min", "Mean"):
                out[k] = float(v)
            else:
                out["samples"][int(k)] = float(v)
    return out
