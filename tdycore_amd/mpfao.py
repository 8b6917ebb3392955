"""Host-side mirror of TDycore's MPFA-O Richards interface over the HIP library.

``MPFAO`` plays the role of a ``TDy`` set up with ``TDySetMode(RICHARDS)`` and
``TDySetDiscretization(MPFA_O)``: construction is ``TDySetup``
(src/fv/mpfao/tdympfao.c:2277), and the methods are the callbacks PETSc's
TS/SNES invoke:

===========================  =====================================================
method                       reference
===========================  =====================================================
``IFunction(t,U,U_t,R)``     TDyMPFAOIFunction   src/fv/mpfao/tdympfao_ts.c:153
``IJacobian(t,U,U_t,a,J)``   TDyMPFAOIJacobian   src/fv/mpfao/tdympfao_ts.c:520
``SetDtime(dt)``             TDySetDtimeForSNESSolver  src/tdycore.c:1575
``SNESPreSolve(Uprev)``      TDyMPFAOSNESPreSolve  src/fv/mpfao/tdympfao_snes.c:32
``SNESFunction(U,R)``        TDyMPFAOSNESFunction  src/fv/mpfao/tdympfao_snes.c:67
``SNESJacobian(U,J)``        TDyMPFAOSNESJacobian  src/fv/mpfao/tdympfao_snes.c:143
===========================  =====================================================

Vectors are float64 torch tensors on the GPU; ``J`` holds the values of the
vertex-adjacency CSR pattern (``jacobian_pattern()``).  Like the reference,
``IJacobian`` uses the state of the preceding ``IFunction`` call.
Errors raise ``TdyHipError`` carrying the library's message (the role of a
nonzero PetscErrorCode).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from .meshgen import Mesh

DENSITY = {"constant": 0, "exponential": 1}
GMATRIX = {"mpfao": 0, "tpf": 1}


def _ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())




class MPFAO:
    def __init__(self, mesh: Mesh, face_bc, K, porosity, Sr, *, vg_m=0.8, vg_alpha=1e-4,
                 poly=(0.995, 0.99, 1.0, 0.005), gravity=9.8068, density="constant",
                 gmatrix="mpfao", pref=101325.0, p_bnd_fn: Optional[np.ndarray] = None):
        L = _lib.load()
        self.mesh = mesh
        self.nc, self.nf, self.nbnd = mesh.nc, mesh.nf, mesh.nbnd
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a.ctypes.data

        ms = _lib.Mesh(mesh.nc, mesh.nf, mesh.nv, mesh.nbnd,
                       arr(mesh.cell_verts, np.int32), arr(mesh.cell_faces, np.int32),
                       arr(mesh.face_verts, np.int32), arr(mesh.face_cells, np.int32),
                       arr(face_bc, np.int32), arr(mesh.vert_xyz, np.float64),
                       arr(mesh.cell_cen, np.float64), arr(mesh.cell_vol, np.float64),
                       arr(mesh.face_cen, np.float64), arr(mesh.face_area, np.float64),
                       arr(mesh.face_normal, np.float64))
        mt = _lib.Materials(vg_m, vg_alpha, poly[0], poly[1], poly[2], poly[3], gravity, pref,
                            DENSITY[density], GMATRIX[gmatrix], arr(K, np.float64),
                            arr(porosity, np.float64), arr(Sr, np.float64),
                            None if p_bnd_fn is None else arr(p_bnd_fn, np.float64))
        h = C.c_void_p()
        _lib.check(L.tdyhip_create(C.byref(ms), C.byref(mt), C.byref(h)))
        self._h = h
        # The context works on the stream it created for itself (never a
        # torch pool stream: torch.cuda.Stream() hands out pooled streams that
        # another context could be using, and the solver captures its launches
        # on this stream into graphs).  Every call is ordered after the
        # caller's current stream and the caller's stream after the call.
        import torch
        hs = C.c_void_p()
        _lib.check(L.tdyhip_get_stream(self._h, C.byref(hs)))
        self._stream = torch.cuda.ExternalStream(hs.value)
        sz = np.zeros(4, np.int64)
        _lib.check(L.tdyhip_sizes(self._h, sz.ctypes.data))
        self.jac_nnz = int(sz[1])

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().tdyhip_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @property
    def stream(self):
        """The context's own stream (a torch.cuda.ExternalStream)."""
        return self._stream

    def _run(self, fn, *args):
        """fn(ctx, *args, stream=NULL -> the context's stream), ordered after
        and before the caller's current torch stream."""
        import torch
        cur = torch.cuda.current_stream()
        self._stream.wait_stream(cur)
        _lib.check(fn(self._h, *args))
        cur.wait_stream(self._stream)

    # ---- pattern / diagnostics ----
    def jacobian_pattern(self):
        rp = np.zeros(self.nc + 1, np.int32)
        cols = np.zeros(self.jac_nnz, np.int32)
        _lib.check(_lib.load().tdyhip_jacobian_pattern(self._h, rp.ctypes.data, cols.ctypes.data))
        return rp, cols

    def trans_csr(self):
        L = _lib.load()
        nnz = C.c_int64()
        _lib.check(L.tdyhip_trans_nnz(self._h, C.byref(nnz)))
        rp = np.zeros(4 * self.nf + 1, np.int32)
        cols = np.zeros(nnz.value, np.int32)
        vals = np.zeros(nnz.value, np.float64)
        _lib.check(L.tdyhip_trans_csr(self._h, rp.ctypes.data, cols.ctypes.data, vals.ctypes.data))
        return rp, cols, vals

    def gravdis(self):
        out = np.zeros(4 * self.nf)
        _lib.check(_lib.load().tdyhip_gravdis(self._h, out.ctypes.data))
        return out

    def setup_info(self):
        """Device times (ms) of the setup phases and the interaction-region
        counts: {'gmatrix_ms', 'local_systems_ms', 'gravity_ms', 'buckets'
        (regions per 16/32/64-lane bucket; hexahedral meshes: all in the
        first), 'max_cells', 'max_faces', 'trans_nnz', 'local_systems_flops',
        'local_systems_bytes', 'lanes' (team width used per bucket)}."""
        ms = np.zeros(3)
        cnt = np.zeros(11, np.int64)
        _lib.check(_lib.load().tdyhip_setup_info(self._h, ms.ctypes.data, cnt.ctypes.data))
        return {"gmatrix_ms": float(ms[0]), "local_systems_ms": float(ms[1]), "gravity_ms": float(ms[2]),
                "buckets": [int(x) for x in cnt[:3]], "max_cells": int(cnt[3]), "max_faces": int(cnt[4]),
                "trans_nnz": int(cnt[5]), "local_systems_flops": int(cnt[6]), "local_systems_bytes": int(cnt[7]),
                "lanes": [int(x) for x in cnt[8:11]]}

    def setup_plan(self):
        """The autotuned interaction-region launch: {'lanes' (team width per
        bucket), 'largest_first', 'chunk', 'width_ms' (3x3 tuning runs,
        bucket x width 16/32/64, 0 = not run), 'schedule_ms' (order
        smallest/largest first x chunk 1/8)}."""
        plan = np.zeros(5, np.int32)
        ms = np.zeros(13)
        _lib.check(_lib.load().tdyhip_setup_plan(self._h, plan.ctypes.data, ms.ctypes.data))
        return {"lanes": [int(x) for x in plan[:3]], "largest_first": bool(plan[3]), "chunk": int(plan[4]),
                "width_ms": ms[:9].reshape(3, 3).tolist(), "schedule_ms": ms[9:].tolist()}

    # ---- TS callbacks ----
    def IFunction(self, t, U, U_t, R):
        self._run(_lib.load().tdyhip_ifunction, float(t), _ptr(U), _ptr(U_t), _ptr(R), None)

    def IFunctionNoAlloc(self, t, U, U_t, R):
        """IFunction with its flux-row scratch in the library's fixed per-GPU
        array instead of the context's buffer: allocates nothing; meshes of at
        most 2,097,152 faces; calls that overlap on one GPU race."""
        self._run(_lib.load().tdyhip_ifunction_noalloc, float(t), _ptr(U), _ptr(U_t), _ptr(R), None)

    def IJacobian(self, t, U, U_t, shift, J):
        self._run(_lib.load().tdyhip_ijacobian, float(t), _ptr(U), _ptr(U_t), float(shift), _ptr(J), None)

    def IFunctionIJacobian(self, t, U, U_t, shift, R, J):
        self._run(_lib.load().tdyhip_ifunction_ijacobian, float(t), _ptr(U), _ptr(U_t), float(shift), _ptr(R),
                  _ptr(J), None)

    def SetSourceSink(self, Q=None):
        """Per-cell source_sink (None: zero), the reference's mpfao->source_sink:
        IFunction subtracts Q*V (tdympfao_ts.c:211), SNESFunction Q
        (tdympfao_snes.c:122)."""
        L = _lib.load()
        if Q is None:
            _lib.check(L.tdyhip_set_source_sink(self._h, None))
        else:
            q = np.ascontiguousarray(Q, dtype=np.float64)
            if q.shape != (self.nc,):
                raise ValueError("source_sink needs one value per cell")
            _lib.check(L.tdyhip_set_source_sink(self._h, q.ctypes.data))

    # ---- SNES callbacks ----
    def SetDtime(self, dt):
        _lib.check(_lib.load().tdyhip_set_dtime(self._h, float(dt)))

    def SNESPreSolve(self, Uprev):
        self._run(_lib.load().tdyhip_snes_presolve, _ptr(Uprev), None)

    def SNESFunction(self, U, R):
        self._run(_lib.load().tdyhip_snes_function, _ptr(U), _ptr(R), None)

    def SNESJacobian(self, U, J):
        self._run(_lib.load().tdyhip_snes_jacobian, _ptr(U), _ptr(J), None)

    def Diagnostics(self, U):
        """(S, liquid mass rho*V, water mass rho*phi*S*V) per cell at U
        (TDyUpdateDiagnostics_MPFAO, src/fv/mpfao/tdympfao.c:2759-2786)."""
        import torch
        S, M, W = (torch.empty_like(U) for _ in range(3))
        self._run(_lib.load().tdyhip_diagnostics, _ptr(U), _ptr(S), _ptr(M), _ptr(W), None)
        return S, M, W

    def RecoverVelocity(self, U):
        """(vel, count) per face at U (TDyMPFAORecoverVelocity,
        src/fv/mpfao/tdympfao_utils.c:661-688): vel[f] sums -(T_row . Phi)/A_f over
        the face's flux rows with Phi = p + rho |g| z, count[f] is the number of
        rows summed."""
        import torch
        vel = torch.empty(self.nf, dtype=torch.float64, device=U.device)
        cnt = torch.empty(self.nf, dtype=torch.int32, device=U.device)
        self._run(_lib.load().tdyhip_recover_velocity, _ptr(U), _ptr(vel), _ptr(cnt), None)
        return vel, cnt

    def device_bytes(self) -> int:
        """Device memory held by this context (setup arrays + solver workspace)."""
        b = C.c_int64()
        _lib.check(_lib.load().tdyhip_device_bytes(self._h, C.byref(b)))
        return b.value

    # ---- per-kernel timing ----
    def kernel_timers(self, enable: bool):
        """Start (reset) or stop HIP-event timing of every kernel launch."""
        _lib.check(_lib.load().tdyhip_kernel_timers(self._h, 1 if enable else 0))

    def kernel_timer_report(self):
        """List of {kernel, launches, ms, bytes} since timers were enabled."""
        import json
        buf = C.create_string_buffer(1 << 16)
        _lib.check(_lib.load().tdyhip_kernel_timer_report(self._h, buf, len(buf)))
        return json.loads(buf.value.decode())

    # ---- Newton / time loops on the device ----
    def set_pc_partition(self, block_of_cell=None):
        """Block-Jacobi blocks (block id per cell, each block <= pc_block cells);
        None restores the breadth-first default."""
        L = _lib.load()
        if block_of_cell is None:
            _lib.check(L.tdyhip_set_pc_partition(self._h, None))
        else:
            a = np.ascontiguousarray(block_of_cell, dtype=np.int32)
            _lib.check(L.tdyhip_set_pc_partition(self._h, a.ctypes.data))

    @staticmethod
    def solver_options(**kw):
        """PETSc-default solver options (SNES rtol 1e-8, GMRES(30), KSP rtol 1e-5,
        block-Jacobi ILU(0)); keyword overrides by field name."""
        o = _lib.SolverOpts()
        _lib.load().tdyhip_solver_defaults(C.byref(o))
        for k, v in kw.items():
            if not hasattr(o, k):
                raise TypeError(f"unknown solver option {k}")
            setattr(o, k, v)
        return o

    @staticmethod
    def _stats_dict(st):
        return {"newton_its": st.newton_its, "linear_its": st.linear_its, "steps": st.steps,
                "cuts": st.cuts, "status": st.status, "krylov_fallbacks": st.krylov_fallbacks,
                "last_fnorm": st.last_fnorm}

    def newton_snes(self, U, opts=None):
        o = opts or self.solver_options()
        st = _lib.SolverStats()
        self._run(_lib.load().tdyhip_newton_snes, _ptr(U), C.byref(o), C.byref(st))
        return self._stats_dict(st)

    def run_snes_steps(self, U, dt, nsteps, opts=None):
        o = opts or self.solver_options()
        st = _lib.SolverStats()
        self._run(_lib.load().tdyhip_run_snes_steps, _ptr(U), float(dt), int(nsteps), C.byref(o), C.byref(st))
        return self._stats_dict(st)

    def run_ts_be(self, U, dt, nsteps, opts=None):
        o = opts or self.solver_options()
        st = _lib.SolverStats()
        self._run(_lib.load().tdyhip_run_ts_be, _ptr(U), float(dt), int(nsteps), C.byref(o), C.byref(st))
        return self._stats_dict(st)

    def run_tdyti(self, U, dt_init, growth, dt_max, final_time, opts=None):
        o = opts or self.solver_options()
        st = _lib.SolverStats()
        self._run(_lib.load().tdyhip_run_tdyti, _ptr(U), float(dt_init), float(growth), float(dt_max),
                  float(final_time), C.byref(o), C.byref(st))
        return self._stats_dict(st)

    def cancel(self, on: bool = True):
        """Ask the solve running on this context (called from another thread)
        to stop: it raises TdyHipError with code ERR_CANCELLED at its next
        Newton iteration or GMRES restart.  Stays set until cancel(False)."""
        if getattr(self, "_h", None) and self._h.value:
            _lib.check(_lib.load().tdyhip_cancel(self._h, 1 if on else 0))
