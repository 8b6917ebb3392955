"""Fused evaluation of many requests over the C ABI.

``Fuser`` (tdyhip_fuser_*): contexts attached to it keep their own solver
threads and pace; their SNES-form residual and Jacobian evaluations are
launched by the fuser's thread, all evaluations pending at once as one
segmented launch per kernel.  This is what the ensemble server uses.

``FusedBatch`` (tdyhip_batch_*), lockstep:

Several requests - MPFAO contexts of hexahedral meshes with their block-Jacobi
partitions set, device states and time-stepping parameters - advance in
lockstep, one Newton iteration per ``round()``: the state, residual and
Jacobian evaluations of all of them run as one segmented launch each, their
Krylov solves side by side on the contexts' own streams.  Each request's
result is the same bit for bit alone, in any mix, and through
``MPFAO.run_tdyti`` (TDyTimeIntegratorRunToTime, src/tdyti.c:100-196).
"""
from __future__ import annotations

import ctypes as C

from . import _lib

MAX_REQUESTS = 32


class Fuser:
    def __init__(self, window_us: int = 0):
        """window_us: how long the first pending evaluation waits for others
        (or until every attached context has one pending)."""
        h = C.c_void_p()
        _lib.check(_lib.load().tdyhip_fuser_create(C.byref(h)))
        self._h = h
        if window_us:
            _lib.check(_lib.load().tdyhip_fuser_set_window(self._h, int(window_us)))

    def attach(self, g):
        _lib.check(_lib.load().tdyhip_set_fuser(g._h, self._h))

    @staticmethod
    def detach(g):
        _lib.check(_lib.load().tdyhip_set_fuser(g._h, None))

    def stats(self):
        """{'launches': fused launch groups, 'evaluations', 'largest_group'}."""
        a = (C.c_int64 * 3)()
        _lib.check(_lib.load().tdyhip_fuser_stats(self._h, a))
        return {"launches": int(a[0]), "evaluations": int(a[1]), "largest_group": int(a[2])}

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.check(_lib.load().tdyhip_fuser_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FusedBatch:
    def __init__(self):
        h = C.c_void_p()
        _lib.check(_lib.load().tdyhip_batch_create(C.byref(h)))
        self._h = h
        self._members = {}  # id -> (context, state, options): kept alive while in the batch

    def add(self, g, U, opts, dt_init, growth, dt_max, final_time) -> int:
        """Add a request: U (float64 device tensor) is advanced in place."""
        rid = C.c_int32()
        _lib.check(_lib.load().tdyhip_batch_add(self._h, g._h, C.c_void_p(U.data_ptr()), C.byref(opts),
                                                float(dt_init), float(growth), float(dt_max), float(final_time),
                                                C.byref(rid)))
        self._members[rid.value] = (g, U, opts)
        return rid.value

    def remove(self, rid: int):
        _lib.check(_lib.load().tdyhip_batch_remove(self._h, int(rid)))
        self._members.pop(rid, None)

    def __len__(self):
        n = C.c_int32()
        _lib.check(_lib.load().tdyhip_batch_size(self._h, C.byref(n)))
        return n.value

    def round(self):
        """One round; returns [(id, rc, stats dict)] of the requests that left
        the batch (rc 0: reached the final time; else the error code)."""
        cap = max(1, len(self._members))
        ids = (C.c_int32 * cap)()
        rcs = (C.c_int32 * cap)()
        sts = (_lib.SolverStats * cap)()
        nd = C.c_int32()
        _lib.check(_lib.load().tdyhip_batch_round(self._h, cap, ids, rcs, sts, C.byref(nd)))
        out = []
        for k in range(nd.value):
            st = sts[k]
            out.append((ids[k], rcs[k], {"newton_its": st.newton_its, "linear_its": st.linear_its,
                                         "steps": st.steps, "cuts": st.cuts, "status": st.status,
                                         "krylov_fallbacks": st.krylov_fallbacks, "last_fnorm": st.last_fnorm}))
            self._members.pop(ids[k], None)
        return out

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.check(_lib.load().tdyhip_batch_destroy(self._h))
            self._h = C.c_void_p()
        self._members.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
