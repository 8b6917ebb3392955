"""Continuous-batching server for ensembles of Richards solves on one GPU.

Requests (a box size, a permeability seed and sigma, a number of
backward-Euler steps) arrive at any time.  Each one goes through three
stages:

1. **build** - a host thread pool generates the mesh, the materials and the
   initial state.  This is the part of the reference that DMPlex and
   ``InitMaterials`` do (src/fv/mpfao/tdympfao.c:630-763);
2. **admit** - the request waits, in the order its build finished, until a GPU slot is free
   and its device footprint fits in the HBM budget beside the requests
   already on the device;
3. **solve** - a slot thread sets up the transmissibilities on the device and
   runs the backward-Euler steps to the request's final time with the
   time-integrator rules of ``TDyTimeIntegratorRunToTime`` (src/tdyti.c:71-196):
   a Newton failure cuts dt and retries.  Each step is a pre-solve plus a
   Newton solve: FP64 residual, FP32 or FP64 Krylov inner solve.  The request's state stays
   in HBM from setup to its last step.  Each context drives the
   non-blocking HIP stream it created for itself, so the setup and Newton kernels of different
   requests overlap on the device, and a request that arrives while others
   are solving starts as soon as a slot and memory free up.

Results are bit-for-bit those of the same request run alone: every kernel of
the path (residual and Jacobian gathers, Krylov reductions) sums in a fixed
order, whatever else runs on the device.

With ``fused=True`` every context is attached to one evaluation fuser
(``batch.Fuser``, the library's tdyhip_fuser_*): each slot thread still runs
its own request's solver at its own pace, and the SNES residual and Jacobian
evaluations of all slots that are pending at once are launched together, one
segmented launch per kernel.  Results are still bit for bit those of the
request alone.  (The library's lockstep batch, tdyhip_batch_*, advances all
requests one Newton iteration per round; on requests of very different
Krylov costs every round waits for the slowest, so the server does not use
it.)

A request may hold a slot for at most ``deadline_s`` seconds of solving: past
that a watchdog cancels its context (the solver stops at its next Newton
iteration or GMRES restart) and the request's future fails with
``DeadlineExceeded``.  Every failed future's exception carries the request's
``rid``, ``t_arrive``, ``t_done`` and ``latency_s``, so failures count in the
latency of responses.
"""
from __future__ import annotations

import concurrent.futures as cf
import dataclasses
import queue
import threading
import time
from typing import Callable, Optional

import numpy as np

from . import workloads
from .meshgen import brick_partition


class DeadlineExceeded(TimeoutError):
    """The request's solve ran past the server's deadline and was cancelled."""


class SolveHandle:
    """Given to a solve function, which attaches the call that stops its solve
    (``MPFAO.cancel``) while the solve runs; the watchdog calls ``cancel``."""

    def __init__(self):
        self._lock = threading.Lock()
        self._stop = None
        self.cancelled = False

    def attach(self, stop: Optional[Callable]):
        with self._lock:
            self._stop = stop
            if stop is not None and self.cancelled:
                stop()

    def cancel(self):
        with self._lock:
            self.cancelled = True
            if self._stop is not None:
                self._stop()


@dataclasses.dataclass
class Request:
    rid: int
    nx: int
    ny: int
    nz: int
    seed: int
    sigma: float
    steps: int = 5
    dt: float = 3600.0


def block_rows(nx, ny, nz, brick):
    """Rows of the block-Jacobi layout: one 64-, 256- or 512-wide slab per brick."""
    bx, by, bz = brick
    nb = -(-nx // bx) * -(-ny // by) * -(-nz // bz)
    bs = bx * by * bz
    return nb * (64 if bs <= 64 else 256 if bs <= 256 else 512)


def footprint_bytes(nc, nf, nv, nbnd, rows, krylov_fp32=True, restart=30):
    """Upper bound of the device bytes a request holds while it solves (the
    library's setup arrays, its setup temporaries and the Newton/Krylov
    workspace), from the mesh counts.  Hexahedral meshes: at most 27 Jacobian
    entries per row."""
    tv = 8  # Krylov buffers are sized for FP64 (the fallback of a failed FP32 solve)
    ctx = 533 * nc + 609 * nf + 32 * nv + 36 * nbnd
    setup_tmp = 768 * nc + 64 * nf + 264 * nv
    # Newton vectors, block-ordered Jacobian and its columns, Krylov basis,
    # block-Jacobi tables and factor (rows = the block layout's row count)
    ws = 64 * nc + 4 * nc + rows * 27 * (8 + 4) + (restart + 5) * nc * tv + rows * (548 + 27 * tv)
    ws += rows * 27 * 4  # FP32 copy of the Jacobian (kept in both precisions)
    return int(ctx + max(setup_tmp, ws) + (1 << 20))


class _Job:
    __slots__ = ("req", "future", "t_arrive", "t_built", "t_admit", "t_done", "problem", "need", "handle")

    def __init__(self, req, future):
        self.req = req
        self.future = future
        self.t_arrive = time.perf_counter()
        self.t_built = self.t_admit = self.t_done = None
        self.problem = None
        self.need = 0
        self.handle = SolveHandle()


class EnsembleServer:
    """Serve Richards solve requests with continuous admission on one GPU.

    slots        concurrent requests on the device (one host thread and one
                 HIP stream each)
    hbm_budget   bytes of device memory the in-flight requests may hold;
                 default 80 % of the free HBM at start
    build_workers  host threads generating problems
    deadline_s   longest a request may solve before it is cancelled (None:
                 no limit)
    solve_fn     (problem, request, handle) -> (pressure ndarray, stats
                 dict); the default runs the device Newton solver.  It
                 attaches its stop call to the SolveHandle while it runs.
    """

    def __init__(self, slots: int = 4, hbm_budget: Optional[int] = None, build_workers: int = 8,
                 krylov_fp32: bool = True, snes_rtol: float = 1e-10, brick=(8, 8, 4),
                 solve_fn: Optional[Callable] = None, pc_coarse: bool = True,
                 deadline_s: Optional[float] = None, fused: bool = False, fuse_window_us: int = 200):
        self.fused = fused
        self.fuser = None
        if fused and solve_fn is None:
            from .batch import Fuser
            self.fuser = Fuser(fuse_window_us)
        self.slots = slots
        self.deadline_s = deadline_s
        self.pc_coarse = pc_coarse
        self.brick = tuple(brick)
        self.krylov_fp32 = krylov_fp32
        self.snes_rtol = snes_rtol
        if hbm_budget is None:
            import torch
            free, _ = torch.cuda.mem_get_info()
            hbm_budget = int(0.8 * free)
        self.hbm_budget = int(hbm_budget)
        self._solve = solve_fn or self._gpu_solve
        self._build = cf.ThreadPoolExecutor(max_workers=build_workers, thread_name_prefix="build")
        self._ready: "queue.Queue" = queue.Queue()
        self._cv = threading.Condition()
        self._used = 0
        self._inflight = 0
        self.peak_used = 0
        self.peak_inflight = 0
        self._closed = False
        self._draining = False
        self._running = {}  # slot -> job on the device
        self._workers = [threading.Thread(target=self._slot, args=(i,), daemon=True, name=f"slot{i}")
                         for i in range(slots)]
        for w in self._workers:
            w.start()
        self._watch_stop = threading.Event()
        self._watch = None
        if deadline_s is not None:
            self._watch = threading.Thread(target=self._watchdog, daemon=True, name="deadline")
            self._watch.start()

    # ---- client side ----
    def submit(self, req: Request) -> cf.Future:
        """Queue a request; the future resolves to a result dict."""
        fut: cf.Future = cf.Future()
        job = _Job(req, fut)
        with self._cv:
            if self._closed:
                raise RuntimeError("server is closed")
        self._build.submit(self._build_job, job)
        return fut

    def drain(self, cancel_running: bool = False):
        """Stop serving requests that are not on the device yet: their futures
        are cancelled.  Requests already solving finish, or with
        cancel_running are cancelled too (their futures fail with
        DeadlineExceeded)."""
        self._draining = True
        if cancel_running:
            with self._cv:
                running = list(self._running.values())
            for job in running:
                job.handle.cancel()
        while True:
            try:
                job = self._ready.get_nowait()
            except queue.Empty:
                break
            if job is None:
                self._ready.put(None)
                break
            job.future.cancel()

    def close(self):
        """Finish the queued requests and stop the slot threads."""
        self._build.shutdown(wait=True)
        with self._cv:
            self._closed = True
        for _ in self._workers:
            self._ready.put(None)
        for w in self._workers:
            w.join()
        self._watch_stop.set()
        if self._watch is not None:
            self._watch.join()
        if self.fuser is not None:
            self.fuser_stats = self.fuser.stats()
            self.fuser.close()
            self.fuser = None

    def _watchdog(self):
        while not self._watch_stop.wait(0.05):
            now = time.perf_counter()
            with self._cv:
                late = [j for j in self._running.values()
                        if j.t_admit is not None and now - j.t_admit > self.deadline_s and not j.handle.cancelled]
            for job in late:
                job.handle.cancel()

    # ---- stages ----
    def _build_job(self, job: _Job):
        r = job.req
        try:
            P = workloads.ensemble_box_problem(r.nx, r.ny, r.nz, r.seed, r.sigma)
            m = P["mesh"]
            rows = block_rows(r.nx, r.ny, r.nz, self.brick)
            job.need = footprint_bytes(m.nc, m.nf, m.nv, m.nbnd, rows, self.krylov_fp32)
            if job.need > self.hbm_budget:
                raise MemoryError(f"request {r.rid} needs {job.need} B, above the HBM budget {self.hbm_budget} B")
            if self._draining:
                job.future.cancel()
                return
            job.problem = P
            job.t_built = time.perf_counter()
            self._ready.put(job)
        except BaseException as e:  # noqa: BLE001 - delivered to the caller's future
            job.future.set_exception(e)

    def _admit(self, job: _Job):
        with self._cv:
            while self._used + job.need > self.hbm_budget:
                self._cv.wait()
            self._used += job.need
            self._inflight += 1
            self.peak_used = max(self.peak_used, self._used)
            self.peak_inflight = max(self.peak_inflight, self._inflight)
        job.t_admit = time.perf_counter()

    def _release(self, job: _Job):
        with self._cv:
            self._used -= job.need
            self._inflight -= 1
            self._cv.notify_all()

    def _slot(self, wid: int):
        while True:
            job = self._ready.get()
            if job is None:
                return
            if self._draining:
                job.future.cancel()
                continue
            self._admit(job)
            with self._cv:
                self._running[wid] = job
            try:
                U, st = self._solve(job.problem, job.req, job.handle)
                job.t_done = time.perf_counter()
                job.future.set_result({
                    "rid": job.req.rid, "nc": job.problem["mesh"].nc, "U": U, "stats": st,
                    "latency_s": job.t_done - job.t_arrive, "build_s": job.t_built - job.t_arrive,
                    "queue_s": job.t_admit - job.t_built, "solve_s": job.t_done - job.t_admit,
                    "t_arrive": job.t_arrive, "t_done": job.t_done, "slot": wid})
            except BaseException as e:  # noqa: BLE001
                job.t_done = time.perf_counter()
                if job.handle.cancelled and not isinstance(e, DeadlineExceeded):
                    e = DeadlineExceeded(f"request {job.req.rid} cancelled after "
                                         f"{job.t_done - job.t_admit:.1f} s of solving ({e})")
                e.rid, e.t_arrive, e.t_done = job.req.rid, job.t_arrive, job.t_done
                e.latency_s = job.t_done - job.t_arrive
                job.future.set_exception(e)
            finally:
                with self._cv:
                    self._running.pop(wid, None)
                job.problem = None
                self._release(job)

    def _gpu_solve(self, P, req: Request, handle: Optional[SolveHandle] = None):
        import torch
        from .mpfao import MPFAO
        g = MPFAO(P["mesh"], P["face_bc"], P["K"], P["porosity"], P["Sr"], **P["kw"])
        if handle is not None:
            handle.attach(g.cancel)
        try:
            if self.fuser is not None:
                self.fuser.attach(g)
            g.set_pc_partition(brick_partition(P["mesh"], *self.brick))
            bs = self.brick[0] * self.brick[1] * self.brick[2]
            o = g.solver_options(snes_rtol=self.snes_rtol, krylov_fp32=1 if self.krylov_fp32 else 0, pc_block=bs,
                                 pc_coarse=1 if self.pc_coarse else 0)
            # all device work of this request stays on the context's own stream
            # (never a torch pool stream, which another slot could be using)
            with torch.cuda.stream(g.stream):
                U = torch.tensor(P["U0"], device="cuda", dtype=torch.float64)
                # TDyTimeIntegratorRunToTime semantics: steps of req.dt, a failed
                # Newton solve cuts dt by 4 and regrows by 1.25 up to req.dt
                st = g.run_tdyti(U, req.dt, 1.25, req.dt, req.steps * req.dt, o)
                out = U.cpu().numpy()
            del U
            st["device_bytes"] = g.device_bytes()
            return out, st
        finally:
            if handle is not None:
                handle.attach(None)
            if self.fuser is not None:
                self.fuser.detach(g)
            g.close()


def latency_summary(results):
    """p50/p99 latency and throughput over a list of result dicts."""
    lat = np.array([r["latency_s"] for r in results])
    t0 = min(r["t_arrive"] for r in results)
    t1 = max(r["t_done"] for r in results)
    return {"requests": len(results), "solves_per_s": len(results) / (t1 - t0),
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()), "makespan_s": t1 - t0,
            "mean_solve_s": float(np.mean([r["solve_s"] for r in results])),
            "mean_queue_s": float(np.mean([r["queue_s"] for r in results])),
            "mean_build_s": float(np.mean([r["build_s"] for r in results]))}
