#!/usr/bin/env python3
"""MPFA-O Richards Newton throughput on one MI355X.

Workload: the 128x128x64 layered, jittered hex box (1,048,576 cells,
1000x1000x100 m), backward-Euler steps of dt = 1 day driven by the device
Newton solver (SNESNEWTONLS + backtracking line search, snes_rtol 1e-10) with
GMRES(30) + block-Jacobi ILU(0) (8x8x4 cell bricks) with a two-level
coarse correction (one value per brick, bricks aggregated into at most 128
groups solved exactly; --pc-coarse 0 runs one-level block Jacobi) as the
inner solve.

A step is one backward-Euler time step: SNES pre-solve (accumulation of the
previous state) and the Newton solve to convergence, every residual,
Jacobian, factorisation and Krylov iteration included.  W warmup steps run
from the initial state, the state is reset, then exactly K steps are timed.

Printed value: Newton cells/s = cells x Newton iterations / wall time of the
K timed steps (FP64 residual, Jacobian, Krylov), timed with the kernel timers
off and each GMRES restart cycle replayed as a hipGraph.  The JSON line also
carries:
  roofline      from a second pass over the same K steps with HIP events
                around every launch on the library's stream (graphs off; the
                iterations are the same, bit for bit): the kernel with the
                largest share of device time, its algorithmic bytes per launch
                / mean launch duration against 8 TB/s, over the launches that
                did work (launches past a GMRES cycle's last Arnoldi step
                return at once and are booked apart); whole_step = all
                kernels' algorithmic bytes / the headline wall time;
  kernels       the same figures for every kernel of the path;
  setup         device times of the transmissibility setup phases and the
                FP64 and HBM fractions of the batched local systems;
  fp32_krylov   the same K steps with FP32 Jacobian storage and FP32 GMRES /
                ILU(0) (FP64 residual and Newton update), its speedup and its
                max relative pressure difference from the FP64 run, plus an
                untimed step-by-step pass: per step the FP32-vs-FP64 pressure
                difference, both forms' mass-balance defects and iteration
                counts;
  one_level_fp64  with the coarse correction on (default), the same K FP64
                steps with PETSc's default one-level block Jacobi, for reference;
  cfg0_latency  the 16^3 jittered cube: residual, Jacobian and per-step
                latency (cache-resident, so latency rather than bandwidth);
  cpu_baseline  the oracle (C restatement of the reference algorithm) timed on
                this host's cores;
  cfg2_prisms   the 4M-prism configuration (200,000 random Delaunay triangles
                x 20 layers): its one-launch region setup and Newton cells/s
                over three 1 h steps (bench.py --prisms runs it in full);
  serving_sample  a bounded sample of the serving ensemble (the first
                --serve-sample requests, Poisson arrivals): solves/s and
                p50/p99 latency of the server with the evaluation fuser
                (bench.py --serve runs the full ensemble, --fused 0 without
                fusion).
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on
# this workload (scaled by the counters' reading of a 1 GiB device copy)
TRAFFIC_FILE = os.path.join(HERE, "profiles", "pmc_traffic_layered_box_128x128x64.json")


def parse():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed backward-Euler steps (K)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed steps before the timed region (W)")
    ap.add_argument("--mesh", type=int, nargs=3, default=[128, 128, 64], metavar=("NX", "NY", "NZ"))
    ap.add_argument("--dt", type=float, default=86400.0)
    ap.add_argument("--brick", type=int, nargs=3, default=[8, 8, 4], help="block-Jacobi brick (cells)")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-Krylov variant")
    ap.add_argument("--no-cpu-baseline", action="store_true", help="skip the CPU oracle sample")
    ap.add_argument("--no-timers", action="store_true", help="skip the per-kernel timer pass")
    ap.add_argument("--no-cfg0", action="store_true", help="skip the 16^3 latency figures")
    ap.add_argument("--prisms", action="store_true",
                    help="measure the prism-mesh transmissibility setup instead (random Delaunay, extruded)")
    ap.add_argument("--prism-tri", type=int, default=200_000, help="prisms: triangles of the 2-D mesh")
    ap.add_argument("--prism-layers", type=int, default=20, help="prisms: layers")
    ap.add_argument("--prism-steps", type=int, default=10, help="prisms: backward-Euler steps of 1 h after the setup")
    ap.add_argument("--prism-setups", type=int, default=3,
                    help="prisms: setups run (the first tunes the region launch; the rest are timed)")
    ap.add_argument("--serve", action="store_true",
                    help="run the serving ensemble instead (Poisson arrivals of box requests)")
    ap.add_argument("--fused", type=int, default=1,
                    help="serving: fused batch engine (1: one segmented launch per evaluation over all requests "
                         "on the device) or one context per slot (0)")
    ap.add_argument("--no-prisms", action="store_true", help="default run: skip the prism-mesh (cfg2) figures")
    ap.add_argument("--no-serve", action="store_true", help="default run: skip the bounded serving sample")
    ap.add_argument("--serve-sample", type=int, default=24,
                    help="default run: requests of the bounded serving sample (fused server)")
    ap.add_argument("--requests", type=int, default=512, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=8.0, help="serving: Poisson arrival rate (requests/s)")
    ap.add_argument("--slots", type=int, default=8, help="serving: concurrent requests on the device")
    ap.add_argument("--build-workers", type=int, default=12, help="serving: host threads building problems")
    ap.add_argument("--pc-coarse", type=int, default=1, help="coarse correction on block Jacobi (1) or not (0)")
    ap.add_argument("--serve-fp32", type=int, default=0, help="serving: FP32 Krylov (1) or FP64 (0)")
    ap.add_argument("--deadline", type=float, default=0.0,
                    help="serving: seconds a request may solve before it is cancelled (0: no limit)")
    ap.add_argument("--max-seconds", type=float, default=0.0,
                    help="serving: after this long, requests not yet on the device are dropped (0: no limit)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


EXITED = " (exited at once)"


def kernel_table(rep):
    """Per-kernel rows from the library's timer report.  Launches of a GMRES
    cycle past its last Arnoldi step return at once on the cycle's stop flag;
    the library books them under "<kernel> (exited at once)" with no bytes, so
    a kernel's own row averages only launches that did its work."""
    out = []
    for r in rep:
        if r["launches"] == 0 or r["ms"] <= 0:
            continue
        avg_ms = r["ms"] / r["launches"]
        avg_b = r["bytes"] / r["launches"]
        out.append({"kernel": r["kernel"], "launches": r["launches"], "total_ms": round(r["ms"], 3),
                    "avg_us": round(1e3 * avg_ms, 2), "bytes_per_launch": avg_b,
                    "GB_s": round(avg_b / (avg_ms * 1e-3) / 1e9, 1)})
    out.sort(key=lambda r: -r["total_ms"])
    return out


def dominant(kernels):
    """The working kernel with the largest share of device time."""
    return next(k for k in kernels if not k["kernel"].endswith(EXITED))


def whole_step(rep, wall_s):
    """All kernels' algorithmic bytes over the wall time of the same steps."""
    total_b = sum(r["bytes"] for r in rep)
    kernel_ms = sum(r["ms"] for r in rep)
    gbs = total_b / wall_s / 1e9
    return {"bytes": total_b, "wall_s": round(wall_s, 4), "GB_s": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4), "kernel_ms_timer_pass": round(kernel_ms, 2)}


def run_steps(g, U, U0, dt, K, W, opts, timers):
    import torch
    U.copy_(U0)
    if W > 0:
        g.run_snes_steps(U, dt, W, opts)
    U.copy_(U0)
    torch.cuda.synchronize()
    if timers:
        g.kernel_timers(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = g.run_snes_steps(U, dt, K, opts)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    rep = None
    if timers:
        rep = g.kernel_timer_report()
        g.kernel_timers(False)
    return el, st, rep


def accuracy_pass(g, U0, dt, K, o64, o32):
    """Untimed: the K steps again, one at a time, FP64 and FP32-Krylov side by
    side.  After each step: max relative pressure difference of the FP32 form
    from the FP64 form, and each form's mass-balance defect
    |sum_c R_c(U^n)| dt / sum_c rho phi S V, with R the SNES residual of the
    step (fluxes + accumulation: internal fluxes cancel, so the sum is the
    step's global water balance)."""
    import torch
    U64, U32 = U0.clone(), U0.clone()
    R = torch.empty_like(U0)
    rows = []

    def defect(U):
        g.SNESFunction(U, R)
        W = g.Diagnostics(U)[2]
        return abs(R.sum().item()) * dt / W.sum().item()

    for k in range(K):
        s64 = g.run_snes_steps(U64, dt, 1, o64)
        mb64 = defect(U64)
        s32 = g.run_snes_steps(U32, dt, 1, o32)
        mb32 = defect(U32)
        rel = ((U32 - U64).abs() / U64.abs()).max().item()
        rows.append({"step": k + 1, "max_rel_p_diff": rel, "mass_balance_fp64": mb64, "mass_balance_fp32": mb32,
                     "newton_fp64": s64["newton_its"], "newton_fp32": s32["newton_its"],
                     "krylov_fp64": s64["linear_its"], "krylov_fp32": s32["linear_its"],
                     "fp64_fallbacks": s32["krylov_fallbacks"]})
    return rows


def cfg0_latency(reps=200):
    """Latency on the 16^3 jittered cube (4,096 cells, cache-resident): one
    TS residual, one TS Jacobian, and each of the configuration's 10
    backward-Euler SNES steps (dt = 3600 s, snes_rtol 1e-10), each call
    timed from its launch to the end of a device synchronisation."""
    import numpy as np
    import torch
    from tdycore_amd import workloads
    from tdycore_amd.mpfao import MPFAO
    P = workloads.jittered_cube_problem(16)
    g = MPFAO(P["mesh"], P["face_bc"], P["K"], P["porosity"], P["Sr"], **P["kw"])
    nc = P["mesh"].nc
    rng = np.random.default_rng(3)
    U = torch.tensor(P["U0"] + 1e3 * (rng.random(nc) - 0.5), device="cuda")
    Ud = torch.tensor(1e-2 * (rng.random(nc) - 0.5), device="cuda")
    R = torch.empty_like(U)
    J = torch.empty(g.jac_nnz, dtype=torch.float64, device="cuda")

    def lat(fn):
        for _ in range(20):
            fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        ts = np.array(ts) * 1e6
        return {"median_us": round(float(np.median(ts)), 1), "p99_us": round(float(np.percentile(ts, 99)), 1)}

    res = lat(lambda: g.IFunction(0.0, U, Ud, R))
    jac = lat(lambda: g.IJacobian(0.0, U, Ud, 1.0 / 3600.0, J))
    opts = g.solver_options(snes_rtol=1e-10, ksp_rtol=1e-5, pc_block=64, pc_coarse=1)
    Us = torch.tensor(P["U0"], device="cuda")
    g.run_snes_steps(Us, 3600.0, 1, opts)  # warm-up step: workspace, graphs
    Us = torch.tensor(P["U0"], device="cuda")
    torch.cuda.synchronize()
    steps, newton, lin = [], 0, 0
    for _ in range(10):
        t0 = time.perf_counter()
        st = g.run_snes_steps(Us, 3600.0, 1, opts)
        torch.cuda.synchronize()
        steps.append(time.perf_counter() - t0)
        newton += st["newton_its"]
        lin += st["linear_its"]
        if st["status"] != 0:
            raise RuntimeError(f"16^3 step failed: {st}")
    g.close()
    steps = np.array(steps) * 1e3
    return {"workload": "16x16x16 jittered unit cube (4096 cells), rotated tensor K, 10 backward-Euler SNES steps "
                        "of 3600 s, snes_rtol 1e-10, GMRES(30) + bjacobi ILU(0) 64-row blocks + coarse correction",
            "residual": res, "jacobian": jac,
            "step_ms": {"median": round(float(np.median(steps)), 3), "max": round(float(steps.max()), 3),
                        "total_10_steps": round(float(steps.sum()), 3)},
            "newton_its": newton, "linear_its": lin,
            "newton_iteration_ms": round(float(steps.sum()) / newton, 3)}


def stream_copy_gbs(nbytes=1 << 30, reps=10):
    """Device-to-device copy bandwidth (read + write bytes / time) of a
    1 GiB FP64 buffer, measured on torch's stream with torch events."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del a, b
    return gbs


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(P, dt):
    """The oracle (reference-algorithm CPU restatement, FP64, OpenMP) on this
    host: one full backward-Euler step of the same mesh and dt on all the
    host's threads (OMP_NUM_THREADS, else every CPU) with GMRES(30) and
    block-Jacobi ILU(0), one block per thread (PETSc's parallel default),
    plus its residual and Jacobian evaluation times; and the first Newton
    iteration of the step on one thread."""
    import numpy as np
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    from pyoracle import Oracle
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    t = time.perf_counter()
    o = Oracle(P["mesh"], P["face_bc"], P["K"], P["porosity"], P["Sr"], **P["kw"])
    setup_s = time.perf_counter() - t
    log(f"cpu oracle setup {setup_s:.1f} s")
    nc = P["mesh"].nc
    o.set_threads(threads)
    U = P["U0"].copy()
    o.set_dtime(dt)
    o.snes_presolve(U)
    t = time.perf_counter()
    o.snes_function(U)
    res_s = time.perf_counter() - t
    t = time.perf_counter()
    o.snes_jacobian(U)
    jac_s = time.perf_counter() - t
    t = time.perf_counter()
    _, st, status = o.newton_snes(U, lin="gmres_bjacobi", rtol=1e-10)
    el = time.perf_counter() - t
    log(f"cpu baseline, {threads} threads: one step {el:.2f} s, {int(st[0])} Newton, {int(st[1])} GMRES")
    o.set_threads(1)
    o.set_newton_maxit(1)
    o.snes_presolve(U)
    t = time.perf_counter()
    _, st1, _ = o.newton_snes(U, lin="gmres_bjacobi", rtol=1e-10)
    el1 = time.perf_counter() - t
    log(f"cpu baseline, 1 thread: one Newton iteration {el1:.2f} s, {int(st1[1])} GMRES")
    return {"value": nc * int(st[0]) / el, "unit": "Newton cells/s", "cores": threads, "kind": "port",
            "label": "reference-algorithm CPU restatement (oracle/tdy_oracle.c, FP64, OpenMP)",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"one backward-Euler step of dt={dt:g} s on the same {nc}-cell mesh: {int(st[0])} Newton, "
                      f"{int(st[1])} GMRES(30) iterations, block-Jacobi ILU(0) with {threads} blocks, "
                      f"{el:.2f} s on {threads} threads (status {status})",
            "residual_ms": round(1e3 * res_s, 2), "jacobian_ms": round(1e3 * jac_s, 2),
            "setup_s": round(setup_s, 2),
            "one_core": {"value": nc * int(st1[0]) / el1, "cores": 1,
                         "sample": f"the step's first Newton iteration on one thread: {int(st1[1])} GMRES "
                                   f"iterations, {el1:.2f} s"}}


def prisms(a):
    """Setup of the interaction regions of a right-prism mesh (random Delaunay
    triangles on a 10 x 10 km square extruded to layers of U(0.5, 5) m): G
    matrices, size-bucketed local systems, gravity term.  A step is one
    setup; the value is interaction regions per second of the local-system
    kernels, with the FP64 roofline of that phase."""
    import numpy as np
    import torch
    from tdycore_amd import workloads
    from tdycore_amd.mpfao import MPFAO
    torch.cuda.init()
    t = time.perf_counter()
    per = max(1, 64 // a.prism_layers)  # block-Jacobi blocks: vertical columns of `per` triangles
    P = workloads.prism_problem(n_tri=a.prism_tri, nlayers=a.prism_layers, infiltration=True, block_tris=per)
    m = P["mesh"]
    log(f"prism mesh: {m.nc} cells, {m.nf} faces, {m.nv} vertices, built in {time.perf_counter() - t:.1f} s")
    runs = []
    plan = None
    nsetup = max(2, a.prism_setups)
    for it in range(nsetup):
        t = time.perf_counter()
        g = MPFAO(m, P["face_bc"], P["K"], P["porosity"], P["Sr"], **P["kw"])
        wall = time.perf_counter() - t
        info = g.setup_info()
        if it == 0:
            plan = g.setup_plan()  # the tuning runs are recorded by the first setup
        g.close()
        del g
        if it >= 1:
            runs.append((wall, info))
        log(f"setup {it}: wall {wall:.2f} s, device {info['gmatrix_ms']:.2f} + {info['local_systems_ms']:.2f} + "
            f"{info['gravity_ms']:.2f} ms, buckets {info['buckets']}, lanes {info['lanes']}")
    info = runs[-1][1]
    log(f"region launch plan: lanes {plan['lanes']}, largest first {plan['largest_first']}, chunk {plan['chunk']}; "
        f"tuning ms widths {plan['width_ms']}, schedules {plan['schedule_ms']}")
    # Newton steps: SNES backward Euler, dt = 1 h, GMRES(30) + block-Jacobi ILU(0)
    # on vertical columns of 3 triangles (60 cells for 20 layers)
    g = MPFAO(m, P["face_bc"], P["K"], P["porosity"], P["Sr"], **P["kw"])
    g.set_pc_partition(P["pc_part"])
    opts = g.solver_options(snes_rtol=1e-10, pc_block=64)
    U = torch.tensor(P["U0"], device="cuda", dtype=torch.float64)
    g.run_snes_steps(U, 3600.0, 1, opts)  # warmup step (not timed): workspace, factor, graphs
    U = torch.tensor(P["U0"], device="cuda", dtype=torch.float64)
    torch.cuda.synchronize()
    t = time.perf_counter()
    st = g.run_snes_steps(U, 3600.0, a.prism_steps, opts)
    torch.cuda.synchronize()
    newton_s = time.perf_counter() - t
    dU = float((U - torch.tensor(P["U0"], device="cuda")).abs().max())
    g.close()
    log(f"prism Newton: {a.prism_steps} steps in {newton_s:.2f} s, {st['newton_its']} Newton, "
        f"{st['linear_its']} GMRES, status {st['status']}, max |dp| {dU:.1f} Pa")
    ls_s = np.mean([r[1]["local_systems_ms"] for r in runs]) / 1e3
    dev_s = np.mean([r[1]["gmatrix_ms"] + r[1]["local_systems_ms"] + r[1]["gravity_ms"] for r in runs]) / 1e3
    wall = np.mean([r[0] for r in runs])
    tflops = info["local_systems_flops"] / ls_s / 1e12
    line = {
        "metric": "MPFA-O Newton cells/s on prisms (FP64)", "value": m.nc * st["newton_its"] / newton_s,
        "unit": "Newton cells/s",
        "n_gpus": 1, "steps": a.prism_steps, "warmup": 1, "ms_per_step": 1e3 * newton_s / a.prism_steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (jittered-lattice random Delaunay, layered permeability)",
        "config": {"workload": f"{a.prism_tri} random Delaunay triangles on 10x10 km extruded to "
                               f"{a.prism_layers} layers ({m.nc} prisms, {m.nv} vertices); "
                               f"{a.prism_steps} backward-Euler steps of 1 h, ponded infiltration from the top",
                   "cells": m.nc, "faces": m.nf, "vertices": m.nv, "buckets_16_32_64": info["buckets"],
                   "max_cells_at_vertex": info["max_cells"], "max_faces_at_vertex": info["max_faces"],
                   "lanes_16_32_64_buckets": info["lanes"],
                   "trans_nnz": info["trans_nnz"]},
        "roofline": {"bound": "fp64", "kernel": "k_ir_persistent (all buckets, one launch)", "achieved": round(tflops, 3),
                     "peak": 78.6, "unit": "TFLOP/s", "frac": round(tflops / 78.6, 4),
                     "hbm_GB_s": round(info["local_systems_bytes"] / ls_s / 1e9, 1), "traffic": None},
        "phases_ms": {"gmatrix": info["gmatrix_ms"], "local_systems": info["local_systems_ms"],
                      "gravity": info["gravity_ms"]},
        "setup": {"regions_per_s": m.nv / ls_s, "device_ms": 1e3 * dev_s, "wall_s": wall, "setups_timed": len(runs),
                  "plan": plan},
        "newton": {"steps": a.prism_steps, "dt_s": 3600.0, "newton_its": st["newton_its"],
                   "linear_its": st["linear_its"], "status": st["status"], "seconds": newton_s,
                   "max_dp_Pa": dU, "preconditioner": f"block-Jacobi ILU(0), columns of {per} triangles"},
    }
    if not a.no_cpu_baseline:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        from pyoracle import Oracle
        S = workloads.prism_problem(n_tri=20_000, nlayers=a.prism_layers, infiltration=True)
        t = time.perf_counter()
        o = Oracle(S["mesh"], S["face_bc"], S["K"], S["porosity"], S["Sr"], **S["kw"])
        el_setup = time.perf_counter() - t
        Uc = S["U0"].copy()
        o.set_dtime(3600.0)
        o.snes_presolve(Uc)
        o.set_newton_maxit(1)
        t = time.perf_counter()
        _, sto, _ = o.newton_snes(Uc, lin="gmres", rtol=1e-10)
        el = time.perf_counter() - t
        ncs = S["mesh"].nc
        line["cpu_baseline"] = {"value": ncs * int(sto[0]) / el, "unit": "Newton cells/s", "cores": 1, "kind": "port",
                                "sample": f"1 Newton iteration (residual, Jacobian, ILU(0), GMRES(30) {int(sto[1])} its, "
                                          f"line search) of step 1 on a 20,000-triangle x {a.prism_layers}-layer prism "
                                          f"mesh ({ncs} cells): {el:.2f} s",
                                "setup_regions_per_s": S["mesh"].nv / el_setup}
    return line


def serve(a, n_requests=None, fused=None):
    """Serving ensemble: requests arrive as a Poisson stream and are admitted by
    HBM footprint into a fixed number of device slots."""
    import numpy as np
    import torch
    from tdycore_amd import serving, workloads
    torch.cuda.init()
    fused = a.fused if fused is None else fused
    specs = workloads.ensemble_request_specs(n_requests or a.requests)
    srv = serving.EnsembleServer(slots=a.slots, build_workers=a.build_workers, pc_coarse=bool(a.pc_coarse),
                                 krylov_fp32=bool(a.serve_fp32), deadline_s=a.deadline or None, fused=bool(fused))
    # warm the library and the device once (not timed)
    srv.submit(serving.Request(rid=-1, nx=16, ny=16, nz=8, seed=1, sigma=1.0, steps=1)).result()
    rng = np.random.default_rng(20261207)
    gaps = rng.exponential(1.0 / a.rate, size=len(specs))
    futs = []
    ndone = [0]

    done_res = []

    def report(f):
        ndone[0] += 1
        if f.cancelled():
            return
        if f.exception() is None:
            r = f.result()
            done_res.append(r)
            st = r["stats"]
            log(f"done {ndone[0]}: rid {r['rid']} cells {r['nc']} newton {st['newton_its']} linear {st['linear_its']} "
                f"steps {st['steps']} cuts {st['cuts']} solve {r['solve_s']:.2f}s latency {r['latency_s']:.2f}s")
            if len(done_res) % 32 == 0:
                log("partial summary: " + json.dumps(serving.latency_summary(list(done_res))))
        else:
            e = f.exception()
            kind = "timed out" if isinstance(e, serving.DeadlineExceeded) else "failed"
            log(f"done {ndone[0]}: {kind} {e}")

    t0 = time.perf_counter()
    t_next = t0
    for sp, gap in zip(specs, gaps):
        t_next += gap
        dt = t_next - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        futs.append(srv.submit(serving.Request(**sp)))
        futs[-1].add_done_callback(report)
        if a.max_seconds and time.perf_counter() - t0 > a.max_seconds:
            break
    if a.max_seconds:
        while time.perf_counter() - t0 < a.max_seconds and not all(f.done() for f in futs):
            time.sleep(1.0)
        srv.drain(cancel_running=True)
    res, failed, timed_out, dropped, responses = [], [], [], [], []
    for sp, f in zip(specs, futs):
        if f.cancelled():
            dropped.append(sp["rid"])
            continue
        e = f.exception()
        if e is None:
            res.append(f.result())
            responses.append(res[-1])
            continue
        if hasattr(e, "latency_s"):
            responses.append({"latency_s": e.latency_s, "t_arrive": e.t_arrive, "t_done": e.t_done})
        if isinstance(e, serving.DeadlineExceeded):
            timed_out.append(sp["rid"])
        else:  # e.g. the time integrator's 25 dt cuts: reported, not retried
            failed.append({"rid": sp["rid"], "error": str(e)})
    dropped += [sp["rid"] for sp in specs[len(futs):]]
    srv.close()
    summ = serving.latency_summary(res)
    lat_all = np.array([r["latency_s"] for r in responses])
    bad = [r["rid"] for r in res if r["stats"]["status"] != 0] + [f["rid"] for f in failed]
    cells = sum(r["nc"] for r in res)
    newton = sum(r["stats"]["newton_its"] for r in res)
    line = {
        "metric": "ensemble Richards solves/s (p50/p99 latency)", "value": summ["solves_per_s"], "unit": "solves/s",
        "n_gpus": 1, "steps": len(res), "warmup": 1, "ms_per_step": 1e3 * summ["makespan_s"] / len(res),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded jittered boxes, rotated log-normal permeability)",
        "config": {"workload": f"{len(res)} requests, Poisson arrivals at {a.rate:g}/s, nx,ny~U{{16..96}}, "
                               f"nz~U{{8..48}}, sigma~U(0.5,1.5), 5 backward-Euler steps of 1 h, FP64 residual, "
                               f"{'FP32' if a.serve_fp32 else 'FP64'} Krylov, {a.slots} device slots, "
                               + ("evaluation fuser (the SNES residual and Jacobian evaluations pending at once "
                                  "across slots run as one segmented launch per kernel)" if fused
                                  else "one context per slot, no fusion"),
                   "fused": bool(fused),
                   "total_cells": cells, "newton_its": newton, "failed": bad, "failures": failed[:20],
                   "timed_out": timed_out, "deadline_s": a.deadline or None,
                   "not_served": len(dropped), "served": len(res), "max_seconds": a.max_seconds,
                   "pc_coarse": a.pc_coarse, "krylov_fp32": a.serve_fp32},
        "responses": {"count": len(responses), "p50_latency_s": float(np.percentile(lat_all, 50)),
                      "p99_latency_s": float(np.percentile(lat_all, 99))},
        "p50_latency_s": summ["p50_latency_s"], "p99_latency_s": summ["p99_latency_s"], "serving": summ,
        "peak_inflight": srv.peak_inflight, "peak_admitted_bytes": srv.peak_used, "hbm_budget": srv.hbm_budget,
        "fuser": getattr(srv, "fuser_stats", None),
    }
    return line


def main():
    a = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}, WORLD_SIZE {world}); "
              "the path is measured on one GPU's full mesh", file=sys.stderr)
        sys.exit(2)
    if a.serve:
        print(json.dumps(serve(a)), flush=True)
        return
    if a.prisms:
        print(json.dumps(prisms(a)), flush=True)
        return
    import numpy as np
    import torch
    from tdycore_amd import workloads
    from tdycore_amd.meshgen import brick_partition
    from tdycore_amd.mpfao import MPFAO

    nx, ny, nz = a.mesh
    t = time.perf_counter()
    P = workloads.layered_box_problem(nx, ny, nz)
    nc = P["mesh"].nc
    log(f"problem {nx}x{ny}x{nz} ({nc} cells) built in {time.perf_counter() - t:.1f} s")
    t = time.perf_counter()
    g = MPFAO(P["mesh"], P["face_bc"], P["K"], P["porosity"], P["Sr"], **P["kw"])
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t
    log(f"device setup {setup_s:.2f} s")
    g.set_pc_partition(brick_partition(P["mesh"], *a.brick))
    bs = a.brick[0] * a.brick[1] * a.brick[2]
    U0 = torch.tensor(P["U0"], device="cuda", dtype=torch.float64)
    U = U0.clone()
    timers = not a.no_timers
    pcc = 1 if a.pc_coarse else 0
    o64 = g.solver_options(snes_rtol=1e-10, pc_block=bs, krylov_fp32=0, pc_coarse=pcc)
    # headline: kernel timers off, GMRES cycles replayed as hipGraphs
    el, st, _ = run_steps(g, U, U0, a.dt, a.steps, a.warmup, o64, False)
    if st["status"] != 0:
        raise RuntimeError(f"FP64 Newton stepping failed: {st}")
    log(f"fp64: {a.steps} steps {el:.3f} s newton {st['newton_its']} linear {st['linear_its']}")
    value = nc * st["newton_its"] / el
    U64 = U.clone()
    # per-kernel figures from a second pass over the same steps with HIP events
    # around every launch (graphs off); the iterations are the same, bit for bit
    kernels, roof, wstep, timer_pass = [], None, None, None
    copy_gbs = stream_copy_gbs()
    log(f"stream copy {copy_gbs:.0f} GB/s")
    if timers:
        elt, stt, rep = run_steps(g, U, U0, a.dt, a.steps, 0, o64, True)
        same = (stt["newton_its"], stt["linear_its"]) == (st["newton_its"], st["linear_its"])
        timer_pass = {"wall_s": round(elt, 4), "same_iterations": same,
                      "same_state": bool(torch.equal(U, U64))}
        log(f"timer pass: {elt:.3f} s, same iterations {same}")
        kernels = kernel_table(rep)
        wstep = whole_step(rep, el)
        k0 = dominant(kernels)
        ach = k0["GB_s"]
        traffic = None
        if (nx, ny, nz) == (128, 128, 64) and os.path.exists(TRAFFIC_FILE):
            pmc = json.load(open(TRAFFIC_FILE))["kernels"].get(k0["kernel"])
            traffic = pmc["traffic_bytes"] if pmc else None
        roof = {"bound": "hbm", "kernel": k0["kernel"], "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                "bytes_per_launch": k0["bytes_per_launch"], "avg_us": k0["avg_us"], "launches": k0["launches"],
                "stream_copy_GB_s": round(copy_gbs, 1), "frac_of_copy": round(ach / copy_gbs, 4),
                "whole_step": wstep}

    info = g.setup_info()
    ls_s = info["local_systems_ms"] * 1e-3
    setup = {"gmatrix_ms": round(info["gmatrix_ms"], 3), "local_systems_ms": round(info["local_systems_ms"], 3),
             "gravity_ms": round(info["gravity_ms"], 3), "regions": info["buckets"][0],
             "local_systems_fp64_TFLOP_s": round(info["local_systems_flops"] / ls_s / 1e12, 3) if ls_s else None,
             "local_systems_fp64_frac": round(info["local_systems_flops"] / ls_s / 78.6e12, 4) if ls_s else None,
             "local_systems_GB_s": round(info["local_systems_bytes"] / ls_s / 1e9, 1) if ls_s else None,
             "local_systems_hbm_frac": round(info["local_systems_bytes"] / ls_s / 8e12, 4) if ls_s else None,
             "wall_s": round(setup_s, 3)}

    fp32 = None
    if not a.no_fp32:
        o32 = g.solver_options(snes_rtol=1e-10, pc_block=bs, krylov_fp32=1, pc_coarse=pcc)
        el32, st32, _ = run_steps(g, U, U0, a.dt, a.steps, a.warmup, o32, False)
        if st32["status"] != 0:
            raise RuntimeError(f"FP32-Krylov Newton stepping failed: {st32}")
        rel = ((U - U64).abs() / U64.abs().clamp_min(1e-300)).max().item()
        v32 = nc * st32["newton_its"] / el32
        k32 = []
        if timers:
            _, _, rep32 = run_steps(g, U, U0, a.dt, a.steps, 0, o32, True)
            k32 = kernel_table(rep32)[:8]
        per_step = accuracy_pass(g, U0, a.dt, a.steps, o64, o32)
        log("per-step FP32 vs FP64: max rel p diff %.2e, mass balance fp64 %.2e fp32 %.2e" % (
            max(r["max_rel_p_diff"] for r in per_step), max(r["mass_balance_fp64"] for r in per_step),
            max(r["mass_balance_fp32"] for r in per_step)))
        fp32 = {"value": v32, "ms_per_step": 1e3 * el32 / a.steps, "speedup_vs_fp64": round(v32 / value, 3),
                "wall_speedup_vs_fp64": round(el / el32, 3),
                "newton_its": st32["newton_its"], "linear_its": st32["linear_its"], "max_rel_p_diff_vs_fp64": rel,
                "kernels": k32, "per_step": per_step}
        log(f"fp32 krylov: {el32:.3f} s newton {st32['newton_its']} linear {st32['linear_its']} rel {rel:.3e}")

    one_level = None
    if pcc:
        o1 = g.solver_options(snes_rtol=1e-10, pc_block=bs, krylov_fp32=0, pc_coarse=0)
        el1, st1, _ = run_steps(g, U, U0, a.dt, a.steps, a.warmup, o1, False)
        one_level = {"value": nc * st1["newton_its"] / el1, "ms_per_step": 1e3 * el1 / a.steps,
                     "newton_its": st1["newton_its"], "linear_its": st1["linear_its"],
                     "note": "FP64, PETSc-default one-level block-Jacobi ILU(0), no kernel timers"}
        log(f"one-level fp64: {el1:.3f} s newton {st1['newton_its']} linear {st1['linear_its']}")
    g.close()
    del g

    small = None
    if not a.no_cfg0:
        small = cfg0_latency()
        log(f"16^3 latency: residual {small['residual']['median_us']} us, Jacobian {small['jacobian']['median_us']} us, "
            f"step {small['step_ms']['median']} ms")

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(P, a.dt)
        log(f"cpu baseline {cpu['value']:.3e} Newton cells/s")

    # the 4M-prism configuration: two setups (the first tunes the region
    # launch) and three backward-Euler steps of 1 h
    cfg2 = None
    if not a.no_prisms:
        import argparse as _ap
        pa = _ap.Namespace(**vars(a))
        pa.prism_setups, pa.prism_steps, pa.no_cpu_baseline = 2, 3, True
        pl = prisms(pa)
        cfg2 = {k: pl[k] for k in ("metric", "value", "unit", "ms_per_step", "config", "roofline", "phases_ms",
                                   "setup", "newton")}
        log(f"cfg2 prisms: {cfg2['value']:.3e} Newton cells/s")
    # a bounded serving sample: the first requests of the ensemble, fused
    # batch engine and one context per slot on the same arrivals
    srv_sample = None
    if not a.no_serve:
        sf = serve(a, n_requests=a.serve_sample, fused=1)
        srv_sample = {k: sf[k] for k in ("metric", "value", "unit", "p50_latency_s", "p99_latency_s", "responses",
                                          "serving", "peak_inflight", "fuser")}
        srv_sample["config"] = {k: sf["config"][k] for k in ("workload", "served", "failed", "timed_out",
                                                              "not_served", "total_cells", "newton_its")}
        log(f"serving sample: {sf['value']:.2f} solves/s, p50 {sf['p50_latency_s']:.2f} s, "
            f"p99 {sf['p99_latency_s']:.2f} s, fuser {sf['fuser']}")

    line = {
        "metric": "MPFA-O Newton cells/s at 1M hex (FP64)", "value": value, "unit": "Newton cells/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * el / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded layered permeability, jittered hex mesh)",
        "config": {"workload": f"layered jittered hex box {nx}x{ny}x{nz} ({nc} cells), 1000x1000x100 m, "
                               f"backward Euler dt={a.dt:g} s, snes_rtol 1e-10, GMRES(30)+bjacobi ILU(0) "
                               f"{a.brick[0]}x{a.brick[1]}x{a.brick[2]} bricks"
                               + (" + coarse correction" if pcc else ""),
                   "pc_coarse": pcc,
                   "cells": nc, "dt": a.dt, "newton_its": st["newton_its"], "linear_its": st["linear_its"],
                   "setup_s": round(setup_s, 3), "headline_kernel_timers": False, "hip_graphs": True},
        "roofline": roof, "cpu_baseline": cpu, "kernels": kernels, "timer_pass": timer_pass, "setup": setup,
        "fp32_krylov": fp32, "one_level_fp64": one_level, "cfg0_latency": small,
        "cfg2_prisms": cfg2, "serving_sample": srv_sample,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
