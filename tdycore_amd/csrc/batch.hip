// Fused serving batch: several independent requests (contexts of hexahedral
// meshes, each with its own state, solver workspace and stream) advanced in
// lockstep, one Newton iteration per round, with their state, residual and
// Jacobian evaluations fused into one segmented launch each (k_seg_*).
//
// Per request the arithmetic is exactly that of tdyhip_run_tdyti on the same
// context: the fused kernels compute each segment as the single-request
// kernels do, the Krylov solve, the Newton tail's reductions and the host
// decisions are the request's own, so a request's result is the same bit for
// bit alone, in any mix, and through tdyhip_run_tdyti.  The Krylov solves of
// the requests run concurrently on their streams (one host thread each per
// round).  Time stepping follows TDyTimeIntegratorRunToTime
// (src/tdyti.c:71-196), Newton PETSc's SNESNEWTONLS with the backtracking line
// search (as in the single-request solver).

namespace tdyhip {

struct BatchMember {
  int32_t id = 0;
  tdyhip_ctx *ctx = nullptr;
  double *U = nullptr;
  tdyhip_solver_opts o{};
  double growth = 1.0, dt_max = 0.0, final_time = 0.0;
  double time = 0.0, dt = 0.0, dt_save = 0.0;
  int icut = 0;
  enum Phase { kStep, kAttempt, kIter, kDone } phase = kStep;
  // the Newton solve in progress
  int it = 0;
  double fnorm = 0.0, fnorm0 = 0.0, ttol = 0.0;
  tdyhip_solver_stats st{};  // the whole run (newton_its, linear_its, steps, cuts, fallbacks; last status)
  int rc = 0;                // error code once failed
  std::string err;
  double fnorm2_h = 0.0;     // ||F||^2 read back at the start of a solve
};

static void parallel_members(std::vector<BatchMember *> &ms, const std::function<void(BatchMember *)> &fn) {
  if (ms.size() == 1) { fn(ms[0]); return; }
  std::vector<std::thread> th;
  th.reserve(ms.size());
  for (BatchMember *m : ms) th.emplace_back(fn, m);
  for (auto &t : th) t.join();
}

static std::string thread_error() { return g_err; }

}  // namespace tdyhip

struct tdyhip_batch {
  hipStream_t st = nullptr;      // the fused launches
  hipEvent_t done = nullptr;     // end of the last fused launch
  std::vector<tdyhip::BatchMember *> members;
  tdyhip::SegProblem *d_prob[3] = {nullptr, nullptr, nullptr};  // segment tables (start F, Jacobian, full-step F)
  tdyhip::SegProblem *h_prob[3] = {nullptr, nullptr, nullptr};  // pinned staging of the same
  int32_t next_id = 0;
};

namespace tdyhip {

// One fused evaluation over `ms` (hexahedral contexts with a solver
// workspace): the residual of the SNES form at X(m) into R(m) (state, boundary
// state, flux rows, gather), or the Jacobian at X(m) into the workspace's
// block-ordered ELL array (from the state of the last residual at X(m)).
// The batch stream waits for each request's stream, and each request's
// stream then waits for the launch.
template <class XF, class RF, class SF>
static int fused_eval(tdyhip_batch *b, int table, std::vector<BatchMember *> &ms, bool jac, XF xof, RF rof,
                      SF stopof) {
  const int n = (int)ms.size();
  if (n == 0) return 0;
  SegProblem *hp = b->h_prob[table];
  SegBlocks s_state{}, s_bnd{}, s_rows{}, s_gather{}, s_jac{};
  s_state.nseg = s_bnd.nseg = s_rows.nseg = s_gather.nseg = s_jac.nseg = n;
  for (int i = 0; i < n; i++) {
    tdyhip_ctx *ctx = ms[i]->ctx;
    SolverWS &ws = *ctx->ws;
    const double *X = xof(ms[i]);
    SegProblem &p = hp[i];
    p.cs = cell_state_args(ctx, 1, X, nullptr);
    p.bs = bnd_state_args(ctx, 1, X);
    p.ra = row_args(ctx, X);
    p.ga = jac ? gather_args(ctx, 1, 0.0, nullptr, nullptr, ws.J, true, ws.d_perm, ws.bj.bs)
               : gather_args(ctx, 1, 0.0, nullptr, rof(ms[i]), nullptr, false, nullptr, 0);
    p.cj = cell_jac_args(ctx, ws.bj.d_row, ws.NB);
    p.stop = stopof(ms[i]);
    s_state.start[i + 1] = s_state.start[i] + nblk(ctx->nc, 256);
    s_bnd.start[i + 1] = s_bnd.start[i] + nblk(ctx->nbnd, 256);
    s_rows.start[i + 1] = s_rows.start[i] + nblk((int64_t)4 * ctx->nf, kRowBlock);
    s_gather.start[i + 1] = s_gather.start[i] + nblk(ctx->nc, kGatherBlock);
    s_jac.start[i + 1] = s_jac.start[i] + nblk(ws.NB, kCellJacCells);
    hipEvent_t e = ctx->ws_event;
    SCHK(hipEventRecord(e, ctx->stream));
    SCHK(hipStreamWaitEvent(b->st, e, 0));
  }
  SCHK(hipMemcpyAsync(b->d_prob[table], hp, sizeof(SegProblem) * n, hipMemcpyHostToDevice, b->st));
  const SegProblem *dp = b->d_prob[table];
  if (jac) {
    k_seg_cell_jac<<<(unsigned)s_jac.start[n], kCellJacBlock, 0, b->st>>>(dp, s_jac);
  } else {
    k_seg_state<<<(unsigned)s_state.start[n], 256, 0, b->st>>>(dp, s_state);
    if (s_bnd.start[n] > 0) k_seg_bnd_state<<<(unsigned)s_bnd.start[n], 256, 0, b->st>>>(dp, s_bnd);
    k_seg_rows<<<(unsigned)s_rows.start[n], kRowBlock, 0, b->st>>>(dp, s_rows);
    k_seg_cell_gather<<<(unsigned)s_gather.start[n], kGatherBlock, 0, b->st>>>(dp, s_gather);
  }
  SCHK(hipGetLastError());
  SCHK(hipEventRecord(b->done, b->st));
  for (int i = 0; i < n; i++) SCHK(hipStreamWaitEvent(ms[i]->ctx->stream, b->done, 0));
  return 0;
}

// The end of a Newton solve (status, iterations): the time integrator's
// bookkeeping, then the member's next phase.
static void finish_newton(BatchMember *m, int status, int its) {
  tdyhip_ctx *ctx = m->ctx;
  SolverWS &ws = *ctx->ws;
  m->st.newton_its += its;
  m->st.status = status;
  m->st.last_fnorm = m->fnorm;
  hipStream_t st = ctx->stream;
  if (status < 0) {
    m->icut++;
    m->st.cuts++;
    m->dt *= 0.25;
    (void)tdyhip_set_dtime(ctx, m->dt);
    if (hipMemcpyAsync(m->U, ws.Usave, sizeof(double) * ctx->nc, hipMemcpyDeviceToDevice, st) != hipSuccess) {
      m->rc = TDYHIP_ERR_HIP; m->err = "hipMemcpyAsync failed"; m->phase = BatchMember::kDone; return;
    }
    if (m->icut == 25) { m->rc = TDYHIP_ERR_NOCONV; m->err = "time step cut 25 times"; m->phase = BatchMember::kDone; return; }
    m->phase = BatchMember::kAttempt;
    return;
  }
  m->time += m->dt;
  m->st.steps++;
  if (m->dt_save > 0.) { m->time = m->final_time; m->dt = m->dt_save; m->dt_save = 0.; }
  m->dt *= m->growth;
  if (m->dt > m->dt_max) m->dt = m->dt_max;
  m->phase = m->time < m->final_time ? BatchMember::kStep : BatchMember::kDone;
}

// The Krylov solve of a Newton iteration up to the full step (runs in the
// member's host thread): GMRES on J Y = F in the requested precision, then
// the tail up to W = X + Y.
static void iter_solve(BatchMember *m) {
  tdyhip_ctx *ctx = m->ctx;
  SolverWS &ws = *ctx->ws;
  hipStream_t st = ctx->stream;
  tdyhip_solver_opts ok = m->o;
  if (ok.krylov_fp32) ok.ksp_max_it = std::min(m->o.ksp_max_it, kFp32MaxIt);
  int its = 0;
  const int rc = ok.krylov_fp32 ? gmres<float>(ctx, ok, ws.F, ws.Y, &its, st, true)
                                : gmres<double>(ctx, ok, ws.F, ws.Y, &its, st, true);
  if (rc && rc != TDYHIP_ERR_SINGULAR && rc != TDYHIP_ERR_NOCONV) { m->rc = rc; m->err = thread_error(); return; }
  newton_tail_pre(ctx, m->U, st);
  if (hipGetLastError() != hipSuccess) { m->rc = TDYHIP_ERR_HIP; m->err = "Newton tail launch failed"; }
}

// After the fused residual at the full step: the rest of the Newton
// iteration (runs in the member's host thread), as in the single-request
// solver: the tail, the FP64 retry of a failed FP32 solve, the line search
// and the convergence tests.
static void iter_finish(BatchMember *m) {
  tdyhip_ctx *ctx = m->ctx;
  SolverWS &ws = *ctx->ws;
  hipStream_t st = ctx->stream;
  const tdyhip_solver_opts &o = m->o;
  const Form fm{1, nullptr, 0.0};
  const int n = ctx->nc;
  const unsigned g = nblk(n, 256);
  int rc;
  auto fail = [&](int code) { m->rc = code; m->err = thread_error(); m->phase = BatchMember::kDone; };
  newton_tail_post(ctx, m->U, st);
  NewtonDev h;
  if ((rc = read_newton(ctx, h, st))) return fail(rc);
  int lits = h.its;
  ws.cycles_hint = std::max(1, (h.its + o.ksp_restart - 1) / o.ksp_restart);
  if (o.krylov_fp32 && !h.lin_ok) {
    m->st.krylov_fallbacks++;
    int its = 0;
    rc = gmres<double>(ctx, o, ws.F, ws.Y, &its, st, true);
    if (rc && rc != TDYHIP_ERR_SINGULAR && rc != TDYHIP_ERR_NOCONV) return fail(rc);
    if ((rc = newton_tail(ctx, fm, m->U, st)) || (rc = read_newton(ctx, h, st))) return fail(rc);
    lits += its;
  }
  m->st.linear_its += lits;
  const int it = m->it;
  int status = 0;
  bool stop = false;
  if (!h.lin_ok || h.ystate == 2) { status = -3; stop = true; }
  else if (h.ystate == 1) { stop = true; }
  else {
    double ynorm = h.ynorm, lambda = 1.0, gnorm = h.gnorm, xnorm = h.xnorm;
    bool ls_done = false;
    if (h.not_accepted) {
      auto trial = [&](double lam, double *gn) -> int {
        KT(ctx, st, "k_waxpy", (double)n * 24, k_waxpy<<<g, 256, 0, st>>>(n, lam, m->U, ws.Y, ws.W));
        int r2 = eval_F(ctx, fm, ws.W, ws.G, st);
        return r2 ? r2 : dev_norm(ctx, ws.G, gn, st);
      };
      auto cancel = [&]() -> int {
        if (!ctx->cancel.load(std::memory_order_relaxed)) return 0;
        set_error("solve cancelled");
        return TDYHIP_ERR_CANCELLED;
      };
      BTParams btp;
      btp.alpha = kBTAlpha;
      btp.minlambda = 1e-12;
      BTResult bt;
      if ((rc = bt_search(m->fnorm, h.initslope, btp, true, h.gnorm, trial, cancel, bt))) return fail(rc);
      if (!bt.accepted) {
        status = (ynorm < o.snes_stol * h.xnorm) ? 0 : -6;
        stop = true;
        ls_done = true;
      } else {
        lambda = bt.lambda;
        gnorm = bt.gnorm;
        if (hipMemcpyAsync(m->U, ws.W, sizeof(double) * n, hipMemcpyDeviceToDevice, st) != hipSuccess ||
            hipMemcpyAsync(ws.F, ws.G, sizeof(double) * n, hipMemcpyDeviceToDevice, st) != hipSuccess)
          return fail(TDYHIP_ERR_HIP);
        k_set_scalar<<<1, 1, 0, st>>>(&ws.nd->fnorm, gnorm);
        if ((rc = dev_norm(ctx, m->U, &xnorm, st))) return fail(rc);
      }
    }
    if (!ls_done) {
      m->fnorm = gnorm;
      ynorm *= lambda;
      if (!std::isfinite(m->fnorm)) { status = -4; stop = true; }
      else if (m->fnorm < o.snes_atol) stop = true;
      else if (m->fnorm <= m->ttol) stop = true;
      else if (ynorm < o.snes_stol * xnorm) stop = true;
      else if (m->fnorm > 1e4 * m->fnorm0) { status = -2; stop = true; }
    }
  }
  if (stop) return finish_newton(m, status, it);
  if (it >= o.snes_max_it) return finish_newton(m, -5, o.snes_max_it);
  m->it = it + 1;
}

// The start of a time step (TDyTimeIntegratorRunToTime's outer loop) and of
// an attempt at it: save the state, the SNES pre-solve, the solver's dt.
static int begin_attempt(BatchMember *m) {
  tdyhip_ctx *ctx = m->ctx;
  SolverWS &ws = *ctx->ws;
  hipStream_t st = ctx->stream;
  int rc;
  if (m->phase == BatchMember::kStep) {
    const double dt_for_sync = m->final_time - m->time;
    if (m->dt > dt_for_sync) { m->dt_save = m->dt; m->dt = dt_for_sync; }
    if ((rc = tdyhip_set_dtime(ctx, m->dt))) return rc;
  }
  SCHK(hipMemcpyAsync(ws.Usave, m->U, sizeof(double) * ctx->nc, hipMemcpyDeviceToDevice, st));
  if ((rc = tdyhip_snes_presolve(ctx, m->U, st))) return rc;
  k_set_dyn<<<1, 1, 0, st>>>(ws.dyn, ctx->dtime, 0.0);
  SCHK(hipGetLastError());
  return 0;
}

}  // namespace tdyhip

extern "C" {

int tdyhip_batch_create(tdyhip_batch **out) {
  using namespace tdyhip;
  if (!out) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  *out = nullptr;
  tdyhip_batch *b = new tdyhip_batch();
  auto fail = [&](const char *what) {
    set_error(what);
    tdyhip_batch_destroy(b);
    return TDYHIP_ERR_HIP;
  };
  if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) { b->st = nullptr; return fail("hipStreamCreate failed"); }
  if (hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess) { b->done = nullptr; return fail("hipEventCreate failed"); }
  for (int t = 0; t < 3; t++) {
    if (hipMalloc((void **)&b->d_prob[t], sizeof(SegProblem) * kMaxSeg) != hipSuccess) { b->d_prob[t] = nullptr; return fail("hipMalloc failed"); }
    if (hipHostMalloc((void **)&b->h_prob[t], sizeof(SegProblem) * kMaxSeg, hipHostMallocDefault) != hipSuccess) {
      b->h_prob[t] = nullptr;
      return fail("hipHostMalloc failed");
    }
  }
  *out = b;
  return 0;
}

int tdyhip_batch_destroy(tdyhip_batch *b) {
  using namespace tdyhip;
  if (!b) return 0;
  if (b->st) (void)hipStreamSynchronize(b->st);
  for (BatchMember *m : b->members) {
    (void)hipStreamSynchronize(m->ctx->stream);
    delete m;
  }
  for (int t = 0; t < 3; t++) {
    if (b->d_prob[t]) (void)hipFree(b->d_prob[t]);
    if (b->h_prob[t]) (void)hipHostFree(b->h_prob[t]);
  }
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->st) (void)hipStreamDestroy(b->st);
  delete b;
  return 0;
}

int tdyhip_batch_add(tdyhip_batch *b, tdyhip_ctx *ctx, double *U, const tdyhip_solver_opts *o, double dt_init,
                     double growth, double dt_max, double final_time, int32_t *id) {
  using namespace tdyhip;
  if (!b || !ctx || !U || !o || !id) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  if (!(dt_init > 0) || !(growth > 0) || !(dt_max > 0) || !(final_time > 0)) { set_error("bad time stepping"); return TDYHIP_ERR_ARG; }
  if (ctx->general) { set_error("a batch takes hexahedral meshes within the fixed-size limits"); return TDYHIP_ERR_ARG; }
  if (ctx->fuser) { set_error("the context is attached to an evaluation fuser"); return TDYHIP_ERR_ARG; }
  if ((int)b->members.size() >= kMaxSeg) { set_error("batch is full (32 requests)"); return TDYHIP_ERR_ARG; }
  for (BatchMember *m : b->members)
    if (m->ctx == ctx) { set_error("context already in the batch"); return TDYHIP_ERR_ARG; }
  int rc = check_opts(o);
  if (rc || (rc = ensure_ws(ctx, *o))) return rc;
  if (!ctx->ws_event && hipEventCreateWithFlags(&ctx->ws_event, hipEventDisableTiming) != hipSuccess) {
    ctx->ws_event = nullptr;
    set_error("hipEventCreate failed");
    return TDYHIP_ERR_HIP;
  }
  BatchMember *m = new BatchMember();
  m->id = b->next_id++;
  m->ctx = ctx;
  m->U = U;
  m->o = *o;
  m->growth = growth;
  m->dt_max = dt_max;
  m->final_time = final_time;
  m->dt = dt_init;
  b->members.push_back(m);
  *id = m->id;
  return 0;
}

int tdyhip_batch_remove(tdyhip_batch *b, int32_t id) {
  using namespace tdyhip;
  if (!b) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  for (size_t i = 0; i < b->members.size(); i++)
    if (b->members[i]->id == id) {
      (void)hipStreamSynchronize(b->members[i]->ctx->stream);
      delete b->members[i];
      b->members.erase(b->members.begin() + i);
      return 0;
    }
  set_error("no such request in the batch");
  return TDYHIP_ERR_ARG;
}

int tdyhip_batch_size(const tdyhip_batch *b, int32_t *n) {
  if (!b || !n) { tdyhip::set_error("null argument"); return TDYHIP_ERR_ARG; }
  *n = (int32_t)b->members.size();
  return 0;
}

int tdyhip_batch_round(tdyhip_batch *b, int32_t cap, int32_t *ids, int32_t *rcs, tdyhip_solver_stats *stats,
                       int32_t *ndone) {
  using namespace tdyhip;
  if (!b || !ids || !rcs || !stats || !ndone) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  if (cap < (int32_t)b->members.size()) { set_error("result arrays shorter than the batch"); return TDYHIP_ERR_ARG; }
  *ndone = 0;
  int rc;
  auto member_fail = [](BatchMember *m, int code) {
    m->rc = code;
    m->err = thread_error();
    m->phase = BatchMember::kDone;
  };
  // 1. requests starting a step or a retry: pre-solve, then the fused
  //    residual at U and its norm (the start of the Newton solve)
  std::vector<BatchMember *> starting;
  for (BatchMember *m : b->members) {
    if (m->phase != BatchMember::kStep && m->phase != BatchMember::kAttempt) continue;
    if (m->ctx->cancel.load(std::memory_order_relaxed)) { set_error("solve cancelled"); member_fail(m, TDYHIP_ERR_CANCELLED); continue; }
    if ((rc = begin_attempt(m))) { member_fail(m, rc); continue; }
    m->phase = BatchMember::kAttempt;
    starting.push_back(m);
  }
  if ((rc = fused_eval(b, 0, starting, false, [](BatchMember *m) { return (const double *)m->U; },
                       [](BatchMember *m) { return m->ctx->ws->F; }, [](BatchMember *) { return (const int *)nullptr; })))
    return rc;
  for (BatchMember *m : starting) {
    SolverWS &ws = *m->ctx->ws;
    hipStream_t st = m->ctx->stream;
    dot_to(m->ctx, ws.F, ws.F, 0, nullptr, st);
    k_nt_set_fnorm<<<1, 1, 0, st>>>(ws.nd, ws.nh);
    if (hipMemcpyAsync(&m->fnorm2_h, ws.nh, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess) {
      member_fail(m, TDYHIP_ERR_HIP);
    }
  }
  for (BatchMember *m : starting) {
    if (m->phase == BatchMember::kDone) continue;
    if (hipStreamSynchronize(m->ctx->stream) != hipSuccess) { set_error("solve failed on the device"); member_fail(m, TDYHIP_ERR_HIP); continue; }
    m->fnorm = std::sqrt(m->fnorm2_h);
    m->fnorm0 = m->fnorm;
    m->ttol = m->o.snes_rtol * m->fnorm;
    if (!std::isfinite(m->fnorm)) finish_newton(m, -4, 0);          // SNES_DIVERGED_FNORM_NAN
    else if (m->fnorm < m->o.snes_atol) finish_newton(m, 0, 0);
    else { m->phase = BatchMember::kIter; m->it = 1; }
  }
  // 2. one Newton iteration of every request in its solve: the fused
  //    Jacobian, the Krylov solves side by side, the fused residual at the
  //    full steps, the rest of each iteration side by side
  std::vector<BatchMember *> iter;
  for (BatchMember *m : b->members) {
    if (m->phase != BatchMember::kIter) continue;
    if (m->ctx->cancel.load(std::memory_order_relaxed)) { set_error("solve cancelled"); member_fail(m, TDYHIP_ERR_CANCELLED); continue; }
    iter.push_back(m);
  }
  if (!iter.empty()) {
    if ((rc = fused_eval(b, 1, iter, true, [](BatchMember *m) { return (const double *)m->U; },
                         [](BatchMember *) { return (double *)nullptr; }, [](BatchMember *) { return (const int *)nullptr; })))
      return rc;
    parallel_members(iter, iter_solve);
    std::vector<BatchMember *> ok;
    for (BatchMember *m : iter) {
      if (m->rc) m->phase = BatchMember::kDone;
      else ok.push_back(m);
    }
    if ((rc = fused_eval(b, 2, ok, false, [](BatchMember *m) { return (const double *)m->ctx->ws->W; },
                         [](BatchMember *m) { return m->ctx->ws->G; },
                         [](BatchMember *m) { return (const int *)&m->ctx->ws->nd->skip; })))
      return rc;
    parallel_members(ok, iter_finish);
  }
  // 3. requests that finished (or failed) leave the batch
  std::vector<BatchMember *> keep;
  for (BatchMember *m : b->members) {
    if (m->phase != BatchMember::kDone) { keep.push_back(m); continue; }
    (void)hipStreamSynchronize(m->ctx->stream);
    const int k = *ndone;
    ids[k] = m->id;
    rcs[k] = m->rc;
    stats[k] = m->st;
    (*ndone)++;
    if (m->rc) set_error(m->err);
    delete m;
  }
  b->members.swap(keep);
  SCHK(hipStreamSynchronize(b->st));
  return 0;
}

}  // extern "C"
