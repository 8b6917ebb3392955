// Evaluation fuser: the SNES-form residual and Jacobian evaluations of the
// contexts attached to it (tdyhip_set_fuser) are launched by one thread,
// which takes every evaluation pending when it wakes and runs them as
// segmented launches (k_seg_*, one per kernel and kind) on its own stream.
// Each attached context keeps its own solver, thread and pace: an evaluation
// enqueues a request and waits only until the fused launch is enqueued and
// the context's stream made to wait for it.  Every segment computes what the
// single-context kernels compute, so results are the same bit for bit as
// without the fuser.

struct tdyhip_fuser {
  struct Req {
    tdyhip_ctx *ctx;
    int kind;  // 0 residual at X into out, 1 Jacobian at X (the solver's ELL array)
    const double *X;
    double *out;
    const int *stop;
    std::atomic<int> launched{0};
    int rc = 0;
  };
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::vector<Req *> pending;
  int attached = 0;     // contexts attached
  int window_us = 0;    // how long the first pending evaluation waits for others
  bool quit = false;
  std::thread th;
  hipStream_t st = nullptr;
  static constexpr int kRing = 8;
  tdyhip::SegProblem *d_tab[kRing] = {};
  tdyhip::SegProblem *h_tab[kRing] = {};
  hipEvent_t ring_ev[kRing] = {};
  int ring_pos = 0;
  // statistics
  std::atomic<int64_t> launches{0}, evals{0}, max_segments{0};
};

namespace tdyhip {

// one fused launch group of `rs` (all of one kind) from the fuser thread
static int fuser_launch(tdyhip_fuser *fz, std::vector<tdyhip_fuser::Req *> &rs) {
  const int n = (int)rs.size();
  const int slot = fz->ring_pos;
  fz->ring_pos = (fz->ring_pos + 1) % tdyhip_fuser::kRing;
  SCHK(hipEventSynchronize(fz->ring_ev[slot]));  // the slot's previous launch has read its table
  SegProblem *hp = fz->h_tab[slot];
  const bool jac = rs[0]->kind == 1;
  SegBlocks s_state{}, s_bnd{}, s_rows{}, s_gather{}, s_jac{};
  s_state.nseg = s_bnd.nseg = s_rows.nseg = s_gather.nseg = s_jac.nseg = n;
  for (int i = 0; i < n; i++) {
    tdyhip_ctx *ctx = rs[i]->ctx;
    SolverWS &ws = *ctx->ws;
    const double *X = rs[i]->X;
    SegProblem &p = hp[i];
    p.cs = cell_state_args(ctx, 1, X, nullptr);
    p.bs = bnd_state_args(ctx, 1, X);
    p.ra = row_args(ctx, X);
    p.ga = jac ? gather_args(ctx, 1, 0.0, nullptr, nullptr, rs[i]->out, true, ws.d_perm, ws.bj.bs)
               : gather_args(ctx, 1, 0.0, nullptr, rs[i]->out, nullptr, false, nullptr, 0);
    p.cj = cell_jac_args(ctx, ws.bj.d_row, ws.NB);
    p.stop = rs[i]->stop;
    s_state.start[i + 1] = s_state.start[i] + nblk(ctx->nc, 256);
    s_bnd.start[i + 1] = s_bnd.start[i] + nblk(ctx->nbnd, 256);
    s_rows.start[i + 1] = s_rows.start[i] + nblk((int64_t)4 * ctx->nf, kRowBlock);
    s_gather.start[i + 1] = s_gather.start[i] + nblk(ctx->nc, kGatherBlock);
    s_jac.start[i + 1] = s_jac.start[i] + nblk(ws.NB, kCellJacCells);
    SCHK(hipStreamWaitEvent(fz->st, ctx->ws_event, 0));
  }
  SCHK(hipMemcpyAsync(fz->d_tab[slot], hp, sizeof(SegProblem) * n, hipMemcpyHostToDevice, fz->st));
  const SegProblem *dp = fz->d_tab[slot];
  if (jac) {
    k_seg_cell_jac<<<(unsigned)s_jac.start[n], kCellJacBlock, 0, fz->st>>>(dp, s_jac);
  } else {
    k_seg_state<<<(unsigned)s_state.start[n], 256, 0, fz->st>>>(dp, s_state);
    if (s_bnd.start[n] > 0) k_seg_bnd_state<<<(unsigned)s_bnd.start[n], 256, 0, fz->st>>>(dp, s_bnd);
    k_seg_rows<<<(unsigned)s_rows.start[n], kRowBlock, 0, fz->st>>>(dp, s_rows);
    k_seg_cell_gather<<<(unsigned)s_gather.start[n], kGatherBlock, 0, fz->st>>>(dp, s_gather);
  }
  SCHK(hipGetLastError());
  SCHK(hipEventRecord(fz->ring_ev[slot], fz->st));
  // each request's stream waits for the launch before its next work (the
  // request's thread is blocked meanwhile, so nothing else enqueues there)
  for (int i = 0; i < n; i++) SCHK(hipStreamWaitEvent(rs[i]->ctx->stream, fz->ring_ev[slot], 0));
  fz->launches++;
  fz->evals += n;
  int64_t mx = fz->max_segments.load();
  while (n > mx && !fz->max_segments.compare_exchange_weak(mx, n)) {
  }
  return 0;
}

static void fuser_loop(tdyhip_fuser *fz) {
  std::unique_lock<std::mutex> lk(fz->mu);
  for (;;) {
    fz->cv_work.wait(lk, [&] { return fz->quit || !fz->pending.empty(); });
    if (fz->pending.empty()) return;  // quit
    // collect the evaluations that arrive within the window, or until every
    // attached context has one pending
    if (fz->window_us > 0)
      fz->cv_work.wait_until(lk, std::chrono::steady_clock::now() + std::chrono::microseconds(fz->window_us),
                             [&] { return fz->quit || (int)fz->pending.size() >= fz->attached; });
    std::vector<tdyhip_fuser::Req *> take;
    take.swap(fz->pending);
    lk.unlock();
    // one launch group per kind, at most kMaxSeg requests each
    for (int kind = 0; kind < 2; kind++) {
      std::vector<tdyhip_fuser::Req *> group;
      for (tdyhip_fuser::Req *r : take) {
        if (r->kind != kind) continue;
        group.push_back(r);
        if ((int)group.size() == kMaxSeg) {
          const int rc = fuser_launch(fz, group);
          for (tdyhip_fuser::Req *g : group) g->rc = rc;
          group.clear();
        }
      }
      if (!group.empty()) {
        const int rc = fuser_launch(fz, group);
        for (tdyhip_fuser::Req *g : group) g->rc = rc;
      }
    }
    lk.lock();
    for (tdyhip_fuser::Req *r : take) r->launched.store(1, std::memory_order_release);
    fz->cv_done.notify_all();
  }
}

// An evaluation of a context attached to a fuser (called from the context's
// solver thread in place of its own launches): the residual (kind 0) at X
// into out, or the Jacobian (kind 1) at X into the solver's ELL array out;
// skipped when *stop is set.
static int fuser_eval(tdyhip_ctx *ctx, int kind, const double *X, double *out, const int *stop) {
  tdyhip_fuser *fz = ctx->fuser;
  SCHK(hipEventRecord(ctx->ws_event, ctx->stream));
  tdyhip_fuser::Req r;
  r.ctx = ctx; r.kind = kind; r.X = X; r.out = out; r.stop = stop;
  std::unique_lock<std::mutex> lk(fz->mu);
  fz->pending.push_back(&r);
  fz->cv_work.notify_one();
  fz->cv_done.wait(lk, [&] { return r.launched.load(std::memory_order_acquire) != 0; });
  if (r.rc) set_error("fused evaluation launch failed");
  return r.rc;
}

}  // namespace tdyhip

extern "C" {

int tdyhip_fuser_create(tdyhip_fuser **out) {
  using namespace tdyhip;
  if (!out) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  *out = nullptr;
  tdyhip_fuser *fz = new tdyhip_fuser();
  auto fail = [&](const char *what) {
    set_error(what);
    tdyhip_fuser_destroy(fz);
    return TDYHIP_ERR_HIP;
  };
  if (hipStreamCreateWithFlags(&fz->st, hipStreamNonBlocking) != hipSuccess) { fz->st = nullptr; return fail("hipStreamCreate failed"); }
  for (int i = 0; i < tdyhip_fuser::kRing; i++) {
    if (hipEventCreateWithFlags(&fz->ring_ev[i], hipEventDisableTiming) != hipSuccess) { fz->ring_ev[i] = nullptr; return fail("hipEventCreate failed"); }
    if (hipMalloc((void **)&fz->d_tab[i], sizeof(SegProblem) * kMaxSeg) != hipSuccess) { fz->d_tab[i] = nullptr; return fail("hipMalloc failed"); }
    if (hipHostMalloc((void **)&fz->h_tab[i], sizeof(SegProblem) * kMaxSeg, hipHostMallocDefault) != hipSuccess) {
      fz->h_tab[i] = nullptr;
      return fail("hipHostMalloc failed");
    }
  }
  fz->th = std::thread(fuser_loop, fz);
  *out = fz;
  return 0;
}

int tdyhip_fuser_destroy(tdyhip_fuser *fz) {
  if (!fz) return 0;
  if (fz->th.joinable()) {
    {
      std::lock_guard<std::mutex> lk(fz->mu);
      fz->quit = true;
    }
    fz->cv_work.notify_all();
    fz->th.join();
  }
  if (fz->st) (void)hipStreamSynchronize(fz->st);
  for (int i = 0; i < tdyhip_fuser::kRing; i++) {
    if (fz->ring_ev[i]) (void)hipEventDestroy(fz->ring_ev[i]);
    if (fz->d_tab[i]) (void)hipFree(fz->d_tab[i]);
    if (fz->h_tab[i]) (void)hipHostFree(fz->h_tab[i]);
  }
  if (fz->st) (void)hipStreamDestroy(fz->st);
  delete fz;
  return 0;
}

int tdyhip_set_fuser(tdyhip_ctx *ctx, tdyhip_fuser *fz) {
  using namespace tdyhip;
  if (!ctx) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  if (fz && ctx->general) { set_error("the fuser takes hexahedral meshes within the fixed-size limits"); return TDYHIP_ERR_ARG; }
  if (fz && !ctx->ws_event && hipEventCreateWithFlags(&ctx->ws_event, hipEventDisableTiming) != hipSuccess) {
    ctx->ws_event = nullptr;
    set_error("hipEventCreate failed");
    return TDYHIP_ERR_HIP;
  }
  if (ctx->fuser == fz) return 0;
  if (ctx->fuser) {
    std::lock_guard<std::mutex> lk(ctx->fuser->mu);
    ctx->fuser->attached--;
  }
  if (fz) {
    std::lock_guard<std::mutex> lk(fz->mu);
    fz->attached++;
  }
  ctx->fuser = fz;
  return 0;
}

int tdyhip_fuser_set_window(tdyhip_fuser *fz, int32_t us) {
  if (!fz || us < 0) { tdyhip::set_error("bad argument"); return TDYHIP_ERR_ARG; }
  std::lock_guard<std::mutex> lk(fz->mu);
  fz->window_us = us;
  return 0;
}

int tdyhip_fuser_stats(const tdyhip_fuser *fz, int64_t *out) {
  if (!fz || !out) { tdyhip::set_error("null argument"); return TDYHIP_ERR_ARG; }
  out[0] = fz->launches.load();
  out[1] = fz->evals.load();
  out[2] = fz->max_segments.load();
  return 0;
}

}  // extern "C"
