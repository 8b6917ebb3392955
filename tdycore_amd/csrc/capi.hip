// C ABI of the MI355X MPFA-O Richards library (include/tdyhip.h).
// Unity build: the kernel files are compiled into this translation unit.
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <functional>
#include <map>
#include <mutex>
#include <thread>

#include "flow_kernels.hip"
#include "setup_kernels.hip"
#include "setup_general.hip"
#include "tdyhip_internal.h"

namespace tdyhip {
struct SolverWS;
static void ws_destroy(SolverWS *ws);
}  // namespace tdyhip

namespace tdyhip {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

struct HipError {};

#define HIPCHK(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) {                                                               \
      set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                      \
      return e_ == hipErrorOutOfMemory ? TDYHIP_ERR_MEM : TDYHIP_ERR_HIP;                 \
    }                                                                                     \
  } while (0)

// Host<->device copies are ordered on the context's own (non-blocking) stream
// and waited for there; the legacy null stream is never used, so a copy in one
// context cannot interfere with a graph capture running in another thread.
static hipError_t copy_sync(const tdyhip_ctx *ctx, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, ctx->stream);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(ctx->stream);
}

template <class T>
static int dalloc(tdyhip_ctx *ctx, T **p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? TDYHIP_ERR_MEM : TDYHIP_ERR_HIP;
  }
  ctx->allocations.push_back(*p);
  ctx->dev_bytes += (int64_t)(n * sizeof(T));
  return 0;
}

template <class T>
static int dupload(tdyhip_ctx *ctx, T **p, const T *h, size_t n) {
  int rc = dalloc(ctx, p, n);
  if (rc) return rc;
  if (n) HIPCHK(copy_sync(ctx, *p, h, n * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// cell-centroid heights and face areas, kept for velocity recovery
static int upload_geometry_for_velocity(tdyhip_ctx *ctx, const tdyhip_mesh *m) {
  std::vector<double> z((size_t)m->nc);
  for (int c = 0; c < m->nc; c++) z[c] = m->cell_cen[(size_t)3 * c + 2];
  int rc = dupload(ctx, &ctx->d_cell_z, z.data(), z.size());
  return rc ? rc : dupload(ctx, &ctx->d_face_area, m->face_area, (size_t)m->nf);
}

// ---- kernel timers (events on the launching stream) ----
static hipEvent_t kt_event(tdyhip_ctx *ctx) {
  hipEvent_t e = nullptr;
  if (!ctx->kt.pool.empty()) {
    e = ctx->kt.pool.back();
    ctx->kt.pool.pop_back();
  } else if (hipEventCreate(&e) != hipSuccess) {
    e = nullptr;
  }
  return e;
}
static hipEvent_t kt_begin(tdyhip_ctx *ctx, hipStream_t st) {
  if (!ctx->kt.on) return nullptr;
  hipEvent_t e = kt_event(ctx);
  if (e) (void)hipEventRecord(e, st);
  return e;
}
static void kt_end(tdyhip_ctx *ctx, hipStream_t st, const char *name, double bytes, hipEvent_t e0) {
  if (!e0) return;
  hipEvent_t e1 = kt_event(ctx);
  if (!e1) return;
  (void)hipEventRecord(e1, st);
  KernelTimers::Rec &r = ctx->kt.get(name);
  r.launches++;
  r.bytes += bytes;
  r.pending.push_back({e0, e1, ctx->kt.tag, bytes});
}
// A GMRES cycle's launches are settled once its step count is known: those
// tagged with a step >= `working` returned at once on the stop flag and move
// to "<kernel> (exited at once)" with no bytes; so do the launches of the
// next cycle's start (tag kTagRestart) when the solve had finished.  The
// rest lose their tag.
constexpr int kTagRestart = 1 << 20;
static void kt_settle(tdyhip_ctx *ctx, int working, bool restart_worked) {
  if (!ctx->kt.on) return;
  std::vector<std::pair<std::string, KernelTimers::Pending>> moved;
  for (auto &nr : ctx->kt.recs) {
    auto &pend = nr.second.pending;
    size_t keep = 0;
    for (size_t i = 0; i < pend.size(); i++) {
      KernelTimers::Pending pe = pend[i];
      const bool exited = pe.tag == kTagRestart ? !restart_worked : pe.tag >= working;
      if (exited) {
        nr.second.launches--;
        nr.second.bytes -= pe.bytes;
        pe.tag = -1;
        pe.bytes = 0.0;
        moved.emplace_back(nr.first + " (exited at once)", pe);
      } else {
        pe.tag = -1;
        pend[keep++] = pe;
      }
    }
    pend.resize(keep);
  }
  for (auto &mp : moved) {
    KernelTimers::Rec &r = ctx->kt.get(mp.first);
    r.launches++;
    r.pending.push_back(mp.second);
  }
}
static void kt_collect(tdyhip_ctx *ctx) {
  for (auto &nr : ctx->kt.recs) {
    for (auto &pe : nr.second.pending) {
      float ms = 0.f;
      if (hipEventSynchronize(pe.e1) == hipSuccess && hipEventElapsedTime(&ms, pe.e0, pe.e1) == hipSuccess)
        nr.second.ms += ms;
      ctx->kt.pool.push_back(pe.e0);
      ctx->kt.pool.push_back(pe.e1);
    }
    nr.second.pending.clear();
  }
}
#define KT(ctx, st, name, bytes, ...)              \
  do {                                             \
    hipEvent_t kt_e0_ = kt_begin(ctx, st);         \
    __VA_ARGS__;                                   \
    kt_end(ctx, st, name, (double)(bytes), kt_e0_); \
  } while (0)

// algorithmic bytes per launch (each array counted once)
static double bytes_state(const tdyhip_ctx *ctx) {
  return (double)ctx->nc * (8 * 5 + 8 * 7) + (double)ctx->nbnd * (4 + 8 + 8 + 8 + 8);
}
// flux rows: T, G, row vertices, face cells and flags, vertex columns, U,
// rho, kr/mu at cells and boundary faces, P_bnd, and the rows' fluxes out
static double bytes_rows(const tdyhip_ctx *ctx) {
  return (double)ctx->trans_nnz * 8 + (double)ctx->active_rows * (8 + 4 + 8) + (double)ctx->nf * (8 + 1) +
         (double)ctx->nv * kMaxCols * 4 + (double)ctx->nc * 8 * 3 + (double)ctx->nbnd * 8 * 3;
}
// residual gather: cell faces, face cells, fluxes, accumulation, R out
static double bytes_gather(const tdyhip_ctx *ctx) {
  return (double)ctx->nc * (24 + 8 + 8) + (double)ctx->nf * (8 + 32);
}
// single-pass Jacobian (+ residual): the flux rows' inputs once (T, G, row
// vertices, up/down region indices, face cells and flags, vertex columns and
// slot map, U, rho, kr/mu and their derivatives, boundary values), cell faces,
// the accumulation inputs of the form, the diagonal slot and J written once
static double bytes_cell_jac(const tdyhip_ctx *ctx, int form, bool R) {
  double b = (double)ctx->trans_nnz * 8 + (double)ctx->active_rows * (8 + 4 + 2) + (double)ctx->nf * (8 + 1) +
             (double)ctx->nv * (kMaxCols * 4 + 64) + (double)ctx->nc * (8 * 5 + 24 + 1) +
             (double)ctx->nbnd * 8 * 3 + (double)ctx->nc * 8 * (form == 0 ? 4 : 2) + (double)ctx->jnnz * 8;
  if (R) b += (double)ctx->nc * (8 + 8);
  return b;
}

static StateParams state_params(const tdyhip_ctx *ctx, int form, double shift) {
  StateParams sp;
  sp.vg_m = ctx->vg_m; sp.vg_alpha = ctx->vg_alpha; sp.x0 = ctx->x0; sp.x1 = ctx->x1; sp.dx = ctx->dx;
  sp.pref = ctx->pref; sp.dtime = ctx->dtime; sp.shift = shift;
  sp.c0 = ctx->mualem_c[0]; sp.c1 = ctx->mualem_c[1]; sp.c2 = ctx->mualem_c[2]; sp.c3 = ctx->mualem_c[3];
  sp.density_type = ctx->density_type; sp.form = form;
  sp.dyn = ctx->launch_dyn; sp.stop = ctx->launch_stop;
  return sp;
}

// Mualem cubic coefficients: the 4x4 system of CubicPolynomialSetup
// (src/materials/tdycharacteristiccurves.c:509-535) solved by Gaussian
// elimination with partial pivoting.
static int mualem_cubic(double m, double x0, double x1, double x2, double dx, double *c) {
  double Kr, dKr;
  mualem_kr(m, x0, 2.0, dx, c, x1, Kr, dKr);  // x1 < 2: unsmoothed branch
  double b[4] = {1.0, Kr, 0.0, dKr};
  const double t1 = (x1 - x0) / dx, t2 = (x2 - x0) / dx;
  double A[4][4] = {{1.0, t2, t2 * t2, t2 * t2 * t2},
                    {1.0, t1, t1 * t1, t1 * t1 * t1},
                    {0.0, 1.0 / dx, 2.0 * t2 / dx, 3.0 * t2 * t2 / dx},
                    {0.0, 1.0 / dx, 2.0 * t1 / dx, 3.0 * t1 * t1 / dx}};
  for (int k = 0; k < 4; k++) {
    int p = k;
    for (int i = k + 1; i < 4; i++)
      if (fabs(A[i][k]) > fabs(A[p][k])) p = i;
    if (A[p][k] == 0.0) return 1;
    if (p != k) {
      for (int j = 0; j < 4; j++) std::swap(A[k][j], A[p][j]);
      std::swap(b[k], b[p]);
    }
    for (int i = k + 1; i < 4; i++) {
      const double l = A[i][k] / A[k][k];
      for (int j = k; j < 4; j++) A[i][j] -= l * A[k][j];
      b[i] -= l * b[k];
    }
  }
  for (int i = 3; i >= 0; i--) {
    double s = b[i];
    for (int j = i + 1; j < 4; j++) s -= A[i][j] * c[j];
    c[i] = s / A[i][i];
  }
  return 0;
}

static int check_stream(tdyhip_ctx *ctx, tdyhip_stream s, hipStream_t *out) {
  *out = s ? (hipStream_t)s : ctx->stream;
  return 0;
}

static RowArgs row_args(tdyhip_ctx *ctx, const double *U) {
  RowArgs a;
  a.nc = ctx->nc; a.nrows = (int64_t)4 * ctx->nf; a.pref = ctx->pref;
  a.face_cells = ctx->d_face_cells; a.face_verts = ctx->d_face_verts; a.face_flags = ctx->d_face_flags;
  a.vcols = ctx->d_vcols; a.T = ctx->d_T; a.G = ctx->d_G; a.U = U; a.Pbnd = ctx->d_Pbnd;
  a.rho = ctx->d_rho; a.ukvr = ctx->d_ukvr; a.drho = ctx->d_drho; a.dukvr = ctx->d_dukvr;
  a.rho_bnd = ctx->d_rho_bnd; a.ukvr_bnd = ctx->d_ukvr_bnd;
  a.flux = ctx->launch_flux ? ctx->launch_flux : ctx->d_flux;
  a.stop = ctx->launch_stop;
  return a;
}

static GatherArgs gather_args(tdyhip_ctx *ctx, int form, double shift, const double *Udot, double *R, double *J,
                              bool ell, const int32_t *perm, int ell_bs) {
  GatherArgs a;
  a.nc = ctx->nc; a.nrows = (int64_t)4 * ctx->nf; a.form = form; a.shift = shift; a.dtime = ctx->dtime;
  a.cell_faces = ctx->d_cell_faces; a.face_cells = ctx->d_face_cells;
  a.face_flags = ctx->d_face_flags;
  a.jdslot = ctx->d_jdslot; a.flux = ctx->launch_flux ? ctx->launch_flux : ctx->d_flux;
  a.acc = ctx->d_acc; a.dM = ctx->d_dM; a.d2M = ctx->d_d2M;
  a.src = ctx->has_source ? ctx->d_source : nullptr;
  a.Udot = Udot; a.vol = ctx->d_cell_vol;
  a.jrp = ell ? nullptr : ctx->d_jrp; a.jbs = ell ? ell_bs : 1; a.ell_w = ctx->jell_w;
  a.perm = ell ? perm : nullptr;
  a.R = R; a.J = J;
  a.dyn = ctx->launch_dyn; a.stop = ctx->launch_stop;
  return a;
}

static CellStateArgs cell_state_args(tdyhip_ctx *ctx, int form, const double *U, const double *Udot) {
  CellStateArgs a;
  a.nc = ctx->nc; a.sp = state_params(ctx, form, 0.0);
  a.U = U; a.Udot = Udot; a.phi = ctx->d_phi; a.Sr = ctx->d_Sr; a.vol = ctx->d_cell_vol; a.accum_prev = ctx->d_accum_prev;
  a.rho = ctx->d_rho; a.ukvr = ctx->d_ukvr; a.drho = ctx->d_drho; a.dukvr = ctx->d_dukvr;
  a.dM = ctx->d_dM; a.d2M = ctx->d_d2M; a.acc = ctx->d_acc;
  return a;
}
static BndStateArgs bnd_state_args(tdyhip_ctx *ctx, int form, const double *U) {
  BndStateArgs a;
  a.nbnd = ctx->nbnd; a.sp = state_params(ctx, form, 0.0);
  a.U = U; a.bnd_cell = ctx->d_bnd_cell; a.Sr = ctx->d_Sr; a.pbnd_fn = ctx->d_pbnd_fn;
  a.Pbnd = ctx->d_Pbnd; a.ukvr_bnd = ctx->d_ukvr_bnd;
  return a;
}

// state (+boundary) at U
static int launch_state(tdyhip_ctx *ctx, int form, const double *U, const double *Udot, hipStream_t st) {
  KT(ctx, st, "k_cell_state", bytes_state(ctx),
     k_cell_state<<<nblk(ctx->nc, 256), 256, 0, st>>>(cell_state_args(ctx, form, U, Udot)));
  if (ctx->nbnd > 0) k_bnd_state<<<nblk(ctx->nbnd, 256), 256, 0, st>>>(bnd_state_args(ctx, form, U));
  HIPCHK(hipGetLastError());
  return 0;
}

static CellJacArgs cell_jac_args(tdyhip_ctx *ctx, const int32_t *order, int64_t norder) {
  CellJacArgs cj;
  cj.order = order;
  cj.npos = order ? norder : ctx->nc;
  cj.jslot = ctx->d_jslot; cj.crow = ctx->d_crow; cj.cell_verts = ctx->d_cell_verts;
  cj.xcd_groups = 0;
  return cj;
}

// Residual (R) and / or Jacobian (J) at U.  The residual alone: flux rows,
// then the per-cell gather.  With J: one pass per cell (k_cell_jac).
// ell = true: J is the solver's block-contiguous ELL array (row c at position
// perm[c] of blocks of ell_bs rows), else CSR values in the pattern order.
// order (ELL only): the block layout's row list (position -> cell, -1
// padding, norder entries); the Jacobian's threads follow it so that a
// workgroup covers one block.  Every entry of J is written (no prior clear).
static int launch_flux_general(tdyhip_ctx *ctx, int form, double shift, const double *U, const double *Udot,
                               double *R, double *J, hipStream_t st, const int32_t *perm, int ell_bs);
static int launch_flux(tdyhip_ctx *ctx, int form, double shift, const double *U, const double *Udot, double *R,
                       double *J, hipStream_t st, bool ell = false, const int32_t *perm = nullptr,
                       int ell_bs = 0, const int32_t *order = nullptr, int64_t norder = 0) {
  if (ctx->general) return launch_flux_general(ctx, form, shift, U, Udot, R, J, st, ell ? perm : nullptr, ell_bs);
  const int64_t nrows = (int64_t)4 * ctx->nf;
  RowArgs a = row_args(ctx, U);
  GatherArgs ga = gather_args(ctx, form, shift, Udot, R, J, ell, perm, ell_bs);
  if (!J) {
    KT(ctx, st, "k_rows", bytes_rows(ctx), k_rows<<<nblk(nrows, kRowBlock), kRowBlock, 0, st>>>(a));
    KT(ctx, st, "k_cell_gather", bytes_gather(ctx),
       k_cell_gather<<<nblk(ctx->nc, kGatherBlock), kGatherBlock, 0, st>>>(ga));
  } else {
    CellJacArgs cj = cell_jac_args(ctx, ell ? order : nullptr, norder);
    const int64_t nb = (cj.npos + kCellJacCells - 1) / kCellJacCells;
    cj.xcd_groups = ctx->cj_xcd ? (int)((nb + 7) / 8) : 0;
    const unsigned g = cj.xcd_groups ? (unsigned)(8 * cj.xcd_groups) : (unsigned)nb;
    if (R)
      KT(ctx, st, "k_cell_jac_R", bytes_cell_jac(ctx, form, true), k_cell_jac<true><<<g, kCellJacBlock, 0, st>>>(a, ga, cj));
    else
      KT(ctx, st, "k_cell_jac", bytes_cell_jac(ctx, form, false), k_cell_jac<false><<<g, kCellJacBlock, 0, st>>>(a, ga, cj));
  }
  HIPCHK(hipGetLastError());
  return 0;
}

}  // namespace tdyhip

using namespace tdyhip;

namespace tdyhip {

// One persistent launch of the interaction-region solves (k_ir_persistent)
// over the buckets in `order` (bucket ids), each at its team width; the
// counter is zeroed first.  Every schedule, chunk and width computes the
// same values (one lane per equation, the same operations in the same order).
struct IRPlan {
  int lanes[3] = {0, 0, 0};
  int order[3] = {0, 1, 2};
  int chunk = 1;
};

static int ir_lds_bytes(const GenArgs (&ga)[3], const int (&lanes)[3]) {
  size_t lds = 8;
  for (int b = 0; b < 3; b++)
    if (ga[b].n > 0 && lanes[b]) lds = std::max(lds, (size_t)(64 / lanes[b]) * ga[b].maxU * ga[b].pitch * sizeof(double));
  return (int)lds;
}

static int launch_ir_persistent(const GenArgs (&ga)[3], const IRPlan &plan, int *d_counter, int grid,
                                hipStream_t st) {
  IRSchedule sc;
  for (int b = 0; b < 3; b++) { sc.ga[b] = ga[b]; sc.lanes[b] = plan.lanes[b]; }
  sc.nseg = 0;
  int items = 0;
  for (int k = 0; k < 3; k++) {
    const int b = plan.order[k];
    if (b < 0 || ga[b].n == 0 || !plan.lanes[b]) continue;
    const int teams = 64 / plan.lanes[b];
    const int nblk = (ga[b].n + teams - 1) / teams;
    items += (nblk + plan.chunk - 1) / plan.chunk;
    sc.seg_bucket[sc.nseg] = b;
    sc.seg_end[sc.nseg] = items;
    sc.nseg++;
  }
  if (sc.nseg == 0) return 0;
  sc.chunk = plan.chunk;
  sc.counter = d_counter;
  const int lds = ir_lds_bytes(ga, plan.lanes);
  if (lds > 65536 && hipFuncSetAttribute(reinterpret_cast<const void *>(&k_ir_persistent),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess) {
    set_error("interaction regions too large for the LDS of one workgroup");
    return TDYHIP_ERR_ARG;
  }
  if (hipMemsetAsync(d_counter, 0, sizeof(int), st) != hipSuccess) return TDYHIP_ERR_HIP;
  k_ir_persistent<<<(unsigned)std::max(1, std::min(grid, items)), 64, lds, st>>>(sc);
  return hipGetLastError() == hipSuccess ? 0 : TDYHIP_ERR_HIP;
}

// workgroups of k_ir_persistent resident at once on the device (its grid)
static int ir_grid(const GenArgs (&ga)[3], const IRPlan &plan) {
  int dev = 0, ncu = 0, per_cu = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0) ncu = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(&k_ir_persistent), 64,
                                                  (size_t)ir_lds_bytes(ga, plan.lanes)) != hipSuccess || per_cu <= 0)
    per_cu = 1;
  return ncu * per_cu;
}

// Autotuning of the launch, kept per device and bucket shapes for the rest of
// the process: first the team width of each bucket of at least kTuneMin
// regions (that bucket alone at every width that holds its largest region,
// 16, 32, 64 lanes; a wider team runs fewer regions per wavefront), then the
// bucket order (largest regions first or last) and the chunk of workgroup
// steps per work item (1 or 8) over all buckets together.  The tuning runs
// leave T complete.
constexpr int kTuneMin = 4096;
static std::mutex g_tune_mu;
static std::map<std::array<int, 10>, IRPlan> g_tune;

// Transmissibility setup of a general (prism) mesh: G matrices, the
// size-bucketed local systems and the gravity term, with HIP-event times of
// the three phases in ctx->setup_ms.
static int setup_general(tdyhip_ctx *ctx, const tdyhip_mesh *m, const tdyhip_materials *mat) {
  HostGeneral *g = new HostGeneral();
  ctx->hgen = g;
  ctx->general = true;
  std::string msg = build_topology_general(*m, *g);
  if (!msg.empty()) { set_error(msg); return TDYHIP_ERR_ARG; }
  const int nc = m->nc, nf = m->nf;
  int rc;
  if ((rc = dupload(ctx, &ctx->d_face_cells, m->face_cells, (size_t)2 * nf))) return rc;
  if ((rc = dupload(ctx, &ctx->d_face_flags, g->face_flags.data(), nf))) return rc;
  if ((rc = dalloc(ctx, &ctx->d_G, (size_t)4 * nf))) return rc;
  if ((rc = dalloc(ctx, &ctx->d_Tgen, (size_t)std::max<int64_t>(g->tsize, 1)))) return rc;
  std::vector<void *> tmp;
  auto done = [&](int code) {
    for (void *p : tmp) (void)hipFree(p);
    return code;
  };
  auto up = [&](auto **ptr, const auto *src, size_t n) -> int {
    hipError_t e = hipMalloc((void **)ptr, std::max<size_t>(n, 1) * sizeof(**ptr));
    if (e != hipSuccess) { set_error(std::string("hipMalloc: ") + hipGetErrorString(e)); return e == hipErrorOutOfMemory ? TDYHIP_ERR_MEM : TDYHIP_ERR_HIP; }
    tmp.push_back((void *)*ptr);
    if (src && n && copy_sync(ctx, *ptr, src, n * sizeof(**ptr), hipMemcpyHostToDevice) != hipSuccess) {
      set_error("hipMemcpy failed");
      return TDYHIP_ERR_HIP;
    }
    return 0;
  };
  int32_t *d_scf = nullptr, *d_vdesc = nullptr, *d_csub = nullptr, *d_csck = nullptr, *d_fa = nullptr,
          *d_fb = nullptr, *d_vl[3] = {nullptr, nullptr, nullptr};
  uint8_t *d_fnv = nullptr;
  double *d_cc = nullptr, *d_fc = nullptr, *d_fa2 = nullptr, *d_fn = nullptr, *d_K = nullptr, *d_Gsub = nullptr;
  int *d_status = nullptr;
  if ((rc = up(&d_scf, g->sc_faces.data(), g->sc_faces.size())) || (rc = up(&d_vdesc, g->vdesc.data(), g->vdesc.size())) ||
      (rc = up(&d_csub, g->csub.data(), g->csub.size())) || (rc = up(&d_csck, g->csck.data(), g->csck.size())) ||
      (rc = up(&d_fa, g->fa.data(), g->fa.size())) || (rc = up(&d_fb, g->fb.data(), g->fb.size())) ||
      (rc = up(&d_fnv, g->face_nv.data(), g->face_nv.size())) || (rc = up(&d_cc, m->cell_cen, (size_t)3 * nc)) ||
      (rc = up(&d_fc, m->face_cen, (size_t)3 * nf)) || (rc = up(&d_fa2, m->face_area, (size_t)nf)) ||
      (rc = up(&d_fn, m->face_normal, (size_t)3 * nf)) || (rc = up(&d_K, mat->K, (size_t)9 * nc)) ||
      (rc = up(&d_Gsub, (const double *)nullptr, (size_t)nc * 72)) || (rc = up(&d_status, (const int *)nullptr, 1)))
    return done(rc);
  for (int b = 0; b < 3; b++)
    if ((rc = up(&d_vl[b], g->bucket[b].data(), g->bucket[b].size()))) return done(rc);
  hipStream_t st = ctx->stream;
  hipEvent_t ev[4];
  for (auto &e : ev)
    if (hipEventCreate(&e) != hipSuccess) { set_error("hipEventCreate failed"); return done(TDYHIP_ERR_HIP); }
  (void)hipMemsetAsync(d_status, 0, sizeof(int), st);
  (void)hipEventRecord(ev[0], st);
  k_gmatrix<<<nblk((int64_t)nc * 8, 256), 256, 0, st>>>(nc, ctx->gmatrix, d_scf, ctx->d_face_cells, d_cc, d_fc, d_fa2,
                                                        d_fn, d_K, d_fnv, d_Gsub);
  (void)hipEventRecord(ev[1], st);
  (void)hipEventSynchronize(ev[1]);
  float gm_ms = 0.f;
  (void)hipEventElapsedTime(&gm_ms, ev[0], ev[1]);
  GenArgs ga[3];
  std::array<int, 10> key{};
  (void)hipGetDevice(&key[0]);
  for (int b = 0; b < 3; b++) {
    ga[b] = GenArgs{d_vl[b], d_vdesc, d_csub, d_csck, d_fa, d_fb, d_Gsub, ctx->d_Tgen, d_status,
                    (int)g->bucket[b].size(), 1, 1};
    int mu = 1, mb = 1;
    for (int v : g->bucket[b]) {
      mu = std::max(mu, g->vdesc[(size_t)v * 8 + 2]);
      mb = std::max(mb, g->vdesc[(size_t)v * 8 + 0] + g->vdesc[(size_t)v * 8 + 3]);
    }
    ga[b].maxU = mu;
    ga[b].pitch = (mu + mb) | 1;
    key[1 + 3 * b] = ga[b].n >= kTuneMin;  // buckets below the tuning size take the default width
    key[2 + 3 * b] = mu;
    key[3 + 3 * b] = mb;
  }
  int *d_counter = nullptr;
  if ((rc = up(&d_counter, (const int *)nullptr, 1))) return done(rc);
  IRPlan plan;
  bool cached = false;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) { plan = it->second; cached = true; }
  }
  for (int b = 0; b < 3; b++) {
    g->lanes[b] = 0;
    if (!cached) plan.lanes[b] = ga[b].n ? 16 << b : 0;
  }
  // TDYHIP_IR_PLAN="w0,w1,w2,largest_first,chunk" replaces the tuned plan
  // (widths of empty buckets are ignored; a width must hold its bucket)
  if (const char *e = getenv("TDYHIP_IR_PLAN")) {
    int w[3], desc, chunk;
    if (sscanf(e, "%d,%d,%d,%d,%d", &w[0], &w[1], &w[2], &desc, &chunk) != 5 || chunk < 1) {
      set_error("TDYHIP_IR_PLAN: expected w0,w1,w2,largest_first,chunk");
      return done(TDYHIP_ERR_ARG);
    }
    for (int b = 0; b < 3; b++) {
      if (!ga[b].n) continue;
      if (w[b] != 16 && w[b] != 32 && w[b] != 64) { set_error("TDYHIP_IR_PLAN: widths are 16, 32 or 64"); return done(TDYHIP_ERR_ARG); }
      if (w[b] < (16 << b)) { set_error("TDYHIP_IR_PLAN: a width smaller than its bucket's regions"); return done(TDYHIP_ERR_ARG); }
      plan.lanes[b] = w[b];
    }
    for (int k = 0; k < 3; k++) plan.order[k] = desc ? 2 - k : k;
    plan.chunk = chunk;
    cached = true;  // no tuning, nothing cached
  }
  auto timed = [&](const IRPlan &p, float &ms) -> int {
    hipEvent_t t0, t1;
    if (hipEventCreate(&t0) != hipSuccess) return TDYHIP_ERR_HIP;
    if (hipEventCreate(&t1) != hipSuccess) { (void)hipEventDestroy(t0); return TDYHIP_ERR_HIP; }
    (void)hipEventRecord(t0, st);
    int r = launch_ir_persistent(ga, p, d_counter, ir_grid(ga, p), st);
    (void)hipEventRecord(t1, st);
    (void)hipEventSynchronize(t1);
    ms = 0.f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return r;
  };
  if (!cached) {
    for (int b = 0; b < 3; b++) {
      if (ga[b].n < kTuneMin) continue;
      float best = 1e30f;
      for (int w = 16 << b; w <= 64; w *= 2) {
        IRPlan one;
        one.lanes[b] = w;
        one.order[0] = b; one.order[1] = -1; one.order[2] = -1;
        float ms;
        if ((rc = timed(one, ms))) return done(rc);
        g->tune_ms[b][w == 16 ? 0 : w == 32 ? 1 : 2] = ms;
        if (ms < best) { best = ms; plan.lanes[b] = w; }
      }
    }
    int nbig = 0;
    for (int b = 0; b < 3; b++) nbig += ga[b].n >= kTuneMin;
    if (nbig > 0) {
      float best = 1e30f;
      IRPlan base = plan;
      for (int desc = 0; desc < 2; desc++)
        for (int chunk : {1, 8}) {
          IRPlan p = base;
          for (int k = 0; k < 3; k++) p.order[k] = desc ? 2 - k : k;
          p.chunk = chunk;
          float ms;
          if ((rc = timed(p, ms))) return done(rc);
          g->sched_ms[desc * 2 + (chunk == 8)] = ms;
          if (ms < best) { best = ms; plan = p; }
        }
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = plan;
  }
  for (int b = 0; b < 3; b++) g->lanes[b] = plan.lanes[b];
  g->ir_desc = plan.order[0] == 2;
  g->ir_chunk = plan.chunk;
  (void)hipEventRecord(ev[1], st);
  if ((rc = launch_ir_persistent(ga, plan, d_counter, ir_grid(ga, plan), st))) return done(rc);
  (void)hipEventRecord(ev[2], st);
  k_gravdis<<<nblk(nf, 256), 256, 0, st>>>(nf, ctx->gmatrix, ctx->gravity, ctx->d_face_cells, ctx->d_face_flags, d_cc,
                                           d_fc, d_fa2, d_fn, d_K, d_fnv, ctx->d_G);
  (void)hipEventRecord(ev[3], st);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
    for (auto &e : ev) (void)hipEventDestroy(e);
    set_error("general setup kernels failed");
    return done(TDYHIP_ERR_HIP);
  }
  for (int i = 0; i < 3; i++) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
    ctx->setup_ms[i] = ms;
  }
  ctx->setup_ms[0] = gm_ms;  // the second event pair brackets the final local-system launches only
  for (auto &e : ev) (void)hipEventDestroy(e);
  int status = 0;
  if (copy_sync(ctx, &status, d_status, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return done(TDYHIP_ERR_HIP);
  if (status) { set_error("singular interaction-region system (zero LU pivot)"); return done(TDYHIP_ERR_SINGULAR); }
  ctx->h_face_verts.assign(m->face_verts, m->face_verts + (size_t)4 * nf);
  ctx->h_face_flags = g->face_flags;
  // residual and Jacobian: per-cell data, state, flux rows, pattern
  const int nbnd = m->nbnd;
  if ((rc = dupload(ctx, &ctx->d_cell_faces, m->cell_faces, (size_t)6 * nc)) ||
      (rc = dupload(ctx, &ctx->d_cell_vol, m->cell_vol, nc)) || (rc = dupload(ctx, &ctx->d_phi, mat->porosity, nc)) ||
      (rc = dupload(ctx, &ctx->d_Sr, mat->Sr, nc)) || (rc = dupload(ctx, &ctx->d_bnd_cell, g->bnd_cell.data(), nbnd)) ||
      (rc = upload_geometry_for_velocity(ctx, m)) ||
      (rc = dupload(ctx, &ctx->d_rtoff, g->rtoff.data(), g->rtoff.size())) ||
      (rc = dupload(ctx, &ctx->d_rinfo, g->rinfo.data(), g->rinfo.size())) ||
      (rc = dupload(ctx, &ctx->d_roc, g->roc.data(), g->roc.size())) ||
      (rc = dupload(ctx, &ctx->d_rdc, g->rdc.data(), g->rdc.size())) ||
      (rc = dupload(ctx, &ctx->d_ccell, g->ccell.data(), g->ccell.size())) ||
      (rc = dupload(ctx, &ctx->d_vdcol, g->vdcol.data(), std::max<size_t>(g->vdcol.size(), 1))) ||
      (rc = dupload(ctx, &ctx->d_jrp, g->jrp.data(), g->jrp.size())) ||
      (rc = dupload(ctx, &ctx->d_jcols, g->jcols.data(), g->jcols.size())) ||
      (rc = dupload(ctx, &ctx->d_jdiag, g->jdiag.data(), g->jdiag.size())) ||
      (rc = dalloc(ctx, &ctx->d_Dgen, (size_t)std::max<int64_t>(g->tsize, 1))) ||
      (rc = dalloc(ctx, &ctx->d_gslot, (size_t)2 * std::max<int64_t>(g->tsize, 1))) ||
      (rc = dalloc(ctx, &ctx->d_flux, (size_t)4 * nf)))
    return done(rc);
  {
    // vertex blocks for the flux-row kernel: T offset, rows, cells, Dirichlet
    // columns, cell offset, Dirichlet-column offset, first block row
    const int nv = m->nv;
    std::vector<int32_t> vb((size_t)nv * 7), br;
    int32_t od = 0;
    bool fits = true;
    for (int v = 0; v < nv && fits; v++) {
      const int32_t *d = g->vdesc.data() + (size_t)v * 8;
      const int nface = d[1], of = d[5], nrow = d[7];
      int32_t *o = vb.data() + (size_t)v * 7;
      o[0] = d[6]; o[1] = nrow; o[2] = d[0]; o[3] = d[3]; o[4] = d[4]; o[5] = od; o[6] = (int32_t)br.size();
      if (nrow > kBlkRows || br.size() + (size_t)nrow > (size_t)INT32_MAX) fits = false;
      br.resize(br.size() + nrow, -1);
      for (int k = 0; k < nface; k++) {
        const int trow = (g->fb[of + k] >> 16) - 1;
        if (trow >= 0 && trow < nrow) br[o[6] + trow] = g->frow[of + k];
      }
      od += d[3];
    }
    if (fits && !br.empty() && std::find(br.begin(), br.end(), -1) == br.end() && od == (int32_t)g->vdcol.size()) {
      if ((rc = dupload(ctx, &ctx->d_vblk, vb.data(), vb.size())) ||
          (rc = dupload(ctx, &ctx->d_brow, br.data(), br.size())))
        return done(rc);
      // rows without a block are never written by the block kernel: their flux stays 0
      k_fill<<<nblk((int64_t)4 * nf, 256), 256, 0, st>>>((int64_t)4 * nf, ctx->d_flux, 0.0);
    }
  }
  k_gen_slot_table<<<nblk((int64_t)4 * nf, 256), 256, 0, st>>>((int64_t)4 * nf, ctx->d_face_cells, ctx->d_face_flags,
                                                                ctx->d_rtoff, ctx->d_rinfo, ctx->d_roc, ctx->d_ccell,
                                                                ctx->d_jrp, ctx->d_jcols, ctx->d_gslot);
  if (mat->p_bnd_fn) {
    ctx->has_pbnd_fn = true;
    if ((rc = dupload(ctx, &ctx->d_pbnd_fn, mat->p_bnd_fn, nbnd))) return done(rc);
  }
  for (double **q : {&ctx->d_rho, &ctx->d_ukvr, &ctx->d_drho, &ctx->d_dukvr, &ctx->d_dM, &ctx->d_d2M, &ctx->d_acc,
                     &ctx->d_accum_prev})
    if ((rc = dalloc(ctx, q, nc))) return done(rc);
  for (double **q : {&ctx->d_Pbnd, &ctx->d_ukvr_bnd, &ctx->d_rho_bnd})
    if ((rc = dalloc(ctx, q, std::max(nbnd, 1)))) return done(rc);
  {
    double rb, d1, d2;  // boundary rho frozen at (Pref, Tref) (tdympfao.c:297-336)
    water_density(ctx->density_type, ctx->pref, rb, d1, d2);
    if (nbnd > 0) k_fill<<<nblk(nbnd, 256), 256, 0, st>>>(nbnd, ctx->d_rho_bnd, rb);
  }
  if (hipStreamSynchronize(st) != hipSuccess) { set_error("general setup failed"); return done(TDYHIP_ERR_HIP); }
  ctx->jnnz = g->jrp[nc];
  ctx->jell_w = 0;
  for (int c = 0; c < nc; c++) ctx->jell_w = std::max(ctx->jell_w, g->jrp[c + 1] - g->jrp[c]);
  if (ctx->jell_w > kGenJSlots) { set_error("more than 64 Jacobian entries in a row"); return done(TDYHIP_ERR_ARG); }
  ctx->h_jrp = g->jrp;
  ctx->h_jcols = g->jcols;
  return done(0);
}

// Residual / Jacobian of a general mesh (TS or SNES form): flux rows from the
// vertex blocks, the residual by k_cell_gather, the Jacobian as CSR values in
// the pattern order.
static int launch_flux_general(tdyhip_ctx *ctx, int form, double shift, const double *U, const double *Udot,
                               double *R, double *J, hipStream_t st, const int32_t *perm, int ell_bs) {
  GenRowArgs a;
  a.nc = ctx->nc; a.nrows = (int64_t)4 * ctx->nf; a.pref = ctx->pref;
  a.face_cells = ctx->d_face_cells; a.face_flags = ctx->d_face_flags;
  a.rtoff = ctx->d_rtoff; a.rinfo = ctx->d_rinfo; a.roc = ctx->d_roc; a.rdc = ctx->d_rdc;
  a.ccell = ctx->d_ccell; a.vdcol = ctx->d_vdcol;
  a.T = ctx->d_Tgen; a.G = ctx->d_G; a.U = U; a.Pbnd = ctx->d_Pbnd; a.rho = ctx->d_rho; a.ukvr = ctx->d_ukvr;
  a.drho = ctx->d_drho; a.dukvr = ctx->d_dukvr; a.rho_bnd = ctx->d_rho_bnd; a.ukvr_bnd = ctx->d_ukvr_bnd;
  a.flux = ctx->launch_flux ? ctx->launch_flux : ctx->d_flux; a.D = ctx->d_Dgen;
  a.stop = ctx->launch_stop;
  if (ctx->d_vblk) {
    GenBlkArgs b;
    b.pref = ctx->pref; b.vblk = ctx->d_vblk; b.brow = ctx->d_brow;
    b.face_cells = ctx->d_face_cells; b.face_flags = ctx->d_face_flags; b.ccell = ctx->d_ccell; b.vdcol = ctx->d_vdcol;
    b.T = ctx->d_Tgen; b.G = ctx->d_G; b.U = U; b.Pbnd = ctx->d_Pbnd; b.rho = ctx->d_rho; b.ukvr = ctx->d_ukvr;
    b.drho = ctx->d_drho; b.dukvr = ctx->d_dukvr; b.rho_bnd = ctx->d_rho_bnd; b.ukvr_bnd = ctx->d_ukvr_bnd;
    b.flux = ctx->launch_flux ? ctx->launch_flux : ctx->d_flux; b.D = ctx->d_Dgen;
    b.stop = ctx->launch_stop;
    const unsigned gv = (unsigned)ctx->nv;
    if (R && J) k_rows_general_blk<true, true><<<gv, 64, 0, st>>>(b);
    else if (R) k_rows_general_blk<true, false><<<gv, 64, 0, st>>>(b);
    else k_rows_general_blk<false, true><<<gv, 64, 0, st>>>(b);
  } else {
    const unsigned gr = nblk(a.nrows, 256);
    if (R && J) k_rows_general<true, true><<<gr, 256, 0, st>>>(a);
    else if (R) k_rows_general<true, false><<<gr, 256, 0, st>>>(a);
    else k_rows_general<false, true><<<gr, 256, 0, st>>>(a);
  }
  if (R) {
    GatherArgs ga = gather_args(ctx, form, shift, Udot, R, nullptr, false, nullptr, 0);
    k_cell_gather<<<nblk(ctx->nc, kGatherBlock), kGatherBlock, 0, st>>>(ga);
  }
  if (J) {
    GenGatherArgs gg;
    gg.nc = ctx->nc; gg.form = form; gg.shift = shift; gg.dtime = ctx->dtime;
    gg.ell_w = ctx->jell_w; gg.perm = perm; gg.jbs = perm ? ell_bs : 1;
    gg.cell_faces = ctx->d_cell_faces; gg.face_cells = ctx->d_face_cells; gg.rtoff = ctx->d_rtoff;
    gg.rinfo = ctx->d_rinfo; gg.roc = ctx->d_roc; gg.ccell = ctx->d_ccell; gg.jrp = ctx->d_jrp;
    gg.jcols = ctx->d_jcols; gg.jdiag = ctx->d_jdiag; gg.face_flags = ctx->d_face_flags; gg.gslot = ctx->d_gslot;
    gg.D = ctx->d_Dgen; gg.dM = ctx->d_dM; gg.d2M = ctx->d_d2M; gg.Udot = Udot; gg.vol = ctx->d_cell_vol; gg.J = J;
    gg.dyn = ctx->launch_dyn; gg.stop = ctx->launch_stop;
    k_gather_general_J<<<nblk(ctx->nc, kGenJBlock / kGenJLanes), kGenJBlock, 0, st>>>(gg);  // kGenJLanes lanes per cell
  }
  return hipGetLastError() == hipSuccess ? 0 : TDYHIP_ERR_HIP;
}

// Trans as CSR (rows face*4+subface, columns sorted: cells, then nc + boundary
// index) from the per-vertex blocks of a general mesh.
static int trans_csr_general(const tdyhip_ctx *ctx, int32_t *rowptr, int32_t *cols, double *vals) {
  const HostGeneral &g = *ctx->hgen;
  const int64_t nrows = (int64_t)4 * ctx->nf;
  std::vector<double> T((size_t)std::max<int64_t>(g.tsize, 1));
  if (hipStreamSynchronize(ctx->stream) != hipSuccess ||
      copy_sync(const_cast<tdyhip_ctx *>(ctx), T.data(), ctx->d_Tgen, (size_t)g.tsize * sizeof(double),
                hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("copy of the transmissibility blocks failed");
    return TDYHIP_ERR_HIP;
  }
  std::vector<int64_t> cnt(nrows + 1, 0);
  for (int v = 0; v < ctx->nv; v++) {
    const int32_t *d = g.vdesc.data() + (size_t)v * 8;
    for (int k = 0; k < d[1]; k++)
      if ((g.fb[d[5] + k] >> 16) > 0) cnt[g.frow[d[5] + k] + 1] += d[0] + d[3];
  }
  for (int64_t r = 0; r < nrows; r++) cnt[r + 1] += cnt[r];
  if (cnt[nrows] > INT32_MAX) { set_error("transmissibility has more than 2^31 entries"); return TDYHIP_ERR_ARG; }
  for (int64_t r = 0; r <= nrows; r++) rowptr[r] = (int32_t)cnt[r];
  std::vector<int32_t> dcol;
  for (int v = 0; v < ctx->nv; v++) {
    const int32_t *d = g.vdesc.data() + (size_t)v * 8;
    const int ncell = d[0], nface = d[1], ndir = d[3], oc = d[4], of = d[5], ot = d[6], nB = ncell + ndir;
    dcol.assign(ndir, -1);
    for (int k = 0; k < nface; k++)
      if ((g.fa[of + k] & 3) == 2) dcol[(g.fa[of + k] >> 8) & 0xff] = g.fcol[of + k];
    for (int k = 0; k < nface; k++) {
      const int trow = (g.fb[of + k] >> 16) - 1;
      if (trow < 0) continue;
      const int64_t base = cnt[g.frow[of + k]];
      int64_t n = 0;
      for (int j = 0; j < nB; j++) {  // insertion sort by column
        const int col = j < ncell ? g.ccell[oc + j] : dcol[j - ncell];
        const double val = T[(size_t)ot + (size_t)trow * nB + j];
        int64_t p = n;
        while (p > 0 && cols[base + p - 1] > col) { cols[base + p] = cols[base + p - 1]; vals[base + p] = vals[base + p - 1]; p--; }
        cols[base + p] = col;
        vals[base + p] = val;
        n++;
      }
    }
  }
  return 0;
}

}  // namespace tdyhip

namespace tdyhip {

// ---- the allocation-free residual (tdyhip_ifunction_noalloc) ----
// Its flux-row scratch lives in this file-scope device array, one copy per
// GPU for the life of the process.  Host code never uses the array's name as
// a device address: each device's address comes from hipGetSymbolAddress,
// taken once per device at set-up (tdyhip_create) with that device current,
// and is passed to the kernels like ctx->d_flux.
constexpr size_t kNoAllocRows = (size_t)8 << 20;  // flux rows (4 per face): 64 MiB of FP64
__device__ __align__(256) double scratch[kNoAllocRows];
static std::mutex g_scratch_mu;
static std::map<int, double *> g_scratch_ptr;  // device -> address of its `scratch`

// The current device's scratch address (looked up and kept the first time).
static int scratch_address(double **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) { *out = it->second; return 0; }
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    return TDYHIP_ERR_HIP;
  }
  g_scratch_ptr[dev] = static_cast<double *>(p);
  *out = static_cast<double *>(p);
  return 0;
}

}  // namespace tdyhip

extern "C" {

const char *tdyhip_last_error(void) { return g_err.c_str(); }

int tdyhip_create(const tdyhip_mesh *m, const tdyhip_materials *mat, tdyhip_ctx **out) {
  if (!m || !mat || !out) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  *out = nullptr;
  if (m->nc <= 0 || m->nf <= 0 || m->nv <= 0 || m->nbnd < 0) { set_error("empty mesh"); return TDYHIP_ERR_ARG; }
  if ((int64_t)m->nf * 4 * kMaxCols > ((int64_t)1 << 40)) { set_error("mesh too large"); return TDYHIP_ERR_ARG; }
  // Hexahedral meshes take the fixed-size kernels (at most 8 cells, 12 faces and
  // 8 transmissibility columns at a vertex, 27 Jacobian entries in a row);
  // other cells (prisms), and hexahedral meshes beyond those limits, take the
  // general path (regions of up to 64 unknowns, rows of up to 64 entries).
  bool hex = mesh_is_hex(*m);
  HostTopology topo;
  if (hex) {
    std::string msg = build_topology(*m, topo);
    if (!msg.empty() && !topo.over_limit) { set_error(msg); return TDYHIP_ERR_ARG; }
    if (!msg.empty()) {
      hex = false;
    } else {
      for (int c = 0; c < m->nc && hex; c++) hex = topo.jrp[c + 1] - topo.jrp[c] <= kJacSlots;
    }
    if (!hex) topo = HostTopology();
  }

  tdyhip_ctx *ctx = new tdyhip_ctx();
  ctx->nc = m->nc; ctx->nf = m->nf; ctx->nv = m->nv; ctx->nbnd = m->nbnd;
  ctx->vg_m = mat->vg_m; ctx->vg_alpha = mat->vg_alpha; ctx->x0 = mat->poly_x0; ctx->x1 = mat->poly_x1;
  ctx->x2 = mat->poly_x2; ctx->dx = mat->poly_dx; ctx->gravity = mat->gravity; ctx->pref = mat->pref;
  ctx->density_type = mat->density_type; ctx->gmatrix = mat->gmatrix_method;
  if (const char *e = getenv("TDYHIP_CELLJAC_XCD")) ctx->cj_xcd = atoi(e) != 0;
  if (mualem_cubic(ctx->vg_m, ctx->x0, ctx->x1, ctx->x2, ctx->dx, ctx->mualem_c)) {
    set_error("Mualem smoothing polynomial is singular");
    delete ctx;
    return TDYHIP_ERR_ARG;
  }
  int rc = 0;
  auto fail = [&](int code) {
    tdyhip_destroy(ctx);
    return code;
  };
  hipError_t he = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (he != hipSuccess) { set_error(std::string("hipStreamCreate: ") + hipGetErrorString(he)); delete ctx; return TDYHIP_ERR_HIP; }
  {
    double *sc = nullptr;  // this device's scratch address, for tdyhip_ifunction_noalloc
    if ((rc = scratch_address(&sc))) { (void)hipStreamDestroy(ctx->stream); delete ctx; return rc; }
  }
  ctx->own_stream = true;
  if (!hex) {
    if ((rc = setup_general(ctx, m, mat))) return fail(rc);
    *out = ctx;
    return 0;
  }
  const int nc = m->nc, nf = m->nf, nv = m->nv, nbnd = m->nbnd;
  const int64_t nrows = (int64_t)4 * nf;
  ctx->jnnz = topo.jrp[nc];
  for (int f = 0; f < nf; f++) {
    if (topo.face_flags[f] & kFaceNoFlow) continue;
    ctx->active_rows += 4;
    for (int q = 0; q < 4; q++) {
      const int v = m->face_verts[f * 4 + q];
      for (int k = 0; k < kMaxCols; k++) ctx->trans_nnz += topo.vcols[(size_t)v * kMaxCols + k] >= 0;
    }
  }

  if ((rc = dupload(ctx, &ctx->d_face_cells, m->face_cells, (size_t)2 * nf))) return fail(rc);
  if ((rc = dupload(ctx, &ctx->d_face_verts, m->face_verts, (size_t)4 * nf))) return fail(rc);
  if ((rc = dupload(ctx, &ctx->d_cell_faces, m->cell_faces, (size_t)6 * nc))) return fail(rc);
  if ((rc = dupload(ctx, &ctx->d_face_flags, topo.face_flags.data(), nf))) return fail(rc);
  if ((rc = dupload(ctx, &ctx->d_cell_vol, m->cell_vol, nc))) return fail(rc);
  if ((rc = dupload(ctx, &ctx->d_phi, mat->porosity, nc))) return fail(rc);
  if ((rc = dupload(ctx, &ctx->d_Sr, mat->Sr, nc))) return fail(rc);
  if (mat->p_bnd_fn) {
    ctx->has_pbnd_fn = true;
    if ((rc = dupload(ctx, &ctx->d_pbnd_fn, mat->p_bnd_fn, nbnd))) return fail(rc);
  }
  if ((rc = dupload(ctx, &ctx->d_bnd_cell, topo.bnd_cell.data(), nbnd))) return fail(rc);
  if ((rc = upload_geometry_for_velocity(ctx, m))) return fail(rc);
  if ((rc = dupload(ctx, &ctx->d_vcols, topo.vcols.data(), topo.vcols.size()))) return fail(rc);
  if ((rc = dupload(ctx, &ctx->d_jslot, topo.jslot.data(), topo.jslot.size()))) return fail(rc);
  {
    // each corner's three flux rows (one per face of the cell at the corner),
    // the cell's index in the corner vertex's region and its side of each row
    std::vector<uint32_t> crow((size_t)nc * 8, 0);
    for (int c = 0; c < nc; c++) {
      int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int lf = 0; lf < 6; lf++) {
        const int f = m->cell_faces[(size_t)c * 6 + lf];
        for (int q = 0; q < 4; q++) {
          const int v = m->face_verts[(size_t)f * 4 + q];
          int j = 0;
          while (j < 8 && m->cell_verts[(size_t)c * 8 + j] != v) j++;
          if (j == 8 || cnt[j] == 3) { set_error("hexahedron with a corner not on exactly three of its faces"); return fail(TDYHIP_ERR_ARG); }
          uint32_t &w = crow[(size_t)c * 8 + j];
          w |= (uint32_t)(lf * 4 + q) << (5 * cnt[j]);
          if (m->face_cells[(size_t)2 * f] == c) w |= 1u << (18 + cnt[j]);
          cnt[j]++;
        }
      }
      for (int j = 0; j < 8; j++) {
        const int v = m->cell_verts[(size_t)c * 8 + j];
        int li = 0;
        while (li < 8 && topo.vcols[(size_t)v * kMaxCols + li] != c) li++;
        if (li == 8) { set_error("cell missing from its vertex's region"); return fail(TDYHIP_ERR_ARG); }
        crow[(size_t)c * 8 + j] |= (uint32_t)li << 15;
      }
    }
    if ((rc = dupload(ctx, &ctx->d_crow, crow.data(), crow.size()))) return fail(rc);
    if ((rc = dupload(ctx, &ctx->d_cell_verts, m->cell_verts, (size_t)nc * 8))) return fail(rc);
  }
  if ((rc = dupload(ctx, &ctx->d_jrp, topo.jrp.data(), topo.jrp.size()))) return fail(rc);
  if ((rc = dupload(ctx, &ctx->d_jcols, topo.jcols.data(), topo.jcols.size()))) return fail(rc);
  {
    std::vector<uint8_t> dslot(nc);
    int W = 0;
    for (int c = 0; c < nc; c++) {
      dslot[c] = (uint8_t)(topo.jdiag[c] - topo.jrp[c]);
      W = std::max(W, topo.jrp[c + 1] - topo.jrp[c]);
    }
    if (W > kJacSlots) { set_error("more than 27 Jacobian entries in a row"); return fail(TDYHIP_ERR_ARG); }
    ctx->jell_w = W;
    std::vector<int32_t> ecols((size_t)W * nc);
    for (int c = 0; c < nc; c++)
      for (int k = 0; k < W; k++) {
        const int p = topo.jrp[c] + k;
        ecols[(size_t)k * nc + c] = p < topo.jrp[c + 1] ? topo.jcols[p] : c;
      }
    if ((rc = dupload(ctx, &ctx->d_jdslot, dslot.data(), dslot.size()))) return fail(rc);
    if ((rc = dupload(ctx, &ctx->d_ecols, ecols.data(), ecols.size()))) return fail(rc);
  }
  if ((rc = dalloc(ctx, &ctx->d_T, (size_t)kMaxCols * nrows))) return fail(rc);
  if ((rc = dalloc(ctx, &ctx->d_G, (size_t)nrows))) return fail(rc);
  for (double **p : {&ctx->d_rho, &ctx->d_ukvr, &ctx->d_drho, &ctx->d_dukvr, &ctx->d_dM, &ctx->d_d2M, &ctx->d_acc,
                     &ctx->d_accum_prev})
    if ((rc = dalloc(ctx, p, nc))) return fail(rc);
  for (double **p : {&ctx->d_Pbnd, &ctx->d_ukvr_bnd, &ctx->d_rho_bnd})
    if ((rc = dalloc(ctx, p, nbnd))) return fail(rc);
  if ((rc = dalloc(ctx, &ctx->d_flux, (size_t)nrows))) return fail(rc);

  // setup temporaries
  int32_t *d_scf = nullptr; IRDesc *d_ir = nullptr; double *d_Gsub = nullptr, *d_K = nullptr, *d_cc = nullptr,
          *d_fc = nullptr, *d_fa = nullptr, *d_fn = nullptr; int *d_status = nullptr;
  auto tmp_fail = [&](int code) {
    for (void *p : {(void *)d_scf, (void *)d_ir, (void *)d_Gsub, (void *)d_K, (void *)d_cc, (void *)d_fc,
                    (void *)d_fa, (void *)d_fn, (void *)d_status})
      if (p) (void)hipFree(p);
    return fail(code);
  };
#define TMPALLOC(ptr, n, src)                                                                       \
  do {                                                                                              \
    hipError_t e_ = hipMalloc((void **)&ptr, (size_t)(n) * sizeof(*ptr));                           \
    if (e_ != hipSuccess) { set_error(std::string("hipMalloc: ") + hipGetErrorString(e_));          \
      return tmp_fail(e_ == hipErrorOutOfMemory ? TDYHIP_ERR_MEM : TDYHIP_ERR_HIP); }               \
    if (src) { e_ = copy_sync(ctx, ptr, src, (size_t)(n) * sizeof(*ptr), hipMemcpyHostToDevice);   \
      if (e_ != hipSuccess) { set_error(std::string("hipMemcpy: ") + hipGetErrorString(e_)); return tmp_fail(TDYHIP_ERR_HIP); } } \
  } while (0)
  TMPALLOC(d_scf, topo.sc_faces.size(), topo.sc_faces.data());
  TMPALLOC(d_ir, topo.ir.size(), topo.ir.data());
  TMPALLOC(d_Gsub, (size_t)nc * 72, (double *)nullptr);
  TMPALLOC(d_K, (size_t)nc * 9, mat->K);
  TMPALLOC(d_cc, (size_t)nc * 3, m->cell_cen);
  TMPALLOC(d_fc, (size_t)nf * 3, m->face_cen);
  TMPALLOC(d_fa, (size_t)nf, m->face_area);
  TMPALLOC(d_fn, (size_t)nf * 3, m->face_normal);
  TMPALLOC(d_status, 1, (int *)nullptr);
#undef TMPALLOC
  hipStream_t st = ctx->stream;
  k_fill<<<nblk((int64_t)kMaxCols * nrows, 256), 256, 0, st>>>((int64_t)kMaxCols * nrows, ctx->d_T, 0.0);
  if (hipMemsetAsync(d_status, 0, sizeof(int), st) != hipSuccess) return tmp_fail(TDYHIP_ERR_HIP);
  // device times of the three setup phases (G matrices, local systems, gravity term)
  hipEvent_t sev[6] = {};
  for (auto &e : sev)
    if (hipEventCreate(&e) != hipSuccess) e = nullptr;
  auto rec = [&](int i) { if (sev[i]) (void)hipEventRecord(sev[i], st); };
  rec(0);
  k_gmatrix<<<nblk((int64_t)nc * 8, 256), 256, 0, st>>>(nc, ctx->gmatrix, d_scf, ctx->d_face_cells, d_cc, d_fc, d_fa,
                                                        d_fn, d_K, nullptr, d_Gsub);
  rec(1);
  rec(2);
  k_local_systems<<<nblk(nv, kTeamsPerBlock), kTeam * kTeamsPerBlock, 0, st>>>(nv, d_ir, d_Gsub, ctx->d_T, nrows,
                                                                               d_status);
  rec(3);
  rec(4);
  k_gravdis<<<nblk(nf, 256), 256, 0, st>>>(nf, ctx->gmatrix, ctx->gravity, ctx->d_face_cells, ctx->d_face_flags, d_cc,
                                           d_fc, d_fa, d_fn, d_K, nullptr, ctx->d_G);
  rec(5);
  // boundary rho, mu frozen at (Pref, Tref) (AllocateMemoryForBoundaryValues, tdympfao.c:297-336)
  {
    double rb, d1, d2;
    water_density(ctx->density_type, ctx->pref, rb, d1, d2);
    if (nbnd > 0) k_fill<<<nblk(nbnd, 256), 256, 0, st>>>(nbnd, ctx->d_rho_bnd, rb);
  }
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
    set_error("setup kernels failed");
    for (auto &e : sev)
      if (e) (void)hipEventDestroy(e);
    return tmp_fail(TDYHIP_ERR_HIP);
  }
  for (int i = 0; i < 3; i++) {
    float ms = 0.f;
    if (sev[2 * i] && sev[2 * i + 1] && hipEventElapsedTime(&ms, sev[2 * i], sev[2 * i + 1]) == hipSuccess)
      ctx->setup_ms[i] = ms;
  }
  for (auto &e : sev)
    if (e) (void)hipEventDestroy(e);
  // algorithmic work of the local systems (the same model as the general path's)
  for (int v = 0; v < nv; v++) {
    const IRDesc &d = topo.ir[v];
    const double nU = d.nint + d.nneu, nB = d.ncell + d.ndir, nrow = d.nint + d.ndir;
    ctx->ls_flops += 2.0 / 3.0 * nU * nU * nU + 3.0 * nU * nU * nB + 6.0 * nrow * nB;
    ctx->ls_bytes += (double)sizeof(IRDesc) + d.ncell * (3 * 9 * 8.0) + nrow * nB * 8.0;
    ctx->ir_max_cells = std::max(ctx->ir_max_cells, (int)d.ncell);
    ctx->ir_max_faces = std::max(ctx->ir_max_faces, (int)d.nface);
  }
  int status = 0;
  if (copy_sync(ctx, &status, d_status, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return tmp_fail(TDYHIP_ERR_HIP);
  for (void *p : {(void *)d_scf, (void *)d_ir, (void *)d_Gsub, (void *)d_K, (void *)d_cc, (void *)d_fc, (void *)d_fa,
                  (void *)d_fn, (void *)d_status})
    (void)hipFree(p);
  if (status) {
    set_error("singular interaction-region system (zero LU pivot)");
    tdyhip_destroy(ctx);
    return TDYHIP_ERR_SINGULAR;
  }
  ctx->h_jrp = std::move(topo.jrp);
  ctx->h_jcols = std::move(topo.jcols);
  ctx->h_vcols = std::move(topo.vcols);
  ctx->h_face_verts.assign(m->face_verts, m->face_verts + (size_t)4 * nf);
  ctx->h_face_flags = std::move(topo.face_flags);
  *out = ctx;
  return 0;
}

int tdyhip_destroy(tdyhip_ctx *ctx) {
  if (!ctx) return 0;
  if (ctx->fuser) (void)tdyhip_set_fuser(ctx, nullptr);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  tdyhip::ws_destroy(ctx->ws);
  ctx->ws = nullptr;
  if (ctx->ws_event) (void)hipEventDestroy(ctx->ws_event);
  for (void *p : ctx->allocations) (void)hipFree(p);
  if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx->hgen;
  delete ctx;
  return 0;
}

int tdyhip_set_stream(tdyhip_ctx *ctx, tdyhip_stream stream) {
  if (!ctx) { set_error("null context"); return TDYHIP_ERR_ARG; }
  if (!stream) { set_error("null stream"); return TDYHIP_ERR_ARG; }
  if (ctx->own_stream) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    ctx->own_stream = false;
  }
  ctx->stream = (hipStream_t)stream;
  return 0;
}

int tdyhip_get_stream(const tdyhip_ctx *ctx, tdyhip_stream *stream) {
  if (!ctx || !stream) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  *stream = (tdyhip_stream)ctx->stream;
  return 0;
}

int tdyhip_sizes(const tdyhip_ctx *ctx, int64_t *out) {
  if (!ctx || !out) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  out[0] = ctx->nc; out[1] = ctx->jnnz; out[2] = ctx->nf; out[3] = ctx->nbnd;
  return 0;
}

int tdyhip_jacobian_pattern(const tdyhip_ctx *ctx, int32_t *rowptr, int32_t *cols) {
  if (!ctx || !rowptr || !cols) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  std::memcpy(rowptr, ctx->h_jrp.data(), sizeof(int32_t) * (ctx->nc + 1));
  std::memcpy(cols, ctx->h_jcols.data(), sizeof(int32_t) * ctx->jnnz);
  return 0;
}

int tdyhip_jacobian_pattern_device(const tdyhip_ctx *ctx, const int32_t **rowptr, const int32_t **cols) {
  if (!ctx || !rowptr || !cols) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  *rowptr = ctx->d_jrp;
  *cols = ctx->d_jcols;
  return 0;
}

int tdyhip_ifunction(tdyhip_ctx *ctx, double t, const double *U, const double *Udot, double *R, tdyhip_stream s) {
  (void)t;
  if (!ctx || !U || !Udot || !R) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  hipStream_t st; check_stream(ctx, s, &st);
  int rc = launch_state(ctx, 0, U, Udot, st);
  if (rc) return rc;
  return launch_flux(ctx, 0, 0.0, U, Udot, R, nullptr, st);
}

int tdyhip_ifunction_noalloc(tdyhip_ctx *ctx, double t, const double *U, const double *Udot, double *R,
                             tdyhip_stream s) {
  (void)t;
  if (!ctx || !U || !Udot || !R) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  if ((size_t)4 * ctx->nf > kNoAllocRows) {
    set_error("tdyhip_ifunction_noalloc: the mesh has " + std::to_string((long long)4 * ctx->nf) +
              " flux rows, more than the " + std::to_string((long long)kNoAllocRows) +
              " its fixed scratch holds (use tdyhip_ifunction)");
    return TDYHIP_ERR_ARG;
  }
  double *sc = nullptr;
  int rc = scratch_address(&sc);
  if (rc) return rc;
  hipStream_t st; check_stream(ctx, s, &st);
  const bool timers = ctx->kt.on;  // timer events are created on demand: none here
  ctx->kt.on = false;
  // rows the general path's flux kernels do not write (a triangle's fourth
  // sub-face) must read zero; the array keeps whatever the last call left
  if (ctx->general) k_fill<<<nblk((int64_t)4 * ctx->nf, 256), 256, 0, st>>>((int64_t)4 * ctx->nf, sc, 0.0);
  rc = launch_state(ctx, 0, U, Udot, st);
  if (!rc) {
    ctx->launch_flux = sc;
    rc = launch_flux(ctx, 0, 0.0, U, Udot, R, nullptr, st);
    ctx->launch_flux = nullptr;
  }
  ctx->kt.on = timers;
  return rc;
}

int tdyhip_ijacobian(tdyhip_ctx *ctx, double t, const double *U, const double *Udot, double shift, double *J,
                     tdyhip_stream s) {
  (void)t;
  if (!ctx || !U || !Udot || !J) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  hipStream_t st; check_stream(ctx, s, &st);
  return launch_flux(ctx, 0, shift, U, Udot, nullptr, J, st);
}

int tdyhip_ifunction_ijacobian(tdyhip_ctx *ctx, double t, const double *U, const double *Udot, double shift,
                               double *R, double *J, tdyhip_stream s) {
  (void)t;
  if (!ctx || !U || !Udot || !R || !J) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  hipStream_t st; check_stream(ctx, s, &st);
  int rc = launch_state(ctx, 0, U, Udot, st);
  if (rc) return rc;
  return launch_flux(ctx, 0, shift, U, Udot, R, J, st);
}

int tdyhip_set_source_sink(tdyhip_ctx *ctx, const double *Q) {
  if (!ctx) { set_error("null context"); return TDYHIP_ERR_ARG; }
  if (!Q) { ctx->has_source = false; return 0; }
  for (int c = 0; c < ctx->nc; c++)
    if (!std::isfinite(Q[c])) { set_error("source_sink is not finite"); return TDYHIP_ERR_ARG; }
  int rc;
  if (!ctx->d_source && (rc = dalloc(ctx, &ctx->d_source, (size_t)ctx->nc))) return rc;
  HIPCHK(hipStreamSynchronize(ctx->stream));  // earlier launches may still read the old values
  HIPCHK(copy_sync(ctx, ctx->d_source, Q, sizeof(double) * ctx->nc, hipMemcpyHostToDevice));
  ctx->has_source = true;
  return 0;
}

int tdyhip_set_dtime(tdyhip_ctx *ctx, double dt) {
  if (!ctx || !(dt > 0.0)) { set_error("bad time step"); return TDYHIP_ERR_ARG; }
  ctx->dtime = dt;
  return 0;
}

int tdyhip_snes_presolve(tdyhip_ctx *ctx, const double *Uprev, tdyhip_stream s) {
  if (!ctx || !Uprev) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  hipStream_t st; check_stream(ctx, s, &st);
  StateParams sp = state_params(ctx, 1, 0.0);
  k_snes_presolve<<<nblk(ctx->nc, 256), 256, 0, st>>>(ctx->nc, sp, Uprev, ctx->d_phi, ctx->d_Sr, ctx->d_cell_vol,
                                                      ctx->d_accum_prev);
  HIPCHK(hipGetLastError());
  return 0;
}

int tdyhip_snes_function(tdyhip_ctx *ctx, const double *U, double *R, tdyhip_stream s) {
  if (!ctx || !U || !R) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  hipStream_t st; check_stream(ctx, s, &st);
  int rc = launch_state(ctx, 1, U, nullptr, st);
  if (rc) return rc;
  return launch_flux(ctx, 1, 0.0, U, nullptr, R, nullptr, st);
}

int tdyhip_diagnostics(tdyhip_ctx *ctx, const double *U, double *S, double *liquid_mass, double *water_mass,
                       tdyhip_stream s) {
  if (!ctx || !U) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  hipStream_t st; check_stream(ctx, s, &st);
  StateParams sp = state_params(ctx, 1, 0.0);
  k_diagnostics<<<nblk(ctx->nc, 256), 256, 0, st>>>(ctx->nc, sp, U, ctx->d_phi, ctx->d_Sr, ctx->d_cell_vol, S,
                                                    liquid_mass, water_mass);
  HIPCHK(hipGetLastError());
  return 0;
}

int tdyhip_recover_velocity(tdyhip_ctx *ctx, const double *U, double *vel, int32_t *count, tdyhip_stream s) {
  if (!ctx || !U || !vel) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  hipStream_t st; check_stream(ctx, s, &st);
  VelArgs a{};
  a.nc = ctx->nc; a.nf = ctx->nf; a.general = ctx->general ? 1 : 0; a.has_fn = ctx->has_pbnd_fn ? 1 : 0;
  a.density_type = ctx->density_type; a.nrows = (int64_t)4 * ctx->nf; a.gabs = fabs(ctx->gravity);
  a.face_flags = ctx->d_face_flags;
  a.face_verts = ctx->d_face_verts; a.vcols = ctx->d_vcols; a.T = ctx->d_T;
  a.rtoff = ctx->d_rtoff; a.rinfo = ctx->d_rinfo; a.roc = ctx->d_roc; a.rdc = ctx->d_rdc;
  a.ccell = ctx->d_ccell; a.vdcol = ctx->d_vdcol; a.Tgen = ctx->d_Tgen;
  a.bnd_cell = ctx->d_bnd_cell; a.U = U; a.cell_z = ctx->d_cell_z; a.pbnd_fn = ctx->d_pbnd_fn;
  a.face_area = ctx->d_face_area; a.vel = vel; a.count = count;
  k_velocity<<<nblk(ctx->nf, 256), 256, 0, st>>>(a);
  HIPCHK(hipGetLastError());
  return 0;
}

int tdyhip_snes_jacobian(tdyhip_ctx *ctx, const double *U, double *J, tdyhip_stream s) {
  if (!ctx || !U || !J) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  hipStream_t st; check_stream(ctx, s, &st);
  return launch_flux(ctx, 1, 0.0, U, nullptr, nullptr, J, st);
}

int tdyhip_trans_nnz(const tdyhip_ctx *ctx, int64_t *nnz) {
  if (!ctx || !nnz) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  if (ctx->general) {
    *nnz = ctx->hgen->tsize;
    return 0;
  }
  int64_t n = 0;
  for (int f = 0; f < ctx->nf; f++) {
    if (ctx->h_face_flags[f] & kFaceNoFlow) continue;
    for (int q = 0; q < 4; q++) {
      const int v = ctx->h_face_verts[f * 4 + q];
      for (int k = 0; k < kMaxCols; k++) n += ctx->h_vcols[(size_t)v * kMaxCols + k] >= 0;
    }
  }
  *nnz = n;
  return 0;
}

int tdyhip_trans_csr(const tdyhip_ctx *ctx, int32_t *rowptr, int32_t *cols, double *vals) {
  if (!ctx || !rowptr || !cols || !vals) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  if (ctx->general) return trans_csr_general(ctx, rowptr, cols, vals);
  const int64_t nrows = (int64_t)4 * ctx->nf;
  std::vector<double> T((size_t)kMaxCols * nrows);
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(copy_sync(ctx, T.data(), ctx->d_T, T.size() * sizeof(double), hipMemcpyDeviceToHost));
  int64_t n = 0;
  rowptr[0] = 0;
  for (int64_t r = 0; r < nrows; r++) {
    const int f = (int)(r >> 2);
    if (!(ctx->h_face_flags[f] & kFaceNoFlow)) {
      const int v = ctx->h_face_verts[r];
      // insertion-sort by column
      int64_t base = n;
      for (int k = 0; k < kMaxCols; k++) {
        const int col = ctx->h_vcols[(size_t)v * kMaxCols + k];
        if (col < 0) continue;
        const double val = T[(size_t)k * nrows + r];
        int64_t p = n;
        while (p > base && cols[p - 1] > col) { cols[p] = cols[p - 1]; vals[p] = vals[p - 1]; p--; }
        cols[p] = col; vals[p] = val; n++;
      }
    }
    rowptr[r + 1] = (int32_t)n;
  }
  return 0;
}

int tdyhip_setup_info(const tdyhip_ctx *ctx, double *ms, int64_t *counts) {
  if (!ctx || !ms || !counts) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  for (int i = 0; i < 3; i++) ms[i] = ctx->setup_ms[i];
  for (int i = 0; i < 11; i++) counts[i] = 0;
  if (ctx->general) {
    const HostGeneral &g = *ctx->hgen;
    for (int b = 0; b < 3; b++) counts[b] = (int64_t)g.bucket[b].size();
    counts[3] = g.max_cells;
    counts[4] = g.max_faces;
    counts[5] = g.tsize;
    // algorithmic work of the local systems: elimination on [A | B]
    // (2/3 nU^3 + 2 nU^2 nB), back substitution (nU^2 nB), T = C X (6 nrow nB);
    // bytes: descriptors, the region's G rows and its T block
    double fl = 0.0, by = 0.0;
    for (int v = 0; v < ctx->nv; v++) {
      const int32_t *d = g.vdesc.data() + (size_t)v * 8;
      const double nU = d[2], nB = d[0] + d[3], nrow = d[7];
      fl += 2.0 / 3.0 * nU * nU * nU + 3.0 * nU * nU * nB + 6.0 * nrow * nB;
      by += 32.0 + d[0] * (8.0 + 3 * 9 * 8.0) + d[1] * 8.0 + nrow * nB * 8.0;
    }
    counts[6] = (int64_t)fl;
    counts[7] = (int64_t)by;
    for (int b = 0; b < 3; b++) counts[8 + b] = g.lanes[b];
  } else {
    // hexahedral meshes: one 16-lane team per region (k_local_systems)
    counts[0] = ctx->nv;
    counts[3] = ctx->ir_max_cells;
    counts[4] = ctx->ir_max_faces;
    counts[5] = ctx->trans_nnz;
    counts[6] = (int64_t)ctx->ls_flops;
    counts[7] = (int64_t)ctx->ls_bytes;
    counts[8] = kTeam;
  }
  return 0;
}

int tdyhip_setup_plan(const tdyhip_ctx *ctx, int32_t *plan, double *ms) {
  if (!ctx || !plan || !ms) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  for (int i = 0; i < 13; i++) ms[i] = 0.0;
  if (ctx->general) {
    const HostGeneral &g = *ctx->hgen;
    for (int b = 0; b < 3; b++) plan[b] = g.lanes[b];
    plan[3] = g.ir_desc ? 1 : 0;
    plan[4] = g.ir_chunk;
    for (int b = 0; b < 3; b++)
      for (int w = 0; w < 3; w++) ms[b * 3 + w] = g.tune_ms[b][w];
    for (int i = 0; i < 4; i++) ms[9 + i] = g.sched_ms[i];
  } else {
    plan[0] = kTeam; plan[1] = plan[2] = plan[3] = 0; plan[4] = 1;
  }
  return 0;
}

int tdyhip_gravdis(const tdyhip_ctx *ctx, double *out) {
  if (!ctx || !out) { set_error("null argument"); return TDYHIP_ERR_ARG; }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(copy_sync(ctx, out, ctx->d_G, sizeof(double) * 4 * ctx->nf, hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"

#include "solver.hip"
#include "batch.hip"
#include "fuser.hip"

namespace tdyhip {
static void ws_destroy(SolverWS *ws) {
  if (!ws) return;
  ws_free(*ws);
  delete ws;
}
}  // namespace tdyhip
