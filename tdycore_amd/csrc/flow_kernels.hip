// Residual and Jacobian of the MPFA-O Richards path (subsystems (b), (c)).
//
//  k_cell_state   per cell: van Genuchten / Mualem / EOS state, rho, Kr/mu and their
//                 pressure derivatives, and the accumulation term (TS or SNES form)
//                 (TDyUpdateState_Richards_MPFAO, src/fv/mpfao/tdympfao.c:2527-2594;
//                  tdympfao_ts.c:191-212; tdympfao_snes.c:19-30, 112-123)
//  k_bnd_state    per boundary face: P_bnd and Kr/mu on the boundary
//                 (tdympfao_utils.c:24-99, 690-759)
//  k_rows         per flux row (face*4 + subface): T.P from the vertex's columns,
//                 upwinded rho^2 kr/mu flux (TDyMPFAOIFunction_Vertices, tdympfao_ts.c:12-150)
//  k_cell_gather  per cell: R = sum over its 6 faces x 4 subfaces (+/- by orientation)
//                 + accumulation
//  k_cell_jac     per cell, one pass: the Jacobian row from its 24 flux rows'
//                 derivatives (TDyMPFAOIJacobian_Vertices, 235-456) + the
//                 accumulation diagonal, optionally the residual as well
//
// Flux rows are stored face-major (row = face*4 + subface), the reference's
// Trans_mat row numbering; T is an 8-wide ELL block stored column-major so that
// consecutive lanes read consecutive addresses.
#include "tdyhip_internal.h"
#include "tdyhip_math.h"

namespace tdyhip {

// Launch-time inputs that a replayed hipGraph must not freeze: when `dyn` is
// set, dtime = dyn[0] and shift = dyn[1] are read from device memory (the
// solver updates them between replays); when `stop` is set and nonzero the
// launch returns at once (the Newton solver's device-side gating).
struct StateParams {
  double vg_m, vg_alpha, x0, x1, dx, pref, dtime, shift;
  double c0, c1, c2, c3;
  int density_type, form;  // form: 0 TS, 1 SNES
  const double *dyn;
  const int *stop;
};

struct CellStateArgs {
  int nc;
  StateParams sp;
  const double *U, *Udot, *phi, *Sr, *vol, *accum_prev;
  double *rho, *ukvr, *drho, *dukvr, *dM, *d2M, *acc;
};

__device__ inline void cell_state(const CellStateArgs &a, int c) {
  const StateParams &sp = a.sp;
  const double p = a.U[c];
  const double sr = a.Sr[c];
  double S, dS, d2S;
  vg_saturation(sp.vg_m, sp.vg_alpha, sr, sp.pref - p, S, dS, d2S);
  const double Se = (S - sr) / (1.0 - sr);
  double Kr, dKr;
  const double cc[4] = {sp.c0, sp.c1, sp.c2, sp.c3};
  mualem_kr(sp.vg_m, sp.x0, sp.x1, sp.dx, cc, Se, Kr, dKr);
  dKr *= 1.0 / (1.0 - sr);
  double rho, drho, d2rho;
  water_density(sp.density_type, p, rho, drho, d2rho);
  const double vis = kViscosity, dvis = 0.0;
  a.rho[c] = rho;
  a.ukvr[c] = Kr / vis;
  a.drho[c] = drho;
  a.dukvr[c] = dKr * dS / vis - Kr / (vis * vis) * dvis;
  const double ph = a.phi[c];
  const double dM = drho * ph * S + rho * ph * dS;
  const double d2M = 2.0 * dS * drho * ph + S * d2rho * ph + d2S * rho * ph;
  a.dM[c] = dM;
  a.d2M[c] = d2M;
  if (sp.form == 0) {
    a.acc[c] = dM * a.Udot[c] * a.vol[c];
  } else {
    a.acc[c] = rho * ph * S * a.vol[c] / (sp.dyn ? sp.dyn[0] : sp.dtime) - a.accum_prev[c];
  }
}

__global__ void k_cell_state(CellStateArgs a) {
  if (a.sp.stop && *a.sp.stop) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < a.nc) cell_state(a, c);
}

// rho phi S V / dt at Uprev (TDyMPFAOSNESPreSolve, tdympfao_snes.c:32-64)
// Diagnostics of TDyUpdateDiagnostics_MPFAO (src/fv/mpfao/tdympfao.c:2759-2786):
// liquid saturation and the reference's "liquid mass" rho*V per cell, plus the
// water mass rho*phi*S*V that a mass balance needs.
__global__ void k_diagnostics(int nc, StateParams sp, const double *__restrict__ U, const double *__restrict__ phi,
                              const double *__restrict__ Sr, const double *__restrict__ vol, double *__restrict__ S_o,
                              double *__restrict__ mass_o, double *__restrict__ water_o) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const double p = U[c];
  double S, dS, d2S;
  vg_saturation(sp.vg_m, sp.vg_alpha, Sr[c], sp.pref - p, S, dS, d2S);
  double rho, drho, d2rho;
  water_density(sp.density_type, p, rho, drho, d2rho);
  if (S_o) S_o[c] = S;
  if (mass_o) mass_o[c] = rho * vol[c];
  if (water_o) water_o[c] = rho * phi[c] * S * vol[c];
}

__global__ void k_snes_presolve(int nc, StateParams sp, const double *__restrict__ U,
                                const double *__restrict__ phi, const double *__restrict__ Sr,
                                const double *__restrict__ vol, double *__restrict__ accum_prev) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double S, dS, d2S, rho, drho, d2rho;
  vg_saturation(sp.vg_m, sp.vg_alpha, Sr[c], sp.pref - U[c], S, dS, d2S);
  water_density(sp.density_type, U[c], rho, drho, d2rho);
  accum_prev[c] = rho * phi[c] * S * vol[c] / sp.dtime;
}

struct BndStateArgs {
  int nbnd;
  StateParams sp;
  const double *U;
  const int32_t *bnd_cell;
  const double *Sr, *pbnd_fn;
  double *Pbnd, *ukvr_bnd;
};

__device__ inline void bnd_state(const BndStateArgs &a, int b) {
  const StateParams &sp = a.sp;
  const int c = a.bnd_cell[b];
  const double P = a.pbnd_fn ? a.pbnd_fn[b] : a.U[c];
  a.Pbnd[b] = P;
  const double sr = a.Sr[c];
  double S, dS, d2S;
  vg_saturation(sp.vg_m, sp.vg_alpha, sr, sp.pref - P, S, dS, d2S);
  const double Se = (S - sr) / (1.0 - sr);
  double Kr, dKr;
  const double cc[4] = {sp.c0, sp.c1, sp.c2, sp.c3};
  mualem_kr(sp.vg_m, sp.x0, sp.x1, sp.dx, cc, Se, Kr, dKr);
  a.ukvr_bnd[b] = Kr / kViscosity;
}

__global__ void k_bnd_state(BndStateArgs a) {
  if (a.sp.stop && *a.sp.stop) return;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < a.nbnd) bnd_state(a, b);
}

struct RowArgs {
  int nc;
  int64_t nrows;
  double pref;
  const int32_t *face_cells;
  const int32_t *face_verts;
  const uint8_t *face_flags;
  const int32_t *vcols;
  const double *T;
  const double *G;
  const double *U;
  const double *Pbnd;
  const double *rho, *ukvr, *drho, *dukvr;
  const double *rho_bnd, *ukvr_bnd;
  double *flux;
  const int *stop;
};

constexpr int kRowBlock = 256;  // k_rows workgroup size

// The arithmetic of one flux row (TDyMPFAOIFunction_Vertices,
// tdympfao_ts.c:12-150; TDyMPFAOIJacobian_Vertices, 235-456), shared by the
// residual's row kernel and the Jacobian's cell kernel so that both take the
// same T.P and the same upwind side.  Inputs: the row's 8 transmissibilities
// T and the vertex's columns ck with their pressures P, rho at the two sides,
// the gravity term G.
__device__ inline double row_tp(const double (&T)[kMaxCols], const int (&ck)[kMaxCols],
                                const double (&P)[kMaxCols]) {
  double TP = 0.0;
#pragma unroll
  for (int k = 0; k < kMaxCols; k++) TP = ck[k] >= 0 ? fma(T[k], P[k], TP) : TP;
  return TP;
}
__device__ inline double row_den(double rho_up, double rho_dn) {
  double den = 0.0;
  den += rho_up;
  den += rho_dn;
  return den * 0.5;
}
__device__ inline bool row_upwind_up(double TP, double den, double G) { return fma(den, G, TP) < 0.0; }
__device__ inline double row_flux(double TP, double den, double G, double ukvr, bool zero) {
  double fluxm = den * ukvr * (-TP);
  fluxm += -(den * den) * ukvr * G;
  return zero ? 0.0 : fluxm;
}
// d(flux)/d(p of column k), negated (the up cell's residual gets -Jac, the
// down cell's +Jac, tdympfao_ts.c:439-444); 0 for padding and boundary-pressure
// columns, which never enter J, and for a row that carries no flux
__device__ inline void row_derivs(int nc, const double (&T)[kMaxCols], const int (&ck)[kMaxCols], int up, int dn,
                                  double TP, double den, double G, double ukvr, bool upwind_up, bool zero,
                                  double dr_up, double dr_dn, double du, double (&jr)[kMaxCols]) {
  const double dden_up = up >= 0 ? 0.5 * dr_up : 0.0;
  const double dden_dn = dn >= 0 ? 0.5 * dr_dn : 0.0;
  const double dukvr_up = (upwind_up && up >= 0) ? du : 0.0;
  const double dukvr_dn = (!upwind_up && dn >= 0) ? du : 0.0;
  // the up (down) cell's column adds the derivatives of rho_bar and kr/mu
  const double Aup = dden_up * ukvr * TP + den * dukvr_up * TP;
  const double Dup = 2 * den * dden_up * ukvr * G, Eup = (den * den) * dukvr_up * G;
  const double Adn = dden_dn * ukvr * TP + den * dukvr_dn * TP;
  const double Ddn = 2 * den * dden_dn * ukvr * G, Edn = (den * den) * dukvr_dn * G;
#pragma unroll
  for (int k = 0; k < kMaxCols; k++) {
    const int col = ck[k];
    const double C = den * ukvr * T[k];
    const double Ju = ((Aup + C) + Dup) + Eup, Jd = ((Adn + C) + Ddn) + Edn;
    const double Jac = col == up ? Ju : (col == dn ? Jd : C);
    jr[k] = (zero || col < 0 || col >= nc) ? 0.0 : -Jac;
  }
}

// the columns of vertex v and their pressures (padding reads U[0], unused)
__device__ inline void vertex_columns(const RowArgs &a, int v, int (&ck)[kMaxCols], double (&P)[kMaxCols]) {
  const int4 c0 = *reinterpret_cast<const int4 *>(a.vcols + (int64_t)v * kMaxCols);
  const int4 c1 = *reinterpret_cast<const int4 *>(a.vcols + (int64_t)v * kMaxCols + 4);
  ck[0] = c0.x; ck[1] = c0.y; ck[2] = c0.z; ck[3] = c0.w;
  ck[4] = c1.x; ck[5] = c1.y; ck[6] = c1.z; ck[7] = c1.w;
#pragma unroll
  for (int k = 0; k < kMaxCols; k++) {
    const int col = ck[k];
    P[k] = *(col < a.nc ? a.U + (col < 0 ? 0 : col) : a.Pbnd + (col - a.nc));
  }
}

// Thread per flux row: the residual's flux rows (the Jacobian is assembled by
// k_cell_jac, one pass per cell).
__device__ inline void flux_row(const RowArgs &a, int64_t r) {
  const int f = (int)(r >> 2);
  const uint8_t fl = a.face_flags[f];
  double flux = 0.0;
  if (!(fl & kFaceNoFlow)) {
    const int2 ud = *reinterpret_cast<const int2 *>(a.face_cells + 2 * f);
    const int up = ud.x, dn = ud.y;
    int ck[kMaxCols];
    double P[kMaxCols], T[kMaxCols];
    vertex_columns(a, a.face_verts[r], ck, P);
#pragma unroll
    for (int k = 0; k < kMaxCols; k++) T[k] = a.T[(int64_t)k * a.nrows + r];
    const double G = a.G[r];
    const double den = row_den(*(up >= 0 ? a.rho + up : a.rho_bnd + (-up - 1)),
                               *(dn >= 0 ? a.rho + dn : a.rho_bnd + (-dn - 1)));
    const double TP = row_tp(T, ck, P);
    const bool upwind_up = row_upwind_up(TP, den, G);
    const int uw = upwind_up ? up : dn;
    const double ukvr = *(uw >= 0 ? a.ukvr + uw : a.ukvr_bnd + (-uw - 1));
    // a seepage face's boundary side carries no flux while its pressure is at
    // or below the reference pressure
    const bool zero = uw < 0 && (fl & kFaceSeepage) && a.Pbnd[-uw - 1] <= a.pref;
    flux = row_flux(TP, den, G, ukvr, zero);
  }
  a.flux[r] = flux;
}

__global__ void __launch_bounds__(kRowBlock) k_rows(RowArgs a) {
  if (a.stop && *a.stop) return;  // uniform over the workgroup
  const int64_t r = (int64_t)blockIdx.x * kRowBlock + threadIdx.x;
  if (r < a.nrows) flux_row(a, r);
}

// ---------------------------------------------------------------- general (prism) meshes
// Flux rows whose vertex has any number of columns: the row's T entries sit in
// its vertex's block (row-major, columns = the region's cells, then its
// Dirichlet faces), and its derivatives D in the same place for the cell
// columns.  The physics is k_rows' (tdympfao_ts.c:12-150, 235-456).
struct GenRowArgs {
  int nc;
  int64_t nrows;
  double pref;
  const int32_t *face_cells;
  const uint8_t *face_flags;
  const int32_t *rtoff, *rinfo, *roc, *rdc, *ccell, *vdcol;
  const double *T, *G, *U, *Pbnd, *rho, *ukvr, *drho, *dukvr, *rho_bnd, *ukvr_bnd;
  double *flux, *D;
  const int *stop;
};

template <bool WITH_R, bool WITH_J>
__global__ void __launch_bounds__(256) k_rows_general(GenRowArgs a) {
  if (a.stop && *a.stop) return;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.nrows) return;
  const int f = (int)(r >> 2);
  const uint8_t fl = a.face_flags[f];
  const int to = a.rtoff[r];
  if (to < 0 || (fl & kFaceNoFlow)) {
    if (WITH_R) a.flux[r] = 0.0;
    return;
  }
  const int ncell = a.rinfo[r] & 0xff, ndir = (a.rinfo[r] >> 8) & 0xff;
  const int32_t *cols = a.ccell + a.roc[r];
  const int32_t *dcols = a.vdcol + a.rdc[r];
  const double *Tr = a.T + to;
  double TP = 0.0;
#pragma unroll 4
  for (int j = 0; j < ncell; j++) TP += Tr[j] * a.U[cols[j]];
  for (int j = 0; j < ndir; j++) TP += Tr[ncell + j] * a.Pbnd[dcols[j]];
  const int up = a.face_cells[2 * f], dn = a.face_cells[2 * f + 1];
  double den = 0.0;
  den += up >= 0 ? a.rho[up] : a.rho_bnd[-up - 1];
  den += dn >= 0 ? a.rho[dn] : a.rho_bnd[-dn - 1];
  den *= 0.5;
  const double G = a.G[r];
  const bool upwind_up = (TP + den * G < 0.0);
  const int uw = upwind_up ? up : dn;
  double ukvr;
  bool zero = false;
  if (uw >= 0) {
    ukvr = a.ukvr[uw];
  } else {
    ukvr = a.ukvr_bnd[-uw - 1];
    if ((fl & kFaceSeepage) && a.Pbnd[-uw - 1] <= a.pref) zero = true;
  }
  if (WITH_R) {
    double fluxm = 0.0;
    if (!zero) {
      fluxm = den * ukvr * (-TP);
      fluxm += -(den * den) * ukvr * G;
    }
    a.flux[r] = fluxm;
  }
  if (WITH_J) {
    const double dden_up = up >= 0 ? 0.5 * a.drho[up] : 0.0;
    const double dden_dn = dn >= 0 ? 0.5 * a.drho[dn] : 0.0;
    double dukvr_up = 0.0, dukvr_dn = 0.0;
    if (upwind_up) { if (up >= 0) dukvr_up = a.dukvr[up]; }
    else { if (dn >= 0) dukvr_dn = a.dukvr[dn]; }
#pragma unroll 4
    for (int j = 0; j < ncell; j++) {
      const int col = cols[j];
      double Jac;
      if (zero) Jac = 0.0;
      else if (up >= 0 && col == up)
        Jac = dden_up * ukvr * TP + den * dukvr_up * TP + den * ukvr * Tr[j] + 2 * den * dden_up * ukvr * G +
              (den * den) * dukvr_up * G;
      else if (dn >= 0 && col == dn)
        Jac = dden_dn * ukvr * TP + den * dukvr_dn * TP + den * ukvr * Tr[j] + 2 * den * dden_dn * ukvr * G +
              (den * den) * dukvr_dn * G;
      else
        Jac = den * ukvr * Tr[j];
      a.D[to + j] = -Jac;  // the up cell's residual gets -Jac, the down cell's +Jac (tdympfao_ts.c:439-444)
    }
  }
}

// Jacobian row of each cell (CSR in the vertex-adjacency pattern, or the
// solver's ELL): the derivatives of its faces' flux rows in a fixed (face,
// sub-face) row order, then the accumulation diagonal.  kGenJLanes lanes own
// one row, summed in LDS (up to kGenJSlots entries), every slot written once.
// Lane l first reads row l's metadata (face lf = l/4, sub-face l%4); then a
// lane per column loads the rows' derivatives eight rows at a time and adds
// them row by row.  A flux row adds to each slot at most once (its columns
// are distinct cells), so every slot is summed in row order.
constexpr int kGenJBlock = 256;
constexpr int kGenJLanes = 32;  // lanes per cell: columns of a flux row (a row of more loops)
constexpr int kGenJSlots = 64;
struct GenGatherArgs {
  int nc, form, ell_w;
  int jbs;               // ELL: block size of the block-contiguous layout (see jac_row_addr)
  const int32_t *perm;   // ELL: position of row c (nullptr: CSR at jrp[c])
  double shift, dtime;
  const int32_t *cell_faces, *face_cells, *rtoff, *rinfo, *roc, *ccell, *jrp, *jcols, *jdiag;
  const uint8_t *face_flags, *gslot;
  const double *D, *dM, *d2M, *Udot, *vol;
  double *J;
  const double *dyn;  // device dtime / shift (see StateParams), or nullptr
  const int *stop;
};

// The same flux rows, one wavefront per vertex block: first a lane per block
// row computes the row's flux and its scalars (T.P, mean density, upwind
// mobility and their derivatives) into LDS, then the lanes sweep the block's
// cell entries in memory order and write the derivatives D, so T and D are
// read and written in contiguous runs.  Rows without a block (a triangle's
// fourth sub-face, no-flow faces) keep the zero flux written at setup.  Same
// arithmetic as k_rows_general, entry by entry.
constexpr int kBlkRows = 64;  // most flux rows of one vertex block (checked at setup)
struct GenBlkArgs {
  double pref;
  const int32_t *vblk;  // per vertex: T offset, rows, cells, Dirichlet columns, cell offset, Dirichlet offset, row offset
  const int32_t *brow;  // flux row of each block row, blocks in vertex order
  const int32_t *face_cells;
  const uint8_t *face_flags;
  const int32_t *ccell, *vdcol;
  const double *T, *G, *U, *Pbnd, *rho, *ukvr, *drho, *dukvr, *rho_bnd, *ukvr_bnd;
  double *flux, *D;
  const int *stop;
};

template <bool WITH_R, bool WITH_J>
__global__ void __launch_bounds__(64) k_rows_general_blk(GenBlkArgs a) {
  if (a.stop && *a.stop) return;  // uniform over the workgroup
  __shared__ double s_tp[kBlkRows], s_den[kBlkRows], s_ukvr[kBlkRows], s_G[kBlkRows];
  __shared__ double s_ddu[kBlkRows], s_ddd[kBlkRows], s_dku[kBlkRows], s_dkd[kBlkRows];
  __shared__ int s_up[kBlkRows], s_dn[kBlkRows], s_flag[kBlkRows];  // flag: 1 no D, 2 zero D
  const int v = blockIdx.x, t = threadIdx.x;
  const int32_t *d = a.vblk + (int64_t)v * 7;
  const int ot = d[0], nrow = d[1], ncell = d[2], ndir = d[3];
  const int32_t *cols = a.ccell + d[4];
  const int32_t *dcols = a.vdcol + d[5];
  const int nB = ncell + ndir;
  if (t < nrow) {
    const int64_t r = a.brow[d[6] + t];
    const int f = (int)(r >> 2);
    const uint8_t fl = a.face_flags[f];
    int flag = 0;
    if (fl & kFaceNoFlow) {
      if (WITH_R) a.flux[r] = 0.0;
      flag = 1;
    } else {
      const double *Tr = a.T + ot + (int64_t)t * nB;
      double TP = 0.0;
#pragma unroll 4
      for (int j = 0; j < ncell; j++) TP += Tr[j] * a.U[cols[j]];
      for (int j = 0; j < ndir; j++) TP += Tr[ncell + j] * a.Pbnd[dcols[j]];
      const int up = a.face_cells[2 * f], dn = a.face_cells[2 * f + 1];
      double den = 0.0;
      den += up >= 0 ? a.rho[up] : a.rho_bnd[-up - 1];
      den += dn >= 0 ? a.rho[dn] : a.rho_bnd[-dn - 1];
      den *= 0.5;
      const double G = a.G[r];
      const bool upwind_up = (TP + den * G < 0.0);
      const int uw = upwind_up ? up : dn;
      double ukvr;
      bool zero = false;
      if (uw >= 0) {
        ukvr = a.ukvr[uw];
      } else {
        ukvr = a.ukvr_bnd[-uw - 1];
        if ((fl & kFaceSeepage) && a.Pbnd[-uw - 1] <= a.pref) zero = true;
      }
      if (WITH_R) {
        double fluxm = 0.0;
        if (!zero) {
          fluxm = den * ukvr * (-TP);
          fluxm += -(den * den) * ukvr * G;
        }
        a.flux[r] = fluxm;
      }
      if (WITH_J) {
        double dukvr_up = 0.0, dukvr_dn = 0.0;
        if (upwind_up) { if (up >= 0) dukvr_up = a.dukvr[up]; }
        else { if (dn >= 0) dukvr_dn = a.dukvr[dn]; }
        s_tp[t] = TP; s_den[t] = den; s_ukvr[t] = ukvr; s_G[t] = G;
        s_ddu[t] = up >= 0 ? 0.5 * a.drho[up] : 0.0;
        s_ddd[t] = dn >= 0 ? 0.5 * a.drho[dn] : 0.0;
        s_dku[t] = dukvr_up; s_dkd[t] = dukvr_dn;
        s_up[t] = up; s_dn[t] = dn;
      }
      flag = zero ? 2 : 0;
    }
    if (WITH_J) s_flag[t] = flag;
  }
  if (!WITH_J) return;
  __syncthreads();
  const int ne = nrow * ncell;
  for (int e = t; e < ne; e += 64) {
    const int i = e / ncell, j = e - i * ncell;
    const int flag = s_flag[i];
    if (flag & 1) continue;
    const int64_t q = (int64_t)ot + (int64_t)i * nB + j;
    const int col = cols[j];
    const int up = s_up[i], dn = s_dn[i];
    const double TP = s_tp[i], den = s_den[i], ukvr = s_ukvr[i], G = s_G[i];
    double Jac;
    if (flag & 2) Jac = 0.0;
    else if (up >= 0 && col == up)
      Jac = s_ddu[i] * ukvr * TP + den * s_dku[i] * TP + den * ukvr * a.T[q] + 2 * den * s_ddu[i] * ukvr * G +
            (den * den) * s_dku[i] * G;
    else if (dn >= 0 && col == dn)
      Jac = s_ddd[i] * ukvr * TP + den * s_dkd[i] * TP + den * ukvr * a.T[q] + 2 * den * s_ddd[i] * ukvr * G +
            (den * den) * s_dkd[i] * G;
    else
      Jac = den * ukvr * a.T[q];
    a.D[q] = -Jac;  // the up cell's residual gets -Jac, the down cell's +Jac (tdympfao_ts.c:439-444)
  }
}

// Slot table of the general Jacobian gather, built once at setup: entry
// to + j of flux row r (its cell column j) has the column's slot in the up
// cell's Jacobian row at gslot[2*(to+j)] and in the down cell's at
// gslot[2*(to+j)+1] (0xff for a boundary side or a no-flow face).  One
// thread per flux row; the slot is found by binary search in the sorted row.
__global__ void __launch_bounds__(256) k_gen_slot_table(int64_t nrows, const int32_t *__restrict__ face_cells,
                                                        const uint8_t *__restrict__ face_flags,
                                                        const int32_t *__restrict__ rtoff,
                                                        const int32_t *__restrict__ rinfo,
                                                        const int32_t *__restrict__ roc,
                                                        const int32_t *__restrict__ ccell,
                                                        const int32_t *__restrict__ jrp,
                                                        const int32_t *__restrict__ jcols,
                                                        uint8_t *__restrict__ gslot) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const int to = rtoff[r];
  if (to < 0) return;
  const int f = (int)(r >> 2);
  const int ncell = rinfo[r] & 0xff;
  const int32_t *cols = ccell + roc[r];
  const bool noflow = (face_flags[f] & kFaceNoFlow) != 0;
  for (int side = 0; side < 2; side++) {
    const int c = face_cells[2 * f + side];
    for (int j = 0; j < ncell; j++) {
      uint8_t s = 0xff;
      if (c >= 0 && !noflow) {
        const int r0 = jrp[c];
        int lo = r0, hi = jrp[c + 1] - 1;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (jcols[mid] < cols[j]) lo = mid + 1; else hi = mid;
        }
        s = (uint8_t)(lo - r0);
      }
      gslot[2 * ((int64_t)to + j) + side] = s;
    }
  }
}

__global__ void __launch_bounds__(kGenJBlock) k_gather_general_J(GenGatherArgs a) {
  constexpr int CPB = kGenJBlock / kGenJLanes;  // cells per workgroup
  constexpr int NR = 24;                        // flux rows of a cell: 6 faces x 4 sub-faces
  constexpr int RB = 8;                         // rows whose loads are in flight together
  static_assert(kGenJLanes >= NR && 64 % kGenJLanes == 0, "a cell's lanes hold its rows and share a wavefront");
  __shared__ double acc[CPB][kGenJSlots];
  if (a.stop && *a.stop) return;
  const int g = threadIdx.x / kGenJLanes, l = threadIdx.x % kGenJLanes;
  const int base = (threadIdx.x & 63) - l;  // the cell's first lane within the wavefront
  const int c = blockIdx.x * CPB + g;
  if (c >= a.nc) return;  // uniform over the cell's lanes (the wavefront's other cell is independent)
  const int r0 = a.jrp[c], len = a.jrp[c + 1] - r0;
  for (int k = l; k < len; k += kGenJLanes) acc[g][k] = 0.0;
  // row l's metadata: its derivatives at D[to ..], its slots, the sign by the cell's side
  int to = -1, nce = 0, side = 0;
  bool up = false;
  if (l < NR) {
    const int f = a.cell_faces[c * 6 + (l >> 2)];
    if (f >= 0 && !(a.face_flags[f] & kFaceNoFlow)) {
      const int64_t r = (int64_t)f * 4 + (l & 3);
      to = a.rtoff[r];
      nce = to >= 0 ? (a.rinfo[r] & 0xff) : 0;
      up = a.face_cells[2 * f] == c;
      side = up ? 0 : 1;
    }
  }
  for (int i0 = 0; i0 < NR; i0 += RB) {
    double dv[RB];
    int sv[RB];
#pragma unroll
    for (int k = 0; k < RB; k++) {
      const int ti = __shfl(to, base + i0 + k, 64), ni = __shfl(nce, base + i0 + k, 64);
      const int si = __shfl(side, base + i0 + k, 64);
      const bool on = ti >= 0 && l < ni;
      dv[k] = on ? a.D[ti + l] : 0.0;
      sv[k] = on ? a.gslot[2 * ((int64_t)ti + l) + si] : -1;
    }
#pragma unroll
    for (int k = 0; k < RB; k++) {
      const int ti = __shfl(to, base + i0 + k, 64), ni = __shfl(nce, base + i0 + k, 64);
      const int si = __shfl(side, base + i0 + k, 64);
      const bool ui = __shfl((int)up, base + i0 + k, 64) != 0;
      if (sv[k] >= 0) acc[g][sv[k]] += ui ? dv[k] : -dv[k];
      // columns past the lanes (regions of more than kGenJLanes cells): the rest of the same row
      for (int j = l + kGenJLanes; ti >= 0 && j < ni; j += kGenJLanes) {
        const double d = a.D[ti + j];
        acc[g][a.gslot[2 * ((int64_t)ti + j) + si]] += ui ? d : -d;
      }
    }
  }
  if (l == 0) {
    const double shift = a.dyn ? a.dyn[1] : a.shift, dtime = a.dyn ? a.dyn[0] : a.dtime;
    const double dvc = (a.form == 0) ? (shift * a.dM[c] + a.d2M[c] * a.Udot[c]) * a.vol[c]
                                     : a.dM[c] * a.vol[c] * (1.0 / dtime);
    acc[g][a.jdiag[c]] += dvc;
  }
  int64_t st = 1;
  double *row = a.J + r0;
  if (a.perm) {
    const int64_t p = a.perm[c], bs = a.jbs;
    row = a.J + (p / bs * a.ell_w) * bs + p % bs;
    st = bs;
  }
  const int nslot = a.perm ? a.ell_w : len;  // ELL rows are padded to ell_w slots with zeros
  for (int k = l; k < nslot; k += kGenJLanes) row[k * st] = k < len ? acc[g][k] : 0.0;
}

struct GatherArgs {
  int nc;
  int64_t nrows;
  int form;  // 0 TS, 1 SNES
  double shift, dtime;
  const int32_t *cell_faces, *face_cells;
  const uint8_t *face_flags, *jdslot;
  const double *flux, *acc, *dM, *d2M, *Udot, *vol;
  const double *src;   // source_sink per cell, or nullptr (TS form: scaled by the volume)
  const int32_t *jrp;  // CSR row starts, or nullptr for the slot-major ELL layout
  const int32_t *perm; // ELL: position of row c (nullptr: c)
  int jbs;             // ELL: block size bs of the block-contiguous layout (slot k of position p at
                       // (p/bs*ell_w + k)*bs + p%bs)
  int ell_w;           // ELL width (slots written per row in the ELL layout)
  double *R, *J;
  const double *dyn;   // device dtime / shift (see StateParams), or nullptr
  const int *stop;
};

constexpr int kGatherBlock = 128;
constexpr int kJacSlots = 27;  // vertex-adjacency row length of a hexahedral mesh

// Where cell c's Jacobian row starts and the distance between its slots: CSR
// values at jrp[c] (stride 1), or the solver's block-contiguous ELL.
__device__ inline void jac_row_addr(const GatherArgs &a, int c, int64_t &base, int64_t &stride) {
  if (a.jrp) { base = a.jrp[c]; stride = 1; return; }
  const int64_t p = a.perm ? a.perm[c] : c, bs = a.jbs;
  base = (p / bs * a.ell_w) * bs + p % bs;
  stride = bs;
}

// R[c] from the flux rows: the cell's faces' four rows summed in sub-face
// order, +/- by orientation, then the accumulation and the source
// (tdympfao_ts.c:191-212, tdympfao_snes.c:112-123).
__device__ inline double cell_residual(const GatherArgs &a, int c, double s) {
  double r = s + a.acc[c];
  if (a.src) r -= (a.form == 0) ? a.src[c] * a.vol[c] : a.src[c];
  return r;
}

// Per-cell residual gather (thread per cell): every cell owns its entry, so
// the result does not depend on scheduling.
__device__ inline void cell_gather(const GatherArgs &a, int c) {
  double s = 0.0;
#pragma unroll
  for (int lf = 0; lf < 6; lf++) {
    const int f = a.cell_faces[c * 6 + lf];
    if (f < 0) continue;  // a prism has 5 faces (a triangle's fourth flux row is zero)
    const double4 q = *reinterpret_cast<const double4 *>(a.flux + (int64_t)f * 4);
    const double fs = ((q.x + q.y) + q.z) + q.w;
    s += (a.face_cells[2 * f] == c) ? fs : -fs;
  }
  a.R[c] = cell_residual(a, c, s);
}

__global__ void __launch_bounds__(kGatherBlock) k_cell_gather(GatherArgs a) {
  if (a.stop && *a.stop) return;
  const int c = blockIdx.x * kGatherBlock + threadIdx.x;
  if (c < a.nc) cell_gather(a, c);
}

// Single-pass Jacobian (and, WITH_R, residual) of a hexahedral mesh
// (TDyMPFAOIJacobian_Vertices, tdympfao_ts.c:235-456, evaluated from the
// cell's side): eight lanes per cell, one per corner.  The three faces of the
// cell that meet at a corner have one flux row each at that corner's vertex;
// the three rows share the vertex's 8 columns and pressures (gathered once)
// and the 8 slots of the cell's Jacobian row those columns occupy, so the lane
// sums the rows' derivatives in registers (face order) and stages the 8 sums
// and their slots in LDS.  After a barrier the corner-0 lane adds the eight
// corners into the row in corner order (a fixed summation order, so the result
// does not depend on scheduling) plus the accumulation diagonal, and the eight
// lanes write the row.  Each flux row is evaluated by both cells of its face;
// the second read of T hits the cache when the two cells run close together,
// which `order` (position -> cell: the block-Jacobi layout's row list, -1
// padding) and the XCD-grouped workgroup numbering arrange.
//   crow[c*8 + j]: the corner's three rows as 5-bit (face slot * 4 + sub-face)
//                  codes in face order (bits 0-4, 5-9, 10-14), the cell's index
//                  in the region of the corner's vertex (bits 15-17), and
//                  whether the cell is each row's up cell (bits 18, 19, 20)
//   cell_verts[c*8 + j]: the corner's vertex
//   jslot[v*64 + li*8 + k]: slot of column k of vertex v's region in the row
//                  of the region's cell li (255: not a cell column)
// A no-flow face's row is evaluated like the others, at the cell itself on
// both sides (every read in range), and dropped.
struct CellJacArgs {
  const int32_t *order;  // position -> cell, or nullptr (position i is cell i)
  int64_t npos;
  int xcd_groups;        // 0: workgroups in launch order; else the launch is 8 * xcd_groups workgroups
  const uint32_t *crow;
  const int32_t *cell_verts;
  const uint8_t *jslot;
};

constexpr int kCellJacBlock = 256;                // lanes per workgroup
constexpr int kCellJacCells = kCellJacBlock / 8;  // cells per workgroup

// block blk of the cells (in `order`): kCellJacCells cells, 8 lanes each
template <bool WITH_R>
__device__ inline void cell_jac_block(const RowArgs &ra, const GatherArgs &a, const CellJacArgs &cj, int64_t blk) {
  __shared__ double sd[kMaxCols][kCellJacBlock];     // corner sums: column k of lane t at sd[k][t]
  __shared__ uint64_t sw[kCellJacBlock];             // their slots (one byte per column)
  __shared__ double sf[WITH_R ? 24 : 1][kCellJacCells];  // flux of row lf*4+q of cell cl
  __shared__ double acc[kJacSlots + 1][kCellJacCells];   // the row (slot kJacSlots: non-cell columns)
  const int t = threadIdx.x, cl = t >> 3, corner = t & 7;
  const int64_t i = blk * kCellJacCells + cl;
  const int c = i < cj.npos ? (cj.order ? cj.order[i] : (int)i) : -1;  // -1: no cell
  if (c >= 0) {
    for (int sl = corner; sl < kJacSlots; sl += 8) acc[sl][cl] = 0.0;
    const unsigned code = cj.crow[(int64_t)c * 8 + corner];
    const int v = cj.cell_verts[(int64_t)c * 8 + corner];
    const int li = (code >> 15) & 7;
    int fr[3], lfq[3];
#pragma unroll
    for (int m = 0; m < 3; m++) {
      lfq[m] = (code >> (5 * m)) & 31;
      fr[m] = a.cell_faces[c * 6 + (lfq[m] >> 2)];
    }
    int ck[kMaxCols];
    double P[kMaxCols];
    vertex_columns(ra, v, ck, P);
    const uint64_t w = *reinterpret_cast<const uint64_t *>(cj.jslot + (int64_t)v * 64 + li * 8);
    // the three rows' loads in two batches: what the faces fix, then what the
    // upwind side fixes (a no-flow face's row runs at the cell on both sides,
    // every read in range, and is dropped)
    int up[3], dn[3];
    bool noflow[3], seep[3];
    double T[3][kMaxCols], G[3], rup[3], rdn[3], drup[3], drdn[3];
#pragma unroll
    for (int m = 0; m < 3; m++) {
      const int f = fr[m];
      const int64_t r = (int64_t)f * 4 + (lfq[m] & 3);
      const uint8_t fl = a.face_flags[f];
      const int2 ud = *reinterpret_cast<const int2 *>(a.face_cells + 2 * f);
      noflow[m] = (fl & kFaceNoFlow) != 0;
      seep[m] = (fl & kFaceSeepage) != 0;
      up[m] = noflow[m] ? c : ud.x;
      dn[m] = noflow[m] ? c : ud.y;
#pragma unroll
      for (int k = 0; k < kMaxCols; k++) T[m][k] = ra.T[(int64_t)k * ra.nrows + r];
      G[m] = ra.G[r];
      rup[m] = *(up[m] >= 0 ? ra.rho + up[m] : ra.rho_bnd + (-up[m] - 1));
      rdn[m] = *(dn[m] >= 0 ? ra.rho + dn[m] : ra.rho_bnd + (-dn[m] - 1));
      drup[m] = ra.drho[up[m] >= 0 ? up[m] : 0];
      drdn[m] = ra.drho[dn[m] >= 0 ? dn[m] : 0];
    }
    double TP[3], den[3], ukvr[3], du[3], pb[3];
    bool upw[3];
    int uw[3];
#pragma unroll
    for (int m = 0; m < 3; m++) {
      TP[m] = row_tp(T[m], ck, P);
      den[m] = row_den(rup[m], rdn[m]);
      upw[m] = row_upwind_up(TP[m], den[m], G[m]);
      uw[m] = upw[m] ? up[m] : dn[m];
    }
#pragma unroll
    for (int m = 0; m < 3; m++) {
      ukvr[m] = *(uw[m] >= 0 ? ra.ukvr + uw[m] : ra.ukvr_bnd + (-uw[m] - 1));
      du[m] = ra.dukvr[uw[m] >= 0 ? uw[m] : 0];
      pb[m] = ra.Pbnd[uw[m] < 0 ? -uw[m] - 1 : 0];  // Pbnd always holds at least one entry
    }
    double sum[kMaxCols];
#pragma unroll
    for (int k = 0; k < kMaxCols; k++) sum[k] = 0.0;
#pragma unroll
    for (int m = 0; m < 3; m++) {
      const bool zero = (uw[m] < 0) & seep[m] & (pb[m] <= ra.pref);
      double jr[kMaxCols];
      row_derivs(ra.nc, T[m], ck, up[m], dn[m], TP[m], den[m], G[m], ukvr[m], upw[m], zero, drup[m], drdn[m], du[m], jr);
      const bool is_up = (code >> (18 + m)) & 1;
#pragma unroll
      for (int k = 0; k < kMaxCols; k++) sum[k] = noflow[m] ? sum[k] : (is_up ? sum[k] + jr[k] : sum[k] - jr[k]);
      if (WITH_R) sf[lfq[m]][cl] = noflow[m] ? 0.0 : row_flux(TP[m], den[m], G[m], ukvr[m], zero);
    }
#pragma unroll
    for (int k = 0; k < kMaxCols; k++) sd[k][t] = sum[k];
    sw[t] = w;
  }
  __syncthreads();
  if (c >= 0 && corner == 0) {
    for (int j = 0; j < 8; j++) {
      // a region's cells are distinct, so a corner's 8 slots are distinct (apart
      // from kJacSlots, which is never read): all reads before the writes
      const uint64_t wj = sw[t + j];
      unsigned sb[kMaxCols];
      double cur[kMaxCols];
#pragma unroll
      for (int k = 0; k < kMaxCols; k++) {
        sb[k] = (unsigned)(wj >> (8 * k)) & 0xffu;
        if (sb[k] == 0xffu) sb[k] = kJacSlots;
        cur[k] = acc[sb[k]][cl];
      }
#pragma unroll
      for (int k = 0; k < kMaxCols; k++) acc[sb[k]][cl] = cur[k] + sd[k][t + j];
    }
    const double shift = a.dyn ? a.dyn[1] : a.shift, dtime = a.dyn ? a.dyn[0] : a.dtime;
    const double dv = (a.form == 0) ? (shift * a.dM[c] + a.d2M[c] * a.Udot[c]) * a.vol[c]
                                    : a.dM[c] * a.vol[c] * (1.0 / dtime);
    const int ds = a.jdslot[c];
    acc[ds][cl] = acc[ds][cl] + dv;
    if (WITH_R) {
      // fluxes, then accumulation, then the source (tdympfao_ts.c:191-212, tdympfao_snes.c:112-123)
      double s = 0.0;
      for (int lf = 0; lf < 6; lf++) {
        const int f = a.cell_faces[c * 6 + lf];
        const double fs = ((sf[lf * 4][cl] + sf[lf * 4 + 1][cl]) + sf[lf * 4 + 2][cl]) + sf[lf * 4 + 3][cl];
        s += (a.face_cells[2 * f] == c) ? fs : -fs;
      }
      a.R[c] = cell_residual(a, c, s);
    }
  }
  __syncthreads();
  if (c < 0) return;
  int64_t base = c, stride = 1;
  jac_row_addr(a, c, base, stride);
  const int len = a.jrp ? a.jrp[c + 1] - a.jrp[c] : a.ell_w;
  for (int sl = corner; sl < len; sl += 8) a.J[base + (int64_t)sl * stride] = acc[sl][cl];
}

template <bool WITH_R>
__global__ void __launch_bounds__(kCellJacBlock) k_cell_jac(RowArgs ra, GatherArgs a, CellJacArgs cj) {
  if (a.stop && *a.stop) return;  // uniform over the workgroup
  // with xcd_groups, hardware round-robin places workgroup w on XCD w % 8:
  // give XCD x the contiguous run of blocks [x * xcd_groups, (x + 1) * xcd_groups)
  const int64_t blk = cj.xcd_groups ? (int64_t)(blockIdx.x % 8) * cj.xcd_groups + blockIdx.x / 8 : blockIdx.x;
  cell_jac_block<WITH_R>(ra, a, cj, blk);
}

// ---------------------------------------------------------------- fused multi-request launches
// One launch runs the same evaluation for several independent problems (the
// requests of a serving batch), each a segment of the grid: a workgroup
// belongs to the segment whose block range holds it and computes exactly
// what the single-problem kernel computes for its local block, so each
// request's result is the same bit for bit whatever else shares the launch.
// A segment whose stop flag is set does nothing.
constexpr int kMaxSeg = 32;
struct SegBlocks {
  int nseg;
  int64_t start[kMaxSeg + 1];  // first global block of each segment; start[nseg] = the grid
};
__device__ inline int seg_find(const SegBlocks &sb, int64_t b) {
  int s = 0;
  while (s + 1 < sb.nseg && b >= sb.start[s + 1]) s++;
  return s;
}
struct SegProblem {  // one request's arguments of every fused launch
  CellStateArgs cs;
  BndStateArgs bs;
  RowArgs ra;
  GatherArgs ga;
  CellJacArgs cj;
  const int *stop;
};

__global__ void __launch_bounds__(256) k_seg_state(const SegProblem *__restrict__ P, SegBlocks sb) {
  const int s = seg_find(sb, blockIdx.x);
  if (P[s].stop && *P[s].stop) return;
  const int c = (int)((blockIdx.x - sb.start[s]) * 256 + threadIdx.x);
  if (c < P[s].cs.nc) cell_state(P[s].cs, c);
}
__global__ void __launch_bounds__(256) k_seg_bnd_state(const SegProblem *__restrict__ P, SegBlocks sb) {
  const int s = seg_find(sb, blockIdx.x);
  if (P[s].stop && *P[s].stop) return;
  const int b = (int)((blockIdx.x - sb.start[s]) * 256 + threadIdx.x);
  if (b < P[s].bs.nbnd) bnd_state(P[s].bs, b);
}
__global__ void __launch_bounds__(kRowBlock) k_seg_rows(const SegProblem *__restrict__ P, SegBlocks sb) {
  const int s = seg_find(sb, blockIdx.x);
  if (P[s].stop && *P[s].stop) return;
  const int64_t r = (blockIdx.x - sb.start[s]) * kRowBlock + threadIdx.x;
  if (r < P[s].ra.nrows) flux_row(P[s].ra, r);
}
__global__ void __launch_bounds__(kGatherBlock) k_seg_cell_gather(const SegProblem *__restrict__ P, SegBlocks sb) {
  const int s = seg_find(sb, blockIdx.x);
  if (P[s].stop && *P[s].stop) return;
  const int c = (int)((blockIdx.x - sb.start[s]) * kGatherBlock + threadIdx.x);
  if (c < P[s].ga.nc) cell_gather(P[s].ga, c);
}
__global__ void __launch_bounds__(kCellJacBlock) k_seg_cell_jac(const SegProblem *__restrict__ P, SegBlocks sb) {
  // the segment is uniform over the workgroup: its arguments are copied out
  // once (scalar loads) rather than read through the table at every use
  const int s = __builtin_amdgcn_readfirstlane(seg_find(sb, blockIdx.x));
  if (P[s].stop && *P[s].stop) return;
  const RowArgs ra = P[s].ra;
  const GatherArgs ga = P[s].ga;
  const CellJacArgs cj = P[s].cj;
  cell_jac_block<false>(ra, ga, cj, blockIdx.x - sb.start[s]);
}

// ELL SpMV over the slot-major layout: slot s of row i at s*n + i.
template <class TV>
__device__ inline TV ell_row(int n, int W, int i, const int32_t *__restrict__ ecol, const TV *__restrict__ a,
                            const TV *__restrict__ x) {
  TV s = 0;
  // FP32 rows of a hexahedral mesh: the whole row's loads issued at once (in
  // FP64 the 81 extra registers cost more occupancy than they gain)
  if (sizeof(TV) == 4 && W == kJacSlots) {
    int c[kJacSlots];
    TV v[kJacSlots];
#pragma unroll
    for (int k = 0; k < kJacSlots; k++) {
      c[k] = ecol[(int64_t)k * n + i];
      v[k] = a[(int64_t)k * n + i];
    }
#pragma unroll
    for (int k = 0; k < kJacSlots; k++) s += v[k] * x[c[k]];
  } else {
    for (int k = 0; k < W; k++) {
      const int64_t p = (int64_t)k * n + i;
      s += a[p] * x[ecol[p]];
    }
  }
  return s;
}

template <class TV>
__global__ void __launch_bounds__(256) k_spmv_ell(int n, int W, const int32_t *__restrict__ ecol,
                                                  const TV *__restrict__ a, const TV *__restrict__ x,
                                                  TV *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  y[i] = ell_row<TV>(n, W, i, ecol, a, x);
}

__global__ void k_fill(int64_t n, double *__restrict__ x, double val) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = val;
}

// Velocity recovery (TDyMPFAORecoverVelocity, tdympfao_utils.c:103-217, 219-532):
// the potential of each cell is Phi = p + rho(p) |g| z_c, and every flux row
// q of face f adds -(T_row . Phi) / A_f to vel[f].  A boundary column takes
// the boundary-pressure function value (no gravity term) when one is set,
// else Phi of the face's cell (tdympfao_utils.c:314-324).  One thread per
// face sums its rows in sub-face order, so the result is deterministic.
struct VelArgs {
  int nc, nf, general, has_fn, density_type;
  int64_t nrows;
  double gabs;  // |g|
  const uint8_t *face_flags;
  const int32_t *face_verts, *vcols;                          // hexahedra: rows by vertex columns
  const double *T;                                            // hexahedra: slot-major over rows
  const int32_t *rtoff, *rinfo, *roc, *rdc, *ccell, *vdcol;   // general meshes: per-row blocks
  const double *Tgen;
  const int32_t *bnd_cell;
  const double *U, *cell_z, *pbnd_fn, *face_area;
  double *vel;
  int32_t *count;
};

__device__ inline double vel_phi_cell(const VelArgs &a, int c) {
  double rho, drho, d2rho;
  water_density(a.density_type, a.U[c], rho, drho, d2rho);
  return a.U[c] + rho * a.gabs * a.cell_z[c];
}

__device__ inline double vel_phi_bnd(const VelArgs &a, int b) {
  return a.has_fn ? a.pbnd_fn[b] : vel_phi_cell(a, a.bnd_cell[b]);
}

__global__ void __launch_bounds__(256) k_velocity(VelArgs a) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= a.nf) return;
  double s = 0.0;
  int n = 0;
  if (!(a.face_flags[f] & kFaceNoFlow)) {
    const double area = a.face_area[f];
    for (int q = 0; q < 4; q++) {
      const int64_t r = (int64_t)4 * f + q;
      double tp = 0.0;
      if (a.general) {
        const int to = a.rtoff[r];
        if (to < 0) continue;  // the fourth row of a triangle
        const int ncell = a.rinfo[r] & 0xff, ndir = (a.rinfo[r] >> 8) & 0xff;
        const int32_t *cols = a.ccell + a.roc[r];
        const int32_t *dcols = a.vdcol + a.rdc[r];
        const double *Tr = a.Tgen + to;
        for (int j = 0; j < ncell; j++) tp += Tr[j] * vel_phi_cell(a, cols[j]);
        for (int j = 0; j < ndir; j++) tp += Tr[ncell + j] * vel_phi_bnd(a, dcols[j]);
      } else {
        const int32_t *cols = a.vcols + (int64_t)a.face_verts[r] * kMaxCols;
        for (int k = 0; k < kMaxCols; k++) {
          const int col = cols[k];
          if (col < 0) continue;
          const double phi = col < a.nc ? vel_phi_cell(a, col) : vel_phi_bnd(a, col - a.nc);
          tp += a.T[(int64_t)k * a.nrows + r] * phi;
        }
      }
      s += -tp / area;
      n++;
    }
  }
  a.vel[f] = s;
  if (a.count) a.count[f] = n;
}

}  // namespace tdyhip

