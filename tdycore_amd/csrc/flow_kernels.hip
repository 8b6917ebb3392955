// Residual and Jacobian of the MPFA-O Richards path (subsystems (b), (c)).
//
//  k_cell_state   per cell: van Genuchten / Mualem / EOS state, rho, Kr/mu and their
//                 pressure derivatives, and the accumulation term (TS or SNES form)
//                 (TDyUpdateState_Richards_MPFAO, src/fv/mpfao/tdympfao.c:2527-2594;
//                  tdympfao_ts.c:191-212; tdympfao_snes.c:19-30, 112-123)
//  k_bnd_state    per boundary face: P_bnd and Kr/mu on the boundary
//                 (tdympfao_utils.c:24-99, 690-759)
//  k_rows         per flux row (face*4 + subface): T.P from the vertex's columns,
//                 upwinded rho^2 kr/mu flux (TDyMPFAOIFunction_Vertices, tdympfao_ts.c:12-150),
//                 optionally the row's Jacobian entries (TDyMPFAOIJacobian_Vertices, 235-456)
//  k_cell_gather  per cell: R = sum over its 6 faces x 4 subfaces (+/- by orientation)
//                 + accumulation; Jacobian diagonal accumulation term
//
// Flux rows are stored face-major (row = face*4 + subface), the reference's
// Trans_mat row numbering; T is an 8-wide ELL block stored column-major so that
// consecutive lanes read consecutive addresses.
#include "tdyhip_internal.h"
#include "tdyhip_math.h"

namespace tdyhip {

struct StateParams {
  double vg_m, vg_alpha, x0, x1, dx, pref, dtime, shift;
  double c0, c1, c2, c3;
  int density_type, form;  // form: 0 TS, 1 SNES
};

__global__ void k_cell_state(int nc, StateParams sp, const double *__restrict__ U,
                             const double *__restrict__ Udot, const double *__restrict__ phi,
                             const double *__restrict__ Sr, const double *__restrict__ vol,
                             const double *__restrict__ accum_prev, double *__restrict__ rho_o,
                             double *__restrict__ ukvr_o, double *__restrict__ drho_o,
                             double *__restrict__ dukvr_o, double *__restrict__ dM_o,
                             double *__restrict__ d2M_o, double *__restrict__ acc_o) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const double p = U[c];
  const double sr = Sr[c];
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
  rho_o[c] = rho;
  ukvr_o[c] = Kr / vis;
  drho_o[c] = drho;
  dukvr_o[c] = dKr * dS / vis - Kr / (vis * vis) * dvis;
  const double ph = phi[c];
  const double dM = drho * ph * S + rho * ph * dS;
  const double d2M = 2.0 * dS * drho * ph + S * d2rho * ph + d2S * rho * ph;
  dM_o[c] = dM;
  d2M_o[c] = d2M;
  if (sp.form == 0) {
    acc_o[c] = dM * Udot[c] * vol[c];
  } else {
    acc_o[c] = rho * ph * S * vol[c] / sp.dtime - accum_prev[c];
  }
}

// rho phi S V / dt at Uprev (TDyMPFAOSNESPreSolve, tdympfao_snes.c:32-64)
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

__global__ void k_bnd_state(int nbnd, StateParams sp, const double *__restrict__ U,
                            const int32_t *__restrict__ bnd_cell, const double *__restrict__ Sr,
                            const double *__restrict__ pbnd_fn, double *__restrict__ Pbnd,
                            double *__restrict__ ukvr_bnd) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbnd) return;
  const int c = bnd_cell[b];
  const double P = pbnd_fn ? pbnd_fn[b] : U[c];
  Pbnd[b] = P;
  const double sr = Sr[c];
  double S, dS, d2S;
  vg_saturation(sp.vg_m, sp.vg_alpha, sr, sp.pref - P, S, dS, d2S);
  const double Se = (S - sr) / (1.0 - sr);
  double Kr, dKr;
  const double cc[4] = {sp.c0, sp.c1, sp.c2, sp.c3};
  mualem_kr(sp.vg_m, sp.x0, sp.x1, sp.dx, cc, Se, Kr, dKr);
  ukvr_bnd[b] = Kr / kViscosity;
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
  const uint8_t *row_ul;
  const uint8_t *jslot;
  const int32_t *jrp;  // CSR row starts, or nullptr for the slot-major ELL layout
  int64_t jstride;     // distance between consecutive slots of a row: 1 (CSR) or n (ELL)
  double *flux;
  double *J;
};

template <bool WITH_R, bool WITH_J>
__global__ void __launch_bounds__(256) k_rows(RowArgs a) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= a.nrows) return;
  const int f = (int)(r >> 2);
  const uint8_t fl = a.face_flags[f];
  if (fl & kFaceNoFlow) {
    if (WITH_R) a.flux[r] = 0.0;
    return;
  }
  const int up = a.face_cells[2 * f], dn = a.face_cells[2 * f + 1];
  const int v = a.face_verts[r];
  const int32_t *cols = a.vcols + (int64_t)v * kMaxCols;
  double Tk[kMaxCols];
  int ck[kMaxCols];
  double TP = 0.0;
#pragma unroll
  for (int k = 0; k < kMaxCols; k++) {
    ck[k] = cols[k];
    Tk[k] = a.T[(int64_t)k * a.nrows + r];
  }
#pragma unroll
  for (int k = 0; k < kMaxCols; k++) {
    const int col = ck[k];
    if (col >= 0) {
      const double P = (col < a.nc) ? a.U[col] : a.Pbnd[col - a.nc];
      TP += Tk[k] * P;
    }
  }
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
    const uint8_t ul = a.row_ul[2 * r], dl = a.row_ul[2 * r + 1];
    const uint8_t *slot = a.jslot + (int64_t)v * 64;
    const int64_t rup = up >= 0 ? (a.jrp ? a.jrp[up] : up) : 0, rdn = dn >= 0 ? (a.jrp ? a.jrp[dn] : dn) : 0;
#pragma unroll
    for (int k = 0; k < kMaxCols; k++) {
      const int col = ck[k];
      if (col < 0 || col >= a.nc) continue;
      double Jac;
      if (zero) Jac = 0.0;
      else if (up >= 0 && col == up)
        Jac = dden_up * ukvr * TP + den * dukvr_up * TP + den * ukvr * Tk[k] + 2 * den * dden_up * ukvr * G +
              (den * den) * dukvr_up * G;
      else if (dn >= 0 && col == dn)
        Jac = dden_dn * ukvr * TP + den * dukvr_dn * TP + den * ukvr * Tk[k] + 2 * den * dden_dn * ukvr * G +
              (den * den) * dukvr_dn * G;
      else
        Jac = den * ukvr * Tk[k];
      Jac = -Jac;
      if (up >= 0) unsafeAtomicAdd(a.J + rup + (int64_t)slot[ul * 8 + k] * a.jstride, Jac);
      if (dn >= 0) unsafeAtomicAdd(a.J + rdn + (int64_t)slot[dl * 8 + k] * a.jstride, -Jac);
    }
  }
}

template <bool WITH_R, bool WITH_J>
__global__ void k_cell_gather(int nc, int form, double shift, double dtime, const int32_t *__restrict__ cell_faces,
                              const int32_t *__restrict__ face_cells, const double *__restrict__ flux,
                              const double *__restrict__ acc, const double *__restrict__ dM,
                              const double *__restrict__ d2M, const double *__restrict__ Udot,
                              const double *__restrict__ vol, const int32_t *__restrict__ jrp,
                              const uint8_t *__restrict__ jdslot, int64_t jstride, double *__restrict__ R,
                              double *__restrict__ J) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  if (WITH_R) {
    double s = 0.0;
#pragma unroll
    for (int lf = 0; lf < 6; lf++) {
      const int f = cell_faces[c * 6 + lf];
      const double4 q = *reinterpret_cast<const double4 *>(flux + (int64_t)f * 4);
      const double fs = ((q.x + q.y) + q.z) + q.w;
      s += (face_cells[2 * f] == c) ? fs : -fs;
    }
    R[c] = s + acc[c];
  }
  if (WITH_J) {
    const double v = (form == 0) ? (shift * dM[c] + d2M[c] * Udot[c]) * vol[c] : dM[c] * vol[c] * (1.0 / dtime);
    J[(jrp ? jrp[c] : (int64_t)c) + (int64_t)jdslot[c] * jstride] += v;
  }
}

// ELL SpMV over the slot-major layout: slot s of row i at s*n + i.
template <class TV>
__global__ void __launch_bounds__(256) k_spmv_ell(int n, int W, const int32_t *__restrict__ ecol,
                                                  const TV *__restrict__ a, const TV *__restrict__ x,
                                                  TV *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  TV s = 0;
  for (int k = 0; k < W; k++) {
    const int64_t p = (int64_t)k * n + i;
    s += a[p] * x[ecol[p]];
  }
  y[i] = s;
}

__global__ void k_fill(int64_t n, double *__restrict__ x, double val) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = val;
}

}  // namespace tdyhip
