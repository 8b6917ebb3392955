// Transmissibility setup on the device (subsystem (a)).
//
//  k_gmatrix        one thread per subcell: the 3x3 sub-cell G matrix
//                   (ComputeGMatrix_MPFAO / _TPF, src/fv/mpfao/tdympfao.c:39-126, 128-274)
//  k_local_systems  one thread per vertex: assemble the interaction-region system
//                   A x = B [p_cells; p_dir] from the G rows, LU with partial pivoting,
//                   T = C A^-1 B - F (+ C_dir)   (tdympfao.c:1126-1513, 1515-1756)
//  k_gravdis        one thread per face: A_sub (n . K_face u)(|g| u_z)
//                   (ComputeGravityDiscretization, tdympfao.c:2169-2274)
#include "tdyhip_internal.h"

namespace tdyhip {

__device__ inline void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

__global__ void k_gmatrix(int nc, int gmatrix, const int32_t *__restrict__ sc_faces,
                          const int32_t *__restrict__ face_cells, const double *__restrict__ cell_cen,
                          const double *__restrict__ face_cen, const double *__restrict__ face_area,
                          const double *__restrict__ face_normal, const double *__restrict__ K0,
                          double *__restrict__ Gsub) {
  const int sc = blockIdx.x * blockDim.x + threadIdx.x;
  if (sc >= nc * 8) return;
  const int c = sc >> 3;
  int fid[3];
  for (int k = 0; k < 3; k++) fid[k] = sc_faces[sc * 3 + k];
  double cc[3], Kc[9];
  for (int d = 0; d < 3; d++) cc[d] = cell_cen[3 * c + d];
  for (int i = 0; i < 9; i++) Kc[i] = K0[c * 9 + i];
  double e[3][3];
  for (int k = 0; k < 3; k++)
    for (int d = 0; d < 3; d++) e[k][d] = face_cen[3 * fid[k] + d] - cc[d];
  double x[3];
  cross3(e[0], e[1], x);
  const double Tsub = fabs(x[0] * e[2][0] + x[1] * e[2][1] + x[2] * e[2][2]);
  double nu[3][3];
  for (int k = 0; k < 3; k++) cross3(e[(k + 1) % 3], e[(k + 2) % 3], nu[k]);
  double *G = Gsub + (size_t)sc * 9;
  for (int i = 0; i < 3; i++) {
    const int f = fid[i];
    const double area = face_area[f] / 4.0;
    const double n0 = face_normal[3 * f], n1 = face_normal[3 * f + 1], n2 = face_normal[3 * f + 2];
    if (gmatrix == 0) {
      for (int j = 0; j < 3; j++) {
        const double kv0 = Kc[0] * nu[j][0] + Kc[1] * nu[j][1] + Kc[2] * nu[j][2];
        const double kv1 = Kc[3] * nu[j][0] + Kc[4] * nu[j][1] + Kc[5] * nu[j][2];
        const double kv2 = Kc[6] * nu[j][0] + Kc[7] * nu[j][1] + Kc[8] * nu[j][2];
        double g = n0 * kv0 + n1 * kv1 + n2 * kv2;
        G[i * 3 + j] = g * (1.0 / Tsub * area);
      }
    } else {
      const int up = face_cells[2 * f], dn = face_cells[2 * f + 1];
      const int nb = (up == c) ? dn : up;
      double ncen[3], Kn[9], dist = 0.0;
      if (nb >= 0) {
        for (int d = 0; d < 3; d++) ncen[d] = cell_cen[3 * nb + d];
        for (int k = 0; k < 9; k++) Kn[k] = K0[nb * 9 + k];
      } else {
        for (int d = 0; d < 3; d++) ncen[d] = face_cen[3 * f + d];
        for (int k = 0; k < 9; k++) Kn[k] = Kc[k];
      }
      double u[3], un = 0.0;
      for (int d = 0; d < 3; d++) { u[d] = cc[d] - ncen[d]; dist += u[d] * u[d]; }
      dist = sqrt(dist);
      un = dist;
      for (int d = 0; d < 3; d++) u[d] /= un;
      if (nb >= 0) dist *= 0.5;
      const double dp = n0 * u[0] + n1 * u[1] + n2 * u[2];
      double kn = 0.0, kv = 0.0;
      for (int k = 0; k < 3; k++) { kn += u[k] * u[k] / Kn[k * 4]; kv += u[k] * u[k] / Kc[k * 4]; }
      const double kav = 0.5 * (1.0 / kv) + 0.5 * (1.0 / kn);
      for (int j = 0; j < 3; j++) G[i * 3 + j] = (i == j) ? area * dp * kav / dist : 0.0;
    }
  }
}

// C row of subcell i, local face l over the columns [x (nU) | Dirichlet (ndir)],
// and F (the coefficient of the subcell's own cell) = -sum_m G[l][m]
// (ComputeCandFmatrix, tdympfao.c:1000-1076).
__device__ inline double crow(const IRDesc &d, const double *__restrict__ Gsub, int nU, int i, int l,
                              double *out) {
  const double *G = Gsub + (size_t)d.subcell[i] * 9 + l * 3;
  for (int q = 0; q < nU + d.ndir; q++) out[q] = 0.0;
  for (int m = 0; m < 3; m++) {
    const int km = d.sck[i][m];
    const int col = (d.role[km] == 2) ? nU + d.fidx[km] : d.fidx[km];
    out[col] = -G[m];
  }
  return -G[0] - G[1] - G[2];
}

__global__ void __launch_bounds__(64) k_local_systems(int nv, const IRDesc *__restrict__ irs,
                                                      const double *__restrict__ Gsub, double *__restrict__ T,
                                                      int64_t nrows, int *__restrict__ status) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const IRDesc d = irs[v];
  const int nU = d.nint + d.nneu, ncell = d.ncell, ndir = d.ndir, nB = ncell + ndir;
  double A[kMaxIRFaces * kMaxIRFaces];
  double X[kMaxIRFaces * kMaxCols];
  double Cu[kMaxIRFaces + kMaxCols], Cd[kMaxIRFaces + kMaxCols];
  int piv[kMaxIRFaces];
  if (nU > 0) {
    for (int q = 0; q < nU * nU; q++) A[q] = 0.0;
    for (int q = 0; q < nU * nB; q++) X[q] = 0.0;
    for (int k = 0; k < d.nface; k++) {
      if (d.role[k] == 0) {
        const int row = d.fidx[k];
        const double Fu = crow(d, Gsub, nU, d.fup_i[k], d.fup_l[k], Cu);
        const double Fd = crow(d, Gsub, nU, d.fdn_i[k], d.fdn_l[k], Cd);
        for (int j = 0; j < nU; j++) A[row + j * nU] = -Cu[j] + Cd[j];
        X[row + d.fup_i[k] * nU] += -Fu;
        X[row + d.fdn_i[k] * nU] += Fd;
        for (int j = 0; j < ndir; j++) X[row + (ncell + j) * nU] = Cu[nU + j] - Cd[nU + j];
      } else if (d.role[k] == 1) {
        const int row = d.fidx[k];
        const double Fu = crow(d, Gsub, nU, d.fup_i[k], d.fup_l[k], Cu);
        for (int j = 0; j < nU; j++) A[row + j * nU] = -Cu[j];
        X[row + d.fup_i[k] * nU] += -Fu;
        for (int j = 0; j < ndir; j++) X[row + (ncell + j) * nU] = Cu[nU + j];
      }
    }
    // LU with partial pivoting (getrf), then X <- A^-1 X (getrs)
    for (int k = 0; k < nU; k++) {
      int p = k;
      double mx = fabs(A[k + k * nU]);
      for (int i = k + 1; i < nU; i++) {
        const double a = fabs(A[i + k * nU]);
        if (a > mx) { mx = a; p = i; }
      }
      piv[k] = p;
      if (mx == 0.0) { atomicOr(status, 1); return; }
      if (p != k)
        for (int j = 0; j < nU; j++) { const double t = A[k + j * nU]; A[k + j * nU] = A[p + j * nU]; A[p + j * nU] = t; }
      const double inv = 1.0 / A[k + k * nU];
      for (int i = k + 1; i < nU; i++) A[i + k * nU] *= inv;
      for (int j = k + 1; j < nU; j++) {
        const double akj = A[k + j * nU];
        for (int i = k + 1; i < nU; i++) A[i + j * nU] -= A[i + k * nU] * akj;
      }
    }
    for (int r = 0; r < nB; r++) {
      double *b = X + r * nU;
      for (int k = 0; k < nU; k++)
        if (piv[k] != k) { const double t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
      for (int k = 0; k < nU; k++) { const double s = b[k]; for (int i = k + 1; i < nU; i++) b[i] -= A[i + k * nU] * s; }
      for (int k = nU - 1; k >= 0; k--) {
        b[k] /= A[k + k * nU];
        const double s = b[k];
        for (int i = 0; i < k; i++) b[i] -= A[i + k * nU] * s;
      }
    }
  }
  // flux rows: internal faces and Dirichlet faces, flux from the up-side subcell
  for (int k = 0; k < d.nface; k++) {
    if (d.role[k] == 1) continue;
    const double F = crow(d, Gsub, nU, d.fup_i[k], d.fup_l[k], Cu);
    const int64_t r = d.row[k];
    for (int j = 0; j < nB; j++) {
      double s = 0.0;
      for (int q = 0; q < nU; q++) s += Cu[q] * X[q + j * nU];
      const double t = (j < ncell) ? s + ((j == d.fup_i[k]) ? -F : 0.0) : s + Cu[nU + (j - ncell)];
      T[(int64_t)j * nrows + r] = t;
    }
  }
}

__device__ inline void inv3(const double *a, double *r) {
  const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                     a[2] * (a[3] * a[7] - a[4] * a[6]);
  r[0] = +1.0 / det * (a[4] * a[8] - a[7] * a[5]);
  r[1] = -1.0 / det * (a[1] * a[8] - a[7] * a[2]);
  r[2] = +1.0 / det * (a[1] * a[5] - a[4] * a[2]);
  r[3] = -1.0 / det * (a[3] * a[8] - a[6] * a[5]);
  r[4] = +1.0 / det * (a[0] * a[8] - a[6] * a[2]);
  r[5] = -1.0 / det * (a[0] * a[5] - a[3] * a[2]);
  r[6] = +1.0 / det * (a[3] * a[7] - a[6] * a[4]);
  r[7] = -1.0 / det * (a[0] * a[7] - a[6] * a[1]);
  r[8] = +1.0 / det * (a[0] * a[4] - a[3] * a[1]);
}

__global__ void k_gravdis(int nf, int gmatrix, double gravity, const int32_t *__restrict__ face_cells,
                          const uint8_t *__restrict__ face_flags, const double *__restrict__ cell_cen,
                          const double *__restrict__ face_cen, const double *__restrict__ face_area,
                          const double *__restrict__ face_normal, const double *__restrict__ K0,
                          double *__restrict__ Gout) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  double val = 0.0;
  if (!(face_flags[f] & kFaceNoFlow)) {
    const int up = face_cells[2 * f], dn = face_cells[2 * f + 1];
    double area = face_area[f] / 4.0;
    double fc[3], cu[3], cd[3], n[3], u[3];
    for (int d = 0; d < 3; d++) {
      fc[d] = face_cen[3 * f + d];
      n[d] = face_normal[3 * f + d];
      cu[d] = up >= 0 ? cell_cen[3 * up + d] : fc[d];
      cd[d] = dn >= 0 ? cell_cen[3 * dn + d] : fc[d];
    }
    double ul = 0.0;
    for (int d = 0; d < 3; d++) { u[d] = cd[d] - cu[d]; ul += u[d] * u[d]; }
    ul = sqrt(ul);
    for (int d = 0; d < 3; d++) u[d] /= ul;
    double dup = 0.0, ddn = 0.0;
    if (up >= 0) { for (int d = 0; d < 3; d++) dup += (cu[d] - fc[d]) * (cu[d] - fc[d]); dup = sqrt(dup); }
    if (dn >= 0) { for (int d = 0; d < 3; d++) ddn += (cd[d] - fc[d]) * (cd[d] - fc[d]); ddn = sqrt(ddn); }
    const int cup = up >= 0 ? up : dn, cdn = dn >= 0 ? dn : up;
    double Kup[9], Kdn[9];
    for (int kk = 0; kk < 3; kk++)
      for (int mm = 0; mm < 3; mm++) {
        Kup[mm * 3 + kk] = K0[cup * 9 + kk * 3 + mm];
        Kdn[mm * 3 + kk] = K0[cdn * 9 + kk * 3 + mm];
      }
    const double wt = dup / (dup + ddn);
    double kval;
    if (gmatrix == 0) {
      double Ki[9], Kj[9], Kfi[9], Kf[9];
      inv3(Kup, Ki);
      inv3(Kdn, Kj);
      for (int i = 0; i < 9; i++) Kfi[i] = wt * Ki[i] + (1.0 - wt) * Kj[i];
      inv3(Kfi, Kf);
      for (int i = 0; i < 9; i++)
        if (Kf[i] < 0.0 || fabs(Kf[i]) < 2.220446049250313e-16) Kf[i] = 0.0;
      kval = 0.0;
      for (int i = 0; i < 3; i++) kval += n[i] * (Kf[i * 3] * u[0] + Kf[i * 3 + 1] * u[1] + Kf[i * 3 + 2] * u[2]);
    } else {
      double ku = 0.0, kd = 0.0;
      for (int k = 0; k < 3; k++) { ku += u[k] * u[k] / Kup[k * 4]; kd += u[k] * u[k] / Kdn[k * 4]; }
      ku = 1.0 / ku; kd = 1.0 / kd;
      kval = (ku * kd) / (wt * kd + (1.0 - wt) * ku);
      area *= n[0] * u[0] + n[1] * u[1] + n[2] * u[2];
    }
    const double gz = fabs(gravity) * u[2];
    val = area * kval * gz;
  }
  for (int q = 0; q < 4; q++) Gout[(int64_t)f * 4 + q] = val;
}

}  // namespace tdyhip
