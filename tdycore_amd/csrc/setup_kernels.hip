// Transmissibility setup on the device (subsystem (a)).
//
// Every setup kernel and helper below is compiled without floating-point
// contraction (no fused multiply-adds) and performs the oracle's operations
// in the oracle's order (multipliers as a * (1 / pivot), the sums of C X in
// column order, the sub-cell volume as |det| / 6 * 6), so the results agree
// bit for bit, also on high-contrast meshes whose local systems amplify any
// rounding difference by their condition number.
//
//  k_gmatrix        one thread per subcell: the 3x3 sub-cell G matrix
//                   (ComputeGMatrix_MPFAO / _TPF, src/fv/mpfao/tdympfao.c:39-126, 128-274)
//  k_local_systems  one 16-lane team per vertex: assemble the interaction-region system
//                   A x = B [p_cells; p_dir] from the G rows, LU with partial pivoting,
//                   T = C A^-1 B - F (+ C_dir)   (tdympfao.c:1126-1513, 1515-1756)
//  k_gravdis        one thread per face: A_sub (n . K_face u)(|g| u_z)
//                   (ComputeGravityDiscretization, tdympfao.c:2169-2274)
#include "tdyhip_internal.h"

namespace tdyhip {

__device__ inline void cross3(const double *a, const double *b, double *c) {
#pragma clang fp contract(off)
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

__global__ void k_gmatrix(int nc, int gmatrix, const int32_t *__restrict__ sc_faces,
                          const int32_t *__restrict__ face_cells, const double *__restrict__ cell_cen,
                          const double *__restrict__ face_cen, const double *__restrict__ face_area,
                          const double *__restrict__ face_normal, const double *__restrict__ K0,
                          const uint8_t *__restrict__ face_nv, double *__restrict__ Gsub) {
#pragma clang fp contract(off)
  const int sc = blockIdx.x * blockDim.x + threadIdx.x;
  if (sc >= nc * 8) return;
  const int c = sc >> 3;
  int fid[3];
  for (int k = 0; k < 3; k++) fid[k] = sc_faces[sc * 3 + k];
  if (fid[0] < 0) return;  // absent corner of a prism
  double cc[3], Kc[9];
  for (int d = 0; d < 3; d++) cc[d] = cell_cen[3 * c + d];
  for (int i = 0; i < 9; i++) Kc[i] = K0[c * 9 + i];
  double e[3][3];
  for (int k = 0; k < 3; k++)
    for (int d = 0; d < 3; d++) e[k][d] = face_cen[3 * fid[k] + d] - cc[d];
  double x[3];
  cross3(e[0], e[1], x);
  // six times the volume of the tetrahedron (cell centroid, three face centroids)
  const double Tsub = fabs(x[0] * e[2][0] + x[1] * e[2][1] + x[2] * e[2][2]) / 6.0 * 6.0;
  double nu[3][3];
  for (int k = 0; k < 3; k++) cross3(e[(k + 1) % 3], e[(k + 2) % 3], nu[k]);
  double *G = Gsub + (size_t)sc * 9;
  for (int i = 0; i < 3; i++) {
    const int f = fid[i];
    // sub-face = face area over the face's vertex count (tdymeshplex.c:2937-2943)
    const double area = face_area[f] / (face_nv ? (double)face_nv[f] : 4.0);
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
#pragma clang fp contract(off)
  const double *G = Gsub + (size_t)d.subcell[i] * 9 + l * 3;
  for (int q = 0; q < nU + d.ndir; q++) out[q] = 0.0;
  for (int m = 0; m < 3; m++) {
    const int km = d.sck[i][m];
    const int col = (d.role[km] == 2) ? nU + d.fidx[km] : d.fidx[km];
    out[col] = -G[m];
  }
  return -G[0] - G[1] - G[2];
}

// One interaction region per 16-lane team (4 per wavefront, 16 per block).
// Lane k owns the equation of the vertex's face k (a row of A and of the
// right-hand side B, in registers); partial-pivoting elimination picks the
// pivot lane by a team max-reduction and broadcasts the pivot row with team
// shuffles, so rows are never moved.  Back substitution leaves X = A^-1 B in
// LDS, and each face lane forms its flux row T = C X - F (+ C_dir) from the
// three nonzeros of its C row.
constexpr int kTeam = 16;
constexpr int kTeamsPerBlock = 16;

__device__ inline double team_bcast(double x, int src) { return __shfl(x, src, kTeam); }

__global__ void __launch_bounds__(kTeam * kTeamsPerBlock)
    k_local_systems(int nv, const IRDesc *__restrict__ irs, const double *__restrict__ Gsub, double *__restrict__ T,
                    int64_t nrows, int *__restrict__ status) {
#pragma clang fp contract(off)
  __shared__ double Xs[kTeamsPerBlock][kMaxIRFaces][kMaxCols];
  const int team = threadIdx.x / kTeam, lane = threadIdx.x % kTeam;
  const int v = blockIdx.x * kTeamsPerBlock + team;
  if (v >= nv) return;  // whole teams leave together: shuffles stay inside a team
  const IRDesc *d = irs + v;
  const int nface = d->nface, ncell = d->ncell, ndir = d->ndir;
  const int nU = d->nint + d->nneu;
  const int nB = ncell + ndir;
  const bool has_face = lane < nface;
  const int role = has_face ? d->role[lane] : 2;
  // up- and down-side subcell C rows of this face: three (column, -G) pairs each
  int cu[3] = {-1, -1, -1}, cd[3] = {-1, -1, -1};
  double gu[3] = {0, 0, 0}, gd[3] = {0, 0, 0};
  int fup_i = 0, fdn_i = 0;
  if (has_face) {
    fup_i = d->fup_i[lane];
    const int lu = d->fup_l[lane];
    const double *G = Gsub + (size_t)d->subcell[fup_i] * 9 + lu * 3;
    for (int m = 0; m < 3; m++) {
      const int km = d->sck[fup_i][m];
      cu[m] = (d->role[km] == 2) ? nU + d->fidx[km] : d->fidx[km];
      gu[m] = G[m];
    }
    if (role == 0) {
      fdn_i = d->fdn_i[lane];
      const int ld = d->fdn_l[lane];
      const double *Gd = Gsub + (size_t)d->subcell[fdn_i] * 9 + ld * 3;
      for (int m = 0; m < 3; m++) {
        const int km = d->sck[fdn_i][m];
        cd[m] = (d->role[km] == 2) ? nU + d->fidx[km] : d->fidx[km];
        gd[m] = Gd[m];
      }
    }
  }
  const double gus = gu[0] + gu[1] + gu[2], gds = gd[0] + gd[1] + gd[2];
  // the up-side pairs in column order: C X sums its terms by column
  if (cu[0] > cu[1]) { int t = cu[0]; cu[0] = cu[1]; cu[1] = t; double w = gu[0]; gu[0] = gu[1]; gu[1] = w; }
  if (cu[1] > cu[2]) { int t = cu[1]; cu[1] = cu[2]; cu[2] = t; double w = gu[1]; gu[1] = gu[2]; gu[2] = w; }
  if (cu[0] > cu[1]) { int t = cu[0]; cu[0] = cu[1]; cu[1] = t; double w = gu[0]; gu[0] = gu[1]; gu[1] = w; }
  // equation of an internal (role 0) or no-flow (role 1) face: A row and B row
  const bool eq = has_face && role != 2;
  double a[kMaxIRFaces], b[kMaxCols];
#pragma unroll
  for (int j = 0; j < kMaxIRFaces; j++) {
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < 3; m++) {
      s += (cu[m] == j) ? gu[m] : 0.0;   // -C_up
      s -= (cd[m] == j) ? gd[m] : 0.0;   // +C_dn
    }
    a[j] = (eq && j < nU) ? s : 0.0;
  }
#pragma unroll
  for (int j = 0; j < kMaxCols; j++) {
    double s = 0.0;
    if (j < ncell) {
      s += (j == fup_i) ? gus : 0.0;                  // -F_up
      if (role == 0) s -= (j == fdn_i) ? gds : 0.0;   // +F_dn
    } else {
#pragma unroll
      for (int m = 0; m < 3; m++) {
        s -= (cu[m] == nU + j - ncell) ? gu[m] : 0.0;  // C_up,dir
        s += (cd[m] == nU + j - ncell) ? gd[m] : 0.0;  // -C_dn,dir
      }
    }
    b[j] = (eq && j < nB) ? s : 0.0;
  }
  // forward elimination with partial pivoting over the team's lanes
  int kpiv = -1;
  int pl[kMaxIRFaces];
#pragma unroll
  for (int k = 0; k < kMaxIRFaces; k++) {
    pl[k] = 0;
    if (k < nU) {
      double mag = (eq && kpiv < 0) ? fabs(a[k]) : -1.0;
      int who = lane;
#pragma unroll
      for (int off = kTeam / 2; off > 0; off >>= 1) {
        const double om = __shfl_xor(mag, off, kTeam);
        const int ow = __shfl_xor(who, off, kTeam);
        if (om > mag || (om == mag && ow < who)) { mag = om; who = ow; }
      }
      pl[k] = who;
      if (!(mag > 0.0)) {
        if (lane == 0) atomicOr(status, 1);
        return;  // uniform over the team
      }
      const double pk = team_bcast(a[k], who);
      if (lane == who) kpiv = k;
      const bool elim = eq && kpiv < 0;
      const double inv = 1.0 / pk;
      const double l = elim ? a[k] * inv : 0.0;
#pragma unroll
      for (int j = k + 1; j < kMaxIRFaces; j++) {
        const double pj = team_bcast(a[j], who);
        if (elim) a[j] -= l * pj;
      }
#pragma unroll
      for (int j = 0; j < kMaxCols; j++) {
        const double pj = team_bcast(b[j], who);
        if (elim) b[j] -= l * pj;
      }
      if (elim) a[k] = 0.0;
    }
  }
  // back substitution: unknown k from the lane that eliminated column k
#pragma unroll
  for (int k = kMaxIRFaces - 1; k >= 0; k--) {
    if (k < nU) {
      const int who = pl[k];
      double y[kMaxCols];
#pragma unroll
      for (int j = 0; j < kMaxCols; j++) {
        y[j] = team_bcast(b[j] / a[k], who);
        if (kpiv >= 0 && kpiv < k) b[j] -= a[k] * y[j];
      }
      if (lane == who) {
#pragma unroll
        for (int j = 0; j < kMaxCols; j++) Xs[team][k][j] = y[j];
      }
    }
  }
  // the team lives in one wavefront: a wavefront-scope fence orders the LDS
  // writes above before the reads below
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  // flux rows of internal and Dirichlet faces: T = C_up X - F_up (+ C_up,dir)
  if (has_face && role != 1) {
    const int64_t r = d->row[lane];
    for (int j = 0; j < nB; j++) {
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < 3; m++)
        if (cu[m] >= 0 && cu[m] < nU) s -= gu[m] * Xs[team][cu[m]][j];
      double t;
      if (j < ncell) {
        t = s + ((j == fup_i) ? gus : 0.0);
      } else {
        double cdir = 0.0;
#pragma unroll
        for (int m = 0; m < 3; m++) cdir -= (cu[m] == nU + j - ncell) ? gu[m] : 0.0;
        t = s + cdir;
      }
      T[(int64_t)j * nrows + r] = t;
    }
  }
}

__device__ inline void inv3(const double *a, double *r) {
#pragma clang fp contract(off)
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
                          const uint8_t *__restrict__ face_nv, double *__restrict__ Gout) {
#pragma clang fp contract(off)
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  double val = 0.0;
  if (!(face_flags[f] & kFaceNoFlow)) {
    const int up = face_cells[2 * f], dn = face_cells[2 * f + 1];
    double area = face_area[f] / (face_nv ? (double)face_nv[f] : 4.0);
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
  // a triangle's fourth row is unused (rows stay face*4 + local vertex)
  const int nvf = face_nv ? face_nv[f] : 4;
  for (int q = 0; q < 4; q++) Gout[(int64_t)f * 4 + q] = q < nvf ? val : 0.0;
}

}  // namespace tdyhip
