/*
 * ORACLE -- test infrastructure only.  FP64 CPU restatement of TDycore's
 * MPFA-O Richards path.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the shipped library never does.
 *
 * What each part restates (reference file:line):
 *   subcell faces, nu vectors, T      src/fv/share/tdymeshplex.c:2829-3000
 *   G matrix (MPFA-O / TPF)           src/fv/mpfao/tdympfao.c:13-126, 128-274
 *   vertex local systems, Trans       src/fv/mpfao/tdympfao.c:1000-1513, 1515-1756, 1801-1860
 *   gravity discretisation            src/fv/mpfao/tdympfao.c:1869-2274, tdympfao_utils.c:11-21
 *   boundary rho/mu frozen at Pref    src/fv/mpfao/tdympfao.c:297-336
 *   state update                      src/fv/mpfao/tdympfao.c:2527-2594
 *   van Genuchten / Mualem (+cubic)   src/materials/tdycharacteristiccurves.c:468-717
 *   EOS                               src/materials/tdyeos.c:5-37, 87-98
 *   boundary pressure / state         src/fv/mpfao/tdympfao_utils.c:24-99, 690-759
 *   TS residual / Jacobian            src/fv/mpfao/tdympfao_ts.c:12-561
 *   SNES residual / Jacobian          src/fv/mpfao/tdympfao_snes.c:19-203
 *   Jacobian pattern (vertex adj.)    src/tdydiscretization.c:156-181, tdympfao.c:808
 *   time loop                         src/tdyti.c:71-196
 *   Newton + line search + tests      PETSc 3.17 SNESNEWTONLS / SNESLINESEARCHBT /
 *                                     SNESConvergedDefault (not in the reference tree)
 *
 * Parity pins (tests/test_oracle_*.py):
 *   - reference golds: transient-mpfaof90{,-neumann,-tpf-neumann},
 *     transient-snes-mpfaof90{,-dirichlet,-seepage,-tdydriver},
 *     richards-driver-snes-prob1 (1e-12 relative on pressure);
 *   - closed forms: orthogonal-grid G = diag(K A/(h/2)); MPFA-O == TPF on
 *     K-orthogonal grids; Trans row sums vanish; linear pressure is
 *     reproduced exactly on parallelepiped grids; hydrostatic zero flux;
 *     VG/Mualem/EOS against closed forms and finite differences; the Mualem
 *     cubic's four end conditions;
 *   - Jacobian against central finite differences of the residual where the
 *     reference Jacobian is exact.
 * Per-face boundary types generalise the reference's single global bc_type;
 * with one type on every face they reduce to it exactly.  Where a vertex has
 * both Dirichlet and Neumann faces (never built by the reference) the
 * Neumann rows use +C_s,dir in B (the reference's tdympfao.c:1278-1285 has
 * -C_32 there, a block it never fills) and Dirichlet flux rows include the
 * Neumann interface unknowns.
 */
#include "tdy_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MACH_EPS 2.220446049250313e-16

static char g_err[512];
const char *ora_last_error(void) { return g_err; }
static void set_err(const char *m) { snprintf(g_err, sizeof g_err, "%s", m); }

typedef struct {
  int nrow, ncol; /* T is nrow x ncol (row-major) */
  int ncell;      /* first ncell columns are IR cells */
  int off_T;      /* offset into Tvals */
  int off_cell;   /* offset into ir_cells */
  int off_face;   /* offset into ir_faces */
  int nface;      /* faces at vertex */
  int nint;       /* internal faces */
  int off_bcol;   /* offset into ir_bcol (dir columns: boundary index) */
  int nbcol;
} vtx_ir;

struct ora_ctx {
  int nc, nf, nv, nbnd;
  int *cell_verts, *cell_faces, *face_verts, *face_cells, *face_bc;
  double *vxyz, *ccen, *cvol, *fcen, *farea, *fnorm;
  double *K0, *phi, *Sr;
  double vg_m, vg_alpha, x0, x1, x2, dx, gravity, pref;
  int density_type, gmatrix;
  double coeffs[4]; /* Mualem cubic */
  double *p_bnd_fn;
  int newton_maxit; /* SNES max_it (PETSc default 50) */

  /* IR */
  vtx_ir *ir;
  int *ir_cells, *ir_subc; /* cell id, subcell (local vertex) idx */
  int *ir_faces, *ir_subface;
  int *ir_bcol;
  int *ir_rowface; /* per T row: face id */
  double *Tvals;
  int *sc_faces; /* nc*8*3 */
  double *G;     /* nc*8*9 */

  /* Trans matrix rows 4*nf, cols nc+nbnd */
  int *tr_rowptr, *tr_cols;
  double *tr_vals;
  double *gravdis; /* 4*nf */

  /* jacobian pattern */
  int *j_rowptr, *j_cols;

  /* state */
  double *S, *dS, *d2S, *Kr, *dKr, *rho, *drho, *d2rho, *vis, *dvis;
  double *P_bnd, *Kr_bnd, *rho_bnd, *vis_bnd, *S_bnd;
  int *bnd_cell;
  double *Pvec, *TP;
  double *accum_prev;
  double *source; /* per-cell source_sink (tdympfao.c:364-384; zero unless set) */
  double dtime;

  /* threads of every parallel loop (default 1).  The residual and Jacobian
   * are formed per flux row and then summed per cell in a fixed order, so
   * their values do not depend on the thread count. */
  int nthreads;
  int *ir_rowstart;  /* nv+1: first T row of each vertex (index into ir_rowface) */
  int *row_T;        /* 4nf: offset of flux row face*4+q's T row in Tvals, -1: no T row */
  int *row_v;        /* 4nf: vertex of the flux row */
  double *flux_row;  /* 4nf fluxes of the last residual */
  double *D;         /* Tvals-sized: d(flux row)/d(cell column) of the last Jacobian */
};

/* ------------------------------------------------------------------ */
/* characteristic curves and EOS                                       */
/* ------------------------------------------------------------------ */

/* PressureSaturation_VanGenuchten, src/materials/tdycharacteristiccurves.c:692-717 */
static void vg_sat(double m, double alpha, double Sr, double Pc, double *S,
                   double *dS_dP, double *d2S_dP2) {
  if (Pc <= 0) {
    *S = 1; *dS_dP = 0; *d2S_dP2 = 0.0;
  } else {
    double n = 1 / (1 - m);
    double pc_alpha = Pc * alpha;
    double pc_alpha_n = pow(pc_alpha, n);
    double one_plus = 1 + pc_alpha_n;
    double Se = pow(one_plus, -m);
    *S = (1.0 - Sr) * Se + Sr;
    double dSe_dPc = -m * n * alpha * pc_alpha_n / (pc_alpha * pow(one_plus, m + 1));
    *dS_dP = -dSe_dPc * (1.0 - Sr);
    double d2Se = m * n * (pc_alpha_n) * pow(one_plus, -m - 2.0) *
                  ((m * n + 1.0) * pc_alpha_n - n + 1) / pow(Pc, 2.0);
    *d2S_dP2 = (1.0 - Sr) * d2Se;
  }
}

/* RelativePermeability_Mualem_Unsmoothed, tdycharacteristiccurves.c:468-485 */
static void mualem_unsmoothed(double m, double Se, double *Kr, double *dKr) {
  *Kr = 1.0; *dKr = 0.0;
  if (Se >= 1.0) return;
  double s1m = pow(Se, 1 / m);
  double tmp = pow(1 - s1m, m);
  *Kr = sqrt(Se);
  *Kr *= (1 - tmp) * (1 - tmp);
  *dKr = 0.5 * (*Kr) / Se;
  *dKr += 2 * pow(Se, 1 / m - 0.5) * pow(1 - s1m, m - 1) * (1 - pow(1 - s1m, m));
}

/* dgesv-equivalent 4x4 solve with partial pivoting (column-major A) */
static int solve4(double *A, double *b) {
  int n = 4;
  for (int k = 0; k < n; k++) {
    int p = k; double mx = fabs(A[k + k * n]);
    for (int i = k + 1; i < n; i++) if (fabs(A[i + k * n]) > mx) { mx = fabs(A[i + k * n]); p = i; }
    if (mx == 0.0) return 1;
    if (p != k) {
      for (int j = 0; j < n; j++) { double t = A[k + j * n]; A[k + j * n] = A[p + j * n]; A[p + j * n] = t; }
      double t = b[k]; b[k] = b[p]; b[p] = t;
    }
    for (int i = k + 1; i < n; i++) {
      double l = A[i + k * n] / A[k + k * n];
      A[i + k * n] = l;
      for (int j = k + 1; j < n; j++) A[i + j * n] -= l * A[k + j * n];
      b[i] -= l * b[k];
    }
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = b[i];
    for (int j = i + 1; j < n; j++) s -= A[i + j * n] * b[j];
    b[i] = s / A[i + i * n];
  }
  return 0;
}

/* CubicPolynomialSetup + RelativePermeability_Mualem_GetSmoothingCoeffs,
 * tdycharacteristiccurves.c:509-596 */
static int mualem_coeffs(double m, double x0, double x1, double x2, double dx, double *c) {
  double Kr, dKr;
  mualem_unsmoothed(m, x1, &Kr, &dKr);
  c[0] = 1.0; c[1] = Kr; c[2] = 0.0; c[3] = dKr;
  double xt1 = (x1 - x0) / dx, xt2 = (x2 - x0) / dx;
  double A[16];
  A[0] = 1.0; A[1] = 1.0; A[2] = 0.0; A[3] = 0.0;
  A[4] = xt2; A[5] = xt1; A[6] = 1.0 / dx; A[7] = 1.0 / dx;
  A[8] = xt2 * xt2; A[9] = xt1 * xt1; A[10] = 2.0 * xt2 / dx; A[11] = 2.0 * xt1 / dx;
  A[12] = xt2 * xt2 * xt2; A[13] = xt1 * xt1 * xt1; A[14] = 3.0 * xt2 * xt2 / dx; A[15] = 3.0 * xt1 * xt1 / dx;
  return solve4(A, c);
}

/* RelativePermeability_Mualem, tdycharacteristiccurves.c:607-614; cubic in
 * CubicPolynomialEvaluate 551-562 */
static void mualem(const ora_ctx *ctx, double Se, double *Kr, double *dKr) {
  if (Se > ctx->x1) {
    double xt = (Se - ctx->x0) / ctx->dx;
    const double *c = ctx->coeffs;
    *Kr = c[0] + xt * (c[1] + xt * (c[2] + xt * c[3]));
    *dKr = (c[1] + xt * (2.0 * c[2] + xt * 3.0 * c[3])) / ctx->dx;
  } else {
    mualem_unsmoothed(ctx->vg_m, Se, Kr, dKr);
  }
}

/* ComputeWaterDensity_Constant / _Exponential, src/materials/tdyeos.c:5-37 */
static void density(int type, double p, double *rho, double *drho, double *d2rho) {
  if (type == 0) {
    *rho = 998.0; *drho = 0.0; *d2rho = 0.0;
  } else {
    const double rho0 = 997.16e0, p0 = 101325.e0, beta = 1.e-8;
    if (p < p0) { *rho = rho0; *drho = 0.0; *d2rho = 0.0; }
    else { *rho = rho0 * exp(beta * (p - p0)); *drho = beta * (*rho); *d2rho = beta * (*drho); }
  }
}
/* ComputeWaterViscosity_Constant, src/materials/tdyeos.c:87-98 */
static void viscosity(double p, double *vis, double *dvis) { (void)p; *vis = 9.94e-4; *dvis = 0.0; }

/* ------------------------------------------------------------------ */
/* small vector helpers (src/tdyutils.c:239-470, 683-718)             */
/* ------------------------------------------------------------------ */
static void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double *a, const double *b) {
  double s = 0.0;
  for (int d = 0; d < 3; d++) s += a[d] * b[d];
  return s;
}
static double length3(const double *a, const double *b) {
  double l = 0.0;
  for (int d = 0; d < 3; d++) l += pow(a[d] - b[d], 2.0);
  return pow(l, 0.5);
}
static void unit_from_to(const double *from, const double *to, double *u) {
  double n = 0.0;
  for (int d = 0; d < 3; d++) { u[d] = to[d] - from[d]; n += u[d] * u[d]; }
  n = sqrt(n);
  for (int d = 0; d < 3; d++) u[d] = u[d] / n;
}
static void inv3(const double *a, double *r) {
  double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
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

/* ------------------------------------------------------------------ */
/* dense LU (LAPACK getrf/getrs semantics, column-major)              */
/* ------------------------------------------------------------------ */
static int lu_factor(int n, double *A, int *piv) {
  for (int k = 0; k < n; k++) {
    int p = k; double mx = fabs(A[k + (size_t)k * n]);
    for (int i = k + 1; i < n; i++) {
      double v = fabs(A[i + (size_t)k * n]);
      if (v > mx) { mx = v; p = i; }
    }
    piv[k] = p;
    if (mx == 0.0) return k + 1;
    if (p != k)
      for (int j = 0; j < n; j++) {
        double t = A[k + (size_t)j * n]; A[k + (size_t)j * n] = A[p + (size_t)j * n]; A[p + (size_t)j * n] = t;
      }
    double inv = 1.0 / A[k + (size_t)k * n];
    for (int i = k + 1; i < n; i++) A[i + (size_t)k * n] *= inv;
    for (int j = k + 1; j < n; j++) {
      double akj = A[k + (size_t)j * n];
      if (akj == 0.0) continue;
      for (int i = k + 1; i < n; i++) A[i + (size_t)j * n] -= A[i + (size_t)k * n] * akj;
    }
  }
  return 0;
}
static void lu_solve(int n, const double *A, const int *piv, int nrhs, double *B) {
  for (int r = 0; r < nrhs; r++) {
    double *b = B + (size_t)r * n;
    for (int k = 0; k < n; k++) if (piv[k] != k) { double t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
    for (int k = 0; k < n; k++) { double s = b[k]; for (int i = k + 1; i < n; i++) b[i] -= A[i + (size_t)k * n] * s; }
    for (int k = n - 1; k >= 0; k--) {
      b[k] /= A[k + (size_t)k * n];
      double s = b[k];
      for (int i = 0; i < k; i++) b[i] -= A[i + (size_t)k * n] * s;
    }
  }
}

/* ------------------------------------------------------------------ */
/* setup                                                                */
/* ------------------------------------------------------------------ */
static int *dupi(const int *a, size_t n) { int *r = malloc(n * sizeof(int)); memcpy(r, a, n * sizeof(int)); return r; }
static double *dupd(const double *a, size_t n) { double *r = malloc(n * sizeof(double)); memcpy(r, a, n * sizeof(double)); return r; }

static int face_has_vertex(const ora_ctx *ctx, int f, int v) {
  for (int q = 0; q < 4; q++) if (ctx->face_verts[f * 4 + q] == v) return q;
  return -1;
}
/* Vertices of face f: 4 for a quadrilateral, 3 for a triangle (face_verts
 * padded with -1).  A sub-face is the face area over this count
 * (tdymeshplex.c:2937-2943). */
static int face_nverts(const ora_ctx *ctx, int f) {
  int n = 0;
  for (int q = 0; q < 4; q++) n += ctx->face_verts[f * 4 + q] >= 0;
  return n;
}
static int is_bnd_face(const ora_ctx *ctx, int f) { return ctx->face_cells[2 * f] < 0 || ctx->face_cells[2 * f + 1] < 0; }
static int face_bc_of(const ora_ctx *ctx, int f) { return ctx->face_bc[f]; }
static int is_dirichlet_like(int bc) { return bc == 0 || bc == 2; }

/* Subcell faces ordered so that the (f0,f1,f2) centroid triangle normal
 * points toward the cell centroid (FindFaceIDsOfACellCommonToAVertex and
 * UpdateFaceOrientationAroundAVertex, tdymeshplex.c:236-275, 1384-1412),
 * nu vectors and T = 6 * tet volume (SetupSubcells, tdymeshplex.c:2883-2966),
 * and G (ComputeGMatrix_MPFAO tdympfao.c:39-126 / _TPF 128-274). */
static int build_subcells(ora_ctx *ctx) {
  int nc = ctx->nc;
  ctx->sc_faces = malloc((size_t)nc * 8 * 3 * sizeof(int));
  ctx->G = calloc((size_t)nc * 8 * 9, sizeof(double));
  for (int c = 0; c < nc; c++) {
    const double *cc = ctx->ccen + 3 * c;
    double Kc[3][3];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) Kc[i][j] = ctx->K0[c * 9 + i * 3 + j];
    for (int s = 0; s < 8; s++) {
      int v = ctx->cell_verts[c * 8 + s];
      if (v < 0) continue; /* prism: 6 of 8 vertex slots */
      int fid[3], nfound = 0;
      for (int lf = 0; lf < 6; lf++) {
        int f = ctx->cell_faces[c * 6 + lf];
        if (f < 0) continue; /* prism: 5 of 6 face slots */
        if (face_has_vertex(ctx, f, v) >= 0) { if (nfound < 3) fid[nfound] = f; nfound++; }
      }
      if (nfound != 3) { set_err("subcell does not have 3 faces"); return 1; }
      double a[3], b[3], cv[3], axb[3];
      for (int d = 0; d < 3; d++) {
        a[d] = ctx->fcen[3 * fid[1] + d] - ctx->fcen[3 * fid[0] + d];
        b[d] = ctx->fcen[3 * fid[2] + d] - ctx->fcen[3 * fid[0] + d];
        cv[d] = cc[d] - ctx->vxyz[3 * v + d];
      }
      cross3(a, b, axb);
      if (dot3(axb, cv) < 0) { int t = fid[1]; fid[1] = fid[2]; fid[2] = t; }
      for (int k = 0; k < 3; k++) ctx->sc_faces[(c * 8 + s) * 3 + k] = fid[k];

      /* T = 6 * |vol(cell_cen, f0, f1, f2)| */
      double e1[3], e2[3], e3[3], x[3];
      for (int d = 0; d < 3; d++) {
        e1[d] = ctx->fcen[3 * fid[0] + d] - cc[d];
        e2[d] = ctx->fcen[3 * fid[1] + d] - cc[d];
        e3[d] = ctx->fcen[3 * fid[2] + d] - cc[d];
      }
      cross3(e1, e2, x);
      double Tsub = fabs(dot3(x, e3)) / 6.0 * 6.0;
      double nu[3][3];
      for (int k = 0; k < 3; k++) {
        int f1 = (k + 1) % 3, f2 = (k + 2) % 3;
        double p1[3], p2[3];
        for (int d = 0; d < 3; d++) {
          p1[d] = ctx->fcen[3 * fid[f1] + d] - cc[d];
          p2[d] = ctx->fcen[3 * fid[f2] + d] - cc[d];
        }
        cross3(p1, p2, nu[k]);
      }
      double *G = ctx->G + (size_t)(c * 8 + s) * 9;
      for (int i = 0; i < 3; i++) {
        int f = fid[i];
        int nverts = face_nverts(ctx, f);
        double area = ctx->farea[f] / nverts;
        const double *n = ctx->fnorm + 3 * f;
        for (int j = 0; j < 3; j++) {
          if (ctx->gmatrix == 0) {
            double Kv[3], g = 0.0;
            for (int ii = 0; ii < 3; ii++) {
              Kv[ii] = 0.0;
              for (int jj = 0; jj < 3; jj++) Kv[ii] += Kc[ii][jj] * nu[j][jj];
            }
            for (int ii = 0; ii < 3; ii++) g += n[ii] * Kv[ii];
            g *= 1.0 / (Tsub) * area;
            G[i * 3 + j] = g;
          } else {
            if (i != j) { G[i * 3 + j] = 0.0; continue; }
            int up = ctx->face_cells[2 * f], dn = ctx->face_cells[2 * f + 1];
            int nb = (up == c) ? dn : up;
            double ncen[3], dist, Kn[3][3];
            if (nb >= 0) {
              for (int d = 0; d < 3; d++) ncen[d] = ctx->ccen[3 * nb + d];
              dist = length3(ncen, cc) * 0.50;
              for (int ii = 0; ii < 3; ii++) for (int jj = 0; jj < 3; jj++) Kn[ii][jj] = ctx->K0[nb * 9 + ii * 3 + jj];
            } else {
              for (int d = 0; d < 3; d++) ncen[d] = ctx->fcen[3 * f + d];
              dist = length3(ncen, cc);
              for (int ii = 0; ii < 3; ii++) for (int jj = 0; jj < 3; jj++) Kn[ii][jj] = Kc[ii][jj];
            }
            double u[3];
            unit_from_to(ncen, cc, u);
            double dp = dot3(n, u);
            double kn = 0.0, kv = 0.0;
            for (int kk = 0; kk < 3; kk++) {
              kn += pow(u[kk], 2.0) / Kn[kk][kk];
              kv += pow(u[kk], 2.0) / Kc[kk][kk];
            }
            kn = 1.0 / kn; kv = 1.0 / kv;
            double kav = 0.5 * kv + 0.5 * kn;
            G[i * 3 + j] = area * (dp)*kav / (dist);
          }
        }
      }
    }
  }
  return 0;
}

/* Vertex interaction regions and their local systems.  Generalises
 * ComputeTransmissibilityMatrix_ForNonCornerVertex (tdympfao.c:1126-1513)
 * and ..._ForBoundaryVertex_NotSharedWithInternalVertices (1515-1756) to
 * per-face boundary types; dispatch as in ComputeTransmissibilityMatrix
 * (1801-1860). */
static int build_trans(ora_ctx *ctx) {
  int nc = ctx->nc, nf = ctx->nf, nv = ctx->nv;
  /* vertex -> cells */
  int *vc_cnt = calloc(nv + 1, sizeof(int));
  for (int c = 0; c < nc; c++)
    for (int s = 0; s < 8; s++)
      if (ctx->cell_verts[c * 8 + s] >= 0) vc_cnt[ctx->cell_verts[c * 8 + s] + 1]++;
  for (int v = 0; v < nv; v++) vc_cnt[v + 1] += vc_cnt[v];
  int ncell_tot = vc_cnt[nv];
  ctx->ir_cells = malloc(ncell_tot * sizeof(int));
  ctx->ir_subc = malloc(ncell_tot * sizeof(int));
  int *fill = calloc(nv, sizeof(int));
  for (int c = 0; c < nc; c++)
    for (int s = 0; s < 8; s++) {
      int v = ctx->cell_verts[c * 8 + s];
      if (v < 0) continue;
      int k = vc_cnt[v] + fill[v]++;
      ctx->ir_cells[k] = c; ctx->ir_subc[k] = s;
    }
  /* vertex -> faces (internal first, then boundary, each by face id) */
  int *vf_cnt = calloc(nv + 1, sizeof(int));
  for (int f = 0; f < nf; f++)
    for (int q = 0; q < 4; q++)
      if (ctx->face_verts[f * 4 + q] >= 0) vf_cnt[ctx->face_verts[f * 4 + q] + 1]++;
  for (int v = 0; v < nv; v++) vf_cnt[v + 1] += vf_cnt[v];
  int nface_tot = vf_cnt[nv];
  ctx->ir_faces = malloc(nface_tot * sizeof(int));
  ctx->ir_subface = malloc(nface_tot * sizeof(int));
  memset(fill, 0, nv * sizeof(int));
  for (int pass = 0; pass < 2; pass++)
    for (int f = 0; f < nf; f++) {
      if (is_bnd_face(ctx, f) != pass) continue;
      for (int q = 0; q < 4; q++) {
        int v = ctx->face_verts[f * 4 + q];
        if (v < 0) continue; /* triangle: 3 of 4 vertex slots */
        int k = vf_cnt[v] + fill[v]++;
        ctx->ir_faces[k] = f; ctx->ir_subface[k] = q;
      }
    }
  free(fill);

  ctx->ir = calloc(nv, sizeof(vtx_ir));
  size_t Tsize = 0; int bcol_tot = 0, row_tot = 0;
  for (int v = 0; v < nv; v++) {
    vtx_ir *r = &ctx->ir[v];
    r->off_cell = vc_cnt[v]; r->ncell = vc_cnt[v + 1] - vc_cnt[v];
    r->off_face = vf_cnt[v]; r->nface = vf_cnt[v + 1] - vf_cnt[v];
    int nint = 0, ndir = 0;
    for (int k = 0; k < r->nface; k++) {
      int f = ctx->ir_faces[r->off_face + k];
      if (!is_bnd_face(ctx, f)) nint++;
      else if (is_dirichlet_like(face_bc_of(ctx, f))) ndir++;
    }
    r->nint = nint; r->nbcol = ndir;
    r->nrow = nint + ndir; r->ncol = r->ncell + ndir;
    r->off_T = (int)Tsize; Tsize += (size_t)r->nrow * r->ncol;
    r->off_bcol = bcol_tot; bcol_tot += ndir;
    row_tot += r->nrow;
  }
  ctx->Tvals = calloc(Tsize ? Tsize : 1, sizeof(double));
  ctx->ir_bcol = malloc((bcol_tot ? bcol_tot : 1) * sizeof(int));
  ctx->ir_rowface = malloc((row_tot ? row_tot : 1) * sizeof(int));

  /* scratch */
  int maxf = 0, maxc = 0;
  for (int v = 0; v < nv; v++) { if (ctx->ir[v].nface > maxf) maxf = ctx->ir[v].nface; if (ctx->ir[v].ncell > maxc) maxc = ctx->ir[v].ncell; }
  int *frole = malloc(maxf * sizeof(int));   /* 0 int, 1 neu, 2 dir */
  int *fidx = malloc(maxf * sizeof(int));    /* unknown or dir index */
  int *fup = malloc(maxf * sizeof(int));     /* local cell index of the subcell that is face up (or the only one) */
  int *fupl = malloc(maxf * sizeof(int));    /* its local face in subcell */
  int *fdn = malloc(maxf * sizeof(int));
  int *fdnl = malloc(maxf * sizeof(int));
  int maxU = maxf, maxB = maxc + maxf;
  double *A = malloc((size_t)maxU * maxU * sizeof(double));
  double *B = malloc((size_t)maxU * maxB * sizeof(double));
  double *Crow = malloc((size_t)(maxU + maxf) * sizeof(double));
  int *piv = malloc(maxU * sizeof(int));

  int rowcur = 0;
  for (int v = 0; v < nv; v++) {
    vtx_ir *r = &ctx->ir[v];
    int nint = r->nint, nface = r->nface, ncell = r->ncell;
    int nneu = 0, ndir = 0;
    for (int k = 0; k < nface; k++) {
      int f = ctx->ir_faces[r->off_face + k];
      fup[k] = fdn[k] = -1;
      if (!is_bnd_face(ctx, f)) { frole[k] = 0; fidx[k] = k; }
      else if (is_dirichlet_like(face_bc_of(ctx, f))) { frole[k] = 2; fidx[k] = ndir; ctx->ir_bcol[r->off_bcol + ndir] = -ctx->face_cells[2 * f] - 1; ndir++; }
      else { frole[k] = 1; fidx[k] = nint + nneu; nneu++; }
    }
    int nU = nint + nneu, nB = ncell + ndir;
    /* find subcells touching each face */
    for (int i = 0; i < ncell; i++) {
      int c = ctx->ir_cells[r->off_cell + i], s = ctx->ir_subc[r->off_cell + i];
      for (int l = 0; l < 3; l++) {
        int f = ctx->sc_faces[(c * 8 + s) * 3 + l];
        int k;
        for (k = 0; k < nface; k++) if (ctx->ir_faces[r->off_face + k] == f) break;
        if (k == nface) { set_err("face not found in vertex list"); return 1; }
        if (ctx->face_cells[2 * f] == c || is_bnd_face(ctx, f)) { fup[k] = i; fupl[k] = l; }
        else { fdn[k] = i; fdnl[k] = l; }
      }
    }
    /* C row of subcell i, local face l: coefficients over [x (nU) | dir (ndir)],
     * F value for cell i = -sum_m G[l][m] (ComputeCandFmatrix 1000-1076) */
#define CROW(i_, l_, out_, Fout_)                                                     \
    do {                                                                              \
      int c_ = ctx->ir_cells[r->off_cell + (i_)], s_ = ctx->ir_subc[r->off_cell + (i_)]; \
      const double *G_ = ctx->G + (size_t)(c_ * 8 + s_) * 9;                          \
      for (int q_ = 0; q_ < nU + ndir; q_++) (out_)[q_] = 0.0;                        \
      double F_ = 0.0;                                                                \
      F_ = -G_[(l_) * 3 + 0] - G_[(l_) * 3 + 1] - G_[(l_) * 3 + 2];                   \
      for (int m_ = 0; m_ < 3; m_++) {                                                \
        int fm_ = ctx->sc_faces[(c_ * 8 + s_) * 3 + m_];                              \
        int km_;                                                                      \
        for (km_ = 0; km_ < nface; km_++) if (ctx->ir_faces[r->off_face + km_] == fm_) break; \
        int col_ = (frole[km_] == 2) ? nU + fidx[km_] : fidx[km_];                    \
        (out_)[col_] = -G_[(l_) * 3 + m_];                                            \
      }                                                                               \
      (Fout_) = F_;                                                                   \
    } while (0)

    double *Tv = ctx->Tvals + r->off_T;
    if (nU > 0) {
      memset(A, 0, sizeof(double) * nU * nU);
      memset(B, 0, sizeof(double) * nU * nB);
      double *Cu = Crow, Fu, Fd;
      double *Cd = malloc(sizeof(double) * (nU + ndir));
      for (int k = 0; k < nface; k++) {
        if (frole[k] == 0) {
          int row = fidx[k];
          CROW(fup[k], fupl[k], Cu, Fu);
          CROW(fdn[k], fdnl[k], Cd, Fd);
          for (int j = 0; j < nU; j++) A[row + j * nU] = -Cu[j] + Cd[j];
          B[row + fup[k] * nU] += -Fu;
          B[row + fdn[k] * nU] += Fd;
          for (int j = 0; j < ndir; j++) B[row + (ncell + j) * nU] = Cu[nU + j] - Cd[nU + j];
        } else if (frole[k] == 1) {
          int row = fidx[k];
          CROW(fup[k], fupl[k], Cu, Fu);
          for (int j = 0; j < nU; j++) A[row + j * nU] = -Cu[j];
          B[row + fup[k] * nU] += -Fu;
          for (int j = 0; j < ndir; j++) B[row + (ncell + j) * nU] = Cu[nU + j];
        }
      }
      free(Cd);
      int info = lu_factor(nU, A, piv);
      if (info) { set_err("singular interaction-region system (LU zero pivot)"); return 1; }
      lu_solve(nU, A, piv, nB, B); /* B <- A^-1 B */
    }
    /* T rows: internal faces, then Dirichlet faces (vertex face order) */
    int trow = 0;
    for (int pass = 0; pass < 2; pass++)
      for (int k = 0; k < nface; k++) {
        if ((pass == 0 && frole[k] != 0) || (pass == 1 && frole[k] != 2)) continue;
        double Fs;
        CROW(fup[k], fupl[k], Crow, Fs);
        double *Trow = Tv + (size_t)trow * r->ncol;
        for (int j = 0; j < nB; j++) {
          double s = 0.0;
          for (int q = 0; q < nU; q++) s += Crow[q] * B[q + j * nU];
          if (j < ncell) Trow[j] = s + ((j == fup[k]) ? -Fs : 0.0);
          else Trow[j] = s + Crow[nU + (j - ncell)];
        }
        ctx->ir_rowface[rowcur + trow] = k;
        trow++;
      }
#undef CROW
    rowcur += trow;
  }
  free(frole); free(fidx); free(fup); free(fupl); free(fdn); free(fdnl);
  free(A); free(B); free(Crow); free(piv);
  free(vc_cnt); free(vf_cnt);

  /* Trans CSR: rows face*4+subface, cols cell | nc + bnd (MatSetValues 1451-1466) */
  int nrows = 4 * nf;
  int *cnt = calloc(nrows + 1, sizeof(int));
  rowcur = 0;
  for (int v = 0; v < nv; v++) {
    vtx_ir *r = &ctx->ir[v];
    for (int t = 0; t < r->nrow; t++) {
      int k = ctx->ir_rowface[rowcur + t];
      int f = ctx->ir_faces[r->off_face + k], q = ctx->ir_subface[r->off_face + k];
      cnt[f * 4 + q + 1] += r->ncol;
    }
    rowcur += r->nrow;
  }
  for (int i = 0; i < nrows; i++) cnt[i + 1] += cnt[i];
  ctx->tr_rowptr = cnt;
  ctx->tr_cols = malloc((cnt[nrows] ? cnt[nrows] : 1) * sizeof(int));
  ctx->tr_vals = malloc((cnt[nrows] ? cnt[nrows] : 1) * sizeof(double));
  rowcur = 0;
  for (int v = 0; v < nv; v++) {
    vtx_ir *r = &ctx->ir[v];
    for (int t = 0; t < r->nrow; t++) {
      int k = ctx->ir_rowface[rowcur + t];
      int f = ctx->ir_faces[r->off_face + k], q = ctx->ir_subface[r->off_face + k];
      int base = cnt[f * 4 + q];
      /* insertion sort by column */
      int n = 0;
      for (int j = 0; j < r->ncol; j++) {
        int col = (j < r->ncell) ? ctx->ir_cells[r->off_cell + j] : nc + ctx->ir_bcol[r->off_bcol + j - r->ncell];
        double val = ctx->Tvals[r->off_T + (size_t)t * r->ncol + j];
        int p = n;
        while (p > 0 && ctx->tr_cols[base + p - 1] > col) { ctx->tr_cols[base + p] = ctx->tr_cols[base + p - 1]; ctx->tr_vals[base + p] = ctx->tr_vals[base + p - 1]; p--; }
        ctx->tr_cols[base + p] = col; ctx->tr_vals[base + p] = val; n++;
      }
    }
    rowcur += r->nrow;
  }
  return 0;
}

/* ComputeGravityDiscretization, tdympfao.c:2169-2274 (+ helpers 1869-2143) */
static void build_gravdis(ora_ctx *ctx) {
  int nf = ctx->nf;
  ctx->gravdis = calloc((size_t)4 * nf, sizeof(double));
  double grav[3] = {0.0, 0.0, ctx->gravity};
  for (int v = 0; v < ctx->nv; v++) {
    vtx_ir *r = &ctx->ir[v];
    for (int k = 0; k < r->nface; k++) {
      int f = ctx->ir_faces[r->off_face + k], q = ctx->ir_subface[r->off_face + k];
      int up = ctx->face_cells[2 * f], dn = ctx->face_cells[2 * f + 1];
      if ((up < 0 || dn < 0) && ctx->face_bc[f] == 1) continue; /* no-flow face */
      double area = ctx->farea[f] / face_nverts(ctx, f);
      const double *fc = ctx->fcen + 3 * f;
      const double *n = ctx->fnorm + 3 * f;
      double cu[3], cd[3], u[3];
      for (int d = 0; d < 3; d++) {
        cu[d] = up >= 0 ? ctx->ccen[3 * up + d] : fc[d];
        cd[d] = dn >= 0 ? ctx->ccen[3 * dn + d] : fc[d];
      }
      unit_from_to(cu, cd, u);
      double dist_up = up >= 0 ? length3(cu, fc) : 0.0;
      double dist_dn = dn >= 0 ? length3(cd, fc) : 0.0;
      double Kup[9], Kdn[9];
      for (int kk = 0; kk < 3; kk++)
        for (int mm = 0; mm < 3; mm++) {
          Kup[mm * 3 + kk] = ctx->K0[(up >= 0 ? up : dn) * 9 + kk * 3 + mm];
          Kdn[mm * 3 + kk] = ctx->K0[(dn >= 0 ? dn : up) * 9 + kk * 3 + mm];
        }
      double wt_up = dist_up / (dist_up + dist_dn);
      double kval;
      if (ctx->gmatrix == 0) {
        double Kui[9], Kdi[9], Kfi[9], Kf[9];
        inv3(Kup, Kui); inv3(Kdn, Kdi);
        for (int i = 0; i < 9; i++) Kfi[i] = wt_up * Kui[i] + (1.0 - wt_up) * Kdi[i];
        inv3(Kfi, Kf);
        for (int i = 0; i < 9; i++) if (Kf[i] < 0.0 || fabs(Kf[i]) < MACH_EPS) Kf[i] = 0.0;
        double Ku[3];
        for (int ii = 0; ii < 3; ii++) { Ku[ii] = 0.0; for (int jj = 0; jj < 3; jj++) Ku[ii] += Kf[ii * 3 + jj] * u[jj]; }
        kval = dot3(n, Ku);
      } else {
        double ku = 0.0, kd = 0.0;
        for (int kk = 0; kk < 3; kk++) { ku += pow(u[kk], 2.0) / Kup[kk * 3 + kk]; kd += pow(u[kk], 2.0) / Kdn[kk * 3 + kk]; }
        ku = 1.0 / ku; kd = 1.0 / kd;
        kval = (ku * kd) / (wt_up * kd + (1.0 - wt_up) * ku);
        area *= dot3(n, u);
      }
      double gz = 0.0;
      for (int d = 0; d < 3; d++) gz += fabs(grav[d]) * u[d];
      ctx->gravdis[f * 4 + q] = area * kval * gz;
    }
  }
}

static void build_jac_pattern(ora_ctx *ctx) {
  int nc = ctx->nc, nv = ctx->nv;
  /* cells sharing a vertex (DMSetBasicAdjacency(dm,TRUE,TRUE)) */
  int *mark = malloc(nc * sizeof(int));
  for (int i = 0; i < nc; i++) mark[i] = -1;
  int *rp = calloc(nc + 1, sizeof(int));
  int maxc = 0;
  for (int v = 0; v < nv; v++) if (ctx->ir[v].ncell > maxc) maxc = ctx->ir[v].ncell;
  int *tmp = malloc((size_t)(maxc > 0 ? maxc : 1) * 8 * sizeof(int));
  for (int pass = 0; pass < 2; pass++) {
    for (int c = 0; c < nc; c++) {
      int n = 0;
      for (int s = 0; s < 8; s++) {
        if (ctx->cell_verts[c * 8 + s] < 0) continue;
        vtx_ir *r = &ctx->ir[ctx->cell_verts[c * 8 + s]];
        for (int i = 0; i < r->ncell; i++) {
          int cc = ctx->ir_cells[r->off_cell + i];
          if (mark[cc] != c) { mark[cc] = c; tmp[n++] = cc; }
        }
      }
      if (pass == 0) rp[c + 1] = n;
      else {
        /* sort */
        for (int i = 1; i < n; i++) { int t = tmp[i], j = i; while (j > 0 && tmp[j - 1] > t) { tmp[j] = tmp[j - 1]; j--; } tmp[j] = t; }
        memcpy(ctx->j_cols + rp[c], tmp, n * sizeof(int));
      }
    }
    if (pass == 0) {
      for (int c = 0; c < nc; c++) rp[c + 1] += rp[c];
      ctx->j_cols = malloc(rp[nc] * sizeof(int));
      for (int i = 0; i < nc; i++) mark[i] = -1;
    }
  }
  ctx->j_rowptr = rp;
  free(mark); free(tmp);
}

ora_ctx *ora_create(int nc, int nf, int nv, int nbnd, const int *cell_verts, const int *cell_faces,
                    const int *face_verts, const int *face_cells, const double *vert_xyz,
                    const double *cell_cen, const double *cell_vol, const double *face_cen,
                    const double *face_area, const double *face_normal, const int *face_bc,
                    const double *K, const double *porosity, const double *Sr,
                    const double *params, const double *p_bnd_fn) {
  ora_ctx *ctx = calloc(1, sizeof(ora_ctx));
  ctx->nc = nc; ctx->nf = nf; ctx->nv = nv; ctx->nbnd = nbnd;
  ctx->cell_verts = dupi(cell_verts, (size_t)nc * 8);
  ctx->cell_faces = dupi(cell_faces, (size_t)nc * 6);
  ctx->face_verts = dupi(face_verts, (size_t)nf * 4);
  ctx->face_cells = dupi(face_cells, (size_t)nf * 2);
  ctx->face_bc = dupi(face_bc, nf);
  ctx->vxyz = dupd(vert_xyz, (size_t)nv * 3);
  ctx->ccen = dupd(cell_cen, (size_t)nc * 3);
  ctx->cvol = dupd(cell_vol, nc);
  ctx->fcen = dupd(face_cen, (size_t)nf * 3);
  ctx->farea = dupd(face_area, nf);
  ctx->fnorm = dupd(face_normal, (size_t)nf * 3);
  ctx->K0 = dupd(K, (size_t)nc * 9);
  ctx->phi = dupd(porosity, nc);
  ctx->Sr = dupd(Sr, nc);
  ctx->vg_m = params[ORA_P_VG_M]; ctx->vg_alpha = params[ORA_P_VG_ALPHA];
  ctx->x0 = params[ORA_P_X0]; ctx->x1 = params[ORA_P_X1]; ctx->x2 = params[ORA_P_X2]; ctx->dx = params[ORA_P_DX];
  ctx->gravity = params[ORA_P_GRAVITY];
  ctx->density_type = (int)params[ORA_P_DENSITY];
  ctx->gmatrix = (int)params[ORA_P_GMATRIX];
  ctx->newton_maxit = 50;
  ctx->pref = params[ORA_P_PREF];
  ctx->p_bnd_fn = p_bnd_fn ? dupd(p_bnd_fn, nbnd) : NULL;
  if (mualem_coeffs(ctx->vg_m, ctx->x0, ctx->x1, ctx->x2, ctx->dx, ctx->coeffs)) { set_err("cubic setup failed"); ora_destroy(ctx); return NULL; }
  if (build_subcells(ctx) || build_trans(ctx)) { ora_destroy(ctx); return NULL; }
  build_gravdis(ctx);
  build_jac_pattern(ctx);
  ctx->nthreads = 1;
  ctx->ir_rowstart = malloc(sizeof(int) * (nv + 1));
  ctx->row_T = malloc(sizeof(int) * 4 * (size_t)nf);
  ctx->row_v = malloc(sizeof(int) * 4 * (size_t)nf);
  ctx->flux_row = calloc((size_t)4 * nf, sizeof(double));
  for (size_t i = 0; i < (size_t)4 * nf; i++) { ctx->row_T[i] = -1; ctx->row_v[i] = -1; }
  {
    int rowcur = 0;
    size_t tsz = 0;
    for (int v = 0; v < nv; v++) {
      const vtx_ir *r = &ctx->ir[v];
      ctx->ir_rowstart[v] = rowcur;
      for (int k = 0; k < r->nface; k++) {
        int f = ctx->ir_faces[r->off_face + k], q = ctx->ir_subface[r->off_face + k];
        ctx->row_v[f * 4 + q] = v;
      }
      for (int t = 0; t < r->nrow; t++) {
        int k = ctx->ir_rowface[rowcur + t];
        int f = ctx->ir_faces[r->off_face + k], q = ctx->ir_subface[r->off_face + k];
        ctx->row_T[f * 4 + q] = r->off_T + t * r->ncol;
      }
      rowcur += r->nrow;
      tsz = (size_t)r->off_T + (size_t)r->nrow * r->ncol > tsz ? (size_t)r->off_T + (size_t)r->nrow * r->ncol : tsz;
    }
    ctx->ir_rowstart[nv] = rowcur;
    ctx->D = calloc(tsz ? tsz : 1, sizeof(double));
  }
  size_t n = nc;
  ctx->S = calloc(n, 8); ctx->dS = calloc(n, 8); ctx->d2S = calloc(n, 8); ctx->Kr = calloc(n, 8);
  ctx->dKr = calloc(n, 8); ctx->rho = calloc(n, 8); ctx->drho = calloc(n, 8); ctx->d2rho = calloc(n, 8);
  ctx->vis = calloc(n, 8); ctx->dvis = calloc(n, 8);
  ctx->P_bnd = calloc(nbnd + 1, 8); ctx->Kr_bnd = calloc(nbnd + 1, 8); ctx->S_bnd = calloc(nbnd + 1, 8);
  ctx->rho_bnd = calloc(nbnd + 1, 8); ctx->vis_bnd = calloc(nbnd + 1, 8);
  ctx->bnd_cell = calloc(nbnd + 1, sizeof(int));
  ctx->Pvec = calloc(nc + nbnd, 8); ctx->TP = calloc((size_t)4 * nf, 8);
  ctx->accum_prev = calloc(n, 8);
  ctx->source = calloc(n, 8);
  ctx->dtime = 1.0;
  for (int f = 0; f < nf; f++) {
    int up = ctx->face_cells[2 * f], dn = ctx->face_cells[2 * f + 1];
    if (up < 0) ctx->bnd_cell[-up - 1] = dn;
    else if (dn < 0) ctx->bnd_cell[-dn - 1] = up;
  }
  /* AllocateMemoryForBoundaryValues: rho, mu at (Pref, Tref), tdympfao.c:297-336 */
  for (int b = 0; b < nbnd; b++) {
    double d1, d2, dv;
    density(ctx->density_type, ctx->pref, &ctx->rho_bnd[b], &d1, &d2);
    viscosity(ctx->pref, &ctx->vis_bnd[b], &dv);
  }
  return ctx;
}

void ora_destroy(ora_ctx *ctx) {
  if (!ctx) return;
  void *ps[] = {ctx->cell_verts, ctx->cell_faces, ctx->face_verts, ctx->face_cells, ctx->face_bc,
                ctx->vxyz, ctx->ccen, ctx->cvol, ctx->fcen, ctx->farea, ctx->fnorm, ctx->K0, ctx->phi,
                ctx->Sr, ctx->p_bnd_fn, ctx->ir, ctx->ir_cells, ctx->ir_subc, ctx->ir_faces,
                ctx->ir_subface, ctx->ir_bcol, ctx->ir_rowface, ctx->Tvals, ctx->sc_faces, ctx->G,
                ctx->tr_rowptr, ctx->tr_cols, ctx->tr_vals, ctx->gravdis, ctx->j_rowptr, ctx->j_cols,
                ctx->S, ctx->dS, ctx->d2S, ctx->Kr, ctx->dKr, ctx->rho, ctx->drho, ctx->d2rho,
                ctx->vis, ctx->dvis, ctx->P_bnd, ctx->Kr_bnd, ctx->S_bnd, ctx->rho_bnd, ctx->vis_bnd,
                ctx->bnd_cell, ctx->Pvec, ctx->TP, ctx->accum_prev, ctx->source, ctx->ir_rowstart,
                ctx->row_T, ctx->row_v, ctx->flux_row, ctx->D};
  for (size_t i = 0; i < sizeof ps / sizeof ps[0]; i++) free(ps[i]);
  free(ctx);
}

void ora_sizes(const ora_ctx *ctx, long long *out) {
  out[0] = ctx->tr_rowptr[4 * ctx->nf];
  out[1] = ctx->j_rowptr[ctx->nc];
  int mc = 0, mf = 0;
  for (int v = 0; v < ctx->nv; v++) { if (ctx->ir[v].ncell > mc) mc = ctx->ir[v].ncell; if (ctx->ir[v].nface > mf) mf = ctx->ir[v].nface; }
  out[2] = mc; out[3] = mf;
}
void ora_trans_csr(const ora_ctx *ctx, int *rowptr, int *cols, double *vals) {
  int nr = 4 * ctx->nf, nnz = ctx->tr_rowptr[nr];
  memcpy(rowptr, ctx->tr_rowptr, (nr + 1) * sizeof(int));
  memcpy(cols, ctx->tr_cols, nnz * sizeof(int));
  memcpy(vals, ctx->tr_vals, nnz * sizeof(double));
}
void ora_gravdis(const ora_ctx *ctx, double *out) { memcpy(out, ctx->gravdis, sizeof(double) * 4 * ctx->nf); }
void ora_jac_pattern(const ora_ctx *ctx, int *rowptr, int *cols) {
  memcpy(rowptr, ctx->j_rowptr, (ctx->nc + 1) * sizeof(int));
  memcpy(cols, ctx->j_cols, ctx->j_rowptr[ctx->nc] * sizeof(int));
}

/* ------------------------------------------------------------------ */
/* state                                                                */
/* ------------------------------------------------------------------ */

/* TDyUpdateState_Richards_MPFAO, tdympfao.c:2527-2594 */
void ora_update_state(ora_ctx *ctx, const double *p) {
#pragma omp parallel for num_threads(ctx->nthreads) schedule(static)
  for (int c = 0; c < ctx->nc; c++) {
    double Pc = ctx->pref - p[c];
    vg_sat(ctx->vg_m, ctx->vg_alpha, ctx->Sr[c], Pc, &ctx->S[c], &ctx->dS[c], &ctx->d2S[c]);
    double Se = (ctx->S[c] - ctx->Sr[c]) / (1.0 - ctx->Sr[c]);
    mualem(ctx, Se, &ctx->Kr[c], &ctx->dKr[c]);
    double dSe_dS = 1.0 / (1.0 - ctx->Sr[c]);
    ctx->dKr[c] *= dSe_dS;
    double P = ctx->pref - Pc;
    density(ctx->density_type, P, &ctx->rho[c], &ctx->drho[c], &ctx->d2rho[c]);
    viscosity(P, &ctx->vis[c], &ctx->dvis[c]);
    ctx->Pvec[c] = P;
  }
}
void ora_get_state(const ora_ctx *ctx, int which, double *out) {
  const double *src[] = {ctx->S, ctx->dS, ctx->d2S, ctx->Kr, ctx->dKr, ctx->rho, ctx->drho, ctx->d2rho, ctx->vis};
  memcpy(out, src[which], sizeof(double) * ctx->nc);
}
void ora_boundary_pressure(const ora_ctx *ctx, double *out) { memcpy(out, ctx->P_bnd, sizeof(double) * ctx->nbnd); }

/* TDyMPFAO_SetBoundaryPressure (tdympfao_utils.c:690-759) and
 * TDyMPFAOUpdateBoundaryState (tdympfao_utils.c:24-99) */
static void boundary_state(ora_ctx *ctx, const double *p) {
#pragma omp parallel for num_threads(ctx->nthreads) schedule(static)
  for (int b = 0; b < ctx->nbnd; b++) {
    int c = ctx->bnd_cell[b];
    ctx->P_bnd[b] = ctx->p_bnd_fn ? ctx->p_bnd_fn[b] : p[c];
    ctx->Pvec[ctx->nc + b] = ctx->P_bnd[b];
    double S, dS, d2S, Kr, dKr;
    vg_sat(ctx->vg_m, ctx->vg_alpha, ctx->Sr[c], ctx->pref - ctx->P_bnd[b], &S, &dS, &d2S);
    double Se = (S - ctx->Sr[c]) / (1.0 - ctx->Sr[c]);
    mualem(ctx, Se, &Kr, &dKr);
    ctx->S_bnd[b] = S; ctx->Kr_bnd[b] = Kr;
  }
}

/* MatMult(Trans_mat, P_vec, TtimesP_vec), tdympfao_ts.c:184 */
static void trans_matmult(ora_ctx *ctx) {
  int nr = 4 * ctx->nf;
#pragma omp parallel for num_threads(ctx->nthreads) schedule(static)
  for (int i = 0; i < nr; i++) {
    double s = 0.0;
    for (int k = ctx->tr_rowptr[i]; k < ctx->tr_rowptr[i + 1]; k++) s += ctx->tr_vals[k] * ctx->Pvec[ctx->tr_cols[k]];
    ctx->TP[i] = s;
  }
}

/* TDyMPFAOIFunction_Vertices, tdympfao_ts.c:12-150: the flux of every
 * sub-face row of every vertex (each row belongs to one vertex), then each
 * cell sums the rows of its faces (+ for the up cell, - for the down cell) in
 * a fixed (face, sub-face) order. */
static void flux_vertices(ora_ctx *ctx, double *R) {
#pragma omp parallel for num_threads(ctx->nthreads) schedule(dynamic, 256)
  for (int v = 0; v < ctx->nv; v++) {
    vtx_ir *r = &ctx->ir[v];
    for (int k = 0; k < r->nface; k++) {
      int f = ctx->ir_faces[r->off_face + k], q = ctx->ir_subface[r->off_face + k];
      int up = ctx->face_cells[2 * f], dn = ctx->face_cells[2 * f + 1];
      int set_zero = 0;
      double TP = ctx->TP[f * 4 + q];
      double den = 0.0;
      den += up >= 0 ? ctx->rho[up] : ctx->rho_bnd[-up - 1];
      den += dn >= 0 ? ctx->rho[dn] : ctx->rho_bnd[-dn - 1];
      den *= 0.5;
      double G = ctx->gravdis[f * 4 + q];
      double ukvr;
      if (TP + den * G < 0.0) {
        if (up >= 0) ukvr = ctx->Kr[up] / ctx->vis[up];
        else {
          ukvr = ctx->Kr_bnd[-up - 1] / ctx->vis_bnd[-up - 1];
          if (ctx->face_bc[f] == 2 && ctx->P_bnd[-up - 1] <= ctx->pref) set_zero = 1;
        }
      } else {
        if (dn >= 0) ukvr = ctx->Kr[dn] / ctx->vis[dn];
        else {
          ukvr = ctx->Kr_bnd[-dn - 1] / ctx->vis_bnd[-dn - 1];
          if (ctx->face_bc[f] == 2 && ctx->P_bnd[-dn - 1] <= ctx->pref) set_zero = 1;
        }
      }
      double fluxm = 0.0;
      if (!set_zero) {
        fluxm = den * ukvr * (-TP);
        fluxm += -pow(den, 2.0) * ukvr * G;
      }
      ctx->flux_row[f * 4 + q] = fluxm;
    }
  }
#pragma omp parallel for num_threads(ctx->nthreads) schedule(static)
  for (int c = 0; c < ctx->nc; c++) {
    double s = 0.0;
    for (int lf = 0; lf < 6; lf++) {
      int f = ctx->cell_faces[c * 6 + lf];
      if (f < 0) continue;
      const double *fr = ctx->flux_row + (size_t)f * 4;
      double fs = ((fr[0] + fr[1]) + fr[2]) + fr[3];
      s += (ctx->face_cells[2 * f] == c) ? fs : -fs;
    }
    R[c] += s;
  }
}

static void residual_common(ora_ctx *ctx, const double *U, double *R) {
  memset(R, 0, sizeof(double) * ctx->nc);
  ora_update_state(ctx, U);
  boundary_state(ctx, U);
  trans_matmult(ctx);
  flux_vertices(ctx, R);
}

/* TDyMPFAOIFunction, tdympfao_ts.c:153-227 */
int ora_ifunction(ora_ctx *ctx, const double *U, const double *Udot, double *R) {
  residual_common(ctx, U, R);
  for (int c = 0; c < ctx->nc; c++) {
    double rho = ctx->rho[c], drho = ctx->drho[c], phi = ctx->phi[c], dphi = 0.0, S = ctx->S[c], dS = ctx->dS[c];
    double dm = rho * dphi * S + drho * phi * S + rho * phi * dS;
    R[c] += dm * Udot[c] * ctx->cvol[c];
    R[c] -= ctx->source[c] * ctx->cvol[c]; /* tdympfao_ts.c:211 */
  }
  return 0;
}

/* position of column col in J row */
static int jpos(const ora_ctx *ctx, int row, int col) {
  int lo = ctx->j_rowptr[row], hi = ctx->j_rowptr[row + 1] - 1;
  while (lo <= hi) {
    int mid = (lo + hi) / 2;
    if (ctx->j_cols[mid] == col) return mid;
    if (ctx->j_cols[mid] < col) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

/* TDyMPFAOIJacobian_Vertices, tdympfao_ts.c:235-456 (uses the state of the
 * preceding residual evaluation, like the reference) */
static void jac_vertices(ora_ctx *ctx, double *J) {
#pragma omp parallel for num_threads(ctx->nthreads) schedule(dynamic, 256)
  for (int v = 0; v < ctx->nv; v++) {
    vtx_ir *r = &ctx->ir[v];
    const int rowcur = ctx->ir_rowstart[v];
    /* map face k -> T row */
    for (int t = 0; t < r->nrow; t++) {
      int k = ctx->ir_rowface[rowcur + t];
      int f = ctx->ir_faces[r->off_face + k], q = ctx->ir_subface[r->off_face + k];
      int up = ctx->face_cells[2 * f], dn = ctx->face_cells[2 * f + 1];
      int set_zero = 0;
      double den = 0.0, dden_up = 0.0, dden_dn = 0.0, ukvr, dukvr_up = 0.0, dukvr_dn = 0.0;
      if (up >= 0) { den += 0.5 * ctx->rho[up]; dden_up = 0.5 * ctx->drho[up]; }
      else { den += 0.5 * ctx->rho_bnd[-up - 1]; dden_up = 0.0; }
      if (dn >= 0) { den += 0.5 * ctx->rho[dn]; dden_dn = 0.5 * ctx->drho[dn]; }
      else { den += 0.5 * ctx->rho_bnd[-dn - 1]; dden_dn = 0.0; }
      double G = ctx->gravdis[f * 4 + q];
      double TP = ctx->TP[f * 4 + q];
      if (TP + den * G < 0.0) {
        if (up >= 0) {
          double Kr = ctx->Kr[up], dKr = ctx->dKr[up], vis = ctx->vis[up], dvis = ctx->dvis[up], dS = ctx->dS[up];
          ukvr = Kr / vis; dukvr_up = dKr * dS / vis - Kr / (vis * vis) * dvis;
        } else {
          ukvr = ctx->Kr_bnd[-up - 1] / ctx->vis_bnd[-up - 1]; dukvr_up = 0.0;
          if (ctx->face_bc[f] == 2 && ctx->P_bnd[-up - 1] <= ctx->pref) set_zero = 1;
        }
      } else {
        if (dn >= 0) {
          double Kr = ctx->Kr[dn], dKr = ctx->dKr[dn], vis = ctx->vis[dn], dvis = ctx->dvis[dn], dS = ctx->dS[dn];
          ukvr = Kr / vis; dukvr_dn = dKr * dS / vis - Kr / (vis * vis) * dvis;
        } else {
          ukvr = ctx->Kr_bnd[-dn - 1] / ctx->vis_bnd[-dn - 1]; dukvr_dn = 0.0;
          if (ctx->face_bc[f] == 2 && ctx->P_bnd[-dn - 1] <= ctx->pref) set_zero = 1;
        }
      }
      const double *Trow = ctx->Tvals + r->off_T + (size_t)t * r->ncol;
      for (int icol = 0; icol < r->ncell; icol++) {
        int cell = ctx->ir_cells[r->off_cell + icol];
        double T = Trow[icol], Jac;
        if (set_zero) Jac = 0.0;
        else if (up > -1 && cell == up)
          Jac = dden_up * ukvr * TP + den * dukvr_up * TP + den * ukvr * T + 2 * den * dden_up * ukvr * G +
                pow(den, 2.0) * dukvr_up * G;
        else if (dn > -1 && cell == dn)
          Jac = dden_dn * ukvr * TP + den * dukvr_dn * TP + den * ukvr * T + 2 * den * dden_dn * ukvr * G +
                pow(den, 2.0) * dukvr_dn * G;
        else Jac = den * ukvr * T;
        ctx->D[r->off_T + (size_t)t * r->ncol + icol] = -Jac;
      }
    }
  }
  /* each cell sums its faces' rows: + for the up cell, - for the down cell
   * (the reference's MatSetValuesLocal calls, tdympfao_ts.c:439-444) */
#pragma omp parallel for num_threads(ctx->nthreads) schedule(static)
  for (int c = 0; c < ctx->nc; c++) {
    for (int lf = 0; lf < 6; lf++) {
      int f = ctx->cell_faces[c * 6 + lf];
      if (f < 0) continue;
      const int is_up = ctx->face_cells[2 * f] == c;
      for (int q = 0; q < 4; q++) {
        const int rt = ctx->row_T[f * 4 + q];
        if (rt < 0) continue;
        const vtx_ir *r = &ctx->ir[ctx->row_v[f * 4 + q]];
        for (int icol = 0; icol < r->ncell; icol++) {
          int cell = ctx->ir_cells[r->off_cell + icol];
          double d = ctx->D[rt + icol];
          J[jpos(ctx, c, cell)] += is_up ? d : -d;
        }
      }
    }
  }
}

/* TDyMPFAOIJacobian + _Accumulation, tdympfao_ts.c:459-561 */
int ora_ijacobian(ora_ctx *ctx, const double *U, const double *Udot, double shift, double *J) {
  (void)U;
  memset(J, 0, sizeof(double) * ctx->j_rowptr[ctx->nc]);
  jac_vertices(ctx, J);
  for (int c = 0; c < ctx->nc; c++) {
    double rho = ctx->rho[c], drho = ctx->drho[c], d2rho = ctx->d2rho[c], phi = ctx->phi[c];
    double S = ctx->S[c], dS = ctx->dS[c], d2S = ctx->d2S[c], dphi = 0.0, d2phi = 0.0;
    double dm = rho * dphi * S + drho * phi * S + rho * phi * dS;
    double d2m = (dS * rho * dphi + S * drho * dphi + S * rho * d2phi + dS * drho * phi + S * d2rho * phi +
                  S * drho * dphi + d2S * rho * phi + dS * drho * phi + dS * rho * dphi);
    double value = (shift * dm + d2m * Udot[c]) * ctx->cvol[c];
    J[jpos(ctx, c, c)] += value;
  }
  return 0;
}

/* SNES form, tdympfao_snes.c:19-203 */
void ora_set_dtime(ora_ctx *ctx, double dt) { ctx->dtime = dt; }
int ora_snes_presolve(ora_ctx *ctx, const double *Uprev) {
  ora_update_state(ctx, Uprev);
  for (int c = 0; c < ctx->nc; c++)
    ctx->accum_prev[c] = ctx->rho[c] * ctx->phi[c] * ctx->S[c] * ctx->cvol[c] / ctx->dtime;
  return 0;
}
int ora_snes_function(ora_ctx *ctx, const double *U, double *R) {
  residual_common(ctx, U, R);
  for (int c = 0; c < ctx->nc; c++) {
    double acc = ctx->rho[c] * ctx->phi[c] * ctx->S[c] * ctx->cvol[c] / ctx->dtime;
    R[c] += acc - ctx->accum_prev[c];
    R[c] -= ctx->source[c]; /* not scaled by the volume, tdympfao_snes.c:122 */
  }
  return 0;
}
int ora_snes_jacobian(ora_ctx *ctx, const double *U, double *J) {
  (void)U;
  memset(J, 0, sizeof(double) * ctx->j_rowptr[ctx->nc]);
  jac_vertices(ctx, J);
  double dtInv = 1.0 / ctx->dtime;
  for (int c = 0; c < ctx->nc; c++) {
    double dm = ctx->rho[c] * 0.0 * ctx->S[c] + ctx->drho[c] * ctx->phi[c] * ctx->S[c] + ctx->rho[c] * ctx->phi[c] * ctx->dS[c];
    J[jpos(ctx, c, c)] += dm * ctx->cvol[c] * dtInv;
  }
  return 0;
}

/* the flow equation's source_sink per cell (zero by default, as the
 * reference's AllocateMemoryForSourceSinkValues leaves it); NULL clears it */
void ora_set_source_sink(ora_ctx *ctx, const double *Q) {
  for (int c = 0; c < ctx->nc; c++) ctx->source[c] = Q ? Q[c] : 0.0;
}

void ora_liquid_mass(ora_ctx *ctx, const double *U, double *out) {
  ora_update_state(ctx, U);
  for (int c = 0; c < ctx->nc; c++) out[c] = ctx->rho[c] * ctx->cvol[c];
}

/* ------------------------------------------------------------------ */
/* linear solvers                                                       */
/* ------------------------------------------------------------------ */
static double vdot(const ora_ctx *ctx, int n, const double *x, const double *y) {
  double s = 0.0;
#pragma omp parallel for num_threads(ctx->nthreads) schedule(static) reduction(+ : s)
  for (int i = 0; i < n; i++) s += x[i] * y[i];
  return s;
}
static double nrm2c(const ora_ctx *ctx, int n, const double *x) { return sqrt(vdot(ctx, n, x, x)); }

static int dense_solve(const ora_ctx *ctx, const double *J, const double *F, double *Y) {
  int n = ctx->nc;
  double *A = calloc((size_t)n * n, sizeof(double));
  int *piv = malloc(n * sizeof(int));
  for (int i = 0; i < n; i++)
    for (int k = ctx->j_rowptr[i]; k < ctx->j_rowptr[i + 1]; k++) A[i + (size_t)ctx->j_cols[k] * n] = J[k];
  memcpy(Y, F, sizeof(double) * n);
  int info = lu_factor(n, A, piv);
  if (!info) lu_solve(n, A, piv, 1, Y);
  free(A); free(piv);
  return info;
}

/* ILU(0) in place on the CSR values (IKJ), natural ordering */
static int ilu0(const ora_ctx *ctx, double *LU, int *diag) {
  int n = ctx->nc;
  const int *rp = ctx->j_rowptr, *cj = ctx->j_cols;
  for (int i = 0; i < n; i++) { diag[i] = jpos(ctx, i, i); }
  int *pos = malloc(n * sizeof(int));
  for (int i = 0; i < n; i++) pos[i] = -1;
  for (int i = 0; i < n; i++) {
    for (int k = rp[i]; k < rp[i + 1]; k++) pos[cj[k]] = k;
    for (int k = rp[i]; k < rp[i + 1] && cj[k] < i; k++) {
      int col = cj[k];
      double l = LU[k] / LU[diag[col]];
      LU[k] = l;
      for (int kk = diag[col] + 1; kk < rp[col + 1]; kk++) {
        int p = pos[cj[kk]];
        if (p >= 0) LU[p] -= l * LU[kk];
      }
    }
    for (int k = rp[i]; k < rp[i + 1]; k++) pos[cj[k]] = -1;
    if (LU[diag[i]] == 0.0) { free(pos); return 1; }
  }
  free(pos);
  return 0;
}
static void ilu_apply(const ora_ctx *ctx, const double *LU, const int *diag, const double *b, double *x) {
  int n = ctx->nc;
  const int *rp = ctx->j_rowptr, *cj = ctx->j_cols;
  for (int i = 0; i < n; i++) {
    double s = b[i];
    for (int k = rp[i]; k < diag[i]; k++) s -= LU[k] * x[cj[k]];
    x[i] = s;
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = x[i];
    for (int k = diag[i] + 1; k < rp[i + 1]; k++) s -= LU[k] * x[cj[k]];
    x[i] = s / LU[diag[i]];
  }
}
/* Block-Jacobi ILU(0) (PETSc's parallel default, PCBJACOBI with PCILU in
 * each block): B contiguous row blocks, entries coupling two blocks dropped,
 * ILU(0) of each block in natural order; blocks are factored and applied in
 * parallel. */
static int bj_ilu0(const ora_ctx *ctx, int B, double *LU, int *diag) {
  int n = ctx->nc;
  const int *rp = ctx->j_rowptr, *cj = ctx->j_cols;
  for (int i = 0; i < n; i++) diag[i] = jpos(ctx, i, i);
  int bad = 0;
#pragma omp parallel for num_threads(ctx->nthreads) schedule(static, 1) reduction(| : bad)
  for (int b = 0; b < B; b++) {
    int lo = (int)((long long)b * n / B), hi = (int)((long long)(b + 1) * n / B);
    int *pos = malloc(sizeof(int) * (size_t)(hi - lo + 1));
    for (int i = 0; i < hi - lo; i++) pos[i] = -1;
    for (int i = lo; i < hi; i++) {
      for (int k = rp[i]; k < rp[i + 1]; k++) if (cj[k] >= lo && cj[k] < hi) pos[cj[k] - lo] = k;
      for (int k = rp[i]; k < rp[i + 1] && cj[k] < i; k++) {
        int col = cj[k];
        if (col < lo) continue;
        double l = LU[k] / LU[diag[col]];
        LU[k] = l;
        for (int kk = diag[col] + 1; kk < rp[col + 1]; kk++) {
          int c2 = cj[kk];
          if (c2 < lo || c2 >= hi) continue;
          int p = pos[c2 - lo];
          if (p >= 0) LU[p] -= l * LU[kk];
        }
      }
      for (int k = rp[i]; k < rp[i + 1]; k++) if (cj[k] >= lo && cj[k] < hi) pos[cj[k] - lo] = -1;
      if (LU[diag[i]] == 0.0) bad = 1;
    }
    free(pos);
  }
  return bad;
}
static void bj_apply(const ora_ctx *ctx, int B, const double *LU, const int *diag, const double *bv, double *x) {
  int n = ctx->nc;
  const int *rp = ctx->j_rowptr, *cj = ctx->j_cols;
#pragma omp parallel for num_threads(ctx->nthreads) schedule(static, 1)
  for (int b = 0; b < B; b++) {
    int lo = (int)((long long)b * n / B), hi = (int)((long long)(b + 1) * n / B);
    for (int i = lo; i < hi; i++) {
      double s = bv[i];
      for (int k = rp[i]; k < diag[i]; k++) if (cj[k] >= lo) s -= LU[k] * x[cj[k]];
      x[i] = s;
    }
    for (int i = hi - 1; i >= lo; i--) {
      double s = x[i];
      for (int k = diag[i] + 1; k < rp[i + 1]; k++) if (cj[k] < hi) s -= LU[k] * x[cj[k]];
      x[i] = s / LU[diag[i]];
    }
  }
}

static void csr_mv(const ora_ctx *ctx, const double *A, const double *x, double *y) {
#pragma omp parallel for num_threads(ctx->nthreads) schedule(static)
  for (int i = 0; i < ctx->nc; i++) {
    double s = 0.0;
    for (int k = ctx->j_rowptr[i]; k < ctx->j_rowptr[i + 1]; k++) s += A[k] * x[ctx->j_cols[k]];
    y[i] = s;
  }
}

/* GMRES(30), left preconditioned with ILU(0) (B = 0: global, natural order;
 * B > 0: block Jacobi with B blocks), classical Gram-Schmidt, stopping on the
 * preconditioned residual (PETSc KSP defaults: rtol 1e-5, atol 1e-50,
 * max_it 10000) */
static int gmres_ilu(const ora_ctx *ctx, const double *J, const double *F, double *Y, int *its_out, int B) {
  int n = ctx->nc, m = 30;
  size_t nnz = ctx->j_rowptr[n];
  double *LU = malloc(nnz * sizeof(double));
  int *diag = malloc(n * sizeof(int));
  memcpy(LU, J, nnz * sizeof(double));
  if (B > 0 ? bj_ilu0(ctx, B, LU, diag) : ilu0(ctx, LU, diag)) { free(LU); free(diag); return 1; }
#define PC_APPLY(b_, x_) (B > 0 ? bj_apply(ctx, B, LU, diag, (b_), (x_)) : ilu_apply(ctx, LU, diag, (b_), (x_)))
  double *V = malloc((size_t)(m + 1) * n * sizeof(double));
  double *H = calloc((size_t)(m + 1) * m, sizeof(double));
  double *cs = malloc(m * sizeof(double)), *sn = malloc(m * sizeof(double)), *g = malloc((m + 1) * sizeof(double));
  double *w = malloc(n * sizeof(double)), *t = malloc(n * sizeof(double));
  memset(Y, 0, n * sizeof(double));
  PC_APPLY(F, t);
  double r0 = nrm2c(ctx, n, t), tol = 1e-5 * r0;
  int its = 0, status = 0;
  if (r0 == 0.0) goto done;
  for (;;) {
    /* r = M^-1 (F - J Y) */
    csr_mv(ctx, J, Y, w);
#pragma omp parallel for num_threads(ctx->nthreads) schedule(static)
    for (int i = 0; i < n; i++) w[i] = F[i] - w[i];
    PC_APPLY(w, t);
    double beta = nrm2c(ctx, n, t);
    if (beta <= tol || beta < 1e-50) break;
#pragma omp parallel for num_threads(ctx->nthreads) schedule(static)
    for (int i = 0; i < n; i++) V[i] = t[i] / beta;
    memset(g, 0, (m + 1) * sizeof(double)); g[0] = beta;
    int j;
    for (j = 0; j < m; j++) {
      csr_mv(ctx, J, V + (size_t)j * n, w);
      PC_APPLY(w, t);
      for (int i = 0; i <= j; i++) H[i + j * (m + 1)] = vdot(ctx, n, t, V + (size_t)i * n);
#pragma omp parallel for num_threads(ctx->nthreads) schedule(static)
      for (int q = 0; q < n; q++) {
        double tq = t[q];
        for (int i = 0; i <= j; i++) tq -= H[i + j * (m + 1)] * V[(size_t)i * n + q];
        t[q] = tq;
      }
      double hn = nrm2c(ctx, n, t);
      H[j + 1 + j * (m + 1)] = hn;
      if (hn > 0) {
#pragma omp parallel for num_threads(ctx->nthreads) schedule(static)
        for (int q = 0; q < n; q++) V[(size_t)(j + 1) * n + q] = t[q] / hn;
      }
      for (int i = 0; i < j; i++) {
        double a = H[i + j * (m + 1)], b = H[i + 1 + j * (m + 1)];
        H[i + j * (m + 1)] = cs[i] * a + sn[i] * b;
        H[i + 1 + j * (m + 1)] = -sn[i] * a + cs[i] * b;
      }
      double a = H[j + j * (m + 1)], b = H[j + 1 + j * (m + 1)];
      double d = sqrt(a * a + b * b);
      cs[j] = a / d; sn[j] = b / d;
      H[j + j * (m + 1)] = d; H[j + 1 + j * (m + 1)] = 0.0;
      g[j + 1] = -sn[j] * g[j]; g[j] = cs[j] * g[j];
      its++;
      if (fabs(g[j + 1]) <= tol || its >= 10000 || hn == 0.0) { j++; break; }
    }
    /* back substitution */
    double *yk = malloc(j * sizeof(double));
    for (int i = j - 1; i >= 0; i--) {
      double s = g[i];
      for (int k = i + 1; k < j; k++) s -= H[i + k * (m + 1)] * yk[k];
      yk[i] = s / H[i + i * (m + 1)];
    }
#pragma omp parallel for num_threads(ctx->nthreads) schedule(static)
    for (int q = 0; q < n; q++) {
      double yq = Y[q];
      for (int i = 0; i < j; i++) yq += yk[i] * V[(size_t)i * n + q];
      Y[q] = yq;
    }
    free(yk);
    if (fabs(g[j]) <= tol) break;
    /* KSP_DIVERGED_ITS is a linear-solve failure for SNES (SNESCheckKSPSolve) */
    if (its >= 10000) { status = 2; break; }
  }
done:
  *its_out = its;
  free(LU); free(diag); free(V); free(H); free(cs); free(sn); free(g); free(w); free(t);
  return status;
#undef PC_APPLY
}

/* ------------------------------------------------------------------ */
/* Newton (SNESNEWTONLS + SNESLINESEARCHBT + SNESConvergedDefault)      */
/* ------------------------------------------------------------------ */
typedef int (*fun_t)(ora_ctx *, const double *, double *, void *);
typedef int (*jac_t)(ora_ctx *, const double *, double *, void *);

typedef struct { const double *Uprev; double dt; } ts_aux;
static int ts_fun(ora_ctx *ctx, const double *U, double *R, void *a_) {
  ts_aux *a = a_;
  double *Ud = malloc(sizeof(double) * ctx->nc);
  for (int i = 0; i < ctx->nc; i++) Ud[i] = (U[i] - a->Uprev[i]) / a->dt;
  ora_ifunction(ctx, U, Ud, R);
  free(Ud);
  return 0;
}
static int ts_jac(ora_ctx *ctx, const double *U, double *J, void *a_) {
  ts_aux *a = a_;
  double *Ud = malloc(sizeof(double) * ctx->nc);
  for (int i = 0; i < ctx->nc; i++) Ud[i] = (U[i] - a->Uprev[i]) / a->dt;
  ora_ijacobian(ctx, U, Ud, 1.0 / a->dt, J);
  free(Ud);
  return 0;
}
static int sn_fun(ora_ctx *ctx, const double *U, double *R, void *a) { (void)a; return ora_snes_function(ctx, U, R); }
static int sn_jac(ora_ctx *ctx, const double *U, double *J, void *a) { (void)a; return ora_snes_jacobian(ctx, U, J); }

static int newton(ora_ctx *ctx, double *X, fun_t fun, jac_t jac, void *aux, int lin, double rtol, int *stats) {
  int n = ctx->nc;
  size_t nnz = ctx->j_rowptr[n];
  double *F = malloc(n * sizeof(double)), *Y = malloc(n * sizeof(double));
  double *W = malloc(n * sizeof(double)), *Gv = malloc(n * sizeof(double));
  double *Jv = malloc(nnz * sizeof(double)), *JY = malloc(n * sizeof(double));
  const double atol = 1e-50, stol = 1e-8, divtol = 1e4, alpha = 1e-4, minlambda = 1e-12, maxstep = 1e8;
  const int maxit = ctx->newton_maxit;
  fun(ctx, X, F, aux);
  double fnorm = nrm2c(ctx, n, F), fnorm0 = fnorm, ttol = rtol * fnorm;
  int status = 0, it = 0, linits = 0;
  if (!(fnorm == fnorm)) { status = -4; goto out; }
  if (fnorm < atol) goto out;
  for (it = 1; it <= maxit; it++) {
    jac(ctx, X, Jv, aux);
    int li = 0;
    int err = (lin == 0) ? dense_solve(ctx, Jv, F, Y) : gmres_ilu(ctx, Jv, F, Y, &li, lin == 2 ? ctx->nthreads : 0);
    linits += (lin == 0) ? 1 : li;
    if (err) { status = -3; break; }
    /* line search bt */
    double ynorm = nrm2c(ctx, n, Y);
    if (ynorm == 0.0) break; /* SNESLineSearchApply_BT: zero step -> converged on snorm */
    if (ynorm > maxstep) {
#pragma omp parallel for num_threads(ctx->nthreads) schedule(static)
      for (int i = 0; i < n; i++) Y[i] *= maxstep / ynorm;
      ynorm = maxstep;
    }
    /* initslope = -(F, J Y) */
    csr_mv(ctx, Jv, Y, JY);
    double initslope = vdot(ctx, n, F, JY);
    if (initslope > 0.0) initslope = -initslope;
    if (initslope == 0.0) initslope = -1.0;
    double lambda = 1.0;
    {
#pragma omp parallel for num_threads(ctx->nthreads) schedule(static)
      for (int i = 0; i < n; i++) W[i] = X[i] - lambda * Y[i];
    }
    fun(ctx, W, Gv, aux);
    double gnorm = nrm2c(ctx, n, Gv);
    if (!(0.5 * gnorm * gnorm <= 0.5 * fnorm * fnorm + lambda * alpha * initslope)) {
      double lambdaprev = lambda, gnormprev = gnorm;
      double lt = -initslope / (gnorm * gnorm - fnorm * fnorm - 2.0 * lambda * initslope);
      if (lt > 0.5 * lambda) lt = 0.5 * lambda;
      if (lt <= 0.1 * lambda) lambda = 0.1 * lambda; else lambda = lt;
      {
#pragma omp parallel for num_threads(ctx->nthreads) schedule(static)
      for (int i = 0; i < n; i++) W[i] = X[i] - lambda * Y[i];
    }
      fun(ctx, W, Gv, aux);
      gnorm = nrm2c(ctx, n, Gv);
      while (!(0.5 * gnorm * gnorm <= 0.5 * fnorm * fnorm + lambda * alpha * initslope)) {
        if (lambda <= minlambda) { status = -6; break; }
        double t1 = 0.5 * (gnorm * gnorm - fnorm * fnorm) - lambda * initslope;
        double t2 = 0.5 * (gnormprev * gnormprev - fnorm * fnorm) - lambdaprev * initslope;
        double a = (t1 / (lambda * lambda) - t2 / (lambdaprev * lambdaprev)) / (lambda - lambdaprev);
        double b = (-lambdaprev * t1 / (lambda * lambda) + lambda * t2 / (lambdaprev * lambdaprev)) / (lambda - lambdaprev);
        double d = b * b - 3 * a * initslope;
        if (d < 0.0) d = 0.0;
        if (a == 0.0) lt = -initslope / (2.0 * b); else lt = (-b + sqrt(d)) / (3.0 * a);
        lambdaprev = lambda; gnormprev = gnorm;
        if (lt > 0.5 * lambda) lt = 0.5 * lambda;
        if (lt <= 0.1 * lambda) lambda = 0.1 * lambda; else lambda = lt;
        {
#pragma omp parallel for num_threads(ctx->nthreads) schedule(static)
      for (int i = 0; i < n; i++) W[i] = X[i] - lambda * Y[i];
    }
        fun(ctx, W, Gv, aux);
        gnorm = nrm2c(ctx, n, Gv);
      }
      if (status) {
        /* SNESSolve_NEWTONLS: a failed line search whose step is below
         * stol*||x|| counts as SNES_CONVERGED_SNORM_RELATIVE, X unchanged */
        if (ynorm < stol * nrm2c(ctx, n, X)) status = 0;
        break;
      }
    }
    memcpy(X, W, n * sizeof(double));
    memcpy(F, Gv, n * sizeof(double));
    fnorm = gnorm;
    double xnorm = nrm2c(ctx, n, X);
    ynorm *= lambda;
    if (!(fnorm == fnorm)) { status = -4; break; }
    if (fnorm < atol) break;
    if (fnorm <= ttol) break;
    if (ynorm < stol * xnorm) break;
    if (fnorm > divtol * fnorm0) { status = -2; break; }
  }
  if (it > maxit) { status = -5; it = maxit; }
out:
  if (stats) { stats[0] += it; stats[1] += linits; }
  free(F); free(Y); free(W); free(Gv); free(Jv); free(JY);
  return status;
}

void ora_set_newton_maxit(ora_ctx *ctx, int maxit) { ctx->newton_maxit = maxit > 0 ? maxit : 50; }
void ora_set_threads(ora_ctx *ctx, int n) { ctx->nthreads = n > 0 ? n : 1; }

int ora_newton_snes(ora_ctx *ctx, double *U, int lin, double rtol, int *stats) {
  return newton(ctx, U, sn_fun, sn_jac, NULL, lin, rtol, stats);
}

/* TS backward Euler with fixed dt (TSBEULER, driven as in
 * demo/transient/transient_mpfaof90.F90:104-175) */
int ora_run_ts_be(ora_ctx *ctx, double *U, double dt, int nsteps, int lin, double rtol, int *stats) {
  double *Uprev = malloc(sizeof(double) * ctx->nc);
  int st = 0;
  if (stats) memset(stats, 0, 4 * sizeof(int));
  for (int s = 0; s < nsteps; s++) {
    memcpy(Uprev, U, sizeof(double) * ctx->nc);
    ts_aux a = {Uprev, dt};
    st = newton(ctx, U, ts_fun, ts_jac, &a, lin, rtol, stats);
    if (stats) stats[2]++;
    if (st) break;
  }
  if (stats) stats[3] = st;
  free(Uprev);
  return st;
}

/* fixed-dt SNES stepping (demo/transient/transient_snes_mpfaof90.F90:406-465) */
int ora_run_snes_steps(ora_ctx *ctx, double *U, double dt, int nsteps, int lin, double rtol, int *stats) {
  int st = 0;
  if (stats) memset(stats, 0, 4 * sizeof(int));
  ora_set_dtime(ctx, dt);
  for (int s = 0; s < nsteps; s++) {
    ora_snes_presolve(ctx, U);
    st = newton(ctx, U, sn_fun, sn_jac, NULL, lin, rtol, stats);
    if (stats) stats[2]++;
    if (st) break;
  }
  if (stats) stats[3] = st;
  return st;
}

/* TDyTimeIntegratorRunToTime (SNES branch), src/tdyti.c:71-196 with the
 * intended cut semantics: on a Newton failure dt *= 0.25 and the step is
 * retried from the previous solution (the reference's TDyTimeCut raises for
 * MPFA-O, src/tdycore.c:2139-2141). */
int ora_run_tdyti(ora_ctx *ctx, double *U, double dt_init, double growth, double dt_max, double final_time,
                  int lin, double rtol, int *stats) {
  double time = 0.0, dt = dt_init, dt_save = 0.0;
  int st = 0, icut = 0;
  double *Uprev = malloc(sizeof(double) * ctx->nc);
  if (stats) memset(stats, 0, 4 * sizeof(int));
  while (time < final_time) {
    double dt_for_sync = final_time - time;
    if (dt > dt_for_sync) { dt_save = dt; dt = dt_for_sync; }
    ora_set_dtime(ctx, dt);
    int conv = 0;
    while (!conv) {
      memcpy(Uprev, U, sizeof(double) * ctx->nc);
      ora_snes_presolve(ctx, U);
      st = newton(ctx, U, sn_fun, sn_jac, NULL, lin, rtol, stats);
      if (st < 0) {
        icut++;
        dt *= 0.25;
        ora_set_dtime(ctx, dt);
        memcpy(U, Uprev, sizeof(double) * ctx->nc);
        if (icut == 25) { free(Uprev); if (stats) stats[3] = st; return st; }
      } else {
        conv = 1;
        time += dt;
        if (stats) stats[2]++;
      }
    }
    if (dt_save > 0.) { time = final_time; dt = dt_save; dt_save = 0.; }
    dt *= growth;
    if (dt > dt_max) dt = dt_max;
  }
  free(Uprev);
  if (stats) stats[3] = 0;
  return 0;
}
