/*
 * CPU restatement of TDycore's MPFA-O Richards path (FP64), used ONLY as the
 * test oracle and as the CPU baseline of the benchmark.  Nothing in the
 * shipped library links, loads or calls this code.
 *
 * Functions follow the reference file:line noted beside each entry point in
 * tdy_oracle.c.  Parity pins are listed in the header comment of that file.
 */
#ifndef TDY_ORACLE_H
#define TDY_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_ctx ora_ctx;

/* params[] indices */
enum {
  ORA_P_VG_M = 0, ORA_P_VG_ALPHA, ORA_P_X0, ORA_P_X1, ORA_P_X2, ORA_P_DX,
  ORA_P_GRAVITY, ORA_P_DENSITY, ORA_P_GMATRIX, ORA_P_PREF, ORA_P_NPARAM
};

/* face_bc codes: -1 internal, 0 Dirichlet, 1 Neumann (no flow), 2 seepage */
ora_ctx *ora_create(int nc, int nf, int nv, int nbnd,
                    const int *cell_verts, const int *cell_faces,
                    const int *face_verts, const int *face_cells,
                    const double *vert_xyz, const double *cell_cen,
                    const double *cell_vol, const double *face_cen,
                    const double *face_area, const double *face_normal,
                    const int *face_bc, const double *K, const double *porosity,
                    const double *Sr, const double *params,
                    const double *p_bnd_fn /* nbnd or NULL */);
void ora_destroy(ora_ctx *ctx);
const char *ora_last_error(void);

/* sizes: [0]=trans nnz, [1]=jac nnz, [2]=max IR cells, [3]=max IR faces */
void ora_sizes(const ora_ctx *ctx, long long *out);
void ora_trans_csr(const ora_ctx *ctx, int *rowptr, int *cols, double *vals);
void ora_gravdis(const ora_ctx *ctx, double *out);
void ora_jac_pattern(const ora_ctx *ctx, int *rowptr, int *cols);

/* state: which = 0 S,1 dS_dP,2 d2S_dP2,3 Kr,4 dKr_dS,5 rho,6 drho_dP,
 * 7 d2rho_dP2, 8 vis */
void ora_update_state(ora_ctx *ctx, const double *p);
void ora_get_state(const ora_ctx *ctx, int which, double *out);
void ora_boundary_pressure(const ora_ctx *ctx, double *out);

/* TS form (tdympfao_ts.c) */
int ora_ifunction(ora_ctx *ctx, const double *U, const double *Udot, double *R);
int ora_ijacobian(ora_ctx *ctx, const double *U, const double *Udot, double shift,
                  double *Jvals);
/* SNES form (tdympfao_snes.c) */
void ora_set_dtime(ora_ctx *ctx, double dt);
int ora_snes_presolve(ora_ctx *ctx, const double *Uprev);
int ora_snes_function(ora_ctx *ctx, const double *U, double *R);
int ora_snes_jacobian(ora_ctx *ctx, const double *U, double *Jvals);

/* threads of every parallel loop (OpenMP; default 1).  Results do not depend on
 * it except through the order of the Krylov solver's dot products. */
void ora_set_threads(ora_ctx *ctx, int n);

/* drivers.  lin: 0 dense LU, 1 GMRES(30)+ILU(0) in natural order, 2
 * GMRES(30)+block-Jacobi ILU(0) with one contiguous row block per thread
 * (PETSc's parallel default).  stats (int[4]):
 * newton iterations, linear iterations, steps, status (0 ok) */
int ora_run_ts_be(ora_ctx *ctx, double *U, double dt, int nsteps, int lin,
                  double snes_rtol, int *stats);
int ora_run_snes_steps(ora_ctx *ctx, double *U, double dt, int nsteps, int lin,
                       double snes_rtol, int *stats);
int ora_run_tdyti(ora_ctx *ctx, double *U, double dt_init, double growth,
                  double dt_max, double final_time, int lin, double snes_rtol,
                  int *stats);
/* one Newton solve of the SNES form at the current dt/accum_prev, starting
 * from U (in/out): returns status, fills stats */
int ora_newton_snes(ora_ctx *ctx, double *U, int lin, double snes_rtol, int *stats);
/* SNES max_it for every driver (default 50; at the cap status is -5) */
void ora_set_newton_maxit(ora_ctx *ctx, int maxit);

void ora_liquid_mass(ora_ctx *ctx, const double *U, double *out);
/* source_sink per cell: the TS residual subtracts Q*V, the SNES residual Q
 * (tdympfao_ts.c:211, tdympfao_snes.c:122); NULL clears it */
void ora_set_source_sink(ora_ctx *ctx, const double *Q);

#ifdef __cplusplus
}
#endif
#endif
