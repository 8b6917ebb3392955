/*
 * tdyhip.h -- C ABI of the MI355X MPFA-O Richards path.
 *
 * Drop-in boundary: the entry points below replace TDycore's MPFA-O
 * callbacks and setup for the Richards mode.  All array arguments of the
 * evaluation calls are DEVICE pointers (HBM); setup takes HOST arrays.
 * Every function returns 0 on success and a nonzero TDYHIP_ERR_* code on
 * failure (the PetscErrorCode convention); tdyhip_last_error() describes
 * the last failure of the calling thread.  No call retains a caller pointer
 * past its return.  Calls on one context are not thread-safe (the reference
 * callbacks are called synchronously from SNES/TS, src/tdycore.c:1892-2034).
 */
#ifndef TDYHIP_H
#define TDYHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tdyhip_ctx tdyhip_ctx;
typedef void *tdyhip_stream; /* a hipStream_t; NULL = the context's stream */

enum {
  TDYHIP_OK = 0,
  TDYHIP_ERR_ARG = 1,      /* bad argument / unsupported mesh */
  TDYHIP_ERR_HIP = 2,      /* HIP runtime failure */
  TDYHIP_ERR_SINGULAR = 3, /* singular interaction-region system (PETSC_ERR_MAT_LU_ZRPVT) */
  TDYHIP_ERR_NOCONV = 4,   /* Newton / Krylov did not converge */
  TDYHIP_ERR_MEM = 5,      /* device memory exhausted */
  TDYHIP_ERR_CANCELLED = 6 /* the solve was stopped by tdyhip_cancel */
};

/* boundary types per face (face_bc): -1 internal, 0 Dirichlet, 1 Neumann
 * (no flow), 2 seepage.  One type on every boundary face is the reference's
 * global -tdy_mpfao_boundary_condition_type (src/fv/mpfao/tdympfao.c:588-596). */
enum { TDYHIP_BC_INTERNAL = -1, TDYHIP_BC_DIRICHLET = 0, TDYHIP_BC_NEUMANN = 1, TDYHIP_BC_SEEPAGE = 2 };

/* Mesh in the form TDyMeshCreateFromPlex produces (src/fv/share/tdymeshplex.c:3685):
 * hexahedra, 8 vertices per cell, 6 faces per cell, 4 vertices per face (right prisms:
 * 6 vertices, 5 faces and triangles of 3 vertices, each array padded with -1; see
 * tdyhip_setup_info)
 * (cyclic), face_cells = (up, dn) with the unit normal from up to dn and a
 * boundary face stored as (-(b+1), cell).  Host pointers. */
typedef struct {
  int32_t nc, nf, nv, nbnd;
  const int32_t *cell_verts;  /* nc*8 */
  const int32_t *cell_faces;  /* nc*6 */
  const int32_t *face_verts;  /* nf*4 */
  const int32_t *face_cells;  /* nf*2 */
  const int32_t *face_bc;     /* nf */
  const double *vert_xyz;     /* nv*3 */
  const double *cell_cen;     /* nc*3 */
  const double *cell_vol;     /* nc   */
  const double *face_cen;     /* nf*3 */
  const double *face_area;    /* nf   */
  const double *face_normal;  /* nf*3 */
} tdyhip_mesh;

/* Material/constitutive inputs (InitMaterials, src/fv/mpfao/tdympfao.c:630-763;
 * option defaults src/tdycore.c:260-296).  Host pointers. */
typedef struct {
  double vg_m, vg_alpha;                       /* van Genuchten m, alpha [1/Pa] */
  double poly_x0, poly_x1, poly_x2, poly_dx;   /* Mualem cubic smoothing */
  double gravity;                              /* |g| (z axis) */
  double pref;                                 /* reference pressure (101325) */
  int32_t density_type;                        /* 0 constant, 1 exponential */
  int32_t gmatrix_method;                      /* 0 MPFAO, 1 TPF */
  const double *K;                             /* nc*9 row-major tensors */
  const double *porosity;                      /* nc */
  const double *Sr;                            /* nc residual saturation */
  const double *p_bnd_fn;                      /* nbnd boundary-pressure function values
                                                  at boundary face centroids, or NULL
                                                  (then P_bnd = p of the adjacent cell) */
} tdyhip_materials;

const char *tdyhip_last_error(void);

/* Setup: TDySetup_Richards_MPFAO (src/fv/mpfao/tdympfao.c:2277-2331).
 * Topology on the host, then the G matrices, the per-vertex local systems
 * (batched LU) and the gravity term on the device. */
int tdyhip_create(const tdyhip_mesh *mesh, const tdyhip_materials *mat, tdyhip_ctx **out);
int tdyhip_destroy(tdyhip_ctx *ctx);
/* The context's stream: by default one it created for itself (non-blocking, default
 * priority) and never shares with another context.  A stream given to set_stream
 * must not be shared with another context either: each solver call may capture its
 * launches on the stream into a hipGraph. */
int tdyhip_set_stream(tdyhip_ctx *ctx, tdyhip_stream stream);
int tdyhip_get_stream(const tdyhip_ctx *ctx, tdyhip_stream *stream);
int tdyhip_sizes(const tdyhip_ctx *ctx, int64_t *out /* [0]=nc [1]=jac nnz [2]=nf [3]=nbnd */);

/* Jacobian sparsity (vertex adjacency, TDyDiscretizationCreateJacobianMatrix,
 * src/tdydiscretization.c:156-181): HOST arrays rowptr[nc+1], cols[nnz]. */
int tdyhip_jacobian_pattern(const tdyhip_ctx *ctx, int32_t *rowptr, int32_t *cols);
/* The same pattern already on the device (read-only, owned by ctx). */
int tdyhip_jacobian_pattern_device(const tdyhip_ctx *ctx, const int32_t **rowptr, const int32_t **cols);

/* TS residual: replaces TDyMPFAOIFunction (src/fv/mpfao/tdympfao_ts.c:153). R is overwritten. */
int tdyhip_ifunction(tdyhip_ctx *ctx, double t, const double *U, const double *Udot, double *R,
                     tdyhip_stream stream);
/* The TS residual of tdyhip_ifunction, allocating nothing: no counterpart in the reference.
 * It is for callers that cannot allocate device memory (code under hipGraph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The same kernels and the same result, bit
 * for bit, but the per-flux-row fluxes go to a fixed scratch array in device memory that the
 * library holds for the life of the process, one per GPU, instead of the context's own buffer.
 * Meshes of at most 2,097,152 faces (8,388,608 flux rows, 64 MiB); larger meshes get
 * TDYHIP_ERR_ARG.  Every call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (from two streams or two contexts) race. */
int tdyhip_ifunction_noalloc(tdyhip_ctx *ctx, double t, const double *U, const double *Udot, double *R,
                             tdyhip_stream stream);
/* TS Jacobian: replaces TDyMPFAOIJacobian (src/fv/mpfao/tdympfao_ts.c:520).  Uses the
 * state of the preceding tdyhip_ifunction, like the reference.  Jvals (nnz, in the
 * pattern order) is overwritten. */
int tdyhip_ijacobian(tdyhip_ctx *ctx, double t, const double *U, const double *Udot, double shift,
                     double *Jvals, tdyhip_stream stream);
/* TS residual and Jacobian from one pass over the transmissibilities. */
int tdyhip_ifunction_ijacobian(tdyhip_ctx *ctx, double t, const double *U, const double *Udot,
                               double shift, double *R, double *Jvals, tdyhip_stream stream);

/* The flow equation's source_sink per cell (mpfao->source_sink, allocated zero by
 * AllocateMemoryForSourceSinkValues, src/fv/mpfao/tdympfao.c:364-384): the TS residual
 * subtracts Q*V (tdympfao_ts.c:211), the SNES residual Q (tdympfao_snes.c:122).  HOST array
 * of nc entries, copied; NULL restores zero.  The reference fills it only through a
 * forcing function (TDySetForcingFunction, src/tdycore.c:1314); manufactured solutions
 * need it. */
int tdyhip_set_source_sink(tdyhip_ctx *ctx, const double *Q);

/* SNES form: TDySetDtimeForSNESSolver (src/tdycore.c:1575), TDyMPFAOSNESPreSolve,
 * TDyMPFAOSNESFunction, TDyMPFAOSNESJacobian (src/fv/mpfao/tdympfao_snes.c:32,67,143). */
int tdyhip_set_dtime(tdyhip_ctx *ctx, double dt);
int tdyhip_snes_presolve(tdyhip_ctx *ctx, const double *Uprev, tdyhip_stream stream);
int tdyhip_snes_function(tdyhip_ctx *ctx, const double *U, double *R, tdyhip_stream stream);
int tdyhip_snes_jacobian(tdyhip_ctx *ctx, const double *U, double *Jvals, tdyhip_stream stream);

/* Diagnostics: TDyUpdateDiagnostics_MPFAO (src/fv/mpfao/tdympfao.c:2759-2786) -- liquid
 * saturation S and the reference's liquid mass rho*V per cell -- plus the water mass
 * rho*phi*S*V per cell (for mass balances).  Device arrays of nc entries; any output
 * may be NULL. */
int tdyhip_diagnostics(tdyhip_ctx *ctx, const double *U, double *S, double *liquid_mass, double *water_mass,
                       tdyhip_stream stream);

/* Velocity recovery: TDyMPFAORecoverVelocity (src/fv/mpfao/tdympfao_utils.c:661-688).  With the
 * cell potential Phi = p + rho(p)*|g|*z_c, every flux row of face f adds -(T_row . Phi)/A_f to
 * vel[f] and 1 to count[f] (the reference's vel and vel_count, tdympfao_utils.c:103-217).  A
 * boundary column takes the boundary-pressure function value (no gravity term) when one is set,
 * else Phi of the face's cell (tdympfao_utils.c:314-324).  Boundary faces sum all their sub-face
 * rows, as interior faces do; the reference fills them only from the single-cell corner vertices
 * (534-659).  No-flow faces get 0.  The reference's error norm against a prescribed velocity
 * field (MMS runs) is not computed.  Device arrays of nf entries; count may be NULL. */
int tdyhip_recover_velocity(tdyhip_ctx *ctx, const double *U, double *vel, int32_t *count, tdyhip_stream stream);

/* ---- Newton / Krylov (PETSc SNESNEWTONLS + KSPGMRES + PCBJACOBI/ILU(0) defaults) ---- */
typedef struct {
  double snes_rtol, snes_atol, snes_stol; /* 1e-8, 1e-50, 1e-8 */
  int32_t snes_max_it;                    /* 50 */
  double ksp_rtol, ksp_atol;              /* 1e-5, 1e-50 (preconditioned residual) */
  int32_t ksp_max_it, ksp_restart;        /* 10000, 30 */
  int32_t krylov_fp32;                    /* 0: FP64 Jacobian + Krylov; 1: FP32 Jacobian copy,
                                             FP32 ILU(0) and Krylov vectors, FP32 accumulation,
                                             FP64 repeat of a failed FP32 solve */
  int32_t pc_block;                       /* rows per block-Jacobi ILU(0) block, 1..512 */
  int32_t hip_graphs;                     /* 1: replay each GMRES restart cycle as a captured
                                             hipGraph (not while kernel timers run); 0: launch */
  int32_t pc_coarse;                      /* 1: add a coarse correction to block Jacobi (one value
                                             per block, blocks aggregated into at most 128 groups
                                             solved exactly); not in the reference, whose PETSc
                                             default is one-level block Jacobi.  Used only when the
                                             mesh fits its limits (<= 32 neighbour blocks per block,
                                             <= 8192 blocks).  Default 0. */
} tdyhip_solver_opts;

typedef struct {
  int64_t newton_its, linear_its, steps, cuts;
  int32_t status;                         /* 0 converged, <0 SNESConvergedReason-like failure */
  int32_t krylov_fallbacks;               /* FP32 linear solves repeated in FP64 (krylov_fp32 = 1:
                                             an FP32 GMRES gets at most 1000 iterations; if it does
                                             not converge the correction is solved again in FP64) */
  double last_fnorm;
} tdyhip_solver_stats;

void tdyhip_solver_defaults(tdyhip_solver_opts *o);
/* Block-Jacobi partition: block id per cell (HOST array, nc entries, every block at
 * most pc_block cells), the role of PCBJacobiSetLocalBlocks; NULL restores the
 * default breadth-first partition over the Jacobian graph. */
int tdyhip_set_pc_partition(tdyhip_ctx *ctx, const int32_t *block_of_cell);
/* One SNESSolve of the SNES form at the current dt / accumulation (U device, in/out). */
int tdyhip_newton_snes(tdyhip_ctx *ctx, double *U, const tdyhip_solver_opts *o, tdyhip_solver_stats *st);
/* Fixed-dt SNES stepping: per step TDyMPFAOSNESPreSolve, SNESSolve, U_prev = U. */
int tdyhip_run_snes_steps(tdyhip_ctx *ctx, double *U, double dt, int32_t nsteps, const tdyhip_solver_opts *o,
                          tdyhip_solver_stats *st);
/* TS backward Euler with fixed dt (TSBEULER over TDyMPFAOIFunction/IJacobian). */
int tdyhip_run_ts_be(tdyhip_ctx *ctx, double *U, double dt, int32_t nsteps, const tdyhip_solver_opts *o,
                     tdyhip_solver_stats *st);
/* TDyTimeIntegratorRunToTime (src/tdyti.c:100-196): dt growth/cap, sync to final time; a
 * failed Newton solve cuts dt by 4 and retries from the previous solution (at most 25 cuts). */
int tdyhip_run_tdyti(tdyhip_ctx *ctx, double *U, double dt_init, double growth, double dt_max,
                     double final_time, const tdyhip_solver_opts *o, tdyhip_solver_stats *st);
/* Cancellation (on = 1) of the solve running on ctx: tdyhip_newton_snes, run_snes_steps,
 * run_ts_be and run_tdyti return TDYHIP_ERR_CANCELLED at their next Newton iteration or
 * GMRES restart, with U part-way.  Callable from any thread while the solve runs; the
 * request stays set until tdyhip_cancel(ctx, 0), so a solve started while it is set stops
 * at its first check.  No counterpart in the reference (PETSc solves run to their own
 * criteria): it is for servers that bound the time a request may hold the device. */
int tdyhip_cancel(tdyhip_ctx *ctx, int32_t on);

/* Fused serving batch.  No counterpart in the reference (PETSc solves one problem per SNES):
 * it is for servers running an ensemble of requests at once.  A batch holds up to 32 requests,
 * each a context of a hexahedral mesh within the fixed-size limits with its block-Jacobi
 * partition set, its device state U (in/out) and the parameters of
 * TDyTimeIntegratorRunToTime (src/tdyti.c:100-196, as tdyhip_run_tdyti).  Each round advances
 * every request by one Newton iteration (or starts its next time step): the state, residual
 * and Jacobian evaluations of all requests run as one segmented launch each, the Krylov
 * solves side by side on the requests' own streams.  A request's results are the same bit
 * for bit alone, in any mix, and through tdyhip_run_tdyti.  Requests may be added between
 * rounds; a request that reaches its final time or fails leaves the batch in that round and
 * is reported (id, rc: 0 or the error code tdyhip_run_tdyti would return, its stats; the
 * last error message is that of the last failed request).  The caller keeps ownership of the
 * contexts and states and must not use them while they are in the batch; cancellation
 * (tdyhip_cancel) ends a request at its next round with TDYHIP_ERR_CANCELLED. */
/* Evaluation fuser.  No counterpart in the reference: it is for servers solving many requests at
 * once, each with its own thread and solver (tdyhip_run_tdyti and the other drivers).  The
 * SNES-form residual and Jacobian evaluations of every context attached to a fuser are launched
 * by the fuser's thread, which runs all evaluations pending when it wakes as one segmented launch
 * per kernel; each context keeps its own pace and its results are the same bit for bit as
 * without the fuser.  Attached contexts run their Newton iterations without whole-iteration
 * graphs.  Only hexahedral meshes within the fixed-size limits can be attached; NULL detaches.
 * Destroy a fuser only after every context attached to it is detached or destroyed.
 * stats[3] = fused launch groups, evaluations, largest group. */
typedef struct tdyhip_fuser tdyhip_fuser;
int tdyhip_fuser_create(tdyhip_fuser **out);
int tdyhip_fuser_destroy(tdyhip_fuser *fz);
int tdyhip_set_fuser(tdyhip_ctx *ctx, tdyhip_fuser *fz);
int tdyhip_fuser_stats(const tdyhip_fuser *fz, int64_t *stats /* 3 */);
/* The first pending evaluation waits up to `us` microseconds (default 0) for others, or until
 * every attached context has one pending, before the group is launched. */
int tdyhip_fuser_set_window(tdyhip_fuser *fz, int32_t us);

typedef struct tdyhip_batch tdyhip_batch;
int tdyhip_batch_create(tdyhip_batch **out);
int tdyhip_batch_destroy(tdyhip_batch *b);
int tdyhip_batch_add(tdyhip_batch *b, tdyhip_ctx *ctx, double *U, const tdyhip_solver_opts *o, double dt_init,
                     double growth, double dt_max, double final_time, int32_t *id);
int tdyhip_batch_remove(tdyhip_batch *b, int32_t id);
int tdyhip_batch_size(const tdyhip_batch *b, int32_t *n);
/* cap: length of ids, rcs and stats (at least the batch size); ndone: entries written */
int tdyhip_batch_round(tdyhip_batch *b, int32_t cap, int32_t *ids, int32_t *rcs, tdyhip_solver_stats *stats,
                       int32_t *ndone);

/* Device memory held by the context (setup arrays plus the solver workspace), in
 * bytes: what an HBM-footprint admission policy accounts per problem.  No
 * counterpart in the reference (PETSc's PetscMallocGetCurrentUsage plays that
 * role on the CPU). */
int tdyhip_device_bytes(const tdyhip_ctx *ctx, int64_t *bytes);

/* Kernel timers: HIP events around every launch of the path on its stream.
 * enable=1 resets and starts, 0 stops.  The report is a JSON array of
 * {kernel, launches, ms, bytes} (bytes: algorithmic bytes, each array counted
 * once per launch). */
int tdyhip_kernel_timers(tdyhip_ctx *ctx, int32_t enable);
int tdyhip_kernel_timer_report(tdyhip_ctx *ctx, char *buf, int64_t buflen);

/* Diagnostics: transmissibility (rows face*4+subface, columns cell | nc+bnd) and the
 * gravity term, copied to HOST arrays, for inspection and parity tests. */
int tdyhip_trans_nnz(const tdyhip_ctx *ctx, int64_t *nnz);
int tdyhip_trans_csr(const tdyhip_ctx *ctx, int32_t *rowptr, int32_t *cols, double *vals);
int tdyhip_gravdis(const tdyhip_ctx *ctx, double *out /* 4*nf */);
/* Setup of a general mesh (any cell other than a hexahedron: right prisms, given in the
 * hexahedral arrays padded with -1; and hexahedral meshes with more than 8 cells, 12 faces or
 * 8 transmissibility columns at a vertex or more than 27 Jacobian entries in a row).  Replaces
 * ComputeTransmissibilityMatrix and ComputeGravityDiscretization for vertices of any degree
 * (src/fv/mpfao/tdympfao.c:1801-1860, 1126-1513, 2169-2274); tdyhip_create runs it.  Every other
 * call works on such a context (rows of up to 64 Jacobian entries); its Newton solver needs a
 * block-Jacobi partition with at most 27 in-block entries per row (tdyhip_set_pc_partition, e.g.
 * vertical columns).
 * Setup report of any mesh: ms[3] = device times (HIP events) of the G matrices, the local
 * systems and the gravity term; counts[11] = interaction regions in the 16-, 32- and 64-lane
 * buckets (hexahedral meshes: all in the first), the most cells and faces at one vertex, the
 * transmissibility entries, the algorithmic FP64 flops and bytes of the local systems, and the
 * team width used for each bucket (0 for an empty bucket). */
int tdyhip_setup_info(const tdyhip_ctx *ctx, double *ms /* 3 */, int64_t *counts /* 11 */);
/* The autotuned plan of a general mesh's interaction-region launch (no counterpart in the
 * reference, whose solves run one region at a time: tdympfao.c:1126-1513).  plan[5] = team width
 * of each bucket (0: empty), 1 when the buckets are drained largest regions first, work-item
 * chunk in workgroup steps; ms[13] = the tuning runs (ms, 0 when not run): bucket x width
 * 16/32/64 (9), then order (smallest / largest first) x chunk (1 / 8) (4).  Plans are cached per
 * device and bucket shapes, so a later setup of the same shapes runs no tuning.  A hexahedral
 * mesh within the fixed-size limits has one kernel and reports plan = {16, 0, 0, 0, 1}. */
int tdyhip_setup_plan(const tdyhip_ctx *ctx, int32_t *plan /* 5 */, double *ms /* 13 */);

#ifdef __cplusplus
}
#endif
#endif
