/*
 * tdympfao_hip.c -- reference-side binding of libtdyhip (include/tdyhip.h) for a
 * PETSc-built TDycore: the MPFA-O Richards TS and SNES callbacks with the
 * reference's exact signatures, running on one MI355X.
 *
 * This file is not part of this library's build and was not compiled here:
 * PETSc is absent from this image.  It is what a TDycore maintainer adds to the
 * reference tree as src/fv/mpfao/tdympfao_hip.c, compiled with -DTDYCORE_HIP and
 * linked with -ltdyhip -lamdhip64; the registration changes it needs in
 * src/tdycore.c are listed at the end of this file.  PETSc 3.17 API (the
 * version the reference pins, tools/Dockerfile.petsc:41).  Serial (one rank):
 * the library holds the whole mesh.
 *
 * Vectors and matrices stay the reference's host objects (VECSEQ, MATSEQAIJ):
 * each callback stages its vectors through device buffers owned here, on the
 * library's own stream, and installs the Jacobian with MatSetValuesCOO in the
 * vertex-adjacency pattern the library reports (the pattern
 * TDyCreateJacobian builds, src/tdydiscretization.c:156-181).  A caller that
 * keeps everything on the device runs the whole solve through
 * tdyhip_run_tdyti / tdyhip_newton_snes instead.
 */
#include <private/tdycoreimpl.h>
#include <private/tdymeshimpl.h>
#include <private/tdympfaoimpl.h>
#include <private/tdyconditionsimpl.h>

#include <hip/hip_runtime_api.h>
#include <tdyhip.h>

typedef struct {
  tdyhip_ctx *g;
  hipStream_t stream;  /* the context's own stream: every copy is ordered on it */
  PetscInt nc, nnz;
  PetscInt *coo_i, *coo_j;
  double *d_u, *d_ut, *d_r, *d_j; /* device staging: U, U_t, R (nc), Jacobian values (nnz) */
  double *h_j;                    /* host copy of the Jacobian values (pattern order) */
} TDyHip;

static const char kKey[] = "TDyHip";

static PetscErrorCode hipchk(int rc) {
  if (rc) SETERRQ(PETSC_COMM_SELF, PETSC_ERR_LIB, "tdyhip error %d: %s", rc, tdyhip_last_error());
  return 0;
}

static PetscErrorCode hiprtchk(hipError_t e) {
  if (e != hipSuccess) SETERRQ(PETSC_COMM_SELF, PETSC_ERR_LIB, "HIP: %s", hipGetErrorString(e));
  return 0;
}

static PetscErrorCode TDyHipFree(void *p) {
  TDyHip *h = (TDyHip *)p;
  if (!h) return 0;
  if (h->g) tdyhip_destroy(h->g);
  (void)hipFree(h->d_u); (void)hipFree(h->d_ut); (void)hipFree(h->d_r); (void)hipFree(h->d_j);
  PetscFree(h->h_j);
  PetscFree(h->coo_i);
  PetscFree(h->coo_j);
  return PetscFree(h);
}

/* The TDyHip of a TDy, composed on the TDy object (a PetscObject through
 * PETSCHEADER, include/private/tdycoreimpl.h:82-83): struct _p_TDy itself is
 * not changed. */
static PetscErrorCode TDyHipGet(TDy tdy, TDyHip **h) {
  PetscContainer c = NULL;
  PetscFunctionBegin;
  PetscCall(PetscObjectQuery((PetscObject)tdy, kKey, (PetscObject *)&c));
  if (!c) SETERRQ(PETSC_COMM_SELF, PETSC_ERR_ARG_WRONGSTATE, "TDyHipSetup has not run on this TDy");
  PetscCall(PetscContainerGetPointer(c, (void **)h));
  PetscFunctionReturn(0);
}

/* TDyMesh (include/private/tdymeshimpl.h:35-219) -> tdyhip_mesh, and the
 * materials of TDyMPFAO (include/private/tdympfaoimpl.h:13-117) ->
 * tdyhip_materials.  The TDyMesh already has what the library expects:
 *   - a boundary face stores cell_ids = {-(b+1), cell} with its normal pointing
 *     from cell_ids[0] to cell_ids[1] (tdymeshplex.c:416-424, 3126-3168);
 *   - face vertex order is the one the flux rows face*4 + local vertex use
 *     (tdympfao.c:1451-1455);
 *   - hexahedra have 8 vertices and 6 faces, right prisms 6 and 5 with
 *     triangular faces of 3 vertices (padded with -1 here).
 * Boundary pressures are the BC function at the boundary-face centroids when
 * one is set (TDyMPFAO_SetBoundaryPressure, tdympfao_utils.c:690-759), else
 * NULL (the library then uses the adjacent cell's pressure, as the reference). */
typedef struct {
  int32_t *cell_verts, *cell_faces, *face_verts, *face_cells, *face_bc;
  double *vert_xyz, *cell_cen, *cell_vol, *face_cen, *face_area, *face_normal, *p_bnd;
} MeshArrays;

static PetscErrorCode MeshArraysFree(MeshArrays *a) {
  PetscFunctionBegin;
  PetscCall(PetscFree(a->cell_verts)); PetscCall(PetscFree(a->cell_faces)); PetscCall(PetscFree(a->face_verts));
  PetscCall(PetscFree(a->face_cells)); PetscCall(PetscFree(a->face_bc)); PetscCall(PetscFree(a->vert_xyz));
  PetscCall(PetscFree(a->cell_cen)); PetscCall(PetscFree(a->cell_vol)); PetscCall(PetscFree(a->face_cen));
  PetscCall(PetscFree(a->face_area)); PetscCall(PetscFree(a->face_normal)); PetscCall(PetscFree(a->p_bnd));
  PetscFunctionReturn(0);
}

static PetscErrorCode TDyHipMesh(TDy tdy, MeshArrays *a, tdyhip_mesh *m, tdyhip_materials *mat) {
  TDyMPFAO *mpfao = (TDyMPFAO *)tdy->context;
  TDyMesh *mesh = mpfao->mesh;
  TDyCell *cells = &mesh->cells;
  TDyFace *faces = &mesh->faces;
  TDyVertex *verts = &mesh->vertices;
  const PetscInt nc = mesh->num_cells, nf = mesh->num_faces, nv = mesh->num_vertices;
  const PetscInt nbnd = mesh->num_boundary_faces;
  PetscFunctionBegin;
  PetscCall(PetscMalloc1(8 * nc, &a->cell_verts)); PetscCall(PetscMalloc1(6 * nc, &a->cell_faces));
  PetscCall(PetscMalloc1(4 * nf, &a->face_verts)); PetscCall(PetscMalloc1(2 * nf, &a->face_cells));
  PetscCall(PetscMalloc1(nf, &a->face_bc)); PetscCall(PetscMalloc1(3 * nv, &a->vert_xyz));
  PetscCall(PetscMalloc1(3 * nc, &a->cell_cen)); PetscCall(PetscMalloc1(nc, &a->cell_vol));
  PetscCall(PetscMalloc1(3 * nf, &a->face_cen)); PetscCall(PetscMalloc1(nf, &a->face_area));
  PetscCall(PetscMalloc1(3 * nf, &a->face_normal));
  a->p_bnd = NULL;
  for (PetscInt c = 0; c < nc; c++) {
    const PetscInt ncv = cells->num_vertices[c], ncf = cells->num_faces[c];
    if (!((ncv == 8 && ncf == 6) || (ncv == 6 && ncf == 5)))
      SETERRQ(PETSC_COMM_SELF, PETSC_ERR_SUP, "cell %" PetscInt_FMT ": only hexahedra and right prisms", c);
    for (PetscInt k = 0; k < 8; k++) a->cell_verts[8 * c + k] = k < ncv ? (int32_t)cells->vertex_ids[cells->vertex_offset[c] + k] : -1;
    for (PetscInt k = 0; k < 6; k++) a->cell_faces[6 * c + k] = k < ncf ? (int32_t)cells->face_ids[cells->face_offset[c] + k] : -1;
    for (PetscInt d = 0; d < 3; d++) a->cell_cen[3 * c + d] = cells->centroid[c].X[d];
    a->cell_vol[c] = cells->volume[c];
  }
  for (PetscInt f = 0; f < nf; f++) {
    const PetscInt nfv = faces->num_vertices[f];
    if (nfv != 3 && nfv != 4) SETERRQ(PETSC_COMM_SELF, PETSC_ERR_SUP, "face with %" PetscInt_FMT " vertices", nfv);
    for (PetscInt k = 0; k < 4; k++) a->face_verts[4 * f + k] = k < nfv ? (int32_t)faces->vertex_ids[faces->vertex_offset[f] + k] : -1;
    a->face_cells[2 * f] = (int32_t)faces->cell_ids[faces->cell_offset[f] + 0];
    a->face_cells[2 * f + 1] = (int32_t)faces->cell_ids[faces->cell_offset[f] + 1];
    /* one boundary type for every boundary face, as the reference (tdympfao.c:588-596):
     * DIRICHLET_BC 0, NEUMANN_BC 1, SEEPAGE_BC 2 are the library's codes too */
    a->face_bc[f] = faces->is_internal[f] ? TDYHIP_BC_INTERNAL : (int32_t)mpfao->bc_type;
    for (PetscInt d = 0; d < 3; d++) {
      a->face_cen[3 * f + d] = faces->centroid[f].X[d];
      a->face_normal[3 * f + d] = faces->normal[f].V[d];
    }
    a->face_area[f] = faces->area[f];
  }
  for (PetscInt v = 0; v < nv; v++)
    for (PetscInt d = 0; d < 3; d++) a->vert_xyz[3 * v + d] = verts->coordinate[v].X[d];
  if (ConditionsHasBoundaryPressure(tdy->conditions)) {
    PetscCall(PetscMalloc1(nbnd, &a->p_bnd));
    for (PetscInt f = 0; f < nf; f++) {
      if (faces->is_internal[f]) continue;
      const PetscInt *ci = &faces->cell_ids[faces->cell_offset[f]];
      const PetscInt b = ci[0] < 0 ? -ci[0] - 1 : -ci[1] - 1;
      PetscCall(ConditionsComputeBoundaryPressure(tdy->conditions, 1, faces->centroid[f].X, &a->p_bnd[b]));
    }
  }
  m->nc = (int32_t)nc; m->nf = (int32_t)nf; m->nv = (int32_t)nv; m->nbnd = (int32_t)nbnd;
  m->cell_verts = a->cell_verts; m->cell_faces = a->cell_faces; m->face_verts = a->face_verts;
  m->face_cells = a->face_cells; m->face_bc = a->face_bc; m->vert_xyz = a->vert_xyz;
  m->cell_cen = a->cell_cen; m->cell_vol = a->cell_vol; m->face_cen = a->face_cen;
  m->face_area = a->face_area; m->face_normal = a->face_normal;
  /* materials: InitMaterials (tdympfao.c:630-763) already filled these per cell */
  mat->vg_m = mpfao->vangenuchten_m;
  mat->vg_alpha = mpfao->vangenuchten_alpha;
  mat->poly_x0 = mpfao->mualem_poly_x0;
  mat->poly_x1 = mpfao->mualem_poly_x1;
  mat->poly_x2 = mpfao->mualem_poly_x2;
  mat->poly_dx = mpfao->mualem_poly_dx;
  mat->gravity = PetscAbsReal(mpfao->gravity[2]);  /* g = (0, 0, +|g|), tdympfao.c:619 */
  mat->pref = mpfao->Pref;
  mat->density_type = tdy->eos.density_type == WATER_DENSITY_EXPONENTIAL ? 1 : 0;
  mat->gmatrix_method = mpfao->gmatrix_method == MPFAO_GMATRIX_TPF ? 1 : 0;
  mat->K = mpfao->K0;
  mat->porosity = mpfao->porosity;
  mat->Sr = mpfao->Sr;
  mat->p_bnd_fn = a->p_bnd;
  PetscFunctionReturn(0);
}

/* Creates the library context for a TDy set up with RICHARDS and MPFA_O
 * (after TDySetup_Richards_MPFAO, tdympfao.c:2277-2331), the device staging
 * buffers, and installs the Jacobian pattern in tdy->J as COO. */
PetscErrorCode TDyHipSetup(TDy tdy) {
  TDyMPFAO *mpfao = (TDyMPFAO *)tdy->context;
  TDyHip *h;
  MeshArrays a;
  tdyhip_mesh m;
  tdyhip_materials mat;
  PetscMPIInt size;
  PetscContainer c;
  PetscFunctionBegin;
  PetscCallMPI(MPI_Comm_size(PetscObjectComm((PetscObject)tdy), &size));
  if (size > 1) SETERRQ(PetscObjectComm((PetscObject)tdy), PETSC_ERR_SUP, "tdyhip runs one rank per mesh");
  PetscCall(PetscNew(&h));
  PetscCall(TDyHipMesh(tdy, &a, &m, &mat));
  PetscCall(hipchk(tdyhip_create(&m, &mat, &h->g)));
  PetscCall(MeshArraysFree(&a));
  /* the reference's source_sink (zero unless something filled it, tdympfao.c:364-384) */
  PetscCall(hipchk(tdyhip_set_source_sink(h->g, mpfao->source_sink)));
  void *s = NULL;
  PetscCall(hipchk(tdyhip_get_stream(h->g, &s)));
  h->stream = (hipStream_t)s;
  int64_t sz[4];
  PetscCall(hipchk(tdyhip_sizes(h->g, sz)));
  h->nc = (PetscInt)sz[0];
  h->nnz = (PetscInt)sz[1];
  int32_t *rp, *cj;
  PetscCall(PetscMalloc1(h->nc + 1, &rp));
  PetscCall(PetscMalloc1(h->nnz, &cj));
  PetscCall(hipchk(tdyhip_jacobian_pattern(h->g, rp, cj)));
  PetscCall(PetscMalloc1(h->nnz, &h->coo_i));
  PetscCall(PetscMalloc1(h->nnz, &h->coo_j));
  for (PetscInt i = 0; i < h->nc; i++)
    for (int32_t k = rp[i]; k < rp[i + 1]; k++) { h->coo_i[k] = i; h->coo_j[k] = cj[k]; }
  PetscCall(PetscFree(rp));
  PetscCall(PetscFree(cj));
  PetscCall(PetscMalloc1(h->nnz, &h->h_j));
  PetscCall(hiprtchk(hipMalloc((void **)&h->d_u, sizeof(double) * h->nc)));
  PetscCall(hiprtchk(hipMalloc((void **)&h->d_ut, sizeof(double) * h->nc)));
  PetscCall(hiprtchk(hipMalloc((void **)&h->d_r, sizeof(double) * h->nc)));
  PetscCall(hiprtchk(hipMalloc((void **)&h->d_j, sizeof(double) * h->nnz)));
  if (!tdy->J) PetscCall(TDyCreateJacobian(tdy));
  PetscCall(MatSetPreallocationCOO(tdy->J, h->nnz, h->coo_i, h->coo_j));
  PetscCall(PetscContainerCreate(PETSC_COMM_SELF, &c));
  PetscCall(PetscContainerSetPointer(c, h));
  PetscCall(PetscContainerSetUserDestroy(c, TDyHipFree));
  PetscCall(PetscObjectCompose((PetscObject)tdy, kKey, (PetscObject)c));
  PetscCall(PetscContainerDestroy(&c));
  PetscFunctionReturn(0);
}

static PetscErrorCode to_dev(TDyHip *h, Vec v, double *d) {
  const PetscScalar *p;
  PetscFunctionBegin;
  PetscCall(VecGetArrayRead(v, &p));
  PetscCall(hiprtchk(hipMemcpyAsync(d, p, sizeof(double) * h->nc, hipMemcpyHostToDevice, h->stream)));
  PetscCall(hiprtchk(hipStreamSynchronize(h->stream)));
  PetscCall(VecRestoreArrayRead(v, &p));
  PetscFunctionReturn(0);
}

static PetscErrorCode from_dev_R(TDyHip *h, Vec R) {
  PetscScalar *r;
  PetscFunctionBegin;
  PetscCall(VecGetArrayWrite(R, &r));
  PetscCall(hiprtchk(hipMemcpyAsync(r, h->d_r, sizeof(double) * h->nc, hipMemcpyDeviceToHost, h->stream)));
  PetscCall(hiprtchk(hipStreamSynchronize(h->stream)));
  PetscCall(VecRestoreArrayWrite(R, &r));
  PetscFunctionReturn(0);
}

/* B = the device Jacobian (values in the COO order installed by TDyHipSetup);
 * A assembled only if it is not B (tdympfao_ts.c:532, 544-547). */
static PetscErrorCode set_jacobian(TDyHip *h, Mat A, Mat B) {
  PetscFunctionBegin;
  PetscCall(hiprtchk(hipMemcpyAsync(h->h_j, h->d_j, sizeof(double) * h->nnz, hipMemcpyDeviceToHost, h->stream)));
  PetscCall(hiprtchk(hipStreamSynchronize(h->stream)));
  PetscCall(MatSetValuesCOO(B, h->h_j, INSERT_VALUES));
  if (A != B) {
    PetscCall(MatAssemblyBegin(A, MAT_FINAL_ASSEMBLY));
    PetscCall(MatAssemblyEnd(A, MAT_FINAL_ASSEMBLY));
  }
  PetscFunctionReturn(0);
}

/* TDyMPFAOIFunction (include/private/tdympfaoimpl.h:150; tdympfao_ts.c:153) */
PetscErrorCode TDyMPFAOIFunction_HIP(TS ts, PetscReal t, Vec U, Vec U_t, Vec R, void *ctx) {
  TDyHip *h;
  PetscFunctionBegin;
  PetscCall(TDyHipGet((TDy)ctx, &h));
  PetscCall(to_dev(h, U, h->d_u));
  PetscCall(to_dev(h, U_t, h->d_ut));
  PetscCall(hipchk(tdyhip_ifunction(h->g, t, h->d_u, h->d_ut, h->d_r, NULL)));
  PetscCall(from_dev_R(h, R));
  PetscFunctionReturn(0);
}

/* TDyMPFAOIJacobian (include/private/tdympfaoimpl.h:153; tdympfao_ts.c:520): uses
 * the state of the preceding IFunction, like the reference */
PetscErrorCode TDyMPFAOIJacobian_HIP(TS ts, PetscReal t, Vec U, Vec U_t, PetscReal shift, Mat A, Mat B, void *ctx) {
  TDyHip *h;
  PetscFunctionBegin;
  PetscCall(TDyHipGet((TDy)ctx, &h));
  PetscCall(to_dev(h, U, h->d_u));
  PetscCall(to_dev(h, U_t, h->d_ut));
  PetscCall(hipchk(tdyhip_ijacobian(h->g, t, h->d_u, h->d_ut, shift, h->d_j, NULL)));
  PetscCall(set_jacobian(h, A, B));
  PetscFunctionReturn(0);
}

/* TDyMPFAOSNESPreSolve (tdympfao_snes.c:32), reached through TDyPreSolveSNESSolver
 * (src/tdycore.c:2074-2120); dt from TDySetDtimeForSNESSolver (src/tdycore.c:1575) */
PetscErrorCode TDyMPFAOSNESPreSolve_HIP(TDy tdy) {
  TDyHip *h;
  PetscFunctionBegin;
  PetscCall(TDyHipGet(tdy, &h));
  PetscCall(hipchk(tdyhip_set_dtime(h->g, tdy->dtime)));
  PetscCall(to_dev(h, tdy->soln_prev, h->d_u));
  PetscCall(hipchk(tdyhip_snes_presolve(h->g, h->d_u, NULL)));
  PetscFunctionReturn(0);
}

/* TDyMPFAOSNESFunction (include/private/tdympfaoimpl.h:157; tdympfao_snes.c:67) */
PetscErrorCode TDyMPFAOSNESFunction_HIP(SNES snes, Vec U, Vec R, void *ctx) {
  TDy tdy = (TDy)ctx;
  TDyHip *h;
  PetscFunctionBegin;
  PetscCall(TDyHipGet(tdy, &h));
  PetscCall(hipchk(tdyhip_set_dtime(h->g, tdy->dtime)));
  PetscCall(to_dev(h, U, h->d_u));
  PetscCall(hipchk(tdyhip_snes_function(h->g, h->d_u, h->d_r, NULL)));
  PetscCall(from_dev_R(h, R));
  PetscFunctionReturn(0);
}

/* TDyMPFAOSNESJacobian (include/private/tdympfaoimpl.h:158; tdympfao_snes.c:143) */
PetscErrorCode TDyMPFAOSNESJacobian_HIP(SNES snes, Vec U, Mat A, Mat B, void *ctx) {
  TDyHip *h;
  PetscFunctionBegin;
  PetscCall(TDyHipGet((TDy)ctx, &h));
  PetscCall(to_dev(h, U, h->d_u));
  PetscCall(hipchk(tdyhip_snes_jacobian(h->g, h->d_u, h->d_j, NULL)));
  PetscCall(set_jacobian(h, A, B));
  PetscFunctionReturn(0);
}

/*
 * Registration changes in src/tdycore.c (each under #if defined(TDYCORE_HIP) and
 * only for tdy->options.discretization == MPFA_O && tdy->options.mode == RICHARDS):
 *
 *   TDySetIFunction (src/tdycore.c:1892):
 *     PetscCall(TDyHipSetup(tdy));
 *     PetscCall(TSSetIFunction(ts, NULL, TDyMPFAOIFunction_HIP, tdy));
 *   TDySetIJacobian (src/tdycore.c:1933-1936), after TDyCreateJacobian:
 *     PetscCall(TSSetIJacobian(ts, tdy->J, tdy->J, TDyMPFAOIJacobian_HIP, tdy));
 *   TDySetSNESFunction (src/tdycore.c:1987):
 *     PetscCall(TDyHipSetup(tdy));
 *     PetscCall(SNESSetFunction(snes, tdy->residual, TDyMPFAOSNESFunction_HIP, tdy));
 *   TDySetSNESJacobian (src/tdycore.c:2034):
 *     PetscCall(SNESSetJacobian(snes, tdy->J, tdy->J, TDyMPFAOSNESJacobian_HIP, tdy));
 *   TDyPreSolveSNESSolver (src/tdycore.c:2074-2120), RICHARDS / MPFA_O branch:
 *     PetscCall(TDyMPFAOSNESPreSolve_HIP(tdy));
 *
 * and the five prototypes above in include/private/tdympfaoimpl.h beside the
 * CPU ones (tdympfaoimpl.h:150-160).  TDyHipSetup runs once per TDy (calling it
 * again composes a new context in place of the old one).
 */
