/*
 * pmx_pcshell.c -- PCSHELL adapter: the functions of the reference's
 * PreMixFEM_3D.h implemented over the GPU library's C ABI.
 *
 * Host programs register PC_setup / PC_apply_vec with PCShellSetSetUp /
 * PCShellSetApply exactly as src/Expe_scal.c:104-112 does; PETSc's CG then
 * runs on the host and every preconditioner apply copies x to the GPU and y
 * back (the host-CG drop-in).  pmx_solve runs the whole PCG on the GPU
 * instead.
 *
 * Build (PETSc present):  cc -c pmx_pcshell.c -I<repo>/include $(PETSc cflags)
 * and link with -lpremixfem_gpu.
 */
#include <stdio.h>
#include <string.h>

#include "PreMixFEM_3D.h"

#define PMX_CALL(expr)                                                    \
  do {                                                                    \
    const int pmx_rc_ = (expr);                                           \
    if (pmx_rc_) SETERRQ(PETSC_COMM_SELF, pmx_rc_, "%s", pmx_last_error()); \
  } while (0)

static void opt_append(char *buf, size_t cap, const char *fmt, const char *name, const char *val) {
  const size_t len = strlen(buf);
  if (len + 1 < cap) snprintf(buf + len, cap - len, fmt, name, val);
}

/* Options the GPU core understands beyond the reference's own set. */
static const char *kValueOptions[] = {"-sdx",      "-sdy",       "-sdz",    "-pc_precision",
                                      "-pin",      "-eig_tol",   "-eig_maxit", "-eig_sigma",
                                      "-seed",     "-warm_start", "-cg_chunk", "-procs_virtual"};
static const char *kFlagOptions[] = {"-kernel_timing", "-log_stages"};

/* src/PreMixFEM_3D.c:1650-1741: option values with the reference's defaults,
 * the DMDA of the cells, then the GPU context. */
PetscErrorCode PC_init(PCCtx *s_ctx, PetscScalar *dom, PetscInt *mesh) {
  PetscMPIInt size;
  char opts[2048], val[256];
  PetscBool set;
  size_t i;

  PetscFunctionBeginUser;
  PetscCallMPI(MPI_Comm_size(PETSC_COMM_WORLD, &size));
  PetscCheck(size == 1, PETSC_COMM_WORLD, PETSC_ERR_SUP,
             "the GPU preconditioner runs one MPI rank per GPU");
  PetscCheck(dom[0] > 0.0 && dom[1] > 0.0 && dom[2] > 0.0, PETSC_COMM_WORLD, PETSC_ERR_ARG_WRONG,
             "Errors in dom(L, W, H)=[%.5f, %.5f, %.5f].", (double)dom[0], (double)dom[1],
             (double)dom[2]);
  PetscCheck(mesh[0] > 0 && mesh[1] > 0 && mesh[2] > 0, PETSC_COMM_WORLD, PETSC_ERR_ARG_WRONG,
             "Errors in mesh(M, N, P)=[%d, %d, %d].", (int)mesh[0], (int)mesh[1], (int)mesh[2]);
  memset(s_ctx, 0, sizeof(*s_ctx));
  s_ctx->L = dom[0];
  s_ctx->W = dom[1];
  s_ctx->H = dom[2];
  s_ctx->M = mesh[0];
  s_ctx->N = mesh[1];
  s_ctx->P = mesh[2];
  s_ctx->H_x = dom[0] / (PetscScalar)mesh[0];
  s_ctx->H_y = dom[1] / (PetscScalar)mesh[1];
  s_ctx->H_z = dom[2] / (PetscScalar)mesh[2];
  s_ctx->smoothing_iters_lv1 = 1;
  s_ctx->smoothing_iters_lv2 = 1;
  s_ctx->sub_domains = 2;
  s_ctx->max_eigen_num_lv1 = 4;
  s_ctx->max_eigen_num_lv2 = 4;
  s_ctx->lv2_eigen_op = EIG_OP_MOD;
  s_ctx->eigen_bd_lv1 = 1.0 / NIL;
  s_ctx->eigen_bd_lv2 = 1.0 / NIL;
  PetscCall(PetscOptionsGetInt(NULL, NULL, "-si_lv1", &s_ctx->smoothing_iters_lv1, NULL));
  PetscCall(PetscOptionsGetInt(NULL, NULL, "-si_lv2", &s_ctx->smoothing_iters_lv2, NULL));
  PetscCall(PetscOptionsGetInt(NULL, NULL, "-sd", &s_ctx->sub_domains, NULL));
  PetscCall(PetscOptionsGetInt(NULL, NULL, "-en_lv1", &s_ctx->max_eigen_num_lv1, NULL));
  PetscCall(PetscOptionsGetInt(NULL, NULL, "-en_lv2", &s_ctx->max_eigen_num_lv2, NULL));
  PetscCall(PetscOptionsGetInt(NULL, NULL, "-eg_op", &s_ctx->lv2_eigen_op, NULL));
  PetscCall(PetscOptionsGetScalar(NULL, NULL, "-eb_lv1", &s_ctx->eigen_bd_lv1, NULL));
  PetscCall(PetscOptionsGetScalar(NULL, NULL, "-eb_lv2", &s_ctx->eigen_bd_lv2, NULL));
  PetscCall(PetscOptionsHasName(NULL, NULL, "-use_W_cycle", &s_ctx->use_W_cycle));
  PetscCall(PetscOptionsHasName(NULL, NULL, "-no_shift_A_cc", &s_ctx->no_shift_A_cc));
  PetscCall(PetscOptionsHasName(NULL, NULL, "-use_full_Cholesky_lv1",
                                &s_ctx->use_full_Cholesky_lv1));
  PetscCall(PetscOptionsHasName(NULL, NULL, "-use_2level", &s_ctx->use_2level));
  if (s_ctx->sub_domains == 1) s_ctx->sub_domains = 2;

  snprintf(opts, sizeof(opts),
           "-si_lv1 %d -si_lv2 %d -sd %d -en_lv1 %d -en_lv2 %d -eg_op %d -eb_lv1 %.17g "
           "-eb_lv2 %.17g%s%s%s%s",
           (int)s_ctx->smoothing_iters_lv1, (int)s_ctx->smoothing_iters_lv2,
           (int)s_ctx->sub_domains, (int)s_ctx->max_eigen_num_lv1,
           (int)s_ctx->max_eigen_num_lv2, (int)s_ctx->lv2_eigen_op,
           (double)s_ctx->eigen_bd_lv1, (double)s_ctx->eigen_bd_lv2,
           s_ctx->use_W_cycle ? " -use_W_cycle" : "", s_ctx->no_shift_A_cc ? " -no_shift_A_cc" : "",
           s_ctx->use_full_Cholesky_lv1 ? " -use_full_Cholesky_lv1" : "",
           s_ctx->use_2level ? " -use_2level" : "");
  for (i = 0; i < sizeof(kValueOptions) / sizeof(kValueOptions[0]); ++i) {
    PetscCall(PetscOptionsGetString(NULL, NULL, kValueOptions[i], val, sizeof(val), &set));
    if (set) opt_append(opts, sizeof(opts), " %s %s", kValueOptions[i], val);
  }
  for (i = 0; i < sizeof(kFlagOptions) / sizeof(kFlagOptions[0]); ++i) {
    PetscCall(PetscOptionsHasName(NULL, NULL, kFlagOptions[i], &set));
    if (set) opt_append(opts, sizeof(opts), " %s%s", kFlagOptions[i], "");
  }

  PetscCall(DMDACreate3d(PETSC_COMM_WORLD, DM_BOUNDARY_NONE, DM_BOUNDARY_NONE, DM_BOUNDARY_NONE,
                         DMDA_STENCIL_STAR, s_ctx->M, s_ctx->N, s_ctx->P, PETSC_DECIDE,
                         PETSC_DECIDE, PETSC_DECIDE, 1, 1, NULL, NULL, NULL, &s_ctx->dm));
  PetscCall(DMSetUp(s_ctx->dm));
  {
    double d[3] = {(double)dom[0], (double)dom[1], (double)dom[2]};
    int m[3] = {(int)mesh[0], (int)mesh[1], (int)mesh[2]};
    const int rc = pmx_init(&s_ctx->gpu, d, m, opts);
    if (rc) {
      PetscCall(DMDestroy(&s_ctx->dm));
      SETERRQ(PETSC_COMM_SELF, rc, "%s", pmx_last_error());
    }
  }
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* src/PreMixFEM_3D.c:1743-1755 */
PetscErrorCode PC_print_info(PCCtx *s_ctx) {
  PetscFunctionBeginUser;
  PMX_CALL(pmx_print_info(s_ctx->gpu));
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* src/PreMixFEM_3D.c:1757-1824: hands kappa to the GPU, which computes the
 * face coefficients; the seven diagonals come back and fill a DMDA matrix. */
PetscErrorCode PC_create_A(PCCtx *s_ctx, Mat *A) {
  const PetscScalar *k[DIM];
  PetscScalar *diag, *off;
  PetscInt ex, ey, ez, n = s_ctx->M * s_ctx->N * s_ctx->P;
  int d;

  PetscFunctionBeginUser;
  for (d = 0; d < DIM; ++d) PetscCall(VecGetArrayRead(s_ctx->kappa[d], &k[d]));
  {
    const int rc = pmx_set_kappa(s_ctx->gpu, k[0], k[1], k[2]);
    for (d = 0; d < DIM; ++d) PetscCall(VecRestoreArrayRead(s_ctx->kappa[d], &k[d]));
    if (rc) SETERRQ(PETSC_COMM_SELF, rc, "%s", pmx_last_error());
  }
  PetscCall(PetscMalloc2(n, &diag, 6 * n, &off));
  PMX_CALL(pmx_create_A(s_ctx->gpu, diag, off));
  PetscCall(DMCreateMatrix(s_ctx->dm, A));
  for (ez = 0; ez < s_ctx->P; ++ez)
    for (ey = 0; ey < s_ctx->N; ++ey)
      for (ex = 0; ex < s_ctx->M; ++ex) {
        static const int dx[6] = {-1, 1, 0, 0, 0, 0}, dy[6] = {0, 0, -1, 1, 0, 0},
                         dz[6] = {0, 0, 0, 0, -1, 1};
        const PetscInt c = (ez * s_ctx->N + ey) * s_ctx->M + ex;
        MatStencil row, col[7];
        PetscScalar val[7];
        PetscInt nc = 0;
        row.i = ex;
        row.j = ey;
        row.k = ez;
        row.c = 0;
        col[nc] = row;
        val[nc++] = diag[c];
        for (d = 0; d < 6; ++d) {
          if (off[6 * c + d] == 0.0) continue;
          col[nc].i = ex + dx[d];
          col[nc].j = ey + dy[d];
          col[nc].k = ez + dz[d];
          col[nc].c = 0;
          val[nc++] = off[6 * c + d];
        }
        PetscCall(MatSetValuesStencil(*A, 1, &row, nc, col, val, INSERT_VALUES));
      }
  PetscCall(MatAssemblyBegin(*A, MAT_FINAL_ASSEMBLY));
  PetscCall(MatAssemblyEnd(*A, MAT_FINAL_ASSEMBLY));
  PetscCall(PetscFree2(diag, off));
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* Pushes the caller's kappa vectors to the GPU context (the reference's
 * _PC_setup reads s_ctx->kappa at every setup, src/PreMixFEM_3D.c:1398-1402). */
static PetscErrorCode push_kappa(PCCtx *s_ctx) {
  const PetscScalar *k[DIM];
  int d, rc;

  PetscFunctionBeginUser;
  for (d = 0; d < DIM; ++d) PetscCall(VecGetArrayRead(s_ctx->kappa[d], &k[d]));
  rc = pmx_set_kappa(s_ctx->gpu, k[0], k[1], k[2]);
  for (d = 0; d < DIM; ++d) PetscCall(VecRestoreArrayRead(s_ctx->kappa[d], &k[d]));
  if (rc) SETERRQ(PETSC_COMM_SELF, rc, "%s", pmx_last_error());
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* The reference's apply takes A from PCGetOperators (src/PreMixFEM_3D.c:1852)
 * for its residuals; the GPU builds the same operator from kappa.  Every row
 * of A must equal the GPU's (PC_create_A's) within 1e-12 of its diagonal, or
 * the setup is refused: a preconditioner of another operator would be built
 * silently otherwise. */
static PetscErrorCode check_operator(PCCtx *s_ctx, Mat A) {
  static const int dx[6] = {-1, 1, 0, 0, 0, 0}, dy[6] = {0, 0, -1, 1, 0, 0},
                   dz[6] = {0, 0, 0, 0, -1, 1};
  const PetscInt M = s_ctx->M, N = s_ctx->N, P = s_ctx->P, n = M * N * P;
  PetscScalar *diag, *off;
  PetscInt row, bad = -1;

  PetscFunctionBeginUser;
  PetscCall(PetscMalloc2(n, &diag, 6 * n, &off));
  {
    const int rc = pmx_create_A(s_ctx->gpu, diag, off);
    if (rc) {
      PetscCall(PetscFree2(diag, off));
      SETERRQ(PETSC_COMM_SELF, rc, "%s", pmx_last_error());
    }
  }
  for (row = 0; row < n && bad < 0; ++row) {
    const PetscInt ex = row % M, ey = (row / M) % N, ez = row / (M * N);
    const PetscScalar tol = 1.0e-12 * (diag[row] > 0.0 ? diag[row] : -diag[row]);
    PetscScalar want[7];
    PetscInt want_col[7], nw = 0, ncols, j, d;
    const PetscInt *cols;
    const PetscScalar *vals;
    want_col[nw] = row;
    want[nw++] = diag[row];
    for (d = 0; d < 6; ++d) {
      const PetscInt x = ex + dx[d], y = ey + dy[d], z = ez + dz[d];
      if (x < 0 || x >= M || y < 0 || y >= N || z < 0 || z >= P) continue;
      want_col[nw] = (z * N + y) * M + x;
      want[nw++] = off[6 * row + d];
    }
    PetscCall(MatGetRow(A, row, &ncols, &cols, &vals));
    /* every stored entry of the row must be one of the 7-point entries with
     * the GPU's value; entries the row does not store must be zero there */
    for (j = 0; j < ncols && bad < 0; ++j) {
      PetscInt w = 0;
      while (w < nw && want_col[w] != cols[j]) ++w;
      if (w == nw ? (vals[j] != 0.0) : (vals[j] - want[w] > tol || want[w] - vals[j] > tol))
        bad = row;
      else if (w < nw)
        want_col[w] = -1 - want_col[w]; /* matched */
    }
    for (d = 0; d < nw && bad < 0; ++d)
      if (want_col[d] >= 0 && (want[d] > tol || -want[d] > tol)) bad = row;
    PetscCall(MatRestoreRow(A, row, &ncols, &cols, &vals));
  }
  PetscCall(PetscFree2(diag, off));
  PetscCheck(bad < 0, PETSC_COMM_WORLD, PETSC_ERR_ARG_WRONG,
             "PC_setup: the operator of the PC differs from the one built from s_ctx->kappa "
             "(first at row %d); the GPU preconditioner supports PC_create_A's operator only",
             (int)bad);
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* src/PreMixFEM_3D.c:1826-1838 (PCShellSetSetUp callback): kappa is read
 * again (as _PC_setup does, :1398-1402), the PC's operator is checked
 * against it, and the GPU preconditioner is built. */
PetscErrorCode PC_setup(PC pc) {
  PCCtx *s_ctx;
  Mat A;

  PetscFunctionBeginUser;
  PetscCall(PCShellGetContext(pc, &s_ctx));
  PetscCall(PCGetOperators(pc, &A, NULL));
  PetscCheck(s_ctx->kappa[0] && s_ctx->kappa[1] && s_ctx->kappa[2], PETSC_COMM_WORLD,
             PETSC_ERR_ARG_WRONG, "PC_setup: s_ctx->kappa has not been created");
  PetscCall(push_kappa(s_ctx));
  PetscCall(check_operator(s_ctx, A));
  PMX_CALL(pmx_setup(s_ctx->gpu));
  s_ctx->op = A;
  PetscCall(PetscObjectStateGet((PetscObject)A, &s_ctx->op_state));
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* src/PreMixFEM_3D.c:1840-1936 (PCShellSetApply callback): y = B x.  The
 * residuals of the cycle use the operator PC_setup checked; an operator
 * that has been replaced or modified since is refused. */
PetscErrorCode PC_apply_vec(PC pc, Vec x, Vec y) {
  PCCtx *s_ctx;
  const PetscScalar *xa;
  PetscScalar *ya;
  Mat A;
  PetscObjectState st;
  int rc;

  PetscFunctionBeginUser;
  PetscCall(PCShellGetContext(pc, &s_ctx));
  PetscCall(PCGetOperators(pc, &A, NULL));
  PetscCall(PetscObjectStateGet((PetscObject)A, &st));
  PetscCheck(A == s_ctx->op && st == s_ctx->op_state, PETSC_COMM_WORLD, PETSC_ERR_ARG_WRONG,
             "PC_apply_vec: the operator changed after PC_setup");
  PetscCall(VecGetArrayRead(x, &xa));
  PetscCall(VecGetArray(y, &ya));
  rc = pmx_apply(s_ctx->gpu, xa, ya);
  PetscCall(VecRestoreArray(y, &ya));
  PetscCall(VecRestoreArrayRead(x, &xa));
  if (rc) SETERRQ(PETSC_COMM_SELF, rc, "%s", pmx_last_error());
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* src/PreMixFEM_3D.c:1938-1999 helper: [min, max] over ranks on rank 0. */
PetscErrorCode PC_get_range(const void *sendbuff, void *recvbuff, MPI_Datatype datatype) {
  PetscFunctionBeginUser;
  if (datatype == MPI_DOUBLE) {
    PetscCallMPI(MPI_Reduce(sendbuff, &((double *)recvbuff)[0], 1, datatype, MPI_MIN, 0,
                            PETSC_COMM_WORLD));
    PetscCallMPI(MPI_Reduce(sendbuff, &((double *)recvbuff)[1], 1, datatype, MPI_MAX, 0,
                            PETSC_COMM_WORLD));
  } else if (datatype == MPI_INT) {
    PetscCallMPI(MPI_Reduce(sendbuff, &((int *)recvbuff)[0], 1, datatype, MPI_MIN, 0,
                            PETSC_COMM_WORLD));
    PetscCallMPI(MPI_Reduce(sendbuff, &((int *)recvbuff)[1], 1, datatype, MPI_MAX, 0,
                            PETSC_COMM_WORLD));
  } else {
    SETERRQ(PETSC_COMM_WORLD, PETSC_ERR_SUP, "PC_get_range: MPI_DOUBLE or MPI_INT only");
  }
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* src/PreMixFEM_3D.c:1938-1999 */
PetscErrorCode PC_print_stat(PCCtx *s_ctx) {
  pmx_stats st;
  int i;

  PetscFunctionBeginUser;
  PMX_CALL(pmx_get_stats(s_ctx->gpu, &st));
  for (i = 0; i < MAX_LOG_STATES; ++i) s_ctx->t_stages[i] = st.t_stages[i];
  s_ctx->eigen_num_lv2 = st.eigen_num_lv2;
  s_ctx->eigen_min_lv2 = st.eigen_min_lv2;
  s_ctx->eigen_max_lv2 = st.eigen_max_lv2;
  PMX_CALL(pmx_print_stat(s_ctx->gpu));
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* src/PreMixFEM_3D.c:2037-2046 */
PetscErrorCode PC_final_default(PCCtx *s_ctx) {
  PetscFunctionBeginUser;
  if (s_ctx->dm) PetscCall(DMDestroy(&s_ctx->dm));
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* src/PreMixFEM_3D.c:2001-2035 */
PetscErrorCode PC_final(PCCtx *s_ctx) {
  PetscFunctionBeginUser;
  if (s_ctx->gpu) {
    const int rc = pmx_final(s_ctx->gpu);
    s_ctx->gpu = NULL;
    if (rc) SETERRQ(PETSC_COMM_SELF, rc, "%s", pmx_last_error());
  }
  PetscCall(PC_final_default(s_ctx));
  PetscFunctionReturn(PETSC_SUCCESS);
}

int mkl_serv_intel_cpu_true(void) { return 1; }
