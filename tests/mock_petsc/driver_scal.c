/* Host driver with the call sequence of src/Expe_scal.c:66-160: PC_init,
 * caller-filled kappa, PC_create_A, a PCSHELL inside PETSc CG
 * (KSP_NORM_UNPRECONDITIONED), PC_print_stat, PC_final.  It is compiled
 * against premixfem_amd/petsc (the adapter) and either the GPU library or a
 * stub core.  With -compare_gpu_pcg it also runs pmx_solve on the same
 * system and reports the relative difference of the two solutions.
 *
 * Output: one line "RESULT its=<n> reason=<r> rnorm=<x> [gpu_its=<n>
 * gpu_reason=<r> rel_diff=<x>]".
 *
 * -kappa_rescale: after the first solve, kappa is doubled in place and the
 * operator scaled by 2 (the operator of 2 kappa, exactly), and the system is
 * solved again: PC_setup must read the new kappa.  Prints "RESULT2 its=<n>
 * reason=<r> ratio=<|x1|/|x2|>" (2 when the second solve used 2 kappa).
 * -operator_mismatch: the operator alone is scaled before the second solve,
 * which PC_setup must refuse.  Test infrastructure only. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "PreMixFEM_3D.h"

static double cell_kappa(PetscInt ex, PetscInt ey, PetscInt ez, PetscInt M) {
  /* channels of contrast 1e3 along x and z, a smooth background elsewhere */
  const int chan = (ey % 5 == 2 && ez % 4 == 1) || (ex % 6 == 3 && ey % 3 == 0);
  return chan ? 1.0e3 : exp(0.5 * sin(0.7 * ex + 0.3 * ey) * cos(0.4 * ez + 0.2 * M));
}

static PetscErrorCode create_kappa(PCCtx *s_ctx) {
  PetscInt sx, sy, sz, nx, ny, nz, ex, ey, ez;
  int d;
  PetscFunctionBeginUser;
  PetscCall(DMDAGetCorners(s_ctx->dm, &sx, &sy, &sz, &nx, &ny, &nz));
  for (d = 0; d < DIM; ++d) {
    PetscScalar *a;
    PetscCall(DMCreateGlobalVector(s_ctx->dm, &s_ctx->kappa[d]));
    PetscCall(VecGetArray(s_ctx->kappa[d], &a));
    for (ez = sz; ez < sz + nz; ++ez)
      for (ey = sy; ey < sy + ny; ++ey)
        for (ex = sx; ex < sx + nx; ++ex)
          a[((ez - sz) * ny + (ey - sy)) * nx + (ex - sx)] =
              cell_kappa(ex, ey, ez, s_ctx->M) * (d == 2 ? 0.5 : 1.0);
    PetscCall(VecRestoreArray(s_ctx->kappa[d], &a));
  }
  PetscFunctionReturn(PETSC_SUCCESS);
}

/* +|E| at the first cell, -|E| at the last one (a consistent right-hand
 * side for the pure-Neumann operator). */
static PetscErrorCode create_rhs(PCCtx *s_ctx, Vec *rhs) {
  PetscScalar *a;
  const PetscScalar meas = s_ctx->H_x * s_ctx->H_y * s_ctx->H_z;
  PetscInt n;
  PetscFunctionBeginUser;
  PetscCall(DMCreateGlobalVector(s_ctx->dm, rhs));
  PetscCall(VecSet(*rhs, 0.0));
  PetscCall(VecGetSize(*rhs, &n));
  PetscCall(VecGetArray(*rhs, &a));
  a[0] = meas;
  a[n - 1] = -meas;
  PetscCall(VecRestoreArray(*rhs, &a));
  PetscFunctionReturn(PETSC_SUCCESS);
}

static PetscErrorCode run(void) {
  PetscScalar dom[3] = {1.0, 1.0, 1.0};
  PetscInt mesh[3] = {16, 16, 16}, its, n;
  PetscBool compare = PETSC_FALSE, rescale = PETSC_FALSE, mismatch = PETSC_FALSE;
  PCCtx s_ctx;
  Mat A;
  Vec rhs, u;
  KSP ksp;
  PC pc;
  KSPConvergedReason reason;
  PetscReal rnorm;

  PetscFunctionBeginUser;
  PetscCall(PetscOptionsGetInt(NULL, NULL, "-mesh_x", &mesh[0], NULL));
  PetscCall(PetscOptionsGetInt(NULL, NULL, "-mesh_y", &mesh[1], NULL));
  PetscCall(PetscOptionsGetInt(NULL, NULL, "-mesh_z", &mesh[2], NULL));
  PetscCall(PetscOptionsHasName(NULL, NULL, "-compare_gpu_pcg", &compare));
  PetscCall(PetscOptionsHasName(NULL, NULL, "-kappa_rescale", &rescale));
  PetscCall(PetscOptionsHasName(NULL, NULL, "-operator_mismatch", &mismatch));
  PetscCall(PC_init(&s_ctx, &dom[0], &mesh[0]));
  PetscCall(PC_print_info(&s_ctx));
  PetscCall(create_kappa(&s_ctx));
  PetscCall(PC_create_A(&s_ctx, &A));
  PetscCall(create_rhs(&s_ctx, &rhs));
  PetscCall(DMCreateGlobalVector(s_ctx.dm, &u));
  PetscCall(VecSet(u, 0.0));

  PetscCall(KSPCreate(PETSC_COMM_WORLD, &ksp));
  PetscCall(KSPSetOperators(ksp, A, A));
  PetscCall(KSPGetPC(ksp, &pc));
  PetscCall(PCSetType(pc, PCSHELL));
  PetscCall(PCShellSetContext(pc, &s_ctx));
  PetscCall(PCShellSetSetUp(pc, PC_setup));
  PetscCall(PCShellSetApply(pc, PC_apply_vec));
  PetscCall(PCShellSetName(pc, "3levels-MG-via-GMsFEM-with-velocity-elimination"));
  PetscCall(KSPSetNormType(ksp, KSP_NORM_UNPRECONDITIONED));
  PetscCall(KSPSetFromOptions(ksp));
  PetscCall(KSPSetUp(ksp));
  PetscCall(KSPSolve(ksp, rhs, u));
  PetscCall(KSPGetIterationNumber(ksp, &its));
  PetscCall(KSPGetConvergedReason(ksp, &reason));
  PetscCall(KSPGetResidualNorm(ksp, &rnorm));
  printf("RESULT its=%d reason=%d rnorm=%.6e", (int)its, (int)reason, (double)rnorm);
  if (compare) {
    const PetscScalar *b, *x;
    PetscScalar rtol = 1e-5;
    double *xg, num = 0.0, den = 0.0, grn = 0.0;
    int gits = 0, rc;
    PetscInt i;
    pmx_stats st;
    PetscCall(PetscOptionsGetScalar(NULL, NULL, "-ksp_rtol", &rtol, NULL));
    PetscCall(VecGetSize(u, &n));
    xg = calloc((size_t)n, sizeof(double));
    PetscCall(VecGetArrayRead(rhs, &b));
    rc = pmx_solve(s_ctx.gpu, b, xg, 0, rtol, 10000, &gits, &grn);
    PetscCall(VecRestoreArrayRead(rhs, &b));
    if (rc) SETERRQ(PETSC_COMM_SELF, rc, "%s", pmx_last_error());
    PetscCall(pmx_get_stats(s_ctx.gpu, &st));
    PetscCall(VecGetArrayRead(u, &x));
    for (i = 0; i < n; ++i) {
      num += (xg[i] - x[i]) * (xg[i] - x[i]);
      den += x[i] * x[i];
    }
    PetscCall(VecRestoreArrayRead(u, &x));
    printf(" gpu_its=%d gpu_reason=%d rel_diff=%.3e", gits, st.reason, sqrt(num / den));
    free(xg);
  }
  printf("\n");
  if (rescale || mismatch) {
    PetscReal n1, n2;
    PetscCall(VecNorm(u, NORM_2, &n1));
    if (rescale)
      for (int d = 0; d < DIM; ++d) {
        PetscScalar *a;
        PetscInt i, nk;
        PetscCall(VecGetSize(s_ctx.kappa[d], &nk));
        PetscCall(VecGetArray(s_ctx.kappa[d], &a));
        for (i = 0; i < nk; ++i) a[i] *= 2.0;
        PetscCall(VecRestoreArray(s_ctx.kappa[d], &a));
      }
    PetscCall(MatScale(A, 2.0));
    PetscCall(VecSet(u, 0.0));
    PetscCall(KSPSolve(ksp, rhs, u));
    PetscCall(KSPGetIterationNumber(ksp, &its));
    PetscCall(KSPGetConvergedReason(ksp, &reason));
    PetscCall(VecNorm(u, NORM_2, &n2));
    printf("RESULT2 its=%d reason=%d ratio=%.15e\n", (int)its, (int)reason, (double)(n1 / n2));
  }
  PetscCall(PC_print_stat(&s_ctx));
  PetscCall(KSPDestroy(&ksp));
  PetscCall(VecDestroy(&rhs));
  PetscCall(VecDestroy(&u));
  PetscCall(MatDestroy(&A));
  for (int d = 0; d < DIM; ++d) PetscCall(VecDestroy(&s_ctx.kappa[d]));
  PetscCall(PC_final(&s_ctx));
  PetscFunctionReturn(PETSC_SUCCESS);
}

int main(int argc, char **argv) {
  int rc;
  PetscCall(PetscInitialize(&argc, &argv, NULL, NULL));
  rc = run();
  PetscFinalize();
  return rc;
}
