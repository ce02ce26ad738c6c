/*
 * PreMixFEM_3D.h -- the PCSHELL interface of the reference
 * (src/PreMixFEM_3D.h:32-67) over the GPU library's C ABI (premixfem_gpu.h).
 *
 * A PETSc host program written against the reference (src/Expe_scal.c,
 * src/Expe_frac.c, src/spe10.c) compiles unchanged against this header and
 * links against pmx_pcshell.c + libpremixfem_gpu.so instead of
 * PreMixFEM_3D.c.  The context keeps the fields the drivers read or write
 * (dm, kappa[], H_x.., M/N/P, the option values and stage times); the
 * preconditioner's own state lives in the GPU context `gpu`.
 *
 * One MPI rank drives one GPU: the adapter refuses communicators with more
 * than one rank (PETSC_ERR_SUP).
 */
#ifndef PMX_PETSC_PREMIXFEM_3D_H
#define PMX_PETSC_PREMIXFEM_3D_H

#include <mpi.h>
#include <petscksp.h>

#include "premixfem_gpu.h"

#define DIM 3
#define NIL 1.0e-12
#define MAX_LOG_STATES PMX_MAX_LOG_STAGES
#define STAGE_SU_LV1 PMX_STAGE_SU_LV1
#define STAGE_SU_LV2_1 PMX_STAGE_SU_LV2_1
#define STAGE_SU_LV2_2 PMX_STAGE_SU_LV2_2
#define STAGE_SU_LV3_1 PMX_STAGE_SU_LV3_1
#define STAGE_SU_LV3_2 PMX_STAGE_SU_LV3_2
#define STAGE_AV_LV1 PMX_STAGE_AV_LV1
#define STAGE_AV_LV2 PMX_STAGE_AV_LV2
#define STAGE_AV_LV3 PMX_STAGE_AV_LV3
#define STAGE_SU PMX_STAGE_SU
#define STAGE_AV PMX_STAGE_AV
#define EIG_OP_MOD PMX_EIG_OP_MOD
#define EIG_OP_MEAN PMX_EIG_OP_MEAN
#define EIG_OP_DIAG PMX_EIG_OP_DIAG

typedef struct preconditioner_context {
  DM dm;             /* DMDA of the pressure cells (created by PC_init) */
  Vec kappa[DIM];    /* created and filled by the caller after PC_init */
  PetscInt smoothing_iters_lv1, smoothing_iters_lv2, sub_domains, lv2_eigen_op;
  PetscInt max_eigen_num_lv1, max_eigen_num_lv2, eigen_num_lv2;
  PetscScalar H_x, H_y, H_z, L, W, H;
  PetscScalar eigen_bd_lv1, eigen_bd_lv2, eigen_min_lv2, eigen_max_lv2;
  PetscScalar t_stages[MAX_LOG_STATES];
  PetscBool use_W_cycle, no_shift_A_cc, use_full_Cholesky_lv1, use_2level;
  PetscInt M, N, P;
  pmx_ctx *gpu;      /* the GPU preconditioner */
  Mat op;            /* the PC operator PC_setup checked against kappa ... */
  PetscObjectState op_state;  /* ... and its state then */
} PCCtx;

PetscErrorCode PC_init(PCCtx *s_ctx, PetscScalar *dom, PetscInt *mesh);
PetscErrorCode PC_print_info(PCCtx *s_ctx);
PetscErrorCode PC_setup(PC pc);
PetscErrorCode PC_create_A(PCCtx *s_ctx, Mat *A);
PetscErrorCode PC_apply_vec(PC pc, Vec x, Vec y);
PetscErrorCode PC_get_range(const void *sendbuff, void *recvbuff, MPI_Datatype datatype);
PetscErrorCode PC_print_stat(PCCtx *s_ctx);
PetscErrorCode PC_final_default(PCCtx *s_ctx);
PetscErrorCode PC_final(PCCtx *s_ctx);
int mkl_serv_intel_cpu_true(void);

#endif /* PMX_PETSC_PREMIXFEM_3D_H */
