/* Minimal PETSc stand-in (single rank, sequential) with just the API the
 * PCSHELL adapter and the test driver use, so both compile and run here
 * without PETSc.  Semantics follow the PETSc manual pages of the same
 * functions; KSPSolve implements CG with KSP_NORM_UNPRECONDITIONED and the
 * default convergence test.  Test infrastructure only. */
#ifndef MOCK_PETSCKSP_H
#define MOCK_PETSCKSP_H

#include <stddef.h>

#include "mpi.h"

typedef int PetscErrorCode;
typedef int PetscInt;
typedef int PetscMPIInt;
typedef double PetscScalar;
typedef double PetscReal;
typedef enum { PETSC_FALSE = 0, PETSC_TRUE = 1 } PetscBool;
typedef long long PetscObjectState;
typedef void *PetscObject;

#define PETSC_SUCCESS 0
#define PETSC_COMM_WORLD 0
#define PETSC_COMM_SELF 1
#define PETSC_DECIDE (-1)
#define PETSC_ERR_MEM 55
#define PETSC_ERR_SUP 56
#define PETSC_ERR_ARG_WRONG 62
#define PETSC_ERR_ARG_OUTOFRANGE 63
#define PETSC_ERR_LIB 76
#define PETSC_ERR_USER 83
#define PETSC_ERR_NOT_CONVERGED 91

int mock_error(int code, const char *file, int line, const char *fmt, ...);
void mock_trace(int code, const char *file, int line);

#define PetscFunctionBeginUser
#define PetscFunctionReturn(x) return (x)
#define PetscCall(expr)                                   \
  do {                                                    \
    const int mock_e_ = (expr);                           \
    if (mock_e_) {                                        \
      mock_trace(mock_e_, __FILE__, __LINE__);            \
      return mock_e_;                                     \
    }                                                     \
  } while (0)
#define PetscCallMPI(expr) PetscCall(expr)
#define SETERRQ(comm, code, ...) return mock_error((code), __FILE__, __LINE__, __VA_ARGS__)
#define PetscCheck(cond, comm, code, ...) \
  do {                                    \
    if (!(cond)) SETERRQ(comm, code, __VA_ARGS__); \
  } while (0)

PetscErrorCode PetscInitialize(int *argc, char ***argv, const char *file, const char *help);
PetscErrorCode PetscFinalize(void);
PetscErrorCode PetscPrintf(MPI_Comm comm, const char *fmt, ...);
PetscErrorCode PetscOptionsGetInt(void *opts, const char *pre, const char *name, PetscInt *v,
                                  PetscBool *set);
PetscErrorCode PetscOptionsGetScalar(void *opts, const char *pre, const char *name,
                                     PetscScalar *v, PetscBool *set);
PetscErrorCode PetscOptionsGetString(void *opts, const char *pre, const char *name, char *v,
                                     size_t len, PetscBool *set);
PetscErrorCode PetscOptionsHasName(void *opts, const char *pre, const char *name,
                                   PetscBool *set);

PetscErrorCode mock_malloc2(size_t b1, void *p1, size_t b2, void *p2);
PetscErrorCode mock_free2(void *p1, void *p2);
#define PetscMalloc2(n1, p1, n2, p2) \
  mock_malloc2((size_t)(n1) * sizeof(**(p1)), (p1), (size_t)(n2) * sizeof(**(p2)), (p2))
#define PetscFree2(a, b) mock_free2((a), (b))

/* DMDA */
typedef enum { DM_BOUNDARY_NONE } DMBoundaryType;
typedef enum { DMDA_STENCIL_STAR, DMDA_STENCIL_BOX } DMDAStencilType;
typedef struct mock_dm *DM;
typedef struct mock_vec *Vec;
typedef struct mock_mat *Mat;
PetscErrorCode DMDACreate3d(MPI_Comm comm, DMBoundaryType bx, DMBoundaryType by,
                            DMBoundaryType bz, DMDAStencilType st, PetscInt M, PetscInt N,
                            PetscInt P, PetscInt m, PetscInt n, PetscInt p, PetscInt dof,
                            PetscInt s, const PetscInt *lx, const PetscInt *ly,
                            const PetscInt *lz, DM *da);
PetscErrorCode DMSetUp(DM dm);
PetscErrorCode DMDestroy(DM *dm);
PetscErrorCode DMCreateGlobalVector(DM dm, Vec *v);
PetscErrorCode DMCreateMatrix(DM dm, Mat *A);
PetscErrorCode DMDAGetCorners(DM dm, PetscInt *x, PetscInt *y, PetscInt *z, PetscInt *m,
                              PetscInt *n, PetscInt *p);

/* Vec */
typedef enum { NORM_2 } NormType;
PetscErrorCode VecDuplicate(Vec v, Vec *w);
PetscErrorCode VecDestroy(Vec *v);
PetscErrorCode VecSet(Vec v, PetscScalar a);
PetscErrorCode VecGetSize(Vec v, PetscInt *n);
PetscErrorCode VecGetArray(Vec v, PetscScalar **a);
PetscErrorCode VecRestoreArray(Vec v, PetscScalar **a);
PetscErrorCode VecGetArrayRead(Vec v, const PetscScalar **a);
PetscErrorCode VecRestoreArrayRead(Vec v, const PetscScalar **a);
PetscErrorCode VecNorm(Vec v, NormType t, PetscReal *nrm);

/* Mat */
typedef struct {
  PetscInt k, j, i, c;
} MatStencil;
typedef enum { INSERT_VALUES, ADD_VALUES } InsertMode;
typedef enum { MAT_FLUSH_ASSEMBLY, MAT_FINAL_ASSEMBLY } MatAssemblyType;
PetscErrorCode MatSetValuesStencil(Mat A, PetscInt m, const MatStencil *row, PetscInt n,
                                   const MatStencil *col, const PetscScalar *v, InsertMode mode);
PetscErrorCode MatAssemblyBegin(Mat A, MatAssemblyType t);
PetscErrorCode MatAssemblyEnd(Mat A, MatAssemblyType t);
PetscErrorCode MatMult(Mat A, Vec x, Vec y);
PetscErrorCode MatScale(Mat A, PetscScalar a);
PetscErrorCode MatGetRow(Mat A, PetscInt row, PetscInt *ncols, const PetscInt **cols,
                         const PetscScalar **vals);
PetscErrorCode MatRestoreRow(Mat A, PetscInt row, PetscInt *ncols, const PetscInt **cols,
                             const PetscScalar **vals);
PetscErrorCode MatDestroy(Mat *A);
/* Object state: increases with every change of the object's values (the
 * mock knows one kind of object with a state, Mat). */
PetscErrorCode PetscObjectStateGet(PetscObject obj, PetscObjectState *state);

/* KSP / PC */
typedef struct mock_ksp *KSP;
typedef struct mock_pc *PC;
typedef const char *PCType;
#define PCSHELL "shell"
typedef enum { KSP_NORM_NONE, KSP_NORM_PRECONDITIONED, KSP_NORM_UNPRECONDITIONED } KSPNormType;
typedef int KSPConvergedReason;
PetscErrorCode KSPCreate(MPI_Comm comm, KSP *ksp);
PetscErrorCode KSPSetOperators(KSP ksp, Mat A, Mat P);
PetscErrorCode KSPGetPC(KSP ksp, PC *pc);
PetscErrorCode KSPSetNormType(KSP ksp, KSPNormType t);
PetscErrorCode KSPSetFromOptions(KSP ksp);
PetscErrorCode KSPSetUp(KSP ksp);
PetscErrorCode KSPSolve(KSP ksp, Vec b, Vec x);
PetscErrorCode KSPGetIterationNumber(KSP ksp, PetscInt *its);
PetscErrorCode KSPGetConvergedReason(KSP ksp, KSPConvergedReason *r);
PetscErrorCode KSPGetResidualNorm(KSP ksp, PetscReal *rn);
PetscErrorCode KSPDestroy(KSP *ksp);
PetscErrorCode PCSetType(PC pc, PCType t);
PetscErrorCode PCGetOperators(PC pc, Mat *Amat, Mat *Pmat);
PetscErrorCode PCShellSetContext(PC pc, void *ctx);
PetscErrorCode PCShellGetContext(PC pc, void *ctx);
PetscErrorCode PCShellSetSetUp(PC pc, PetscErrorCode (*f)(PC));
PetscErrorCode PCShellSetApply(PC pc, PetscErrorCode (*f)(PC, Vec, Vec));
PetscErrorCode PCShellSetName(PC pc, const char *name);

#endif
