/* Sequential implementation of the PETSc/MPI subset declared in
 * petscksp.h / mpi.h.  Test infrastructure only. */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "petscksp.h"

static int g_argc;
static char **g_argv;

int MPI_Comm_size(MPI_Comm comm, int *size) {
  (void)comm;
  *size = 1;
  return MPI_SUCCESS;
}

int MPI_Reduce(const void *send, void *recv, int count, MPI_Datatype type, MPI_Op op, int root,
               MPI_Comm comm) {
  (void)op;
  (void)root;
  (void)comm;
  memcpy(recv, send, (size_t)count * (type == MPI_DOUBLE ? sizeof(double) : sizeof(int)));
  return MPI_SUCCESS;
}

int mock_error(int code, const char *file, int line, const char *fmt, ...) {
  va_list ap;
  fprintf(stderr, "[mock-petsc] error %d at %s:%d: ", code, file, line);
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  fprintf(stderr, "\n");
  return code;
}

void mock_trace(int code, const char *file, int line) {
  fprintf(stderr, "[mock-petsc]   from %s:%d (code %d)\n", file, line, code);
}

PetscErrorCode PetscInitialize(int *argc, char ***argv, const char *file, const char *help) {
  (void)file;
  (void)help;
  g_argc = *argc;
  g_argv = *argv;
  return PETSC_SUCCESS;
}

PetscErrorCode PetscFinalize(void) { return PETSC_SUCCESS; }

PetscErrorCode PetscPrintf(MPI_Comm comm, const char *fmt, ...) {
  va_list ap;
  (void)comm;
  va_start(ap, fmt);
  vprintf(fmt, ap);
  va_end(ap);
  return PETSC_SUCCESS;
}

static const char *opt_find(const char *name, PetscBool *found) {
  int i;
  *found = PETSC_FALSE;
  for (i = 1; i < g_argc; ++i) {
    if (strcmp(g_argv[i], name) == 0) {
      *found = PETSC_TRUE;
      if (i + 1 < g_argc && !(g_argv[i + 1][0] == '-' && g_argv[i + 1][1] >= 'a' &&
                              g_argv[i + 1][1] <= 'z'))
        return g_argv[i + 1];
      return "";
    }
  }
  return NULL;
}

PetscErrorCode PetscOptionsGetInt(void *opts, const char *pre, const char *name, PetscInt *v,
                                  PetscBool *set) {
  PetscBool f;
  const char *s = opt_find(name, &f);
  (void)opts;
  (void)pre;
  if (set) *set = PETSC_FALSE;
  if (f && s && *s) {
    char *end;
    long x = strtol(s, &end, 10);
    if (*end) return mock_error(PETSC_ERR_ARG_WRONG, __FILE__, __LINE__, "bad int %s", s);
    *v = (PetscInt)x;
    if (set) *set = PETSC_TRUE;
  }
  return PETSC_SUCCESS;
}

PetscErrorCode PetscOptionsGetScalar(void *opts, const char *pre, const char *name,
                                     PetscScalar *v, PetscBool *set) {
  PetscBool f;
  const char *s = opt_find(name, &f);
  (void)opts;
  (void)pre;
  if (set) *set = PETSC_FALSE;
  if (f && s && *s) {
    char *end;
    double x = strtod(s, &end);
    if (*end) return mock_error(PETSC_ERR_ARG_WRONG, __FILE__, __LINE__, "bad real %s", s);
    *v = x;
    if (set) *set = PETSC_TRUE;
  }
  return PETSC_SUCCESS;
}

PetscErrorCode PetscOptionsGetString(void *opts, const char *pre, const char *name, char *v,
                                     size_t len, PetscBool *set) {
  PetscBool f;
  const char *s = opt_find(name, &f);
  (void)opts;
  (void)pre;
  *set = PETSC_FALSE;
  if (f && s) {
    snprintf(v, len, "%s", s);
    *set = PETSC_TRUE;
  }
  return PETSC_SUCCESS;
}

PetscErrorCode PetscOptionsHasName(void *opts, const char *pre, const char *name,
                                   PetscBool *set) {
  (void)opts;
  (void)pre;
  opt_find(name, set);
  return PETSC_SUCCESS;
}

PetscErrorCode mock_malloc2(size_t b1, void *p1, size_t b2, void *p2) {
  *(void **)p1 = malloc(b1 ? b1 : 1);
  *(void **)p2 = malloc(b2 ? b2 : 1);
  if (!*(void **)p1 || !*(void **)p2) return PETSC_ERR_MEM;
  return PETSC_SUCCESS;
}

PetscErrorCode mock_free2(void *p1, void *p2) {
  free(p1);
  free(p2);
  return PETSC_SUCCESS;
}

/* ---- DMDA: one rank owns the whole M x N x P box, natural order ---- */
struct mock_dm {
  PetscInt M, N, P;
};
struct mock_vec {
  PetscInt n;
  double *a;
};
struct mock_mat {
  PetscObjectState state; /* first member: PetscObjectStateGet reads it */
  PetscInt M, N, P;
  double *v; /* [cell][7]: diag, -x, +x, -y, +y, -z, +z */
  PetscInt row_cols[7];   /* MatGetRow's output of the current row */
  PetscScalar row_vals[7];
};

PetscErrorCode DMDACreate3d(MPI_Comm comm, DMBoundaryType bx, DMBoundaryType by,
                            DMBoundaryType bz, DMDAStencilType st, PetscInt M, PetscInt N,
                            PetscInt P, PetscInt m, PetscInt n, PetscInt p, PetscInt dof,
                            PetscInt s, const PetscInt *lx, const PetscInt *ly,
                            const PetscInt *lz, DM *da) {
  (void)comm, (void)bx, (void)by, (void)bz, (void)st, (void)m, (void)n, (void)p, (void)lx,
      (void)ly, (void)lz;
  if (dof != 1 || s != 1) return PETSC_ERR_SUP;
  *da = calloc(1, sizeof(struct mock_dm));
  (*da)->M = M;
  (*da)->N = N;
  (*da)->P = P;
  return PETSC_SUCCESS;
}

PetscErrorCode DMSetUp(DM dm) {
  (void)dm;
  return PETSC_SUCCESS;
}

PetscErrorCode DMDestroy(DM *dm) {
  free(*dm);
  *dm = NULL;
  return PETSC_SUCCESS;
}

static Vec vec_new(PetscInt n) {
  Vec v = calloc(1, sizeof(struct mock_vec));
  v->n = n;
  v->a = calloc((size_t)n, sizeof(double));
  return v;
}

PetscErrorCode DMCreateGlobalVector(DM dm, Vec *v) {
  *v = vec_new(dm->M * dm->N * dm->P);
  return PETSC_SUCCESS;
}

PetscErrorCode DMCreateMatrix(DM dm, Mat *A) {
  *A = calloc(1, sizeof(struct mock_mat));
  (*A)->M = dm->M;
  (*A)->N = dm->N;
  (*A)->P = dm->P;
  (*A)->v = calloc((size_t)dm->M * dm->N * dm->P * 7, sizeof(double));
  return PETSC_SUCCESS;
}

PetscErrorCode DMDAGetCorners(DM dm, PetscInt *x, PetscInt *y, PetscInt *z, PetscInt *m,
                              PetscInt *n, PetscInt *p) {
  *x = *y = *z = 0;
  *m = dm->M;
  *n = dm->N;
  *p = dm->P;
  return PETSC_SUCCESS;
}

PetscErrorCode VecDuplicate(Vec v, Vec *w) {
  *w = vec_new(v->n);
  return PETSC_SUCCESS;
}

PetscErrorCode VecDestroy(Vec *v) {
  if (*v) free((*v)->a);
  free(*v);
  *v = NULL;
  return PETSC_SUCCESS;
}

PetscErrorCode VecSet(Vec v, PetscScalar a) {
  PetscInt i;
  for (i = 0; i < v->n; ++i) v->a[i] = a;
  return PETSC_SUCCESS;
}

PetscErrorCode VecGetSize(Vec v, PetscInt *n) {
  *n = v->n;
  return PETSC_SUCCESS;
}

PetscErrorCode VecGetArray(Vec v, PetscScalar **a) {
  *a = v->a;
  return PETSC_SUCCESS;
}

PetscErrorCode VecRestoreArray(Vec v, PetscScalar **a) {
  (void)v;
  *a = NULL;
  return PETSC_SUCCESS;
}

PetscErrorCode VecGetArrayRead(Vec v, const PetscScalar **a) {
  *a = v->a;
  return PETSC_SUCCESS;
}

PetscErrorCode VecRestoreArrayRead(Vec v, const PetscScalar **a) {
  (void)v;
  *a = NULL;
  return PETSC_SUCCESS;
}

static double dot(const double *x, const double *y, PetscInt n) {
  double s = 0.0;
  PetscInt i;
  for (i = 0; i < n; ++i) s += x[i] * y[i];
  return s;
}

PetscErrorCode VecNorm(Vec v, NormType t, PetscReal *nrm) {
  (void)t;
  *nrm = sqrt(dot(v->a, v->a, v->n));
  return PETSC_SUCCESS;
}

PetscErrorCode MatSetValuesStencil(Mat A, PetscInt m, const MatStencil *row, PetscInt n,
                                   const MatStencil *col, const PetscScalar *v, InsertMode mode) {
  PetscInt r, c;
  for (r = 0; r < m; ++r) {
    const PetscInt cell = (row[r].k * A->N + row[r].j) * A->M + row[r].i;
    for (c = 0; c < n; ++c) {
      const int dx = col[c].i - row[r].i, dy = col[c].j - row[r].j, dz = col[c].k - row[r].k;
      int slot;
      if (!dx && !dy && !dz)
        slot = 0;
      else if (dx == -1 && !dy && !dz)
        slot = 1;
      else if (dx == 1 && !dy && !dz)
        slot = 2;
      else if (!dx && dy == -1 && !dz)
        slot = 3;
      else if (!dx && dy == 1 && !dz)
        slot = 4;
      else if (!dx && !dy && dz == -1)
        slot = 5;
      else if (!dx && !dy && dz == 1)
        slot = 6;
      else
        return mock_error(PETSC_ERR_SUP, __FILE__, __LINE__, "entry outside the 7-point star");
      if (mode == INSERT_VALUES)
        A->v[(size_t)cell * 7 + slot] = v[r * n + c];
      else
        A->v[(size_t)cell * 7 + slot] += v[r * n + c];
    }
  }
  A->state++;
  return PETSC_SUCCESS;
}

PetscErrorCode MatScale(Mat A, PetscScalar a) {
  size_t i, n = (size_t)A->M * A->N * A->P * 7;
  for (i = 0; i < n; ++i) A->v[i] *= a;
  A->state++;
  return PETSC_SUCCESS;
}

/* The stored entries of one row in increasing column order (the 7-point
 * star inside the grid). */
PetscErrorCode MatGetRow(Mat A, PetscInt row, PetscInt *ncols, const PetscInt **cols,
                         const PetscScalar **vals) {
  const PetscInt M = A->M, N = A->N, P = A->P;
  const PetscInt ex = row % M, ey = (row / M) % N, ez = row / (M * N);
  const double *v = A->v + (size_t)row * 7;
  PetscInt k = 0;
  if (row < 0 || row >= M * N * P) return mock_error(PETSC_ERR_ARG_OUTOFRANGE, __FILE__, __LINE__, "row");
#define MOCK_ROW_ENTRY(cond, col, slot) \
  if (cond) {                          \
    A->row_cols[k] = (col);            \
    A->row_vals[k++] = v[slot];        \
  }
  MOCK_ROW_ENTRY(ez > 0, row - M * N, 5)
  MOCK_ROW_ENTRY(ey > 0, row - M, 3)
  MOCK_ROW_ENTRY(ex > 0, row - 1, 1)
  MOCK_ROW_ENTRY(1, row, 0)
  MOCK_ROW_ENTRY(ex < M - 1, row + 1, 2)
  MOCK_ROW_ENTRY(ey < N - 1, row + M, 4)
  MOCK_ROW_ENTRY(ez < P - 1, row + M * N, 6)
#undef MOCK_ROW_ENTRY
  *ncols = k;
  *cols = A->row_cols;
  *vals = A->row_vals;
  return PETSC_SUCCESS;
}

PetscErrorCode MatRestoreRow(Mat A, PetscInt row, PetscInt *ncols, const PetscInt **cols,
                             const PetscScalar **vals) {
  (void)A, (void)row;
  *ncols = 0;
  *cols = NULL;
  *vals = NULL;
  return PETSC_SUCCESS;
}

PetscErrorCode PetscObjectStateGet(PetscObject obj, PetscObjectState *state) {
  *state = ((struct mock_mat *)obj)->state;
  return PETSC_SUCCESS;
}

PetscErrorCode MatAssemblyBegin(Mat A, MatAssemblyType t) {
  (void)A;
  (void)t;
  return PETSC_SUCCESS;
}

PetscErrorCode MatAssemblyEnd(Mat A, MatAssemblyType t) {
  (void)A;
  (void)t;
  return PETSC_SUCCESS;
}

PetscErrorCode MatMult(Mat A, Vec x, Vec y) {
  PetscInt ex, ey, ez;
  const PetscInt M = A->M, N = A->N, P = A->P;
  for (ez = 0; ez < P; ++ez)
    for (ey = 0; ey < N; ++ey)
      for (ex = 0; ex < M; ++ex) {
        const PetscInt c = (ez * N + ey) * M + ex;
        const double *v = A->v + (size_t)c * 7;
        double s = v[0] * x->a[c];
        if (ex > 0) s += v[1] * x->a[c - 1];
        if (ex < M - 1) s += v[2] * x->a[c + 1];
        if (ey > 0) s += v[3] * x->a[c - M];
        if (ey < N - 1) s += v[4] * x->a[c + M];
        if (ez > 0) s += v[5] * x->a[c - M * N];
        if (ez < P - 1) s += v[6] * x->a[c + M * N];
        y->a[c] = s;
      }
  return PETSC_SUCCESS;
}

PetscErrorCode MatDestroy(Mat *A) {
  if (*A) free((*A)->v);
  free(*A);
  *A = NULL;
  return PETSC_SUCCESS;
}

/* ---- KSP (CG) and PCSHELL ---- */
struct mock_pc {
  Mat A;                      /* operator (KSPSetOperators) */
  PetscObjectState setup_state; /* its state at the last setup */
  void *ctx;
  PetscErrorCode (*setup)(PC);
  PetscErrorCode (*apply)(PC, Vec, Vec);
  int is_shell, set_up;
};
struct mock_ksp {
  Mat A;
  struct mock_pc pc;
  KSPNormType norm;
  double rtol, abstol, dtol, rnorm;
  PetscInt max_it, its;
  KSPConvergedReason reason;
};

PetscErrorCode KSPCreate(MPI_Comm comm, KSP *ksp) {
  (void)comm;
  *ksp = calloc(1, sizeof(struct mock_ksp));
  (*ksp)->rtol = 1e-5;
  (*ksp)->abstol = 1e-50;
  (*ksp)->dtol = 1e5;
  (*ksp)->max_it = 10000;
  (*ksp)->norm = KSP_NORM_PRECONDITIONED;
  return PETSC_SUCCESS;
}

PetscErrorCode KSPSetOperators(KSP ksp, Mat A, Mat P) {
  (void)P;
  ksp->A = A;
  ksp->pc.A = A;
  ksp->pc.set_up = 0;
  return PETSC_SUCCESS;
}

PetscErrorCode PCGetOperators(PC pc, Mat *Amat, Mat *Pmat) {
  if (Amat) *Amat = pc->A;
  if (Pmat) *Pmat = pc->A;
  return PETSC_SUCCESS;
}

PetscErrorCode KSPGetPC(KSP ksp, PC *pc) {
  *pc = &ksp->pc;
  return PETSC_SUCCESS;
}

PetscErrorCode KSPSetNormType(KSP ksp, KSPNormType t) {
  ksp->norm = t;
  return PETSC_SUCCESS;
}

PetscErrorCode KSPSetFromOptions(KSP ksp) {
  PetscCall(PetscOptionsGetScalar(NULL, NULL, "-ksp_rtol", &ksp->rtol, NULL));
  PetscCall(PetscOptionsGetScalar(NULL, NULL, "-ksp_atol", &ksp->abstol, NULL));
  PetscCall(PetscOptionsGetInt(NULL, NULL, "-ksp_max_it", &ksp->max_it, NULL));
  return PETSC_SUCCESS;
}

/* PCSetUp semantics: the shell's setup runs again when the operator was
 * set anew or its values changed since the last setup. */
PetscErrorCode KSPSetUp(KSP ksp) {
  if (ksp->pc.is_shell && ksp->pc.setup &&
      (!ksp->pc.set_up || ksp->A->state != ksp->pc.setup_state)) {
    PetscCall(ksp->pc.setup(&ksp->pc));
    ksp->pc.set_up = 1;
    ksp->pc.setup_state = ksp->A->state;
  }
  return PETSC_SUCCESS;
}

/* CG with the unpreconditioned residual norm and the default test
 * ||r_k|| <= max(rtol ||r_0||, abstol), divergence at dtol ||r_0||. */
PetscErrorCode KSPSolve(KSP ksp, Vec b, Vec x) {
  const PetscInt n = b->n;
  Vec r = vec_new(n), z = vec_new(n), p = vec_new(n), w = vec_new(n);
  double beta = 0.0, betaold = 1.0, dp, rnorm0, ttol;
  PetscInt i, it;
  if (ksp->norm != KSP_NORM_UNPRECONDITIONED)
    return mock_error(PETSC_ERR_SUP, __FILE__, __LINE__, "mock KSP: unpreconditioned norm only");
  PetscCall(KSPSetUp(ksp));
  PetscCall(MatMult(ksp->A, x, r));
  for (i = 0; i < n; ++i) r->a[i] = b->a[i] - r->a[i];
  dp = sqrt(dot(r->a, r->a, n));
  rnorm0 = dp;
  ttol = fmax(ksp->rtol * rnorm0, ksp->abstol);
  ksp->reason = 0;
  ksp->its = 0;
  ksp->rnorm = dp;
  if (dp <= ttol) ksp->reason = dp <= ksp->abstol ? 3 : 2;
  for (it = 0; !ksp->reason && it < ksp->max_it; ++it) {
    double dpi, a;
    if (ksp->pc.is_shell) {
      PetscCall(ksp->pc.apply(&ksp->pc, r, z));
    } else {
      memcpy(z->a, r->a, (size_t)n * sizeof(double));
    }
    beta = dot(z->a, r->a, n);
    if (beta < 0.0) {
      ksp->reason = -8;
      break;
    }
    if (it == 0) {
      memcpy(p->a, z->a, (size_t)n * sizeof(double));
    } else {
      const double bb = beta / betaold;
      for (i = 0; i < n; ++i) p->a[i] = z->a[i] + bb * p->a[i];
    }
    betaold = beta;
    PetscCall(MatMult(ksp->A, p, w));
    dpi = dot(p->a, w->a, n);
    if (dpi <= 0.0) {
      ksp->reason = -8;
      break;
    }
    a = beta / dpi;
    for (i = 0; i < n; ++i) {
      x->a[i] += a * p->a[i];
      r->a[i] -= a * w->a[i];
    }
    dp = sqrt(dot(r->a, r->a, n));
    ksp->its = it + 1;
    ksp->rnorm = dp;
    if (dp <= ttol)
      ksp->reason = dp <= ksp->abstol ? 3 : 2;
    else if (dp >= ksp->dtol * rnorm0)
      ksp->reason = -4;
  }
  if (!ksp->reason) ksp->reason = -3;
  VecDestroy(&r);
  VecDestroy(&z);
  VecDestroy(&p);
  VecDestroy(&w);
  return PETSC_SUCCESS;
}

PetscErrorCode KSPGetIterationNumber(KSP ksp, PetscInt *its) {
  *its = ksp->its;
  return PETSC_SUCCESS;
}

PetscErrorCode KSPGetConvergedReason(KSP ksp, KSPConvergedReason *r) {
  *r = ksp->reason;
  return PETSC_SUCCESS;
}

PetscErrorCode KSPGetResidualNorm(KSP ksp, PetscReal *rn) {
  *rn = ksp->rnorm;
  return PETSC_SUCCESS;
}

PetscErrorCode KSPDestroy(KSP *ksp) {
  free(*ksp);
  *ksp = NULL;
  return PETSC_SUCCESS;
}

PetscErrorCode PCSetType(PC pc, PCType t) {
  pc->is_shell = strcmp(t, PCSHELL) == 0;
  return PETSC_SUCCESS;
}

PetscErrorCode PCShellSetContext(PC pc, void *ctx) {
  pc->ctx = ctx;
  return PETSC_SUCCESS;
}

PetscErrorCode PCShellGetContext(PC pc, void *ctx) {
  *(void **)ctx = pc->ctx;
  return PETSC_SUCCESS;
}

PetscErrorCode PCShellSetSetUp(PC pc, PetscErrorCode (*f)(PC)) {
  pc->setup = f;
  return PETSC_SUCCESS;
}

PetscErrorCode PCShellSetApply(PC pc, PetscErrorCode (*f)(PC, Vec, Vec)) {
  pc->apply = f;
  return PETSC_SUCCESS;
}

PetscErrorCode PCShellSetName(PC pc, const char *name) {
  (void)pc;
  (void)name;
  return PETSC_SUCCESS;
}
