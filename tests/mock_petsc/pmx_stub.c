/* CPU stand-in for the GPU library's C ABI, for exercising the PCSHELL
 * adapter where no GPU is present: a 7-point operator from kappa (harmonic
 * face averages) and a Jacobi preconditioner.  It records the options string
 * it was given (printed as "STUB options: ...").  Test infrastructure only. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "premixfem_gpu.h"

struct pmx_ctx {
  int M, N, P;
  double *kx, *ky, *kz, *diag;
  int applies;
};

static char g_err[256];

const char *pmx_last_error(void) { return g_err; }

int pmx_init(pmx_ctx **out, const double *dom, const int *mesh, const char *options) {
  (void)dom;
  printf("STUB options: %s\n", options ? options : "");
  if (options && strstr(options, "-en_lv1 0")) {
    snprintf(g_err, sizeof(g_err), "Error in max_eigen_num_lv1=0 for the level-1 problem.");
    return PMX_ERR_ARG_WRONG;
  }
  *out = calloc(1, sizeof(pmx_ctx));
  (*out)->M = mesh[0];
  (*out)->N = mesh[1];
  (*out)->P = mesh[2];
  return PMX_SUCCESS;
}

int pmx_print_info(const pmx_ctx *c) {
  printf("STUB mesh %d %d %d\n", c->M, c->N, c->P);
  return PMX_SUCCESS;
}

static double *dup(const double *a, size_t n) {
  double *b = malloc(n * sizeof(double));
  memcpy(b, a, n * sizeof(double));
  return b;
}

int pmx_set_kappa(pmx_ctx *c, const double *kx, const double *ky, const double *kz) {
  const size_t n = (size_t)c->M * c->N * c->P;
  free(c->kx);
  free(c->ky);
  free(c->kz);
  c->kx = dup(kx, n);
  c->ky = dup(ky, n);
  c->kz = dup(kz, n);
  return PMX_SUCCESS;
}

static double harm(double a, double b) { return 2.0 / (1.0 / a + 1.0 / b); }

int pmx_create_A(pmx_ctx *c, double *diag, double *off) {
  const int M = c->M, N = c->N, P = c->P;
  const size_t n = (size_t)M * N * P;
  int ex, ey, ez;
  if (!c->kx) {
    snprintf(g_err, sizeof(g_err), "kappa not set");
    return PMX_ERR_ORDER;
  }
  memset(off, 0, 6 * n * sizeof(double));
  for (ez = 0; ez < P; ++ez)
    for (ey = 0; ey < N; ++ey)
      for (ex = 0; ex < M; ++ex) {
        const size_t i = ((size_t)ez * N + ey) * M + ex;
        if (ex > 0) off[6 * i + 0] = -harm(c->kx[i], c->kx[i - 1]);
        if (ex < M - 1) off[6 * i + 1] = -harm(c->kx[i], c->kx[i + 1]);
        if (ey > 0) off[6 * i + 2] = -harm(c->ky[i], c->ky[i - M]);
        if (ey < N - 1) off[6 * i + 3] = -harm(c->ky[i], c->ky[i + M]);
        if (ez > 0) off[6 * i + 4] = -harm(c->kz[i], c->kz[i - (size_t)M * N]);
        if (ez < P - 1) off[6 * i + 5] = -harm(c->kz[i], c->kz[i + (size_t)M * N]);
        diag[i] = 0.0;
        for (int d = 0; d < 6; ++d) diag[i] -= off[6 * i + d];
      }
  free(c->diag);
  c->diag = dup(diag, n);
  return PMX_SUCCESS;
}

int pmx_setup(pmx_ctx *c) {
  if (!c->diag) {
    snprintf(g_err, sizeof(g_err), "PC_create_A must precede PC_setup");
    return PMX_ERR_ORDER;
  }
  return PMX_SUCCESS;
}

int pmx_apply(pmx_ctx *c, const double *x, double *y) {
  const size_t n = (size_t)c->M * c->N * c->P;
  for (size_t i = 0; i < n; ++i) y[i] = x[i] / c->diag[i];
  c->applies++;
  return PMX_SUCCESS;
}

int pmx_solve(pmx_ctx *c, const double *b, double *x, int x0_nonzero, double rtol, int max_it,
              int *its, double *rnorm) {
  (void)c, (void)b, (void)x, (void)x0_nonzero, (void)rtol, (void)max_it, (void)its, (void)rnorm;
  snprintf(g_err, sizeof(g_err), "the stub has no PCG");
  return PMX_ERR_SUP;
}

int pmx_get_stats(const pmx_ctx *c, pmx_stats *st) {
  memset(st, 0, sizeof(*st));
  st->M = c->M;
  st->N = c->N;
  st->P = c->P;
  return PMX_SUCCESS;
}

int pmx_print_stat(const pmx_ctx *c) {
  printf("STUB applies %d\n", c->applies);
  return PMX_SUCCESS;
}

int pmx_final(pmx_ctx *c) {
  free(c->kx);
  free(c->ky);
  free(c->kz);
  free(c->diag);
  free(c);
  return PMX_SUCCESS;
}
