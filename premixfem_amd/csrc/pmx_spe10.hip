// The reservoir loop of the reference's spe10 fork (src/spe10.c) on the
// device: total-mobility permeability, Peaceman producers and the injector
// (the pressure system's well term and right-hand side), the PCSHELL pressure
// solve with the previous pressure as initial guess, and the explicit upwind
// saturation update.  All fields stay in HBM in the reference's natural
// order arr[ez][ey][ex]; the preconditioner context receives kappa and the
// well term through its device entry points, so an IMPES step never copies
// a field to the host.
#include <chrono>
#include <cmath>
#include <vector>

#include "pmx_kernels.h"

namespace pmx {

constexpr double kPoroMin = 0.05;        // PORO_MIN (src/spe10.c:44)
constexpr double kSwBoundMin = 0.2;      // S_W_BOUND_MIN (:45)
constexpr double kSwBoundMax = 0.8;      // S_W_BOUND_MAX (:46)
constexpr int kSpBlock = 256;

// Fluid and well constants of SPE10_init (src/spe10.c:345-405).
struct SpPar {
  double mu_w, mu_o, S_wc, S_or, S_wi;
  double q_inj;     // injection_rate * B_w
  double bhp;       // bottomhole pressure
  double wi;        // well index
  double Hx, Hy, Hz, meas_elem;
  int M, N, P;
};

// get_mobi (:802-814): Corey exponent 2.
__device__ __forceinline__ void mobility(const SpPar& q, double S, double& lw, double& lt) {
  const double s = (S - q.S_wc) / (1.0 - q.S_wc - q.S_or);
  const double krw = s * s, kro = (1.0 - s) * (1.0 - s);
  lw = krw / q.mu_w;
  lt = krw / q.mu_w + kro / q.mu_o;
}

__device__ __forceinline__ bool is_producer(const SpPar& q, int x, int y) {
  return (x == 0 || x == q.M - 1) && (y == 0 || y == q.N - 1);
}

__device__ __forceinline__ bool is_injector(const SpPar& q, int x, int y) {
  return (x == q.M / 2 - 1 || x == q.M / 2) && (y == q.N / 2 - 1 || y == q.N / 2);
}

// INTER_upd_kappa (:620-648) and get_linear_system (:723-800) per cell:
// kappa_d = perm_d lambda_t(S_w); producers add |E| perm_x lambda_t WI to
// the diagonal and that times BHP to the rhs; the injector cells receive
// q_inj / (4 P).
__global__ void k_sp_system(SpPar q, const double* __restrict__ px, const double* __restrict__ py,
                            const double* __restrict__ pz, const double* __restrict__ sw,
                            double* __restrict__ kx, double* __restrict__ ky,
                            double* __restrict__ kz, double* __restrict__ wd,
                            double* __restrict__ rhs) {
  const size_t n = (size_t)q.M * q.N * q.P;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(c % q.M), y = (int)((c / q.M) % q.N);
    double lw, lt;
    mobility(q, sw[c], lw, lt);
    kx[c] = px[c] * lt;
    ky[c] = py[c] * lt;
    kz[c] = pz[c] * lt;
    double w = 0.0, r = 0.0;
    if (is_producer(q, x, y)) {
      w = q.meas_elem * px[c] * lt * q.wi;
      r = w * q.bhp;
    }
    if (is_injector(q, x, y)) r += q.q_inj / (4 * q.P);
    wd[c] = w;
    rhs[c] = r;
  }
}

// Saturation rate of get_next_S_w_multisteps (:1960-2040) for one cell:
// injection, producer outflow and the six upwind face fluxes
// f_w(upwind) * kbar * (p_nb - p_c) / H^2, divided by the porosity.  The
// face coefficient kbar uses kappa of the step (fixed over the substeps),
// f_w = lambda_w / lambda_t the current saturation.  Writes the rate and the
// block maximum.
__global__ void __launch_bounds__(kSpBlock) k_sp_rate(
    SpPar q, const double* __restrict__ permx, const double* __restrict__ kx,
    const double* __restrict__ ky, const double* __restrict__ kz, const double* __restrict__ poro,
    const double* __restrict__ p, const double* __restrict__ sw, double* __restrict__ rate,
    double* __restrict__ bmax) {
  __shared__ double red[kSpBlock / 64];
  const size_t n = (size_t)q.M * q.N * q.P;
  double mx = -1.0e300;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(c % q.M), y = (int)((c / q.M) % q.N), z = (int)(c / ((size_t)q.M * q.N));
    double r = 0.0;
    if (is_injector(q, x, y)) r += q.q_inj / (4 * q.P) / q.meas_elem;
    double lw, lt;
    mobility(q, sw[c], lw, lt);
    const double fw = lw / lt;
    const double pc = p[c];
    if (is_producer(q, x, y)) r += permx[c] * lw * q.wi * (q.bhp - pc);
    const double* kk[3] = {kx, ky, kz};
    const double H[3] = {q.Hx, q.Hy, q.Hz};
    const int co[3] = {x, y, z}, ext[3] = {q.M, q.N, q.P};
    const size_t st[3] = {1, (size_t)q.M, (size_t)q.M * q.N};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
      for (int sgn = 0; sgn < 2; ++sgn) {
        if (sgn == 0 ? co[a] == 0 : co[a] == ext[a] - 1) continue;
        const size_t o = sgn == 0 ? c - st[a] : c + st[a];
        const double kb = 2.0 / (1.0 / kk[a][o] + 1.0 / kk[a][c]);
        double lwo, lto;
        mobility(q, sw[o], lwo, lto);
        const double fup = pc > p[o] ? fw : lwo / lto;
        r += fup * kb * (p[o] - pc) / H[a] / H[a];
      }
    }
    const double ph = poro[c];
    r = ph > 0.0 ? r / ph : 0.0;
    rate[c] = r;
    mx = fmax(mx, r);
  }
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, red[w]);
    bmax[blockIdx.x] = m;
  }
}

// Max over the block maxima; the substep dt = DS_max / max is added to
// scal[0]; scal[2] = 1 when the maximum is negative (the reference's "water
// saturation is decreasing everywhere" error), and the update is skipped.
__global__ void k_sp_dt(const double* __restrict__ bmax, int nb, double DS_max,
                        double* __restrict__ scal) {
  __shared__ double red[16];
  double m = -1.0e300;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) m = fmax(m, bmax[i]);
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = fmax(r, red[w]);
    if (r < 0.0 || scal[2] != 0.0) {
      scal[1] = 0.0;
      scal[2] = 1.0;
    } else {
      scal[1] = DS_max / r;
      scal[0] += DS_max / r;
    }
  }
}

// S_w += dt * rate (VecAXPY of :2050).
__global__ void k_sp_update(const double* __restrict__ rate, const double* __restrict__ scal,
                            double* __restrict__ sw, size_t n) {
  const double dt = scal[1];
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x)
    sw[c] += dt * rate[c];
}

// Block min / max of S_w (VecMin / VecMax of :248-249).
__global__ void __launch_bounds__(kSpBlock) k_sp_range(const double* __restrict__ sw, size_t n,
                                                       double* __restrict__ part) {
  __shared__ double rmin[kSpBlock / 64], rmax[kSpBlock / 64];
  double lo = 1.0e300, hi = -1.0e300;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    lo = fmin(lo, sw[c]);
    hi = fmax(hi, sw[c]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o, 64));
    hi = fmax(hi, __shfl_xor(hi, o, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    rmin[threadIdx.x >> 6] = lo;
    rmax[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      lo = fmin(lo, rmin[w]);
      hi = fmax(hi, rmax[w]);
    }
    lo = fmin(lo, rmin[0]);
    hi = fmax(hi, rmax[0]);
    part[2 * blockIdx.x] = lo;
    part[2 * blockIdx.x + 1] = hi;
  }
}

__global__ void k_sp_fill(double* __restrict__ a, size_t n, double v, double lo) {
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x)
    a[c] = lo > 0.0 ? fmax(a[c], lo) : v;
}

}  // namespace pmx

using namespace pmx;

struct pmx_spe10 {
  SpPar q;
  size_t n = 0;
  int nblk = 0;
  hipStream_t stream = nullptr;
  DBuf<double> perm[3], poro, sw, p, kappa[3], wd, rhs, rate, bmax, scal, part;
  std::vector<double> hpart;
  pmx_spe10_step_fn hook = nullptr;
  void* hook_user = nullptr;
};

namespace {

double wall_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int sp_grid(const pmx_spe10* sp) { return sp->nblk; }

int spe10_system(pmx_spe10* sp) {
  hipLaunchKernelGGL(k_sp_system, dim3(sp_grid(sp)), dim3(kSpBlock), 0, sp->stream, sp->q,
                     sp->perm[0].p, sp->perm[1].p, sp->perm[2].p, sp->sw.p, sp->kappa[0].p,
                     sp->kappa[1].p, sp->kappa[2].p, sp->wd.p, sp->rhs.p);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

}  // namespace

extern "C" {

int pmx_spe10_create(pmx_spe10** out, const int* mesh, const double* H, const double* perm_x,
                     const double* perm_y, const double* perm_z, const double* poro) {
  if (!out || !mesh || !H || !perm_x || !perm_y || !perm_z || !poro) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  *out = nullptr;
  if (!(mesh[0] > 1 && mesh[1] > 1 && mesh[2] > 0) || !(H[0] > 0.0 && H[1] > 0.0 && H[2] > 0.0)) {
    set_error("spe10: bad mesh or cell size");
    return PMX_ERR_ARG_WRONG;
  }
  const size_t n = (size_t)mesh[0] * mesh[1] * mesh[2];
  const double* src[3] = {perm_x, perm_y, perm_z};
  for (int a = 0; a < 3; ++a)
    for (size_t i = 0; i < n; ++i)
      if (!(src[a][i] > 0.0) || !std::isfinite(src[a][i])) {
        set_error("spe10: perm[%d][%zu]=%g is not positive and finite", a, i, src[a][i]);
        return PMX_ERR_ARG_WRONG;
      }
  pmx_spe10* sp = new pmx_spe10();
  auto fail = [&](int rc) {
    pmx_spe10_destroy(sp);
    return rc;
  };
  SpPar& q = sp->q;
  q.M = mesh[0];
  q.N = mesh[1];
  q.P = mesh[2];
  q.Hx = H[0];
  q.Hy = H[1];
  q.Hz = H[2];
  q.meas_elem = H[0] * H[1] * H[2];
  // SPE10_init (:345-405).
  const double cp = 1.6786805555555556e-12;  // 1 cP in psi * day
  q.mu_w = 0.3 * cp;
  q.mu_o = 3.0 * cp;
  q.S_wc = 0.2;
  q.S_or = 0.2;
  q.S_wi = 0.2;
  q.q_inj = 5.0e3 * 5.61458 * 1.01;
  q.bhp = 4.0e3;
  const double rw = 1.0;
  const double r1 = std::sqrt(9.0 * q.Hx * q.Hx + q.Hy * q.Hy) / 2.0;
  const double r2 = std::sqrt(q.Hx * q.Hx + 9.0 * q.Hy * q.Hy) / 2.0;
  double t = q.Hy / q.Hx * std::log(r1 / rw) + q.Hx / q.Hy * std::log(r2 / rw);
  t *= 2.0 / M_PI;
  t -= 1.0;
  q.wi = (1.0 / q.Hx / q.Hx + 1.0 / q.Hy / q.Hy) / t;
  sp->n = n;
  sp->nblk = (int)std::min<size_t>((n + kSpBlock - 1) / kSpBlock, 2048);
  if (hipStreamCreate(&sp->stream) != hipSuccess) {
    set_error("spe10: hipStreamCreate failed");
    return fail(PMX_ERR_LIB);
  }
  int rc = PMX_SUCCESS;
  for (DBuf<double>* b : {&sp->perm[0], &sp->perm[1], &sp->perm[2], &sp->poro, &sp->sw, &sp->p,
                          &sp->kappa[0], &sp->kappa[1], &sp->kappa[2], &sp->wd, &sp->rhs,
                          &sp->rate})
    if ((rc = b->alloc(n))) return fail(rc);
  if ((rc = sp->bmax.alloc(sp->nblk)) || (rc = sp->part.alloc(2 * (size_t)sp->nblk)) ||
      (rc = sp->scal.alloc(4)))
    return fail(rc);
  for (int a = 0; a < 3; ++a)
    if (hipMemcpy(sp->perm[a].p, src[a], n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return fail(PMX_ERR_LIB);
  if (hipMemcpy(sp->poro.p, poro, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return fail(PMX_ERR_LIB);
  // get_modified_poro (:2084-2101), get_init_S_w (:2065-2082), P = 0.
  hipLaunchKernelGGL(k_sp_fill, dim3(sp->nblk), dim3(kSpBlock), 0, sp->stream, sp->poro.p, n, 0.0,
                     kPoroMin);
  hipLaunchKernelGGL(k_sp_fill, dim3(sp->nblk), dim3(kSpBlock), 0, sp->stream, sp->sw.p, n, q.S_wi,
                     0.0);
  hipLaunchKernelGGL(k_sp_fill, dim3(sp->nblk), dim3(kSpBlock), 0, sp->stream, sp->p.p, n, 0.0,
                     0.0);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(sp->stream) != hipSuccess) {
    set_error("spe10: initialisation kernels failed");
    return fail(PMX_ERR_LIB);
  }
  *out = sp;
  return PMX_SUCCESS;
}

int pmx_spe10_destroy(pmx_spe10* sp) {
  if (!sp) return PMX_SUCCESS;
  if (sp->stream) (void)hipStreamSynchronize(sp->stream);
  for (DBuf<double>* b : {&sp->perm[0], &sp->perm[1], &sp->perm[2], &sp->poro, &sp->sw, &sp->p,
                          &sp->kappa[0], &sp->kappa[1], &sp->kappa[2], &sp->wd, &sp->rhs,
                          &sp->rate, &sp->bmax, &sp->scal, &sp->part})
    b->release();
  if (sp->stream) (void)hipStreamDestroy(sp->stream);
  delete sp;
  return PMX_SUCCESS;
}

static int check_ctx(pmx_spe10* sp, pmx_ctx* ctx) {
  if (!sp || !ctx) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  if (ctx->mesh[0] != sp->q.M || ctx->mesh[1] != sp->q.N || ctx->mesh[2] != sp->q.P) {
    set_error("spe10: the preconditioner context has another mesh");
    return PMX_ERR_ARG_WRONG;
  }
  return PMX_SUCCESS;
}

int pmx_spe10_assemble(pmx_spe10* sp, pmx_ctx* ctx) {
  PMX_TRY(check_ctx(sp, ctx));
  PMX_TRY(spe10_system(sp));
  PMX_HIP(hipStreamSynchronize(sp->stream));
  PMX_TRY(pmx_set_kappa_device(ctx, sp->kappa[0].p, sp->kappa[1].p, sp->kappa[2].p));
  PMX_TRY(pmx_set_well_term_device(ctx, sp->wd.p));
  return PMX_SUCCESS;
}

int pmx_spe10_get_system(pmx_spe10* sp, double* kx, double* ky, double* kz, double* wd,
                         double* rhs) {
  if (!sp) return PMX_ERR_ARG_NULL;
  PMX_TRY(spe10_system(sp));
  PMX_HIP(hipStreamSynchronize(sp->stream));
  double* dst[5] = {kx, ky, kz, wd, rhs};
  const double* srcs[5] = {sp->kappa[0].p, sp->kappa[1].p, sp->kappa[2].p, sp->wd.p, sp->rhs.p};
  for (int i = 0; i < 5; ++i)
    if (dst[i])
      PMX_HIP(hipMemcpy(dst[i], srcs[i], sp->n * sizeof(double), hipMemcpyDeviceToHost));
  return PMX_SUCCESS;
}

int pmx_spe10_solve(pmx_spe10* sp, pmx_ctx* ctx, double rtol, int max_it, int* its,
                    double* rnorm) {
  PMX_TRY(check_ctx(sp, ctx));
  PMX_TRY(pmx_setup(ctx));
  return pmx_solve_device(ctx, sp->rhs.p, sp->p.p, 1, rtol, max_it, its, rnorm);
}

int pmx_spe10_transport(pmx_spe10* sp, int steps, double DS_max, double* delta_t) {
  if (!sp) return PMX_ERR_ARG_NULL;
  if (steps < 0 || !(DS_max > 0.0)) {
    set_error("spe10: transport needs steps >= 0 and DS_max > 0");
    return PMX_ERR_ARG_WRONG;
  }
  PMX_HIP(hipMemsetAsync(sp->scal.p, 0, 4 * sizeof(double), sp->stream));
  for (int s = 0; s < steps; ++s) {
    hipLaunchKernelGGL(k_sp_rate, dim3(sp->nblk), dim3(kSpBlock), 0, sp->stream, sp->q,
                       sp->perm[0].p, sp->kappa[0].p, sp->kappa[1].p, sp->kappa[2].p, sp->poro.p,
                       sp->p.p, sp->sw.p, sp->rate.p, sp->bmax.p);
    hipLaunchKernelGGL(k_sp_dt, dim3(1), dim3(1024), 0, sp->stream, sp->bmax.p, sp->nblk, DS_max,
                       sp->scal.p);
    hipLaunchKernelGGL(k_sp_update, dim3(sp->nblk), dim3(kSpBlock), 0, sp->stream, sp->rate.p,
                       sp->scal.p, sp->sw.p, sp->n);
    PMX_HIP(hipGetLastError());
  }
  double h[4];
  PMX_HIP(hipMemcpyAsync(h, sp->scal.p, sizeof(h), hipMemcpyDeviceToHost, sp->stream));
  PMX_HIP(hipStreamSynchronize(sp->stream));
  if (h[2] != 0.0) {
    set_error("Water saturation is decreasing everywhere, this is impossible!");
    return PMX_ERR_USER;
  }
  if (delta_t) *delta_t = h[0];
  return PMX_SUCCESS;
}

int pmx_spe10_saturation_range(pmx_spe10* sp, double* smin, double* smax) {
  if (!sp || !smin || !smax) return PMX_ERR_ARG_NULL;
  hipLaunchKernelGGL(k_sp_range, dim3(sp->nblk), dim3(kSpBlock), 0, sp->stream, sp->sw.p, sp->n,
                     sp->part.p);
  PMX_HIP(hipGetLastError());
  sp->hpart.resize(2 * (size_t)sp->nblk);
  PMX_HIP(hipMemcpyAsync(sp->hpart.data(), sp->part.p, sp->hpart.size() * sizeof(double),
                         hipMemcpyDeviceToHost, sp->stream));
  PMX_HIP(hipStreamSynchronize(sp->stream));
  double lo = 1.0e300, hi = -1.0e300;
  for (int b = 0; b < sp->nblk; ++b) {
    lo = std::min(lo, sp->hpart[2 * b]);
    hi = std::max(hi, sp->hpart[2 * b + 1]);
  }
  *smin = lo;
  *smax = hi;
  return PMX_SUCCESS;
}

int pmx_spe10_get_state(const pmx_spe10* sp, double* S_w, double* P) {
  if (!sp) return PMX_ERR_ARG_NULL;
  if (S_w) PMX_HIP(hipMemcpy(S_w, sp->sw.p, sp->n * sizeof(double), hipMemcpyDeviceToHost));
  if (P) PMX_HIP(hipMemcpy(P, sp->p.p, sp->n * sizeof(double), hipMemcpyDeviceToHost));
  return PMX_SUCCESS;
}

int pmx_spe10_set_state(pmx_spe10* sp, const double* S_w, const double* P) {
  if (!sp) return PMX_ERR_ARG_NULL;
  if (S_w) PMX_HIP(hipMemcpy(sp->sw.p, S_w, sp->n * sizeof(double), hipMemcpyHostToDevice));
  if (P) PMX_HIP(hipMemcpy(sp->p.p, P, sp->n * sizeof(double), hipMemcpyHostToDevice));
  return PMX_SUCCESS;
}

int pmx_spe10_set_step_hook(pmx_spe10* sp, pmx_spe10_step_fn cb, void* user) {
  if (!sp) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  sp->hook = cb;
  sp->hook_user = user;
  return PMX_SUCCESS;
}

int pmx_spe10_run(pmx_spe10* sp, pmx_ctx* ctx, int max_steps, const double* simu_time_stages,
                  const double* DS_max_stages, const int* steps_stages, double rtol, int max_it,
                  pmx_spe10_log* log, int* steps_done) {
  PMX_TRY(check_ctx(sp, ctx));
  if (!simu_time_stages || !DS_max_stages || !steps_stages) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  constexpr int kStages = 6;  // SIMU_STAGES_NUM (src/spe10.c:36)
  double simu_time = 0.0;
  int stage = 1;
  if (steps_done) *steps_done = 0;
  for (int ind = 0; ind < max_steps; ++ind) {
    pmx_spe10_log row{};
    const double t0 = wall_seconds();
    PMX_TRY(pmx_spe10_assemble(sp, ctx));
    const double t1 = wall_seconds();
    int its = 0;
    double rnorm = 0.0;
    PMX_TRY(pmx_setup(ctx));
    const double t2 = wall_seconds();
    const int rc = pmx_solve_device(ctx, sp->rhs.p, sp->p.p, 1, rtol, max_it, &its, &rnorm);
    if (rc != PMX_SUCCESS && rc != PMX_ERR_NOT_CONVERGED) return rc;
    const double t3 = wall_seconds();
    double smin = 0.0, smax = 0.0;
    PMX_TRY(pmx_spe10_saturation_range(sp, &smin, &smax));
    row.step = ind;
    row.simu_time = simu_time;
    row.stage = stage;
    row.DS_max = DS_max_stages[stage];
    row.S_w_steps = steps_stages[stage];
    row.t_assemble = t1 - t0;
    row.t_setup = t2 - t1;
    row.t_solve = t3 - t2;
    row.its = its;
    row.reason = ctx->last_reason;
    row.S_w_min = smin;
    row.S_w_max = smax;
    if (!(kSwBoundMin - kNil < smin && kSwBoundMax + kNil > smax)) {
      set_error("S_w=[%.5E, %.5E] exceeds the bounds.", smin, smax);
      if (log) log[ind] = row;
      if (steps_done) *steps_done = ind + 1;
      return PMX_ERR_USER;
    }
    if (sp->hook) {
      PMX_HIP(hipStreamSynchronize(sp->stream));
      const int hr = sp->hook(sp->hook_user, &row);
      if (hr != 0) {
        set_error("IMPES step hook returned %d at step %d", hr, ind);
        if (log) log[ind] = row;
        if (steps_done) *steps_done = ind + 1;
        return hr;
      }
    }
    if (simu_time > simu_time_stages[kStages]) {
      if (log) log[ind] = row;
      if (steps_done) *steps_done = ind + 1;
      break;
    }
    double dt = 0.0;
    const double t4 = wall_seconds();
    PMX_TRY(pmx_spe10_transport(sp, steps_stages[stage], DS_max_stages[stage], &dt));
    row.t_transport = wall_seconds() - t4;
    row.delta_t = dt;
    simu_time += dt;
    if (simu_time > simu_time_stages[stage]) stage = std::min(stage + 1, kStages);
    if (log) log[ind] = row;
    if (steps_done) *steps_done = ind + 1;
  }
  return PMX_SUCCESS;
}

}  // extern "C"
