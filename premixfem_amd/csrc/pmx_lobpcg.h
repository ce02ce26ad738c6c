// Shared pieces of the batched subdomain LOBPCG kernels (pmx_lobpcg.hip:
// the VALU kernel; pmx_lobpcg_mf.hip: the matrix-core kernel): launch
// arguments, limits and the per-workgroup building blocks both use.
#pragma once

#include <cstdint>

#include "pmx_dense.h"
#include "pmx_kernels.h"

namespace pmx {

constexpr int kLobCh = 4;      // preconditioner columns per sweep pair
// Threads per subdomain workgroup.  Large subdomains (>= kLobBigN cells) use
// 512 threads: 8 waves share the CU's register budget (one workgroup per
// CU), so a slow subdomain left running alone streams its blocks with twice
// the waves (MB = 8 at 16^3: 23 % less time per iteration; MB = 10 spills a
// few registers at this size and is still faster than with 256 threads).
// Small subdomains (<= kLobSmallN cells) use 128 threads: twice the
// subdomains in flight per CU (8^3 batch: 38 % less time; 64 threads were
// slower again).
constexpr int kLobWaves = 8;   // most waves per workgroup (LDS reduction slots)
constexpr int kLobBigN = 2048;
constexpr int kLobSmallN = 512;   // at most this many cells: 128-thread workgroups
constexpr int kStagIts = 40;   // iterations without progress before the stagnation test
// AX and AP follow X and P by block updates whose rounding errors
// accumulate; they are recomputed from A every kRefreshIts iterations once
// the best residual is within kRefreshNear x tol.
constexpr int kRefreshIts = 10;
constexpr double kRefreshNear = 1e3;

struct LobArgs {
  int kwant, maxit, kmax;
  int stride;   // workgroup b solves subdomain b * stride (> 1: autotuner sample)
  double tol;
  uint64_t seed;
  double* X;
  double* AX;
  double* W;
  double* AW;
  double* P;
  double* AP;
  const float4* pf;
  const float4* pb;
  const double* Tx;
  const double* Ty;
  const double* Tz;
  const double* mass;
  const double* x0;   // optional warm start rows [cell][MB]
  double* phi;        // output rows [cell][kmax]
  double* eig;        // output [nsub][kmax]
  int* its;           // output [nsub]
  int* status;        // output [nsub] 0 = converged
  int* fallback;      // output [nsub] 1 = the warm start failed, restarted at random
  int* stagnated;     // output [nsub] 1 = accepted at the stagnation test
  double stag_factor; // stagnation acceptance: worst residual <= stag_factor * tol
  int debug;          // subdomain whose iterations are traced with printf (-1: none)
  const int* order;   // workgroup b solves subdomain order[b] (nullptr: b * stride)
  int ortho_coef;     // matrix-core kernel: orthogonalise W in coefficient space from one
                      // Gram of [X P W] (1) or by two explicit projection passes (0)
};

__device__ __forceinline__ double hash_uniform(uint64_t a) {
  a += 0x9E3779B97F4A7C15ull;
  a = (a ^ (a >> 30)) * 0xBF58476D1CE4E5B9ull;
  a = (a ^ (a >> 27)) * 0x94D049BB133111EBull;
  a ^= a >> 31;
  return (double)(a >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// out[i*ldo + j] = sum_r L_i[r] wgt[r] R_j[r] for i < NL, j < NR (columns
// contiguous with stride n).  red: 4*NL*NR doubles of LDS.
template <int NL, int NR>
__device__ __forceinline__ void wg_multidot(const double* __restrict__ L,
                                            const double* __restrict__ R,
                                            const double* __restrict__ wgt, int n, double* red,
                                            double* out, int ldo) {
  double acc[NL * NR];
#pragma unroll
  for (int e = 0; e < NL * NR; ++e) acc[e] = 0.0;
#pragma unroll 2
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    double lv[NL], rv[NR];
    const double wr = wgt ? wgt[r] : 1.0;
#pragma unroll
    for (int i = 0; i < NL; ++i) lv[i] = L[(size_t)i * n + r] * wr;
#pragma unroll
    for (int j = 0; j < NR; ++j) rv[j] = R[(size_t)j * n + r];
#pragma unroll
    for (int i = 0; i < NL; ++i)
#pragma unroll
      for (int j = 0; j < NR; ++j) acc[i * NR + j] += lv[i] * rv[j];
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
#pragma unroll
  for (int e = 0; e < NL * NR; ++e) {
    double s = acc[e];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) red[wv * NL * NR + e] = s;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < NL * NR; e += blockDim.x) {
    double s = 0.0;
    for (int q = 0; q < nwv; ++q) s += red[q * NL * NR + e];
    out[(e / NR) * ldo + e % NR] = s;
  }
  __syncthreads();
}

// Per-column sums sum_r wgt[r] B_j[r]^2 (j < NB) into out[j].
template <int NB>
__device__ __forceinline__ void wg_colnorm2(const double* __restrict__ B,
                                            const double* __restrict__ wgt, int n, double* red,
                                            double* out) {
  double acc[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) acc[j] = 0.0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    const double wr = wgt[r];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const double x = B[(size_t)j * n + r];
      acc[j] += wr * x * x;
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    double s = acc[j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) red[wv * NB + j] = s;
  }
  __syncthreads();
  if (threadIdx.x < NB) {
    double s = 0.0;
    for (int q = 0; q < nwv; ++q) s += red[q * NB + threadIdx.x];
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

// Y = A_i^N V for the first nb of NB columns (interior faces only).
template <int NB>
__device__ inline void wg_apply_AN(const BrickView& v, const double* __restrict__ Tx,
                                   const double* __restrict__ Ty, const double* __restrict__ Tz,
                                   const double* __restrict__ wd, const double* __restrict__ V,
                                   double* __restrict__ Y, int nb) {
  const int n = v.n;
  for_cells(v, [&](int c, int lx, int ly, int lz) {
    const int li = c - v.off;
    const double tl0 = lx > 0 ? Tx[c] : 0.0, th0 = lx < v.nx - 1 ? Tx[c + 1] : 0.0;
    const double tl1 = ly > 0 ? Ty[c] : 0.0, th1 = ly < v.ny - 1 ? Ty[c + v.nx] : 0.0;
    const double tl2 = lz > 0 ? Tz[c] : 0.0, th2 = lz < v.nz - 1 ? Tz[c + v.nxny] : 0.0;
    const int o0 = lx > 0 ? li - 1 : li, o1 = lx < v.nx - 1 ? li + 1 : li;
    const int o2 = ly > 0 ? li - v.nx : li, o3 = ly < v.ny - 1 ? li + v.nx : li;
    const int o4 = lz > 0 ? li - v.nxny : li, o5 = lz < v.nz - 1 ? li + v.nxny : li;
    const double wc = wd ? wd[c] : 0.0;
    // Columns beyond nb read column 0 (in range) and store nothing, so the
    // loads of all NB columns are issued together.
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const double* x = V + (size_t)(j < nb ? j : 0) * n;
      const double xc = x[li];
      const double y = tl0 * (xc - x[o0]) + th0 * (xc - x[o1]) + tl1 * (xc - x[o2]) +
                       th1 * (xc - x[o3]) + tl2 * (xc - x[o4]) + th2 * (xc - x[o5]) + wc * xc;
      if (j < nb) Y[(size_t)j * n + li] = y;
    }
  });
  __syncthreads();
}

// The forward and backward FP32 block ICC(0) sweeps (hyperplane order) on ch
// columns held in LDS (lv, column-major, n per column).
__device__ inline void wg_sweeps(const Brick& b, const BrickView& v, const int* ps, const int* pc,
                                 const float4* __restrict__ pf, const float4* __restrict__ pb,
                                 float* lv, int ch) {
  const int n = v.n;
  for (int dirn = 0; dirn < 2; ++dirn) {
    const bool fwd = dirn == 0;
    const float4* fac = fwd ? pf : pb;
    for (int k = 0; k < b.nplanes; ++k) {
      const int s = fwd ? k : b.nplanes - 1 - k;
      for (int pos = ps[s] + threadIdx.x; pos < ps[s + 1]; pos += blockDim.x) {
        const int pk = pc[pos];
        const int lx = pk & 1023, ly = (pk >> 10) & 1023, lz = pk >> 20;
        const int li = (lz * v.ny + ly) * v.nx + lx;
        const float4 f = fac[v.off + pos];
        for (int q = 0; q < ch; ++q) {
          float* vv = lv + q * n;
          if (fwd) {
            float acc = vv[li];
            if (lx > 0) acc += f.y * vv[li - 1];
            if (ly > 0) acc += f.z * vv[li - v.nx];
            if (lz > 0) acc += f.w * vv[li - v.nxny];
            vv[li] = acc * f.x;
          } else {
            float acc = 0.0f;
            if (lx < v.nx - 1) acc += f.y * vv[li + 1];
            if (ly < v.ny - 1) acc += f.z * vv[li + v.nx];
            if (lz < v.nz - 1) acc += f.w * vv[li + v.nxny];
            vv[li] = vv[li] + acc * f.x;
          }
        }
      }
      __syncthreads();
    }
  }
}

// W <- T W with the FP32 block ICC(0) sweeps, kLobCh columns at a time.
__device__ inline void wg_precondition(const Brick& b, const BrickView& v, const int* ps,
                                       const int* pc, const float4* __restrict__ pf,
                                       const float4* __restrict__ pb, double* W, int nc,
                                       float* lv) {
  const int n = v.n;
  for (int j0 = 0; j0 < nc; j0 += kLobCh) {
    const int ch = min(kLobCh, nc - j0);
    for (int e = threadIdx.x; e < ch * n; e += blockDim.x) lv[e] = (float)W[(size_t)j0 * n + e];
    __syncthreads();
    wg_sweeps(b, v, ps, pc, pf, pb, lv, ch);
    for (int e = threadIdx.x; e < ch * n; e += blockDim.x) W[(size_t)j0 * n + e] = (double)lv[e];
    __syncthreads();
  }
}

// Row-wise in-place transform of an NB-column block: B(:, 0:nout) <- B(:, 0:nin) C.
template <int NB>
__device__ inline void wg_transform(double* B, int n, const double* C, int ldc, int nin,
                                    int nout) {
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    double x[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) x[k] = k < nin ? B[(size_t)k * n + r] : 0.0;
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      if (c >= nout) break;
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < NB; ++k)
        if (k < nin) s += x[k] * C[k * ldc + c];
      B[(size_t)c * n + r] = s;
    }
  }
  __syncthreads();
}


// SVQB of an nb x nb M-Gram G (destroyed): Z = D U_kept theta^{-1/2},
// D = diag(G)^{-1/2}; returns the number of kept columns.  G1 and Q1 are
// scratch.
__device__ inline int wg_svqb(double* G, double* G1, double* Z, double* Q1, double* w, double* rot,
                              int nb, int ld, double drop) {
  __shared__ double dsc[64];
  __shared__ int s_keep, s_first;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    const double d = G[i * ld + i];
    dsc[i] = d > 0.0 ? 1.0 / sqrt(d) : 0.0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
    const int i = e / nb, j = e % nb;
    G[i * ld + j] *= dsc[i] * dsc[j];
  }
  __syncthreads();
  wg_syev2(G, G1, Z, Q1, w, rot, nb, ld);
  if (threadIdx.x == 0) {
    const double tmax = w[nb - 1];
    int first = nb;
    for (int i = 0; i < nb; ++i)
      if (w[i] > drop * tmax) {
        first = i;
        break;
      }
    s_first = first;
    s_keep = nb - first;
  }
  __syncthreads();
  const int keep = s_keep, first = s_first;
  for (int e = threadIdx.x; e < nb * keep; e += blockDim.x) {
    const int i = e / keep, c = e % keep;
    G1[i * ld + c] = dsc[i] * Z[i * ld + first + c] / sqrt(w[first + c]);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nb * keep; e += blockDim.x) {
    const int i = e / keep, c = e % keep;
    Z[i * ld + c] = G1[i * ld + c];
  }
  __syncthreads();
  return keep;
}

// Launches of the matrix-core kernel (pmx_lobpcg_mf.hip) for block size MB in
// {4, 6, 8, 10}: one workgroup per subdomain (nblocks of them), threads per
// workgroup chosen from the largest subdomain (maxn cells) as for the VALU
// kernel; shape 1 runs large subdomains (>= kLobBigN cells) two per CU
// (256 threads, narrower sweep chunks) instead of one per CU.
void launch_lobpcg_mf(int MB, int shape, const Geo& g, int nblocks, hipStream_t st,
                      const LobArgs& a, int maxn);

}  // namespace pmx
