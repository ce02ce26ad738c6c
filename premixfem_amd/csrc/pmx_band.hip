// Exact level-2 coarse factorisation for large coarse spaces: the Cholesky
// factor of A_c in band storage.  A_c (natural subdomain order, eigen index
// inner; src/PreMixFEM_3D.c:40-45) is block 7-point, so its Cholesky factor
// fills in only inside the band of half-width sdx*sdy*k (the +z neighbour).
// The reference factors A_c with MUMPS (:894-906); a dense factor at
// nc = 32768 would be 8.6 GB and 12 TFLOP, the band one is 0.57 GB and
// 0.15 TFLOP.
//
// Storage: 64x64 tiles; block column J keeps the tiles of block rows
// J..J+w (tile (J+d, J) at band[(J*(w+1) + d) * 4096], row-major), plus the
// inverse of every diagonal factor tile.  Right-looking blocked Cholesky:
// per block column one POTRF (+ tile inverse), w TRSMs, w(w+1)/2 tile
// updates.  Solves: forward and backward tile-column sweeps, one launch per
// tile column spread over the w tile rows it updates.
#include <cmath>
#include <vector>

#include "pmx_dense.h"
#include "pmx_kernels.h"

namespace pmx {

constexpr int kBt = 64;              // tile edge
constexpr size_t kBtt = kBt * kBt;   // tile size

__device__ __forceinline__ double* band_tile(double* band, int w, int J, int d) {
  return band + ((size_t)J * (w + 1) + d) * kBtt;
}

// Scatter the lower triangle of A_c (block rows Acb) into the band tiles.
__global__ void k_band_fill(Geo g, const double* __restrict__ Acb, int kmax, int w,
                            double* __restrict__ band) {
  const int s = blockIdx.x;
  const Brick b = g.bricks[s];
  const size_t kk = (size_t)kmax * kmax;
  for (int e = threadIdx.x; e < 7 * kmax * kmax; e += blockDim.x) {
    const int slot = e / (kmax * kmax), i = (e / kmax) % kmax, j = e % kmax;
    const int o = slot == 0 ? s : b.nb[slot - 1];
    if (o < 0) continue;
    const Brick ob = g.bricks[o];
    if (i >= b.neig || j >= ob.neig) continue;
    const int r = b.dof + i, c = ob.dof + j;
    if (c > r) continue;
    const int I = r / kBt, J = c / kBt;
    band_tile(band, w, J, I - J)[(r % kBt) * kBt + c % kBt] = Acb[(size_t)s * 7 * kk + slot * kk + i * kmax + j];
  }
}

// Diagonal tile J: Cholesky in place (lower) and its inverse L^{-1} into
// dinv[J].  The last tile may be partial (nr rows): the padding is identity.
__global__ void __launch_bounds__(256) k_band_potrf(double* __restrict__ band, int w, int J, int nr,
                                                    double* __restrict__ dinv,
                                                    int* __restrict__ fail) {
  __shared__ double A[kBt * kBt], W[kBt * kBt];
  int* fl = reinterpret_cast<int*>(&W[kBtt - 1]);  // W is free until the inverse
  double* T = band_tile(band, w, J, 0);
  for (int e = threadIdx.x; e < (int)kBtt; e += blockDim.x) {
    const int i = e / kBt, j = e % kBt;
    A[e] = (i < nr && j < nr) ? (j <= i ? T[e] : T[j * kBt + i]) : (i == j ? 1.0 : 0.0);
  }
  __syncthreads();
  const bool ok = wg_cholesky(A, kBt, kBt, fl);
  if (!ok && threadIdx.x == 0) atomicMax(fail, J + 1);
  for (int e = threadIdx.x; e < (int)kBtt; e += blockDim.x) T[e] = A[e];
  wg_trinv_lower(A, W, kBt, kBt);
  double* D = dinv + (size_t)J * kBtt;
  for (int e = threadIdx.x; e < (int)kBtt; e += blockDim.x) D[e] = A[e];
}

// C (64x64) = A * op(B) with op(B) = B^T (transB) or B, accumulated into C
// with sign sg (C = sg * A op(B), or C += sg * A op(B) when acc).
__device__ inline void tile_gemm(const double* A, const double* B, double* C, bool transB,
                                 double sg, bool acc,
                                 double (*As)[kBt + 1], double (*Bs)[kBt + 1]) {
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  double r[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) r[a][c] = 0.0;
  for (int k0 = 0; k0 < kBt; k0 += 32) {
    for (int e = threadIdx.x; e < kBt * 32; e += blockDim.x) {
      const int i = e / 32, k = e % 32;
      As[k][i] = A[i * kBt + k0 + k];
      Bs[k][i] = transB ? B[i * kBt + k0 + k] : B[(k0 + k) * kBt + i];
    }
    __syncthreads();
#pragma unroll 8
    for (int k = 0; k < 32; ++k) {
      double av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = As[k][ty + 16 * a];
#pragma unroll
      for (int c = 0; c < 4; ++c) bv[c] = Bs[k][tx + 16 * c];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) r[a][c] += av[a] * bv[c];
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int i = ty + 16 * a, j = tx + 16 * c;
      C[i * kBt + j] = acc ? C[i * kBt + j] + sg * r[a][c] : sg * r[a][c];
    }
}

// TRSM: L(J+d, J) = A(J+d, J) L_JJ^{-T} for d = 1 + blockIdx.x.
// (tile_gemm reads all of A before it writes C, so C may alias A.)
__global__ void __launch_bounds__(256) k_band_trsm(double* __restrict__ band, int w, int J,
                                                   const double* __restrict__ dinv) {
  __shared__ double As[32][kBt + 1], Bs[32][kBt + 1];
  const int d = 1 + blockIdx.x;
  double* T = band_tile(band, w, J, d);
  tile_gemm(T, dinv + (size_t)J * kBtt, T, true, 1.0, false, As, Bs);
}

// Update: A(J+d2, J+d1) -= L(J+d2, J) L(J+d1, J)^T for 1 <= d1 <= d2 <= dw
// (pair index blockIdx.x in row-major lower-triangle order).
__global__ void __launch_bounds__(256) k_band_update(double* __restrict__ band, int w, int J) {
  __shared__ double As[32][kBt + 1], Bs[32][kBt + 1];
  int p = blockIdx.x, d2 = 1;
  while (p >= d2) {
    p -= d2;
    ++d2;
  }
  const int d1 = 1 + p;
  tile_gemm(band_tile(band, w, J, d2), band_tile(band, w, J, d1),
            band_tile(band, w, J + d1, d2 - d1), true, -1.0, true, As, Bs);
}

// Solves A_c d = t by tile columns, one launch per tile column and sweep,
// with (w, ncol) workgroups: workgroup (dd, c) recomputes the pivot tile of
// right-hand side c redundantly and applies it to one tile row.
//   forward  (L x = t):   x_J = L_JJ^{-1} y_J;  y_{J+dd} -= L(J+dd, J) x_J
//   backward (L^T z = x): z_K = L_KK^{-T} x_K;  x_{K-dd} -= L(K, K-dd)^T z_K
// y / x live in yv (updated in place), the finished tiles go to out; the two
// arrays never alias, so no workgroup reads what another writes in a launch.
__global__ void __launch_bounds__(256) k_band_step(const double* __restrict__ band, int w, int nt,
                                                   const double* __restrict__ dinv, int J,
                                                   int fwd, double* __restrict__ yv,
                                                   double* __restrict__ out, int nc) {
  __shared__ double xs[kBt];
  __shared__ double part[4][kBt];
  const int dd = 1 + blockIdx.x;
  double* y = yv + (size_t)blockIdx.y * nc;
  double* o = out + (size_t)blockIdx.y * nc;
  const int row = threadIdx.x % kBt, grp = threadIdx.x / kBt;  // 4 groups
  const double* D = dinv + (size_t)J * kBtt;
  {
    double acc = 0.0;
    for (int k = grp; k < kBt; k += 4) {
      const int r = J * kBt + k;
      const double yk = r < nc ? y[r] : 0.0;
      acc += (fwd ? D[row * kBt + k] : D[k * kBt + row]) * yk;
    }
    part[grp][row] = acc;
  }
  __syncthreads();
  if (threadIdx.x < kBt) {
    const double v = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] +
                     part[3][threadIdx.x];
    xs[threadIdx.x] = v;
    const int r = J * kBt + threadIdx.x;
    if (dd == 1 && r < nc) o[r] = v;
  }
  __syncthreads();
  const int T = fwd ? J + dd : J - dd;
  if (fwd ? T >= nt : T < 0) return;
  // fwd: tile (J+dd, J) rows i; bwd: tile (J, J-dd) transposed, rows = its columns.
  const double* L = fwd ? band + ((size_t)J * (w + 1) + dd) * kBtt
                        : band + ((size_t)T * (w + 1) + dd) * kBtt;
  double acc = 0.0;
  for (int k = grp; k < kBt; k += 4) acc += (fwd ? L[row * kBt + k] : L[k * kBt + row]) * xs[k];
  part[grp][row] = acc;
  __syncthreads();
  if (threadIdx.x < kBt) {
    const int r = T * kBt + threadIdx.x;
    if (r < nc)
      y[r] -= part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] +
              part[3][threadIdx.x];
  }
}

int band_setup(pmx_ctx* ctx) {
  const int nc = ctx->nc;
  // Half-bandwidth from the block structure: the farthest column of any row.
  int bw = 0;
  for (int s = 0; s < ctx->nsub; ++s) {
    const Brick& b = ctx->h_bricks[s];
    for (int dir = 0; dir < 6; ++dir) {
      const int o = b.nb[dir];
      if (o < 0) continue;
      const Brick& ob = ctx->h_bricks[o];
      bw = std::max(bw, std::abs((ob.dof + ob.neig - 1) - b.dof));
      bw = std::max(bw, std::abs((b.dof + b.neig - 1) - ob.dof));
    }
    bw = std::max(bw, b.neig - 1);
  }
  const int nt = (nc + kBt - 1) / kBt;
  const int w = std::min(nt - 1, bw / kBt + 1);
  ctx->band_nt = nt;
  ctx->band_w = w;
  PMX_TRY(ctx->band.alloc((size_t)nt * (w + 1) * kBtt));
  PMX_TRY(ctx->band_dinv.alloc((size_t)nt * kBtt));
  PMX_TRY(ctx->band_tmp.alloc((size_t)nc * 32));  // solve work, so a solve never allocates
  PMX_HIP(hipMemsetAsync(ctx->band.p, 0, ctx->band.n * sizeof(double), ctx->stream));
  hipLaunchKernelGGL(k_band_fill, dim3(ctx->nsub), dim3(256), 0, ctx->stream, ctx->geo, ctx->Acb.p,
                     ctx->en_lv1, w, ctx->band.p);
  DBuf<int> fail;
  PMX_TRY(fail.alloc(1));
  PMX_HIP(hipMemsetAsync(fail.p, 0, sizeof(int), ctx->stream));
  for (int J = 0; J < nt; ++J) {
    const int nr = std::min(kBt, nc - J * kBt);
    hipLaunchKernelGGL(k_band_potrf, dim3(1), dim3(256), 0, ctx->stream, ctx->band.p, w, J, nr,
                       ctx->band_dinv.p, fail.p);
    const int dw = std::min(w, nt - 1 - J);
    if (dw > 0) {
      hipLaunchKernelGGL(k_band_trsm, dim3(dw), dim3(256), 0, ctx->stream, ctx->band.p, w, J,
                         ctx->band_dinv.p);
      hipLaunchKernelGGL(k_band_update, dim3(dw * (dw + 1) / 2), dim3(256), 0, ctx->stream,
                         ctx->band.p, w, J);
    }
  }
  PMX_HIP(hipGetLastError());
  int hf = 0;
  PMX_HIP(hipMemcpyAsync(&hf, fail.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  PMX_HIP(hipStreamSynchronize(ctx->stream));
  if (hf) {
    set_error("band Cholesky of the coarse matrix: non-positive pivot in tile %d", hf - 1);
    return PMX_ERR_NOT_CONVERGED;
  }
  return PMX_SUCCESS;
}

int launch_band_solve(pmx_ctx* ctx, const double* t, double* d, int ncol) {
  const int nt = ctx->band_nt, w = ctx->band_w, nc = ctx->nc;
  PMX_TRY(ctx->band_tmp.alloc((size_t)nc * std::max(ncol, 32)));
  double* y = ctx->band_tmp.p;
  PMX_HIP(hipMemcpyAsync(y, t, (size_t)nc * ncol * sizeof(double), hipMemcpyDeviceToDevice,
                         ctx->stream));
  for (int J = 0; J < nt; ++J) {
    const int dw = std::max(1, std::min(w, nt - 1 - J));
    hipLaunchKernelGGL(k_band_step, dim3(dw, ncol), dim3(256), 0, ctx->stream, ctx->band.p, w, nt,
                       ctx->band_dinv.p, J, 1, y, d, nc);
  }
  // d now holds x: the backward sweep updates it in place and writes z to y.
  for (int K = nt - 1; K >= 0; --K) {
    const int dw = std::max(1, std::min(w, K));
    hipLaunchKernelGGL(k_band_step, dim3(dw, ncol), dim3(256), 0, ctx->stream, ctx->band.p, w, nt,
                       ctx->band_dinv.p, K, 0, d, y, nc);
  }
  PMX_HIP(hipMemcpyAsync(d, y, (size_t)nc * ncol * sizeof(double), hipMemcpyDeviceToDevice,
                         ctx->stream));
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

}  // namespace pmx
