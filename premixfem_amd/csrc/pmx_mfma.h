// FP64 matrix-core (MFMA) tiles for the dense coarse-level algebra.
//
// v_mfma_f64_16x16x4f64 (one wavefront): D(16x16) += A(16x4) B(4x16) with
//   A: lane l holds A[l % 16][l / 16],
//   B: lane l holds B[l / 16][l % 16],
//   D: lane l, item v holds D[l / 16 + 4 v][l % 16].
// A workgroup of 4 waves computes a (WR*RB*16) x (WC*CB*16) tile of
// C = alpha op(A) op(B) + beta C: the waves form a WR x WC grid, each wave
// owns RB x CB blocks of 16 x 16.  K is staged through LDS in chunks of
// kMk, zero-filled outside the matrices, so any M, N, K work.
#pragma once

#include <hip/hip_runtime.h>

namespace pmx {

typedef double mfma_d4 __attribute__((ext_vector_type(4)));

constexpr int kMk = 16;   // K chunk staged per LDS round

template <int WR, int WC, int RB, int CB>
struct MfmaTile {
  static constexpr int TM = WR * RB * 16;
  static constexpr int TN = WC * CB * 16;
  static constexpr int LDA = kMk + 1;   // As[row][k]
  static constexpr int LDB = TN + 1;    // Bs[k][col]
  static constexpr int LDS_DOUBLES = TM * LDA + kMk * LDB;
};

// acc += op(A)[I0:I0+TM, :K] op(B)[:K, J0:J0+TN] for this wave's blocks.
// A is row-major (A[i*lda + k]).  TB = false: B row-major K x N
// (B[k*ldb + j]); TB = true: B given as its transpose (B[j*ldb + k]).
// lds: LDS_DOUBLES doubles.  All 256 threads must call it.  The next K
// chunk is loaded into registers while the current one is multiplied.
template <int WR, int WC, int RB, int CB, bool TB>
__device__ __forceinline__ void mfma_tile_acc(const double* __restrict__ A, int lda,
                                              const double* __restrict__ B, int ldb, int M, int N,
                                              int K, int I0, int J0, double* lds,
                                              mfma_d4 (&acc)[RB][CB]) {
  using T = MfmaTile<WR, WC, RB, CB>;
  constexpr int NA = T::TM * kMk / 256, NB = T::TN * kMk / 256;
  static_assert(NA * 256 == T::TM * kMk && NB * 256 == T::TN * kMk, "tile / K chunk");
  double* As = lds;
  double* Bs = lds + T::TM * T::LDA;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave / WC, wc = wave % WC;
  const int li = lane & 15, lk = lane >> 4;
  double ra[NA], rb[NB];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < NA; ++q) {
      const int e = tid + 256 * q;
      const int r = e / kMk, k = e % kMk;
      const int gi = I0 + r, gk = k0 + k;
      ra[q] = (gi < M && gk < K) ? A[(size_t)gi * lda + gk] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int e = tid + 256 * q;
      if (TB) {
        const int j = e / kMk, k = e % kMk;
        const int gj = J0 + j, gk = k0 + k;
        rb[q] = (gj < N && gk < K) ? B[(size_t)gj * ldb + gk] : 0.0;
      } else {
        const int k = e / T::TN, j = e % T::TN;
        const int gj = J0 + j, gk = k0 + k;
        rb[q] = (gj < N && gk < K) ? B[(size_t)gk * ldb + gj] : 0.0;
      }
    }
  };
  load(0);
  for (int k0 = 0; k0 < K; k0 += kMk) {
#pragma unroll
    for (int q = 0; q < NA; ++q) {
      const int e = tid + 256 * q;
      As[(e / kMk) * T::LDA + e % kMk] = ra[q];
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int e = tid + 256 * q;
      if (TB)
        Bs[(e % kMk) * T::LDB + e / kMk] = rb[q];
      else
        Bs[(e / T::TN) * T::LDB + e % T::TN] = rb[q];
    }
    __syncthreads();
    if (k0 + kMk < K) load(k0 + kMk);
#pragma unroll
    for (int k4 = 0; k4 < kMk; k4 += 4) {
      double af[RB], bf[CB];
#pragma unroll
      for (int r = 0; r < RB; ++r) af[r] = As[((wr * RB + r) * 16 + li) * T::LDA + k4 + lk];
#pragma unroll
      for (int c = 0; c < CB; ++c) bf[c] = Bs[(k4 + lk) * T::LDB + (wc * CB + c) * 16 + li];
#pragma unroll
      for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int c = 0; c < CB; ++c)
          acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[r], bf[c], acc[r][c], 0, 0, 0);
    }
    __syncthreads();
  }
}

// Row and column of item v of block (rb, cb) held by this lane, relative to
// the tile origin.
template <int WR, int WC, int RB, int CB>
__device__ __forceinline__ void mfma_tile_pos(int rb, int cb, int v, int& r, int& c) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave / WC, wc = wave % WC;
  r = (wr * RB + rb) * 16 + (lane >> 4) + 4 * v;
  c = (wc * CB + cb) * 16 + (lane & 15);
}

// C = alpha op(A) op(B) + beta C (beta == 0: C is not read).  Grid:
// (ceil(N / TN), ceil(M / TM)), 256 threads.
template <int WR, int WC, int RB, int CB, bool TB>
__global__ void __launch_bounds__(256) k_gemm_mfma(int M, int N, int K, double alpha,
                                                   const double* __restrict__ A, int lda,
                                                   const double* __restrict__ B, int ldb,
                                                   double beta, double* __restrict__ C, int ldc) {
  using T = MfmaTile<WR, WC, RB, CB>;
  __shared__ double lds[T::LDS_DOUBLES];
  const int I0 = blockIdx.y * T::TM, J0 = blockIdx.x * T::TN;
  mfma_d4 acc[RB][CB];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) acc[rb][cb] = mfma_d4{0.0, 0.0, 0.0, 0.0};
  mfma_tile_acc<WR, WC, RB, CB, TB>(A, lda, B, ldb, M, N, K, I0, J0, lds, acc);
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int r, c;
        mfma_tile_pos<WR, WC, RB, CB>(rb, cb, v, r, c);
        const int i = I0 + r, j = J0 + c;
        if (i < M && j < N) {
          double* dst = C + (size_t)i * ldc + j;
          *dst = beta == 0.0 ? alpha * acc[rb][cb][v] : beta * *dst + alpha * acc[rb][cb][v];
        }
      }
}

}  // namespace pmx

namespace pmx {

// Split-K partial products: workgroup (x, y, z) computes the (y, x) tile of
// op(A)[:, Kz] op(B)[Kz, :] over the z-th K range of Kc (a multiple of kMk)
// into W[z][M][N].  k_gemm_sum adds the partials in fixed order, so the
// result does not depend on scheduling.
template <int WR, int WC, int RB, int CB, bool TB>
__global__ void __launch_bounds__(256) k_gemm_mfma_part(int M, int N, int K, int Kc,
                                                        const double* __restrict__ A, int lda,
                                                        const double* __restrict__ B, int ldb,
                                                        double* __restrict__ W) {
  using T = MfmaTile<WR, WC, RB, CB>;
  __shared__ double lds[T::LDS_DOUBLES];
  const int I0 = blockIdx.y * T::TM, J0 = blockIdx.x * T::TN;
  const int kz = blockIdx.z * Kc;
  const int kl = min(Kc, K - kz);
  mfma_d4 acc[RB][CB];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) acc[rb][cb] = mfma_d4{0.0, 0.0, 0.0, 0.0};
  // The K range is selected by offsetting A's columns and B's K index.
  mfma_tile_acc<WR, WC, RB, CB, TB>(A + kz, lda, TB ? B + kz : B + (size_t)kz * ldb, ldb, M, N,
                                    kl, I0, J0, lds, acc);
  double* Wz = W + (size_t)blockIdx.z * M * N;
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int r, c;
        mfma_tile_pos<WR, WC, RB, CB>(rb, cb, v, r, c);
        const int i = I0 + r, j = J0 + c;
        if (i < M && j < N) Wz[(size_t)i * N + j] = acc[rb][cb][v];
      }
}

// C = alpha sum_z W[z] + beta C (beta == 0: C is not read).
__global__ void k_gemm_sum(const double* __restrict__ W, int S, int M, int N, double alpha,
                           double beta, double* __restrict__ C, int ldc) {
  const size_t mn = (size_t)M * N;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < mn;
       e += (size_t)gridDim.x * blockDim.x) {
    double v = 0.0;
    for (int z = 0; z < S; ++z) v += W[z * mn + e];
    const size_t i = e / N, j = e % N;
    double* dst = C + i * ldc + j;
    *dst = beta == 0.0 ? alpha * v : beta * *dst + alpha * v;
  }
}

}  // namespace pmx
