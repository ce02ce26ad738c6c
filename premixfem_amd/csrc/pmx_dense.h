// Small dense FP64 routines executed cooperatively by one workgroup on
// matrices held in LDS (row-major, leading dimension ld).  Used by the
// Rayleigh-Ritz steps of the level-2 and level-3 eigensolvers.
#pragma once

#include "pmx_internal.h"

namespace pmx {

// Cholesky A = L L^T in place (lower triangle), m <= 64.  Returns false (to
// every thread) if a pivot is not positive.
__device__ inline bool wg_cholesky(double* A, int m, int ld, int* flag) {
  if (threadIdx.x == 0) *flag = 0;
  __syncthreads();
  for (int j = 0; j < m; ++j) {
    if (threadIdx.x == 0) {
      const double d = A[j * ld + j];
      if (!(d > 0.0)) *flag = 1;
      A[j * ld + j] = d > 0.0 ? sqrt(d) : 1.0;
    }
    __syncthreads();
    const double ljj = A[j * ld + j];
    for (int i = j + 1 + threadIdx.x; i < m; i += blockDim.x) A[i * ld + j] /= ljj;
    __syncthreads();
    for (int e = threadIdx.x; e < (m - j - 1) * (m - j - 1); e += blockDim.x) {
      const int i = j + 1 + e / (m - j - 1), k = j + 1 + e % (m - j - 1);
      if (k <= i) A[i * ld + k] -= A[i * ld + j] * A[k * ld + j];
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
    const int i = e / m, k = e % m;
    if (k > i) A[i * ld + k] = 0.0;
  }
  __syncthreads();
  return *flag == 0;
}

// Inverse of a lower-triangular L in place (m <= 64), column by column.
__device__ inline void wg_trinv_lower(double* L, double* W, int m, int ld) {
  // W = L^{-1}, solved column by column by forward substitution, one thread
  // per column.
  for (int c = threadIdx.x; c < m; c += blockDim.x) {
    for (int i = 0; i < m; ++i) {
      double s = (i == c) ? 1.0 : 0.0;
      for (int k = c; k < i; ++k) s -= L[i * ld + k] * W[k * ld + c];
      W[i * ld + c] = (i < c) ? 0.0 : s / L[i * ld + i];
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < m * m; e += blockDim.x) L[(e / m) * ld + e % m] = W[(e / m) * ld + e % m];
  __syncthreads();
}

// C = A * B for m x m matrices in LDS (C distinct from A and B).
__device__ inline void wg_matmul(const double* A, const double* B, double* C, int m, int ld,
                                 bool transA) {
  for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
    const int i = e / m, j = e % m;
    double s = 0.0;
    for (int k = 0; k < m; ++k) s += (transA ? A[k * ld + i] : A[i * ld + k]) * B[k * ld + j];
    C[i * ld + j] = s;
  }
  __syncthreads();
}

// Symmetric eigendecomposition A = Q diag(w) Q^T by parallel cyclic Jacobi
// (round-robin ordering, m/2 disjoint rotations per round), eigenvalues in
// ascending order.  A (destroyed), Q and tmp are m x m in LDS; w has m
// entries; rot is scratch of 4*m doubles.  m <= 64.
__device__ inline void wg_syev(double* A, double* Q, double* tmp, double* w, double* rot, int m,
                               int ld, int max_sweeps = 40) {
  for (int e = threadIdx.x; e < m * m; e += blockDim.x)
    Q[(e / m) * ld + e % m] = (e / m == e % m) ? 1.0 : 0.0;
  __shared__ int s_rot;
  __syncthreads();
  const int mp = (m + 1) & ~1;  // even number of players; index m is a dummy
  double* cc = rot;             // per index: c
  double* ss = rot + m;         // per index: signed s
  int* part = reinterpret_cast<int*>(rot + 2 * m);  // per index: partner
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    if (threadIdx.x == 0) s_rot = 0;
    __syncthreads();
    for (int r = 0; r < mp - 1; ++r) {
      for (int k = threadIdx.x; k < mp / 2; k += blockDim.x) {
        int p, q;
        if (k == 0) {
          p = r;
          q = mp - 1;
        } else {
          p = (r + k) % (mp - 1);
          q = (r - k + (mp - 1)) % (mp - 1);
        }
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        if (q < m) {
          const double app = A[p * ld + p], aqq = A[q * ld + q], apq = A[p * ld + q];
          double c = 1.0, s = 0.0;
          // Rotate only when the coupling is significant relative to the
          // diagonal (threshold Jacobi: high relative accuracy of small
          // eigenvalues); a sweep without rotations ends the iteration.
          if (fabs(apq) > 1e-300 && fabs(apq) > 1e-17 * sqrt(fabs(app * aqq))) {
            s_rot = 1;
            const double th = (aqq - app) / (2.0 * apq);
            const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
            c = 1.0 / sqrt(t * t + 1.0);
            s = t * c;
          }
          cc[p] = c;
          cc[q] = c;
          ss[p] = -s;
          ss[q] = s;
          part[p] = q;
          part[q] = p;
        } else if (p < m) {
          cc[p] = 1.0;
          ss[p] = 0.0;
          part[p] = p;
        }
      }
      __syncthreads();
      for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
        const int i = e / m, j = e % m;
        const int ip = part[i], jp = part[j];
        const double ci = cc[i], si = ss[i], cj = cc[j], sj = ss[j];
        double v;
        if (ip == j && ip != i) {
          v = 0.0;  // the annihilated pair
        } else {
          v = ci * cj * A[i * ld + j] + ci * sj * A[i * ld + jp] + si * cj * A[ip * ld + j] +
              si * sj * A[ip * ld + jp];
        }
        tmp[i * ld + j] = v;
      }
      __syncthreads();
      for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
        const int i = e / m, j = e % m;
        A[i * ld + j] = tmp[i * ld + j];
        tmp[i * ld + j] = cc[j] * Q[i * ld + j] + ss[j] * Q[i * ld + part[j]];
      }
      __syncthreads();
      for (int e = threadIdx.x; e < m * m; e += blockDim.x) Q[(e / m) * ld + e % m] = tmp[(e / m) * ld + e % m];
      __syncthreads();
    }
    if (s_rot == 0) break;
    __syncthreads();
  }
  // Sort ascending (selection on thread 0), then permute Q's columns.
  int* perm = reinterpret_cast<int*>(rot);
  if (threadIdx.x == 0) {
    for (int i = 0; i < m; ++i) {
      w[i] = A[i * ld + i];
      perm[i] = i;
    }
    for (int i = 0; i < m; ++i) {
      int best = i;
      for (int j = i + 1; j < m; ++j)
        if (w[j] < w[best]) best = j;
      const double tw = w[i];
      w[i] = w[best];
      w[best] = tw;
      const int tp = perm[i];
      perm[i] = perm[best];
      perm[best] = tp;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
    const int i = e / m, j = e % m;
    tmp[i * ld + j] = Q[i * ld + perm[j]];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < m * m; e += blockDim.x) Q[(e / m) * ld + e % m] = tmp[(e / m) * ld + e % m];
  __syncthreads();
}

}  // namespace pmx
