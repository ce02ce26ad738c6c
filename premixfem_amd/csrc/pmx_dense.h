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

// 1/x and 1/sqrt(x) from the hardware estimates and two Newton steps.
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  return fma(r, fma(-x, r, 1.0), r);
}
__device__ __forceinline__ double rsq_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = y * fma(-0.5 * x * y, y, 1.5);
  return y * fma(-0.5 * x * y, y, 1.5);
}

// Pair k (0 <= k < mp / 2) of round r (0 <= r < mp - 1) of the round-robin
// ordering over mp (even) indices: (r, mp - 1) for k = 0, else
// (r + k, r - k) mod (mp - 1); returned with p < q.  Over the mp - 1 rounds
// every index pair meets once.  q >= m marks an index without a partner.
__device__ __forceinline__ void jacobi_pair(int r, int k, int mp, int& p, int& q) {
  if (k == 0) {
    p = r;
    q = mp - 1;
  } else {
    p = r + k;
    if (p >= mp - 1) p -= mp - 1;
    q = r - k;
    if (q < 0) q += mp - 1;
  }
  if (p > q) {
    const int t = p;
    p = q;
    q = t;
  }
}

// Symmetric eigendecomposition by parallel cyclic Jacobi with two barriers
// per round (ping-pong buffers).  Input in A0 (destroyed); on return the
// eigenvectors are in Q0 (columns, ascending eigenvalues w).  A1, Q1 are
// scratch; rot holds 3*ld doubles.  Rotations are applied only where the
// coupling is significant relative to the diagonal (threshold Jacobi, which
// keeps small eigenvalues relatively accurate); a sweep without rotations
// ends the iteration.  A round computes the rotations of its mp / 2
// disjoint pairs once, then applies them as A <- R^T A R, Q <- Q R in
// units of a 2 x 2 block (pair a rows, pair b columns) of A and of Q, so a
// unit reads its operands in one batch.
__device__ inline int wg_syev2(double* A0, double* A1, double* Q0, double* Q1, double* w,
                               double* rot, int m, int ld, int max_sweeps = 30) {
  __shared__ int s_rot[2];
  double* pcs = rot;       // pair k: cosine
  double* psn = rot + ld;  // pair k: sine (the second index's; the first takes -s)
  for (int e = threadIdx.x; e < m * m; e += blockDim.x)
    Q0[(e / m) * ld + e % m] = (e / m == e % m) ? 1.0 : 0.0;
  if (threadIdx.x == 0) s_rot[0] = s_rot[1] = 0;
  __syncthreads();
  const int mp = (m + 1) & ~1, np2 = mp >> 1;
  double *Ac = A0, *An = A1, *Qc = Q0, *Qn = Q1;
  int sweeps = 0;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    const int fl = sweep & 1;
    sweeps = sweep + 1;
    for (int r = 0; r < mp - 1; ++r) {
      for (int k = threadIdx.x; k < np2; k += blockDim.x) {
        int p, q;
        jacobi_pair(r, k, mp, p, q);
        double c = 1.0, s = 0.0;
        if (q < m) {
          const double app = Ac[p * ld + p], aqq = Ac[q * ld + q], apq = Ac[p * ld + q];
          // the threshold test and the rotation are computed side by side
          // (no branch between them); apq = 0 takes the identity.  The
          // reciprocals and 1/sqrt are the hardware estimates refined by two
          // Newton steps (c and s to a few ulps; A and Q take the same ones).
          const bool big = fabs(apq) > 1e-300 && fabs(apq) > 1e-17 * sqrt(fabs(app * aqq));
          const double th = (aqq - app) * rcp_nr(2.0 * (apq != 0.0 ? apq : 1.0));
          const double t = (th >= 0.0 ? 1.0 : -1.0) * rcp_nr(fabs(th) + sqrt(th * th + 1.0));
          const double cr = rsq_nr(t * t + 1.0);
          c = big ? cr : 1.0;
          s = big ? t * cr : 0.0;
          if (big) s_rot[fl] = 1;
        }
        pcs[k] = c;
        psn[k] = s;
      }
      __syncthreads();
      // every reader of the previous sweep's flag is past this point
      if (r == 0 && threadIdx.x == 0) s_rot[fl ^ 1] = 0;
      for (int u = threadIdx.x; u < np2 * np2; u += blockDim.x) {
        const int ka = u / np2, kb = u - ka * np2;
        int pa, qa, pb, qb;
        jacobi_pair(r, ka, mp, pa, qa);
        jacobi_pair(r, kb, mp, pb, qb);
        const bool ha = qa < m, hb = qb < m;
        // element (i, j): ci cj a_ij + ci sj a_i,j' + si cj a_i',j + si sj a_i',j'
        // with (c, -s) at the first index of a pair and (c, s) at the second,
        // i' and j' the partners; an index without a partner has (1, 0) and
        // is its own partner
        const int qa1 = ha ? qa : pa, qb1 = hb ? qb : pb;
        const double ca = pcs[ka], sa = psn[ka], cb = pcs[kb], sb = psn[kb];
        const double a00 = Ac[pa * ld + pb], a01 = Ac[pa * ld + qb1];
        const double a10 = Ac[qa1 * ld + pb], a11 = Ac[qa1 * ld + qb1];
        const double q00 = Qc[pa * ld + pb], q01 = Qc[pa * ld + qb1];
        const double q10 = Qc[qa1 * ld + pb], q11 = Qc[qa1 * ld + qb1];
        const double sa0 = ha ? -sa : 0.0, sb0 = hb ? -sb : 0.0;
        auto elem = [](double ci, double si, double cj, double sj, double aij, double aijp,
                       double aipj, double aipjp) {
          return ci * cj * aij + ci * sj * aijp + si * cj * aipj + si * sj * aipjp;
        };
        const bool diag = ka == kb;  // the annihilated couplings
        An[pa * ld + pb] = elem(ca, sa0, cb, sb0, a00, a01, a10, a11);
        Qn[pa * ld + pb] = cb * q00 + sb0 * q01;
        if (hb) {
          An[pa * ld + qb] = (diag && ha) ? 0.0 : elem(ca, sa0, cb, sb, a01, a00, a11, a10);
          Qn[pa * ld + qb] = cb * q01 + sb * q00;
        }
        if (ha) {
          An[qa * ld + pb] = diag ? 0.0 : elem(ca, sa, cb, sb0, a10, a11, a00, a01);
          Qn[qa * ld + pb] = cb * q10 + sb0 * q11;
          if (hb) {
            An[qa * ld + qb] = elem(ca, sa, cb, sb, a11, a10, a01, a00);
            Qn[qa * ld + qb] = cb * q11 + sb * q10;
          }
        }
      }
      __syncthreads();
      double* t = Ac;
      Ac = An;
      An = t;
      t = Qc;
      Qc = Qn;
      Qn = t;
    }
    if (s_rot[fl] == 0) break;
  }
  int* perm = reinterpret_cast<int*>(rot);
  if (threadIdx.x == 0) {
    for (int i = 0; i < m; ++i) {
      w[i] = Ac[i * ld + i];
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
    An[i * ld + j] = Qc[i * ld + perm[j]];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < m * m; e += blockDim.x)
    Q0[(e / m) * ld + e % m] = An[(e / m) * ld + e % m];
  __syncthreads();
  return sweeps;
}

}  // namespace pmx
