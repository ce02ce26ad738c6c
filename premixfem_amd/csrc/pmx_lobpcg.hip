// Level-2 coarse space: every subdomain's generalized eigenproblem
// A_i^N phi = lambda M_i phi (src/PreMixFEM_3D.c:224-421) solved by LOBPCG,
// all subdomains in one persistent launch (one workgroup per subdomain).
//
// * The constant vector is the exact first eigenpair (A_i^N 1 = 0) and is
//   deflated analytically; LOBPCG iterates on the next en_lv1-1 pairs with a
//   block of m = en_lv1-1+guard vectors.
// * Preconditioner: the FP32 block ICC(0) of A_i^N + sigma M_i, applied with
//   the same hyperplane sweeps as level 1.
// * Rayleigh-Ritz on S = [X W P] with SVQB (scaled eigen-orthonormalisation)
//   of the M-Gram, so dependent directions are dropped instead of breaking a
//   Cholesky.  Converged when ||A x - lambda M x||_{M^-1} <= tol |lambda| for
//   the wanted pairs (SLEPc's relative criterion, default tol 1e-8).
// * Output: M-orthonormal eigenvectors with a fixed sign convention, written
//   into the basis rows, eigenvalues ascending with lambda_0 = 0.
#include <cmath>

#include "pmx_dense.h"
#include "pmx_kernels.h"

namespace pmx {

constexpr int kLobM = 10;            // largest LOBPCG block
constexpr int kLobS = 3 * kLobM;     // largest Rayleigh-Ritz basis
constexpr int kLobCh = 4;            // preconditioner columns per sweep

struct LobArgs {
  int m, kwant, maxit, kmax;
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
  const double* x0;   // optional warm start rows [cell][m]
  double* phi;        // output rows [cell][kmax]
  double* eig;        // output [nsub][kmax]
  int* its;           // output [nsub]
  int* status;        // output [nsub] 0 = converged
  int debug;          // subdomain whose iterations are traced with printf (-1: none)
};

__device__ __forceinline__ double hash_uniform(uint64_t a) {
  a += 0x9E3779B97F4A7C15ull;
  a = (a ^ (a >> 30)) * 0xBF58476D1CE4E5B9ull;
  a = (a ^ (a >> 27)) * 0x94D049BB133111EBull;
  a ^= a >> 31;
  return (double)(a >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// Reduce K per-thread values over the workgroup into out[0..K) (LDS).
__device__ inline void wg_reduce_many(double* acc, int K, double* red, double* out) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int k = 0; k < K; ++k) {
    double v = acc[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[w * K + k] = v;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < K; k += blockDim.x) {
    double s = 0.0;
    for (int i = 0; i < nw; ++i) s += red[i * K + k];
    out[k] = s;
  }
  __syncthreads();
}

// out[i*ldo + j] = sum_r L_i[r] wgt[r] R_j[r]  (columns contiguous, stride n).
__device__ inline void wg_multidot(const double* __restrict__ L, int nl,
                                   const double* __restrict__ R, int nr,
                                   const double* __restrict__ wgt, int n, double* red,
                                   double* tmpout, double* out, int ldo) {
  double acc[kLobM * kLobM];
#pragma unroll
  for (int e = 0; e < kLobM * kLobM; ++e) acc[e] = 0.0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    double lv[kLobM], rv[kLobM];
    const double wr = wgt ? wgt[r] : 1.0;
#pragma unroll
    for (int i = 0; i < kLobM; ++i) lv[i] = i < nl ? L[(size_t)i * n + r] * wr : 0.0;
#pragma unroll
    for (int j = 0; j < kLobM; ++j) rv[j] = j < nr ? R[(size_t)j * n + r] : 0.0;
#pragma unroll
    for (int i = 0; i < kLobM; ++i)
#pragma unroll
      for (int j = 0; j < kLobM; ++j) acc[i * kLobM + j] += lv[i] * rv[j];
  }
  // Wave reduction of the needed entries (static register indices), then the
  // sum over the waves.
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const int K = nl * nr;
#pragma unroll
  for (int i = 0; i < kLobM; ++i)
#pragma unroll
    for (int j = 0; j < kLobM; ++j) {
      if (i < nl && j < nr) {
        double v = acc[i * kLobM + j];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) red[wv * K + i * nr + j] = v;
      }
    }
  __syncthreads();
  for (int e = threadIdx.x; e < K; e += blockDim.x) {
    double s = 0.0;
    for (int q = 0; q < nwv; ++q) s += red[q * K + e];
    out[(e / nr) * ldo + e % nr] = s;
  }
  __syncthreads();
  (void)tmpout;
}

// Y = A_i^N V for nc columns (interior faces only).
__device__ inline void wg_apply_AN(const Brick& b, const double* __restrict__ Tx,
                                   const double* __restrict__ Ty, const double* __restrict__ Tz,
                                   const double* __restrict__ V, double* __restrict__ Y, int nc) {
  const int n = b.nx * b.ny * b.nz, nxny = b.nx * b.ny;
  for (int li = threadIdx.x; li < n; li += blockDim.x) {
    const int lx = li % b.nx, ly = (li / b.nx) % b.ny, lz = li / nxny;
    const int c = b.off + li;
    const double tl0 = lx > 0 ? Tx[c] : 0.0, th0 = lx < b.nx - 1 ? Tx[c + 1] : 0.0;
    const double tl1 = ly > 0 ? Ty[c] : 0.0, th1 = ly < b.ny - 1 ? Ty[c + b.nx] : 0.0;
    const double tl2 = lz > 0 ? Tz[c] : 0.0, th2 = lz < b.nz - 1 ? Tz[c + nxny] : 0.0;
    for (int j = 0; j < nc; ++j) {
      const double* v = V + (size_t)j * n;
      const double vc = v[li];
      double a = 0.0;
      if (lx > 0) a += tl0 * (vc - v[li - 1]);
      if (lx < b.nx - 1) a += th0 * (vc - v[li + 1]);
      if (ly > 0) a += tl1 * (vc - v[li - b.nx]);
      if (ly < b.ny - 1) a += th1 * (vc - v[li + b.nx]);
      if (lz > 0) a += tl2 * (vc - v[li - nxny]);
      if (lz < b.nz - 1) a += th2 * (vc - v[li + nxny]);
      Y[(size_t)j * n + li] = a;
    }
  }
  __syncthreads();
}

// W <- T W with the FP32 block ICC(0) sweeps, kLobCh columns at a time.
__device__ inline void wg_precondition(const Brick& b, const int* ps, const int* pc,
                                       const float4* __restrict__ pf,
                                       const float4* __restrict__ pb, double* W, int nc,
                                       float* v) {
  const int n = b.nx * b.ny * b.nz, nxny = b.nx * b.ny;
  for (int j0 = 0; j0 < nc; j0 += kLobCh) {
    const int ch = min(kLobCh, nc - j0);
    for (int e = threadIdx.x; e < ch * n; e += blockDim.x) v[e] = (float)W[(size_t)j0 * n + e];
    __syncthreads();
    for (int s = 0; s < b.nplanes; ++s) {
      for (int pos = ps[s] + threadIdx.x; pos < ps[s + 1]; pos += blockDim.x) {
        const int pk = pc[pos];
        const int lx = pk & 1023, ly = (pk >> 10) & 1023, lz = pk >> 20;
        const int li = (lz * b.ny + ly) * b.nx + lx;
        const float4 f = pf[b.off + pos];
        for (int q = 0; q < ch; ++q) {
          float* vv = v + q * n;
          float acc = vv[li];
          if (lx > 0) acc += f.y * vv[li - 1];
          if (ly > 0) acc += f.z * vv[li - b.nx];
          if (lz > 0) acc += f.w * vv[li - nxny];
          vv[li] = acc * f.x;
        }
      }
      __syncthreads();
    }
    for (int s = b.nplanes - 1; s >= 0; --s) {
      for (int pos = ps[s] + threadIdx.x; pos < ps[s + 1]; pos += blockDim.x) {
        const int pk = pc[pos];
        const int lx = pk & 1023, ly = (pk >> 10) & 1023, lz = pk >> 20;
        const int li = (lz * b.ny + ly) * b.nx + lx;
        const float4 u = pb[b.off + pos];
        for (int q = 0; q < ch; ++q) {
          float* vv = v + q * n;
          float acc = 0.0f;
          if (lx < b.nx - 1) acc += u.y * vv[li + 1];
          if (ly < b.ny - 1) acc += u.z * vv[li + b.nx];
          if (lz < b.nz - 1) acc += u.w * vv[li + nxny];
          vv[li] = vv[li] + acc * u.x;
        }
      }
      __syncthreads();
    }
    for (int e = threadIdx.x; e < ch * n; e += blockDim.x) W[(size_t)j0 * n + e] = (double)v[e];
    __syncthreads();
  }
}

// Row-wise in-place transform: columns [0,nout) of B <- B(:, 0:nin) C.
__device__ inline void wg_transform(double* B, int n, int nin, const double* C, int ldc, int nout) {
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    double v[kLobM];
#pragma unroll
    for (int k = 0; k < kLobM; ++k) v[k] = k < nin ? B[(size_t)k * n + r] : 0.0;
#pragma unroll
    for (int c = 0; c < kLobM; ++c) {
      if (c >= nout) break;
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < kLobM; ++k)
        if (k < nin) s += v[k] * C[k * ldc + c];
      B[(size_t)c * n + r] = s;
    }
  }
  __syncthreads();
}

// SVQB of an nb x nb Gram G (destroyed): Z = D U_kept theta^{-1/2} with
// D = diag(G)^{-1/2}; returns the number of kept columns.
__device__ inline int wg_svqb(double* G, double* Z, double* tmp, double* Q, double* w, double* rot,
                              int nb, int ld, double drop) {
  __shared__ double dsc[kLobS];
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
  wg_syev(G, Q, tmp, w, rot, nb, ld);
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
    Z[i * ld + c] = dsc[i] * Q[i * ld + first + c] / sqrt(w[first + c]);
  }
  __syncthreads();
  return keep;
}

__global__ void __launch_bounds__(kBlock) k_lobpcg(Geo g, LobArgs a) {
  extern __shared__ double lds[];
  const int sidx = blockIdx.x;
  const Brick b = g.bricks[sidx];
  const int n = b.nx * b.ny * b.nz;
  const int m = a.m;
  const int* ps = g.plane_start + g.shape_pstart[b.shape];
  const int* pc = g.plane_cell + g.shape_coff[b.shape];
  const double* mass = a.mass + b.off;
  const size_t base = (size_t)b.off * m;
  double* X = a.X + base;
  double* AX = a.AX + base;
  double* W = a.W + base;
  double* AW = a.AW + base;
  double* P = a.P + base;
  double* AP = a.AP + base;

  // LDS carve: small matrices; the preconditioner reuses the same space.
  constexpr int LD = kLobS;
  double* GA = lds;
  double* GB = GA + LD * LD;
  double* Z = GB + LD * LD;
  double* Q = Z + LD * LD;
  double* TM = Q + LD * LD;
  double* w = TM + LD * LD;
  double* rot = w + LD;
  double* red = rot + 4 * LD;                 // 4 waves x 100 values
  double* tmpout = red + 4 * kLobM * kLobM;
  float* pv = reinterpret_cast<float*>(lds);  // preconditioner buffer
  __shared__ double lam[kLobM], rho[kLobM], msum_s[1];
  __shared__ int s_conv;

  // sum of the mass diagonal (for the constant eigenvector)
  {
    double s = 0.0;
    for (int r = threadIdx.x; r < n; r += blockDim.x) s += mass[r];
    s = block_sum(s, red);
    if (threadIdx.x == 0) msum_s[0] = s;
    __syncthreads();
  }
  const double msum = msum_s[0];
  const int kw = a.kwant;
  int it = 0;
  int status = 0;
  if (kw > 0) {
    for (int attempt = 0; attempt < 2; ++attempt) {
      // ---- initial block ----
      for (int e = threadIdx.x; e < m * n; e += blockDim.x) {
        const int j = e / n, r = e % n;
        double v;
        if (a.x0 && attempt == 0)
          v = a.x0[(size_t)(b.off + r) * m + j];
        else
          v = hash_uniform(a.seed ^ ((uint64_t)sidx << 40) ^ ((uint64_t)j << 32) ^ (uint64_t)r ^
                           ((uint64_t)attempt << 60));
        X[e] = v;
      }
      __syncthreads();
      // deflate the constant: v -= (sum m v / sum m)
      wg_multidot(X, m, mass, 1, nullptr, n, red, tmpout, TM, 1);
      for (int e = threadIdx.x; e < m * n; e += blockDim.x) X[e] -= TM[e / n] / msum;
      __syncthreads();
      // M-orthonormalise X
      wg_multidot(X, m, X, m, mass, n, red, tmpout, GB, LD);
      int keep = wg_svqb(GB, Z, TM, Q, w, rot, m, LD, 1e-14);
      if (keep < m) {  // degenerate start: fall back to random
        if (attempt == 0) continue;
        status = 2;
        break;
      }
      wg_transform(X, n, m, Z, LD, m);
      wg_apply_AN(b, a.Tx, a.Ty, a.Tz, X, AX, m);
      wg_multidot(X, m, AX, m, nullptr, n, red, tmpout, GA, LD);
      for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
        const int i = e / m, j = e % m;
        if (j > i) GA[i * LD + j] = GA[j * LD + i] = 0.5 * (GA[i * LD + j] + GA[j * LD + i]);
      }
      __syncthreads();
      wg_syev(GA, Q, TM, w, rot, m, LD);
      if (threadIdx.x < m) lam[threadIdx.x] = w[threadIdx.x];
      __syncthreads();
      wg_transform(X, n, m, Q, LD, m);
      wg_transform(AX, n, m, Q, LD, m);
      int np = 0;
      status = 1;
      for (it = 0; it < a.maxit; ++it) {
        // ---- residuals R = AX - M X Lambda (into W) and their M^-1 norms ----
        {
          double acc[kLobM];
#pragma unroll
          for (int j = 0; j < kLobM; ++j) acc[j] = 0.0;
          for (int r = threadIdx.x; r < n; r += blockDim.x) {
            const double mr = mass[r];
#pragma unroll
            for (int j = 0; j < kLobM; ++j) {
              if (j < m) {
                const double rv = AX[(size_t)j * n + r] - lam[j] * mr * X[(size_t)j * n + r];
                W[(size_t)j * n + r] = rv;
                acc[j] += rv * rv / mr;
              }
            }
          }
          wg_reduce_many(acc, m, red, tmpout);
          if (threadIdx.x < m) rho[threadIdx.x] = sqrt(tmpout[threadIdx.x]);
          __syncthreads();
        }
        // ---- W = T R, deflated against the constant vector ----
        wg_precondition(b, ps, pc, a.pf, a.pb, W, m, pv);
        wg_multidot(W, m, mass, 1, nullptr, n, red, tmpout, TM, 1);
        for (int e = threadIdx.x; e < m * n; e += blockDim.x) W[e] -= TM[e / n] / msum;
        __syncthreads();
        // ---- convergence: ||T r||_M <= tol for the wanted pairs.  With T
        // approximating (A^N + sigma M)^{-1} this is the residual of the
        // shift-inverted problem relative to its eigenvalue, the quantity
        // SLEPc's relative test bounds under STSINVERT. ----
        {
          double acc[kLobM];
#pragma unroll
          for (int j = 0; j < kLobM; ++j) acc[j] = 0.0;
          for (int r = threadIdx.x; r < n; r += blockDim.x) {
            const double mr = mass[r];
#pragma unroll
            for (int j = 0; j < kLobM; ++j)
              if (j < m) {
                const double wv = W[(size_t)j * n + r];
                acc[j] += wv * wv * mr;
              }
          }
          wg_reduce_many(acc, m, red, tmpout);
          if (threadIdx.x == 0) {
            int conv = 1;
            for (int j = 0; j < kw; ++j)
              if (!(sqrt(tmpout[j]) <= a.tol)) conv = 0;
            s_conv = conv;
            if (sidx == a.debug)
              printf("it %d lam0 %.6e lam%d %.6e rho0 %.3e rho%d %.3e Tr0 %.3e Tr%d %.3e np %d\n",
                     it, lam[0], kw - 1, lam[kw - 1], rho[0], kw - 1, rho[kw - 1], sqrt(tmpout[0]),
                     kw - 1, sqrt(tmpout[kw - 1]), np);
          }
          __syncthreads();
        }
        if (s_conv) {
          status = 0;
          break;
        }
        // ---- W -= X (X^T M W), P (P^T M W) ----
        wg_multidot(X, m, W, m, mass, n, red, tmpout, GB, LD);
        if (np > 0) wg_multidot(P, np, W, m, mass, n, red, tmpout, GB + m * LD, LD);
        for (int r = threadIdx.x; r < n; r += blockDim.x) {
          double xv[kLobM], pv2[kLobM];
#pragma unroll
          for (int i = 0; i < kLobM; ++i) {
            xv[i] = i < m ? X[(size_t)i * n + r] : 0.0;
            pv2[i] = i < np ? P[(size_t)i * n + r] : 0.0;
          }
#pragma unroll
          for (int j = 0; j < kLobM; ++j) {
            if (j >= m) break;
            double s = W[(size_t)j * n + r];
#pragma unroll
            for (int i = 0; i < kLobM; ++i) {
              if (i < m) s -= xv[i] * GB[i * LD + j];
              if (i < np) s -= pv2[i] * GB[(m + i) * LD + j];
            }
            W[(size_t)j * n + r] = s;
          }
        }
        __syncthreads();
        // ---- SVQB(W) ----
        wg_multidot(W, m, W, m, mass, n, red, tmpout, GB, LD);
        int nw = wg_svqb(GB, Z, TM, Q, w, rot, m, LD, 1e-12);
        if (nw == 0) {
          status = 3;
          break;
        }
        wg_transform(W, n, m, Z, LD, nw);
        wg_apply_AN(b, a.Tx, a.Ty, a.Tz, W, AW, nw);
        // ---- Gram matrices of S = [X W P] ----
        const int ns = m + nw + np;
        const double* Sb[3] = {X, W, P};
        const double* ASb[3] = {AX, AW, AP};
        const int sz[3] = {m, nw, np};
        const int so[3] = {0, m, m + nw};
        for (int bi = 0; bi < 3; ++bi)
          for (int bj = bi; bj < 3; ++bj) {
            if (sz[bi] == 0 || sz[bj] == 0) continue;
            wg_multidot(Sb[bi], sz[bi], ASb[bj], sz[bj], nullptr, n, red, tmpout,
                        GA + so[bi] * LD + so[bj], LD);
            wg_multidot(Sb[bi], sz[bi], Sb[bj], sz[bj], mass, n, red, tmpout,
                        GB + so[bi] * LD + so[bj], LD);
          }
        for (int e = threadIdx.x; e < ns * ns; e += blockDim.x) {
          const int i = e / ns, j = e % ns;
          // blocks were computed for block-row <= block-col; mirror and
          // symmetrise the diagonal blocks
          const int bi = i < m ? 0 : (i < m + nw ? 1 : 2);
          const int bj = j < m ? 0 : (j < m + nw ? 1 : 2);
          if (bi > bj) {
            GA[i * LD + j] = GA[j * LD + i];
            GB[i * LD + j] = GB[j * LD + i];
          }
        }
        __syncthreads();
        for (int e = threadIdx.x; e < ns * ns; e += blockDim.x) {
          const int i = e / ns, j = e % ns;
          if (j > i) {
            const double ga = 0.5 * (GA[i * LD + j] + GA[j * LD + i]);
            const double gb = 0.5 * (GB[i * LD + j] + GB[j * LD + i]);
            GA[i * LD + j] = GA[j * LD + i] = ga;
            GB[i * LD + j] = GB[j * LD + i] = gb;
          }
        }
        __syncthreads();
        // ---- Rayleigh-Ritz: SVQB of G_B, then eig of Z^T G_A Z ----
        const int nk = wg_svqb(GB, Z, TM, Q, w, rot, ns, LD, 1e-13);
        if (nk < m) {
          status = 4;
          break;
        }
        // H = Z^T G_A Z  (TM = G_A Z, then GB = Z^T TM)
        for (int e = threadIdx.x; e < ns * nk; e += blockDim.x) {
          const int i = e / nk, c = e % nk;
          double s = 0.0;
          for (int k = 0; k < ns; ++k) s += GA[i * LD + k] * Z[k * LD + c];
          TM[i * LD + c] = s;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < nk * nk; e += blockDim.x) {
          const int i = e / nk, c = e % nk;
          double s = 0.0;
          for (int k = 0; k < ns; ++k) s += Z[k * LD + i] * TM[k * LD + c];
          GB[i * LD + c] = s;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < nk * nk; e += blockDim.x) {
          const int i = e / nk, c = e % nk;
          if (c > i) GB[i * LD + c] = GB[c * LD + i] = 0.5 * (GB[i * LD + c] + GB[c * LD + i]);
        }
        __syncthreads();
        wg_syev(GB, Q, TM, w, rot, nk, LD);
        // C = Z Q(:, 0:m)  -> GA (GA no longer needed)
        for (int e = threadIdx.x; e < ns * m; e += blockDim.x) {
          const int i = e / m, c = e % m;
          double s = 0.0;
          for (int k = 0; k < nk; ++k) s += Z[i * LD + k] * Q[k * LD + c];
          GA[i * LD + c] = s;
        }
        if (threadIdx.x < m) lam[threadIdx.x] = w[threadIdx.x];
        __syncthreads();
        // ---- X = S C, P = [W P] C_WP; same for the A-images ----
        for (int pass = 0; pass < 2; ++pass) {
          double* B0 = pass == 0 ? X : AX;
          const double* B1 = pass == 0 ? W : AW;
          double* B2 = pass == 0 ? P : AP;
          for (int r = threadIdx.x; r < n; r += blockDim.x) {
            double v0[kLobM], v1[kLobM], v2[kLobM];
#pragma unroll
            for (int k = 0; k < kLobM; ++k) {
              v0[k] = k < m ? B0[(size_t)k * n + r] : 0.0;
              v1[k] = k < nw ? B1[(size_t)k * n + r] : 0.0;
              v2[k] = k < np ? B2[(size_t)k * n + r] : 0.0;
            }
#pragma unroll
            for (int c = 0; c < kLobM; ++c) {
              if (c >= m) break;
              double sx = 0.0, sp = 0.0;
#pragma unroll
              for (int k = 0; k < kLobM; ++k) {
                if (k < m) sx += v0[k] * GA[k * LD + c];
                if (k < nw) sp += v1[k] * GA[(m + k) * LD + c];
                if (k < np) sp += v2[k] * GA[(m + nw + k) * LD + c];
              }
              B0[(size_t)c * n + r] = sx + sp;
              B2[(size_t)c * n + r] = sp;
            }
          }
          __syncthreads();
        }
        // ---- keep P M-orthogonal to X and M-orthonormal (robust LOBPCG):
        // P -= X (X^T M P), AP -= AX (X^T M P), then SVQB(P) ----
        wg_multidot(X, m, P, m, mass, n, red, tmpout, GB, LD);
        for (int r = threadIdx.x; r < n; r += blockDim.x) {
          double xv[kLobM], axv[kLobM];
#pragma unroll
          for (int i = 0; i < kLobM; ++i) {
            xv[i] = i < m ? X[(size_t)i * n + r] : 0.0;
            axv[i] = i < m ? AX[(size_t)i * n + r] : 0.0;
          }
#pragma unroll
          for (int j = 0; j < kLobM; ++j) {
            if (j >= m) break;
            double sp = P[(size_t)j * n + r], sa = AP[(size_t)j * n + r];
#pragma unroll
            for (int i = 0; i < kLobM; ++i)
              if (i < m) {
                sp -= xv[i] * GB[i * LD + j];
                sa -= axv[i] * GB[i * LD + j];
              }
            P[(size_t)j * n + r] = sp;
            AP[(size_t)j * n + r] = sa;
          }
        }
        __syncthreads();
        wg_multidot(P, m, P, m, mass, n, red, tmpout, GB, LD);
        np = wg_svqb(GB, Z, TM, Q, w, rot, m, LD, 1e-12);
        if (np > 0) {
          wg_transform(P, n, m, Z, LD, np);
          wg_transform(AP, n, m, Z, LD, np);
        }
      }
      if (status == 0) break;
      if (attempt == 0 && a.x0 == nullptr) break;  // random start already failed
    }
  } else {
    status = 0;
  }
  // ---- outputs ----
  const double phi0 = 1.0 / sqrt(msum);
  double* phi = a.phi + (size_t)b.off * a.kmax;
  for (int r = threadIdx.x; r < n; r += blockDim.x) phi[(size_t)r * a.kmax] = phi0;
  if (kw > 0 && status == 0) {
    // sign: positive inner product with 1 + frac(0.618..(r+1))
    double acc[kLobM];
    for (int j = 0; j < kLobM; ++j) acc[j] = 0.0;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
      const double wgt = 1.0 + fmod(0.6180339887498949 * (r + 1.0), 1.0);
      for (int j = 0; j < kw; ++j) acc[j] += wgt * X[(size_t)j * n + r];
    }
    wg_reduce_many(acc, kw, red, tmpout);
    for (int e = threadIdx.x; e < kw * n; e += blockDim.x) {
      const int j = e / n, r = e % n;
      const double sg = tmpout[j] < 0.0 ? -1.0 : 1.0;
      phi[(size_t)r * a.kmax + 1 + j] = sg * X[(size_t)j * n + r];
    }
  }
  if (threadIdx.x == 0) {
    a.eig[(size_t)sidx * a.kmax] = 0.0;
    for (int j = 0; j < kw; ++j) a.eig[(size_t)sidx * a.kmax + 1 + j] = lam[j];
    a.its[sidx] = it;
    a.status[sidx] = status;
  }
}

int lobpcg_setup(pmx_ctx* ctx, const double* x0_rows) {
  const int kw = ctx->en_lv1 - 1;
  const int m = std::min(kLobM, kw + ctx->eig_guard);
  if (kw > kLobM - 1) {
    set_error("en_lv1=%d exceeds the LOBPCG block limit %d", ctx->en_lv1, kLobM);
    return PMX_ERR_SUP;
  }
  int maxn = 0;
  for (const Brick& b : ctx->h_bricks) maxn = std::max(maxn, b.nx * b.ny * b.nz);
  if (m > maxn - 1) {
    set_error("subdomains of %d cells are too small for %d eigenvectors", maxn, ctx->en_lv1);
    return PMX_ERR_ARG_WRONG;
  }
  // LOBPCG preconditioner: FP32 block ICC(0) of A^N + sigma M.
  DBuf<float> pf, pb;
  PMX_TRY(pf.alloc((size_t)ctx->ncell * 4));
  PMX_TRY(pb.alloc((size_t)ctx->ncell * 4));
  PMX_TRY(launch_dic_factor(ctx, 1, ctx->eig_sigma, pf.p, pb.p, true));
  DBuf<double> buf;
  const size_t blk = (size_t)ctx->ncell * std::max(m, 1);
  PMX_TRY(buf.alloc(6 * blk));
  DBuf<double> eig;
  DBuf<int> its, status;
  PMX_TRY(eig.alloc((size_t)ctx->nsub * ctx->en_lv1));
  PMX_TRY(its.alloc(ctx->nsub));
  PMX_TRY(status.alloc(ctx->nsub));
  LobArgs a;
  a.m = m;
  a.kwant = kw;
  a.maxit = ctx->eig_maxit;
  a.kmax = ctx->en_lv1;
  a.tol = ctx->eig_tol;
  a.seed = ctx->seed;
  a.X = buf.p;
  a.AX = buf.p + blk;
  a.W = buf.p + 2 * blk;
  a.AW = buf.p + 3 * blk;
  a.P = buf.p + 4 * blk;
  a.AP = buf.p + 5 * blk;
  a.pf = (const float4*)pf.p;
  a.pb = (const float4*)pb.p;
  a.Tx = ctx->T[0].p;
  a.Ty = ctx->T[1].p;
  a.Tz = ctx->T[2].p;
  a.mass = ctx->mass.p;
  a.x0 = x0_rows;
  a.phi = ctx->phi.p;
  a.eig = eig.p;
  a.its = its.p;
  a.status = status.p;
  a.debug = ctx->eig_debug;
  const size_t small = (5 * kLobS * kLobS + kLobS + 4 * kLobS + 4 * kLobM * kLobM + kLobM * kLobM) *
                       sizeof(double);
  const size_t prec = (size_t)maxn * kLobCh * sizeof(float);
  const size_t lds = std::max(small, prec);
  hipLaunchKernelGGL(k_lobpcg, dim3(ctx->nsub), dim3(kBlock), lds, ctx->stream, ctx->geo, a);
  PMX_HIP(hipGetLastError());
  std::vector<int> hits(ctx->nsub), hst(ctx->nsub);
  ctx->h_eig_lv1.resize((size_t)ctx->nsub * ctx->en_lv1);
  PMX_HIP(hipMemcpyAsync(hits.data(), its.p, ctx->nsub * sizeof(int), hipMemcpyDeviceToHost,
                         ctx->stream));
  PMX_HIP(hipMemcpyAsync(hst.data(), status.p, ctx->nsub * sizeof(int), hipMemcpyDeviceToHost,
                         ctx->stream));
  PMX_HIP(hipMemcpyAsync(ctx->h_eig_lv1.data(), eig.p, ctx->h_eig_lv1.size() * sizeof(double),
                         hipMemcpyDeviceToHost, ctx->stream));
  PMX_HIP(hipStreamSynchronize(ctx->stream));
  int bad = 0, first_bad = -1;
  ctx->eig_iters_max = 0;
  ctx->eig_iters_sum = 0;
  for (int s = 0; s < ctx->nsub; ++s) {
    ctx->eig_iters_max = std::max(ctx->eig_iters_max, hits[s]);
    ctx->eig_iters_sum += hits[s];
    if (hst[s] != 0) {
      if (first_bad < 0) first_bad = s;
      ++bad;
    }
  }
  if (bad) {
    set_error("eigensolver could not converge %d of %d subdomain eigenproblems (first: %d, "
              "status %d)",
              bad, ctx->nsub, first_bad, hst[first_bad]);
    return PMX_ERR_USER;
  }
  return PMX_SUCCESS;
}

}  // namespace pmx
