// Level-2 coarse space: every subdomain's generalized eigenproblem
// A_i^N phi = lambda M_i phi (src/PreMixFEM_3D.c:224-421) solved by LOBPCG,
// all subdomains in one persistent launch (one workgroup per subdomain).
//
// * The constant vector is the exact first eigenpair (A_i^N 1 = 0) and is
//   deflated analytically; LOBPCG iterates on the next en_lv1-1 pairs with a
//   block of MB vectors (MB >= en_lv1-1 + 3 guard vectors).  A subdomain with
//   a well term (the spe10 producers, src/spe10.c:1005-1012) has no null
//   vector: there LOBPCG iterates on all en_lv1 pairs without deflation.
// * Preconditioner: the FP32 block ICC(0) of A_i^N + sigma M_i, applied with
//   the same hyperplane sweeps as level 1.
// * The basis S = [X W P] is kept M-orthonormal block by block (W and P are
//   projected against X (and P) and SVQB-orthonormalised), so the
//   Rayleigh-Ritz M-Gram is I plus small cross blocks: one Cholesky-reduced
//   symmetric eigenproblem per iteration; if that Cholesky breaks down the
//   iteration restarts without P.
// * Convergence: ||T r||_M <= tol for every wanted pair, T ~ (A^N+sigma M)^{-1}
//   -- the residual of the shift-inverted problem relative to its eigenvalue,
//   the quantity SLEPc's relative test bounds under STSINVERT (default 1e-8).
// * Output: M-orthonormal eigenvectors with a fixed sign convention, written
//   into the basis rows, eigenvalues ascending with lambda_0 = 0.
#include <algorithm>
#include <cmath>
#include <map>

#include "pmx_lobpcg.h"

namespace pmx {

template <int MB, int BS>
__global__ void __launch_bounds__(BS) k_lobpcg(Geo g, LobArgs a) {
  constexpr int LD = 3 * MB;
  extern __shared__ double lds[];
  const int sidx = a.order ? a.order[blockIdx.x] : blockIdx.x * a.stride;
  const Brick b = g.bricks[sidx];
  const BrickView v = load_view(g, sidx);
  const int n = v.n;
  const int* ps = g.plane_start + g.shape_pstart[b.shape];
  const int* pc = g.plane_cell + g.shape_coff[b.shape];
  const double* mass = a.mass + v.off;
  const size_t base = (size_t)v.off * MB;
  double* X = a.X + base;
  double* AX = a.AX + base;
  double* W = a.W + base;
  double* AW = a.AW + base;
  double* P = a.P + base;
  double* AP = a.AP + base;

  // LDS: small matrices (the preconditioner buffer reuses the same space).
  double* GA = lds;
  double* GB = GA + LD * LD;
  double* Z = GB + LD * LD;
  double* Q = Z + LD * LD;
  double* TM = Q + LD * LD;
  double* w = TM + LD * LD;
  double* rot = w + LD;
  double* red = rot + 3 * LD;  // kLobWaves x MB*MB
  float* pv = reinterpret_cast<float*>(lds);
  __shared__ double lam[MB], nrm[MB], msum_s;
  __shared__ int s_conv, s_flag, s_lock[MB], s_itbest, s_stag;
  __shared__ double s_rbest;

  __shared__ int s_well;
  {
    double s = 0.0, wsum = 0.0;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
      s += mass[r];
      if (g.wd) wsum += g.wd[v.off + r];
    }
    s = block_sum(s, red);
    wsum = block_sum(wsum, red);
    if (threadIdx.x == 0) {
      msum_s = s;
      s_well = wsum > 0.0;
    }
    __syncthreads();
  }
  const double msum = msum_s;
  // A subdomain with a well term has no null vector: all en_lv1 pairs are
  // iterated and the constant is not deflated.
  const bool well = s_well != 0;
  const int kw = well ? a.kwant + 1 : a.kwant;
  int it = 0, status = 0, np = 0;
  unsigned long long tph[10] = {0}, tlast = 0;
  const bool trace = sidx == a.debug && threadIdx.x == 0;
#define PMX_PHASE(i)                                                   \
  if (trace) {                                                         \
    const unsigned long long tn = __builtin_amdgcn_s_memrealtime();    \
    if (tlast) tph[i] += tn - tlast;                                   \
    tlast = tn;                                                        \
  }
  // B -= 1 (sum m B / sum m) for the MB columns of B.
  auto deflate = [&](double* B) {
    if (well) return;
    wg_multidot<MB, 1>(B, mass, nullptr, n, red, TM, 1);
    for (int e = threadIdx.x; e < MB * n; e += blockDim.x) B[e] -= TM[e / n] / msum;
    __syncthreads();
  };
  // G(r0.., c0..) = (L^T diag(wgt) R)(0:nl, 0:nr) through TM.
  auto gram = [&](const double* L, int nl, const double* R, int nr, const double* wgt, double* G,
                  int r0, int c0) {
    wg_multidot<MB, MB>(L, R, wgt, n, red, TM, LD);
    for (int e = threadIdx.x; e < nl * nr; e += blockDim.x)
      G[(r0 + e / nr) * LD + c0 + e % nr] = TM[(e / nr) * LD + e % nr];
    __syncthreads();
  };
  int attempts = 0;
  if (kw > 0) {
    for (int attempt = 0; attempt < 2; ++attempt) {
      attempts = attempt;
      // ---- initial block: warm start rows or seeded random ----
      for (int e = threadIdx.x; e < MB * n; e += blockDim.x) {
        const int j = e / n, r = e % n;
        X[e] = (a.x0 && attempt == 0)
                   ? a.x0[(size_t)(v.off + r) * MB + j]
                   : hash_uniform(a.seed ^ ((uint64_t)b.lid << 40) ^ ((uint64_t)j << 32) ^
                                  (uint64_t)r ^ ((uint64_t)attempt << 60));
      }
      __syncthreads();
      deflate(X);
      wg_multidot<MB, MB>(X, X, mass, n, red, GB, LD);
      const int keep = wg_svqb(GB, GA, Z, Q, w, rot, MB, LD, 1e-14);
      if (keep < MB) {
        status = 2;
        continue;
      }
      wg_transform<MB>(X, n, Z, LD, MB, MB);
      wg_apply_AN<MB>(v, a.Tx, a.Ty, a.Tz, g.wd, X, AX, MB);
      wg_multidot<MB, MB>(X, AX, nullptr, n, red, GA, LD);
      for (int e = threadIdx.x; e < MB * MB; e += blockDim.x) {
        const int i = e / MB, j = e % MB;
        if (j > i) GA[i * LD + j] = GA[j * LD + i] = 0.5 * (GA[i * LD + j] + GA[j * LD + i]);
      }
      __syncthreads();
      wg_syev2(GA, GB, Q, Z, w, rot, MB, LD);
      if (threadIdx.x < MB) lam[threadIdx.x] = w[threadIdx.x];
      __syncthreads();
      wg_transform<MB>(X, n, Q, LD, MB, MB);
      wg_transform<MB>(AX, n, Q, LD, MB, MB);
      np = 0;
      status = 1;
      if (threadIdx.x == 0) {
        s_rbest = 1e300;
        s_itbest = 0;
        s_stag = 0;
      }
      __syncthreads();
      for (it = 0; it < a.maxit; ++it) {
        // Near convergence the drift of the updated AX and AP from A X and
        // A P sets the residual's floor and can turn the iteration unstable.
        if (it > 0 && it % kRefreshIts == 0 && s_rbest <= kRefreshNear * a.tol) {
          wg_apply_AN<MB>(v, a.Tx, a.Ty, a.Tz, g.wd, X, AX, MB);
          if (np > 0) wg_apply_AN<MB>(v, a.Tx, a.Ty, a.Tz, g.wd, P, AP, np);
        }
        // ---- W = T (AX - M X Lambda), deflated ----
        for (int e = threadIdx.x; e < MB * n; e += blockDim.x) {
          const int j = e / n, r = e % n;
          W[e] = AX[e] - lam[j] * mass[r] * X[e];
        }
        __syncthreads();
        PMX_PHASE(0)
        wg_precondition(b, v, ps, pc, a.pf, a.pb, W, MB, pv);
        PMX_PHASE(1)
        deflate(W);
        // ---- convergence: ||T r||_M <= tol for the wanted pairs ----
        wg_colnorm2<MB>(W, mass, n, red, nrm);
        if (threadIdx.x == 0) {
          int conv = 1;
          double rmax = 0.0;
          for (int j = 0; j < kw; ++j) {
            const double rj = sqrt(nrm[j]);
            if (!(rj <= a.tol)) conv = 0;
            rmax = fmax(rmax, rj);
          }
          // Stagnation: the residuals of a few hard subdomains (clusters of
          // tiny eigenvalues of 1e5-contrast media) level off just above tol
          // at the rounding floor of T r.  Accept once the worst residual is
          // within stag_factor * tol and has not halved for kStagIts its.
          if (rmax < 0.5 * s_rbest) {
            s_rbest = rmax;
            s_itbest = it;
          }
          if (!conv && it - s_itbest >= kStagIts && rmax <= a.stag_factor * a.tol) {
            conv = 1;
            s_stag = 1;
          }
          s_conv = conv;
          if (sidx == a.debug && (it < 3 || it % 10 == 0))
            printf("it %d lam0 %.6e lam%d %.6e Tr0 %.3e Tr%d %.3e np %d\n", it, lam[0], kw - 1,
                   lam[kw - 1], sqrt(nrm[0]), kw - 1, sqrt(nrm[kw - 1]), np);
        }
        __syncthreads();
        PMX_PHASE(2)
        if (s_conv) {
          status = 0;
          break;
        }
        // ---- soft locking: converged wanted pairs add no search direction ----
        if (threadIdx.x < MB)
          s_lock[threadIdx.x] = threadIdx.x < kw && sqrt(nrm[threadIdx.x]) <= a.tol;
        __syncthreads();
        for (int e = threadIdx.x; e < MB * n; e += blockDim.x)
          if (s_lock[e / n]) W[e] = 0.0;
        __syncthreads();
        // ---- W -= X (X^T M W) + P (P^T M W), two passes (a small residual
        // is mostly cancelled by the first, the second restores
        // orthogonality to working precision) ----
        for (int pass = 0; pass < 2; ++pass) {
          wg_multidot<MB, MB>(X, W, mass, n, red, GB, LD);
          if (np > 0) wg_multidot<MB, MB>(P, W, mass, n, red, GB + MB * LD, LD);
          for (int r = threadIdx.x; r < n; r += blockDim.x) {
            double xv[MB], pv2[MB];
#pragma unroll
            for (int i = 0; i < MB; ++i) {
              xv[i] = X[(size_t)i * n + r];
              pv2[i] = i < np ? P[(size_t)i * n + r] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < MB; ++j) {
              double s = W[(size_t)j * n + r];
#pragma unroll
              for (int i = 0; i < MB; ++i) {
                s -= xv[i] * GB[i * LD + j];
                if (i < np) s -= pv2[i] * GB[(MB + i) * LD + j];
              }
              W[(size_t)j * n + r] = s;
            }
          }
          __syncthreads();
        }
        PMX_PHASE(3)
        // ---- SVQB(W) ----
        wg_multidot<MB, MB>(W, W, mass, n, red, GB, LD);
        const int nw = wg_svqb(GB, GA, Z, Q, w, rot, MB, LD, 1e-12);
        if (nw == 0) {
          status = 3;
          break;
        }
        wg_transform<MB>(W, n, Z, LD, MB, nw);
        PMX_PHASE(4)
        wg_apply_AN<MB>(v, a.Tx, a.Ty, a.Tz, g.wd, W, AW, nw);
        PMX_PHASE(5)
        // ---- Gram blocks of S = [X W P] (compact, ns = MB + nw + np): the
        // diagonal blocks of G_B are I and G_A(X,X) = Lambda by construction;
        // the cross blocks are computed. ----
        const int ns = MB + nw + np;
        for (int e = threadIdx.x; e < ns * ns; e += blockDim.x) {
          const int i = e / ns, j = e % ns;
          GA[i * LD + j] = (i == j && i < MB) ? lam[i] : 0.0;
          GB[i * LD + j] = i == j ? 1.0 : 0.0;
        }
        __syncthreads();
        gram(X, MB, AW, nw, nullptr, GA, 0, MB);
        gram(W, nw, AW, nw, nullptr, GA, MB, MB);
        gram(X, MB, W, nw, mass, GB, 0, MB);
        if (np > 0) {
          gram(X, MB, AP, np, nullptr, GA, 0, MB + nw);
          gram(W, nw, AP, np, nullptr, GA, MB, MB + nw);
          gram(P, np, AP, np, nullptr, GA, MB + nw, MB + nw);
          gram(X, MB, P, np, mass, GB, 0, MB + nw);
          gram(W, nw, P, np, mass, GB, MB, MB + nw);
        }
        for (int e = threadIdx.x; e < ns * ns; e += blockDim.x) {
          const int i = e / ns, j = e % ns;
          if (j > i) {
            GA[j * LD + i] = GA[i * LD + j];
            GB[j * LD + i] = GB[i * LD + j];
          }
        }
        __syncthreads();
        PMX_PHASE(6)
        // ---- Rayleigh-Ritz: G_B = L L^T, H = L^{-1} G_A L^{-T} = Q w Q^T,
        // C = L^{-T} Q(:, 0:MB) ----
        if (!wg_cholesky(GB, ns, LD, &s_flag)) {
          np = 0;
          status = 4;
          continue;
        }
        wg_trinv_lower(GB, TM, ns, LD);      // GB = L^{-1}
        wg_matmul(GB, GA, TM, ns, LD, false);  // TM = L^{-1} G_A
        for (int e = threadIdx.x; e < ns * ns; e += blockDim.x) {
          const int i = e / ns, j = e % ns;
          double s = 0.0;
          for (int k = 0; k <= j; ++k) s += TM[i * LD + k] * GB[j * LD + k];
          Z[i * LD + j] = s;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < ns * ns; e += blockDim.x) {
          const int i = e / ns, j = e % ns;
          if (j > i) Z[i * LD + j] = Z[j * LD + i] = 0.5 * (Z[i * LD + j] + Z[j * LD + i]);
        }
        __syncthreads();
        wg_syev2(Z, TM, Q, GA, w, rot, ns, LD);  // eigenvectors in Q
        for (int e = threadIdx.x; e < ns * MB; e += blockDim.x) {
          const int i = e / MB, c = e % MB;
          double s = 0.0;
          for (int k = i; k < ns; ++k) s += GB[k * LD + i] * Q[k * LD + c];
          TM[i * LD + c] = s;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < ns * MB; e += blockDim.x)
          GA[(e / MB) * LD + e % MB] = TM[(e / MB) * LD + e % MB];
        if (threadIdx.x < MB) lam[threadIdx.x] = w[threadIdx.x];
        __syncthreads();
        PMX_PHASE(7)
        // ---- X = S C, P = [W P] C_WP; same for the A-images ----
        for (int pass = 0; pass < 2; ++pass) {
          double* B0 = pass == 0 ? X : AX;
          const double* B1 = pass == 0 ? W : AW;
          double* B2 = pass == 0 ? P : AP;
          for (int r = threadIdx.x; r < n; r += blockDim.x) {
            double v0[MB], v1[MB], v2[MB];
#pragma unroll
            for (int k = 0; k < MB; ++k) {
              v0[k] = B0[(size_t)k * n + r];
              v1[k] = k < nw ? B1[(size_t)k * n + r] : 0.0;
              v2[k] = k < np ? B2[(size_t)k * n + r] : 0.0;
            }
#pragma unroll
            for (int c = 0; c < MB; ++c) {
              double sx = 0.0, sp = 0.0;
#pragma unroll
              for (int k = 0; k < MB; ++k) {
                sx += v0[k] * GA[k * LD + c];
                if (k < nw) sp += v1[k] * GA[(MB + k) * LD + c];
                if (k < np) sp += v2[k] * GA[(MB + nw + k) * LD + c];
              }
              B0[(size_t)c * n + r] = sx + sp;
              B2[(size_t)c * n + r] = sp;
            }
          }
          __syncthreads();
        }
        PMX_PHASE(8)
        // ---- P -= X (X^T M P), AP -= AX (X^T M P); SVQB(P) ----
        wg_multidot<MB, MB>(X, P, mass, n, red, GB, LD);
        for (int r = threadIdx.x; r < n; r += blockDim.x) {
          double xv[MB], axv[MB];
#pragma unroll
          for (int i = 0; i < MB; ++i) {
            xv[i] = X[(size_t)i * n + r];
            axv[i] = AX[(size_t)i * n + r];
          }
#pragma unroll
          for (int j = 0; j < MB; ++j) {
            double sp = P[(size_t)j * n + r], sa = AP[(size_t)j * n + r];
#pragma unroll
            for (int i = 0; i < MB; ++i) {
              sp -= xv[i] * GB[i * LD + j];
              sa -= axv[i] * GB[i * LD + j];
            }
            P[(size_t)j * n + r] = sp;
            AP[(size_t)j * n + r] = sa;
          }
        }
        __syncthreads();
        wg_multidot<MB, MB>(P, P, mass, n, red, GB, LD);
        np = wg_svqb(GB, GA, Z, Q, w, rot, MB, LD, 1e-12);
        if (np > 0) {
          wg_transform<MB>(P, n, Z, LD, MB, np);
          wg_transform<MB>(AP, n, Z, LD, MB, np);
        }
        PMX_PHASE(9)
      }
      if (status == 0) break;
      if (attempt == 0 && a.x0 == nullptr && status != 2) break;
    }
  }
  if (trace) {
    printf("lobpcg phases (us) its %d: resid %.0f prec %.0f conv %.0f proj %.0f svqbW %.0f AW %.0f "
           "gram %.0f rr %.0f update %.0f portho %.0f\n",
           it, tph[0] * 0.01, tph[1] * 0.01, tph[2] * 0.01, tph[3] * 0.01, tph[4] * 0.01,
           tph[5] * 0.01, tph[6] * 0.01, tph[7] * 0.01, tph[8] * 0.01, tph[9] * 0.01);
  }
#undef PMX_PHASE
  // ---- outputs ----
  const double phi0 = 1.0 / sqrt(msum);
  double* phi = a.phi + (size_t)v.off * a.kmax;
  const int c0 = well ? 0 : 1;  // first basis column LOBPCG fills
  if (!well)
    for (int r = threadIdx.x; r < n; r += blockDim.x) phi[(size_t)r * a.kmax] = phi0;
  if (kw > 0 && status == 0) {
    // sign: positive inner product with 1 + frac(0.618..(r+1))
    double acc[MB];
#pragma unroll
    for (int j = 0; j < MB; ++j) acc[j] = 0.0;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
      const double wgt = 1.0 + fmod(0.6180339887498949 * (r + 1.0), 1.0);
#pragma unroll
      for (int j = 0; j < MB; ++j) acc[j] += wgt * X[(size_t)j * n + r];
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < MB; ++j) {
      double s = acc[j];
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
      if (lane == 0) red[wv * MB + j] = s;
    }
    __syncthreads();
    if (threadIdx.x < MB) {
      double s = 0.0;
      for (int q = 0; q < (int)(blockDim.x >> 6); ++q) s += red[q * MB + threadIdx.x];
      nrm[threadIdx.x] = s < 0.0 ? -1.0 : 1.0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kw * n; e += blockDim.x) {
      const int j = e / n, r = e % n;
      phi[(size_t)r * a.kmax + c0 + j] = nrm[j] * X[(size_t)j * n + r];
    }
  }
  if (threadIdx.x == 0) {
    if (!well) a.eig[(size_t)sidx * a.kmax] = 0.0;
    for (int j = 0; j < kw; ++j) a.eig[(size_t)sidx * a.kmax + c0 + j] = lam[j];
    a.its[sidx] = it;
    a.status[sidx] = status;
    a.fallback[sidx] = a.x0 != nullptr && attempts > 0;
    a.stagnated[sidx] = status == 0 && s_stag;
  }
}

template <int MB>
static size_t lobpcg_lds(int maxn) {
  constexpr int LD = 3 * MB;
  const size_t small = (5 * LD * LD + LD + 3 * LD + kLobWaves * MB * MB + 64) * sizeof(double);
  const size_t prec = (size_t)maxn * kLobCh * sizeof(float);
  return std::max(small, prec);
}

template <int MB>
static void launch_lobpcg_geo(const Geo& g, int nblocks, hipStream_t st, const LobArgs& a,
                              int maxn) {
  if (maxn >= kLobBigN)
    hipLaunchKernelGGL((k_lobpcg<MB, 512>), dim3(nblocks), dim3(512), lobpcg_lds<MB>(maxn), st, g,
                       a);
  else if (maxn <= kLobSmallN)
    hipLaunchKernelGGL((k_lobpcg<MB, 128>), dim3(nblocks), dim3(128), lobpcg_lds<MB>(maxn), st, g,
                       a);
  else
    hipLaunchKernelGGL((k_lobpcg<MB, 256>), dim3(nblocks), dim3(256), lobpcg_lds<MB>(maxn), st, g,
                       a);
}

template <int MB>
static void launch_lobpcg_mb(pmx_ctx* ctx, const LobArgs& a, int maxn) {
  launch_lobpcg_geo<MB>(ctx->geo, (ctx->nsub + a.stride - 1) / a.stride, ctx->stream, a, maxn);
}

int lobpcg_block_size(int en_lv1, int guard) {
  const int mb = en_lv1 - 1 + guard;
  return mb <= 4 ? 4 : (mb <= 6 ? 6 : (mb <= 8 ? 8 : (mb <= 10 ? 10 : 0)));
}

// One batched LOBPCG launch with block size MB over every stride-th
// subdomain.  ms == nullptr: the real solve (results and statistics kept);
// otherwise an autotuner trial, only its duration (*ms) is returned.
static int lobpcg_run(pmx_ctx* ctx, const double* x0_rows, int MB, int stride, float* ms) {
  const int kw = ctx->en_lv1 - 1;
  if (MB == 0) {
    set_error("en_lv1=%d exceeds the LOBPCG block limit", ctx->en_lv1);
    return PMX_ERR_SUP;
  }
  int maxn = 0, minn = 1 << 30;
  for (const Brick& b : ctx->h_bricks) {
    maxn = std::max(maxn, b.nx * b.ny * b.nz);
    minn = std::min(minn, b.nx * b.ny * b.nz);
  }
  if (MB > minn - 1) {
    set_error("subdomains of %d cells are too small for %d eigenvectors", minn, ctx->en_lv1);
    return PMX_ERR_ARG_WRONG;
  }
  if (ctx->eig_kernel == 1 && ctx->eig_ortho == 0 && maxn <= kLobSmallN) {
    // The explicit-projection path in 128-thread workgroups took a GPU
    // memory fault on the batched workload (8^3 subdomains) whose cause is
    // not found; it stays switched off there.
    set_error("-eig_ortho 0 with subdomains of at most %d cells is not available", kLobSmallN);
    return PMX_ERR_SUP;
  }
  if (ctx->have_wd && MB < ctx->en_lv1 + 1) {
    // well subdomains iterate on all en_lv1 pairs and need one guard vector
    set_error("en_lv1=%d with a well term needs -eig_guard >= 2", ctx->en_lv1);
    return PMX_ERR_SUP;
  }
  // LOBPCG preconditioner: FP32 block ICC(0) of A^N + sigma M.
  DBuf<float> pf, pb;
  PMX_TRY(pf.alloc((size_t)ctx->ncell * 4));
  PMX_TRY(pb.alloc((size_t)ctx->ncell * 4));
  PMX_TRY(launch_dic_factor(ctx, 1, ctx->eig_sigma, pf.p, pb.p, true));
  const size_t blk = (size_t)ctx->ncell * MB;
  PMX_TRY(ctx->lob_buf.alloc(6 * blk));
  PMX_HIP(hipMemsetAsync(ctx->lob_buf.p, 0, 6 * blk * sizeof(double), ctx->stream));
  DBuf<double> eig;
  DBuf<int> its, status, fallback, stag;
  PMX_TRY(eig.alloc((size_t)ctx->nsub * ctx->en_lv1));
  PMX_TRY(its.alloc(ctx->nsub));
  PMX_TRY(status.alloc(ctx->nsub));
  PMX_TRY(fallback.alloc(ctx->nsub));
  PMX_TRY(stag.alloc(ctx->nsub));
  LobArgs a;
  a.stride = stride;
  a.kwant = kw;
  a.maxit = ctx->eig_maxit;
  a.kmax = ctx->en_lv1;
  a.tol = ctx->eig_tol;
  a.seed = ctx->seed;
  double* buf = ctx->lob_buf.p;
  a.X = buf;
  a.AX = buf + blk;
  a.W = buf + 2 * blk;
  a.AW = buf + 3 * blk;
  a.P = buf + 4 * blk;
  a.AP = buf + 5 * blk;
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
  a.fallback = fallback.p;
  a.stagnated = stag.p;
  a.stag_factor = ctx->eig_stag;
  a.debug = ctx->eig_debug;
  a.ortho_coef = ctx->eig_ortho;
  // Launch order: one 512-thread workgroup fills a CU, so subdomains beyond
  // the first wave of workgroups start late, and a hard one started late
  // sets the launch time.  A context that was set up before launches its
  // subdomains by decreasing iteration count of that setup (the IMPES loop
  // re-sets up every step on a slowly changing kappa); the first setup runs
  // in index order.  Workgroups are independent, so the results do not
  // depend on the order.
  a.order = nullptr;
  if (!ms && stride == 1 && ctx->eig_order && (int)ctx->h_lob_its.size() == ctx->nsub) {
    std::vector<int> ord(ctx->nsub);
    for (int s = 0; s < ctx->nsub; ++s) ord[s] = s;
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int x, int y) { return ctx->h_lob_its[x] > ctx->h_lob_its[y]; });
    PMX_TRY(ctx->d_lob_order.alloc(ctx->nsub));
    PMX_HIP(hipMemcpyAsync(ctx->d_lob_order.p, ord.data(), ctx->nsub * sizeof(int),
                           hipMemcpyHostToDevice, ctx->stream));
    a.order = ctx->d_lob_order.p;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // Autotuner trials, and the real launch under -kernel_timing, are timed
  // with HIP events on the library stream.
  const bool timed = ms != nullptr || ctx->kernel_timing;
  if (timed) {
    PMX_HIP(hipEventCreate(&e0));
    PMX_HIP(hipEventCreate(&e1));
    PMX_HIP(hipEventRecord(e0, ctx->stream));
  }
  if (ctx->eig_kernel == 1) {
    launch_lobpcg_mf(MB, ctx->eig_shape, ctx->geo, (ctx->nsub + a.stride - 1) / a.stride, ctx->stream, a, maxn);
  } else {
  switch (MB) {
    case 4:
      launch_lobpcg_mb<4>(ctx, a, maxn);
      break;
    case 6:
      launch_lobpcg_mb<6>(ctx, a, maxn);
      break;
    case 8:
      launch_lobpcg_mb<8>(ctx, a, maxn);
      break;
    default:
      launch_lobpcg_mb<10>(ctx, a, maxn);
      break;
  }
  }
  PMX_HIP(hipGetLastError());
  if (timed) {
    PMX_HIP(hipEventRecord(e1, ctx->stream));
    PMX_HIP(hipEventSynchronize(e1));
    float el = 0.0f;
    PMX_HIP(hipEventElapsedTime(&el, e0, e1));
    if (ms) *ms = el;
    else ctx->t_eig_kernel = 1e-3 * el;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  if (ms) return PMX_SUCCESS;
  ctx->eig_block = MB;
  std::vector<int> hits(ctx->nsub), hst(ctx->nsub), hfb(ctx->nsub), hsg(ctx->nsub);
  PMX_HIP(hipMemcpyAsync(hsg.data(), stag.p, ctx->nsub * sizeof(int), hipMemcpyDeviceToHost,
                         ctx->stream));
  PMX_HIP(hipMemcpyAsync(hfb.data(), fallback.p, ctx->nsub * sizeof(int), hipMemcpyDeviceToHost,
                         ctx->stream));
  ctx->h_eig_lv1.resize((size_t)ctx->nsub * ctx->en_lv1);
  PMX_HIP(hipMemcpyAsync(hits.data(), its.p, ctx->nsub * sizeof(int), hipMemcpyDeviceToHost,
                         ctx->stream));
  PMX_HIP(hipMemcpyAsync(hst.data(), status.p, ctx->nsub * sizeof(int), hipMemcpyDeviceToHost,
                         ctx->stream));
  PMX_HIP(hipMemcpyAsync(ctx->h_eig_lv1.data(), eig.p, ctx->h_eig_lv1.size() * sizeof(double),
                         hipMemcpyDeviceToHost, ctx->stream));
  PMX_HIP(hipStreamSynchronize(ctx->stream));
  ctx->h_lob_its = hits;
  int bad = 0, first_bad = -1;
  ctx->eig_iters_max = 0;
  ctx->eig_iters_sum = 0;
  ctx->eig_fallbacks = 0;
  ctx->eig_stagnated = 0;
  for (int s = 0; s < ctx->nsub; ++s) {
    ctx->eig_fallbacks += hfb[s];
    ctx->eig_stagnated += hsg[s];
  }
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

int lobpcg_setup(pmx_ctx* ctx, const double* x0_rows) {
  return lobpcg_run(ctx, x0_rows, lobpcg_block_size(ctx->en_lv1, ctx->eig_guard), 1, nullptr);
}

// Autotuner of the batched LOBPCG (-eig_autotune 1) over the block size and,
// for the matrix-core kernel, the workgroup shape.  The block size MB
// (wanted pairs + guard vectors) trades work per iteration against
// iterations; the shape trades waves per subdomain against subdomains per CU
// (one 512-thread workgroup per CU, or two 256-thread ones with narrower
// sweep chunks on large subdomains; every workgroup held to 256 registers on
// small ones).  The best pair depends on the subdomain size, the eigen count
// and the medium.  Every admissible (MB in {4, 6, 8, 10}, shape) is timed,
// except that a shape more than 10 % slower than the other at one block size
// is left out at the larger ones (on the workloads measured such a shape
// either stayed slower at the larger blocks or won them back by under 10 %
// where the smaller block was faster still), each trial starting
// from the start the setup will use (random, or the learned warm start) on
// the whole batch (a strided sample of 512 subdomains on larger problems,
// one full round of workgroups at two per CU, 1024 for small subdomains,
// whose 128-thread workgroups run more per CU: the launch is timed at the
// occupancy it runs at); the fastest sets
// -eig_guard and -eig_shape, cached per (largest subdomain, en_lv1, well
// term, warm start, kernel) for the process.  Tuning only changes how fast
// LOBPCG reaches the tolerance.
namespace {
constexpr int kTuneSample = 512;   // subdomains timed per trial on larger problems
                                   // (twice that for subdomains of <= kLobSmallN cells)
constexpr float kShapeMargin = 1.1f;   // a shape this much slower at one block size
                                       // is not timed at the larger ones
struct TuneKey {
  int maxn, k, well, warm, kernel;
  bool operator<(const TuneKey& o) const {
    if (maxn != o.maxn) return maxn < o.maxn;
    if (k != o.k) return k < o.k;
    if (well != o.well) return well < o.well;
    return warm != o.warm ? warm < o.warm : kernel < o.kernel;
  }
};
struct TuneChoice {
  int mb, shape;
};
std::map<TuneKey, TuneChoice>& tune_cache() {
  static std::map<TuneKey, TuneChoice> m;
  return m;
}
}  // namespace

int lobpcg_autotune(pmx_ctx* ctx) {
  int maxn = 0, minn = 1 << 30;
  for (const Brick& b : ctx->h_bricks) {
    maxn = std::max(maxn, b.nx * b.ny * b.nz);
    minn = std::min(minn, b.nx * b.ny * b.nz);
  }
  const int k = ctx->en_lv1;
  const bool warm = ctx->warm_start == 1;
  const TuneKey key{maxn, k, ctx->have_wd, warm ? 1 : 0, ctx->eig_kernel};
  auto it = tune_cache().find(key);
  if (it == tune_cache().end()) {
    const int need = ctx->have_wd ? k + 1 : k;   // at least one guard vector
    // small subdomains run more workgroups per CU: twice the sample
    const int sample = maxn <= kLobSmallN ? 2 * kTuneSample : kTuneSample;
    const int stride = std::max(1, (ctx->nsub + sample - 1) / sample);
    TuneChoice best{0, 0};
    float best_ms = 0.0f;
    const int user_shape = ctx->eig_shape;
    const int nshape = ctx->eig_kernel == 1 ? 2 : 1;
    int drop_shape = -1;   // a shape that lost one block size by more than kShapeMargin
    for (int mb : {4, 6, 8, 10}) {
      if (mb < need || mb > minn - 1) continue;
      // with the learned warm start the trials start from it, as the setup will
      DBuf<double> x0;
      if (warm) {
        PMX_TRY(x0.alloc((size_t)ctx->ncell * mb));
        PMX_TRY(cnn_warm_start(ctx, mb, x0.p));
      }
      float shape_ms[2] = {0.0f, 0.0f};
      for (int shape = 0; shape < nshape; ++shape) {
        if (shape == drop_shape) continue;
        float ms = 0.0f;
        ctx->eig_shape = shape;
        const int rc = lobpcg_run(ctx, warm ? x0.p : nullptr, mb, stride, &ms);
        ctx->eig_shape = user_shape;
        PMX_TRY(rc);
        shape_ms[shape] = ms;
        if (!best.mb || ms < best_ms) {
          best = TuneChoice{mb, shape};
          best_ms = ms;
        }
      }
      if (nshape == 2 && drop_shape < 0) {
        if (shape_ms[0] > kShapeMargin * shape_ms[1]) drop_shape = 0;
        else if (shape_ms[1] > kShapeMargin * shape_ms[0]) drop_shape = 1;
      }
    }
    if (!best.mb) return PMX_SUCCESS;   // nothing admissible: keep -eig_guard / -eig_shape
    it = tune_cache().emplace(key, best).first;
  }
  ctx->eig_guard = it->second.mb - (k - 1);
  if (ctx->eig_kernel == 1) ctx->eig_shape = it->second.shape;
  return PMX_SUCCESS;
}

// ---------------------------------------------------------------------------
// Batched setup of several independent problems: the subdomain
// eigenproblems of all of them in ONE persistent LOBPCG launch (one
// workgroup per subdomain of any problem).  The problems' bricks, plane
// tables and fields are concatenated into a batch geometry; a brick's random
// start is seeded by its index inside its own problem, so each problem gets
// bit for bit the basis its own pmx_setup would compute.  Requires equal
// en_lv1, eigensolver options and warm-start mode (none, or the learned
// start for every problem).
// ---------------------------------------------------------------------------
int lobpcg_batch(pmx_ctx** ctxs, int nprob) {
  pmx_ctx* c0 = ctxs[0];
  const int k = c0->en_lv1;
  const int MB = lobpcg_block_size(k, c0->eig_guard);
  if (MB == 0) {
    set_error("en_lv1=%d exceeds the LOBPCG block limit", k);
    return PMX_ERR_SUP;
  }
  std::vector<size_t> coff(nprob + 1, 0);
  std::vector<int> boff(nprob + 1, 0);
  bool any_wd = false, all_cnn = true;
  int maxn = 0, minn = 1 << 30;
  for (int i = 0; i < nprob; ++i) {
    pmx_ctx* c = ctxs[i];
    if (c->en_lv1 != k || c->eig_guard != c0->eig_guard || c->eig_tol != c0->eig_tol ||
        c->eig_maxit != c0->eig_maxit || c->eig_sigma != c0->eig_sigma ||
        c->seed != c0->seed || c->eig_stag != c0->eig_stag || c->eig_kernel != c0->eig_kernel ||
        c->eig_shape != c0->eig_shape || c->eig_ortho != c0->eig_ortho ||
        (c->warm_start == 1) != (c0->warm_start == 1)) {
      set_error("batched setup: problem %d has other en_lv1 / eigensolver options", i);
      return PMX_ERR_ARG_WRONG;
    }
    if (c->eig_kernel == 1 && c->eig_ortho == 0) {
      // see lobpcg_run: the explicit-projection path is off for the batch
      set_error("batched setup: -eig_ortho 0 is not available");
      return PMX_ERR_SUP;
    }
    coff[i + 1] = coff[i] + c->ncell;
    boff[i + 1] = boff[i] + c->nsub;
    any_wd = any_wd || c->have_wd;
    all_cnn = all_cnn && c->warm_start == 1;
    for (const Brick& b : c->h_bricks) {
      maxn = std::max(maxn, b.nx * b.ny * b.nz);
      minn = std::min(minn, b.nx * b.ny * b.nz);
    }
  }
  if (MB > minn - 1 || (any_wd && MB < k + 1)) {
    set_error("batched setup: block size %d does not fit the subdomains", MB);
    return PMX_ERR_ARG_WRONG;
  }
  const size_t N = coff[nprob];
  const int S = boff[nprob];
  // Batch geometry: bricks and views shifted to the concatenated cell space,
  // plane tables appended per problem.
  std::vector<Brick> hb;
  std::vector<BrickView> hv;
  std::vector<int> pstart, pcell, spstart, scoff;
  hb.reserve(S);
  hv.reserve(S);
  for (int i = 0; i < nprob; ++i) {
    pmx_ctx* c = ctxs[i];
    const int sh0 = (int)spstart.size();
    for (int q = 0; q < c->nshape; ++q) {
      spstart.push_back(c->h_shape_pstart[q] + (int)pstart.size());
      scoff.push_back(c->h_shape_coff[q] + (int)pcell.size());
    }
    pstart.insert(pstart.end(), c->h_plane_start.begin(), c->h_plane_start.end());
    pcell.insert(pcell.end(), c->h_plane_cell.begin(), c->h_plane_cell.end());
    for (const Brick& b0 : c->h_bricks) {
      Brick b = b0;
      b.off += (int)coff[i];
      b.shape += sh0;
      b.plane_off = scoff[b.shape];
      for (int d = 0; d < 6; ++d) b.nb[d] = -1;
      hb.push_back(b);
      BrickView v{};
      v.off = b.off;
      v.nx = b.nx;
      v.ny = b.ny;
      v.nz = b.nz;
      v.nxny = b.nx * b.ny;
      v.n = v.nxny * b.nz;
      for (int d = 0; d < 6; ++d) {
        v.noff[d] = -1;
        v.next[d] = 0;
      }
      hv.push_back(v);
    }
  }
  hipStream_t st = c0->stream;
  DBuf<Brick> d_b;
  DBuf<BrickView> d_v;
  DBuf<int> d_ps, d_pc, d_sps, d_sco;
  PMX_TRY(d_b.alloc(S));
  PMX_TRY(d_v.alloc(S));
  PMX_TRY(d_ps.alloc(pstart.size()));
  PMX_TRY(d_pc.alloc(pcell.size()));
  PMX_TRY(d_sps.alloc(spstart.size()));
  PMX_TRY(d_sco.alloc(scoff.size()));
  PMX_HIP(hipMemcpy(d_b.p, hb.data(), S * sizeof(Brick), hipMemcpyHostToDevice));
  PMX_HIP(hipMemcpy(d_v.p, hv.data(), S * sizeof(BrickView), hipMemcpyHostToDevice));
  PMX_HIP(hipMemcpy(d_ps.p, pstart.data(), pstart.size() * sizeof(int), hipMemcpyHostToDevice));
  PMX_HIP(hipMemcpy(d_pc.p, pcell.data(), pcell.size() * sizeof(int), hipMemcpyHostToDevice));
  PMX_HIP(hipMemcpy(d_sps.p, spstart.data(), spstart.size() * sizeof(int), hipMemcpyHostToDevice));
  PMX_HIP(hipMemcpy(d_sco.p, scoff.data(), scoff.size() * sizeof(int), hipMemcpyHostToDevice));
  Geo g = c0->geo;
  g.nsub = S;
  g.ncell = (int)N;
  g.bricks = d_b.p;
  g.views = d_v.p;
  g.order = nullptr;
  g.plane_start = d_ps.p;
  g.plane_cell = d_pc.p;
  g.shape_pstart = d_sps.p;
  g.shape_coff = d_sco.p;
  g.pin = -1;
  g.stop = nullptr;
  // Fields: face coefficients, eigen mass, well term, the FP32 LOBPCG
  // preconditioner factors (built per problem into its slice), warm start.
  DBuf<double> T[3], mass, wd, x0;
  DBuf<float> pf, pb;
  for (int a = 0; a < 3; ++a) PMX_TRY(T[a].alloc(N));
  PMX_TRY(mass.alloc(N));
  PMX_TRY(pf.alloc(N * 4));
  PMX_TRY(pb.alloc(N * 4));
  if (any_wd) {
    PMX_TRY(wd.alloc(N));
    PMX_HIP(hipMemsetAsync(wd.p, 0, N * sizeof(double), st));
  }
  if (all_cnn) PMX_TRY(x0.alloc(N * MB));
  for (int i = 0; i < nprob; ++i) {
    pmx_ctx* c = ctxs[i];
    const size_t o = coff[i], nb = (size_t)c->ncell * sizeof(double);
    PMX_TRY(launch_dic_factor(c, 1, c->eig_sigma, pf.p + o * 4, pb.p + o * 4, true));
    if (all_cnn) PMX_TRY(cnn_warm_start(c, MB, x0.p + o * MB));
    PMX_HIP(hipStreamSynchronize(c->stream));
    for (int a = 0; a < 3; ++a)
      PMX_HIP(hipMemcpyAsync(T[a].p + o, c->T[a].p, nb, hipMemcpyDeviceToDevice, st));
    PMX_HIP(hipMemcpyAsync(mass.p + o, c->mass.p, nb, hipMemcpyDeviceToDevice, st));
    if (c->have_wd)
      PMX_HIP(hipMemcpyAsync(wd.p + o, c->wd.p, nb, hipMemcpyDeviceToDevice, st));
  }
  g.wd = any_wd ? wd.p : nullptr;
  g.wd32 = nullptr;
  DBuf<double> lob, phi, eig;
  DBuf<int> its, status, fallback, stag;
  const size_t blk = N * MB;
  PMX_TRY(lob.alloc(6 * blk));
  PMX_HIP(hipMemsetAsync(lob.p, 0, 6 * blk * sizeof(double), st));
  PMX_TRY(phi.alloc(N * k));
  PMX_TRY(eig.alloc((size_t)S * k));
  PMX_TRY(its.alloc(S));
  PMX_TRY(status.alloc(S));
  PMX_TRY(fallback.alloc(S));
  PMX_TRY(stag.alloc(S));
  LobArgs a;
  a.stride = 1;
  a.kwant = k - 1;
  a.maxit = c0->eig_maxit;
  a.kmax = k;
  a.tol = c0->eig_tol;
  a.seed = c0->seed;
  a.X = lob.p;
  a.AX = lob.p + blk;
  a.W = lob.p + 2 * blk;
  a.AW = lob.p + 3 * blk;
  a.P = lob.p + 4 * blk;
  a.AP = lob.p + 5 * blk;
  a.pf = (const float4*)pf.p;
  a.pb = (const float4*)pb.p;
  a.Tx = T[0].p;
  a.Ty = T[1].p;
  a.Tz = T[2].p;
  a.mass = mass.p;
  a.x0 = all_cnn ? x0.p : nullptr;
  a.phi = phi.p;
  a.eig = eig.p;
  a.its = its.p;
  a.status = status.p;
  a.fallback = fallback.p;
  a.stagnated = stag.p;
  a.stag_factor = c0->eig_stag;
  a.debug = -1;
  a.order = nullptr;
  a.ortho_coef = c0->eig_ortho;
  if (c0->eig_kernel == 1) {
    launch_lobpcg_mf(MB, c0->eig_shape, g, S, st, a, maxn);
  } else {
  switch (MB) {
    case 4:
      launch_lobpcg_geo<4>(g, S, st, a, maxn);
      break;
    case 6:
      launch_lobpcg_geo<6>(g, S, st, a, maxn);
      break;
    case 8:
      launch_lobpcg_geo<8>(g, S, st, a, maxn);
      break;
    default:
      launch_lobpcg_geo<10>(g, S, st, a, maxn);
      break;
  }
  }
  PMX_HIP(hipGetLastError());
  std::vector<int> hits(S), hst(S), hfb(S), hsg(S);
  std::vector<double> heig((size_t)S * k);
  PMX_HIP(hipMemcpyAsync(hits.data(), its.p, S * sizeof(int), hipMemcpyDeviceToHost, st));
  PMX_HIP(hipMemcpyAsync(hst.data(), status.p, S * sizeof(int), hipMemcpyDeviceToHost, st));
  PMX_HIP(hipMemcpyAsync(hfb.data(), fallback.p, S * sizeof(int), hipMemcpyDeviceToHost, st));
  PMX_HIP(hipMemcpyAsync(hsg.data(), stag.p, S * sizeof(int), hipMemcpyDeviceToHost, st));
  PMX_HIP(hipMemcpyAsync(heig.data(), eig.p, heig.size() * sizeof(double), hipMemcpyDeviceToHost,
                         st));
  for (int i = 0; i < nprob; ++i) {
    pmx_ctx* c = ctxs[i];
    PMX_HIP(hipMemcpyAsync(c->phi.p, phi.p + coff[i] * k, (size_t)c->ncell * k * sizeof(double),
                           hipMemcpyDeviceToDevice, st));
  }
  PMX_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < nprob; ++i) {
    pmx_ctx* c = ctxs[i];
    c->eig_block = MB;
    c->h_eig_lv1.assign(heig.begin() + (size_t)boff[i] * k, heig.begin() + (size_t)boff[i + 1] * k);
    c->eig_iters_max = c->eig_iters_sum = c->eig_fallbacks = c->eig_stagnated = 0;
    for (int s = boff[i]; s < boff[i + 1]; ++s) {
      c->eig_iters_max = std::max(c->eig_iters_max, hits[s]);
      c->eig_iters_sum += hits[s];
      c->eig_fallbacks += hfb[s];
      c->eig_stagnated += hsg[s];
      if (hst[s] != 0) {
        set_error("batched setup: problem %d, subdomain %d: eigensolver status %d", i,
                  s - boff[i], hst[s]);
        return PMX_ERR_USER;
      }
    }
  }
  return PMX_SUCCESS;
}

}  // namespace pmx
