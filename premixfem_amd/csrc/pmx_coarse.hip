// Coarse levels: Galerkin assembly of Phi^T A Phi, its dense FP64 SPD
// inversion (block Gauss-Jordan sweep), the coarse-space operators of the
// V-cycle, and the level-3 eigenproblem / coarse-coarse matrix.
#include <algorithm>
#include <cmath>
#include <random>

#include "pmx_dense.h"
#include "pmx_kernels.h"
#include "pmx_mfma.h"

namespace pmx {

// ---------------------------------------------------------------------------
// Galerkin coarse matrix A_c = Phi^T A Phi (src/PreMixFEM_3D.c:737-890) in
// the reference's energy form.  Diagonal block of subdomain s: every face
// inside s adds T (phi_L - phi_R)(phi_L - phi_R)^T, every face to another
// subdomain (and the pin / well term) adds T phi phi^T; off-diagonal block
// with the upper neighbour s' across a face: -sum_face T phi_s,i phi_s',j.  One workgroup per subdomain, rows in chunks
// of 4 to bound registers; only the upper triangle of the diagonal block is
// reduced and mirrored, so A_c is exactly symmetric.  The 4 x kMaxEig sums
// of a chunk are reduced together: a register reduce-scatter across the wave
// leaves lane L with the wave's total of sum L, one LDS exchange adds the
// waves -- one barrier pair per chunk instead of one per entry.
// Output: dense nc x nc (row-major) and block rows Acb[s][slot][kmax][kmax]
// with slot 0 = self, 1..6 = -x,+x,-y,+y,-z,+z.
// ---------------------------------------------------------------------------
//
// Process boxes (_PC_setup_lv2_J, :806-888): the dense A_c, which level 2
// inverts, keeps only the couplings between subdomains of one box, while the
// block rows Acb keep all of them -- they are Phi^T A Phi, which the coarse
// residual updates between stages stand for.  inner = 1 builds the box-local
// Neumann matrix A_i_inner of level 3 (:939-1073) into Ac instead: faces
// between boxes are dropped from the diagonal blocks too, and Acb is not
// written (pass nullptr).  inner = 2 keeps every coupling in Ac as well: the
// 2-level variant's coarse matrix, assembled over all processes with the
// neighbours' dof offsets (_PC_setup_lv2_c, :423-735).
// One stage of the wave reduce-scatter: N values per lane -> N / 2, lane
// bit N / 2 picks the upper or lower half to keep, the partner across that
// bit supplies its copy of the kept half.
template <int N>
__device__ __forceinline__ void rs_stage(double* v, int lane) {
  constexpr int H = N / 2;
  const bool hi = (lane & H) != 0;
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const double send = hi ? v[k] : v[k + H];
    const double keep = hi ? v[k + H] : v[k];
    v[k] = keep + __shfl_xor(send, H, 64);
  }
}

// Sums acc[4][kMaxEig] (64 values per thread) over the workgroup; on return
// threads 0..63 hold the total of entry threadIdx.x = a * kMaxEig + j.
// red holds (blockDim.x / 64) x 64 doubles.  Ends with a barrier, so red can
// be reused at once.
__device__ __forceinline__ double chunk_sum(double (&acc)[4][kMaxEig], double* red) {
  static_assert(4 * kMaxEig == 64, "one accumulator per lane");
  double* v = &acc[0][0];
  const int lane = threadIdx.x & 63;
  rs_stage<64>(v, lane);
  rs_stage<32>(v, lane);
  rs_stage<16>(v, lane);
  rs_stage<8>(v, lane);
  rs_stage<4>(v, lane);
  rs_stage<2>(v, lane);
  red[threadIdx.x] = v[0];
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x < 64)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += red[w * 64 + threadIdx.x];
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(kBlock) k_galerkin(Geo g, const double* __restrict__ phi, int kmax,
                           const double* __restrict__ Tx, const double* __restrict__ Ty,
                           const double* __restrict__ Tz, int nc, double* __restrict__ Ac,
                           double* __restrict__ Acb, int inner) {
  __shared__ double red[kBlock];
  const int s = blockIdx.x;
  const Brick b = g.bricks[s];
  const int n = b.nx * b.ny * b.nz;
  const int ke = b.neig;
  const double* TT3[3] = {Tx, Ty, Tz};
  double* blk = Acb ? Acb + (size_t)s * 7 * kmax * kmax : nullptr;
  // Unrolled so that the row offset i0 is a constant.  The basis rows are
  // read again for every rank-1 term (L1 hits) rather than held, which keeps
  // the 64 accumulators and a few values in registers.
#pragma unroll
  for (int ic = 0; ic < kMaxEig / 4; ++ic) {
    const int i0 = 4 * ic;
    if (i0 >= ke) break;
    double acc[4][kMaxEig];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int j = 0; j < kMaxEig; ++j) acc[a][j] = 0.0;
    // acc[a][j] += w u_(i0+a) u_j with u = pa - pb (pb may be null).
    auto rank1 = [&](double w, const double* pa, const double* pb) {
      double ui[4];
#pragma unroll
      for (int a = 0; a < 4; ++a)
        ui[a] = i0 + a < ke ? w * (pa[i0 + a] - (pb ? pb[i0 + a] : 0.0)) : 0.0;
#pragma unroll
      for (int j = 0; j < kMaxEig; ++j) {
        if (j < ke) {
          const double uj = pa[j] - (pb ? pb[j] : 0.0);
#pragma unroll
          for (int a = 0; a < 4; ++a) acc[a][j] += ui[a] * uj;
        }
      }
    };
    for (int li = threadIdx.x; li < n; li += blockDim.x) {
      const int lx = li % b.nx, ly = (li / b.nx) % b.ny, lz = li / (b.nx * b.ny);
      const int c = b.off + li;
      const double* pc = phi + (size_t)c * kmax;
      // Faces to other subdomains, the pin and the well term act on the
      // diagonal block as d phi phi^T.
      double sdiag = 0.0;
#pragma unroll
      for (int d = 0; d < 6; ++d) {
        const int nbi = nbr_index(g, b, c, lx, ly, lz, d);
        if (nbi < 0) continue;
        const double t = (d & 1) ? TT3[d >> 1][nbi] : TT3[d >> 1][c];
        const int lcoord = (d >> 1) == 0 ? lx : ((d >> 1) == 1 ? ly : lz);
        const int lext = (d >> 1) == 0 ? b.nx : ((d >> 1) == 1 ? b.ny : b.nz);
        const bool inside = (d & 1) ? (lcoord < lext - 1) : (lcoord > 0);
        if (!inside) {
          if (inner != 1 || g.bricks[b.nb[d]].box == b.box) sdiag += t;
        } else if ((d & 1) == 0) {
          // Interior face with the lower neighbour: T (phi_c - phi_nb)(...)^T
          // from the differences of the basis values, as the reference
          // assembles it -- no cancellation near the null space.
          rank1(t, pc, phi + (size_t)nbi * kmax);
        }
      }
      if (c == g.pin) sdiag += g.tau;
      if (g.wd) sdiag += g.wd[c];
      rank1(sdiag, pc, nullptr);
    }
    const double v = chunk_sum(acc, red);
    const int i = i0 + (int)(threadIdx.x / kMaxEig), j = threadIdx.x % kMaxEig;
    if (threadIdx.x < 64 && i < ke && j >= i && j < ke) {
      if (Ac) {
        Ac[(size_t)(b.dof + i) * nc + b.dof + j] = v;
        Ac[(size_t)(b.dof + j) * nc + b.dof + i] = v;
      }
      if (Acb) {
        blk[i * kmax + j] = v;
        blk[j * kmax + i] = v;
      }
    }
  }
  // Faces to the upper neighbours.
  for (int dir = 1; dir < 6; dir += 2) {
    const int sn = g.bricks[s].nb[dir];  // from memory: b stays in registers
    if (sn < 0) continue;
    const Brick o = g.bricks[sn];
    const int axis = dir >> 1;
    const int nf = axis == 0 ? b.ny * b.nz : (axis == 1 ? b.nx * b.nz : b.nx * b.ny);
    for (int i0 = 0; i0 < ke; i0 += 4) {
      double acc[4][kMaxEig];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int j = 0; j < kMaxEig; ++j) acc[a][j] = 0.0;
      for (int f = threadIdx.x; f < nf; f += blockDim.x) {
        int lx, ly, lz;
        if (axis == 0) {
          lx = b.nx - 1;
          ly = f % b.ny;
          lz = f / b.ny;
        } else if (axis == 1) {
          ly = b.ny - 1;
          lx = f % b.nx;
          lz = f / b.nx;
        } else {
          lz = b.nz - 1;
          lx = f % b.nx;
          ly = f / b.nx;
        }
        const int c = b.off + (lz * b.ny + ly) * b.nx + lx;
        const int nbi = nbr_index(g, b, c, lx, ly, lz, dir);
        const double t = TT3[axis][nbi];
        const double* pc = phi + (size_t)c * kmax;
        const double* pn = phi + (size_t)nbi * kmax;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          if (i0 + a >= ke) break;
          const double pi = t * pc[i0 + a];
#pragma unroll
          for (int j = 0; j < kMaxEig; ++j)
            if (j < o.neig) acc[a][j] -= pi * pn[j];
        }
      }
      const double v = chunk_sum(acc, red);
      const int i = i0 + (int)(threadIdx.x / kMaxEig), j = threadIdx.x % kMaxEig;
      if (threadIdx.x < 64 && i < ke && j < o.neig) {
        if (Ac && (o.box == b.box || inner == 2)) {
          Ac[(size_t)(b.dof + i) * nc + o.dof + j] = v;
          Ac[(size_t)(o.dof + j) * nc + b.dof + i] = v;
        }
        if (Acb) {
          blk[(size_t)(dir + 1) * kmax * kmax + i * kmax + j] = v;
          Acb[((size_t)sn * 7 + dir) * kmax * kmax + j * kmax + i] = v;  // slot of -dir
        }
      }
    }
  }
}

int launch_galerkin(pmx_ctx* ctx) {
  if (ctx->Ac.p)
    PMX_HIP(hipMemsetAsync(ctx->Ac.p, 0, sizeof(double) * (size_t)ctx->nc * ctx->nc, ctx->stream));
  PMX_HIP(hipMemsetAsync(ctx->Acb.p, 0, sizeof(double) * ctx->Acb.n, ctx->stream));
  hipLaunchKernelGGL(k_galerkin, dim3(ctx->nsub), dim3(kBlock), 0, ctx->stream, ctx->geo,
                     ctx->phi.p, ctx->en_lv1, ctx->T[0].p, ctx->T[1].p, ctx->T[2].p, ctx->nc,
                     ctx->Ac.p, ctx->Acb.p, ctx->use_2level ? 2 : 0);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

// ---------------------------------------------------------------------------
// Dense SPD inversion by the block sweep operator: for each SW-wide pivot
// block k with D = A_kk^{-1}:  A_ij -= A_ik D A_kj (i,j != k),
// A_ik <- A_ik D, A_kj <- D A_kj, A_kk <- -D.  After all blocks A = -A^{-1}.
// Pivots are Schur complements of an SPD matrix, so no pivoting is needed.
// The pivot width trades the one-workgroup pivot inversions (SW steps of
// one barrier each, every thread reading a row segment of SW / 4 values per
// step) against the passes of the trailing update over the matrix (n / SW):
// 64 up to n = kSwWideN, 128 above.
// ---------------------------------------------------------------------------
constexpr int kSwT = 64;          // trailing-update tile (pivot blocks are whole tiles)
constexpr int kSwWideN = 4096;    // 128-wide pivot blocks above this size

// Gauss-Jordan inverse of the bs x bs pivot block (identity-padded to SW),
// read from src (leading dimension ld): 4 SW threads, thread t holds row
// t/4, columns (SW/4)(t%4) .. in registers; one barrier per elimination step
// (the row / column broadcast buffers alternate).  The unsymmetrised inverse
// goes through Dt (global scratch) to be symmetrised into D.
template <int SW>
__device__ __forceinline__ void sweep_pivot_block(const double* src, int ld, int bs,
                                                  double* __restrict__ D,
                                                  double* __restrict__ Dt) {
  constexpr int CW = SW / 4;   // columns per thread
  __shared__ double2 rowv2[2][SW / 2];
  __shared__ double colv[2][SW];
  const int r = threadIdx.x >> 2, q = threadIdx.x & 3;
  double m[CW];
#pragma unroll
  for (int c = 0; c < CW; ++c) {
    const int j = q * CW + c;
    m[c] = (r < bs && j < bs) ? src[(size_t)r * ld + j] : (r == j ? 1.0 : 0.0);
  }
  for (int tq = 0; tq < 4; ++tq) {
#pragma unroll
    for (int tc = 0; tc < CW; ++tc) {
      const int t = tq * CW + tc, par = t & 1;
      if (q == tq) colv[par][r] = m[tc];
      if (r == t) {
#pragma unroll
        for (int c = 0; c < CW; c += 2) rowv2[par][q * (CW / 2) + c / 2] = double2{m[c], m[c + 1]};
      }
      __syncthreads();
      const double* rowv = reinterpret_cast<const double*>(rowv2[par]);
      const double piv = 1.0 / rowv[t];
      const double f = colv[par][r];
#pragma unroll
      for (int c = 0; c < CW; c += 2) {
        const double2 rv = rowv2[par][q * (CW / 2) + c / 2];
        const int j = q * CW + c;
        const double rt0 = (j == t ? 1.0 : rv.x) * piv;
        const double rt1 = (j + 1 == t ? 1.0 : rv.y) * piv;
        m[c] = r == t ? rt0 : (j == t ? 0.0 : m[c]) - f * rt0;
        m[c + 1] = r == t ? rt1 : (j + 1 == t ? 0.0 : m[c + 1]) - f * rt1;
      }
    }
  }
#pragma unroll
  for (int c = 0; c < CW; ++c) Dt[r * SW + q * CW + c] = m[c];
  __syncthreads();
  for (int e = threadIdx.x; e < SW * SW; e += blockDim.x) {
    const int i = e / SW, j = e % SW;
    // Symmetrise the pivot inverse.
    D[e] = (i < bs && j < bs) ? 0.5 * (Dt[i * SW + j] + Dt[j * SW + i]) : 0.0;
  }
}

template <int SW>
__global__ void __launch_bounds__(4 * SW) k_sweep_pivot(const double* __restrict__ A, int n,
                                                        int k0, int bs,
                                                        double* __restrict__ D,
                                                        double* __restrict__ Dt) {
  sweep_pivot_block<SW>(A + (size_t)k0 * n + k0, n, bs, D, Dt);
}

// P[r][t] = A[r][k0+t] for t < bs, 0 beyond.
template <int SW>
__global__ void k_sweep_panel(const double* __restrict__ A, int n, int k0, int bs,
                              double* __restrict__ P) {
  const size_t tot = (size_t)n * SW;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(e / SW), t = (int)(e % SW);
    P[e] = t < bs ? A[(size_t)r * n + k0 + t] : 0.0;
  }
}

// Trailing update A_ij -= U_i P_j^T on 64x64 tiles (FP64 MFMA, K = bs); the
// tiles of the pivot block rows / columns take U, U^T and -D.
// With SW = kSwT the workgroup of the next pivot block's tile (knext, a
// tile of its own) inverts that block right after updating it, into Dn: the
// next step's pivot then costs no launch and overlaps this update.
template <int SW>
__global__ void __launch_bounds__(256) k_sweep_update(double* __restrict__ A, int n, int k0,
                                                      int bs, const double* __restrict__ P,
                                                      const double* __restrict__ U,
                                                      const double* __restrict__ D, int knext,
                                                      double* __restrict__ Dn,
                                                      double* __restrict__ Dt) {
  using TL = MfmaTile<2, 2, 2, 2>;
  static_assert(TL::TM == kSwT && TL::TN == kSwT && SW % kSwT == 0,
                "pivot blocks are whole tiles");
  __shared__ double lds[TL::LDS_DOUBLES];
  const int I0 = blockIdx.y * kSwT, J0 = blockIdx.x * kSwT;
  const bool prow = I0 >= k0 && I0 < k0 + SW, pcol = J0 >= k0 && J0 < k0 + SW;
  if (prow || pcol) {
    for (int e = threadIdx.x; e < kSwT * kSwT; e += blockDim.x) {
      const int i = I0 + e / kSwT, j = J0 + e % kSwT;
      if (i >= n || j >= n) continue;
      double v;
      if (prow && pcol)
        v = -D[(i - k0) * SW + (j - k0)];
      else if (prow)
        v = U[(size_t)j * SW + (i - k0)];
      else
        v = U[(size_t)i * SW + (j - k0)];
      A[(size_t)i * n + j] = v;
    }
    return;
  }
  mfma_d4 acc[2][2];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) acc[rb][cb] = mfma_d4{0.0, 0.0, 0.0, 0.0};
  // op(A) = U (rows I0..), op(B) = P^T (P row j = column j of op(B)).
  mfma_tile_acc<2, 2, 2, 2, true>(U, SW, P, SW, n, n, bs, I0, J0, lds, acc);
#pragma unroll
  for (int rb = 0; rb < 2; ++rb)
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int r, c;
        mfma_tile_pos<2, 2, 2, 2>(rb, cb, v, r, c);
        const int i = I0 + r, j = J0 + c;
        if (i < n && j < n) A[(size_t)i * n + j] -= acc[rb][cb][v];
      }
  if constexpr (SW == kSwT) {
    if (knext < n && I0 == knext && J0 == knext) {
      // the workgroup reads back its own stores (the barrier's workgroup
      // fences order them; the workgroup's waves share one L1)
      __syncthreads();
      sweep_pivot_block<SW>(A + (size_t)knext * n + knext, n, min(SW, n - knext), Dn, Dt);
    }
  }
}

__global__ void k_copy_shift(const double* __restrict__ A, double* __restrict__ B, int n,
                             double shift, double scale) {
  const size_t nn = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn;
       e += (size_t)gridDim.x * blockDim.x) {
    double v = A[e] * scale;
    if (e / n == e % n) v += shift;
    B[e] = v;
  }
}

int coarse_invert(pmx_ctx* ctx, const double* A, double* Ainv, double shift) {
  return spd_invert(ctx->stream, ctx->nc, A, Ainv, shift);
}

template <int SW>
static int spd_invert_sw(hipStream_t stream, int n, const double* A, double* Ainv,
                         double shift) {
  DBuf<double> D, Dt, P, U;
  PMX_TRY(D.alloc(2 * SW * SW));   // two pivot inverses: this step's and the next
  PMX_TRY(Dt.alloc(SW * SW));
  PMX_TRY(P.alloc((size_t)n * SW));
  PMX_TRY(U.alloc((size_t)n * SW));
  const int nblk = 1024;
  hipLaunchKernelGGL(k_copy_shift, dim3(nblk), dim3(kBlock), 0, stream, A, Ainv, n, shift,
                     1.0);
  PMX_HIP(hipGetLastError());
  const int nt = (n + kSwT - 1) / kSwT;
  const int npb = std::min((n * SW + 255) / 256, 2048);
  // With 64-wide pivots the update of step k inverts the pivot block of step
  // k + 1 (k_sweep_update); only the first one is a launch of its own.
  constexpr bool fused = SW == kSwT;
  for (int k0 = 0, kk = 0; k0 < n; k0 += SW, ++kk) {
    const int bs = std::min(SW, n - k0);
    double* Dk = D.p + (size_t)(kk & 1) * SW * SW;
    double* Dnext = D.p + (size_t)((kk + 1) & 1) * SW * SW;
    if (!fused || kk == 0)
      hipLaunchKernelGGL(k_sweep_pivot<SW>, dim3(1), dim3(4 * SW), 0, stream, Ainv, n, k0, bs,
                         Dk, Dt.p);
    hipLaunchKernelGGL(k_sweep_panel<SW>, dim3(npb), dim3(256), 0, stream, Ainv, n, k0, bs,
                       P.p);
    // U = P D (the pivot columns times the pivot inverse; one 128-column tile)
    hipLaunchKernelGGL((k_gemm_mfma<2, 2, 2, 4, false>), dim3(1, (n + 63) / 64), dim3(256), 0,
                       stream, n, SW, SW, 1.0, P.p, SW, Dk, SW, 0.0, U.p, SW);
    hipLaunchKernelGGL(k_sweep_update<SW>, dim3(nt, nt), dim3(256), 0, stream, Ainv, n, k0, bs,
                       P.p, U.p, Dk, fused ? k0 + SW : n, Dnext, Dt.p);
    PMX_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_copy_shift, dim3(nblk), dim3(kBlock), 0, stream, Ainv, Ainv, n, 0.0,
                     -1.0);
  PMX_HIP(hipGetLastError());
  PMX_HIP(hipStreamSynchronize(stream));
  return PMX_SUCCESS;
}

int spd_invert(hipStream_t stream, int n, const double* A, double* Ainv, double shift) {
  return n <= kSwWideN ? spd_invert_sw<64>(stream, n, A, Ainv, shift)
                       : spd_invert_sw<128>(stream, n, A, Ainv, shift);
}

// ---------------------------------------------------------------------------
// Level-2 coarse solve with the near-null direction handled exactly.
//
// A_c = K + tau g g^T with K = Phi^T A_N Phi + delta I, where A_N is the
// pure-Neumann operator (A_N 1 = 0), so K's Neumann part has the exact null
// vector z = Phi^T M 1 / ||.|| (the coarse coordinates of the constant), g the
// pin cell's basis row (tau = 0 without pin) and delta the NIL shift of the
// 2-level variant.  One pinned cell makes A_c nearly singular (smallest
// eigenvalue ~1e-11 against ~1e-1 for the rest), which an explicit inverse
// cannot resolve; instead B = K + gamma z z^T (gamma = trace/n) is inverted
// (well conditioned) and, with W = [z g], S = diag(-gamma, tau), Woodbury
//   A_c^{-1} t = y - [z/(delta+gamma), h] Cap^{-1} [z.t/(delta+gamma), h.t],
//   y = B^{-1} t, h = B^{-1} g, Cap = S^{-1} + W^T B^{-1} W,
// with B^{-1} z = z/(delta+gamma) in closed form.  Without pin and shift the
// same formula with Cap -> projection gives the pseudo-inverse on z-perp.
// cs = [dg, a00, a01, a11]: d = y - [z/dg, h] [[a00,a01],[a01,a11]] [z.t/dg, h.t].
// ---------------------------------------------------------------------------
__global__ void k_coarse_fix(const double* __restrict__ y, const double* __restrict__ t,
                             const double* __restrict__ z, const double* __restrict__ h, int n,
                             const double* __restrict__ cs, double* __restrict__ d) {
  __shared__ double red[8];
  const int c = blockIdx.x;  // column
  const double* yc = y + (size_t)c * n;
  const double* tc = t + (size_t)c * n;
  double* dc = d + (size_t)c * n;
  double zt = 0.0, ht = 0.0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    zt += z[r] * tc[r];
    ht += h[r] * tc[r];
  }
  zt = block_sum(zt, red);
  ht = block_sum(ht, red);
  const double dg = cs[0];
  const double u0 = zt / dg;
  const double a0 = cs[1] * u0 + cs[2] * ht;
  const double a1 = cs[2] * u0 + cs[3] * ht;
  for (int r = threadIdx.x; r < n; r += blockDim.x) dc[r] = yc[r] - (z[r] / dg) * a0 - h[r] * a1;
}

// B = (A_c - tau g g^T) + delta I + gamma z z^T (dense, row-major).
__global__ void k_coarse_B(const double* __restrict__ A, double* __restrict__ B, int n,
                           const double* __restrict__ z, const double* __restrict__ g, double tau,
                           double delta, double gamma) {
  const size_t nn = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n), j = (int)(e % n);
    double v = A[e] - tau * g[i] * g[j] + gamma * z[i] * z[j];
    if (i == j) v += delta;
    B[e] = v;
  }
}

__global__ void k_trace(const double* __restrict__ A, int n, double* __restrict__ out) {
  __shared__ double red[8];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += A[(size_t)i * n + i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[0] = s;
}

int coarse_solve_setup(pmx_ctx* ctx, double delta) {
  const int n = ctx->nc;
  PMX_TRY(ctx->c_z.alloc(n));
  PMX_TRY(ctx->c_g.alloc(n));
  PMX_TRY(ctx->c_h.alloc(n));
  PMX_TRY(ctx->c_cs.alloc(4));
  if (!ctx->coarse_deflate) {
    // Plain inverse of A_c (+ delta I), no deflation: the Woodbury fix is the
    // identity (z = h = 0).
    PMX_HIP(hipMemsetAsync(ctx->c_z.p, 0, n * sizeof(double), ctx->stream));
    PMX_HIP(hipMemsetAsync(ctx->c_g.p, 0, n * sizeof(double), ctx->stream));
    PMX_HIP(hipMemsetAsync(ctx->c_h.p, 0, n * sizeof(double), ctx->stream));
    const double cs[4] = {1.0, 0.0, 0.0, 0.0};
    PMX_HIP(hipMemcpy(ctx->c_cs.p, cs, sizeof(cs), hipMemcpyHostToDevice));
    PMX_TRY(coarse_invert(ctx, ctx->Ac.p, ctx->Ainv.p, delta));
    ctx->coarse_gamma = 0.0;
    return PMX_SUCCESS;
  }
  // z = Phi^T M 1, normalised.
  PMX_TRY(launch_resid_restrict_f64(ctx, ctx->mass.p, ctx->c_z.p));
  std::vector<double> hz(n), hg(n, 0.0);
  PMX_HIP(hipMemcpy(hz.data(), ctx->c_z.p, n * sizeof(double), hipMemcpyDeviceToHost));
  double nz = 0.0;
  for (double v : hz) nz += v * v;
  nz = std::sqrt(nz);
  for (double& v : hz) v /= nz;
  PMX_HIP(hipMemcpy(ctx->c_z.p, hz.data(), n * sizeof(double), hipMemcpyHostToDevice));
  // g = basis row of the pin cell.
  double tau = 0.0;
  if (ctx->pin_brick >= 0) {
    tau = ctx->tau;
    int s = 0;
    while (s + 1 < ctx->nsub && ctx->h_bricks[s + 1].off <= ctx->pin_brick) ++s;
    const Brick& b = ctx->h_bricks[s];
    std::vector<double> row(ctx->en_lv1);
    PMX_HIP(hipMemcpy(row.data(), ctx->phi.p + (size_t)ctx->pin_brick * ctx->en_lv1,
                      ctx->en_lv1 * sizeof(double), hipMemcpyDeviceToHost));
    for (int j = 0; j < b.neig; ++j) hg[b.dof + j] = row[j];
  }
  PMX_HIP(hipMemcpy(ctx->c_g.p, hg.data(), n * sizeof(double), hipMemcpyHostToDevice));
  double tr = 0.0;
  DBuf<double> tmp;
  PMX_TRY(tmp.alloc(1));
  hipLaunchKernelGGL(k_trace, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->Ac.p, n, tmp.p);
  PMX_HIP(hipGetLastError());
  PMX_HIP(hipMemcpy(&tr, tmp.p, sizeof(double), hipMemcpyDeviceToHost));
  const double gamma = tr / n;
  // B in the Ainv buffer, inverted in place.
  hipLaunchKernelGGL(k_coarse_B, dim3(1024), dim3(kBlock), 0, ctx->stream, ctx->Ac.p, ctx->Ainv.p,
                     n, ctx->c_z.p, ctx->c_g.p, tau, delta, gamma);
  PMX_HIP(hipGetLastError());
  PMX_TRY(coarse_invert(ctx, ctx->Ainv.p, ctx->Ainv.p, 0.0));
  PMX_TRY(launch_gemv(ctx, ctx->Ainv.p, ctx->c_g.p, ctx->c_h.p, n));
  std::vector<double> hh(n);
  PMX_HIP(hipMemcpy(hh.data(), ctx->c_h.p, n * sizeof(double), hipMemcpyDeviceToHost));
  double gz = 0.0, gh = 0.0;
  for (int i = 0; i < n; ++i) {
    gz += hg[i] * hz[i];
    gh += hg[i] * hh[i];
  }
  const double dg = delta + gamma;
  double cs[4] = {dg, 0.0, 0.0, 0.0};
  if (tau > 0.0) {
    const double c00 = -delta / (gamma * dg), c01 = gz / dg, c11 = 1.0 / tau + gh;
    const double det = c00 * c11 - c01 * c01;
    if (!(std::fabs(det) > 0.0)) {
      set_error("coarse solve: singular capacitance matrix (pin outside the coarse space?)");
      return PMX_ERR_USER;
    }
    cs[1] = c11 / det;
    cs[2] = -c01 / det;
    cs[3] = c00 / det;
  } else if (delta > 0.0) {
    cs[1] = 1.0 / (-delta / (gamma * dg));
  } else {
    cs[1] = dg;  // pseudo-inverse: remove the z component
  }
  PMX_HIP(hipMemcpy(ctx->c_cs.p, cs, sizeof(cs), hipMemcpyHostToDevice));
  ctx->coarse_gamma = gamma;
  return PMX_SUCCESS;
}

int launch_coarse_fix(pmx_ctx* ctx, const double* y, const double* t, double* d) {
  hipLaunchKernelGGL(k_coarse_fix, dim3(1), dim3(kBlock), 0, ctx->stream, y, t, ctx->c_z.p,
                     ctx->c_h.p, ctx->nc, ctx->c_cs.p, d);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

int launch_coarse_solve(pmx_ctx* ctx, const double* t, double* d, double* ytmp, int ncol) {
  const int n = ctx->nc;
  for (int c = 0; c < ncol; ++c) PMX_TRY(launch_gemv(ctx, ctx->Ainv.p, t + (size_t)c * n, ytmp + (size_t)c * n, n));
  hipLaunchKernelGGL(k_coarse_fix, dim3(ncol), dim3(kBlock), 0, ctx->stream, ytmp, t, ctx->c_z.p,
                     ctx->c_h.p, n, ctx->c_cs.p, d);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

// ---------------------------------------------------------------------------
// Coarse-space operators
// ---------------------------------------------------------------------------
// y = A x for a dense row-major n x n matrix.  One workgroup per row: every
// lane streams an independent slice of the row with 16-byte loads (many loads
// in flight per lane), then a fixed-order block reduction.  x is re-read by
// every row and stays in L2.
constexpr int kGemvBlock = 256;

__global__ void __launch_bounds__(kGemvBlock) k_gemv(const double* __restrict__ A,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ y, int n,
                                                     const double* stop) {
  PMX_STOP_CHECK(stop);
  __shared__ double red[16];
  const int row = blockIdx.x;
  const double* a = A + (size_t)row * n;
  double s0 = 0.0, s1 = 0.0;
  if ((n & 1) == 0) {
    // rows start 16-byte aligned when n is even
    const double2* a2 = reinterpret_cast<const double2*>(a);
    const double2* x2 = reinterpret_cast<const double2*>(x);
    const int n2 = n >> 1;
#pragma unroll 4
    for (int j = threadIdx.x; j < n2; j += kGemvBlock) {
      const double2 av = a2[j], xv = x2[j];
      s0 += av.x * xv.x;
      s1 += av.y * xv.y;
    }
  } else {
#pragma unroll 4
    for (int j = threadIdx.x; j < n; j += kGemvBlock) s0 += a[j] * x[j];
  }
  const double s = block_sum(s0 + s1, red);
  if (threadIdx.x == 0) y[row] = s;
}

int launch_gemv(pmx_ctx* ctx, const double* A, const double* x, double* y, int n) {
  hipLaunchKernelGGL(k_gemv, dim3(n), dim3(kGemvBlock), 0, ctx->stream, A, x, y, n,
                     ctx->geo.stop);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

// Row r of A_c d from the block rows.
__device__ __forceinline__ double coarse_row(const Geo& g, const double* __restrict__ Acb, int kmax,
                                             int s, int i, const double* __restrict__ d) {
  const Brick& b = g.bricks[s];
  const double* blk = Acb + (size_t)s * 7 * kmax * kmax;
  double acc = 0.0;
  for (int j = 0; j < b.neig; ++j) acc += blk[i * kmax + j] * d[b.dof + j];
  for (int dir = 0; dir < 6; ++dir) {
    const int sn = b.nb[dir];
    if (sn < 0) continue;
    const Brick& o = g.bricks[sn];
    const double* bb = blk + (size_t)(dir + 1) * kmax * kmax;
    for (int j = 0; j < o.neig; ++j) acc += bb[i * kmax + j] * d[o.dof + j];
  }
  return acc;
}

// ---------------------------------------------------------------------------
// Level 2 of the spe10 fork (src/spe10.c:509-516, :1072-1213): PETSc ICC(0)
// of A_c in natural order.  A_c's pattern is block 7-point over the
// subdomains with every block stored whole, so scalar IC(0) on it is the
// block incomplete Cholesky
//   D_s = A_ss - sum_{t in {s-x, s-y, s-z}} A_st D_t^{-1} A_ts,
//   M = (D + L) D^{-1} (D + L^T),  L = block strictly lower part of A_c:
// between s and a lower neighbour t the only common lower neighbour is t, so
// L_st = A_st L_tt^{-T} is exact, and the dropped fill is the one between s
// and the other neighbours of t.  Subdomains depend on their -x, -y, -z
// neighbours only, so the factorisation and both triangular sweeps run by
// hyperplanes sx+sy+sz = p inside one workgroup.
// ---------------------------------------------------------------------------
constexpr int kBiccWaves = 8;

// Factorisation: one wavefront per subdomain of a hyperplane; the waves of
// the workgroup step through the plane in uniform rounds.  dinv[s] holds
// D_s^{-1} (k x k, row stride kmax).  A non-positive pivot sets *fail to the
// subdomain index + 1.
__global__ void __launch_bounds__(64 * kBiccWaves) k_bicc_factor(
    Geo g, const double* __restrict__ Acb, int kmax, const int* __restrict__ plane,
    const int* __restrict__ elem, int nplanes, double* __restrict__ dinv,
    int* __restrict__ fail) {
  __shared__ double Dsh[kBiccWaves][kMaxEig * kMaxEig];
  __shared__ double Csh[kBiccWaves][kMaxEig * kMaxEig];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const size_t kk = (size_t)kmax * kmax;
  double* D = Dsh[wv];
  double* C = Csh[wv];
  for (int p = 0; p < nplanes; ++p) {
    const int cnt = plane[p + 1] - plane[p];
    for (int r0 = 0; r0 < cnt; r0 += kBiccWaves) {
      const bool act = r0 + wv < cnt;
      const int s = act ? elem[plane[p] + r0 + wv] : 0;
      const Brick& b = g.bricks[s];
      const int k = b.neig;
      const double* blk = Acb + (size_t)s * 7 * kk;
      if (act)
        for (int e = lane; e < k * k; e += 64) D[e] = blk[(e / k) * kmax + e % k];
      __syncthreads();
      for (int dir = 0; dir < 6; dir += 2) {
        // process boxes: the ICC(0) of each box's block (no coupling between
        // boxes, spe10.c:1074-1205 assembles A_c per process)
        int t = act ? b.nb[dir] : -1;
        if (t >= 0 && g.bricks[t].box != b.box) t = -1;
        const int kt = t >= 0 ? g.bricks[t].neig : 0;
        const double* B = blk + (size_t)(dir + 1) * kk;  // A_st, columns over t's dofs
        const double* Dt = dinv + (size_t)(t >= 0 ? t : 0) * kk;
        for (int e = lane; e < k * kt; e += 64) {  // C = A_st D_t^{-1}
          const int i = e / kt, j = e % kt;
          double acc = 0.0;
          for (int l = 0; l < kt; ++l) acc += B[i * kmax + l] * Dt[l * kmax + j];
          C[e] = acc;
        }
        __syncthreads();
        for (int e = lane; e < k * k && kt > 0; e += 64) {  // D -= C A_st^T
          const int i = e / k, j = e % k;
          double acc = 0.0;
          for (int l = 0; l < kt; ++l) acc += C[i * kt + l] * B[j * kmax + l];
          D[e] -= acc;
        }
        __syncthreads();
      }
      if (act && lane == 0) {
        // Cholesky D = L L^T (lower triangle of C), then D^{-1} = L^{-T} L^{-1}.
        bool ok = true;
        for (int j = 0; j < k && ok; ++j) {
          double dj = 0.5 * (D[j * k + j] + D[j * k + j]);
          for (int l = 0; l < j; ++l) dj -= C[j * k + l] * C[j * k + l];
          if (!(dj > 0.0)) {
            ok = false;
            break;
          }
          const double ljj = sqrt(dj);
          C[j * k + j] = ljj;
          for (int i = j + 1; i < k; ++i) {
            double v = 0.5 * (D[i * k + j] + D[j * k + i]);
            for (int l = 0; l < j; ++l) v -= C[i * k + l] * C[j * k + l];
            C[i * k + j] = v / ljj;
          }
        }
        if (!ok) {
          atomicMax(fail, s + 1);
        } else {
          // D <- L^{-1} (lower), column by column.
          for (int j = 0; j < k; ++j) {
            for (int i = 0; i < k; ++i) D[i * k + j] = 0.0;
            D[j * k + j] = 1.0 / C[j * k + j];
            for (int i = j + 1; i < k; ++i) {
              double v = 0.0;
              for (int l = j; l < i; ++l) v -= C[i * k + l] * D[l * k + j];
              D[i * k + j] = v / C[i * k + i];
            }
          }
          double* out = dinv + (size_t)s * kk;
          for (int i = 0; i < k; ++i)
            for (int j = 0; j <= i; ++j) {
              double v = 0.0;
              for (int l = i; l < k; ++l) v += D[l * k + i] * D[l * k + j];
              out[i * kmax + j] = v;
              out[j * kmax + i] = v;
            }
        }
      }
      __syncthreads();
    }
  }
}

// d = M^{-1} t with the block ICC(0) factor: forward w_s = D_s^{-1} (t_s -
// sum_{lower o} A_so w_o), backward y_s = w_s - D_s^{-1} sum_{upper o} A_so
// y_o.  A subdomain's lower (upper) neighbours all lie on the previous
// (next) hyperplane, so the values of the last plane live in an LDS ring of
// two plane slots (slot entry pos*kp + row): the dependency chain of the
// sweep never waits on global memory.  Each hyperplane position has a
// precomputed descriptor (subdomain, its dofs, and the ring positions and
// eigen counts of its neighbours), so a round needs one dependent load
// before its matrix rows.  w and y are also written to d.  Lane group of
// kp >= neig lanes per subdomain, one row each.
struct BiccSlot {
  int s, dof, neig, pad;
  int lpos[3], lk[3];   // -x, -y, -z neighbours: ring position (-1: none), eigen count
  int hpos[3], hk[3];   // +x, +y, +z neighbours
};

struct BiccArgs {
  const double* dinv;
  const int* plane;
  const BiccSlot* slot;   // descriptors in hyperplane order
  int nplanes, kp, maxcnt;
};

constexpr int kBiccMaxPlanes = 512;

__global__ void __launch_bounds__(1024) k_bicc_solve(const double* __restrict__ Acb, int kmax,
                                                     BiccArgs bi, const double* __restrict__ tcol,
                                                     double* __restrict__ dcol, int nc) {
  extern __shared__ double ring[];   // [2][maxcnt * kp]
  __shared__ double vsh[2][1024];
  __shared__ int pl[kBiccMaxPlanes + 1];
  const double* t = tcol + (size_t)blockIdx.x * nc;
  double* d = dcol + (size_t)blockIdx.x * nc;
  for (int e = threadIdx.x; e <= bi.nplanes; e += blockDim.x) pl[e] = bi.plane[e];
  __syncthreads();
  const size_t kk = (size_t)kmax * kmax;
  const int kp = bi.kp;
  const int grp = threadIdx.x / kp, i = threadIdx.x % kp;
  const int epr = blockDim.x / kp;
  const int rs = bi.maxcnt * kp;
  int round = 0;
  for (int dirn = 0; dirn < 2; ++dirn) {
    const bool fwd = dirn == 0;
    for (int pp = 0; pp < bi.nplanes; ++pp) {
      const int p = fwd ? pp : bi.nplanes - 1 - pp;
      double* cur = ring + (p & 1) * rs;
      const double* oth = ring + ((p + 1) & 1) * rs;  // plane p-1 (fwd) or p+1 (bwd)
      const int cnt = pl[p + 1] - pl[p];
      for (int r0 = 0; r0 < cnt; r0 += epr, ++round) {
        double* vb = vsh[round & 1];
        const int q = r0 + grp;
        const bool grp_act = q < cnt;
        BiccSlot sl{};
        if (grp_act) sl = bi.slot[pl[p] + q];
        const bool act = grp_act && i < sl.neig;
        double v = 0.0, own = 0.0;
        double di[16];
        if (act) {
          const double* Di = bi.dinv + (size_t)sl.s * kk + (size_t)i * kmax;
#pragma unroll
          for (int j = 0; j < 16; ++j) di[j] = j < sl.neig ? Di[j] : 0.0;
          v = fwd ? t[sl.dof + i] : 0.0;
          if (!fwd) own = d[sl.dof + i];
          const double* blk = Acb + (size_t)sl.s * 7 * kk + (size_t)i * kmax;
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            const int op = fwd ? sl.lpos[a] : sl.hpos[a];
            if (op < 0) continue;
            const int ko = fwd ? sl.lk[a] : sl.hk[a];
            const double* B = blk + (size_t)(2 * a + (fwd ? 1 : 2)) * kk;
            const double* xo = oth + op * kp;
            double acc = 0.0;
            for (int j = 0; j < ko; ++j) acc += B[j] * xo[j];
            v = fwd ? v - acc : v + acc;
          }
        }
        vb[threadIdx.x] = v;
        __syncthreads();
        if (act) {
          const double* vg = vb + grp * kp;
          double acc = 0.0;
#pragma unroll
          for (int j = 0; j < 16; ++j)
            if (j < sl.neig) acc += di[j] * vg[j];
          const double r = fwd ? acc : own - acc;
          cur[q * kp + i] = r;
          d[sl.dof + i] = r;
        }
      }
      __syncthreads();   // this plane's ring slot is complete before the next plane reads it
    }
  }
}

// The same sweeps with the next round's row data loaded ahead: every round
// of the flattened (direction, plane, row group) schedule first loads its
// slot, D^{-1} row, neighbour coupling rows and right-hand side, which do
// not depend on the sweep, so those loads overlap the previous round's
// barriers instead of starting after them.  Only the backward sweep's own
// value of its first plane (written by the last forward round) is read at
// use time.  KP = lanes per subdomain = en_lv1 (8); the arithmetic and its
// order are those of k_bicc_solve.
template <int KP>
__global__ void __launch_bounds__(1024) k_bicc_solve_pf(const double* __restrict__ Acb, int kmax,
                                                        BiccArgs bi,
                                                        const double* __restrict__ tcol,
                                                        double* __restrict__ dcol, int nc) {
  extern __shared__ double ring[];   // [2][maxcnt * KP]
  __shared__ double vsh[2][1024];
  __shared__ int pl[kBiccMaxPlanes + 1];
  const double* t = tcol + (size_t)blockIdx.x * nc;
  double* d = dcol + (size_t)blockIdx.x * nc;
  for (int e = threadIdx.x; e <= bi.nplanes; e += blockDim.x) pl[e] = bi.plane[e];
  __syncthreads();
  const size_t kk = (size_t)kmax * kmax;
  const int grp = threadIdx.x / KP, i = threadIdx.x % KP;
  const int epr = blockDim.x / KP;
  const int rs = bi.maxcnt * KP;
  const int np = bi.nplanes;
  // Row data of the current round (c*) and of the next one (n*).
  bool cact = false, nact = false;
  int cdof = 0, cneig = 0, ndof = 0, nneig = 0;
  int cop[3], cko[3], nop[3], nko[3];
  double cdi[KP], cB[3 * KP], cv0 = 0.0, ndi[KP], nB[3 * KP], nv0 = 0.0;
#define PMX_BICC_FETCH(DIRN, PP, R0)                                                        \
  {                                                                                        \
    const bool ffwd = (DIRN) == 0;                                                         \
    const int fp = ffwd ? (PP) : np - 1 - (PP);                                            \
    const int fq = (R0) + grp;                                                             \
    nact = false;                                                                          \
    if (fq < pl[fp + 1] - pl[fp]) {                                                        \
      const BiccSlot* sl = bi.slot + pl[fp] + fq;                                          \
      nneig = sl->neig;                                                                    \
      nact = i < nneig;                                                                    \
      if (nact) {                                                                          \
        const int fs = sl->s;                                                              \
        ndof = sl->dof;                                                                    \
        const double* Di = bi.dinv + (size_t)fs * kk + (size_t)i * kmax;                   \
        _Pragma("unroll") for (int j = 0; j < KP; ++j) ndi[j] = j < nneig ? Di[j] : 0.0;   \
        nv0 = ffwd ? t[ndof + i] : ((PP) == 0 ? 0.0 : d[ndof + i]);                        \
        const double* blk = Acb + (size_t)fs * 7 * kk + (size_t)i * kmax;                  \
        _Pragma("unroll") for (int a = 0; a < 3; ++a) {                                    \
          nop[a] = ffwd ? sl->lpos[a] : sl->hpos[a];                                       \
          nko[a] = ffwd ? sl->lk[a] : sl->hk[a];                                           \
          const double* B = blk + (size_t)(2 * a + (ffwd ? 1 : 2)) * kk;                   \
          _Pragma("unroll") for (int j = 0; j < KP; ++j) nB[a * KP + j] =                  \
              (nop[a] >= 0 && j < nko[a]) ? B[j] : 0.0;                                    \
        }                                                                                  \
      }                                                                                    \
    }                                                                                      \
  }
  int dirn = 0, pp = 0, r0 = 0, round = 0;
  PMX_BICC_FETCH(0, 0, 0)
  for (;;) {
    cact = nact;
    cdof = ndof;
    cneig = nneig;
    cv0 = nv0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      cop[a] = nop[a];
      cko[a] = nko[a];
    }
#pragma unroll
    for (int j = 0; j < KP; ++j) cdi[j] = ndi[j];
#pragma unroll
    for (int j = 0; j < 3 * KP; ++j) cB[j] = nB[j];
    const bool fwd = dirn == 0;
    const int p = fwd ? pp : np - 1 - pp;
    double* cur = ring + (p & 1) * rs;
    const double* oth = ring + ((p + 1) & 1) * rs;  // plane p-1 (fwd) or p+1 (bwd)
    const int cnt = pl[p + 1] - pl[p];
    double* vb = vsh[round & 1];
    const int q = r0 + grp;
    double v = 0.0;
    if (cact) {
      v = fwd ? cv0 : 0.0;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (cop[a] < 0) continue;
        const double* xo = oth + cop[a] * KP;
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < KP; ++j)
          if (j < cko[a]) acc += cB[a * KP + j] * xo[j];
        v = fwd ? v - acc : v + acc;
      }
    }
    vb[threadIdx.x] = v;
    // Next round of the schedule, and its loads.
    int nd = dirn, npp = pp, nr0 = r0 + epr;
    const bool plane_end = nr0 >= cnt;
    if (plane_end) {
      nr0 = 0;
      if (++npp == np) {
        npp = 0;
        ++nd;
      }
    }
    const bool more = nd < 2;
    nact = false;
    if (more) PMX_BICC_FETCH(nd, npp, nr0)
    __syncthreads();
    if (cact) {
      const double own = fwd ? 0.0 : (pp == 0 ? d[cdof + i] : cv0);
      const double* vg = vb + grp * KP;
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < KP; ++j)
        if (j < cneig) acc += cdi[j] * vg[j];
      const double r = fwd ? acc : own - acc;
      cur[q * KP + i] = r;
      d[cdof + i] = r;
    }
    if (plane_end) __syncthreads();  // this plane's ring slot is complete before the next reads it
    if (!more) break;
    dirn = nd;
    pp = npp;
    r0 = nr0;
    ++round;
  }
#undef PMX_BICC_FETCH
}

static int bicc_lanes(int kmax) {
  int kp = 1;
  while (kp < kmax) kp *= 2;
  return kp;
}

static BiccArgs bicc_args(const pmx_ctx* ctx) {
  return BiccArgs{ctx->bicc_dinv.p, ctx->bicc_plane.p,
                  reinterpret_cast<const BiccSlot*>(ctx->bicc_slot.p), ctx->bicc_nplanes,
                  bicc_lanes(ctx->en_lv1), ctx->bicc_maxcnt};
}

int launch_bicc_solve(pmx_ctx* ctx, const double* t, double* d, int ncol) {
  const BiccArgs bi = bicc_args(ctx);
  const size_t lds = 2 * (size_t)bi.maxcnt * bi.kp * sizeof(double);
  // The prefetching variant reads whole KP-wide rows: only where KP = en_lv1
  // (8, the spe10 fork's count; checked bit-identical to k_bicc_solve).
  if (bi.kp == 8 && ctx->en_lv1 == 8 && ctx->bicc_prefetch)
    hipLaunchKernelGGL(k_bicc_solve_pf<8>, dim3(ncol), dim3(1024), lds, ctx->stream, ctx->Acb.p,
                       ctx->en_lv1, bi, t, d, ctx->nc);
  else
    hipLaunchKernelGGL(k_bicc_solve, dim3(ncol), dim3(1024), lds, ctx->stream, ctx->Acb.p,
                       ctx->en_lv1, bi, t, d, ctx->nc);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

int bicc_setup(pmx_ctx* ctx) {
  const int sdx = ctx->sd[0], sdy = ctx->sd[1], sdz = ctx->sd[2];
  const int np = sdx + sdy + sdz - 2;
  std::vector<int> start(np + 1, 0), elems;
  elems.reserve(ctx->nsub);
  for (int p = 0; p < np; ++p) {
    start[p] = (int)elems.size();
    for (int s = 0; s < ctx->nsub; ++s) {
      const int sx = s % sdx, sy = (s / sdx) % sdy, sz = s / (sdx * sdy);
      if (sx + sy + sz == p) elems.push_back(s);
    }
  }
  start[np] = (int)elems.size();
  std::vector<int> pos(ctx->nsub);
  int maxcnt = 0;
  for (int p = 0; p < np; ++p) {
    maxcnt = std::max(maxcnt, start[p + 1] - start[p]);
    for (int q = start[p]; q < start[p + 1]; ++q) pos[elems[q]] = q - start[p];
  }
  const size_t lds = 2 * (size_t)maxcnt * bicc_lanes(ctx->en_lv1) * sizeof(double);
  if (lds + 2 * 1024 * sizeof(double) + (kBiccMaxPlanes + 1) * sizeof(int) > 65536 ||
      np > kBiccMaxPlanes || ctx->en_lv1 > 16) {
    set_error("-lv2_icc: %d coarse hyperplanes of up to %d subdomains exceed the LDS plane ring",
              np, maxcnt);
    return PMX_ERR_SUP;
  }
  std::vector<BiccSlot> slots(ctx->nsub);
  for (int q = 0; q < ctx->nsub; ++q) {
    const Brick& b = ctx->h_bricks[elems[q]];
    BiccSlot& sl = slots[q];
    sl.s = elems[q];
    sl.dof = b.dof;
    sl.neig = b.neig;
    sl.pad = 0;
    for (int a = 0; a < 3; ++a) {
      // process boxes: neighbours in another box do not couple
      int lo = b.nb[2 * a], hi = b.nb[2 * a + 1];
      if (lo >= 0 && ctx->h_bricks[lo].box != b.box) lo = -1;
      if (hi >= 0 && ctx->h_bricks[hi].box != b.box) hi = -1;
      sl.lpos[a] = lo >= 0 ? pos[lo] : -1;
      sl.lk[a] = lo >= 0 ? ctx->h_bricks[lo].neig : 0;
      sl.hpos[a] = hi >= 0 ? pos[hi] : -1;
      sl.hk[a] = hi >= 0 ? ctx->h_bricks[hi].neig : 0;
    }
  }
  static_assert(sizeof(BiccSlot) % sizeof(int) == 0, "descriptor layout");
  PMX_TRY(ctx->bicc_slot.alloc(ctx->nsub * sizeof(BiccSlot) / sizeof(int)));
  PMX_HIP(hipMemcpy(ctx->bicc_slot.p, slots.data(), ctx->nsub * sizeof(BiccSlot),
                    hipMemcpyHostToDevice));
  ctx->bicc_maxcnt = maxcnt;
  PMX_TRY(ctx->bicc_plane.alloc(np + 1));
  PMX_TRY(ctx->bicc_elem.alloc(ctx->nsub));
  PMX_HIP(hipMemcpy(ctx->bicc_plane.p, start.data(), (np + 1) * sizeof(int),
                    hipMemcpyHostToDevice));
  PMX_HIP(hipMemcpy(ctx->bicc_elem.p, elems.data(), ctx->nsub * sizeof(int),
                    hipMemcpyHostToDevice));
  ctx->bicc_nplanes = np;
  const size_t kk = (size_t)ctx->en_lv1 * ctx->en_lv1;
  PMX_TRY(ctx->bicc_dinv.alloc((size_t)ctx->nsub * kk));
  PMX_HIP(hipMemsetAsync(ctx->bicc_dinv.p, 0, ctx->bicc_dinv.n * sizeof(double), ctx->stream));
  DBuf<int> fail;
  PMX_TRY(fail.alloc(1));
  PMX_HIP(hipMemsetAsync(fail.p, 0, sizeof(int), ctx->stream));
  hipLaunchKernelGGL(k_bicc_factor, dim3(1), dim3(64 * kBiccWaves), 0, ctx->stream, ctx->geo,
                     ctx->Acb.p, ctx->en_lv1, ctx->bicc_plane.p, ctx->bicc_elem.p, np,
                     ctx->bicc_dinv.p, fail.p);
  PMX_HIP(hipGetLastError());
  int hf = 0;
  PMX_HIP(hipMemcpyAsync(&hf, fail.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  PMX_HIP(hipStreamSynchronize(ctx->stream));
  if (hf) {
    set_error("level-2 ICC(0) of the coarse matrix: non-positive pivot in subdomain %d "
              "(PETSc would apply its positive-definite shift; not implemented)", hf - 1);
    return PMX_ERR_NOT_CONVERGED;
  }
  return PMX_SUCCESS;
}

// ---------------------------------------------------------------------------
// Composite coarse operator.  Every stage of the coarse program maps the
// coarse residual t to a correction d = G t (G = A_c^{-1} for level 2,
// V A_cc^{-1} V^T for level 3) and updates t -= A_c d, u += d.  The program
// is therefore one linear map u = F t0, composed here once per setup with
// dense matrices:  T = I, U = 0;  per stage D = G T, U += D, T -= A_c D.
// ---------------------------------------------------------------------------
constexpr int kGt = 64;   // GEMM tile

// S = B^{-1} - [z/dg, h] Cap^{-1} [z/dg, h]^T  (the Woodbury-corrected A_c^{-1}).
__global__ void k_form_S(const double* __restrict__ Binv, const double* __restrict__ z,
                         const double* __restrict__ h, const double* __restrict__ cs,
                         double* __restrict__ S, int n) {
  const size_t nn = (size_t)n * n;
  const double dg = cs[0], a00 = cs[1], a01 = cs[2], a11 = cs[3];
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n), j = (int)(e % n);
    const double zi = z[i] / dg, zj = z[j] / dg;
    S[e] = Binv[e] - zi * (a00 * zj + a01 * h[j]) - h[i] * (a01 * zj + a11 * h[j]);
  }
}

// U = D (first) or U += D; if update: T -= A_c D (block rows of A_c).
__global__ void k_stage_update(Geo g, const double* __restrict__ Acb, int kmax,
                               const double* __restrict__ D, double* __restrict__ T,
                               double* __restrict__ U, int n, int first, int update) {
  const size_t nn = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn;
       e += (size_t)gridDim.x * blockDim.x) {
    U[e] = first ? D[e] : U[e] + D[e];
  }
  if (!update) return;
  // rows are coarse dofs; column j is independent
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)g.nsub * kmax * n;
       e += (size_t)gridDim.x * blockDim.x) {
    const int j = (int)(e % n);
    const int gid = (int)(e / n);
    const int sb = gid / kmax, i = gid % kmax;
    const Brick& b = g.bricks[sb];
    if (i >= b.neig) continue;
    const double* blk = Acb + (size_t)sb * 7 * kmax * kmax;
    double acc = 0.0;
    for (int q = 0; q < b.neig; ++q) acc += blk[i * kmax + q] * D[(size_t)(b.dof + q) * n + j];
    for (int dir = 0; dir < 6; ++dir) {
      const int sn = b.nb[dir];
      if (sn < 0) continue;
      const Brick& o = g.bricks[sn];
      const double* bb = blk + (size_t)(dir + 1) * kmax * kmax;
      for (int q = 0; q < o.neig; ++q) acc += bb[i * kmax + q] * D[(size_t)(o.dof + q) * n + j];
    }
    T[(size_t)(b.dof + i) * n + j] -= acc;
  }
}

// Y = A_cc^{-1} (V^T T), k2 x n (column j per thread).
// Y[:, j] = A_cc^{-1} Y[:, j] (k2 x k2 LU with partial pivoting, row
// interchanges in LU[k2*k2 ..]) for every column j of the k2 x n block Y.
// KM: capacity of the level-3 space (kMaxEig, or kMaxCC with process boxes).
template <int KM>
__global__ void k_lv3_lu_cols(int k2, const double* __restrict__ LU, double* __restrict__ Y,
                              int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double b[KM];
  for (int q = 0; q < k2; ++q) b[q] = Y[(size_t)q * n + j];
  const double* piv = LU + k2 * k2;
  for (int i = 0; i < k2; ++i) {
    const int p = (int)piv[i];
    const double tmp = b[i];
    b[i] = b[p];
    b[p] = tmp;
  }
  for (int i = 0; i < k2; ++i)
    for (int q = 0; q < i; ++q) b[i] -= LU[i * k2 + q] * b[q];
  for (int i = k2 - 1; i >= 0; --i) {
    for (int q = i + 1; q < k2; ++q) b[i] -= LU[i * k2 + q] * b[q];
    b[i] /= LU[i * k2 + i];
  }
  for (int q = 0; q < k2; ++q) Y[(size_t)q * n + j] = b[q];
}

// D = V Y (n x n).
__global__ void k_lv3_right(const double* __restrict__ V, int k2, const double* __restrict__ Y,
                            double* __restrict__ D, int n) {
  const size_t nn = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn;
       e += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(e / n), j = (int)(e % n);
    double acc = 0.0;
    for (int q = 0; q < k2; ++q) acc += V[(size_t)q * n + r] * Y[(size_t)q * n + j];
    D[e] = acc;
  }
}

__global__ void k_set_identity(double* __restrict__ T, int n) {
  const size_t nn = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn;
       e += (size_t)gridDim.x * blockDim.x)
    T[e] = (e / n == e % n) ? 1.0 : 0.0;
}

int coarse_operator_setup(pmx_ctx* ctx, const std::vector<int>& prog) {
  const int n = ctx->nc;
  const size_t nn = (size_t)n * n;
  DBuf<double> S, T, D, Y;
  PMX_TRY(S.alloc(nn));
  PMX_TRY(T.alloc(nn));
  PMX_TRY(D.alloc(nn));
  PMX_TRY(ctx->Fop.alloc(nn));
  const int eb = 2048;
  hipLaunchKernelGGL(k_form_S, dim3(eb), dim3(kBlock), 0, ctx->stream, ctx->Ainv.p, ctx->c_z.p,
                     ctx->c_h.p, ctx->c_cs.p, S.p, n);
  hipLaunchKernelGGL(k_set_identity, dim3(eb), dim3(kBlock), 0, ctx->stream, T.p, n);
  PMX_HIP(hipGetLastError());
  const int nt = (n + kGt - 1) / kGt;
  const int k2 = ctx->eigen_num_lv2;
  if (k2 > 0) PMX_TRY(Y.alloc((size_t)k2 * n));
  bool T_is_I = true;
  for (size_t k = 0; k < prog.size(); ++k) {
    const bool last = k + 1 == prog.size();
    const double* Dp;
    if (prog[k] == 2) {
      if (T_is_I) {
        Dp = S.p;
      } else {
        hipLaunchKernelGGL((k_gemm_mfma<2, 2, 2, 4, false>), dim3((n + 127) / 128, (n + 63) / 64),
                           dim3(256), 0, ctx->stream, n, n, n, 1.0, S.p, n, T.p, n, 0.0, D.p, n);
        Dp = D.p;
      }
    } else {
      // Y = A_cc^{-1} V^T T: V^T T on MFMA (V column-major = V^T row-major),
      // then the k2 x k2 solve per column.
      hipLaunchKernelGGL((k_gemm_mfma<1, 4, 1, 1, false>), dim3((n + 63) / 64, (k2 + 15) / 16),
                         dim3(256), 0, ctx->stream, k2, n, n, 1.0, ctx->V2.p, n, T.p, n, 0.0, Y.p,
                         n);
      hipLaunchKernelGGL(k2 <= kMaxEig ? k_lv3_lu_cols<kMaxEig> : k_lv3_lu_cols<kMaxCC>,
                         dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, k2,
                         ctx->Acc_lu.p, Y.p, n);
      hipLaunchKernelGGL(k_lv3_right, dim3(eb), dim3(kBlock), 0, ctx->stream, ctx->V2.p, k2, Y.p,
                         D.p, n);
      Dp = D.p;
    }
    hipLaunchKernelGGL(k_stage_update, dim3(eb), dim3(kBlock), 0, ctx->stream, ctx->geo,
                       ctx->Acb.p, ctx->en_lv1, Dp, T.p, ctx->Fop.p, n, k == 0, last ? 0 : 1);
    PMX_HIP(hipGetLastError());
    if (!last) T_is_I = false;
  }
  PMX_HIP(hipStreamSynchronize(ctx->stream));
  return PMX_SUCCESS;
}

// Runs the coarse program prog (2 = level 2, 3 = level 3) on t, leaving the
// total correction in u: GEMVs with B^{-1} on the whole device, everything
// else fused per segment.
int launch_coarse_prec(pmx_ctx* ctx, const double* t, double* d, int ncol) {
  if (ctx->use_band)
    return ctx->use_planes ? launch_planes_solve(ctx, t, d, ncol)
                           : launch_band_solve(ctx, t, d, ncol);
  return launch_bicc_solve(ctx, t, d, ncol);
}

int launch_coarse_program(pmx_ctx* ctx, const std::vector<int>& prog, double* t, double* u,
                          double* y, double* d, int kt_gemv, int kt_small,
                          int (*kt_begin)(pmx_ctx*, int), void (*kt_end)(pmx_ctx*, int)) {
  // One stage at a time: the level-2 solve (dense inverse GEMV + Woodbury
  // fix, ICC(0) sweeps or band solves) or the level-3 correction (one
  // workgroup), then u += d and t -= A_c d on the whole device.
  const bool sparse = ctx->lv2_icc || ctx->use_band;
  for (size_t k = 0; k < prog.size(); ++k) {
    const int ev = kt_begin(ctx, prog[k] == 2 ? kt_gemv : kt_small);
    if (prog[k] == 2 && sparse) {
      PMX_TRY(launch_coarse_prec(ctx, t, d, 1));
    } else if (prog[k] == 2) {
      if (ctx->coarse_deflate) {
        PMX_TRY(launch_gemv(ctx, ctx->Ainv.p, t, y, ctx->nc));
        PMX_TRY(launch_coarse_fix(ctx, y, t, d));
      } else {
        PMX_TRY(launch_gemv(ctx, ctx->Ainv.p, t, d, ctx->nc));
      }
    } else {
      PMX_TRY(launch_lv3_apply(ctx, t, d));
    }
    kt_end(ctx, ev);
    PMX_TRY(launch_coarse_update(ctx, d, t, u, k == 0, k + 1 < prog.size()));
  }
  return PMX_SUCCESS;
}

// u = d (first) or u += d; if update_t: t -= A_c d.
__global__ void k_coarse_update(Geo g, const double* __restrict__ Acb, int kmax,
                                const double* __restrict__ d, double* __restrict__ t,
                                double* __restrict__ u, int first, int update_t) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = gid / kmax, i = gid % kmax;
  if (s >= g.nsub) return;
  const Brick& b = g.bricks[s];
  if (i >= b.neig) return;
  const int r = b.dof + i;
  u[r] = first ? d[r] : u[r] + d[r];
  if (update_t) t[r] -= coarse_row(g, Acb, kmax, s, i, d);
}

int launch_coarse_update(pmx_ctx* ctx, const double* d, double* t, double* u, int first,
                         int update_t) {
  const int n = ctx->nsub * ctx->en_lv1;
  hipLaunchKernelGGL(k_coarse_update, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     ctx->stream, ctx->geo, ctx->Acb.p, ctx->en_lv1, d, t, u, first, update_t);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

// Y = A_c X for m column vectors (column-major, stride nc); one thread per
// row and column (grid y = the column), so the columns' rows run in
// parallel rather than one after the other in a thread.
__global__ void k_coarse_spmm(Geo g, const double* __restrict__ Acb, int kmax,
                              const double* __restrict__ X, double* __restrict__ Y, int nc) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = gid / kmax, i = gid % kmax, c = blockIdx.y;
  if (s >= g.nsub) return;
  const Brick& b = g.bricks[s];
  if (i >= b.neig) return;
  Y[(size_t)c * nc + b.dof + i] = coarse_row(g, Acb, kmax, s, i, X + (size_t)c * nc);
}

// d = V (A_cc^{-1} (V^T t)) with the LU of A_cc (level-3 apply,
// _PC_apply_vec_lv3, src/PreMixFEM_3D.c:1597-1648, written in the coarse
// space since Phi_cc = Phi V).  One workgroup.
template <int KM>
__global__ void k_lv3_apply(const double* __restrict__ V, int nc, int k2,
                            const double* __restrict__ LU, const double* __restrict__ t,
                            double* __restrict__ d) {
  __shared__ double red[8];
  __shared__ double c[KM];
  for (int j = 0; j < k2; ++j) {
    double s = 0.0;
    for (int r = threadIdx.x; r < nc; r += blockDim.x) s += V[(size_t)j * nc + r] * t[r];
    s = block_sum(s, red);
    if (threadIdx.x == 0) c[j] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    // LU with row pivots stored after the k2 x k2 factors.
    const double* piv = LU + k2 * k2;
    double b[KM];
    for (int i = 0; i < k2; ++i) b[i] = c[i];
    for (int i = 0; i < k2; ++i) {
      const int p = (int)piv[i];
      const double tmp = b[i];
      b[i] = b[p];
      b[p] = tmp;
    }
    for (int i = 0; i < k2; ++i)
      for (int j = 0; j < i; ++j) b[i] -= LU[i * k2 + j] * b[j];
    for (int i = k2 - 1; i >= 0; --i) {
      for (int j = i + 1; j < k2; ++j) b[i] -= LU[i * k2 + j] * b[j];
      b[i] /= LU[i * k2 + i];
    }
    for (int i = 0; i < k2; ++i) c[i] = b[i];
  }
  __syncthreads();
  for (int r = threadIdx.x; r < nc; r += blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < k2; ++j) s += V[(size_t)j * nc + r] * c[j];
    d[r] = s;
  }
}

int launch_lv3_apply(pmx_ctx* ctx, const double* t, double* d) {
  hipLaunchKernelGGL(ctx->eigen_num_lv2 <= kMaxEig ? k_lv3_apply<kMaxEig> : k_lv3_apply<kMaxCC>,
                     dim3(1), dim3(kBlock), 0, ctx->stream, ctx->V2.p, ctx->nc,
                     ctx->eigen_num_lv2, ctx->Acc_lu.p, t, d);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

// Multi-workgroup block algebra of the level-3 eigensolvers: column-major
// n x m blocks; Gram matrices summed by many workgroups (partials in fixed
// order), the small Cholesky / eigenproblems in one workgroup.
constexpr int kGramRows = 2048;   // rows per workgroup of a partial Gram
constexpr int kLv3Max = 32;       // largest LOBPCG basis [X W P] (3 m)
// Stagnation window of the shift-invert subspace iterations of level 3:
// they gain a constant factor per iteration (4x on the 256^3 fracture
// medium), so kLv3StagWin iterations without a 2x gain mean the residual sits
// on its rounding floor.
constexpr int kLv3StagWin = 3;

// part[(g*ma + i)*mb + j] = sum over rows of chunk g of A[:, i] . B[:, j].
__global__ void __launch_bounds__(256) k_gram_part(const double* __restrict__ A, int ma,
                                                   const double* __restrict__ B, int mb, int n,
                                                   double* __restrict__ part) {
  __shared__ double red[4][kLv3Max];
  const int g = blockIdx.x, i = blockIdx.y;
  const int r0 = g * kGramRows, r1 = min(n, r0 + kGramRows);
  double acc[kLv3Max];
#pragma unroll
  for (int j = 0; j < kLv3Max; ++j) acc[j] = 0.0;
  const double* a = A + (size_t)i * n;
  for (int r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
    const double av = a[r];
#pragma unroll
    for (int j = 0; j < kLv3Max; ++j)
      if (j < mb) acc[j] += av * B[(size_t)j * n + r];
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < kLv3Max; ++j) {
    if (j >= mb) break;
    double v = acc[j];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[wv][j] = v;
  }
  __syncthreads();
  if (threadIdx.x < mb) {
    double v = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v += red[w][threadIdx.x];
    part[((size_t)g * ma + i) * mb + threadIdx.x] = v;
  }
}

__global__ void k_gram_sum(const double* __restrict__ part, int G, int ma, int mb,
                           double* __restrict__ out) {
  for (int e = threadIdx.x; e < ma * mb; e += blockDim.x) {
    double v = 0.0;
    for (int g = 0; g < G; ++g) v += part[(size_t)g * ma * mb + e];
    out[e] = v;
  }
}

// mode 0: T = L^{-T} for the Cholesky factor L of G (ns x ns, row-major);
// *flag = 1 if G is not numerically positive definite.  mode 1: G
// (symmetrised) = Q diag(w) Q^T, T = Q, ascending w.
__global__ void k_lv3_small(const double* __restrict__ Gm, int ns, int mode,
                            double* __restrict__ T, double* __restrict__ wout,
                            int* __restrict__ flag) {
  constexpr int L = kLv3Max;
  __shared__ double A[L * L], Q[L * L], tmp[L * L], Q1[L * L], w[L], rot[4 * L];
  __shared__ int fl;
  // mode 0 works on the column-scaled Gram D G D (D = diag(G)^{-1/2}), so
  // columns of very different norms (small residual directions) stay
  // well-conditioned; T = D L^{-T}.
  if (threadIdx.x < ns) w[threadIdx.x] = Gm[threadIdx.x * ns + threadIdx.x] > 0.0
                                             ? 1.0 / sqrt(Gm[threadIdx.x * ns + threadIdx.x])
                                             : 1.0;
  __syncthreads();
  for (int e = threadIdx.x; e < ns * ns; e += blockDim.x) {
    const int i = e / ns, j = e % ns;
    A[i * L + j] = mode == 0 ? Gm[e] * w[i] * w[j] : 0.5 * (Gm[e] + Gm[j * ns + i]);
  }
  __syncthreads();
  if (mode == 0) {
    const bool ok = wg_cholesky(A, ns, L, &fl);
    if (threadIdx.x == 0 && !ok) *flag = 1;   // sticky: cleared by the caller
    wg_trinv_lower(A, tmp, ns, L);  // A = L^{-1}
    for (int e = threadIdx.x; e < ns * ns; e += blockDim.x) {
      const int k = e / ns, c = e % ns;
      T[e] = w[k] * A[c * L + k];
    }
    return;
  }
  wg_syev2(A, tmp, Q, Q1, w, rot, ns, L);
  for (int e = threadIdx.x; e < ns * ns; e += blockDim.x) T[e] = Q[(e / ns) * L + e % ns];
  if (threadIdx.x < ns) wout[threadIdx.x] = w[threadIdx.x];
}

// out[:, c] = beta out[:, c] + alpha sum_{k < kin} S[:, k] T[(r0 + k) * tld + c0 + c]
// for c < cout.
__global__ void __launch_bounds__(256) k_xform(const double* __restrict__ S, int n, int kin,
                                               const double* __restrict__ T, int tld, int r0,
                                               int c0, int cout, double* __restrict__ out,
                                               double alpha, double beta) {
  __shared__ double Ts[kLv3Max * kLv3Max];
  for (int e = threadIdx.x; e < kin * cout; e += blockDim.x)
    Ts[e] = T[(size_t)(r0 + e / cout) * tld + c0 + e % cout];
  __syncthreads();
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    double sv[kLv3Max];
#pragma unroll
    for (int k = 0; k < kLv3Max; ++k) sv[k] = k < kin ? S[(size_t)k * n + r] : 0.0;
    for (int c = 0; c < cout; ++c) {
      double v = 0.0;
#pragma unroll
      for (int k = 0; k < kLv3Max; ++k)
        if (k < kin) v += sv[k] * Ts[k * cout + c];
      out[(size_t)c * n + r] = beta == 0.0 ? alpha * v : beta * out[(size_t)c * n + r] + alpha * v;
    }
  }
}

// y += alpha x.
__global__ void k_axpy(const double* __restrict__ x, double alpha, double* __restrict__ y,
                       size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] += alpha * x[i];
}

// res[j] = ||X_j|| (one workgroup per column).
__global__ void k_col_norm(const double* __restrict__ X, int n, double* __restrict__ res) {
  __shared__ double red[8];
  const int j = blockIdx.x;
  double s = 0.0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    const double e = X[(size_t)j * n + r];
    s += e * e;
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) res[j] = sqrt(s);
}

// R = AX - X diag(w), res[j] = ||R_j|| (one workgroup per column).
__global__ void k_col_resid(const double* __restrict__ AX, const double* __restrict__ X,
                            const double* __restrict__ w, double* __restrict__ R, int n,
                            double* __restrict__ res) {
  __shared__ double red[8];
  const int j = blockIdx.x;
  const double wj = w[j];
  double s = 0.0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    const double e = AX[(size_t)j * n + r] - wj * X[(size_t)j * n + r];
    R[(size_t)j * n + r] = e;
    s += e * e;
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) res[j] = sqrt(s);
}

namespace {

struct Lv3Work {
  pmx_ctx* ctx;
  int n, ngram;
  DBuf<double> part, G, T, wall;
  DBuf<int> flag;

  int gram(const double* A, int ma, const double* B, int mb) {
    hipLaunchKernelGGL(k_gram_part, dim3(ngram, ma), dim3(256), 0, ctx->stream, A, ma, B, mb, n,
                       part.p);
    hipLaunchKernelGGL(k_gram_sum, dim3(1), dim3(256), 0, ctx->stream, part.p, ngram, ma, mb,
                       G.p);
    PMX_HIP(hipGetLastError());
    return PMX_SUCCESS;
  }
  int xform(const double* S, int kin, int tld, int r0, int c0, int cout, double* out,
            const double* Tm = nullptr, double alpha = 1.0, double beta = 0.0) {
    const int nb = std::min((n + 255) / 256, 2048);
    hipLaunchKernelGGL(k_xform, dim3(nb), dim3(256), 0, ctx->stream, S, n, kin, Tm ? Tm : T.p, tld,
                       r0, c0, cout, out, alpha, beta);
    PMX_HIP(hipGetLastError());
    return PMX_SUCCESS;
  }
  // S (ns columns) <- S L^{-T}: orthonormal; returns 1 in *bad if the Gram
  // matrix is not positive definite.
  int orthonormalize(double* S, int ns, double* tmp, int* bad) {
    PMX_TRY(gram(S, ns, S, ns));
    PMX_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), ctx->stream));
    hipLaunchKernelGGL(k_lv3_small, dim3(1), dim3(256), 0, ctx->stream, G.p, ns, 0, T.p, wall.p,
                       flag.p);
    PMX_HIP(hipGetLastError());
    int h = 0;
    PMX_HIP(hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    PMX_HIP(hipStreamSynchronize(ctx->stream));
    *bad = h;
    if (h) return PMX_SUCCESS;
    PMX_TRY(xform(S, ns, ns, 0, 0, ns, tmp));
    PMX_HIP(hipMemcpyAsync(S, tmp, (size_t)ns * n * sizeof(double), hipMemcpyDeviceToDevice,
                           ctx->stream));
    return PMX_SUCCESS;
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// Level 3 (src/PreMixFEM_3D.c:918-1390): smallest en_lv2 eigenpairs of the
// coarse matrix by block inverse (subspace) iteration on the dense inverse
// with Rayleigh-Ritz, then Phi_cc = Phi V and A_cc = V^T A_c V + NIL I with
// an LU factorisation.
// ---------------------------------------------------------------------------
// Y[:, c] = Ainv X[:, c] for the m columns of column-major n x m blocks:
// as row-major m x n blocks, Y = X Ainv^T (FP64 MFMA, 32 x 32 tiles, K split
// kTallSplit ways into ws: kTallSplit * m * n doubles).
constexpr int kTallSplit = 8;
static void launch_gemm_tall(hipStream_t st, const double* Ainv, const double* X, double* Y, int n,
                             int m, double* ws) {
  const int kc = ((n + kTallSplit - 1) / kTallSplit + kMk - 1) / kMk * kMk;
  const int nz = (n + kc - 1) / kc;
  hipLaunchKernelGGL((k_gemm_mfma_part<2, 2, 1, 1, true>), dim3((n + 31) / 32, (m + 31) / 32, nz),
                     dim3(256), 0, st, m, n, n, kc, X, n, Ainv, n, ws);
  hipLaunchKernelGGL(k_gemm_sum, dim3(std::min((m * n + 255) / 256, 2048)), dim3(256), 0, st, ws,
                     nz, m, n, 1.0, 0.0, Y, n);
}

// ---------------------------------------------------------------------------
// Level-3 coarse-coarse matrix in the reference's energy form
// (_PC_setup_lv3_cc, src/PreMixFEM_3D.c:1177-1360): with Phi_cc = Phi V on
// the fine grid, A_cc[a][b] = sum over all faces T (dphi_a)(dphi_b) + the pin
// / well terms d phi_a phi_b, + NIL I unless -no_shift_A_cc; then its LU.
// ---------------------------------------------------------------------------
// Sign of every column of V: positive inner product with 1 + frac(0.618..(r+1)).
__global__ void k_lv3_sign(double* __restrict__ V, int n, int k2) {
  __shared__ double red[8];
  __shared__ double sg[kMaxCC];
  for (int j = 0; j < k2; ++j) {
    double s = 0.0;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
      const double wgt = 1.0 + fmod(0.6180339887498949 * (r + 1.0), 1.0);
      s += wgt * V[(size_t)j * n + r];
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) sg[j] = s < 0.0 ? -1.0 : 1.0;
  }
  __syncthreads();
  for (int j = 0; j < k2; ++j)
    for (int r = threadIdx.x; r < n; r += blockDim.x) V[(size_t)j * n + r] *= sg[j];
}

// phicc[c * k2 + a] = sum_j phi[c][j] V[a][dof_s + j] for the cells of brick s.
__global__ void k_phicc(Geo g, const double* __restrict__ phi, int kmax,
                        const double* __restrict__ V, int nc, int k2,
                        double* __restrict__ phicc) {
  const Brick b = g.bricks[blockIdx.x];
  const int n = b.nx * b.ny * b.nz;
  for (int e = threadIdx.x; e < n * k2; e += blockDim.x) {
    const int li = e / k2, a = e % k2;
    const size_t c = (size_t)b.off + li;
    double v = 0.0;
    for (int j = 0; j < b.neig; ++j) v += phi[c * kmax + j] * V[(size_t)a * nc + b.dof + j];
    phicc[c * k2 + a] = v;
  }
}

// Row a = blockIdx.y of brick s = blockIdx.x's share of A_cc: the lower
// faces of its cells (to any brick) and the diagonal terms.
template <int KM>
__global__ void __launch_bounds__(256) k_acc_part(Geo g, const double* __restrict__ phicc, int k2,
                                                  const double* __restrict__ Tx,
                                                  const double* __restrict__ Ty,
                                                  const double* __restrict__ Tz,
                                                  double* __restrict__ part) {
  __shared__ double red[4][KM];
  const int s = blockIdx.x, a = blockIdx.y;
  const BrickView v = load_view(g, s);
  const double* TT3[3] = {Tx, Ty, Tz};
  double acc[KM];
#pragma unroll
  for (int q = 0; q < KM; ++q) acc[q] = 0.0;
  for (int li = threadIdx.x; li < v.n; li += blockDim.x) {
    const int lx = li % v.nx, ly = (li / v.nx) % v.ny, lz = li / v.nxny;
    const int c = v.off + li;
    int nb[6];
    nbr6(v, c, lx, ly, lz, nb);
    const double* pc = phicc + (size_t)c * k2;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int o = nb[2 * d];
      if (o < 0) continue;
      const double t = TT3[d][c];
      const double* po = phicc + (size_t)o * k2;
      const double da = t * (pc[a] - po[a]);
#pragma unroll
      for (int q = 0; q < KM; ++q)
        if (q < k2) acc[q] += da * (pc[q] - po[q]);
    }
    double dg = 0.0;
    if (c == g.pin) dg += g.tau;
    if (g.wd) dg += g.wd[c];
    if (dg != 0.0) {
      const double pa = dg * pc[a];
#pragma unroll
      for (int q = 0; q < KM; ++q)
        if (q < k2) acc[q] += pa * pc[q];
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < KM; ++q) {
    if (q >= k2) break;
    double x = acc[q];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lane == 0) red[wv][q] = x;
  }
  __syncthreads();
  if (threadIdx.x < k2) {
    double x = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) x += red[w][threadIdx.x];
    part[((size_t)s * k2 + a) * k2 + threadIdx.x] = x;
  }
}

// A_cc = sum of the brick partials (fixed order), symmetrised, + shift I, and
// its LU with partial pivoting (row-major factors, pivots after them).
template <int KM>
__global__ void k_acc_lu(const double* __restrict__ part, int nsub, int k2, double shift,
                         double* __restrict__ LU) {
  __shared__ double A[KM * KM];
  for (int e = threadIdx.x; e < k2 * k2; e += blockDim.x) {
    double x = 0.0;
    for (int s = 0; s < nsub; ++s) x += part[(size_t)s * k2 * k2 + e];
    A[(e / k2) * KM + e % k2] = x;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int i = 0; i < k2; ++i)
    for (int j = i + 1; j < k2; ++j) {
      const double m = 0.5 * (A[i * KM + j] + A[j * KM + i]);
      A[i * KM + j] = m;
      A[j * KM + i] = m;
    }
  for (int i = 0; i < k2; ++i) A[i * KM + i] += shift;
  double piv[KM];
  for (int k = 0; k < k2; ++k) {
    int p = k;
    for (int i = k + 1; i < k2; ++i)
      if (fabs(A[i * KM + k]) > fabs(A[p * KM + k])) p = i;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < k2; ++j) {
        const double t = A[k * KM + j];
        A[k * KM + j] = A[p * KM + j];
        A[p * KM + j] = t;
      }
    for (int i = k + 1; i < k2; ++i) {
      const double f = A[i * KM + k] / A[k * KM + k];
      A[i * KM + k] = f;
      for (int j = k + 1; j < k2; ++j) A[i * KM + j] -= f * A[k * KM + j];
    }
  }
  for (int i = 0; i < k2; ++i)
    for (int j = 0; j < k2; ++j) LU[i * k2 + j] = A[i * KM + j];
  for (int i = 0; i < k2; ++i) LU[k2 * k2 + i] = piv[i];
}

// Sign-normalises ctx->V2 (num columns) and builds the LU of A_cc from it.
static int lv3_finish_energy(pmx_ctx* ctx, int num) {
  const int n = ctx->nc;
  hipLaunchKernelGGL(k_lv3_sign, dim3(1), dim3(kBlock), 0, ctx->stream, ctx->V2.p, n, num);
  DBuf<double> phicc, part;
  PMX_TRY(phicc.alloc((size_t)ctx->ncell * num));
  PMX_TRY(part.alloc((size_t)ctx->nsub * num * num));
  hipLaunchKernelGGL(k_phicc, dim3(ctx->nsub), dim3(kBlock), 0, ctx->stream, ctx->geo, ctx->phi.p,
                     ctx->en_lv1, ctx->V2.p, n, num, phicc.p);
  const bool small = num <= kMaxEig;
  hipLaunchKernelGGL(small ? k_acc_part<kMaxEig> : k_acc_part<kMaxCC>, dim3(ctx->nsub, num),
                     dim3(256), 0, ctx->stream, ctx->geo, phicc.p, num, ctx->T[0].p, ctx->T[1].p,
                     ctx->T[2].p, part.p);
  PMX_TRY(ctx->Acc_lu.alloc((size_t)num * num + num));
  hipLaunchKernelGGL(small ? k_acc_lu<kMaxEig> : k_acc_lu<kMaxCC>, dim3(1), dim3(kBlock), 0,
                     ctx->stream, part.p, ctx->nsub, num,
                     ctx->no_shift_A_cc ? 0.0 : kNil, ctx->Acc_lu.p);
  PMX_HIP(hipGetLastError());
  PMX_HIP(hipStreamSynchronize(ctx->stream));
  return PMX_SUCCESS;
}

// ---------------------------------------------------------------------------
// Level 3 with process boxes (_PC_setup_lv3_eigen, src/PreMixFEM_3D.c:918-1175,
// for every box at once): box b takes the smallest en_lv2 eigenpairs of its
// box-local Neumann coarse matrix A_i_inner, a dense symmetric block of a few
// hundred unknowns.  All eigenpairs by parallel cyclic Jacobi (round-robin
// ordering, m/2 disjoint rotations per round; threshold rotations, so the
// zero eigenvalue of an unpinned box stays accurate), every box at once:
// one launch per round, A' = J^T A J and Q' = Q J over the whole grid into
// ping-pong buffers, each workgroup recomputing the round's rotations in
// LDS.  Blocks are packed row-major at moff[b], m_b = boff[b+1] - boff[b].
// ---------------------------------------------------------------------------
constexpr int kJacMaxM = 2048;   // largest box block (LDS of the rotation table)
constexpr int kJacTile = 2048;   // matrix entries per workgroup and round

__global__ void k_box_gather(const double* __restrict__ Ain, int nc, const int* __restrict__ dofs,
                             const int* __restrict__ boff, const long long* __restrict__ moff,
                             double* __restrict__ A, double* __restrict__ Q) {
  const int b = blockIdx.y;
  const int m = boff[b + 1] - boff[b];
  const int* d = dofs + boff[b];
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < m * m; e += gridDim.x * blockDim.x) {
    const int i = e / m, j = e % m;
    A[moff[b] + e] = 0.5 * (Ain[(size_t)d[i] * nc + d[j]] + Ain[(size_t)d[j] * nc + d[i]]);
    Q[moff[b] + e] = i == j ? 1.0 : 0.0;
  }
}

// Round r of the sweep for box blockIdx.y: rows/columns i and partner(i) =
// (2r - i) mod (mp - 1) (index mp - 1 pairs with r; an odd m has a dummy
// index m), rotation (c, s) of the pair from the pre-round A.
__global__ void __launch_bounds__(256) k_jacobi_round(const double* __restrict__ A,
                                                      double* __restrict__ A2,
                                                      const double* __restrict__ Q,
                                                      double* __restrict__ Q2,
                                                      const int* __restrict__ boff,
                                                      const long long* __restrict__ moff, int r,
                                                      int* __restrict__ rotated) {
  __shared__ double cc[kJacMaxM], ss[kJacMaxM];
  __shared__ int part[kJacMaxM];
  const int b = blockIdx.y;
  const int m = boff[b + 1] - boff[b];
  const int e0 = blockIdx.x * kJacTile;
  if (e0 >= m * m) return;
  const int mp = (m + 1) & ~1;
  const double* Ab = A + moff[b];
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    int pi;
    if (r >= mp - 1)
      pi = i;  // a smaller box than the largest: identity round
    else if (i == mp - 1)
      pi = r;
    else if (i == r)
      pi = mp - 1;
    else
      pi = ((2 * r - i) % (mp - 1) + (mp - 1)) % (mp - 1);
    double c = 1.0, sv = 0.0;
    if (pi < m && pi != i) {
      const int p = min(i, pi), q = max(i, pi);
      const double app = Ab[(size_t)p * m + p], aqq = Ab[(size_t)q * m + q];
      const double apq = Ab[(size_t)p * m + q];
      if (fabs(apq) > 1e-300 && fabs(apq) > 1e-17 * sqrt(fabs(app * aqq))) {
        const double th = (aqq - app) / (2.0 * apq);
        const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        c = 1.0 / sqrt(t * t + 1.0);
        sv = i == p ? -t * c : t * c;
        if (i == p && blockIdx.x == 0) rotated[b] = 1;
      }
    } else {
      pi = i;
    }
    cc[i] = c;
    ss[i] = sv;
    part[i] = pi;
  }
  __syncthreads();
  const double* Qb = Q + moff[b];
  double* A2b = A2 + moff[b];
  double* Q2b = Q2 + moff[b];
  const int e1 = min(e0 + kJacTile, m * m);
  for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    const int i = e / m, j = e % m;
    const int ip = part[i], jp = part[j];
    const double ci = cc[i], si = ss[i], cj = cc[j], sj = ss[j];
    double v;
    if (ip == j && ip != i) {
      v = 0.0;  // the annihilated pair
    } else {
      v = ci * cj * Ab[(size_t)i * m + j] + ci * sj * Ab[(size_t)i * m + jp] +
          si * cj * Ab[(size_t)ip * m + j] + si * sj * Ab[(size_t)ip * m + jp];
    }
    A2b[e] = v;
    Q2b[e] = cj * Qb[e] + sj * Qb[(size_t)i * m + jp];
  }
}

// Eigenvalues = the diagonal, ranked ascending (ties by index); the k2
// smallest columns of Q go to Vt[(b*k2max + rank)*nc + dof], lambda to
// wout[b*k2max + rank].
__global__ void k_box_pick(const double* __restrict__ A, const double* __restrict__ Q, int nc,
                           const int* __restrict__ dofs, const int* __restrict__ boff,
                           const long long* __restrict__ moff, int k2max,
                           double* __restrict__ Vt, double* __restrict__ wout) {
  const int b = blockIdx.x;
  const int m = boff[b + 1] - boff[b];
  const int* d = dofs + boff[b];
  const double* Ab = A + moff[b];
  const double* Qb = Q + moff[b];
  const int k2 = min(k2max, m);
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const double wi = Ab[(size_t)i * m + i];
    int rank = 0;
    for (int j = 0; j < m; ++j) {
      const double wj = Ab[(size_t)j * m + j];
      rank += (wj < wi || (wj == wi && j < i)) ? 1 : 0;
    }
    if (rank < k2) {
      wout[b * k2max + rank] = wi;
      double* col = Vt + (size_t)(b * k2max + rank) * nc;
      for (int q = 0; q < m; ++q) col[d[q]] = Qb[(size_t)q * m + i];
    }
  }
}

int level3_setup_boxes(pmx_ctx* ctx) {
  const int n = ctx->nc, nbox = ctx->nbox, k2max = ctx->en_lv2;
  // Coarse dofs of every box, in subdomain order.
  std::vector<std::vector<int>> lists(nbox);
  for (const Brick& b : ctx->h_bricks)
    for (int j = 0; j < b.neig; ++j) lists[b.box].push_back(b.dof + j);
  std::vector<int> boff(nbox + 1, 0), dofs;
  std::vector<long long> moff(nbox + 1, 0);
  int mmax = 0;
  for (int q = 0; q < nbox; ++q) {
    const long long m = (long long)lists[q].size();
    boff[q + 1] = boff[q] + (int)m;
    dofs.insert(dofs.end(), lists[q].begin(), lists[q].end());
    moff[q + 1] = moff[q] + m * m;
    mmax = std::max(mmax, (int)m);
  }
  if (mmax > kJacMaxM) {
    set_error("-procs_virtual: a box with %d coarse unknowns (at most %d)", mmax, kJacMaxM);
    return PMX_ERR_SUP;
  }
  DBuf<double> Ain, Aj, Aj2, Qj, Qj2, Vt, wv;
  DBuf<int> d_dofs, d_boff, rot;
  DBuf<long long> d_moff;
  PMX_TRY(Ain.alloc((size_t)n * n));
  for (DBuf<double>* w : {&Aj, &Aj2, &Qj, &Qj2}) PMX_TRY(w->alloc((size_t)moff[nbox]));
  PMX_TRY(rot.alloc(nbox));
  PMX_TRY(Vt.alloc((size_t)n * nbox * k2max));
  PMX_TRY(wv.alloc((size_t)nbox * k2max));
  PMX_TRY(d_dofs.alloc(dofs.size()));
  PMX_TRY(d_boff.alloc(boff.size()));
  PMX_TRY(d_moff.alloc(moff.size()));
  PMX_HIP(hipMemcpyAsync(d_dofs.p, dofs.data(), dofs.size() * sizeof(int), hipMemcpyHostToDevice,
                         ctx->stream));
  PMX_HIP(hipMemcpyAsync(d_boff.p, boff.data(), boff.size() * sizeof(int), hipMemcpyHostToDevice,
                         ctx->stream));
  PMX_HIP(hipMemcpyAsync(d_moff.p, moff.data(), moff.size() * sizeof(long long),
                         hipMemcpyHostToDevice, ctx->stream));
  PMX_HIP(hipMemsetAsync(Ain.p, 0, sizeof(double) * (size_t)n * n, ctx->stream));
  PMX_HIP(hipMemsetAsync(Vt.p, 0, sizeof(double) * (size_t)n * nbox * k2max, ctx->stream));
  PMX_HIP(hipMemsetAsync(wv.p, 0, sizeof(double) * (size_t)nbox * k2max, ctx->stream));
  hipLaunchKernelGGL(k_galerkin, dim3(ctx->nsub), dim3(kBlock), 0, ctx->stream, ctx->geo,
                     ctx->phi.p, ctx->en_lv1, ctx->T[0].p, ctx->T[1].p, ctx->T[2].p, n, Ain.p,
                     (double*)nullptr, 1);
  const int ntile = (int)(((long long)mmax * mmax + kJacTile - 1) / kJacTile);
  hipLaunchKernelGGL(k_box_gather, dim3(std::min(ntile, 256), nbox), dim3(256), 0, ctx->stream,
                     Ain.p, n, d_dofs.p, d_boff.p, d_moff.p, Aj.p, Qj.p);
  PMX_HIP(hipGetLastError());
  // Sweeps until one leaves every box without a rotation (at most 40).
  const int rounds = ((mmax + 1) & ~1) - 1;
  int sweeps = 0;
  std::vector<int> hrot(nbox);
  for (; sweeps < 40; ++sweeps) {
    PMX_HIP(hipMemsetAsync(rot.p, 0, nbox * sizeof(int), ctx->stream));
    for (int r = 0; r < rounds; ++r) {
      hipLaunchKernelGGL(k_jacobi_round, dim3(ntile, nbox), dim3(256), 0, ctx->stream, Aj.p,
                         Aj2.p, Qj.p, Qj2.p, d_boff.p, d_moff.p, r, rot.p);
      std::swap(Aj, Aj2);
      std::swap(Qj, Qj2);
    }
    PMX_HIP(hipGetLastError());
    PMX_HIP(hipMemcpyAsync(hrot.data(), rot.p, nbox * sizeof(int), hipMemcpyDeviceToHost,
                           ctx->stream));
    PMX_HIP(hipStreamSynchronize(ctx->stream));
    if (std::all_of(hrot.begin(), hrot.end(), [](int v) { return v == 0; })) break;
  }
  if (sweeps == 40) {
    int b = 0;
    while (hrot[b] == 0) ++b;
    set_error("level-3 box eigenproblems: the Jacobi sweeps still rotate in box %d after 40 "
              "sweeps", b);
    return PMX_ERR_NOT_CONVERGED;
  }
  hipLaunchKernelGGL(k_box_pick, dim3(nbox), dim3(256), 0, ctx->stream, Aj.p, Qj.p, n, d_dofs.p,
                     d_boff.p, d_moff.p, k2max, Vt.p, wv.p);
  PMX_HIP(hipGetLastError());
  std::vector<double> hw((size_t)nbox * k2max);
  PMX_HIP(hipMemcpyAsync(hw.data(), wv.p, hw.size() * sizeof(double), hipMemcpyDeviceToHost,
                         ctx->stream));
  PMX_HIP(hipStreamSynchronize(ctx->stream));
  // Per-box count against eb_lv2 (:1100-1115); the columns of A_cc are the
  // boxes' vectors in rank order.
  ctx->h_eigen_num_lv2_box.assign(nbox, 0);
  ctx->h_eig_lv2.clear();
  int K = 0;
  double emin = 0.0, emax = 0.0;
  for (int q = 0; q < nbox; ++q) {
    const int k2 = std::min(k2max, boff[q + 1] - boff[q]);
    const double* lam = hw.data() + (size_t)q * k2max;
    int num = k2;
    for (int j = 0; j < k2; ++j)
      if (lam[j] >= ctx->eb_lv2) {
        num = j + 1;
        break;
      }
    ctx->h_eigen_num_lv2_box[q] = num;
    emin = q == 0 ? lam[0] : std::min(emin, lam[0]);
    emax = q == 0 ? lam[num - 1] : std::max(emax, lam[num - 1]);
    ctx->h_eig_lv2.insert(ctx->h_eig_lv2.end(), lam, lam + num);
    K += num;
  }
  if (K > kMaxCC) {
    set_error("-procs_virtual: %d level-3 vectors over all boxes exceed %d", K, kMaxCC);
    return PMX_ERR_SUP;
  }
  PMX_TRY(ctx->V2.alloc((size_t)n * K));
  int col = 0;
  for (int q = 0; q < nbox; ++q) {
    const int num = ctx->h_eigen_num_lv2_box[q];
    PMX_HIP(hipMemcpyAsync(ctx->V2.p + (size_t)col * n, Vt.p + (size_t)q * k2max * n,
                           (size_t)num * n * sizeof(double), hipMemcpyDeviceToDevice,
                           ctx->stream));
    col += num;
  }
  ctx->eigen_num_lv2 = K;
  ctx->eigen_min_lv2 = emin;
  ctx->eigen_max_lv2 = emax;
  ctx->lv3_iters = sweeps + 1;
  ctx->lv3_floor = 0.0;
  return lv3_finish_energy(ctx, K);
}

// Refined shift-invert step, part 1: R = A_c X - X diag(w) (Z = A_c X).
__global__ void k_lv3_refine(const double* __restrict__ Z, const double* __restrict__ X,
                             const double* __restrict__ w, double* __restrict__ R, int n, int m) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < (size_t)n * m;
       e += (size_t)gridDim.x * blockDim.x)
    R[e] = Z[e] - w[e / n] * X[e];
}

// Part 2: with D = A_c^{-1} R in Y, Y <- X - D (= diag(w) A_c^{-1} X in exact
// arithmetic, but the inverse only acts on the small residual, so its
// rounding error no longer limits the accuracy) and res[j] = ||D_j||, the
// shift-inverted residual of the Ritz pair j.  One workgroup per column.
__global__ void k_lv3_update(const double* __restrict__ X, double* __restrict__ Y, int n,
                             double* __restrict__ res) {
  __shared__ double red[8];
  const int j = blockIdx.x;
  double s = 0.0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    const double d = Y[(size_t)j * n + r];
    s += d * d;
    Y[(size_t)j * n + r] = X[(size_t)j * n + r] - d;
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) res[j] = sqrt(s);
}

// res[j] = || w_j Y_j - X_j || for column j = blockIdx.x.
__global__ void k_lv3_sires(const double* __restrict__ Y, const double* __restrict__ X,
                            const double* __restrict__ w, int n, double* __restrict__ res) {
  __shared__ double red[8];
  const int j = blockIdx.x;
  double s = 0.0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    const double e = w[j] * Y[(size_t)j * n + r] - X[(size_t)j * n + r];
    s += e * e;
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) res[j] = sqrt(s);
}

// The level-3 iterations' random start block: count standard normal values
// of the context's seed (mt19937_64 seeded with seed + 7), generated on the
// host once and kept on the device for the setups that follow (the same
// values every setup; generating them took ~10 ms of host time per setup at
// nc = 32768, m = 16).
static int lv3_random_block(pmx_ctx* ctx, size_t count, double* dst) {
  const unsigned long long seed = (unsigned long long)ctx->seed + 7;
  if (ctx->lv3_rand.n != count || ctx->lv3_rand_seed != seed || !ctx->lv3_rand.p) {
    std::vector<double> h(count);
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd;
    for (auto& v : h) v = nd(rng);
    PMX_TRY(ctx->lv3_rand.alloc(count));
    PMX_HIP(hipMemcpyAsync(ctx->lv3_rand.p, h.data(), count * sizeof(double),
                           hipMemcpyHostToDevice, ctx->stream));
    PMX_HIP(hipStreamSynchronize(ctx->stream));   // h is freed on return
    ctx->lv3_rand_seed = seed;
  }
  PMX_HIP(hipMemcpyAsync(dst, ctx->lv3_rand.p, count * sizeof(double), hipMemcpyDeviceToDevice,
                         ctx->stream));
  return PMX_SUCCESS;
}

int level3_setup(pmx_ctx* ctx) {
  const int n = ctx->nc;
  const int k2 = std::min(ctx->en_lv2, n);
  const int m = std::min(std::max(4 * k2, k2 + 8), std::min(n, 2 * kMaxEig));
  // Pure Neumann (no pin, no well term; the reference drivers' system): A_c
  // is singular with the coarse constant z as exact null vector, the first
  // eigenpair (0, z) of the level-3 problem (SLEPc finds it with its shift
  // -1e-12, :1086-1096).  The pseudo-inverse iteration below works on z-perp
  // and supplies the other k2-1 pairs.
  const bool nullz = ctx->pin_brick < 0 && !ctx->have_wd && ctx->coarse_deflate &&
                     ctx->coarse_gamma > 0.0;
  const int k2i = nullz ? k2 - 1 : k2;
  DBuf<double> X, Y, Z, W, R, w, res;
  DBuf<int> fail;
  PMX_TRY(W.alloc((size_t)n * m));
  PMX_TRY(R.alloc((size_t)n * m));
  PMX_TRY(X.alloc((size_t)n * m));
  PMX_TRY(Y.alloc((size_t)n * m));
  PMX_TRY(Z.alloc((size_t)n * m));
  PMX_TRY(w.alloc(m));
  PMX_TRY(res.alloc(m));
  PMX_TRY(fail.alloc(1));
  DBuf<double> ws;
  PMX_TRY(ws.alloc((size_t)kTallSplit * n * m));
  Lv3Work wk;
  wk.ctx = ctx;
  wk.n = n;
  wk.ngram = (n + kGramRows - 1) / kGramRows;
  PMX_TRY(wk.part.alloc((size_t)wk.ngram * kLv3Max * kLv3Max));
  PMX_TRY(wk.G.alloc(kLv3Max * kLv3Max));
  PMX_TRY(wk.T.alloc(kLv3Max * kLv3Max));
  PMX_TRY(wk.wall.alloc(kLv3Max));
  PMX_TRY(wk.flag.alloc(1));
  PMX_TRY(lv3_random_block(ctx, (size_t)n * m, X.p));
  PMX_HIP(hipMemsetAsync(fail.p, 0, sizeof(int), ctx->stream));
  const int nspmm = (ctx->nsub * ctx->en_lv1 + kBlock - 1) / kBlock;
  std::vector<double> hw(m), hres(m), hist;
  int it = 0;
  const int maxit = 500;
  ctx->lv3_floor = 0.0;
  for (it = 1; it <= maxit; ++it) {
    if (it == 1) {
      launch_gemm_tall(ctx->stream, ctx->Ainv.p, X.p, W.p, n, m, ws.p);
      hipLaunchKernelGGL(k_coarse_fix, dim3(m), dim3(kBlock), 0, ctx->stream, W.p, X.p,
                         ctx->c_z.p, ctx->c_h.p, n, ctx->c_cs.p, Y.p);
    } else {
      // Shift-invert step with residual correction: Y = X - A_c^{-1} (A_c X -
      // X diag(w)) for the Ritz pairs (X, w) and Z = A_c X of the previous
      // Rayleigh-Ritz step; res = ||w A_c^{-1} x - x||, the residual of the
      // shift-inverted problem (SLEPc's relative test under STSINVERT with
      // target ~0), computed on the small correction.
      hipLaunchKernelGGL(k_lv3_refine, dim3(std::min((n * m + kBlock - 1) / kBlock, 4096)),
                         dim3(kBlock), 0, ctx->stream, Z.p, X.p, w.p, R.p, n, m);
      launch_gemm_tall(ctx->stream, ctx->Ainv.p, R.p, W.p, n, m, ws.p);
      hipLaunchKernelGGL(k_coarse_fix, dim3(m), dim3(kBlock), 0, ctx->stream, W.p, R.p,
                         ctx->c_z.p, ctx->c_h.p, n, ctx->c_cs.p, Y.p);
      hipLaunchKernelGGL(k_lv3_update, dim3(m), dim3(kBlock), 0, ctx->stream, X.p, Y.p, n,
                         res.p);
      PMX_HIP(hipMemcpyAsync(hres.data(), res.p, k2i * sizeof(double), hipMemcpyDeviceToHost,
                             ctx->stream));
      PMX_HIP(hipStreamSynchronize(ctx->stream));
      bool conv = true;
      double mres = 0.0;
      for (int j = 0; j < k2i; ++j) {
        if (!(hres[j] <= ctx->eig_tol)) conv = false;
        mres = std::max(mres, hres[j]);
      }
      // Rounding floor: the near-null direction of the pinned coarse matrix
      // is amplified ~1/lambda_min by the inverse, which bounds the
      // attainable shift-invert residual (7e-8 on the 128^3 1e5-contrast
      // medium, lambda_min(A_c) ~ 6e-12).  A residual that stagnates (no 2x
      // improvement over the last 10 iterations) within -eig_stag x
      // -eig_tol (100 x 1e-8 by default, the bound of the subdomain
      // eigensolver and of the sparse coarse path) is accepted and reported
      // in lv3_floor; above it the iteration runs on to its cap and fails
      // as SLEPc's nconv < nev does (:1098).
      hist.push_back(mres);
      if (!conv && (int)hist.size() > kLv3StagWin && mres <= ctx->eig_stag * ctx->eig_tol) {
        double best_old = 1e300;
        for (size_t q = 0; q + kLv3StagWin < hist.size(); ++q)
          best_old = std::min(best_old, hist[q]);
        if (mres > 0.5 * best_old) {
          conv = true;
          ctx->lv3_floor = mres;
        }
      }
      if (ctx->eig_debug >= 0 && (it < 20 || it % 25 == 0)) {
        PMX_HIP(hipMemcpy(hw.data(), w.p, m * sizeof(double), hipMemcpyDeviceToHost));
        printf("lv3 it %d res", it);
        for (int j = 0; j < k2i; ++j) printf(" %.3e", hres[j]);
        printf(" w");
        for (int j = 0; j < m; ++j) printf(" %.4e", hw[j]);
        printf("\n");
      }
      if (conv) break;
    }
    // Rayleigh-Ritz: Y <- Y L^{-T} (Cholesky QR of Y^T Y), Z = A_c Y,
    // Y^T Z = Q diag(w) Q^T, X = Y Q, Z <- Z Q.  W is free here (scratch).
    PMX_TRY(wk.gram(Y.p, m, Y.p, m));
    hipLaunchKernelGGL(k_lv3_small, dim3(1), dim3(256), 0, ctx->stream, wk.G.p, m, 0, wk.T.p,
                       wk.wall.p, fail.p);
    PMX_TRY(wk.xform(Y.p, m, m, 0, 0, m, W.p));
    std::swap(Y, W);
    hipLaunchKernelGGL(k_coarse_spmm, dim3(nspmm, m), dim3(kBlock), 0, ctx->stream, ctx->geo,
                       ctx->Acb.p, ctx->en_lv1, Y.p, Z.p, n);
    PMX_TRY(wk.gram(Y.p, m, Z.p, m));
    hipLaunchKernelGGL(k_lv3_small, dim3(1), dim3(256), 0, ctx->stream, wk.G.p, m, 1, wk.T.p,
                       w.p, fail.p);
    PMX_TRY(wk.xform(Y.p, m, m, 0, 0, m, X.p));
    PMX_TRY(wk.xform(Z.p, m, m, 0, 0, m, W.p));
    std::swap(Z, W);
    PMX_HIP(hipGetLastError());
  }
  PMX_HIP(hipMemcpyAsync(hw.data(), w.p, m * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  PMX_HIP(hipStreamSynchronize(ctx->stream));
  int hfail = 0;
  PMX_HIP(hipMemcpy(&hfail, fail.p, sizeof(int), hipMemcpyDeviceToHost));
  if (it > maxit || hfail) {
    set_error("level-3 eigensolver did not converge (%d iterations, fail=%d)", it, hfail);
    return PMX_ERR_USER;
  }
  ctx->lv3_iters = it;
  if (nullz) hw.insert(hw.begin(), 0.0);
  // Adaptive count against eb_lv2 (src/PreMixFEM_3D.c:1100-1115).
  int num = k2;
  for (int j = 0; j < k2; ++j)
    if (hw[j] >= ctx->eb_lv2) {
      num = j + 1;
      break;
    }
  ctx->eigen_num_lv2 = num;
  ctx->eigen_min_lv2 = hw[0];
  ctx->eigen_max_lv2 = hw[num - 1];
  ctx->h_eig_lv2.assign(hw.begin(), hw.begin() + k2);
  PMX_TRY(ctx->V2.alloc((size_t)n * num));
  if (nullz) {
    PMX_HIP(hipMemcpyAsync(ctx->V2.p, ctx->c_z.p, (size_t)n * sizeof(double),
                           hipMemcpyDeviceToDevice, ctx->stream));
    if (num > 1)
      PMX_HIP(hipMemcpyAsync(ctx->V2.p + n, X.p, (size_t)n * (num - 1) * sizeof(double),
                             hipMemcpyDeviceToDevice, ctx->stream));
  } else {
    PMX_HIP(hipMemcpyAsync(ctx->V2.p, X.p, (size_t)n * num * sizeof(double),
                           hipMemcpyDeviceToDevice, ctx->stream));
  }
  PMX_TRY(lv3_finish_energy(ctx, num));
  return PMX_SUCCESS;
}

// ---------------------------------------------------------------------------
// Level 3 of the spe10 fork (src/spe10.c:1255-1417) without a dense
// factorisation: the smallest en_lv2 eigenpairs of A_c by block LOBPCG with
// the level-2 ICC(0) factor M as preconditioner.  Blocks are column-major
// n x m; the Gram matrices are summed by many workgroups (partials in fixed
// order), the ns x ns Cholesky / eigenproblems run in one workgroup.
// ---------------------------------------------------------------------------

// Block LOBPCG (Knyazev 2001) on A_c with the ICC(0) preconditioner:
// S = [X, M^{-1}(A_c X - X diag(w)), P], S orthonormalised by Cholesky QR
// (P dropped when that fails), Rayleigh-Ritz on S^T A_c S, X = S Q_{:,1:m},
// P = S_{:,m+1:} Q_{m+1:,1:m}.  m >= en_lv2 + 4 vectors; converged when
// ||A_c x - w x|| <= eig_tol |w| for the en_lv2 wanted pairs (x of unit
// norm); a residual that stagnates (no 2x gain over 20 iterations) within
// eig_stag x eig_tol is accepted and reported in lv3_floor.  Warm start: the
// previous setup's eigenvectors when the coarse space has the same size.
// With the exact band factor instead of ICC(0) the same loop runs the
// residual-corrected shift-invert subspace iteration of the dense path
// (basis X - A_c^{-1}(A_c X - X diag(w)) only, m = max(4 en_lv2, en_lv2 + 8))
// and tests the shift-inverted residual ||x - w A_c^{-1} x||, whose
// stagnation window is kLv3StagWin iterations.
int level3_setup_icc(pmx_ctx* ctx) {
  const int n = ctx->nc;
  const int k2 = std::min(ctx->en_lv2, n);
  const int m = ctx->use_band ? std::min(std::min(std::max(4 * k2, k2 + 8), kLv3Max), n)
                             : std::min(std::min(k2 + 6, kLv3Max / 3), n);
  if (m < k2) {
    set_error("sparse coarse path: en_lv2=%d exceeds the level-3 block limit", k2);
    return PMX_ERR_SUP;
  }
  Lv3Work wk;
  wk.ctx = ctx;
  wk.n = n;
  wk.ngram = (n + kGramRows - 1) / kGramRows;
  PMX_TRY(wk.part.alloc((size_t)wk.ngram * kLv3Max * kLv3Max));
  PMX_TRY(wk.G.alloc(kLv3Max * kLv3Max));
  PMX_TRY(wk.T.alloc(kLv3Max * kLv3Max));
  PMX_TRY(wk.wall.alloc(kLv3Max));
  PMX_TRY(wk.flag.alloc(1));
  DBuf<double> S, Z, tmp, X, AX, P, R, res;
  for (DBuf<double>* b : {&S, &Z, &tmp}) PMX_TRY(b->alloc((size_t)n * 3 * m));
  for (DBuf<double>* b : {&X, &AX, &P, &R}) PMX_TRY(b->alloc((size_t)n * m));
  PMX_TRY(res.alloc(m));
  const int nspmm = (ctx->nsub * ctx->en_lv1 + kBlock - 1) / kBlock;
  const size_t colb = (size_t)n * sizeof(double);

  // Initial block: previous eigenvectors (if any) and M^{-1} of random columns.
  PMX_TRY(lv3_random_block(ctx, (size_t)n * m, R.p));
  PMX_TRY(launch_coarse_prec(ctx, R.p, S.p, m));
  const bool warm = ctx->lv3_warm && ctx->V2.p && ctx->lv3_prev_n == n && ctx->eigen_num_lv2 > 0;
  if (warm)
    PMX_HIP(hipMemcpyAsync(S.p, ctx->V2.p, (size_t)ctx->eigen_num_lv2 * colb,
                           hipMemcpyDeviceToDevice, ctx->stream));
  std::vector<double> hw(m), hres(m), hist;
  int it = 0, np = 0, ns = m;
  bool conv = false;
  ctx->lv3_floor = 0.0;
  const int maxit = std::max(ctx->eig_maxit, 2000);
  for (it = 0; it <= maxit; ++it) {
    if (it > 0 && ctx->use_band) {
      // Exact factor: residual-corrected shift-invert subspace iteration,
      // S = X - A_c^{-1}(A_c X - X diag(w)) = diag(w) A_c^{-1} X.
      PMX_HIP(hipMemcpyAsync(S.p, X.p, m * colb, hipMemcpyDeviceToDevice, ctx->stream));
      hipLaunchKernelGGL(k_axpy, dim3(std::min((n * m + kBlock - 1) / kBlock, 4096)),
                         dim3(kBlock), 0, ctx->stream, S.p + (size_t)m * n, -1.0, S.p,
                         (size_t)n * m);
      ns = m;
    } else if (it > 0) {
      // X into S[:, 0:m], P into S[:, 2m:3m] (W = M^{-1} R is in S[:, m:2m]).
      PMX_HIP(hipMemcpyAsync(S.p, X.p, m * colb, hipMemcpyDeviceToDevice, ctx->stream));
      if (np) PMX_HIP(hipMemcpyAsync(S.p + (size_t)2 * m * n, P.p, m * colb,
                                     hipMemcpyDeviceToDevice, ctx->stream));
      ns = np ? 3 * m : 2 * m;
    }
    int bad = 0;
    PMX_TRY(wk.orthonormalize(S.p, ns, tmp.p, &bad));
    if (bad && ns == 3 * m) {
      // restart without P (S is unchanged when the Cholesky fails)
      ns = 2 * m;
      np = 0;
      PMX_TRY(wk.orthonormalize(S.p, ns, tmp.p, &bad));
    }
    if (bad && !hist.empty() && hist.back() <= ctx->eig_stag * ctx->eig_tol) {
      // The new directions are rounding noise of converged pairs: accept.
      conv = true;
      ctx->lv3_floor = hist.back();
      break;
    }
    if (bad) {
      set_error("level-3 eigensolver (sparse coarse path): basis lost rank at iteration %d",
                it);
      return PMX_ERR_USER;
    }
    // Rayleigh-Ritz on S^T A_c S.
    hipLaunchKernelGGL(k_coarse_spmm, dim3(nspmm, ns), dim3(kBlock), 0, ctx->stream, ctx->geo,
                       ctx->Acb.p, ctx->en_lv1, S.p, Z.p, n);
    PMX_TRY(wk.gram(S.p, ns, Z.p, ns));
    hipLaunchKernelGGL(k_lv3_small, dim3(1), dim3(256), 0, ctx->stream, wk.G.p, ns, 1, wk.T.p,
                       wk.wall.p, wk.flag.p);
    PMX_HIP(hipGetLastError());
    PMX_TRY(wk.xform(S.p, ns, ns, 0, 0, m, X.p));
    PMX_TRY(wk.xform(Z.p, ns, ns, 0, 0, m, AX.p));
    if (ns > m) {
      PMX_TRY(wk.xform(S.p + (size_t)m * n, ns - m, ns, m, 0, m, P.p));
      np = 1;
    }
    hipLaunchKernelGGL(k_col_resid, dim3(m), dim3(kBlock), 0, ctx->stream, AX.p, X.p, wk.wall.p,
                       R.p, n, res.p);
    // Next search directions W = M^{-1} R.  With the exact band factor
    // ||W_j|| = ||x_j - w_j A_c^{-1} x_j|| is the shift-inverted residual
    // (scale-free, also for the pinned near-null pair): the convergence test.
    double* W = S.p + (size_t)m * n;
    PMX_TRY(launch_coarse_prec(ctx, R.p, W, m));
    if (ctx->use_band)
      hipLaunchKernelGGL(k_col_norm, dim3(m), dim3(kBlock), 0, ctx->stream, W, n, res.p);
    if (!ctx->use_band) {
      // W -= X (X^T W): the new directions orthogonal to the Ritz vectors.
      PMX_TRY(wk.gram(X.p, m, W, m));
      PMX_TRY(wk.xform(X.p, m, m, 0, 0, m, W, wk.G.p, -1.0, 1.0));
    }
    PMX_HIP(hipGetLastError());
    PMX_HIP(hipMemcpyAsync(hres.data(), res.p, m * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
    PMX_HIP(hipMemcpyAsync(hw.data(), wk.wall.p, m * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
    PMX_HIP(hipStreamSynchronize(ctx->stream));
    double mres = 0.0;
    for (int j = 0; j < k2; ++j)
      mres = std::max(mres, ctx->use_band ? hres[j]
                                          : hres[j] / std::max(std::fabs(hw[j]), 1e-300));
    hist.push_back(mres);
    if (ctx->eig_debug >= 0 && (it < 20 || it % 50 == 0))
      printf("lv3 icc it %d ns %d res %.3e w0 %.6e\n", it, ns, mres, hw[0]);
    if (mres <= ctx->eig_tol) {
      conv = true;
      break;
    }
    // Stagnation within -eig_stag x -eig_tol: no 2x gain over the window
    // (kLv3StagWin iterations with the exact factor, 20 for the ICC(0)
    // LOBPCG, whose progress is less regular).
    const int win = ctx->use_band ? kLv3StagWin : 20;
    if ((int)hist.size() > (ctx->use_band ? win : 2 * win) &&
        mres <= ctx->eig_stag * ctx->eig_tol) {
      double best_old = 1e300;
      for (size_t q = 0; q + win < hist.size(); ++q) best_old = std::min(best_old, hist[q]);
      if (mres > 0.5 * best_old) {
        conv = true;
        ctx->lv3_floor = mres;
        break;
      }
    }
  }
  if (!conv) {
    set_error("level-3 eigensolver (sparse coarse path) did not converge (%d iterations)",
              it);
    return PMX_ERR_USER;
  }
  ctx->lv3_iters = it;
  int num = k2;
  for (int j = 0; j < k2; ++j)
    if (hw[j] >= ctx->eb_lv2) {
      num = j + 1;
      break;
    }
  ctx->eigen_num_lv2 = num;
  ctx->eigen_min_lv2 = hw[0];
  ctx->eigen_max_lv2 = hw[num - 1];
  ctx->h_eig_lv2.assign(hw.begin(), hw.begin() + k2);
  PMX_TRY(ctx->V2.alloc((size_t)n * num));
  PMX_HIP(hipMemcpyAsync(ctx->V2.p, X.p, (size_t)n * num * sizeof(double), hipMemcpyDeviceToDevice,
                         ctx->stream));
  PMX_TRY(lv3_finish_energy(ctx, num));
  ctx->lv3_prev_n = n;
  return PMX_SUCCESS;
}

}  // namespace pmx
