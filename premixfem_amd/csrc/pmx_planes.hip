// Exact level-2 coarse solve for large coarse spaces by z-planes of
// subdomains (the reference factors A_c with MUMPS, src/PreMixFEM_3D.c:894-906).
//
// In the natural subdomain order (x fastest, src/PreMixFEM_3D.c:23-27) with
// the eigen index inner (:40-45), the dofs of one z-plane of subdomains are
// contiguous and A_c is block tridiagonal over the planes: diagonal blocks
// D_z (the couplings inside plane z, block 5-point) and C_z = A_{z,z-1},
// whose only non-zero k x k blocks couple a subdomain with the one below it.
// Block LDL^T with explicit inverses of the Schur complements:
//   S_0 = D_0,  S_z = D_z - C_z G_{z-1} C_z^T,  G_z = S_z^{-1},
// and A_c u = t becomes
//   forward   w_0 = t_0,  w_{z+1} = t_{z+1} - C_{z+1} (G_z w_z),
//   backward  u_Z = G_Z w_Z,  u_z = G_z (w_z - C_{z+1}^T u_{z+1}).
// A solve is 2Z-1 dense GEMVs of n_p x n_p (n_p = dofs of a plane), each
// spread over the whole device with one workgroup per subdomain of the
// plane, instead of one small launch per 64-wide tile column of a band
// factor.  The C blocks are read from the block rows Acb (slots -z / +z).
// With process boxes the factored matrix keeps only the couplings inside a
// box (_PC_setup_lv2_J, src/PreMixFEM_3D.c:806-888): blocks between
// subdomains of two boxes are skipped here, while Acb keeps them for the
// coarse residual updates.
// At 256^3 cells with 16^3 subdomains of 16^3 cells and k = 8: 16 planes of
// n_p = 2048, G_z 33.5 MB each (0.54 GB), 31 GEMVs per solve.
#include <cmath>
#include <vector>

#include "pmx_kernels.h"

namespace pmx {

typedef double pl_d4 __attribute__((ext_vector_type(4)));   // an MFMA accumulator

constexpr int kPlMaxCols = 32;    // right-hand sides of one solve (level-3 set-up blocks)

// One plane step for the right-hand sides c0 = NCG * blockIdx.y .. c0+NCG-1
// (vectors of stride nc) and subdomain s = sub0 + blockIdx.x of plane z, one
// workgroup of NT threads per subdomain:
//   out = G_z w_z for the rows of s, then
//   fwd:  w[s + z-neighbour] -= A_c(s_up, s) out           (slot -z of s_up)
//   bwd:  d[s] = out;  w[s - z-neighbour] -= A_c(s_dn, s) out  (slot +z of s_dn)
// The workgroup is wide (NT = 1024 for one right-hand side: at n_p = 2048 every
// thread issues all of its row loads at once), so one workgroup per CU keeps
// enough loads in flight to stream G_z near HBM speed with no exchange between
// workgroups.  Several right-hand sides share every load of G_z: NCG per
// thread, and NGRP groups of threads reading the same G rows in step (the
// level-3 set-up's 16-column block solves stream G_z once, not four times).
// Every
// subdomain of plane z has at most one neighbour in the next plane and no two
// share it, so the in-place updates of w do not collide.
template <int KR, int NCG, int NT, int NGRP = 1>
__global__ void __launch_bounds__(NT) k_pl_step(Geo g, const double* __restrict__ Acb, int kmax,
                                                const double* __restrict__ G, int np, int dof0,
                                                int sub0, int fwd, double* __restrict__ w,
                                                double* __restrict__ d, int nc, int ncol,
                                                const double* stop) {
  PMX_STOP_CHECK(stop);
  constexpr int NW = NT / 64;
  constexpr int TPG = NT / NGRP;           // threads per group of NCG right-hand sides
  static_assert(TPG % 64 == 0, "groups of whole waves");
  constexpr int NO = NGRP * NCG * KR;      // outputs of the workgroup
  __shared__ double red[NW][NCG * KR];
  __shared__ double outv[NO];
  const int s = sub0 + blockIdx.x;
  const Brick b = g.bricks[s];
  const int nr = b.neig, r0 = b.dof - dof0;
  // The NGRP groups take consecutive blocks of NCG right-hand sides and read
  // the same G rows at the same time, so HBM streams G_z once for all of them.
  const int grp = threadIdx.x / TPG, tg = threadIdx.x % TPG;
  const int cb = blockIdx.y * NGRP * NCG;  // first right-hand side of the workgroup
  const int c0 = cb + grp * NCG;
  const int ncg = min(NCG, ncol - c0);     // <= 0: an idle group
  double acc[NCG][KR];
#pragma unroll
  for (int c = 0; c < NCG; ++c)
#pragma unroll
    for (int i = 0; i < KR; ++i) acc[c][i] = 0.0;
  const double* Gr = G + (size_t)r0 * np;
  const double* x[NCG];
#pragma unroll
  for (int c = 0; c < NCG; ++c)
    x[c] = w + (size_t)(ncg > 0 ? c0 + min(c, ncg - 1) : 0) * nc + dof0;
  if (((np | dof0 | nc) & 1) == 0) {  // 16-byte aligned rows and segments
    const int n2 = np >> 1;
    for (int j = tg; j < n2; j += TPG) {
      double2 xv[NCG];
#pragma unroll
      for (int c = 0; c < NCG; ++c) xv[c] = reinterpret_cast<const double2*>(x[c])[j];
#pragma unroll
      for (int i = 0; i < KR; ++i) {
        if (KR == 8 || i < nr) {  // every brick of a KR = 8 instance has 8 rows
          const double2 gv = reinterpret_cast<const double2*>(Gr + (size_t)i * np)[j];
#pragma unroll
          for (int c = 0; c < NCG; ++c) acc[c][i] += gv.x * xv[c].x + gv.y * xv[c].y;
        }
      }
    }
  } else {
    for (int j = tg; j < np; j += TPG) {
      double xv[NCG];
#pragma unroll
      for (int c = 0; c < NCG; ++c) xv[c] = x[c][j];
#pragma unroll
      for (int i = 0; i < KR; ++i) {
        if (KR == 8 || i < nr) {
          const double gv = Gr[(size_t)i * np + j];
#pragma unroll
          for (int c = 0; c < NCG; ++c) acc[c][i] += gv * xv[c];
        }
      }
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < NCG; ++c)
#pragma unroll
    for (int i = 0; i < KR; ++i) {
      double v = acc[c][i];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      if (lane == 0) red[wv][c * KR + i] = v;
    }
  __syncthreads();
  // outv[q * NCG * KR + e]: group q's sum over its waves, in wave order
  constexpr int WPG = TPG / 64;
  for (int t = threadIdx.x; t < NO; t += NT) {
    const int q = t / (NCG * KR), e = t % (NCG * KR);
    double v = 0.0;
    for (int u = 0; u < WPG; ++u) v += red[q * WPG + u][e];
    outv[t] = v;
  }
  __syncthreads();
  // process boxes: A_c keeps no coupling between subdomains of two boxes
  int nbi = b.nb[fwd ? 5 : 4];
  if (nbi >= 0 && g.bricks[nbi].box != b.box) nbi = -1;
  const Brick o = nbi >= 0 ? g.bricks[nbi] : b;
  const int slot_c = fwd ? 5 : 6;  // the block of o that couples it with s
  const double* blk = nbi >= 0 ? Acb + ((size_t)nbi * 7 + slot_c) * kmax * kmax : nullptr;
  const int nout = min(NGRP * NCG, ncol - cb) * KR;
  for (int t = threadIdx.x; t < nout; t += NT) {
    const int c = t / KR, i = t % KR;
    const size_t co = (size_t)(cb + c) * nc;
    if (!fwd && i < nr) d[co + b.dof + i] = outv[c * KR + i];
    if (nbi >= 0 && i < o.neig) {
      double v = 0.0;
      for (int j = 0; j < nr; ++j) v += blk[i * kmax + j] * outv[c * KR + j];
      w[co + o.dof + i] -= v;
    }
  }
}

// D_z of plane z into S (n_p x n_p row-major, zeroed before): the self block
// and the blocks of the in-plane neighbours -x, +x, -y, +y of every
// subdomain of the plane.
__global__ void k_pl_gather(Geo g, const double* __restrict__ Acb, int kmax, int sub0, int dof0,
                            int np, double* __restrict__ S) {
  const int s = sub0 + blockIdx.x;
  const Brick b = g.bricks[s];
  const size_t kk = (size_t)kmax * kmax;
  for (int e = threadIdx.x; e < 5 * kmax * kmax; e += blockDim.x) {
    const int slot = e / (kmax * kmax), i = (e / kmax) % kmax, j = e % kmax;
    const int o = slot == 0 ? s : b.nb[slot - 1];
    if (o < 0) continue;
    const Brick ob = g.bricks[o];
    if (ob.box != b.box) continue;   // process boxes: no coupling between boxes
    if (i >= b.neig || j >= ob.neig) continue;
    S[(size_t)(b.dof - dof0 + i) * np + (ob.dof - dof0 + j)] =
        Acb[(size_t)s * 7 * kk + slot * kk + i * kmax + j];
  }
}

// H = C_z G_{z-1}: the rows of subdomain s of plane z are the rows of its
// lower neighbour's block of G_{z-1} combined by A_c(s, s_dn).
__global__ void k_pl_cg(Geo g, const double* __restrict__ Acb, int kmax, int sub0, int dof0,
                        const double* __restrict__ Gp, int npp, int dof0p,
                        double* __restrict__ H) {
  const int s = sub0 + blockIdx.x;
  const Brick b = g.bricks[s];
  const int sd = b.nb[4];
  if (sd < 0) return;
  const Brick bd = g.bricks[sd];
  if (bd.box != b.box) {
    // process boxes: no coupling with the lower neighbour, a zero row block
    for (int col = threadIdx.x; col < npp; col += blockDim.x)
      for (int i = 0; i < b.neig; ++i) H[(size_t)(b.dof - dof0 + i) * npp + col] = 0.0;
    return;
  }
  const double* blk = Acb + ((size_t)s * 7 + 5) * kmax * kmax;
  for (int col = threadIdx.x; col < npp; col += blockDim.x) {
    double gv[kMaxEig];
    for (int a = 0; a < bd.neig; ++a) gv[a] = Gp[(size_t)(bd.dof - dof0p + a) * npp + col];
    for (int i = 0; i < b.neig; ++i) {
      double v = 0.0;
      for (int a = 0; a < bd.neig; ++a) v += blk[i * kmax + a] * gv[a];
      H[(size_t)(b.dof - dof0 + i) * npp + col] = v;
    }
  }
}

// S_z -= H C_z^T: column j of subdomain s' takes sum_b H[i][s'_dn, b] A_c(s', s'_dn)[j][b].
__global__ void k_pl_schur(Geo g, const double* __restrict__ Acb, int kmax, int sub0, int dof0,
                           int np, const double* __restrict__ H, int npp, int dof0p,
                           double* __restrict__ S) {
  const int sp = sub0 + blockIdx.x;
  const Brick b = g.bricks[sp];
  const int sd = b.nb[4];
  if (sd < 0) return;
  const Brick bd = g.bricks[sd];
  if (bd.box != b.box) return;   // process boxes: no coupling between boxes
  const double* blk = Acb + ((size_t)sp * 7 + 5) * kmax * kmax;
  for (int i = blockIdx.y * blockDim.x + threadIdx.x; i < np; i += gridDim.y * blockDim.x) {
    double hv[kMaxEig];
    for (int q = 0; q < bd.neig; ++q) hv[q] = H[(size_t)i * npp + (bd.dof - dof0p + q)];
    for (int j = 0; j < b.neig; ++j) {
      double v = 0.0;
      for (int q = 0; q < bd.neig; ++q) v += hv[q] * blk[j * kmax + q];
      S[(size_t)i * np + (b.dof - dof0 + j)] -= v;
    }
  }
}

// A = (A + A^T) / 2 for a dense n x n matrix; fail if a diagonal entry is
// not positive and finite (the SPD matrices and inverses of this path).
__global__ void k_pl_symmetrize(double* __restrict__ A, int n, int* __restrict__ fail,
                                int fail_id) {
  const size_t nn = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nn;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n), j = (int)(e % n);
    if (j > i) {
      const double v = 0.5 * (A[e] + A[(size_t)j * n + i]);
      A[e] = v;
      A[(size_t)j * n + i] = v;
    } else if (i == j) {
      const double v = A[e];
      if (!(v > 0.0) || !(v < INFINITY)) atomicMax(fail, fail_id);
    }
  }
}

int planes_setup(pmx_ctx* ctx) {
  const int L = ctx->sd[0] * ctx->sd[1], Z = ctx->sd[2];
  ctx->pl_dof.assign(Z + 1, 0);
  ctx->pl_goff.assign(Z + 1, 0);
  for (int z = 0; z < Z; ++z) {
    ctx->pl_dof[z + 1] = ctx->h_dof[(size_t)(z + 1) * L];
    const size_t np = ctx->pl_dof[z + 1] - ctx->pl_dof[z];
    ctx->pl_goff[z + 1] = ctx->pl_goff[z] + np * np;
  }
  int npmax = 0;
  for (int z = 0; z < Z; ++z) npmax = std::max(npmax, ctx->pl_dof[z + 1] - ctx->pl_dof[z]);
  ctx->pl_bytes = ctx->pl_goff[Z] * sizeof(double);
  PMX_TRY(ctx->pl_G.alloc(ctx->pl_goff[Z]));
  // the solves' work vectors (up to 32 right-hand sides) now, so a solve
  // never allocates
  PMX_TRY(ctx->pl_w.alloc((size_t)ctx->nc * kPlMaxCols));
  DBuf<double> S, H;
  DBuf<int> fail;
  PMX_TRY(S.alloc((size_t)npmax * npmax));
  PMX_TRY(H.alloc((size_t)npmax * npmax));
  PMX_TRY(fail.alloc(1));
  PMX_HIP(hipMemsetAsync(fail.p, 0, sizeof(int), ctx->stream));
  const int kmax = ctx->en_lv1;
  for (int z = 0; z < Z; ++z) {
    const int dof0 = ctx->pl_dof[z], np = ctx->pl_dof[z + 1] - dof0;
    PMX_HIP(hipMemsetAsync(S.p, 0, sizeof(double) * (size_t)np * np, ctx->stream));
    hipLaunchKernelGGL(k_pl_gather, dim3(L), dim3(256), 0, ctx->stream, ctx->geo, ctx->Acb.p,
                       kmax, z * L, dof0, np, S.p);
    if (z > 0) {
      const int dof0p = ctx->pl_dof[z - 1], npp = dof0 - dof0p;
      const double* Gp = ctx->pl_G.p + ctx->pl_goff[z - 1];
      hipLaunchKernelGGL(k_pl_cg, dim3(L), dim3(256), 0, ctx->stream, ctx->geo, ctx->Acb.p, kmax,
                         z * L, dof0, Gp, npp, dof0p, H.p);
      hipLaunchKernelGGL(k_pl_schur, dim3(L, (np + 255) / 256), dim3(256), 0, ctx->stream,
                         ctx->geo, ctx->Acb.p, kmax, z * L, dof0, np, H.p, npp, dof0p, S.p);
    }
    hipLaunchKernelGGL(k_pl_symmetrize, dim3(1024), dim3(256), 0, ctx->stream, S.p, np, fail.p,
                       z + 1);
    PMX_HIP(hipGetLastError());
    double* Gz = ctx->pl_G.p + ctx->pl_goff[z];
    PMX_TRY(spd_invert(ctx->stream, np, S.p, Gz, 0.0));
    hipLaunchKernelGGL(k_pl_symmetrize, dim3(1024), dim3(256), 0, ctx->stream, Gz, np, fail.p,
                       Z + z + 1);
    PMX_HIP(hipGetLastError());
  }
  int hf = 0;
  PMX_HIP(hipMemcpyAsync(&hf, fail.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  PMX_HIP(hipStreamSynchronize(ctx->stream));
  if (hf) {
    const bool inv = hf > Z;
    set_error("coarse solve by planes: %s of plane %d has a non-positive or non-finite diagonal",
              inv ? "the inverse Schur complement" : "the Schur complement",
              (inv ? hf - Z : hf) - 1);
    return PMX_ERR_NOT_CONVERGED;
  }
  return PMX_SUCCESS;
}

// The same plane step for up to 16 right-hand sides on the FP64 matrix
// cores, for bricks of 8 eigenvectors: a workgroup of 4 waves takes two
// subdomains of the plane (16 rows of G_z) and 16 right-hand sides, C(16 x
// 16) = G_z[rows, :] W[:, cols] by v_mfma_f64_16x16x4f64 with the n_p
// columns of G_z split over the waves.  Lane l feeds row l % 16 and column
// l % 16 with four consecutive k of one 16-k block (k = 16 b + 4 (l / 16) +
// j for the j-th product), so each operand is one 32-byte load; the waves'
// partial tiles are added in wave order.  Rows of an absent second
// subdomain and columns beyond ncol are computed on valid data and dropped.
// Requires n_p % 16 == 0 (the plane's column blocks) and 32-byte aligned
// plane and column offsets (multiples of 8 dofs).
__global__ void __launch_bounds__(256) k_pl_step_mfma(Geo g, const double* __restrict__ Acb,
                                                      int kmax, const double* __restrict__ G,
                                                      int np, int dof0, int sub0, int L, int fwd,
                                                      double* __restrict__ w,
                                                      double* __restrict__ d, int nc, int ncol,
                                                      const double* stop) {
  PMX_STOP_CHECK(stop);
  __shared__ double red[4][4][64];    // [wave][item][lane]
  __shared__ double outv[2][16][8];   // [subdomain][right-hand side][row]
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = lane & 15, g4 = lane >> 4;
  const int sA = sub0 + 2 * (int)blockIdx.x;
  const bool has2 = 2 * (int)blockIdx.x + 1 < L;
  const int cb = (int)blockIdx.y * 16;
  // this lane's row of G_z (rows 8..15: the second subdomain, or the first
  // one's again when there is none) and right-hand side
  const int sr = (li < 8 || !has2) ? sA : sA + 1;
  const int rowG = g.bricks[sr].dof - dof0 + (li & 7);
  const bool rowok = li < 8 || has2;
  const int col = cb + li < ncol ? cb + li : cb;
  const double* Gr = G + (size_t)rowG * np;
  const double* Wc = w + (size_t)col * nc + dof0;
  const int nkb = np >> 4, per = (nkb + 3) >> 2;
  const int kb0 = wv * per, kb1 = min(nkb, kb0 + per);
  pl_d4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int kb = kb0; kb < kb1; ++kb) {
    const int k = (kb << 4) + 4 * g4;
    const double4 a = *reinterpret_cast<const double4*>(Gr + k);
    const double4 b = *reinterpret_cast<const double4*>(Wc + k);
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(rowok ? a.x : 0.0, b.x, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(rowok ? a.y : 0.0, b.y, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(rowok ? a.z : 0.0, b.z, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(rowok ? a.w : 0.0, b.w, acc, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) red[wv][v][lane] = acc[v];
  __syncthreads();
  if (wv == 0) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double x = ((red[0][v][lane] + red[1][v][lane]) + red[2][v][lane]) + red[3][v][lane];
      const int r16 = g4 + 4 * v;   // item v of lane l: row l / 16 + 4 v, column l % 16
      outv[r16 >> 3][li][r16 & 7] = x;
    }
  }
  __syncthreads();
  // epilogue as k_pl_step: thread t -> subdomain t / 128, right-hand side
  // (t / 8) % 16, row t % 8
  const int t = threadIdx.x;
  const int half = t >> 7, c = (t >> 3) & 15, i = t & 7;
  if (half == 1 && !has2) return;
  if (cb + c >= ncol) return;
  const Brick b = g.bricks[sA + half];
  const int nr = b.neig;
  int nbi = b.nb[fwd ? 5 : 4];
  if (nbi >= 0 && g.bricks[nbi].box != b.box) nbi = -1;
  const size_t co = (size_t)(cb + c) * nc;
  if (!fwd && i < nr) d[co + b.dof + i] = outv[half][c][i];
  if (nbi >= 0) {
    const Brick o = g.bricks[nbi];
    if (i < o.neig) {
      const int slot_c = fwd ? 5 : 6;
      const double* blk = Acb + ((size_t)nbi * 7 + slot_c) * kmax * kmax;
      double v = 0.0;
      for (int j = 0; j < nr; ++j) v += blk[i * kmax + j] * outv[half][c][j];
      w[co + o.dof + i] -= v;
    }
  }
}

template <int KR, int NCG, int NT, int NGRP = 1>
static void pl_step(pmx_ctx* ctx, int z, int fwd, double* w, double* d, int ncol) {
  const int L = ctx->sd[0] * ctx->sd[1];
  const int dof0 = ctx->pl_dof[z], np = ctx->pl_dof[z + 1] - dof0;
  constexpr int per = NCG * NGRP;
  hipLaunchKernelGGL((k_pl_step<KR, NCG, NT, NGRP>), dim3(L, (ncol + per - 1) / per), dim3(NT), 0,
                     ctx->stream, ctx->geo, ctx->Acb.p, ctx->en_lv1,
                     ctx->pl_G.p + ctx->pl_goff[z], np, dof0, z * L, fwd, w, d, ctx->nc, ncol,
                     ctx->geo.stop);
}

int launch_planes_solve(pmx_ctx* ctx, const double* t, double* d, int ncol) {
  const int Z = ctx->sd[2], nc = ctx->nc;
  if (ncol > kPlMaxCols) {
    set_error("coarse solve by planes: %d right-hand sides (at most %d)", ncol, kPlMaxCols);
    return PMX_ERR_SUP;
  }
  double* w = ctx->pl_w.p;
  PMX_HIP(hipMemcpyAsync(w, t, (size_t)nc * ncol * sizeof(double), hipMemcpyDeviceToDevice,
                         ctx->stream));
  // bricks of exactly 8 eigenvectors take the unguarded 8-row instances
  const bool k8 = ctx->en_lv1 == 8 && nc == 8 * ctx->nsub;
  auto step = [&](int z, int fwd) {
    if (ncol == 1) {
      if (k8) pl_step<8, 1, 1024>(ctx, z, fwd, w, d, ncol);
      else pl_step<kMaxEig, 1, 1024>(ctx, z, fwd, w, d, ncol);
    } else if (ncol <= 4) {
      if (k8) pl_step<8, 4, 512>(ctx, z, fwd, w, d, ncol);
      else pl_step<kMaxEig, 2, 512>(ctx, z, fwd, w, d, ncol);
    } else {
      const int L = ctx->sd[0] * ctx->sd[1];
      const int dof0 = ctx->pl_dof[z], np = ctx->pl_dof[z + 1] - dof0;
      if (k8 && np % 16 == 0 && dof0 % 8 == 0) {
        // 16 right-hand sides per workgroup on the matrix cores
        hipLaunchKernelGGL(k_pl_step_mfma, dim3((L + 1) / 2, (ncol + 15) / 16), dim3(256), 0,
                           ctx->stream, ctx->geo, ctx->Acb.p, ctx->en_lv1,
                           ctx->pl_G.p + ctx->pl_goff[z], np, dof0, z * L, L, fwd, w, d, nc,
                           ncol, ctx->geo.stop);
      } else if (k8) {
        // up to 16 right-hand sides per workgroup in 4 groups of 4
        pl_step<8, 4, 1024, 4>(ctx, z, fwd, w, d, ncol);
      } else {
        pl_step<kMaxEig, 2, 1024, 4>(ctx, z, fwd, w, d, ncol);
      }
    }
  };
  for (int z = 0; z + 1 < Z; ++z) step(z, 1);
  for (int z = Z - 1; z >= 0; --z) step(z, 0);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

}  // namespace pmx
