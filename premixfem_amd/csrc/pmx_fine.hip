// Fine-grid kernels: layout permutation, face coefficients, eigen mass,
// level-1 block ICC(0) (factor and wavefront solve), residual + restriction,
// prolongation and the PCG vector kernels.  All are HBM-bandwidth bound; each
// workgroup owns one subdomain brick so every access is to a contiguous block.
#include "pmx_kernels.h"

namespace pmx {

// ---------------------------------------------------------------------------
// Natural <-> brick permutation.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int brick_of_natural(const Geo& g, int c) {
  const int x = c % g.M;
  const int y = (c / g.M) % g.N;
  const int z = c / (g.M * g.N);
  const int s = g.ax_sub[0][x] + g.sdx * (g.ax_sub[1][y] + g.sdy * g.ax_sub[2][z]);
  const Brick& b = g.bricks[s];
  return b.off + (g.ax_loc[2][z] * b.ny + g.ax_loc[1][y]) * b.nx + g.ax_loc[0][x];
}

template <typename S, typename D>
__global__ void k_nat2brick(Geo g, const S* __restrict__ src, D* __restrict__ dst, int nvec,
                            int src_stride, int dst_stride) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.ncell) return;
  const int bi = brick_of_natural(g, c);
  for (int v = 0; v < nvec; ++v) dst[(size_t)v * dst_stride + bi] = (D)src[(size_t)v * src_stride + c];
}

template <typename S, typename D>
__global__ void k_brick2nat(Geo g, const S* __restrict__ src, D* __restrict__ dst, int nvec,
                            int src_stride, int dst_stride) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.ncell) return;
  const int bi = brick_of_natural(g, c);
  for (int v = 0; v < nvec; ++v) dst[(size_t)v * dst_stride + c] = (D)src[(size_t)v * src_stride + bi];
}

// Row-major [cell][k] basis <-> natural column vectors.
__global__ void k_basis_nat2rows(Geo g, const double* __restrict__ src, double* __restrict__ rows,
                                 int k) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.ncell) return;
  const int bi = brick_of_natural(g, c);
  for (int j = 0; j < k; ++j) rows[(size_t)bi * k + j] = src[(size_t)j * g.ncell + c];
}

__global__ void k_basis_rows2nat(Geo g, const double* __restrict__ rows, double* __restrict__ dst,
                                 int k) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.ncell) return;
  const int bi = brick_of_natural(g, c);
  for (int j = 0; j < k; ++j) dst[(size_t)j * g.ncell + c] = rows[(size_t)bi * k + j];
}

// ---------------------------------------------------------------------------
// Face coefficients T = (|F|^2/|E|) * 2/(1/k_L + 1/k_R)
// (src/PreMixFEM_3D.c:1785-1786, :1797-1798, :1809-1810).  T[d][c] is the
// coefficient of the face between c and its lower neighbour in direction d,
// 0 on the domain boundary (no-flow).
// ---------------------------------------------------------------------------
__global__ void k_trans(Geo g, const double* __restrict__ kx, const double* __restrict__ ky,
                        const double* __restrict__ kz, double cx, double cy, double cz,
                        double* __restrict__ Tx, double* __restrict__ Ty, double* __restrict__ Tz,
                        float* __restrict__ Tx32, float* __restrict__ Ty32,
                        float* __restrict__ Tz32) {
  const Brick b = g.bricks[blockIdx.x];
  const int n = b.nx * b.ny * b.nz;
  for (int li = threadIdx.x; li < n; li += blockDim.x) {
    const int lx = li % b.nx, ly = (li / b.nx) % b.ny, lz = li / (b.nx * b.ny);
    const int c = b.off + li;
    double t[3];
    const double* kk[3] = {kx, ky, kz};
    const double cc[3] = {cx, cy, cz};
    for (int d = 0; d < 3; ++d) {
      const int nb = nbr_index(g, b, c, lx, ly, lz, 2 * d);
      t[d] = nb < 0 ? 0.0 : cc[d] * (2.0 / (1.0 / kk[d][nb] + 1.0 / kk[d][c]));
    }
    Tx[c] = t[0];
    Ty[c] = t[1];
    Tz[c] = t[2];
    if (Tx32) {
      Tx32[c] = (float)t[0];
      Ty32[c] = (float)t[1];
      Tz32[c] = (float)t[2];
    }
  }
}

// ---------------------------------------------------------------------------
// Eigen mass diagonal (src/PreMixFEM_3D.c:265-343).  MOD: per cell, in the
// reference's accumulation order: [lower x if first cell of the row], upper x,
// lower y, upper y, lower z, upper z, [upper x again if not last of the row];
// each term 2*kbar/H/H, then scaled by |E|.  MEAN: kx+ky+kz.  DIAG: diagonal
// of the Neumann block, with the well term (src/spe10.c:990-1012).
// ---------------------------------------------------------------------------
__global__ void k_mass(Geo g, int op, const double* __restrict__ kx, const double* __restrict__ ky,
                       const double* __restrict__ kz, const double* __restrict__ Tx,
                       const double* __restrict__ Ty, const double* __restrict__ Tz, double Hx,
                       double Hy, double Hz, double meas_elem, double* __restrict__ m) {
  const Brick b = g.bricks[blockIdx.x];
  const int n = b.nx * b.ny * b.nz;
  for (int li = threadIdx.x; li < n; li += blockDim.x) {
    const int lx = li % b.nx, ly = (li / b.nx) % b.ny, lz = li / (b.nx * b.ny);
    const int c = b.off + li;
    if (op == PMX_EIG_OP_MEAN) {
      m[c] = kx[c] + ky[c] + kz[c];
      continue;
    }
    if (op == PMX_EIG_OP_DIAG) {
      double v = 0.0;
      if (lx > 0) v += Tx[c];
      if (lx < b.nx - 1) v += Tx[c + 1];
      if (ly > 0) v += Ty[c];
      if (ly < b.ny - 1) v += Ty[c + b.nx];
      if (lz > 0) v += Tz[c];
      if (lz < b.nz - 1) v += Tz[c + b.nx * b.ny];
      if (g.wd) v += g.wd[c];
      m[c] = v;
      continue;
    }
    const double* kk[3] = {kx, ky, kz};
    const double H[3] = {Hx, Hy, Hz};
    double lo[3], hi[3];
    bool has_lo[3], has_hi[3];
    for (int d = 0; d < 3; ++d) {
      const int nl = nbr_index(g, b, c, lx, ly, lz, 2 * d);
      const int nh = nbr_index(g, b, c, lx, ly, lz, 2 * d + 1);
      has_lo[d] = nl >= 0;
      has_hi[d] = nh >= 0;
      lo[d] = has_lo[d] ? 2.0 * (2.0 / (1.0 / kk[d][nl] + 1.0 / kk[d][c])) / H[d] / H[d] : 0.0;
      hi[d] = has_hi[d] ? 2.0 * (2.0 / (1.0 / kk[d][c] + 1.0 / kk[d][nh])) / H[d] / H[d] : 0.0;
    }
    double v = 0.0;
    if (lx == 0 && has_lo[0]) v = lo[0];
    if (has_hi[0]) v = v + hi[0];
    if (has_lo[1]) v = v + lo[1];
    if (has_hi[1]) v = v + hi[1];
    if (has_lo[2]) v = v + lo[2];
    if (has_hi[2]) v = v + hi[2];
    if (lx < b.nx - 1) v = v + hi[0];
    m[c] = v * meas_elem;
  }
}

// ---------------------------------------------------------------------------
// Level-1 block ICC(0) factor by local hyperplanes lx+ly+lz = s.
// The 7-point graph has no triangles, so ICC(0) in natural order keeps only
// the pivots: d = a - cz^2/d_{-z} - cy^2/d_{-y} - cx^2/d_{-x} (couplings inside
// the subdomain only).  mode 0: a = diagonal of A[Omega_i, Omega_i] in the
// accumulation order of _PC_setup_lv1_J (src/PreMixFEM_3D.c:93-197), + pin,
// + well term (src/spe10.c:921-927); mode 1: a = Neumann diagonal (+ well
// term) + sigma * mass (the LOBPCG preconditioner).
// Writes plane-ordered {dinv, cx, cy, cz} for the lower and upper couplings.
// ---------------------------------------------------------------------------
// Diagonal of the level-1 block at cell c in the accumulation order of
// _PC_setup_lv1_J (src/PreMixFEM_3D.c:93-197): lower faces and the upper
// faces on the brick boundary in x, y, z order, then the interior upper
// faces, then the pin and well terms.
__device__ __forceinline__ double lv1_diag_order(const Geo& g, int c, const double lo[3],
                                                 const double hi[3], const bool in_hi[3]) {
  double a = lo[0];
  if (!in_hi[0]) a = a + hi[0];
  a = a + lo[1];
  if (!in_hi[1]) a = a + hi[1];
  a = a + lo[2];
  if (!in_hi[2]) a = a + hi[2];
  if (in_hi[0]) a = a + hi[0];
  if (in_hi[1]) a = a + hi[1];
  if (in_hi[2]) a = a + hi[2];
  if (c == g.pin) a = a + g.tau;
  if (g.wd) a = a + g.wd[c];
  return a;
}

template <typename F4T>
__global__ void k_dic_factor(Geo g, int mode, const double* __restrict__ Tx,
                             const double* __restrict__ Ty, const double* __restrict__ Tz,
                             const double* __restrict__ mass, double sigma,
                             F4T* __restrict__ fac_f, F4T* __restrict__ fac_b) {
  extern __shared__ double dsh[];
  const Brick b = g.bricks[blockIdx.x];
  const int nxny = b.nx * b.ny;
  const int* ps = g.plane_start + g.shape_pstart[b.shape];
  const int* pc = g.plane_cell + g.shape_coff[b.shape];
  const double* TT[3] = {Tx, Ty, Tz};
  for (int s = 0; s < b.nplanes; ++s) {
    for (int pos = ps[s] + threadIdx.x; pos < ps[s + 1]; pos += blockDim.x) {
      const int pk = pc[pos];
      const int lx = pk & 1023, ly = (pk >> 10) & 1023, lz = pk >> 20;
      const int li = (lz * b.ny + ly) * b.nx + lx;
      const int c = b.off + li;
      double lo[3], hi[3];
      bool in_hi[3];
      const int lc[3] = {lx, ly, lz};
      const int ln[3] = {b.nx, b.ny, b.nz};
      for (int d = 0; d < 3; ++d) {
        lo[d] = TT[d][c];  // 0 at the domain boundary
        const int nh = nbr_index(g, b, c, lx, ly, lz, 2 * d + 1);
        hi[d] = nh >= 0 ? TT[d][nh] : 0.0;
        in_hi[d] = lc[d] < ln[d] - 1;
      }
      double a;
      if (mode == 0) {
        a = lv1_diag_order(g, c, lo, hi, in_hi);
      } else {
        a = 0.0;
        if (lx > 0) a += lo[0];
        if (in_hi[0]) a += hi[0];
        if (ly > 0) a += lo[1];
        if (in_hi[1]) a += hi[1];
        if (lz > 0) a += lo[2];
        if (in_hi[2]) a += hi[2];
        if (g.wd) a += g.wd[c];
        a += sigma * mass[c];
      }
      if (lz > 0) a = a - lo[2] * lo[2] / dsh[li - nxny];
      if (ly > 0) a = a - lo[1] * lo[1] / dsh[li - b.nx];
      if (lx > 0) a = a - lo[0] * lo[0] / dsh[li - 1];
      dsh[li] = a;
    }
    __syncthreads();
  }
  const int n = b.nx * b.ny * b.nz;
  for (int pos = threadIdx.x; pos < n; pos += blockDim.x) {
    const int pk = pc[pos];
    const int lx = pk & 1023, ly = (pk >> 10) & 1023, lz = pk >> 20;
    const int li = (lz * b.ny + ly) * b.nx + lx;
    const int c = b.off + li;
    const double dinv = 1.0 / dsh[li];
    F4T f, u;
    f.x = dinv;
    f.y = lx > 0 ? Tx[c] : 0.0;
    f.z = ly > 0 ? Ty[c] : 0.0;
    f.w = lz > 0 ? Tz[c] : 0.0;
    u.x = dinv;
    u.y = lx < b.nx - 1 ? Tx[c + 1] : 0.0;
    u.z = ly < b.ny - 1 ? Ty[c + b.nx] : 0.0;
    u.w = lz < b.nz - 1 ? Tz[c + nxny] : 0.0;
    fac_f[b.off + pos] = f;
    fac_b[b.off + pos] = u;
  }
}

// ---------------------------------------------------------------------------
// Wavefront sweeps of the block ICC(0) solve on a brick held in LDS:
// forward  w = (r + cx w_{-x} + cy w_{-y} + cz w_{-z}) * dinv
// backward y = w + (cx' y_{+x} + cy' y_{+y} + cz' y_{+z}) * dinv
// i.e. y = [(D+L) D^{-1} (D+L^T)]^{-1} r.  Factor data is plane-ordered so
// the loads of one hyperplane are contiguous.  Each thread keeps a ring of
// kSweepDepth planes of factor data in flight: the loads for plane k+depth
// are issued right after plane k is processed, so HBM latency overlaps the
// LDS work and barriers of the planes in between.  The plane table sits in
// LDS.
// ---------------------------------------------------------------------------
constexpr int kSweepDepth = 4;

// Branch-free cell update: a coupling is 0 on a brick face, so the
// out-of-brick neighbour index is replaced by the cell itself and its term
// adds an exact zero (same result as skipping it).
template <typename S, typename F4T, bool FWD>
__device__ __forceinline__ void dic_cell(const BrickView& v, int pk, const F4T& f, S* lv) {
  const int lx = pk & 1023, ly = (pk >> 10) & 1023, lz = pk >> 20;
  const int li = (lz * v.ny + ly) * v.nx + lx;
  if (FWD) {
    const S a0 = lv[li];
    const S ax = lv[lx > 0 ? li - 1 : li];
    const S ay = lv[ly > 0 ? li - v.nx : li];
    const S az = lv[lz > 0 ? li - v.nxny : li];
    S acc = a0;
    acc += (S)f.y * ax;
    acc += (S)f.z * ay;
    acc += (S)f.w * az;
    lv[li] = acc * (S)f.x;
  } else {
    const S ax = lv[lx < v.nx - 1 ? li + 1 : li];
    const S ay = lv[ly < v.ny - 1 ? li + v.nx : li];
    const S az = lv[lz < v.nz - 1 ? li + v.nxny : li];
    const S a0 = lv[li];
    S acc = (S)0;
    acc += (S)f.y * ax;
    acc += (S)f.z * ay;
    acc += (S)f.w * az;
    lv[li] = a0 + acc * (S)f.x;
  }
}

// BIG: some hyperplane has more cells than the block has threads (the extra
// cells are processed without prefetch).
template <typename S, typename F4T, bool FWD, bool BIG>
__device__ __forceinline__ void dic_sweep(const BrickView& v, int nplanes, const int* sps,
                                          const int* __restrict__ pc,
                                          const F4T* __restrict__ fac, S* lv) {
  constexpr int D = kSweepDepth;
  const int t = threadIdx.x;
  F4T f[D];
  int pk[D];
  bool ok[D];
  // Loads are issued unconditionally (clamped to a valid cell) so the
  // compiler can count them across the unrolled steps.
  auto fetch = [&](int k, F4T& ff, int& pp, bool& okk) {
    const int kk = k < nplanes ? k : nplanes - 1;
    const int s = FWD ? kk : nplanes - 1 - kk;
    const int lo = sps[s], hi = sps[s + 1];
    okk = k < nplanes && lo + t < hi;
    const int pos = min(lo + t, hi - 1);
    ff = fac[v.off + pos];
    pp = pc[pos];
  };
  // Issue order f[0], f[1], ... (the scheduler must not reorder them, or the
  // waits in the loop become full drains).
#pragma unroll
  for (int d = 0; d < D; ++d) {
    fetch(d, f[d], pk[d], ok[d]);
    __builtin_amdgcn_sched_barrier(0);
  }
  // The step count is padded to a multiple of D (steps past the last plane
  // only fetch and synchronise) so every path through the unrolled body
  // issues the same loads and the waits stay partial.
  for (int k0 = 0; k0 < nplanes; k0 += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int k = k0 + d;
      // The bounds of plane k + D are read from LDS before the cell update,
      // so their latency overlaps that of the update's neighbour reads.
      const int kn = k + D;
      const int kk = kn < nplanes ? kn : nplanes - 1;
      const int sn = FWD ? kk : nplanes - 1 - kk;
      const int lo = sps[sn], hi = sps[sn + 1];
      if (ok[d]) dic_cell<S, F4T, FWD>(v, pk[d], f[d], lv);
      if (BIG && k < nplanes) {
        const int s = FWD ? k : nplanes - 1 - k;
        for (int pos = sps[s] + t + (int)blockDim.x; pos < sps[s + 1]; pos += blockDim.x)
          dic_cell<S, F4T, FWD>(v, pc[pos], fac[v.off + pos], lv);
      }
      ok[d] = kn < nplanes && lo + t < hi;
      const int pos = min(lo + t, hi - 1);
      f[d] = fac[v.off + pos];
      pk[d] = pc[pos];
      __syncthreads();
    }
  }
}

// Loads the brick shape's plane table into LDS (sps: nplanes+1 ints).
__device__ __forceinline__ void load_planes(const Geo& g, const Brick& b, int* sps) {
  const int* ps = g.plane_start + g.shape_pstart[b.shape];
  for (int i = threadIdx.x; i <= b.nplanes; i += blockDim.x) sps[i] = ps[i];
}

template <typename S, typename F4T, bool BIG>
__device__ __forceinline__ void dic_sweeps(const Geo& g, const Brick& b, const BrickView& v,
                                           const int* sps, const F4T* __restrict__ fac_f,
                                           const F4T* __restrict__ fac_b, S* lv) {
  const int* pc = g.plane_cell + g.shape_coff[b.shape];
  dic_sweep<S, F4T, true, BIG>(v, b.nplanes, sps, pc, fac_f, lv);
  dic_sweep<S, F4T, false, BIG>(v, b.nplanes, sps, pc, fac_b, lv);
}

// ---------------------------------------------------------------------------
// Exact level-1 blocks (-use_full_Cholesky_lv1, src/PreMixFEM_3D.c:205-207:
// PCCHOLESKY instead of PCICC on every block A[Omega_i, Omega_i]).  In the
// brick-local x-fastest order a block is banded with half-bandwidth
// bw = nx*ny, and its Cholesky factor fills exactly that band.  The solve is
// column-oriented: one workgroup barrier per column, the factor column read
// one column ahead.
// ---------------------------------------------------------------------------
constexpr int kCholMaxBw = 1023;   // nx*ny of a brick (4 band entries per thread)

// L y = r then L^T z = y on the brick vector lv (FP64); tmp: n doubles of LDS.
// Every thread of the workgroup calls it; the result replaces lv.
__device__ __forceinline__ void chol_solve_brick(const Geo& g, int sb, const BrickView& v,
                                                 double* lv, double* tmp) {
  const int n = v.n, bw = v.nxny, W = bw + 1;
  const double* L = g.l1L + g.l1off[sb];
  const double* U = g.l1U + g.l1off[sb];
  const int t = threadIdx.x;
  // forward: tmp[j] = y_j, lv[j+d] -= L(j+d, j) y_j
  double lnext[4], pnext = L[0];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int d = 1 + t + q * (int)blockDim.x;
    lnext[q] = d <= bw && d < n ? L[d] : 0.0;
  }
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    double lc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) lc[q] = lnext[q];
    const double piv = pnext;
    if (j + 1 < n) {
      const double* col = L + (size_t)(j + 1) * W;
      pnext = col[0];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int d = 1 + t + q * (int)blockDim.x;
        lnext[q] = d <= bw && j + 1 + d < n ? col[d] : 0.0;
      }
    }
    const double y = lv[j] / piv;
    if (t == 0) tmp[j] = y;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int d = 1 + t + q * (int)blockDim.x;
      if (d <= bw && j + d < n) lv[j + d] -= lc[q] * y;
    }
    __syncthreads();
  }
  // backward: lv[j] = z_j, tmp[j-d] -= L(j, j-d) z_j
  {
    const double* row = U + (size_t)(n - 1) * W;
    pnext = row[0];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int d = 1 + t + q * (int)blockDim.x;
      lnext[q] = d <= bw && n - 1 - d >= 0 ? row[d] : 0.0;
    }
  }
  for (int j = n - 1; j >= 0; --j) {
    double lc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) lc[q] = lnext[q];
    const double piv = pnext;
    if (j > 0) {
      const double* row = U + (size_t)(j - 1) * W;
      pnext = row[0];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int d = 1 + t + q * (int)blockDim.x;
        lnext[q] = d <= bw && j - 1 - d >= 0 ? row[d] : 0.0;
      }
    }
    const double z = tmp[j] / piv;
    if (t == 0) lv[j] = z;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int d = 1 + t + q * (int)blockDim.x;
      if (d <= bw && j - d >= 0) tmp[j - d] -= lc[q] * z;
    }
    __syncthreads();
  }
}

// (A x)_c for one cell: diag*x_c - sum_faces T x_nb, diag = sum of the
// existing faces (+ pin).  T[d][c] is the lower face of c (0 on the domain
// boundary), the upper face is the lower face of the +d neighbour.
template <typename TT, typename XT, typename ACC>
__device__ __forceinline__ ACC stencil_fast(const Geo& g, const BrickView& v, int c, int lx, int ly,
                                            int lz, const TT* __restrict__ Tx,
                                            const TT* __restrict__ Ty,
                                            const TT* __restrict__ Tz, const XT* __restrict__ x) {
  int nb[6];
  nbr6(v, c, lx, ly, lz, nb);
  const TT* TT3[3] = {Tx, Ty, Tz};
  ACC diag = (ACC)0, acc = (ACC)0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int nl = nb[2 * d], nh = nb[2 * d + 1];
    const ACC tl = (ACC)TT3[d][c];  // 0 when nl < 0
    const ACC th = nh >= 0 ? (ACC)TT3[d][nh] : (ACC)0;
    diag += tl + th;
    acc += tl * (ACC)x[nl >= 0 ? nl : c] + th * (ACC)x[nh >= 0 ? nh : c];
  }
  if (c == g.pin) diag += (ACC)g.tau;
  const TT* W = well_term<TT>(g);
  if (W) diag += (ACC)W[c];
  return diag * (ACC)x[c] - acc;
}

// ---------------------------------------------------------------------------
// PCG scalars without separate reduction launches.
//
// Every kernel that needs a scalar of the CG recurrence (beta = z'r,
// dpi = p'w, ||r||^2) sums the per-brick partials of the kernel that
// produced them itself: wave 0 of each block adds them in a fixed order, so
// every block obtains the bitwise same value and takes the same decision.
// Partials of z'r and p'w alternate between two slots by iteration parity
// so the previous value (betaold, dpiold) is still there.  Block 0 records
// the outcome (iterations, reason, norms) in scal[] for the host.
// ---------------------------------------------------------------------------
// The index i of the current iteration lives on the device (scal[kSIter]):
// the prolongation of iteration i advances it to i+1 (no kernel that reads
// it runs concurrently with that write), so a captured chunk of iterations
// replays without host-side arguments.
struct CgCtl {
  double* scal;
  double* zr;  // [2][nparts]
  double* pw;  // [2][nparts]
  double* rr;  // [nparts]
  int nparts;
  double abstol, dtol;
  int max_it;
};

// Fixed-order sums of up to three arrays of n doubles by one wave: lane l
// adds elements l, l+64, l+128, ... in order, then a butterfly.  Eight loads
// per array are in flight per lane, so the sums cost about one memory
// round trip for n <= 512.  Results are valid in every lane of the wave.
__device__ __forceinline__ void wave_sum3(const double* __restrict__ a, const double* __restrict__ b,
                                          const double* __restrict__ c, int n, double& sa,
                                          double& sb, double& sc) {
  const int lane = threadIdx.x & 63;
  sa = sb = sc = 0.0;
  for (int i0 = 0; i0 < n; i0 += 64 * 8) {
    double va[8], vb[8], vc[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int i = i0 + lane + 64 * j;
      va[j] = i < n ? a[i] : 0.0;
      vb[j] = (b && i < n) ? b[i] : 0.0;
      vc[j] = (c && i < n) ? c[i] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      sa += va[j];
      sb += vb[j];
      sc += vc[j];
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    sa += __shfl_xor(sa, o, 64);
    sb += __shfl_xor(sb, o, 64);
    sc += __shfl_xor(sc, o, 64);
  }
}

__device__ __forceinline__ double wave_sum(const double* __restrict__ p, int n) {
  double s, u, v;
  wave_sum3(p, nullptr, nullptr, n, s, u, v);
  return s;
}

__device__ __forceinline__ void cg_finish(const CgCtl& c, int its, int reason, double rnorm2) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    c.scal[kSIts] = its;
    c.scal[kSReason] = reason;
    if (rnorm2 >= 0.0) c.scal[kSRnorm2] = rnorm2;
    c.scal[kSDone] = 1.0;
  }
}

// Iteration i, part 1 (src/Expe_scal.c:114 KSPSolve_CG with
// KSP_NORM_UNPRECONDITIONED):
//   i > 0: ||r|| of the previous update against ttol / dtol / max_it;
//   beta checks (beta = 0 -> CONVERGED_ATOL, sign change -> INDEFINITE_PC);
//   p = z + (beta/betaold) p (p = z when i = 0), w = A p, partial p'w.
// p_old is read at neighbours, so p goes to a second buffer.
// Scalar part of k_cg_p_spmv: the convergence tests of iteration i on the
// previous update's ||r||^2 and the beta checks.  Returns false when the
// solve ends here (every block takes the same decision); otherwise bc =
// beta / betaold (0 at i = 0), first = (i == 0), i = the iteration.
__device__ __forceinline__ bool cg_p_begin(const CgCtl& cg, double* sh, double& bc, bool& first,
                                           int& i) {
  i = (int)cg.scal[kSIter];
  if (threadIdx.x < 64) {  // wave 0 (a block may have only one wave)
    double s0, s1, s2;
    wave_sum3(cg.zr + (size_t)(i & 1) * cg.nparts,
              i > 0 ? cg.zr + (size_t)((i + 1) & 1) * cg.nparts : nullptr,
              i > 0 ? cg.rr : nullptr, cg.nparts, s0, s1, s2);
    if (threadIdx.x == 0) {
      sh[0] = s0;
      sh[1] = s1;
      sh[2] = s2;
    }
  }
  __syncthreads();
  const double beta = sh[0];
  if (i > 0) {
    const double rr = sh[2];
    const double dp = sqrt(rr);
    int reason = 0;
    if (dp <= cg.scal[kSTtol])
      reason = dp < cg.abstol ? PMX_CONVERGED_ATOL : PMX_CONVERGED_RTOL;
    else if (dp >= cg.dtol * sqrt(cg.scal[kSRnorm0]))
      reason = PMX_DIVERGED_DTOL;
    else if (dp != dp)
      reason = PMX_DIVERGED_NANORINF;
    else if (i >= cg.max_it)
      reason = PMX_DIVERGED_ITS;
    if (reason) {
      cg_finish(cg, i, reason, rr);
      return false;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      cg.scal[kSIts] = i;
      cg.scal[kSRnorm2] = rr;
    }
  }
  if (beta == 0.0) {
    cg_finish(cg, i + 1, PMX_CONVERGED_ATOL, -1.0);
    return false;
  }
  if (i > 0 && beta * sh[1] < 0.0) {
    cg_finish(cg, i + 1, PMX_DIVERGED_INDEFINITE_PC, -1.0);
    return false;
  }
  first = i == 0;
  bc = first ? 0.0 : beta / sh[1];
  return true;
}

__global__ void __launch_bounds__(kStreamBlock) k_cg_p_spmv(
    Geo g, const double* __restrict__ z, const double* __restrict__ pold,
    double* __restrict__ pnew, double* __restrict__ w, const double* __restrict__ Tx,
    const double* __restrict__ Ty, const double* __restrict__ Tz, CgCtl cg) {
  PMX_STOP_CHECK(g.stop);
  __shared__ double sh[4];
  __shared__ double red[16];
  double bc;
  bool first;
  int i;
  if (!cg_p_begin(cg, sh, bc, first, i)) return;
  const BrickView v = load_view(g, block_brick(g));
  double acc = 0.0;
  const double* TT3[3] = {Tx, Ty, Tz};
  auto pval = [&](int q) { return first ? z[q] : z[q] + bc * pold[q]; };
  for_cells(v, [&](int c, int lx, int ly, int lz) {
    int nb[6];
    nbr6(v, c, lx, ly, lz, nb);
    const double pc = pval(c);
    double diag = 0.0, off = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int nl = nb[2 * d], nh = nb[2 * d + 1];
      const double tl = TT3[d][c];
      const double th = nh >= 0 ? TT3[d][nh] : 0.0;
      diag += tl + th;
      if (nl >= 0) off += tl * pval(nl);
      if (nh >= 0) off += th * pval(nh);
    }
    if (c == g.pin) diag += g.tau;
    if (g.wd) diag += g.wd[c];
    const double wc = diag * pc - off;
    pnew[c] = pc;
    w[c] = wc;
    acc += pc * wc;
  });
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) cg.pw[(size_t)(i & 1) * cg.nparts + blockIdx.x] = acc;
}

// The same iteration step with four consecutive x-cells per thread, for
// bricks whose x-extent is a multiple of 4 (so every brick offset and row is
// 32-byte aligned): p of the brick is staged in LDS, the interior
// neighbours come from there, the face neighbours from the adjacent bricks;
// z, p, w and the face coefficients move as double4.  Each cell's p and w
// are computed by the same operations in the same order as in k_cg_p_spmv;
// only the block's p'w partial is summed in another order.
// Requires n <= 4 * blockDim.x; LDS: n doubles.
__global__ void __launch_bounds__(kStreamBlock) k_cg_p_spmv_q(
    Geo g, const double* __restrict__ z, const double* __restrict__ pold,
    double* __restrict__ pnew, double* __restrict__ w, const double* __restrict__ Tx,
    const double* __restrict__ Ty, const double* __restrict__ Tz, CgCtl cg) {
  PMX_STOP_CHECK(g.stop);
  extern __shared__ double psh[];
  __shared__ double sh[4];
  __shared__ double red[16];
  double bc;
  bool first;
  int i;
  if (!cg_p_begin(cg, sh, bc, first, i)) return;
  const BrickView v = load_view(g, block_brick(g));
  const int q0 = 4 * (int)threadIdx.x;
  const bool act = q0 < v.n;
  const int c0 = v.off + q0;
  auto ld4 = [](const double* a, int c) { return *reinterpret_cast<const double4*>(a + c); };
  auto comb = [&](double a, double b2) { return first ? a : a + bc * b2; };
  auto pv4 = [&](int c) {
    const double4 zz = ld4(z, c);
    if (first) return zz;
    const double4 po = ld4(pold, c);
    return double4{zz.x + bc * po.x, zz.y + bc * po.y, zz.z + bc * po.z, zz.w + bc * po.w};
  };
  const double4 zero{0.0, 0.0, 0.0, 0.0};
  double4 pc = zero;
  if (act) {
    pc = pv4(c0);
    *reinterpret_cast<double4*>(pnew + c0) = pc;
    *reinterpret_cast<double4*>(psh + q0) = pc;
  }
  __syncthreads();
  double acc = 0.0;
  if (act) {
    const int lx0 = q0 % v.nx, l1 = q0 / v.nx;
    const int ly = l1 % v.ny, lz = l1 / v.ny;
    const double P[4] = {pc.x, pc.y, pc.z, pc.w};
    // Per cell: diag and the off-diagonal sum accumulate direction by
    // direction (x, y, z) in the order of k_cg_p_spmv; a missing neighbour
    // has T = 0 and p = 0, so its terms add +0.
    double diag[4], off[4];
    {
      double pxl = 0.0, pxr = 0.0, txr = 0.0;
      if (lx0 > 0) {
        pxl = psh[q0 - 1];
      } else if (v.noff[0] >= 0) {
        const int h = v.noff[0] + (lz * v.ny + ly) * v.next[0] + v.next[0] - 1;
        pxl = comb(z[h], pold[h]);
      }
      if (lx0 + 4 < v.nx) {
        pxr = psh[q0 + 4];
        txr = Tx[c0 + 4];
      } else if (v.noff[1] >= 0) {
        const int h = v.noff[1] + (lz * v.ny + ly) * v.next[1];
        pxr = comb(z[h], pold[h]);
        txr = Tx[h];
      }
      const double4 t4 = ld4(Tx, c0);
      const double T5[5] = {t4.x, t4.y, t4.z, t4.w, txr};
      const double P6[6] = {pxl, P[0], P[1], P[2], P[3], pxr};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        diag[j] = 0.0;
        off[j] = 0.0;
        diag[j] += T5[j] + T5[j + 1];
        off[j] += T5[j] * P6[j];
        off[j] += T5[j + 1] * P6[j + 2];
      }
    }
    {
      double4 pl = zero, ph = zero, th = zero;
      if (ly > 0) {
        pl = *reinterpret_cast<const double4*>(psh + q0 - v.nx);
      } else if (v.noff[2] >= 0) {
        pl = pv4(v.noff[2] + (lz * v.next[2] + v.next[2] - 1) * v.nx + lx0);
      }
      if (ly < v.ny - 1) {
        ph = *reinterpret_cast<const double4*>(psh + q0 + v.nx);
        th = ld4(Ty, c0 + v.nx);
      } else if (v.noff[3] >= 0) {
        const int h = v.noff[3] + lz * v.next[3] * v.nx + lx0;
        ph = pv4(h);
        th = ld4(Ty, h);
      }
      const double4 tl = ld4(Ty, c0);
      const double TL[4] = {tl.x, tl.y, tl.z, tl.w}, TH[4] = {th.x, th.y, th.z, th.w};
      const double PL[4] = {pl.x, pl.y, pl.z, pl.w}, PH[4] = {ph.x, ph.y, ph.z, ph.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        diag[j] += TL[j] + TH[j];
        off[j] += TL[j] * PL[j];
        off[j] += TH[j] * PH[j];
      }
    }
    {
      double4 pl = zero, ph = zero, th = zero;
      if (lz > 0) {
        pl = *reinterpret_cast<const double4*>(psh + q0 - v.nxny);
      } else if (v.noff[4] >= 0) {
        pl = pv4(v.noff[4] + ((v.next[4] - 1) * v.ny + ly) * v.nx + lx0);
      }
      if (lz < v.nz - 1) {
        ph = *reinterpret_cast<const double4*>(psh + q0 + v.nxny);
        th = ld4(Tz, c0 + v.nxny);
      } else if (v.noff[5] >= 0) {
        const int h = v.noff[5] + ly * v.nx + lx0;
        ph = pv4(h);
        th = ld4(Tz, h);
      }
      const double4 tl = ld4(Tz, c0);
      const double TL[4] = {tl.x, tl.y, tl.z, tl.w}, TH[4] = {th.x, th.y, th.z, th.w};
      const double PL[4] = {pl.x, pl.y, pl.z, pl.w}, PH[4] = {ph.x, ph.y, ph.z, ph.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        diag[j] += TL[j] + TH[j];
        off[j] += TL[j] * PL[j];
        off[j] += TH[j] * PH[j];
      }
    }
    double4 wd = zero;
    if (g.wd) wd = ld4(g.wd, c0);
    const double WD[4] = {wd.x, wd.y, wd.z, wd.w};
    double wv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (c0 + j == g.pin) diag[j] += g.tau;
      if (g.wd) diag[j] += WD[j];
      wv[j] = diag[j] * P[j] - off[j];
    }
    *reinterpret_cast<double4*>(w + c0) = double4{wv[0], wv[1], wv[2], wv[3]};
#pragma unroll
    for (int j = 0; j < 4; ++j) acc += P[j] * wv[j];
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) cg.pw[(size_t)(i & 1) * cg.nparts + blockIdx.x] = acc;
}

// Stage "lv1" of PC_apply_vec with y = 0 (so r = x exactly): y = B1 x.
// UPD: first fuses part 2 of PCG iteration i: dpi = p'w (sign-change check ->
// INDEFINITE_MAT), a = beta/dpi, x += a p, r -= a w (r is the PC input x),
// partial ||r||^2.
template <typename S, typename F4T, bool UPD, bool BIG, bool CHOL = false>
__global__ void __launch_bounds__(kBlock) k_lv1_first(Geo g, double* __restrict__ x,
                                                      S* __restrict__ y,
                                                      const F4T* __restrict__ fac_f,
                                                      const F4T* __restrict__ fac_b,
                                                      double* __restrict__ xs,
                                                      const double* __restrict__ p,
                                                      const double* __restrict__ w, CgCtl cg) {
  PMX_STOP_CHECK(g.stop);
  extern __shared__ char shraw[];
  S* lv = reinterpret_cast<S*>(shraw);
  __shared__ int sps[kMaxPlanes + 1];
  __shared__ double sh[4];
  __shared__ double red[16];
  const int sb = block_brick(g);
  const Brick b = g.bricks[sb];
  const BrickView v = load_view(g, sb);
  load_planes(g, b, sps);
  if (UPD) {
    const int i = (int)cg.scal[kSIter];
    if (threadIdx.x < 64) {
      double s0, s1, s2;
      wave_sum3(cg.pw + (size_t)(i & 1) * cg.nparts, cg.zr + (size_t)(i & 1) * cg.nparts,
                i > 0 ? cg.pw + (size_t)((i + 1) & 1) * cg.nparts : nullptr, cg.nparts, s0, s1,
                s2);
      if (threadIdx.x == 0) {
        sh[0] = s0;
        sh[1] = s1;
        sh[2] = s2;
      }
    }
    __syncthreads();
    const double dpi = sh[0];
    if (dpi == 0.0 || (i > 0 && (dpi > 0.0) != (sh[2] > 0.0))) {
      cg_finish(cg, i + 1, PMX_DIVERGED_INDEFINITE_MAT, -1.0);
      return;
    }
    const double a = sh[1] / dpi;
    double acc = 0.0;
    for (int li = threadIdx.x; li < v.n; li += blockDim.x) {
      const int c = v.off + li;
      xs[c] += a * p[c];
      const double rr = x[c] - a * w[c];
      x[c] = rr;
      acc += rr * rr;
      lv[li] = (S)rr;
    }
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) cg.rr[blockIdx.x] = acc;
  } else {
    for (int li = threadIdx.x; li < v.n; li += blockDim.x) lv[li] = (S)x[v.off + li];
    __syncthreads();
  }
  if constexpr (CHOL) {
    chol_solve_brick(g, sb, v, reinterpret_cast<double*>(lv),
                     reinterpret_cast<double*>(lv) + v.n);
  } else {
    dic_sweeps<S, F4T, BIG>(g, b, v, sps, fac_f, fac_b, lv);
  }
  for (int li = threadIdx.x; li < v.n; li += blockDim.x) y[v.off + li] = lv[li];
}

// Final lv1 stage: r = x - A y, z = y + B1 r, and partial sums of z.x
// (the PCG's beta = z'r, the PC input x being the PCG residual).
template <typename S, typename TT, typename F4T, bool BIG, bool CHOL = false>
__global__ void __launch_bounds__(kBlock) k_lv1_last(Geo g, const double* __restrict__ x,
                                                     const S* __restrict__ y,
                                                     const TT* __restrict__ Tx,
                                                     const TT* __restrict__ Ty,
                                                     const TT* __restrict__ Tz,
                                                     const F4T* __restrict__ fac_f,
                                                     const F4T* __restrict__ fac_b,
                                                     double* __restrict__ z, S* __restrict__ yout,
                                                     CgCtl cg) {
  PMX_STOP_CHECK(g.stop);
  extern __shared__ char shraw[];
  S* lv = reinterpret_cast<S*>(shraw);
  __shared__ int sps[kMaxPlanes + 1];
  __shared__ double red[16];
  const int sb = block_brick(g);
  const Brick b = g.bricks[sb];
  const BrickView v = load_view(g, sb);
  load_planes(g, b, sps);
  for_cells(v, [&](int c, int lx, int ly, int lz) {
    const S ay = stencil_fast<TT, S, S>(g, v, c, lx, ly, lz, Tx, Ty, Tz, y);
    lv[c - v.off] = (S)x[c] - ay;
  });
  __syncthreads();
  if constexpr (CHOL) {
    chol_solve_brick(g, sb, v, reinterpret_cast<double*>(lv),
                     reinterpret_cast<double*>(lv) + v.n);
  } else {
    dic_sweeps<S, F4T, BIG>(g, b, v, sps, fac_f, fac_b, lv);
  }
  double acc = 0.0;
  for (int li = threadIdx.x; li < v.n; li += blockDim.x) {
    const int c = v.off + li;
    const S yn = y[c] + lv[li];
    if (z) {
      const double zz = (double)yn;
      z[c] = zz;
      acc += zz * x[c];
    }
    if (yout) yout[c] = yn;
  }
  if (cg.zr) {  // z'r partials into the slot of the iteration the counter names
    acc = block_sum(acc, red);
    if (threadIdx.x == 0)
      cg.zr[(size_t)((int)cg.scal[kSIter] & 1) * cg.nparts + blockIdx.x] = acc;
  }
}

template <typename S>
struct Vec4T;
template <>
struct Vec4T<float> {
  using type = float4;
};
template <>
struct Vec4T<double> {
  using type = double4;
};

template <typename S>
__device__ __forceinline__ void ld4v(const S* p, S (&o)[4]) {
  const typename Vec4T<S>::type v = *reinterpret_cast<const typename Vec4T<S>::type*>(p);
  o[0] = v.x;
  o[1] = v.y;
  o[2] = v.z;
  o[3] = v.w;
}

template <typename S>
__device__ __forceinline__ void st4v(S* p, const S (&o)[4]) {
  *reinterpret_cast<typename Vec4T<S>::type*>(p) = typename Vec4T<S>::type{o[0], o[1], o[2], o[3]};
}

// k_lv1_last for bricks of quads (x-extent a multiple of 4, at most 4096
// cells, four quads per thread): the residual r = x - A y and the output
// z = y + w work on quads of cells with y staged in LDS (interior neighbours
// from there, face neighbours from the adjacent bricks) and x kept in
// registers from the residual to the z'x partial; the hyperplane sweeps are
// those of k_lv1_last.  Each cell's arithmetic is that of k_lv1_last; the
// z'x partial is summed in another order.  LDS: 2 n values of S.
template <typename S, typename TT, typename F4T>
__global__ void __launch_bounds__(kBlock) k_lv1_last_q(Geo g, const double* __restrict__ x,
                                                       const S* __restrict__ y,
                                                       const TT* __restrict__ Tx,
                                                       const TT* __restrict__ Ty,
                                                       const TT* __restrict__ Tz,
                                                       const F4T* __restrict__ fac_f,
                                                       const F4T* __restrict__ fac_b,
                                                       double* __restrict__ z,
                                                       S* __restrict__ yout, CgCtl cg) {
  PMX_STOP_CHECK(g.stop);
  constexpr int QPT = 4;  // quads per thread (4096-cell bricks, 256 threads)
  static_assert(QPT * 4 * kBlock >= 4 * kStreamBlock, "a quad brick must fit one pass");
  extern __shared__ char shraw[];
  const int sb = block_brick(g);
  const Brick b = g.bricks[sb];
  const BrickView v = load_view(g, sb);
  S* lv = reinterpret_cast<S*>(shraw);
  S* ysh = lv + v.n;
  __shared__ int sps[kMaxPlanes + 1];
  __shared__ double red[16];
  load_planes(g, b, sps);
  double xr[QPT][4];
#pragma unroll
  for (int u = 0; u < QPT; ++u) {
    const int q0 = 4 * ((int)threadIdx.x + u * kBlock);
    if (q0 < v.n) {
      const double4 x4 = *reinterpret_cast<const double4*>(x + v.off + q0);
      xr[u][0] = x4.x;
      xr[u][1] = x4.y;
      xr[u][2] = x4.z;
      xr[u][3] = x4.w;
      S yo[4];
      ld4v(y + v.off + q0, yo);
      st4v(ysh + q0, yo);
    }
  }
  __syncthreads();
  const TT* W = well_term<TT>(g);
#pragma unroll
  for (int u = 0; u < QPT; ++u) {
    const int q0 = 4 * ((int)threadIdx.x + u * kBlock);
    if (q0 >= v.n) continue;
    const int c0 = v.off + q0;
    const int lx0 = q0 % v.nx, l1 = q0 / v.nx;
    const int ly = l1 % v.ny, lz = l1 / v.ny;
    S yo[4], tx[4], ty[4], tz[4], tyh[4] = {}, tzh[4] = {}, wd[4] = {};
    S yl[4] = {}, yh[4] = {}, zl[4] = {}, zh[4] = {};
    ld4v(ysh + q0, yo);
    ld4v(Tx + c0, tx);
    ld4v(Ty + c0, ty);
    ld4v(Tz + c0, tz);
    if (W) ld4v(W + c0, wd);
    // neighbour y: a missing one has T = 0 and takes the cell's own value,
    // as in stencil_fast
    S xl = yo[0], xh = yo[3], txr = S(0);
    if (lx0 > 0) {
      xl = ysh[q0 - 1];
    } else if (v.noff[0] >= 0) {
      xl = y[v.noff[0] + (lz * v.ny + ly) * v.next[0] + v.next[0] - 1];
    }
    if (lx0 + 4 < v.nx) {
      xh = ysh[q0 + 4];
      txr = Tx[c0 + 4];
    } else if (v.noff[1] >= 0) {
      const int h = v.noff[1] + (lz * v.ny + ly) * v.next[1];
      xh = y[h];
      txr = Tx[h];
    }
    bool hyl = true, hyh = true, hzl = true, hzh = true;
    if (ly > 0)
      ld4v(ysh + q0 - v.nx, yl);
    else if (v.noff[2] >= 0)
      ld4v(y + v.noff[2] + (lz * v.next[2] + v.next[2] - 1) * v.nx + lx0, yl);
    else
      hyl = false;
    if (ly < v.ny - 1) {
      ld4v(ysh + q0 + v.nx, yh);
      ld4v(Ty + c0 + v.nx, tyh);
    } else if (v.noff[3] >= 0) {
      const int h = v.noff[3] + lz * v.next[3] * v.nx + lx0;
      ld4v(y + h, yh);
      ld4v(Ty + h, tyh);
    } else {
      hyh = false;
    }
    if (lz > 0)
      ld4v(ysh + q0 - v.nxny, zl);
    else if (v.noff[4] >= 0)
      ld4v(y + v.noff[4] + ((v.next[4] - 1) * v.ny + ly) * v.nx + lx0, zl);
    else
      hzl = false;
    if (lz < v.nz - 1) {
      ld4v(ysh + q0 + v.nxny, zh);
      ld4v(Tz + c0 + v.nxny, tzh);
    } else if (v.noff[5] >= 0) {
      const int h = v.noff[5] + ly * v.nx + lx0;
      ld4v(y + h, zh);
      ld4v(Tz + h, tzh);
    } else {
      hzh = false;
    }
    S r[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const S xc = yo[j];
      const S vxl = j > 0 ? yo[j - 1] : xl;
      const S vxh = j < 3 ? yo[j + 1] : xh;
      const S vyl = hyl ? yl[j] : xc, vyh = hyh ? yh[j] : xc;
      const S vzl = hzl ? zl[j] : xc, vzh = hzh ? zh[j] : xc;
      const S thx = j < 3 ? tx[j + 1] : txr;
      S diag = S(0), a = S(0);
      diag += tx[j] + thx;
      a += tx[j] * vxl + thx * vxh;
      diag += ty[j] + tyh[j];
      a += ty[j] * vyl + tyh[j] * vyh;
      diag += tz[j] + tzh[j];
      a += tz[j] * vzl + tzh[j] * vzh;
      if (c0 + j == g.pin) diag += (S)g.tau;
      if (W) diag += wd[j];
      const S ay = diag * xc - a;
      r[j] = (S)xr[u][j] - ay;
    }
    st4v(lv + q0, r);
  }
  __syncthreads();
  dic_sweeps<S, F4T, false>(g, b, v, sps, fac_f, fac_b, lv);
  double acc = 0.0;
#pragma unroll
  for (int u = 0; u < QPT; ++u) {
    const int q0 = 4 * ((int)threadIdx.x + u * kBlock);
    if (q0 >= v.n) continue;
    const int c0 = v.off + q0;
    S yo[4], w4[4], yn[4];
    ld4v(ysh + q0, yo);
    ld4v(lv + q0, w4);
#pragma unroll
    for (int j = 0; j < 4; ++j) yn[j] = yo[j] + w4[j];
    if (z) {
      double4 z4;
      z4.x = (double)yn[0];
      z4.y = (double)yn[1];
      z4.z = (double)yn[2];
      z4.w = (double)yn[3];
      *reinterpret_cast<double4*>(z + c0) = z4;
      acc += z4.x * xr[u][0];
      acc += z4.y * xr[u][1];
      acc += z4.z * xr[u][2];
      acc += z4.w * xr[u][3];
    }
    if (yout) st4v(yout + c0, yn);
  }
  if (cg.zr) {  // z'r partials into the slot of the iteration the counter names
    acc = block_sum(acc, red);
    if (threadIdx.x == 0)
      cg.zr[(size_t)((int)cg.scal[kSIter] & 1) * cg.nparts + blockIdx.x] = acc;
  }
}

// r = x - A y and c[dof+j] = sum_cells phi[cell][j] r[cell]  (the VecMDot of
// _PC_apply_vec_lv2, src/PreMixFEM_3D.c:1564).  y == nullptr means y = 0.
template <typename S, typename TT>
__global__ void __launch_bounds__(kStreamBlock) k_resid_restrict(
    Geo g, const double* __restrict__ x, const S* __restrict__ y, const TT* __restrict__ Tx,
    const TT* __restrict__ Ty, const TT* __restrict__ Tz, const S* __restrict__ phi, int kmax,
    double* __restrict__ cvec) {
  PMX_STOP_CHECK(g.stop);
  __shared__ double red[16 * kMaxEig];
  const int sb = block_brick(g);
  const Brick b = g.bricks[sb];
  const BrickView v = load_view(g, sb);
  const int ke = b.neig;
  double acc[kMaxEig];
#pragma unroll
  for (int j = 0; j < kMaxEig; ++j) acc[j] = 0.0;
  for_cells(v, [&](int c, int lx, int ly, int lz) {
    double r = x[c];
    if (y) r -= (double)stencil_fast<TT, S, S>(g, v, c, lx, ly, lz, Tx, Ty, Tz, y);
    const S* ph = phi + (size_t)c * kmax;
#pragma unroll
    for (int j = 0; j < kMaxEig; ++j)
      if (j < ke) acc[j] += (double)ph[j] * r;
  });
  // wave sums, then the waves in a fixed order
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < kMaxEig; ++j) {
    if (j < ke) {
      double s = acc[j];
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
      if (lane == 0) red[wv * kMaxEig + j] = s;
    }
  }
  __syncthreads();
  if (threadIdx.x < ke) {
    double s = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) s += red[q * kMaxEig + threadIdx.x];
    cvec[b.dof + threadIdx.x] = s;
  }
}

// y += phi c  (the VecMAXPY of _PC_apply_vec_lv2, src/PreMixFEM_3D.c:1578).
// iter (not null): the PCG iteration counter, advanced by one.
template <typename S>
__global__ void __launch_bounds__(kStreamBlock) k_prolong(Geo g, S* __restrict__ y,
                                                          const S* __restrict__ phi, int kmax,
                                                          const double* __restrict__ cvec,
                                                          double* __restrict__ iter) {
  PMX_STOP_CHECK(g.stop);
  __shared__ double cs[kMaxEig];
  const Brick b = g.bricks[block_brick(g)];
  if (iter && blockIdx.x == 0 && threadIdx.x == 0) *iter += 1.0;
  if (threadIdx.x < kMaxEig) cs[threadIdx.x] = threadIdx.x < b.neig ? cvec[b.dof + threadIdx.x] : 0.0;
  __syncthreads();
  const int n = b.nx * b.ny * b.nz;
  const int ke = b.neig;
  for (int li = threadIdx.x; li < n; li += blockDim.x) {
    const int c = b.off + li;
    const S* ph = phi + (size_t)c * kmax;
    S acc = (S)0;
#pragma unroll
    for (int j = 0; j < kMaxEig; ++j)
      if (j < ke) acc += ph[j] * (S)cs[j];
    y[c] = y[c] + acc;
  }
}

// Quad-cell k_resid_restrict (bricks with x-extent a multiple of 4, n <= 4
// blockDim.x): thread t owns cells 4t..4t+3 of its brick; y is staged in LDS
// for the interior neighbours; all global loads (x, y, faces, the KM basis
// rows of the quad) are issued before the LDS barrier.  Each cell's r is
// computed exactly as stencil_fast does; the restriction sums run in another
// order than in the per-cell kernel.  LDS: n values of S.
template <typename S, typename TT, int KM>
__global__ void __launch_bounds__(kStreamBlock) k_resid_restrict_q(
    Geo g, const double* __restrict__ x, const S* __restrict__ y, const TT* __restrict__ Tx,
    const TT* __restrict__ Ty, const TT* __restrict__ Tz, const S* __restrict__ phi,
    double* __restrict__ cvec) {
  PMX_STOP_CHECK(g.stop);
  extern __shared__ char shraw[];
  S* ysh = reinterpret_cast<S*>(shraw);
  __shared__ double red[16 * kMaxEig];
  const int sb = block_brick(g);
  const Brick b = g.bricks[sb];
  const BrickView v = load_view(g, sb);
  const int ke = b.neig;
  const int q0 = 4 * (int)threadIdx.x;
  const bool act = q0 < v.n;
  const int c0 = v.off + q0;
  const int lx0 = q0 % v.nx, l1 = q0 / v.nx;
  const int ly = l1 % v.ny, lz = l1 / v.ny;
  double xr[4] = {0.0, 0.0, 0.0, 0.0};
  S yo[4] = {}, tx[4] = {}, ty[4] = {}, tz[4] = {}, tyh[4] = {}, tzh[4] = {}, wd[4] = {};
  S hq[4][4] = {};
  S hxl = S(0), hxr = S(0), txr = S(0);
  S ph[4 * KM];
  if (act) {
    {
      const double4 v4 = *reinterpret_cast<const double4*>(x + c0);
      xr[0] = v4.x;
      xr[1] = v4.y;
      xr[2] = v4.z;
      xr[3] = v4.w;
    }
    const typename Vec4T<S>::type* p4 =
        reinterpret_cast<const typename Vec4T<S>::type*>(phi + (size_t)c0 * KM);
#pragma unroll
    for (int u = 0; u < KM; ++u) {
      const typename Vec4T<S>::type t = p4[u];
      ph[4 * u] = t.x;
      ph[4 * u + 1] = t.y;
      ph[4 * u + 2] = t.z;
      ph[4 * u + 3] = t.w;
    }
    if (y) {
      ld4v(y + c0, yo);
      ld4v(Tx + c0, tx);
      ld4v(Ty + c0, ty);
      ld4v(Tz + c0, tz);
      const TT* W = well_term<TT>(g);
      if (W) ld4v(W + c0, wd);
      if (lx0 == 0 && v.noff[0] >= 0)
        hxl = y[v.noff[0] + (lz * v.ny + ly) * v.next[0] + v.next[0] - 1];
      if (lx0 + 4 < v.nx) {
        txr = Tx[c0 + 4];
      } else if (v.noff[1] >= 0) {
        const int h = v.noff[1] + (lz * v.ny + ly) * v.next[1];
        hxr = y[h];
        txr = Tx[h];
      }
      if (ly == 0 && v.noff[2] >= 0)
        ld4v(y + v.noff[2] + (lz * v.next[2] + v.next[2] - 1) * v.nx + lx0, hq[0]);
      if (ly < v.ny - 1) {
        ld4v(Ty + c0 + v.nx, tyh);
      } else if (v.noff[3] >= 0) {
        const int h = v.noff[3] + lz * v.next[3] * v.nx + lx0;
        ld4v(y + h, hq[1]);
        ld4v(Ty + h, tyh);
      }
      if (lz == 0 && v.noff[4] >= 0)
        ld4v(y + v.noff[4] + ((v.next[4] - 1) * v.ny + ly) * v.nx + lx0, hq[2]);
      if (lz < v.nz - 1) {
        ld4v(Tz + c0 + v.nxny, tzh);
      } else if (v.noff[5] >= 0) {
        const int h = v.noff[5] + ly * v.nx + lx0;
        ld4v(y + h, hq[3]);
        ld4v(Tz + h, tzh);
      }
      st4v(ysh + q0, yo);
    }
  }
  if (y) __syncthreads();
  double acc[kMaxEig];
#pragma unroll
  for (int j = 0; j < kMaxEig; ++j) acc[j] = 0.0;
  if (act) {
    double r[4] = {xr[0], xr[1], xr[2], xr[3]};
    if (y) {
      // neighbour y: a missing one has T = 0; its value is then the cell's
      // own, as in stencil_fast
      S yl[4], yh[4], zl[4], zh[4];
      if (ly > 0) ld4v(ysh + q0 - v.nx, yl);
      if (ly < v.ny - 1) ld4v(ysh + q0 + v.nx, yh);
      if (lz > 0) ld4v(ysh + q0 - v.nxny, zl);
      if (lz < v.nz - 1) ld4v(ysh + q0 + v.nxny, zh);
      const bool hl = lx0 > 0 || v.noff[0] >= 0, hr = lx0 + 4 < v.nx || v.noff[1] >= 0;
      const S xl = lx0 > 0 ? ysh[q0 - 1] : hxl;
      const S xh = lx0 + 4 < v.nx ? ysh[q0 + 4] : hxr;
      const bool byl = ly > 0 || v.noff[2] >= 0, byh = ly < v.ny - 1 || v.noff[3] >= 0;
      const bool bzl = lz > 0 || v.noff[4] >= 0, bzh = lz < v.nz - 1 || v.noff[5] >= 0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const S xc = yo[j];
        const S vxl = j > 0 ? yo[j - 1] : (hl ? xl : xc);
        const S vxh = j < 3 ? yo[j + 1] : (hr ? xh : xc);
        const S vyl = byl ? (ly > 0 ? yl[j] : hq[0][j]) : xc;
        const S vyh = byh ? (ly < v.ny - 1 ? yh[j] : hq[1][j]) : xc;
        const S vzl = bzl ? (lz > 0 ? zl[j] : hq[2][j]) : xc;
        const S vzh = bzh ? (lz < v.nz - 1 ? zh[j] : hq[3][j]) : xc;
        const S thx = j < 3 ? tx[j + 1] : txr;
        S diag = S(0), a = S(0);
        diag += tx[j] + thx;
        a += tx[j] * vxl + thx * vxh;
        diag += ty[j] + tyh[j];
        a += ty[j] * vyl + tyh[j] * vyh;
        diag += tz[j] + tzh[j];
        a += tz[j] * vzl + tzh[j] * vzh;
        if (c0 + j == g.pin) diag += (S)g.tau;
        if (well_term<TT>(g)) diag += wd[j];
        r[j] -= (double)(diag * xc - a);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int j = 0; j < KM; ++j)
        if (j < ke) acc[j] += (double)ph[q * KM + j] * r[q];
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < KM; ++j) {
    if (j < ke) {
      double sum = acc[j];
      for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
      if (lane == 0) red[wv * kMaxEig + j] = sum;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < ke) {
    double sum = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) sum += red[q * kMaxEig + threadIdx.x];
    cvec[b.dof + threadIdx.x] = sum;
  }
}

// Quad-cell k_prolong: y += phi c for cells 4t..4t+3 (same arithmetic per
// cell as k_prolong).
template <typename S, int KM>
__global__ void __launch_bounds__(kStreamBlock) k_prolong_q(Geo g, S* __restrict__ y,
                                                            const S* __restrict__ phi,
                                                            const double* __restrict__ cvec,
                                                            double* __restrict__ iter) {
  PMX_STOP_CHECK(g.stop);
  const Brick b = g.bricks[block_brick(g)];
  if (iter && blockIdx.x == 0 && threadIdx.x == 0) *iter += 1.0;
  const int q0 = 4 * (int)threadIdx.x;
  const int n = b.nx * b.ny * b.nz;
  if (q0 >= n) return;
  const int c0 = b.off + q0;
  const int ke = b.neig;
  S cs[KM];
#pragma unroll
  for (int j = 0; j < KM; ++j) cs[j] = j < ke ? (S)cvec[b.dof + j] : S(0);
  S ph[4 * KM];
  const typename Vec4T<S>::type* p4 =
      reinterpret_cast<const typename Vec4T<S>::type*>(phi + (size_t)c0 * KM);
#pragma unroll
  for (int u = 0; u < KM; ++u) {
    const typename Vec4T<S>::type t = p4[u];
    ph[4 * u] = t.x;
    ph[4 * u + 1] = t.y;
    ph[4 * u + 2] = t.z;
    ph[4 * u + 3] = t.w;
  }
  S yv[4];
  ld4v(y + c0, yv);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    S acc = (S)0;
#pragma unroll
    for (int j = 0; j < KM; ++j)
      if (j < ke) acc += ph[q * KM + j] * cs[j];
    yv[q] = yv[q] + acc;
  }
  st4v(y + c0, yv);
}

// r = b - A x, partial r'r; with an initial guess (x != null) also the
// partial b'b, written nb entries further on.
__global__ void __launch_bounds__(kStreamBlock) k_residual(
    Geo g, const double* __restrict__ bvec, const double* __restrict__ x, double* __restrict__ r,
    const double* __restrict__ Tx, const double* __restrict__ Ty, const double* __restrict__ Tz,
    double* __restrict__ partial) {
  __shared__ double red[16];
  const BrickView v = load_view(g, block_brick(g));
  double acc = 0.0, bacc = 0.0;
  for_cells(v, [&](int c, int lx, int ly, int lz) {
    const double bc = bvec[c];
    double rr = bc;
    if (x) rr -= stencil_fast<double, double, double>(g, v, c, lx, ly, lz, Tx, Ty, Tz, x);
    r[c] = rr;
    acc += rr * rr;
    bacc += bc * bc;
  });
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
  if (x) {
    bacc = block_sum(bacc, red);
    if (threadIdx.x == 0) partial[gridDim.x + blockIdx.x] = bacc;
  }
}

// y = A x (brick order), FP64.
__global__ void __launch_bounds__(kStreamBlock) k_spmv(Geo g, const double* __restrict__ x,
                                                       double* __restrict__ y,
                                                       const double* __restrict__ Tx,
                                                       const double* __restrict__ Ty,
                                                       const double* __restrict__ Tz) {
  const BrickView v = load_view(g, block_brick(g));
  for_cells(v, [&](int c, int lx, int ly, int lz) {
    y[c] = stencil_fast<double, double, double>(g, v, c, lx, ly, lz, Tx, Ty, Tz, x);
  });
}

// Start of the PCG: ||r0||^2 from the residual partials, its = 0 and the
// reference norm of KSPConvergedDefault at iteration 0: ||r0|| for a zero
// initial guess, ||b|| for a nonzero one (||r0|| if b = 0); ttol =
// max(rtol * rnorm0, abstol).  bpart: the b'b partials, null for a zero guess.
__global__ void k_cg_start(const double* __restrict__ partial, const double* __restrict__ bpart,
                           int nb, double* __restrict__ scal, double rtol, double abstol) {
  if (threadIdx.x >= 64) return;
  const double acc = wave_sum(partial, nb);
  const double bacc = bpart ? wave_sum(bpart, nb) : 0.0;
  if (threadIdx.x != 0) return;
  const double ref = bacc > 0.0 ? bacc : acc;
  scal[kSRnorm2] = acc;
  scal[kSRnorm0] = ref;
  const double dp = sqrt(acc);
  const double ttol = fmax(rtol * sqrt(ref), abstol);
  scal[kSTtol] = ttol;
  scal[kSIts] = 0.0;
  scal[kSIter] = 0.0;
  scal[kSDone] = dp <= ttol ? 1.0 : 0.0;
  scal[kSReason] = dp <= ttol ? (dp < abstol ? PMX_CONVERGED_ATOL : PMX_CONVERGED_RTOL) : 0.0;
}

// RT0 face fluxes (velocity recovery): F = T (p_L - p_R) across the lower
// x/y/z face of every cell (the face coefficient of PC_create_A,
// src/PreMixFEM_3D.c:1785-1810), 0 on the domain boundary.  p is in brick
// order, the three flux arrays in natural order.
__global__ void __launch_bounds__(kStreamBlock) k_velocity(
    Geo g, const double* __restrict__ p, const double* __restrict__ Tx,
    const double* __restrict__ Ty, const double* __restrict__ Tz, double* __restrict__ fx,
    double* __restrict__ fy, double* __restrict__ fz) {
  const int s = block_brick(g);
  const BrickView v = load_view(g, s);
  const Brick b = g.bricks[s];
  for_cells(v, [&](int c, int lx, int ly, int lz) {
    int nb[6];
    nbr6(v, c, lx, ly, lz, nb);
    const size_t nat = ((size_t)(b.gz + lz) * g.N + (b.gy + ly)) * g.M + (b.gx + lx);
    const double pc = p[c];
    fx[nat] = nb[0] >= 0 ? Tx[c] * (p[nb[0]] - pc) : 0.0;
    fy[nat] = nb[2] >= 0 ? Ty[c] * (p[nb[2]] - pc) : 0.0;
    fz[nat] = nb[4] >= 0 ? Tz[c] * (p[nb[4]] - pc) : 0.0;
  });
}

// First index whose value is not finite or not > 0 (strict) / < 0 (!strict),
// or stays INT_MAX.
__global__ void k_check_field(const double* __restrict__ a, int n, int strict,
                              int* __restrict__ first) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double v = a[i];
    const bool ok = strict ? (v > 0.0 && v <= 1.7976931348623157e308)
                           : (v >= 0.0 && v <= 1.7976931348623157e308);
    if (!ok) atomicMin(first, i);
  }
}

// ---------------------------------------------------------------------------
// Launch wrappers
// ---------------------------------------------------------------------------
size_t brick_lds_bytes(const pmx_ctx* ctx, size_t elem) {
  int mx = 0;
  for (const Brick& b : ctx->h_bricks) mx = std::max(mx, b.nx * b.ny * b.nz);
  return (size_t)mx * elem;
}

// Threads per block for the streaming per-brick kernels: enough waves to keep
// HBM busy (up to 1024), no more than the largest brick needs, and at least
// one x-line.
int stream_threads(const pmx_ctx* ctx) {
  int mx = 0, nx = 0;
  for (const Brick& b : ctx->h_bricks) {
    mx = std::max(mx, b.nx * b.ny * b.nz);
    nx = std::max(nx, b.nx);
  }
  int t = std::min(kStreamBlock, (mx + 63) / 64 * 64);
  t = std::max(t, (nx + 63) / 64 * 64);
  return std::min(t, kStreamBlock);
}

static CgCtl make_cg(pmx_ctx* ctx) {
  CgCtl c;
  c.scal = ctx->cg_scal;
  c.nparts = ctx->nsub;
  c.zr = ctx->cg_part;
  c.pw = ctx->cg_part + 2 * (size_t)ctx->nsub;
  c.rr = ctx->cg_part + 4 * (size_t)ctx->nsub;
  c.abstol = ctx->cg_abstol;
  c.dtol = ctx->cg_dtol;
  c.max_it = ctx->cg_max_it;
  return c;
}

int launch_nat2brick(pmx_ctx* ctx, const double* src, double* dst, int nvec, int sstride,
                     int dstride) {
  const int nb = (ctx->ncell + kBlock - 1) / kBlock;
  hipLaunchKernelGGL((k_nat2brick<double, double>), dim3(nb), dim3(kBlock), 0, ctx->stream,
                     ctx->geo, src, dst, nvec, sstride, dstride);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

int launch_brick2nat(pmx_ctx* ctx, const double* src, double* dst, int nvec, int sstride,
                     int dstride) {
  const int nb = (ctx->ncell + kBlock - 1) / kBlock;
  hipLaunchKernelGGL((k_brick2nat<double, double>), dim3(nb), dim3(kBlock), 0, ctx->stream,
                     ctx->geo, src, dst, nvec, sstride, dstride);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

int launch_basis_nat2rows(pmx_ctx* ctx, const double* src, double* rows, int k) {
  const int nb = (ctx->ncell + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_basis_nat2rows, dim3(nb), dim3(kBlock), 0, ctx->stream, ctx->geo, src,
                     rows, k);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

int launch_basis_rows2nat(pmx_ctx* ctx, const double* rows, double* dst, int k) {
  const int nb = (ctx->ncell + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_basis_rows2nat, dim3(nb), dim3(kBlock), 0, ctx->stream, ctx->geo, rows,
                     dst, k);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

int launch_trans(pmx_ctx* ctx) {
  const bool f32 = ctx->pc_precision == 32;
  hipLaunchKernelGGL(k_trans, dim3(ctx->nsub), dim3(kBlock), 0, ctx->stream, ctx->geo,
                     ctx->kappa[0].p, ctx->kappa[1].p, ctx->kappa[2].p, ctx->coef[0],
                     ctx->coef[1], ctx->coef[2], ctx->T[0].p, ctx->T[1].p, ctx->T[2].p,
                     f32 ? ctx->T32[0].p : nullptr, f32 ? ctx->T32[1].p : nullptr,
                     f32 ? ctx->T32[2].p : nullptr);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

int launch_mass(pmx_ctx* ctx) {
  hipLaunchKernelGGL(k_mass, dim3(ctx->nsub), dim3(kBlock), 0, ctx->stream, ctx->geo, ctx->eg_op,
                     ctx->kappa[0].p, ctx->kappa[1].p, ctx->kappa[2].p, ctx->T[0].p, ctx->T[1].p,
                     ctx->T[2].p, ctx->Hx, ctx->Hy, ctx->Hz, ctx->meas_elem, ctx->mass.p);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

// Band Cholesky of the level-1 block of brick blockIdx.x, left-looking by
// columns: column j of L from A(j:j+bw, j) minus the products with the
// previous bw columns (row j of L staged in LDS), then scaled by the pivot.
// Entries of A: the block diagonal (lv1_diag_order) and -T of the faces
// inside the brick.  *fail = 1 + j on a non-positive pivot.
__global__ void __launch_bounds__(kBlock) k_chol_factor(Geo g, const double* __restrict__ Tx,
                                                        const double* __restrict__ Ty,
                                                        const double* __restrict__ Tz,
                                                        double* __restrict__ Lb,
                                                        int* __restrict__ fail) {
  __shared__ double lrow[kCholMaxBw + 1];
  __shared__ double s_piv;
  const int sb = blockIdx.x;
  const Brick b = g.bricks[sb];
  const int n = b.nx * b.ny * b.nz, bw = b.nx * b.ny, W = bw + 1;
  double* L = Lb + g.l1off[sb];
  const double* TT[3] = {Tx, Ty, Tz};
  for (int j = 0; j < n; ++j) {
    // row j of L: lrow[q] = L(j, j - bw + q), 0 before the first column
    for (int q = threadIdx.x; q < bw; q += blockDim.x) {
      const int k = j - bw + q;
      lrow[q] = k >= 0 ? L[(size_t)k * W + (j - k)] : 0.0;
    }
    __syncthreads();
    const int lx = j % b.nx, ly = (j / b.nx) % b.ny, lz = j / bw;
    double val[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int d = threadIdx.x + q * (int)blockDim.x;
      val[q] = 0.0;
      if (d > bw || j + d >= n) continue;
      double a = 0.0;
      if (d == 0) {
        const int c = b.off + j;
        double lo[3], hi[3];
        bool in_hi[3];
        const int lc[3] = {lx, ly, lz};
        const int ln[3] = {b.nx, b.ny, b.nz};
        for (int e = 0; e < 3; ++e) {
          lo[e] = TT[e][c];
          const int nh = nbr_index(g, b, c, lx, ly, lz, 2 * e + 1);
          hi[e] = nh >= 0 ? TT[e][nh] : 0.0;
          in_hi[e] = lc[e] < ln[e] - 1;
        }
        a = lv1_diag_order(g, c, lo, hi, in_hi);
      } else {
        // the upper neighbours of cell j inside the brick (the lower face of
        // the neighbour is their shared face)
        if (d == 1 && lx + 1 < b.nx) a -= Tx[b.off + j + 1];
        if (d == b.nx && ly + 1 < b.ny) a -= Ty[b.off + j + b.nx];
        if (d == bw && lz + 1 < b.nz) a -= Tz[b.off + j + bw];
      }
      const int i = j + d;
      const int k0 = max(0, i - bw);
      double sacc = 0.0;
      for (int k = k0; k < j; ++k) sacc += L[(size_t)k * W + (i - k)] * lrow[k - (j - bw)];
      val[q] = a - sacc;
    }
    if (threadIdx.x == 0) {
      const double p = val[0];
      if (!(p > 0.0) && *fail == 0) *fail = 1 + j;
      s_piv = p > 0.0 ? sqrt(p) : 1.0;
    }
    __syncthreads();
    const double piv = s_piv;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int d = threadIdx.x + q * (int)blockDim.x;
      if (d > bw) continue;
      L[(size_t)j * W + d] = j + d >= n ? 0.0 : (d == 0 ? piv : val[q] / piv);
    }
    __syncthreads();
  }
}

// Row layout of the factor: U[j*W + d] = L(j, j - d) = L[(j-d)*W + d].
__global__ void k_chol_rows(Geo g, const double* __restrict__ Lb, double* __restrict__ Ub) {
  const int sb = blockIdx.x;
  const Brick b = g.bricks[sb];
  const int n = b.nx * b.ny * b.nz, bw = b.nx * b.ny, W = bw + 1;
  const double* L = Lb + g.l1off[sb];
  double* U = Ub + g.l1off[sb];
  for (size_t e = threadIdx.x; e < (size_t)n * W; e += blockDim.x) {
    const int j = (int)(e / W), d = (int)(e % W);
    U[e] = j - d >= 0 ? L[(size_t)(j - d) * W + d] : 0.0;
  }
}

int chol_lv1_setup(pmx_ctx* ctx) {
  std::vector<long long> off(ctx->nsub + 1, 0);
  for (int s = 0; s < ctx->nsub; ++s) {
    const Brick& b = ctx->h_bricks[s];
    const long long bw = (long long)b.nx * b.ny;
    if (bw > kCholMaxBw) {
      set_error("-use_full_Cholesky_lv1: subdomain x-y faces of %lld cells exceed %d", bw,
                kCholMaxBw);
      return PMX_ERR_SUP;
    }
    off[s + 1] = off[s] + (long long)b.nx * b.ny * b.nz * (bw + 1);
  }
  PMX_TRY(ctx->l1off.alloc(ctx->nsub + 1));
  PMX_HIP(hipMemcpy(ctx->l1off.p, off.data(), off.size() * sizeof(long long),
                    hipMemcpyHostToDevice));
  PMX_TRY(ctx->l1L.alloc((size_t)off[ctx->nsub]));
  PMX_TRY(ctx->l1U.alloc((size_t)off[ctx->nsub]));
  ctx->geo.l1off = ctx->l1off.p;
  ctx->geo.l1L = ctx->l1L.p;
  ctx->geo.l1U = ctx->l1U.p;
  DBuf<int> fail;
  PMX_TRY(fail.alloc(1));
  PMX_HIP(hipMemsetAsync(fail.p, 0, sizeof(int), ctx->stream));
  hipLaunchKernelGGL(k_chol_factor, dim3(ctx->nsub), dim3(kBlock), 0, ctx->stream, ctx->geo,
                     ctx->T[0].p, ctx->T[1].p, ctx->T[2].p, ctx->l1L.p, fail.p);
  hipLaunchKernelGGL(k_chol_rows, dim3(ctx->nsub), dim3(kBlock), 0, ctx->stream, ctx->geo,
                     ctx->l1L.p, ctx->l1U.p);
  PMX_HIP(hipGetLastError());
  int h = 0;
  PMX_HIP(hipMemcpyAsync(&h, fail.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  PMX_HIP(hipStreamSynchronize(ctx->stream));
  if (h) {
    set_error("-use_full_Cholesky_lv1: a level-1 block is not positive definite (column %d)",
              h - 1);
    return PMX_ERR_USER;
  }
  return PMX_SUCCESS;
}

int launch_dic_factor(pmx_ctx* ctx, int mode, double sigma, void* fac_f, void* fac_b,
                      bool f32) {
  const size_t lds = brick_lds_bytes(ctx, sizeof(double));
  if (f32)
    hipLaunchKernelGGL(k_dic_factor<float4>, dim3(ctx->nsub), dim3(kBlock), lds, ctx->stream,
                       ctx->geo, mode, ctx->T[0].p, ctx->T[1].p, ctx->T[2].p, ctx->mass.p, sigma,
                       (float4*)fac_f, (float4*)fac_b);
  else
    hipLaunchKernelGGL(k_dic_factor<double4>, dim3(ctx->nsub), dim3(kBlock), lds, ctx->stream,
                       ctx->geo, mode, ctx->T[0].p, ctx->T[1].p, ctx->T[2].p, ctx->mass.p, sigma,
                       (double4*)fac_f, (double4*)fac_b);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

// y = B1 x.  With upd the kernel first performs the x/r update of the
// current PCG iteration (x = the PCG residual r, updated in place; xs = the
// PCG solution).
int launch_lv1_first(pmx_ctx* ctx, double* x, bool upd, double* xs, const double* p,
                     const double* w) {
  const CgCtl cg = make_cg(ctx);
  if (ctx->use_full_Cholesky_lv1) {
    auto k = upd ? k_lv1_first<double, double4, true, false, true>
                 : k_lv1_first<double, double4, false, false, true>;
    hipLaunchKernelGGL(k, dim3(ctx->nsub), dim3(kBlock), 2 * brick_lds_bytes(ctx, sizeof(double)),
                       ctx->stream, ctx->geo, x, ctx->w_y.p, (const double4*)ctx->fac_f.p,
                       (const double4*)ctx->fac_b.p, xs, p, w, cg);
    PMX_HIP(hipGetLastError());
    return PMX_SUCCESS;
  }
  const bool big = ctx->max_plane > kBlock;
  if (ctx->pc_precision == 32) {
    auto k = upd ? (big ? k_lv1_first<float, float4, true, true>
                        : k_lv1_first<float, float4, true, false>)
                 : (big ? k_lv1_first<float, float4, false, true>
                        : k_lv1_first<float, float4, false, false>);
    hipLaunchKernelGGL(k, dim3(ctx->nsub), dim3(kBlock), brick_lds_bytes(ctx, sizeof(float)),
                       ctx->stream, ctx->geo, x, ctx->w_y32.p, (const float4*)ctx->fac32_f.p,
                       (const float4*)ctx->fac32_b.p, xs, p, w, cg);
  } else {
    auto k = upd ? (big ? k_lv1_first<double, double4, true, true>
                        : k_lv1_first<double, double4, true, false>)
                 : (big ? k_lv1_first<double, double4, false, true>
                        : k_lv1_first<double, double4, false, false>);
    hipLaunchKernelGGL(k, dim3(ctx->nsub), dim3(kBlock), brick_lds_bytes(ctx, sizeof(double)),
                       ctx->stream, ctx->geo, x, ctx->w_y.p, (const double4*)ctx->fac_f.p,
                       (const double4*)ctx->fac_b.p, xs, p, w, cg);
  }
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

// keep_y: y <- y + B1 (x - A y) written to the second buffer (other bricks
// read y as halo while it is being updated), then the buffers swap.
int launch_lv1_last(pmx_ctx* ctx, const double* x, double* z, bool keep_y, bool cg_partial) {
  CgCtl cg = make_cg(ctx);
  if (!cg_partial) cg.zr = nullptr;
  const bool big = ctx->max_plane > kBlock;
  if (ctx->use_full_Cholesky_lv1) {
    hipLaunchKernelGGL((k_lv1_last<double, double, double4, false, true>), dim3(ctx->nsub),
                       dim3(kBlock), 2 * brick_lds_bytes(ctx, sizeof(double)), ctx->stream,
                       ctx->geo, x, ctx->w_y.p, ctx->T[0].p, ctx->T[1].p, ctx->T[2].p,
                       (const double4*)ctx->fac_f.p, (const double4*)ctx->fac_b.p, z,
                       keep_y ? ctx->w_yb.p : nullptr, cg);
    if (keep_y) std::swap(ctx->w_y, ctx->w_yb);
    PMX_HIP(hipGetLastError());
    return PMX_SUCCESS;
  }
  if (ctx->pc_precision == 32 && !big && ctx->quad_ok && quad_threads(ctx) != 0) {
    // quad bricks hold at most 4096 cells: four quads per thread
    hipLaunchKernelGGL((k_lv1_last_q<float, float, float4>), dim3(ctx->nsub), dim3(kBlock),
                       2 * brick_lds_bytes(ctx, sizeof(float)), ctx->stream, ctx->geo, x,
                       ctx->w_y32.p, ctx->T32[0].p, ctx->T32[1].p, ctx->T32[2].p,
                       (const float4*)ctx->fac32_f.p, (const float4*)ctx->fac32_b.p, z,
                       keep_y ? ctx->w_y32b.p : nullptr, cg);
    if (keep_y) std::swap(ctx->w_y32, ctx->w_y32b);
    PMX_HIP(hipGetLastError());
    return PMX_SUCCESS;
  }
  if (ctx->pc_precision == 32) {
    auto k = big ? k_lv1_last<float, float, float4, true> : k_lv1_last<float, float, float4, false>;
    hipLaunchKernelGGL(k, dim3(ctx->nsub), dim3(kBlock), brick_lds_bytes(ctx, sizeof(float)),
                       ctx->stream, ctx->geo, x, ctx->w_y32.p, ctx->T32[0].p, ctx->T32[1].p,
                       ctx->T32[2].p, (const float4*)ctx->fac32_f.p,
                       (const float4*)ctx->fac32_b.p, z, keep_y ? ctx->w_y32b.p : nullptr, cg);
    if (keep_y) std::swap(ctx->w_y32, ctx->w_y32b);
  } else {
    auto k = big ? k_lv1_last<double, double, double4, true>
                 : k_lv1_last<double, double, double4, false>;
    hipLaunchKernelGGL(k, dim3(ctx->nsub), dim3(kBlock), brick_lds_bytes(ctx, sizeof(double)),
                       ctx->stream, ctx->geo, x, ctx->w_y.p, ctx->T[0].p, ctx->T[1].p,
                       ctx->T[2].p, (const double4*)ctx->fac_f.p, (const double4*)ctx->fac_b.p,
                       z, keep_y ? ctx->w_yb.p : nullptr, cg);
    if (keep_y) std::swap(ctx->w_y, ctx->w_yb);
  }
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

template <typename S, typename TT>
static bool launch_resid_restrict_q(pmx_ctx* ctx, const double* x, const S* y, const TT* Tx,
                                    const TT* Ty, const TT* Tz, const S* phi, double* cvec) {
  const int qt = ctx->quad_ok ? quad_threads(ctx) : 0;
  if (!qt) return false;
  size_t lds = 0;
  for (const Brick& b : ctx->h_bricks) lds = std::max(lds, (size_t)b.nx * b.ny * b.nz);
  lds *= sizeof(S);
  switch (ctx->en_lv1) {
    case 4:
      hipLaunchKernelGGL((k_resid_restrict_q<S, TT, 4>), dim3(ctx->nsub), dim3(qt), lds,
                         ctx->stream, ctx->geo, x, y, Tx, Ty, Tz, phi, cvec);
      return true;
    case 6:
      hipLaunchKernelGGL((k_resid_restrict_q<S, TT, 6>), dim3(ctx->nsub), dim3(qt), lds,
                         ctx->stream, ctx->geo, x, y, Tx, Ty, Tz, phi, cvec);
      return true;
    case 8:
      hipLaunchKernelGGL((k_resid_restrict_q<S, TT, 8>), dim3(ctx->nsub), dim3(qt), lds,
                         ctx->stream, ctx->geo, x, y, Tx, Ty, Tz, phi, cvec);
      return true;
    default:
      return false;
  }
}

int launch_resid_restrict(pmx_ctx* ctx, const double* x, bool y_zero, double* cvec) {
  const int nt = stream_threads(ctx);
  // (FP32 only: the FP64 quad variant needs more registers than the
  // occupancy of 1024-thread workgroups allows.)
  const bool done =
      ctx->pc_precision == 32 &&
      launch_resid_restrict_q<float, float>(ctx, x, y_zero ? nullptr : ctx->w_y32.p,
                                            ctx->T32[0].p, ctx->T32[1].p, ctx->T32[2].p,
                                            ctx->phi32.p, cvec);
  if (done) {
    PMX_HIP(hipGetLastError());
    return PMX_SUCCESS;
  }
  if (ctx->pc_precision == 32)
    hipLaunchKernelGGL((k_resid_restrict<float, float>), dim3(ctx->nsub), dim3(nt), 0,
                       ctx->stream, ctx->geo, x, y_zero ? nullptr : ctx->w_y32.p, ctx->T32[0].p,
                       ctx->T32[1].p, ctx->T32[2].p, ctx->phi32.p, ctx->en_lv1, cvec);
  else
    hipLaunchKernelGGL((k_resid_restrict<double, double>), dim3(ctx->nsub), dim3(nt), 0,
                       ctx->stream, ctx->geo, x, y_zero ? nullptr : ctx->w_y.p, ctx->T[0].p,
                       ctx->T[1].p, ctx->T[2].p, ctx->phi.p, ctx->en_lv1, cvec);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

// c = Phi^T x with the FP64 basis (y = 0), whatever the preconditioner precision.
int launch_resid_restrict_f64(pmx_ctx* ctx, const double* x, double* cvec) {
  hipLaunchKernelGGL((k_resid_restrict<double, double>), dim3(ctx->nsub),
                     dim3(stream_threads(ctx)), 0, ctx->stream, ctx->geo, x, nullptr, ctx->T[0].p,
                     ctx->T[1].p, ctx->T[2].p, ctx->phi.p, ctx->en_lv1, cvec);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

template <typename S>
static bool launch_prolong_q(pmx_ctx* ctx, S* y, const S* phi, const double* cvec, double* iter) {
  const int qt = ctx->quad_ok ? quad_threads(ctx) : 0;
  if (!qt) return false;
  switch (ctx->en_lv1) {
    case 4:
      hipLaunchKernelGGL((k_prolong_q<S, 4>), dim3(ctx->nsub), dim3(qt), 0, ctx->stream, ctx->geo,
                         y, phi, cvec, iter);
      return true;
    case 6:
      hipLaunchKernelGGL((k_prolong_q<S, 6>), dim3(ctx->nsub), dim3(qt), 0, ctx->stream, ctx->geo,
                         y, phi, cvec, iter);
      return true;
    case 8:
      hipLaunchKernelGGL((k_prolong_q<S, 8>), dim3(ctx->nsub), dim3(qt), 0, ctx->stream, ctx->geo,
                         y, phi, cvec, iter);
      return true;
    default:
      return false;
  }
}

int launch_prolong(pmx_ctx* ctx, const double* cvec, bool count_iter) {
  double* iter = count_iter ? ctx->cg_scal + kSIter : nullptr;
  const int nt = stream_threads(ctx);
  const bool done = ctx->pc_precision == 32
                        ? launch_prolong_q<float>(ctx, ctx->w_y32.p, ctx->phi32.p, cvec, iter)
                        : launch_prolong_q<double>(ctx, ctx->w_y.p, ctx->phi.p, cvec, iter);
  if (done) {
    PMX_HIP(hipGetLastError());
    return PMX_SUCCESS;
  }
  if (ctx->pc_precision == 32)
    hipLaunchKernelGGL(k_prolong<float>, dim3(ctx->nsub), dim3(nt), 0, ctx->stream, ctx->geo,
                       ctx->w_y32.p, ctx->phi32.p, ctx->en_lv1, cvec, iter);
  else
    hipLaunchKernelGGL(k_prolong<double>, dim3(ctx->nsub), dim3(nt), 0, ctx->stream, ctx->geo,
                       ctx->w_y.p, ctx->phi.p, ctx->en_lv1, cvec, iter);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

// Quad-cell stream kernels apply when every brick's x-extent is a multiple
// of 4 and a brick fits in 4 * kStreamBlock cells; returns their block size
// (0: use the per-cell kernels).
int quad_threads(const pmx_ctx* ctx) {
  int mx = 0;
  for (const Brick& b : ctx->h_bricks) {
    if (b.nx % 4 != 0) return 0;
    mx = std::max(mx, b.nx * b.ny * b.nz);
  }
  if (mx > 4 * kStreamBlock) return 0;
  return std::max(64, (mx / 4 + 63) / 64 * 64);
}

int launch_cg_p_spmv(pmx_ctx* ctx, const double* z, const double* pold, double* pnew, double* w) {
  const int qt = ctx->quad_ok ? quad_threads(ctx) : 0;
  if (qt) {
    size_t lds = 0;
    for (const Brick& b : ctx->h_bricks) lds = std::max(lds, (size_t)b.nx * b.ny * b.nz);
    hipLaunchKernelGGL(k_cg_p_spmv_q, dim3(ctx->nsub), dim3(qt), lds * sizeof(double),
                       ctx->stream, ctx->geo, z, pold, pnew, w, ctx->T[0].p, ctx->T[1].p,
                       ctx->T[2].p, make_cg(ctx));
  } else {
    hipLaunchKernelGGL(k_cg_p_spmv, dim3(ctx->nsub), dim3(stream_threads(ctx)), 0, ctx->stream,
                       ctx->geo, z, pold, pnew, w, ctx->T[0].p, ctx->T[1].p, ctx->T[2].p,
                       make_cg(ctx));
  }
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

// r = b - A x (x == nullptr: x = 0), then the PCG start test.
int launch_residual(pmx_ctx* ctx, const double* b, const double* x, double* r) {
  double* part = ctx->cg_part + 4 * (size_t)ctx->nsub;
  hipLaunchKernelGGL(k_residual, dim3(ctx->nsub), dim3(stream_threads(ctx)), 0, ctx->stream,
                     ctx->geo, b, x, r, ctx->T[0].p, ctx->T[1].p, ctx->T[2].p, part);
  PMX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_cg_start, dim3(1), dim3(64), 0, ctx->stream, part,
                     x ? part + ctx->nsub : nullptr, ctx->nsub, ctx->cg_scal, ctx->cg_rtol,
                     ctx->cg_abstol);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

__global__ void k_to_f32(const double* __restrict__ src, float* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}

int launch_to_f32(pmx_ctx* ctx, const double* src, float* dst, size_t n) {
  const int nb = (int)std::min<size_t>((n + kBlock - 1) / kBlock, 8192);
  hipLaunchKernelGGL(k_to_f32, dim3(std::max(nb, 1)), dim3(kBlock), 0, ctx->stream, src, dst, n);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

int launch_spmv(pmx_ctx* ctx, const double* x, double* y) {
  hipLaunchKernelGGL(k_spmv, dim3(ctx->nsub), dim3(stream_threads(ctx)), 0, ctx->stream,
                     ctx->geo, x, y, ctx->T[0].p, ctx->T[1].p, ctx->T[2].p);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

int launch_velocity(pmx_ctx* ctx, const double* p, double* fx, double* fy, double* fz) {
  hipLaunchKernelGGL(k_velocity, dim3(ctx->nsub), dim3(stream_threads(ctx)), 0, ctx->stream,
                     ctx->geo, p, ctx->T[0].p, ctx->T[1].p, ctx->T[2].p, fx, fy, fz);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

int launch_check_field(pmx_ctx* ctx, const double* a, int n, bool strict, int* bad) {
  DBuf<int> flag;
  PMX_TRY(flag.alloc(1));
  const int init = 0x7fffffff;
  PMX_HIP(hipMemcpyAsync(flag.p, &init, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  const int nb = std::max(1, std::min((n + kBlock - 1) / kBlock, 4096));
  hipLaunchKernelGGL(k_check_field, dim3(nb), dim3(kBlock), 0, ctx->stream, a, n, strict ? 1 : 0,
                     flag.p);
  PMX_HIP(hipGetLastError());
  int h = init;
  PMX_HIP(hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  PMX_HIP(hipStreamSynchronize(ctx->stream));
  *bad = h == init ? -1 : h;
  return PMX_SUCCESS;
}

}  // namespace pmx
