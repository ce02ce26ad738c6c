// Batched subdomain LOBPCG on the FP64 matrix cores (v_mfma_f64_16x16x4f64).
// Same eigenproblems as the VALU kernel in pmx_lobpcg.hip
// (src/PreMixFEM_3D.c:224-421, one workgroup per subdomain, the constant
// deflated, the FP32 block ICC(0) preconditioner, ||T r||_M <= tol, the
// same start blocks, outputs and statistics), with the passes over the
// subdomain's blocks regrouped so that an iteration streams them about 45
// times instead of about 75, and every Gram and block transform runs on MFMA:
//
// * R:  W = T (A X - M X Lambda): the residual is formed while the FP32 sweep
//       buffer is loaded; after the sweeps W is deflated, and W, A W and the
//       M-norms of W come from that buffer (reads X, AX; writes W, AW).
// * O:  W and AW projected against [X P] twice (classical Gram-Schmidt in
//       the M inner product) and SVQB-normalised; the second projection and
//       the SVQB transform are one pass.
// * G:  the full Grams S^T M S and S^T A S of S = [X W P].
// * RR: Cholesky-reduced Rayleigh-Ritz; the new P = [W P] C_p is
//       M-orthogonalised against the new X and SVQB-normalised on its
//       coefficients through S^T M S.
// * U:  X, P and their A-images in one pass over S and A S.
#include <algorithm>
#include <cmath>

#include "pmx_lobpcg.h"

namespace pmx {

typedef double mfma_d4 __attribute__((ext_vector_type(4)));   // one MFMA 16x16 f64 tile per lane

constexpr int kKP = 32;   // rows / columns of the coefficient matrix of a block update
constexpr int kMfPlanes = 255;   // hyperplanes whose offsets are kept in LDS

// Columns per sweep chunk: the whole block up to 8 (one sweep pair per
// iteration), else two halves; the FP32 chunk of a 16^3 subdomain then takes
// at most 128 KB of LDS.
constexpr int lobpcg_mf_ch(int mb) { return mb <= 8 ? mb : (mb + 1) / 2; }

// Up to three column-major blocks (stride n) seen as one set of columns.
struct ColSet {
  const double* p[3];
  int c[3];
};
struct OutSet {
  double* p[2];
  int c[2];
};

__device__ __forceinline__ ColSet colset(const double* a, int na, const double* b = nullptr,
                                         int nb = 0, const double* d = nullptr, int nd = 0) {
  ColSet s;
  s.p[0] = a;
  s.c[0] = na;
  s.p[1] = b;
  s.c[1] = nb;
  s.p[2] = d;
  s.c[2] = nd;
  return s;
}
__device__ __forceinline__ OutSet outset(double* a, int na, double* b = nullptr, int nb = 0) {
  OutSet s;
  s.p[0] = a;
  s.c[0] = na;
  s.p[1] = b;
  s.c[1] = nb;
  return s;
}
// Column j of the set (nullptr beyond its columns).
__device__ __forceinline__ const double* col_of(const ColSet& s, int j, int n) {
  if (j < s.c[0]) return s.p[0] + (size_t)j * n;
  j -= s.c[0];
  if (j < s.c[1]) return s.p[1] + (size_t)j * n;
  j -= s.c[1];
  if (j < s.c[2]) return s.p[2] + (size_t)j * n;
  return nullptr;
}
__device__ __forceinline__ double* col_of(const OutSet& s, int j, int n) {
  if (j < s.c[0]) return s.p[0] + (size_t)j * n;
  j -= s.c[0];
  if (j < s.c[1]) return s.p[1] + (size_t)j * n;
  return nullptr;
}

// out[q] = workgroup sum of v[q], q < cnt.  sh: (waves x NQ) doubles.
template <int NQ>
__device__ __forceinline__ void wg_sum(const double (&v)[NQ], int cnt, double* sh, double* out) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double s = v[q];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) sh[wv * NQ + q] = s;
  }
  __syncthreads();
  if ((int)threadIdx.x < cnt) {
    double s = 0.0;
    for (int q = 0; q < nwv; ++q) s += sh[q * NQ + threadIdx.x];
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

// The block columns live in global memory.  The helpers below are separate
// functions that receive them through ColSet / OutSet references, so the
// compiler cannot infer their address space and would emit flat accesses,
// which count against the LDS counter too: a wait for an LDS read then also
// waits for every global load in flight, and a tile loaded ahead is waited
// for before the current one is used.  Casting to the global address space
// gives global loads and stores, tracked by the vector-memory counter alone.
typedef const double __attribute__((address_space(1))) gcdouble;
typedef double __attribute__((address_space(1))) gdouble;
__device__ __forceinline__ const gcdouble* gptr(const double* p) {
  return (const gcdouble*)p;
}
__device__ __forceinline__ gdouble* gptr(double* p) { return (gdouble*)p; }
typedef double gd2 __attribute__((ext_vector_type(2)));
typedef const gd2 __attribute__((address_space(1))) gcd2;
// two consecutive doubles at a 16-byte aligned global address
__device__ __forceinline__ gd2 ld2(const gcdouble* p) { return *reinterpret_cast<const gcd2*>(p); }

// Steps [s0, s1) of wg_gram on the 16-byte path (all 16 rows present):
// lane group g loads rows r0 + 2g, r0 + 2g + 1, r0 + 8 + 2g, r0 + 9 + 2g of
// its columns as two 16-byte pairs, the left operand weighted when WGT; the
// next step goes into the other buffer before this one is multiplied.
template <int NTL, int NTR, bool WGT, class Mul>
__device__ __forceinline__ void gram_steps(const gcdouble* const (&lp)[NTL],
                                           const gcdouble* const (&rp)[NTR], const gcdouble* wp,
                                           int g, int s0, int s1, Mul& mul) {
  auto load = [&](int st, double (&av)[NTL][4], double (&bv)[NTR][4]) {
    const int ra = (st << 4) + 2 * g, rb = ra + 8;
    gd2 ma = {1.0, 1.0}, mb = {1.0, 1.0};
    if (WGT) {
      ma = ld2(wp + ra);
      mb = ld2(wp + rb);
    }
#pragma unroll
    for (int t = 0; t < NTL; ++t) {
      const gd2 u = ld2(lp[t] + ra), z = ld2(lp[t] + rb);
      av[t][0] = WGT ? u.x * ma.x : u.x;
      av[t][1] = WGT ? u.y * ma.y : u.y;
      av[t][2] = WGT ? z.x * mb.x : z.x;
      av[t][3] = WGT ? z.y * mb.y : z.y;
    }
#pragma unroll
    for (int t = 0; t < NTR; ++t) {
      const gd2 u = ld2(rp[t] + ra), z = ld2(rp[t] + rb);
      bv[t][0] = u.x;
      bv[t][1] = u.y;
      bv[t][2] = z.x;
      bv[t][3] = z.y;
    }
  };
  if ((s1 - s0) & 1) {   // an odd step first, on its own
    double a[NTL][4], b[NTR][4];
    load(s0, a, b);
    mul(a, b);
    ++s0;
  }
  if (s0 >= s1) return;
  // Pairs of steps with both buffers in flight: each refill is issued right
  // after its buffer is multiplied, clamped into the range (the last refills
  // reload steps already used, and are not multiplied), so every wait covers
  // exactly the buffer about to be used.
  double a0[NTL][4], b0[NTR][4], a1[NTL][4], b1[NTR][4];
  load(s0, a0, b0);
  load(s0 + 1, a1, b1);
  for (int st = s0; st < s1; st += 2) {
    mul(a0, b0);
    load(min(st + 2, s1 - 2), a0, b0);
    mul(a1, b1);
    load(min(st + 3, s1 - 1), a1, b1);
  }
}

// G[i * ldg + j] = sum_r L_i[r] wgt[r] R_j[r] for i < nl, j < nr
// (nl <= 16 NTL, nr <= 16 NTR; wgt may be null).  Each wave takes a
// contiguous range of 16-row steps; within a step lane group g = lane / 16
// feeds rows 8 (i / 2) + 2 g + i % 2 to MFMA i (i = 0..3), so every load is
// a 16-byte pair of rows (vec: n and the subdomain offset even, which makes
// every column and the weights 16-byte aligned).  The next step's rows are
// loaded while the current step is multiplied.  Lanes beyond the set's
// columns read its first column (finite data) and their entries are dropped;
// rows beyond n are read clamped and zeroed.  The waves' partial Grams are
// added in a fixed tree order through red ((waves / 2) * NTL * NTR * 256
// doubles), so the result does not depend on timing.
template <int NTL, int NTR>
__device__ void wg_gram(const ColSet& L, int nl, const ColSet& R, int nr,
                        const double* __restrict__ wgt, int n, bool vec, double* red, double* G,
                        int ldg) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const int cl = lane & 15, g = lane >> 4;
  const gcdouble* lp[NTL];
  const gcdouble* rp[NTR];
#pragma unroll
  for (int t = 0; t < NTL; ++t) {
    const int c = cl + 16 * t;
    lp[t] = gptr(col_of(L, c < nl ? c : 0, n));
  }
#pragma unroll
  for (int t = 0; t < NTR; ++t) {
    const int c = cl + 16 * t;
    rp[t] = gptr(col_of(R, c < nr ? c : 0, n));
  }
  const gcdouble* wp = wgt ? gptr(wgt) : nullptr;
  mfma_d4 acc[NTL][NTR];
#pragma unroll
  for (int tl = 0; tl < NTL; ++tl)
#pragma unroll
    for (int tr = 0; tr < NTR; ++tr) acc[tl][tr] = mfma_d4{0.0, 0.0, 0.0, 0.0};
  auto mul = [&](const double (&av)[NTL][4], const double (&bv)[NTR][4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int tl = 0; tl < NTL; ++tl)
#pragma unroll
        for (int tr = 0; tr < NTR; ++tr)
          acc[tl][tr] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[tl][i], bv[tr][i], acc[tl][tr], 0,
                                                             0, 0);
  };
  // Full 16-row steps on the 16-byte path: branch-free loads of the next
  // step into the other of two buffers while this one is multiplied.
  const int full = vec ? n >> 4 : 0;
  if (full > 0) {
    const int per = (full + nwv - 1) / nwv;
    const int s0 = wv * per, s1 = min(full, s0 + per);
    if (wp)
      gram_steps<NTL, NTR, true>(lp, rp, wp, g, s0, s1, mul);
    else
      gram_steps<NTL, NTR, false>(lp, rp, wp, g, s0, s1, mul);
  }
  // Remaining rows (every row when not on the 16-byte path): row by row,
  // clamped and zeroed beyond n, one step per wave in turn.
  const int steps = (n + 15) >> 4;
  for (int st = full + wv; st < steps; st += nwv) {
    const int r0 = st << 4;
    double av[NTL][4], bv[NTR][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = r0 + 8 * (i >> 1) + 2 * g + (i & 1);
      const bool ok = r < n;
      const int rc = ok ? r : n - 1;
      const double m = wp ? wp[rc] : 1.0;
#pragma unroll
      for (int t = 0; t < NTL; ++t) {
        const double x = lp[t][rc];
        av[t][i] = ok ? x * m : 0.0;
      }
#pragma unroll
      for (int t = 0; t < NTR; ++t) {
        const double x = rp[t][rc];
        bv[t][i] = ok ? x : 0.0;
      }
    }
    mul(av, bv);
  }
  constexpr int NE = NTL * NTR * 4;
  for (int h = nwv >> 1; h >= 1; h >>= 1) {
    if (wv >= h && wv < 2 * h) {
      double* dst = red + (size_t)(wv - h) * NE * 64;
#pragma unroll
      for (int tl = 0; tl < NTL; ++tl)
#pragma unroll
        for (int tr = 0; tr < NTR; ++tr)
#pragma unroll
          for (int q = 0; q < 4; ++q) dst[((tl * NTR + tr) * 4 + q) * 64 + lane] = acc[tl][tr][q];
    }
    __syncthreads();
    if (wv < h) {
      const double* src = red + (size_t)wv * NE * 64;
#pragma unroll
      for (int tl = 0; tl < NTL; ++tl)
#pragma unroll
        for (int tr = 0; tr < NTR; ++tr)
#pragma unroll
          for (int q = 0; q < 4; ++q) acc[tl][tr][q] += src[((tl * NTR + tr) * 4 + q) * 64 + lane];
    }
    __syncthreads();
  }
  if (wv == 0) {
#pragma unroll
    for (int tl = 0; tl < NTL; ++tl)
#pragma unroll
      for (int tr = 0; tr < NTR; ++tr)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int i = 16 * tl + g + 4 * q, j = 16 * tr + cl;
          if (i < nl && j < nr) G[i * ldg + j] = acc[tl][tr][q];
        }
  }
  __syncthreads();
}

// GB = S^T diag(wgt) S and GA = S^T (A S) in one pass over S and A S
// (ns <= 16 NT columns each; GB[i * ldg + j], GA[i * ldg + j]): every row of
// S is loaded once and serves as the left operand of both products and,
// weighted, as the right operand of the first.  Same step layout,
// pipelining and fixed-order wave reduction as wg_gram (the two Grams
// reduced one after the other through red).  HAS_A = false computes GB
// alone.
template <int NT, bool HAS_A = true>
__device__ void wg_gram_ma(const ColSet& S, const ColSet& AS, int ns,
                           const double* __restrict__ wgt, int n, bool vec, double* red,
                           double* GB, double* GA, int ldg) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const int cl = lane & 15, g = lane >> 4;
  const gcdouble* sp[NT];
  const gcdouble* ap[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int c = cl + 16 * t < ns ? cl + 16 * t : 0;
    sp[t] = gptr(col_of(S, c, n));
    ap[t] = HAS_A ? gptr(col_of(AS, c, n)) : sp[t];
  }
  const gcdouble* wp = gptr(wgt);
  mfma_d4 accB[NT][NT], accA[NT][NT];
#pragma unroll
  for (int tl = 0; tl < NT; ++tl)
#pragma unroll
    for (int tr = 0; tr < NT; ++tr) {
      accB[tl][tr] = mfma_d4{0.0, 0.0, 0.0, 0.0};
      accA[tl][tr] = mfma_d4{0.0, 0.0, 0.0, 0.0};
    }
  // sv: rows of S, av: rows of A S, m: weights of the four rows
  auto mul = [&](const double (&sv)[NT][4], const double (&av)[NT][4], const double (&m)[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double sw[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) sw[t] = sv[t][i] * m[i];
#pragma unroll
      for (int tl = 0; tl < NT; ++tl)
#pragma unroll
        for (int tr = 0; tr < NT; ++tr) {
          accB[tl][tr] = __builtin_amdgcn_mfma_f64_16x16x4f64(sv[tl][i], sw[tr], accB[tl][tr], 0,
                                                              0, 0);
          if (HAS_A)
            accA[tl][tr] = __builtin_amdgcn_mfma_f64_16x16x4f64(sv[tl][i], av[tr][i],
                                                                accA[tl][tr], 0, 0, 0);
        }
    }
  };
  const int full = vec ? n >> 4 : 0;
  if (full > 0) {
    const int per = (full + nwv - 1) / nwv;
    int s0 = wv * per;
    const int s1 = min(full, s0 + per);
    auto load = [&](int st, double (&sv)[NT][4], double (&av)[NT][4], double (&m)[4]) {
      const int ra = (st << 4) + 2 * g, rb = ra + 8;
      const gd2 ma = ld2(wp + ra), mb = ld2(wp + rb);
      m[0] = ma.x;
      m[1] = ma.y;
      m[2] = mb.x;
      m[3] = mb.y;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const gd2 u = ld2(sp[t] + ra), z = ld2(sp[t] + rb);
        sv[t][0] = u.x;
        sv[t][1] = u.y;
        sv[t][2] = z.x;
        sv[t][3] = z.y;
        if (HAS_A) {
          const gd2 x = ld2(ap[t] + ra), y = ld2(ap[t] + rb);
          av[t][0] = x.x;
          av[t][1] = x.y;
          av[t][2] = y.x;
          av[t][3] = y.y;
        }
      }
    };
    if (s0 < s1 && ((s1 - s0) & 1)) {   // an odd step first, on its own
      double sv[NT][4], av[NT][4], m[4];
      load(s0, sv, av, m);
      mul(sv, av, m);
      ++s0;
    }
    if (s0 < s1) {
      // pairs of steps with both buffers in flight, refills clamped into the
      // range (the last ones reload used steps and are not multiplied)
      double s0v[NT][4], a0v[NT][4], m0[4], s1v[NT][4], a1v[NT][4], m1[4];
      load(s0, s0v, a0v, m0);
      load(s0 + 1, s1v, a1v, m1);
      for (int st = s0; st < s1; st += 2) {
        mul(s0v, a0v, m0);
        load(min(st + 2, s1 - 2), s0v, a0v, m0);
        mul(s1v, a1v, m1);
        load(min(st + 3, s1 - 1), s1v, a1v, m1);
      }
    }
  }
  // remaining rows (every row off the 16-byte path), clamped and zeroed
  const int steps = (n + 15) >> 4;
  for (int st = full + wv; st < steps; st += nwv) {
    const int r0 = st << 4;
    double sv[NT][4], av[NT][4], m[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = r0 + 8 * (i >> 1) + 2 * g + (i & 1);
      const bool ok = r < n;
      const int rc = ok ? r : n - 1;
      m[i] = ok ? wp[rc] : 0.0;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const double x = sp[t][rc];
        sv[t][i] = ok ? x : 0.0;
        if (HAS_A) {
          const double y = ap[t][rc];
          av[t][i] = ok ? y : 0.0;
        }
      }
    }
    mul(sv, av, m);
  }
  constexpr int NE = NT * NT * 4;
  auto reduce_store = [&](mfma_d4 (&acc)[NT][NT], double* G) {
    for (int h = nwv >> 1; h >= 1; h >>= 1) {
      if (wv >= h && wv < 2 * h) {
        double* dst = red + (size_t)(wv - h) * NE * 64;
#pragma unroll
        for (int tl = 0; tl < NT; ++tl)
#pragma unroll
          for (int tr = 0; tr < NT; ++tr)
#pragma unroll
            for (int q = 0; q < 4; ++q) dst[((tl * NT + tr) * 4 + q) * 64 + lane] = acc[tl][tr][q];
      }
      __syncthreads();
      if (wv < h) {
        const double* src = red + (size_t)wv * NE * 64;
#pragma unroll
        for (int tl = 0; tl < NT; ++tl)
#pragma unroll
          for (int tr = 0; tr < NT; ++tr)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[tl][tr][q] += src[((tl * NT + tr) * 4 + q) * 64 + lane];
      }
      __syncthreads();
    }
    if (wv == 0) {
#pragma unroll
      for (int tl = 0; tl < NT; ++tl)
#pragma unroll
        for (int tr = 0; tr < NT; ++tr)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int i = 16 * tl + g + 4 * q, j = 16 * tr + cl;
            if (i < ns && j < ns) G[i * ldg + j] = acc[tl][tr][q];
          }
    }
    __syncthreads();
  };
  reduce_store(accB, GB);
  if (HAS_A) reduce_store(accA, GA);
}

// Out(:, c) = sum_k In(:, k) CF[k * kKP + c] for c < nout (<= 16 NTO),
// k < nin (<= 32), CF in LDS.  Computed as D^T = CF^T In^T: one 16-row tile
// per step, lane l % 16 a row and l / 16 a column group, so loads and stores
// are 128-byte runs of a column.  The coefficients are read into registers
// once; the next tile is loaded before this tile is multiplied and stored.
// Lanes beyond the input columns read its first column (finite data) against
// a zero coefficient; rows beyond n are read clamped and not stored.  Out may
// share columns with In: every row tile is read whole by the wave that owns
// it before that wave writes it.
template <int NTO>
__device__ void wg_update(const ColSet& In, int nin, const double* CF, const OutSet& Out,
                          int nout, int n) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const int cl = lane & 15, g = lane >> 4;
  const gcdouble* ip[8];
  double af[8][NTO];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int k = 4 * q + g;
    ip[q] = gptr(col_of(In, k < nin ? k : 0, n));
#pragma unroll
    for (int t = 0; t < NTO; ++t) {
      const int c = 16 * t + cl;
      af[q][t] = (k < nin && c < nout) ? CF[k * kKP + c] : 0.0;
    }
  }
  gdouble* op[NTO][4];
#pragma unroll
  for (int t = 0; t < NTO; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int c = 16 * t + g + 4 * v;
      op[t][v] = c < nout ? gptr(col_of(Out, c, n)) : nullptr;
    }
  const int tiles = (n + 15) >> 4, per = (tiles + nwv - 1) / nwv;
  const int t0 = wv * per, t1 = min(tiles, t0 + per);
  auto load = [&](int tt, double (&dst)[8]) {
    const int row = min((tt << 4) + cl, n - 1);
#pragma unroll
    for (int q = 0; q < 8; ++q) dst[q] = ip[q][row];
  };
  auto tile = [&](int tt, const double (&bv)[8]) {
    mfma_d4 acc[NTO];
#pragma unroll
    for (int t = 0; t < NTO; ++t) acc[t] = mfma_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 8; ++q)
#pragma unroll
      for (int t = 0; t < NTO; ++t)
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[q][t], bv[q], acc[t], 0, 0, 0);
    const int row = (tt << 4) + cl;
    if (row < n) {
#pragma unroll
      for (int t = 0; t < NTO; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (op[t][v]) op[t][v][row] = acc[t][v];
    }
  };
  int ta = t0;
  if ((t1 - ta) & 1) {   // an odd tile first, on its own
    double b[8];
    load(ta, b);
    tile(ta, b);
    ++ta;
  }
  if (ta < t1) {
    // Pairs of tiles with both buffers in flight: each refill is issued right
    // after its buffer's tile is stored, clamped into the range (the last
    // refills reload tiles already done, and are not used), so every wait
    // covers exactly the buffer about to be multiplied.  A refill reads rows
    // of a later tile than the ones stored before it.
    double b0[8], b1[8];
    load(ta, b0);
    load(ta + 1, b1);
    for (int tt = ta; tt < t1; tt += 2) {
      tile(tt, b0);
      load(min(tt + 2, t1 - 2), b0);
      tile(tt + 1, b1);
      load(min(tt + 3, t1 - 1), b1);
    }
  }
  __syncthreads();
}

// The forward and backward FP32 block ICC(0) sweeps on ch <= CH columns in
// LDS (the recurrences of wg_sweeps).  Each thread loads its cell (packed
// position and factor) of the next hyperplane before working on the current
// one, and reads the neighbours of all its columns before writing any, so a
// plane costs one LDS round trip rather than one per column; pl holds the
// plane offsets (in LDS when they fit).  A plane wider than the workgroup
// loads its extra cells on the spot.
template <int CH>
__device__ void mf_sweeps(const BrickView& v, int nplanes, const int* pl, const int* pc,
                          const float4* __restrict__ pf, const float4* __restrict__ pb, float* lv,
                          int ch) {
  const int n = v.n, tid = threadIdx.x, bs = blockDim.x;
  for (int dirn = 0; dirn < 2; ++dirn) {
    const bool fwd = dirn == 0;
    const float4* fac = (fwd ? pf : pb) + v.off;
    int s = fwd ? 0 : nplanes - 1;
    int pos = pl[s] + tid, end = pl[s + 1];
    int pk = 0;
    float4 f = {0.f, 0.f, 0.f, 0.f};
    if (pos < end) {
      pk = pc[pos];
      f = fac[pos];
    }
    for (int k = 0; k < nplanes; ++k) {
      int npos = 0, nend = 0, npk = 0;
      float4 nf = {0.f, 0.f, 0.f, 0.f};
      if (k + 1 < nplanes) {
        const int s2 = fwd ? k + 1 : nplanes - 2 - k;
        npos = pl[s2] + tid;
        nend = pl[s2 + 1];
        if (npos < nend) {
          npk = pc[npos];
          nf = fac[npos];
        }
      }
      for (int p = pos; p < end; p += bs) {
        int pkk = pk;
        float4 ff = f;
        if (p != pos) {
          pkk = pc[p];
          ff = fac[p];
        }
        const int lx = pkk & 1023, ly = (pkk >> 10) & 1023, lz = pkk >> 20;
        const int li = (lz * v.ny + ly) * v.nx + lx;
        // neighbour offsets (0: the term is absent, its coefficient made 0)
        int oa, ob, oc;
        float ca, cb, cc;
        if (fwd) {
          oa = lx > 0 ? -1 : 0;
          ob = ly > 0 ? -v.nx : 0;
          oc = lz > 0 ? -v.nxny : 0;
        } else {
          oa = lx < v.nx - 1 ? 1 : 0;
          ob = ly < v.ny - 1 ? v.nx : 0;
          oc = lz < v.nz - 1 ? v.nxny : 0;
        }
        ca = oa ? ff.y : 0.0f;
        cb = ob ? ff.z : 0.0f;
        cc = oc ? ff.w : 0.0f;
        float xs[CH], xa[CH], xb[CH], xc[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) {
          if (q < ch) {
            const float* vv = lv + q * n + li;
            xs[q] = vv[0];
            xa[q] = vv[oa];
            xb[q] = vv[ob];
            xc[q] = vv[oc];
          }
        }
#pragma unroll
        for (int q = 0; q < CH; ++q) {
          if (q < ch) {
            float r;
            if (fwd) {
              r = ((xs[q] + ca * xa[q]) + cb * xb[q]) + cc * xc[q];
              r *= ff.x;
            } else {
              const float acc = (ca * xa[q] + cb * xb[q]) + cc * xc[q];
              r = xs[q] + acc * ff.x;
            }
            lv[q * n + li] = r;
          }
        }
      }
      __syncthreads();
      pos = npos;
      end = nend;
      pk = npk;
      f = nf;
    }
  }
}

// C = alpha op(A) op(B) + beta C (m x nn, inner dimension k), row-major LDS
// matrices; C must not share storage with A or B.
__device__ inline void wg_mm(const double* A, int lda, bool ta, const double* B, int ldb, bool tb,
                             double* C, int ldc, int m, int nn, int k, double alpha, double beta) {
  for (int e = threadIdx.x; e < m * nn; e += blockDim.x) {
    const int i = e / nn, j = e % nn;
    double s = 0.0;
    for (int q = 0; q < k; ++q)
      s += (ta ? A[q * lda + i] : A[i * lda + q]) * (tb ? B[j * ldb + q] : B[q * ldb + j]);
    C[i * ldc + j] = alpha * s + (beta != 0.0 ? beta * C[i * ldc + j] : 0.0);
  }
  __syncthreads();
}

// A = (A + A^T) / 2 for the leading m x m block.
__device__ inline void wg_symmetrize(double* A, int m, int ld) {
  for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
    const int i = e / m, j = e % m;
    if (j > i) A[i * ld + j] = A[j * ld + i] = 0.5 * (A[i * ld + j] + A[j * ld + i]);
  }
  __syncthreads();
}

// One subdomain per workgroup of BS threads, the preconditioner sweeps on
// chunks of CH columns.  WPE > 1 holds the registers to what WPE waves per
// SIMD leave each wave (WPE = 1: no limit beyond the launch bounds; at 512
// threads the bounds alone give two waves per SIMD, one workgroup per CU).
template <int MB, int BS, int CH, int WPE>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(WPE)))
k_lobpcg_mf(Geo g, LobArgs a) {
  static_assert(3 * MB <= kKP, "LOBPCG block too wide for the update tiles");
  constexpr int LD = 3 * MB;
  constexpr int NW = BS / 64;
  extern __shared__ double lds[];
  const int sidx = a.order ? a.order[blockIdx.x] : blockIdx.x * a.stride;
  const Brick b = g.bricks[sidx];
  const BrickView v = load_view(g, sidx);
  const int n = v.n;
  const bool vec = ((n | v.off) & 1) == 0;
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

  // LDS: small matrices, the update coefficients and the Gram reduction
  // slots; the FP32 sweep buffer of phase R reuses the same space.
  double* GA = lds;
  double* GB = GA + LD * LD;
  double* Z = GB + LD * LD;
  double* Q = Z + LD * LD;
  double* TM = Q + LD * LD;
  double* CF = TM + LD * LD;   // kKP x kKP (also an LD x LD scratch)
  double* w = CF + kKP * kKP;
  double* rot = w + LD;
  double* red = rot + 3 * LD;  // (NW / 2) * 1024
  float* pv = reinterpret_cast<float*>(lds);
  __shared__ double lam[MB], nrm[MB], msum_s, s_rbest;
  __shared__ double s_red[NW * MB], s_dc[CH];
  __shared__ int s_pl[kMfPlanes + 1];
  __shared__ int s_conv, s_flag, s_lock[MB], s_itbest, s_stag, s_well;

  {
    double s = 0.0, wsum = 0.0;
    for (int r = threadIdx.x; r < n; r += BS) {
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
  // hyperplane offsets of the sweeps, in LDS when they fit
  const int* pl = ps;
  if (b.nplanes <= kMfPlanes) {
    for (int i = threadIdx.x; i <= b.nplanes; i += BS) s_pl[i] = ps[i];
    pl = s_pl;
  }
  __syncthreads();
  const double msum = msum_s;
  // A subdomain with a well term has no null vector: all en_lv1 pairs are
  // iterated and the constant is not deflated.
  const bool well = s_well != 0;
  const int kw = well ? a.kwant + 1 : a.kwant;
  int it = 0, status = 0, np = 0;
  // phase timers of the traced subdomain (thread 0), kept in LDS
  __shared__ unsigned long long tph[16];
  __shared__ int tsw;   // Rayleigh-Ritz Jacobi sweeps of the traced subdomain
  const bool trace = sidx == a.debug && threadIdx.x == 0;
  if (trace) {
    for (int i = 0; i < 16; ++i) tph[i] = 0;
    tsw = 0;
  }
#define PMX_PHASE(i)                                                   \
  if (trace) {                                                         \
    const unsigned long long tn = __builtin_amdgcn_s_memrealtime();    \
    if (tph[15]) tph[i] += tn - tph[15];                               \
    tph[15] = tn;                                                      \
  }
  // CF(0:r, 0:c) = M(0:r, 0:c) for an LD-strided small matrix M.
  auto load_cf = [&](const double* M, int r, int c) {
    for (int e = threadIdx.x; e < r * c; e += BS) CF[(e / c) * kKP + e % c] = M[(e / c) * LD + e % c];
    __syncthreads();
  };
  int attempts = 0;
  if (kw > 0) {
    for (int attempt = 0; attempt < 2; ++attempt) {
      attempts = attempt;
      // ---- initial block: warm start rows or seeded random ----
      for (int e = threadIdx.x; e < MB * n; e += BS) {
        const int j = e / n, r = e % n;
        X[e] = (a.x0 && attempt == 0)
                   ? a.x0[(size_t)(v.off + r) * MB + j]
                   : hash_uniform(a.seed ^ ((uint64_t)b.lid << 40) ^ ((uint64_t)j << 32) ^
                                  (uint64_t)r ^ ((uint64_t)attempt << 60));
      }
      __syncthreads();
      if (!well) {
        double acc[MB];
#pragma unroll
        for (int j = 0; j < MB; ++j) acc[j] = 0.0;
        for (int r = threadIdx.x; r < n; r += BS) {
          const double m = mass[r];
#pragma unroll
          for (int j = 0; j < MB; ++j) acc[j] += m * X[(size_t)j * n + r];
        }
        wg_sum<MB>(acc, MB, s_red, TM);
        for (int e = threadIdx.x; e < MB * n; e += BS) X[e] -= TM[e / n] / msum;
        __syncthreads();
      }
      const ColSet sx = colset(X, MB), sax = colset(AX, MB);
      wg_gram<1, 1>(sx, MB, sx, MB, mass, n, vec, red, GB, LD);
      const int keep = wg_svqb(GB, GA, Z, Q, w, rot, MB, LD, 1e-14);
      if (keep < MB) {
        status = 2;
        continue;
      }
      load_cf(Z, MB, MB);
      wg_update<1>(sx, MB, CF, outset(X, MB), MB, n);
      wg_apply_AN<MB>(v, a.Tx, a.Ty, a.Tz, g.wd, X, AX, MB);
      wg_gram<1, 1>(sx, MB, sax, MB, nullptr, n, vec, red, GA, LD);
      wg_symmetrize(GA, MB, LD);
      wg_syev2(GA, GB, Q, Z, w, rot, MB, LD);
      if (threadIdx.x < MB) lam[threadIdx.x] = w[threadIdx.x];
      __syncthreads();
      load_cf(Q, MB, MB);
      wg_update<1>(sx, MB, CF, outset(X, MB), MB, n);
      wg_update<1>(sax, MB, CF, outset(AX, MB), MB, n);
      np = 0;
      status = 1;
      if (threadIdx.x == 0) {
        s_rbest = 1e300;
        s_itbest = 0;
        s_stag = 0;
      }
      __syncthreads();
      for (it = 0; it < a.maxit; ++it) {
        PMX_PHASE(14)
        // Near convergence the drift of the updated AX and AP from A X and
        // A P sets the residual's floor and can turn the iteration unstable.
        if (it > 0 && it % kRefreshIts == 0 && s_rbest <= kRefreshNear * a.tol) {
          wg_apply_AN<MB>(v, a.Tx, a.Ty, a.Tz, g.wd, X, AX, MB);
          if (np > 0) wg_apply_AN<MB>(v, a.Tx, a.Ty, a.Tz, g.wd, P, AP, np);
        }
        // ---- R: W = T (AX - M X Lambda), deflated; AW = A W; ||W||_M ----
        for (int j0 = 0; j0 < MB; j0 += CH) {
          const int ch = min(CH, MB - j0);
          for (int e = threadIdx.x; e < ch * n; e += BS) {
            const int q = e / n, r = e - q * n;
            const size_t o = (size_t)(j0 + q) * n + r;
            pv[e] = (float)(AX[o] - lam[j0 + q] * mass[r] * X[o]);
          }
          __syncthreads();
          PMX_PHASE(0)
          mf_sweeps<CH>(v, b.nplanes, pl, pc, a.pf, a.pb, pv, ch);
          PMX_PHASE(1)
          double acc[CH];
#pragma unroll
          for (int q = 0; q < CH; ++q) acc[q] = 0.0;
          if (!well) {
            for (int r = threadIdx.x; r < n; r += BS) {
              const double m = mass[r];
#pragma unroll
              for (int q = 0; q < CH; ++q)
                if (q < ch) acc[q] += m * (double)pv[q * n + r];
            }
          }
          wg_sum<CH>(acc, ch, s_red, s_dc);
          if ((int)threadIdx.x < ch) s_dc[threadIdx.x] = well ? 0.0 : s_dc[threadIdx.x] / msum;
          __syncthreads();
          PMX_PHASE(2)
#pragma unroll
          for (int q = 0; q < CH; ++q) acc[q] = 0.0;
          for_cells(v, [&](int c, int lx, int ly, int lz) {
            const int li = c - v.off;
            const double tl0 = lx > 0 ? a.Tx[c] : 0.0, th0 = lx < v.nx - 1 ? a.Tx[c + 1] : 0.0;
            const double tl1 = ly > 0 ? a.Ty[c] : 0.0, th1 = ly < v.ny - 1 ? a.Ty[c + v.nx] : 0.0;
            const double tl2 = lz > 0 ? a.Tz[c] : 0.0,
                         th2 = lz < v.nz - 1 ? a.Tz[c + v.nxny] : 0.0;
            const int o0 = lx > 0 ? li - 1 : li, o1 = lx < v.nx - 1 ? li + 1 : li;
            const int o2 = ly > 0 ? li - v.nx : li, o3 = ly < v.ny - 1 ? li + v.nx : li;
            const int o4 = lz > 0 ? li - v.nxny : li, o5 = lz < v.nz - 1 ? li + v.nxny : li;
            const double wc = g.wd ? g.wd[c] : 0.0;
            const double m = mass[li];
#pragma unroll
            for (int q = 0; q < CH; ++q) {
              if (q >= ch) break;
              const float* vv = pv + q * n;
              const double dc = s_dc[q];
              const double xc = (double)vv[li] - dc;
              const double y = tl0 * (xc - ((double)vv[o0] - dc)) +
                               th0 * (xc - ((double)vv[o1] - dc)) +
                               tl1 * (xc - ((double)vv[o2] - dc)) +
                               th1 * (xc - ((double)vv[o3] - dc)) +
                               tl2 * (xc - ((double)vv[o4] - dc)) +
                               th2 * (xc - ((double)vv[o5] - dc)) + wc * xc;
              W[(size_t)(j0 + q) * n + li] = xc;
              AW[(size_t)(j0 + q) * n + li] = y;
              acc[q] += m * xc * xc;
            }
          });
          wg_sum<CH>(acc, ch, s_red, nrm + j0);
          PMX_PHASE(3)
        }
        // ---- convergence: ||T r||_M <= tol for the wanted pairs ----
        if (threadIdx.x == 0) {
          int conv = 1;
          double rmax = 0.0;
          for (int j = 0; j < kw; ++j) {
            const double rj = sqrt(nrm[j]);
            if (!(rj <= a.tol)) conv = 0;
            rmax = fmax(rmax, rj);
          }
          // Stagnation: accept once the worst residual is within
          // stag_factor * tol and has not halved for kStagIts iterations.
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
        if (s_conv) {
          status = 0;
          break;
        }
        // soft locking: converged wanted pairs add no search direction
        if (threadIdx.x < MB)
          s_lock[threadIdx.x] = (int)threadIdx.x < kw && sqrt(nrm[threadIdx.x]) <= a.tol;
        __syncthreads();
        PMX_PHASE(4)
        // ---- O: W -= X (X^T M W) + P (P^T M W) twice; SVQB(W) ----
        const int nxp = MB + np, nk = nxp + MB;
        const ColSet sxpw = colset(X, MB, P, np, W, MB), axpw = colset(AX, MB, AP, np, AW, MB);
        int nw = 0;
        if (a.ortho_coef) {
          // Both projections and the SVQB from one M-Gram G of B = [X P W]
          // (one pass instead of two Grams and two block updates):
          // W' = B K1, K1 = [-G_xw; I] (locked columns zero); the second
          // projection a' = [X P]^T M W' = (G K1)_x; W'^T M W' = K1^T G K1;
          // the Gram after it K1^T G K1 - a'^T a'; W'' = B (K1 - [a'; 0]) Z.
          if (nk > 16)
            wg_gram_ma<2, false>(sxpw, sxpw, nk, mass, n, vec, red, GB, nullptr, LD);
          else
            wg_gram_ma<1, false>(sxpw, sxpw, nk, mass, n, vec, red, GB, nullptr, LD);
          for (int e = threadIdx.x; e < nk * MB; e += BS) {   // K1 -> CF
            const int k = e / MB, c = e % MB;
            double val;
            if (s_lock[c])
              val = 0.0;
            else if (k < nxp)
              val = -GB[k * LD + nxp + c];
            else
              val = (k - nxp == c) ? 1.0 : 0.0;
            CF[k * kKP + c] = val;
          }
          __syncthreads();
          for (int e = threadIdx.x; e < nk * MB; e += BS) {   // T1 = G K1 -> TM
            const int k = e / MB, c = e % MB;
            double v = 0.0;
            for (int l = 0; l < nk; ++l) v += GB[k * LD + l] * CF[l * kKP + c];
            TM[k * LD + c] = v;
          }
          __syncthreads();
          for (int e = threadIdx.x; e < MB * MB; e += BS) {   // K1^T T1 - a'^T a'
            const int i = e / MB, j = e % MB;
            double v = 0.0;
            for (int l = 0; l < nk; ++l) v += CF[l * kKP + i] * TM[l * LD + j];
            for (int k = 0; k < nxp; ++k) v -= TM[k * LD + i] * TM[k * LD + j];
            GA[i * LD + j] = v;
          }
          __syncthreads();
          wg_symmetrize(GA, MB, LD);
          for (int e = threadIdx.x; e < nxp * MB; e += BS) {  // K = K1 - [a'; 0]
            const int k = e / MB, c = e % MB;
            CF[k * kKP + c] -= TM[k * LD + c];
          }
          __syncthreads();
          PMX_PHASE(7)
          nw = wg_svqb(GA, Q, Z, TM, w, rot, MB, LD, 1e-12);
          if (nw == 0) {
            status = 3;
            break;
          }
          for (int e = threadIdx.x; e < nk * nw; e += BS) {  // K Z -> GB (G is done)
            const int k = e / nw, c = e % nw;
            double v = 0.0;
            for (int j = 0; j < MB; ++j) v += CF[k * kKP + j] * Z[j * LD + c];
            GB[k * LD + c] = v;
          }
          __syncthreads();
          for (int e = threadIdx.x; e < nk * nw; e += BS)
            CF[(e / nw) * kKP + e % nw] = GB[(e / nw) * LD + e % nw];
          __syncthreads();
          PMX_PHASE(8)
        } else {
        if (nxp > 16)
          wg_gram<2, 1>(colset(X, MB, P, np), nxp, colset(W, MB), MB, mass, n, vec, red, GB, LD);
        else
          wg_gram<1, 1>(colset(X, MB, P, np), nxp, colset(W, MB), MB, mass, n, vec, red, GB, LD);
        for (int e = threadIdx.x; e < nk * MB; e += BS) {
          const int k = e / MB, c = e % MB;
          double val;
          if (s_lock[c])
            val = 0.0;
          else if (k < nxp)
            val = -GB[k * LD + c];
          else
            val = (k - nxp == c) ? 1.0 : 0.0;
          CF[k * kKP + c] = val;
        }
        __syncthreads();
        PMX_PHASE(5)
        wg_update<1>(sxpw, nk, CF, outset(W, MB), MB, n);
        wg_update<1>(axpw, nk, CF, outset(AW, MB), MB, n);
        PMX_PHASE(6)
        if (nk > 16)
          wg_gram<2, 1>(sxpw, nk, colset(W, MB), MB, mass, n, vec, red, GB, LD);
        else
          wg_gram<1, 1>(sxpw, nk, colset(W, MB), MB, mass, n, vec, red, GB, LD);
        // M-Gram of W after the second projection: W'^T M W' - a'^T a' - b'^T b'
        for (int e = threadIdx.x; e < MB * MB; e += BS) {
          const int i = e / MB, j = e % MB;
          double s = 0.5 * (GB[(nxp + i) * LD + j] + GB[(nxp + j) * LD + i]);
          for (int k = 0; k < nxp; ++k) s -= GB[k * LD + i] * GB[k * LD + j];
          GA[i * LD + j] = s;
        }
        __syncthreads();
        PMX_PHASE(7)
        nw = wg_svqb(GA, Q, Z, TM, w, rot, MB, LD, 1e-12);
        if (nw == 0) {
          status = 3;
          break;
        }
        // CF = [-(a', b') Z; Z]: the second projection and the SVQB in one pass
        for (int e = threadIdx.x; e < nk * nw; e += BS) {
          const int k = e / nw, c = e % nw;
          double val;
          if (k < nxp) {
            val = 0.0;
            for (int j = 0; j < MB; ++j) val -= GB[k * LD + j] * Z[j * LD + c];
          } else {
            val = Z[(k - nxp) * LD + c];
          }
          CF[k * kKP + c] = val;
        }
        __syncthreads();
        PMX_PHASE(8)
        }
        wg_update<1>(sxpw, nk, CF, outset(W, nw), nw, n);
        wg_update<1>(axpw, nk, CF, outset(AW, nw), nw, n);
        PMX_PHASE(9)
        // ---- G: S^T M S and S^T A S, S = [X W P] ----
        const int ns = MB + nw + np;
        const ColSet S = colset(X, MB, W, nw, P, np), AS = colset(AX, MB, AW, nw, AP, np);
        if (ns > 16)
          wg_gram_ma<2>(S, AS, ns, mass, n, vec, red, GB, GA, LD);
        else
          wg_gram_ma<1>(S, AS, ns, mass, n, vec, red, GB, GA, LD);
        wg_symmetrize(GA, ns, LD);
        wg_symmetrize(GB, ns, LD);
        PMX_PHASE(10)
        // ---- Rayleigh-Ritz: G_B = L L^T, L^{-1} G_A L^{-T} = Q w Q^T,
        // C_x = L^{-T} Q(:, 0:MB) ----
        for (int e = threadIdx.x; e < ns * ns; e += BS)
          CF[(e / ns) * LD + e % ns] = GB[(e / ns) * LD + e % ns];
        __syncthreads();
        if (!wg_cholesky(CF, ns, LD, &s_flag)) {
          np = 0;
          status = 4;
          continue;
        }
        wg_trinv_lower(CF, TM, ns, LD);      // CF = L^{-1}
        wg_matmul(CF, GA, TM, ns, LD, false);  // TM = L^{-1} G_A
        for (int e = threadIdx.x; e < ns * ns; e += BS) {
          const int i = e / ns, j = e % ns;
          double s = 0.0;
          for (int k = 0; k <= j; ++k) s += TM[i * LD + k] * CF[j * LD + k];
          Z[i * LD + j] = s;
        }
        __syncthreads();
        wg_symmetrize(Z, ns, LD);
        PMX_PHASE(11)
        {
          const int sw = wg_syev2(Z, TM, Q, GA, w, rot, ns, LD);  // eigenvectors in Q
          if (trace) tsw += sw;
        }
        PMX_PHASE(12)
        for (int e = threadIdx.x; e < ns * MB; e += BS) {
          const int i = e / MB, c = e % MB;
          double s = 0.0;
          for (int k = i; k < ns; ++k) s += CF[k * LD + i] * Q[k * LD + c];
          TM[i * LD + c] = s;
        }
        if (threadIdx.x < MB) lam[threadIdx.x] = w[threadIdx.x];
        __syncthreads();
        // P_new = [W P] C_p (C_p = C_x without its X rows), made M-orthogonal
        // to X_new = S C_x and M-orthonormal on the coefficients:
        // C_p -= C_x (C_x^T G_B C_p), then SVQB of C_p^T G_B C_p.
        for (int e = threadIdx.x; e < ns * MB; e += BS) {
          const int i = e / MB, c = e % MB;
          Z[i * LD + c] = i < MB ? 0.0 : TM[i * LD + c];
        }
        __syncthreads();
        wg_mm(GB, LD, false, Z, LD, false, Q, LD, ns, MB, ns, 1.0, 0.0);
        wg_mm(TM, LD, true, Q, LD, false, GA, LD, MB, MB, ns, 1.0, 0.0);
        wg_mm(TM, LD, false, GA, LD, false, Z, LD, ns, MB, MB, -1.0, 1.0);
        wg_mm(GB, LD, false, Z, LD, false, Q, LD, ns, MB, ns, 1.0, 0.0);
        wg_mm(Z, LD, true, Q, LD, false, GA, LD, MB, MB, ns, 1.0, 0.0);
        wg_symmetrize(GA, MB, LD);
        const int npn = wg_svqb(GA, Q, GB, CF, w, rot, MB, LD, 1e-12);  // Z_p in GB
        const int nout = MB + npn;
        for (int e = threadIdx.x; e < ns * nout; e += BS) {
          const int k = e / nout, c = e % nout;
          double val;
          if (c < MB) {
            val = TM[k * LD + c];
          } else {
            val = 0.0;
            for (int j = 0; j < MB; ++j) val += Z[k * LD + j] * GB[j * LD + c - MB];
          }
          CF[k * kKP + c] = val;
        }
        __syncthreads();
        PMX_PHASE(13)
        // ---- U: [X P] = S [C_x C_p], [AX AP] = A S [C_x C_p] ----
        if (nout > 16) {
          wg_update<2>(S, ns, CF, outset(X, MB, P, npn), nout, n);
          wg_update<2>(AS, ns, CF, outset(AX, MB, AP, npn), nout, n);
        } else {
          wg_update<1>(S, ns, CF, outset(X, MB, P, npn), nout, n);
          wg_update<1>(AS, ns, CF, outset(AX, MB, AP, npn), nout, n);
        }
        np = npn;
      }
      if (status == 0) break;
      if (attempt == 0 && a.x0 == nullptr && status != 2) break;
    }
  }
  if (trace) {
    printf("lobpcg-mf phases (us) its %d: resid %.0f sweeps %.0f deflate %.0f W,AW %.0f "
           "conv %.0f | ortho gram1 %.0f upd1 %.0f gram2 %.0f svqb %.0f upd2 %.0f | gram %.0f | "
           "rr chol %.0f syev %.0f (%d sweeps) P %.0f | update %.0f\n",
           it, tph[0] * 0.01, tph[1] * 0.01, tph[2] * 0.01, tph[3] * 0.01, tph[4] * 0.01,
           tph[5] * 0.01, tph[6] * 0.01, tph[7] * 0.01, tph[8] * 0.01, tph[9] * 0.01,
           tph[10] * 0.01, tph[11] * 0.01, tph[12] * 0.01, tsw, tph[13] * 0.01, tph[14] * 0.01);
  }
#undef PMX_PHASE
  // ---- outputs ----
  const double phi0 = 1.0 / sqrt(msum);
  double* phi = a.phi + (size_t)v.off * a.kmax;
  const int c0 = well ? 0 : 1;  // first basis column LOBPCG fills
  if (!well)
    for (int r = threadIdx.x; r < n; r += BS) phi[(size_t)r * a.kmax] = phi0;
  if (kw > 0 && status == 0) {
    // sign: positive inner product with 1 + frac(0.618..(r+1))
    double acc[MB];
#pragma unroll
    for (int j = 0; j < MB; ++j) acc[j] = 0.0;
    for (int r = threadIdx.x; r < n; r += BS) {
      const double wgt = 1.0 + fmod(0.6180339887498949 * (r + 1.0), 1.0);
#pragma unroll
      for (int j = 0; j < MB; ++j) acc[j] += wgt * X[(size_t)j * n + r];
    }
    wg_sum<MB>(acc, MB, s_red, nrm);
    if (threadIdx.x < MB) nrm[threadIdx.x] = nrm[threadIdx.x] < 0.0 ? -1.0 : 1.0;
    __syncthreads();
    for (int e = threadIdx.x; e < kw * n; e += BS) {
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

// Shape 1: every workgroup held to 256 registers (two waves per SIMD), and
// large subdomains on 256 threads with sweep chunks of at most kMfPairCh
// columns (64 KB of LDS at 4096 cells), so two subdomains share a CU and one's
// latency-bound phases (sweeps, the small Rayleigh-Ritz algebra) overlap the
// other's block passes.
constexpr int kMfPairCh = 4;

template <int MB>
static size_t lobpcg_mf_lds(int maxn, int bs, int ch) {
  constexpr int LD = 3 * MB;
  const size_t small =
      (5 * LD * LD + kKP * kKP + 4 * LD + (size_t)(bs / 128) * 1024) * sizeof(double);
  const size_t prec = (size_t)maxn * ch * sizeof(float);
  return std::max(small, prec);
}

template <int MB, int BS, int CH = lobpcg_mf_ch(MB), int WPE = 1>
static void launch_mf_bs(const Geo& g, int nblocks, hipStream_t st, const LobArgs& a, int maxn) {
  const size_t sh = lobpcg_mf_lds<MB>(maxn, BS, CH);
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_lobpcg_mf<MB, BS, CH, WPE>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
  hipLaunchKernelGGL((k_lobpcg_mf<MB, BS, CH, WPE>), dim3(nblocks), dim3(BS), sh, st, g, a);
}

template <int MB>
static void launch_mf(const Geo& g, int nblocks, hipStream_t st, const LobArgs& a, int maxn,
                      int shape) {
  if (maxn >= kLobBigN) {
    if (shape == 1)
      launch_mf_bs<MB, 256, (MB < kMfPairCh ? MB : kMfPairCh), 2>(g, nblocks, st, a, maxn);
    else
      launch_mf_bs<MB, 512>(g, nblocks, st, a, maxn);
  } else if (maxn <= kLobSmallN) {
    if (shape == 1)
      launch_mf_bs<MB, 128, lobpcg_mf_ch(MB), 2>(g, nblocks, st, a, maxn);
    else
      launch_mf_bs<MB, 128>(g, nblocks, st, a, maxn);
  } else {
    if (shape == 1)
      launch_mf_bs<MB, 256, lobpcg_mf_ch(MB), 2>(g, nblocks, st, a, maxn);
    else
      launch_mf_bs<MB, 256>(g, nblocks, st, a, maxn);
  }
}

void launch_lobpcg_mf(int MB, int shape, const Geo& g, int nblocks, hipStream_t st,
                      const LobArgs& a, int maxn) {
  switch (MB) {
    case 4:
      launch_mf<4>(g, nblocks, st, a, maxn, shape);
      break;
    case 6:
      launch_mf<6>(g, nblocks, st, a, maxn, shape);
      break;
    case 8:
      launch_mf<8>(g, nblocks, st, a, maxn, shape);
      break;
    default:
      launch_mf<10>(g, nblocks, st, a, maxn, shape);
      break;
  }
}

}  // namespace pmx
