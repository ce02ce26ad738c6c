// Learned warm start for the level-2 eigensolver: a 3-layer 3D CNN maps a
// subdomain's log-permeability patch (plus the three cell coordinates) to a
// block of approximate eigenvectors, which LOBPCG then refines to the usual
// eigen-residual tolerance.
//
// Network: conv 3x3x3 (4 -> 16) + ReLU, conv 3x3x3 (16 -> 16) + ReLU,
// conv 3x3x3 (16 -> cout), zero padding at the patch faces.  Weights are
// kept in FP32 (training master copy) and used in BF16; activations are BF16
// [cell][16 channels] (cell = brick order); every product accumulates in
// FP32 on the matrix cores.
//
// Convolution as GEMM on MFMA (v_mfma_f32_16x16x16_bf16): for a row of 16
// cells along x and one of the 27 taps, D[co][cell] += W[co][tap][ci] *
// in[cell + tap][ci] is one 16x16x16 product (M = output channels, N =
// cells, K = the 16 input channels of the tap); a tile sums 27 of them.  A
// workgroup owns one brick and walks it in z-slabs staged in LDS with a
// one-cell halo.  The backward pass reuses the same kernel for the data
// gradient (weights transposed and flipped, ReLU mask in the epilogue) and
// a second MFMA kernel for the weight gradient (K = cells).
#include <cmath>

#include "pmx_kernels.h"

namespace pmx {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kCnnCh = 16;    // channel width of every activation
constexpr int kCnnTaps = 27;  // 3x3x3
constexpr int kCnnSlab = 2;   // output planes per LDS slab
constexpr int kCnnLayerW = kCnnCh * kCnnTaps * kCnnCh;  // weights of one layer
constexpr int kCnnLayer = kCnnLayerW + kCnnCh;          // + bias

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) {
  return __uint_as_float((unsigned int)h << 16);
}

// LDS bytes of one slab tile for bricks with at most nx, ny cells.
static size_t cnn_tile_bytes(int nx, int ny) {
  const int nxp = (nx + 15) / 16 * 16;
  return (size_t)(kCnnSlab + 2) * (ny + 2) * (nxp + 2) * kCnnCh * sizeof(unsigned short);
}

// Stages planes z0-1 .. z0+kCnnSlab of the brick (16 channels, zero outside
// the brick) into the tile [(kCnnSlab+2)][ny+2][nxp+2][16].
__device__ __forceinline__ void cnn_stage(const BrickView& v, const unsigned short* __restrict__ in,
                                          unsigned short* tile, int z0, int TX, int TY) {
  const int nvox = (kCnnSlab + 2) * TY * TX;
  for (int e = threadIdx.x; e < 2 * nvox; e += blockDim.x) {
    const int h = e & 1, q = e >> 1;
    const int tx = q % TX, ty = (q / TX) % TY, tz = q / (TX * TY);
    const int gx = tx - 1, gy = ty - 1, gz = z0 + tz - 1;
    uint4 val = make_uint4(0u, 0u, 0u, 0u);
    if (gx >= 0 && gx < v.nx && gy >= 0 && gy < v.ny && gz >= 0 && gz < v.nz)
      val = *reinterpret_cast<const uint4*>(
          in + (size_t)(v.off + (gz * v.ny + gy) * v.nx + gx) * kCnnCh + 8 * h);
    *reinterpret_cast<uint4*>(tile + (size_t)q * kCnnCh + 8 * h) = val;
  }
}

// One convolution layer over the bricks (one workgroup per brick, 4 waves).
// W: [16 co][27 tap][16 ci] BF16.  Epilogue per output value:
//   + bias[co] (if bias), ReLU (if relu), * (mask[cell][co] > 0) (if mask);
// stored as BF16 to out and/or FP32 to out32 ([cell][16]).
__global__ void __launch_bounds__(256) k_cnn_conv(const BrickView* __restrict__ views,
                                                  const unsigned short* __restrict__ in,
                                                  const unsigned short* __restrict__ W,
                                                  const float* __restrict__ bias, int relu,
                                                  const unsigned short* __restrict__ mask,
                                                  unsigned short* __restrict__ out,
                                                  float* __restrict__ out32) {
  extern __shared__ unsigned short ctile[];
  const BrickView v = views[blockIdx.x];
  const int xt = (v.nx + 15) / 16;
  const int TX = 16 * xt + 2, TY = v.ny + 2;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const int n = lane & 15, kq = lane >> 4;
  // A operands (weights) of all taps: lane holds W[co = n][tap][4kq .. 4kq+3].
  bf16x4 a[kCnnTaps];
#pragma unroll
  for (int t = 0; t < kCnnTaps; ++t)
    a[t] = *reinterpret_cast<const bf16x4*>(W + (n * kCnnTaps + t) * kCnnCh + 4 * kq);
  float bco[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) bco[i] = bias ? bias[4 * kq + i] : 0.0f;
  for (int z0 = 0; z0 < v.nz; z0 += kCnnSlab) {
    __syncthreads();
    cnn_stage(v, in, ctile, z0, TX, TY);
    __syncthreads();
    const int nzh = min(kCnnSlab, v.nz - z0);
    const int ntiles = nzh * v.ny * xt;
    for (int tl = wv; tl < ntiles; tl += nwv) {
      const int xi = tl % xt, y = (tl / xt) % v.ny, zl = tl / (xt * v.ny);
      const int x = 16 * xi + n;
      f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int t = 0; t < kCnnTaps; ++t) {
        const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
        // B operand: lane holds in[cell (x, y, z) + tap][4kq .. 4kq+3]
        const bf16x4 b = *reinterpret_cast<const bf16x4*>(
            ctile + ((size_t)((zl + 1 + dz) * TY + (y + 1 + dy)) * TX + (x + 1 + dx)) * kCnnCh +
            4 * kq);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[t], b, acc, 0, 0, 0);
      }
      // D: lane holds out[cell x][co = 4kq .. 4kq+3]
      if (x < v.nx) {
        const size_t cell = (size_t)v.off + ((z0 + zl) * v.ny + y) * v.nx + x;
        float r[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          float val = acc[i] + bco[i];
          if (relu) val = fmaxf(val, 0.0f);
          if (mask && !(bf2f(mask[cell * kCnnCh + 4 * kq + i]) > 0.0f)) val = 0.0f;
          r[i] = val;
        }
        if (out) {
          ushort4 o;
          o.x = f2bf(r[0]);
          o.y = f2bf(r[1]);
          o.z = f2bf(r[2]);
          o.w = f2bf(r[3]);
          *reinterpret_cast<ushort4*>(out + cell * kCnnCh + 4 * kq) = o;
        }
        if (out32)
          *reinterpret_cast<float4*>(out32 + cell * kCnnCh + 4 * kq) =
              make_float4(r[0], r[1], r[2], r[3]);
      }
    }
  }
}

// Weight and bias gradients of one layer, per brick:
//   dW[co][tap][ci] = sum_cells g[cell][co] in[cell + tap][ci],
//   db[co] = sum_cells g[cell][co].
// MFMA with M = co, N = ci, K = 16 cells of a row (A = g^T, B = shifted
// input).  part: [brick][kCnnLayer] FP32 (fixed-order sums, reduced by
// k_cnn_reduce).
__global__ void __launch_bounds__(256) k_cnn_wgrad(const BrickView* __restrict__ views,
                                                   const unsigned short* __restrict__ in,
                                                   const unsigned short* __restrict__ g,
                                                   float* __restrict__ part) {
  extern __shared__ unsigned short ctile[];
  __shared__ float red[4 * kCnnCh];
  const BrickView v = views[blockIdx.x];
  const int xt = (v.nx + 15) / 16;
  const int TX = 16 * xt + 2, TY = v.ny + 2;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const int n = lane & 15, kq = lane >> 4;
  f32x4 acc[kCnnTaps];
#pragma unroll
  for (int t = 0; t < kCnnTaps; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  float bsum = 0.0f;  // lane n < 16 of wave: channel n partial
  for (int z0 = 0; z0 < v.nz; z0 += kCnnSlab) {
    __syncthreads();
    cnn_stage(v, in, ctile, z0, TX, TY);
    __syncthreads();
    const int nzh = min(kCnnSlab, v.nz - z0);
    const int ntiles = nzh * v.ny * xt;
    for (int tl = wv; tl < ntiles; tl += nwv) {
      const int xi = tl % xt, y = (tl / xt) % v.ny, zl = tl / (xt * v.ny);
      // A operand: lane holds g[cell 16xi + 4kq + j][co = n], j = 0..3
      bf16x4 ag;
      const int z = z0 + zl;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int x = 16 * xi + 4 * kq + j;
        unsigned short gv = 0;
        if (x < v.nx) gv = g[((size_t)v.off + (z * v.ny + y) * v.nx + x) * kCnnCh + n];
        ag[j] = (short)gv;
        bsum += bf2f(gv);
      }
#pragma unroll
      for (int t = 0; t < kCnnTaps; ++t) {
        const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
        // B operand: lane holds in[cell (16xi + 4kq + j) + tap][ci = n]
        bf16x4 bi;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int x = 16 * xi + 4 * kq + j;
          bi[j] = (short)ctile[((size_t)((zl + 1 + dz) * TY + (y + 1 + dy)) * TX +
                                (x + 1 + dx)) * kCnnCh + n];
        }
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ag, bi, acc[t], 0, 0, 0);
      }
    }
  }
  // bias: lanes n of the four lane groups hold partials of channel n
  bsum += __shfl_xor(bsum, 16, 64);
  bsum += __shfl_xor(bsum, 32, 64);
  if (lane < kCnnCh) red[wv * kCnnCh + lane] = bsum;
  // Waves add their D fragments in a fixed order through LDS (reusing the
  // tile as FP32 scratch: kCnnLayerW floats).
  float* dsum = reinterpret_cast<float*>(ctile);
  for (int w = 0; w < nwv; ++w) {
    __syncthreads();
    if (wv == w) {
#pragma unroll
      for (int t = 0; t < kCnnTaps; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          // D[co = 4kq + i][ci = n]
          const int idx = ((4 * kq + i) * kCnnTaps + t) * kCnnCh + n;
          dsum[idx] = (w == 0 ? 0.0f : dsum[idx]) + acc[t][i];
        }
    }
  }
  __syncthreads();
  float* out = part + (size_t)blockIdx.x * kCnnLayer;
  for (int e = threadIdx.x; e < kCnnLayerW; e += blockDim.x) out[e] = dsum[e];
  if (threadIdx.x < kCnnCh) {
    float s = 0.0f;
    for (int w = 0; w < nwv; ++w) s += red[w * kCnnCh + threadIdx.x];
    out[kCnnLayerW + threadIdx.x] = s;
  }
}

// grad[e] (+)= sum_b part[b][e], fixed order.
__global__ void k_cnn_reduce(const float* __restrict__ part, int nb, float* __restrict__ grad,
                             int accumulate) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kCnnLayer) return;
  float s = 0.0f;
  for (int b = 0; b < nb; ++b) s += part[(size_t)b * kCnnLayer + e];
  grad[e] = accumulate ? grad[e] + s : s;
}

// BF16 copies of the three layers' weights (forward layout) and of their
// transposed, tap-flipped versions for the data gradient:
// wt[l][ci][26 - t][co] = w[l][co][t][ci].
__global__ void k_cnn_weights(const float* __restrict__ params, unsigned short* __restrict__ wf,
                              unsigned short* __restrict__ wt) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 3 * kCnnLayerW) return;
  const int l = e / kCnnLayerW, r = e % kCnnLayerW;
  const int co = r / (kCnnTaps * kCnnCh), t = (r / kCnnCh) % kCnnTaps, ci = r % kCnnCh;
  const unsigned short h = f2bf(params[(size_t)l * kCnnLayer + r]);
  wf[e] = h;
  wt[(size_t)l * kCnnLayerW + (ci * kCnnTaps + (kCnnTaps - 1 - t)) * kCnnCh + co] = h;
}

// MSE loss over the first cout channels and its gradient (BF16):
// L = scale * sum (out - target)^2, g = 2 scale (out - target).
__global__ void __launch_bounds__(256) k_cnn_mse(const float* __restrict__ out,
                                                 const float* __restrict__ target, size_t ncell,
                                                 int cout, float scale,
                                                 unsigned short* __restrict__ g,
                                                 double* __restrict__ loss_part) {
  __shared__ double red[16];
  double acc = 0.0;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < ncell * kCnnCh;
       e += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % kCnnCh);
    float d = 0.0f;
    if (c < cout) d = out[e] - target[e];
    acc += (double)d * d;
    g[e] = f2bf(2.0f * scale * d);
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) loss_part[blockIdx.x] = acc * scale;
}

// Adam on the FP32 parameters (entries whose gradient is always zero, the
// channel padding, stay at zero).
__global__ void k_cnn_adam(float* __restrict__ p, const float* __restrict__ g,
                           float* __restrict__ m, float* __restrict__ s, int n, float lr, float b1,
                           float b2, float eps, float c1, float c2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float ge = g[e];
  const float me = b1 * m[e] + (1.0f - b1) * ge;
  const float se = b2 * s[e] + (1.0f - b2) * ge * ge;
  m[e] = me;
  s[e] = se;
  p[e] -= lr * (me / c1) / (sqrtf(se / c2) + eps);
}

// Input activations of one brick: channel 0 = log permeability (geometric
// mean of the three components) standardised over the brick, channels 1-3 =
// cell coordinates in [-1, 1], channels 4-15 = 0.
__global__ void __launch_bounds__(256) k_cnn_input(const BrickView* __restrict__ views,
                                                   const double* __restrict__ kx,
                                                   const double* __restrict__ ky,
                                                   const double* __restrict__ kz,
                                                   unsigned short* __restrict__ act) {
  __shared__ double red[16];
  const BrickView v = views[blockIdx.x];
  double s1 = 0.0, s2 = 0.0;
  for (int li = threadIdx.x; li < v.n; li += blockDim.x) {
    const int c = v.off + li;
    const double lk = (log(kx[c]) + log(ky[c]) + log(kz[c])) / 3.0;
    s1 += lk;
    s2 += lk * lk;
  }
  s1 = block_sum(s1, red);
  s2 = block_sum(s2, red);
  const double mean = s1 / v.n;
  const double sd = sqrt(fmax(s2 / v.n - mean * mean, 0.0));
  const double inv = 1.0 / fmax(sd, 1e-3);
  for (int li = threadIdx.x; li < v.n; li += blockDim.x) {
    const int c = v.off + li;
    const int lz = li / v.nxny, r = li - lz * v.nxny, ly = r / v.nx, lx = r - ly * v.nx;
    const double lk = (log(kx[c]) + log(ky[c]) + log(kz[c])) / 3.0;
    unsigned short h[kCnnCh];
#pragma unroll
    for (int q = 0; q < kCnnCh; ++q) h[q] = 0;
    h[0] = f2bf((float)((lk - mean) * inv));
    h[1] = f2bf(2.0f * (lx + 0.5f) / v.nx - 1.0f);
    h[2] = f2bf(2.0f * (ly + 0.5f) / v.ny - 1.0f);
    h[3] = f2bf(2.0f * (lz + 0.5f) / v.nz - 1.0f);
    uint4* dst = reinterpret_cast<uint4*>(act + (size_t)c * kCnnCh);
    dst[0] = make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16), 0u, 0u);
    dst[1] = make_uint4(0u, 0u, 0u, 0u);
  }
}

// Training targets of one brick from the level-2 basis rows phi[cell][kmax]:
// eigenvectors 1..cout (the constant one is skipped), orthonormalised over
// the brick in the Euclidean inner product (modified Gram-Schmidt) and
// scaled to unit RMS; other channels 0.
__global__ void __launch_bounds__(256) k_cnn_targets(const BrickView* __restrict__ views,
                                                     const double* __restrict__ phi, int kmax,
                                                     int cout, float* __restrict__ tgt) {
  __shared__ double red[16];
  const BrickView v = views[blockIdx.x];
  for (int li = threadIdx.x; li < v.n; li += blockDim.x) {
    const size_t c = v.off + li;
    for (int q = 0; q < kCnnCh; ++q)
      tgt[c * kCnnCh + q] = q < cout ? (float)phi[c * kmax + 1 + q] : 0.0f;
  }
  __syncthreads();
  for (int j = 0; j < cout; ++j) {
    for (int i = 0; i < j; ++i) {
      double s = 0.0;
      for (int li = threadIdx.x; li < v.n; li += blockDim.x) {
        const size_t c = v.off + li;
        s += (double)tgt[c * kCnnCh + i] * tgt[c * kCnnCh + j];
      }
      s = block_sum(s, red) / v.n;  // columns i < j have unit RMS
      for (int li = threadIdx.x; li < v.n; li += blockDim.x) {
        const size_t c = v.off + li;
        tgt[c * kCnnCh + j] -= (float)(s * tgt[c * kCnnCh + i]);
      }
      __syncthreads();
    }
    double s = 0.0;
    for (int li = threadIdx.x; li < v.n; li += blockDim.x) {
      const double x = tgt[(size_t)(v.off + li) * kCnnCh + j];
      s += x * x;
    }
    s = block_sum(s, red);
    const double scl = s > 0.0 ? 1.0 / sqrt(s / v.n) : 0.0;
    for (int li = threadIdx.x; li < v.n; li += blockDim.x) tgt[(size_t)(v.off + li) * kCnnCh + j] *= (float)scl;
    __syncthreads();
  }
}

// Subspace loss of one brick, invariant to sign and rotation of the
// predicted block Y (first kw channels of out) and of the target block U
// (orthogonal columns of unit RMS; Ut = U / sqrt(n) is orthonormal):
//   L_b = kw - tr(C^T Gd^{-1} C),  G = Y^T Y,  Gd = G + delta I,  C = Y^T Ut,
// the squared distance of span(Ut) from span(Y).  Gradient
//   dL/dY = -2 (Ut C^T Gd^{-1} - Y Gd^{-1} C C^T Gd^{-1}).
// lscale multiplies loss and gradient (1 / (bricks * kw)); delta = 1e-6 n.
__global__ void __launch_bounds__(256) k_cnn_subspace(const BrickView* __restrict__ views,
                                                      const float* __restrict__ out,
                                                      const float* __restrict__ target, int kw,
                                                      float lscale,
                                                      unsigned short* __restrict__ g,
                                                      double* __restrict__ loss_part) {
  __shared__ double red[4][16];
  __shared__ double Gm[kCnnCh * kCnnCh], Cm[kCnnCh * kCnnCh], Gi[kCnnCh * kCnnCh];
  __shared__ double Am[kCnnCh * kCnnCh], Km[kCnnCh * kCnnCh];
  __shared__ double s_loss;
  const BrickView v = views[blockIdx.x];
  const int n = v.n;
  const double un = 1.0 / sqrt((double)n);
  const int np = 2 * kw * kw;  // G then C entries
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  for (int p0 = 0; p0 < np; p0 += 16) {
    double acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.0;
    for (int li = threadIdx.x; li < n; li += blockDim.x) {
      const float* yv = out + (size_t)(v.off + li) * kCnnCh;
      const float* uv = target + (size_t)(v.off + li) * kCnnCh;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int p = p0 + q;
        if (p < np) {
          const int isC = p >= kw * kw, r = isC ? p - kw * kw : p;
          const int i = r / kw, j = r % kw;
          acc[q] += (double)yv[i] * (isC ? uv[j] * un : (double)yv[j]);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      double x = acc[q];
      for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
      if (lane == 0) red[wv][q] = x;
    }
    __syncthreads();
    if (threadIdx.x < 16 && p0 + threadIdx.x < np) {
      double x = 0.0;
      for (int w = 0; w < nwv; ++w) x += red[w][threadIdx.x];
      const int p = p0 + threadIdx.x;
      if (p < kw * kw)
        Gm[p] = x;
      else
        Cm[p - kw * kw] = x;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // Gi = (G + delta I)^{-1} by Gauss-Jordan (SPD, no pivoting needed)
    double M[kCnnCh * 2 * kCnnCh];
    const double delta = 1e-6 * n;
    for (int i = 0; i < kw; ++i)
      for (int j = 0; j < 2 * kw; ++j)
        M[i * 2 * kw + j] = j < kw ? Gm[i * kw + j] + (i == j ? delta : 0.0) : (j - kw == i ? 1.0 : 0.0);
    for (int c = 0; c < kw; ++c) {
      const double pv = M[c * 2 * kw + c];
      for (int j = 0; j < 2 * kw; ++j) M[c * 2 * kw + j] /= pv;
      for (int i = 0; i < kw; ++i) {
        if (i == c) continue;
        const double f = M[i * 2 * kw + c];
        for (int j = 0; j < 2 * kw; ++j) M[i * 2 * kw + j] -= f * M[c * 2 * kw + j];
      }
    }
    for (int i = 0; i < kw; ++i)
      for (int j = 0; j < kw; ++j) Gi[i * kw + j] = M[i * 2 * kw + kw + j];
    // A = C^T Gi (A[j][i] multiplies Ut[:, j] into dY[:, i]); K = Gi C C^T Gi
    double GC[kCnnCh * kCnnCh];
    for (int i = 0; i < kw; ++i)
      for (int j = 0; j < kw; ++j) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < kw; ++k) {
          a += Cm[k * kw + j] * Gi[k * kw + i];  // (C^T Gi)[j][i]
          b += Gi[i * kw + k] * Cm[k * kw + j];  // (Gi C)[i][j]
        }
        Am[j * kw + i] = a;
        GC[i * kw + j] = b;
      }
    double tr = 0.0;
    for (int i = 0; i < kw; ++i)
      for (int j = 0; j < kw; ++j) {
        double k = 0.0;
        for (int q = 0; q < kw; ++q) k += GC[i * kw + q] * GC[j * kw + q];
        Km[i * kw + j] = k;
      }
    for (int j = 0; j < kw; ++j)
      for (int k = 0; k < kw; ++k) tr += Cm[k * kw + j] * GC[k * kw + j];
    s_loss = kw - tr;
  }
  __syncthreads();
  for (int li = threadIdx.x; li < n; li += blockDim.x) {
    const size_t c = v.off + li;
    const float* yv = out + c * kCnnCh;
    const float* uv = target + c * kCnnCh;
    unsigned short h[kCnnCh];
#pragma unroll
    for (int i = 0; i < kCnnCh; ++i) {
      double d = 0.0;
      if (i < kw) {
        for (int j = 0; j < kw; ++j) d += uv[j] * un * Am[j * kw + i] - (double)yv[j] * Km[j * kw + i];
        d *= -2.0 * lscale;
      }
      h[i] = f2bf((float)d);
    }
    uint4* dst = reinterpret_cast<uint4*>(g + c * kCnnCh);
    dst[0] = make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16),
                        h[6] | (h[7] << 16));
    dst[1] = make_uint4(h[8] | (h[9] << 16), h[10] | (h[11] << 16), h[12] | (h[13] << 16),
                        h[14] | (h[15] << 16));
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = s_loss * lscale;
}

// LOBPCG initial block x0[cell][mb]: columns j < cout from the network
// output, the others seeded uniform in [-1, 1].
__global__ void k_cnn_x0(const float* __restrict__ out32, size_t ncell, int cout, int mb,
                         uint64_t seed, double* __restrict__ x0) {
  const size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  for (int j = 0; j < mb; ++j) {
    double val;
    if (j < cout) {
      val = out32[c * kCnnCh + j];
    } else {
      uint64_t a = seed ^ (c * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)j << 56);
      a = (a ^ (a >> 30)) * 0xBF58476D1CE4E5B9ull;
      a = (a ^ (a >> 27)) * 0x94D049BB133111EBull;
      a ^= a >> 31;
      val = (double)(a >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
    x0[c * mb + j] = val;
  }
}

}  // namespace pmx

// ---------------------------------------------------------------------------
// Host side: the network object (training and inference on uniform batches
// of bricks) and the warm start inside pmx_setup.
// ---------------------------------------------------------------------------
using namespace pmx;

struct pmx_cnn {
  int cout = 0;
  DBuf<float> params, grad, m, s;        // [3][kCnnLayer] FP32
  DBuf<unsigned short> wf, wt;           // BF16 forward / transposed weights
  // per-batch scratch (grown on demand)
  int nb_cap = 0, n_cap = 0;
  DBuf<BrickView> views;
  DBuf<unsigned short> a1, a2, g3, g2, g1;
  DBuf<float> out32, part;
  DBuf<double> loss_part;
  int geom_nb = -1, geom_nx = 0, geom_ny = 0, geom_nz = 0;
  int adam_t = 0;
  hipStream_t stream = nullptr;
};

namespace pmx {

static int cnn_check_shape(int nx, int ny, int nz) {
  if (nx < 1 || ny < 1 || nz < 1 || nx > 64 || ny > 64 || nz > 256) {
    set_error("CNN: brick shape %dx%dx%d is outside 1..64 x 1..64 x 1..256", nx, ny, nz);
    return PMX_ERR_ARG_OUTOFRANGE;
  }
  if (cnn_tile_bytes(nx, ny) > 65536) {
    set_error("CNN: a %dx%d slab does not fit in LDS", nx, ny);
    return PMX_ERR_SUP;
  }
  return PMX_SUCCESS;
}

// Uniform batch geometry: nb bricks of nx*ny*nz cells, brick b at b*n.
static int cnn_geometry(pmx_cnn* net, int nb, int nx, int ny, int nz) {
  PMX_TRY(cnn_check_shape(nx, ny, nz));
  const int n = nx * ny * nz;
  if (nb > net->nb_cap || n > net->n_cap) {
    const int nbc = std::max(nb, net->nb_cap), nc = std::max(n, net->n_cap);
    const size_t cells = (size_t)nbc * nc * kCnnCh;
    for (DBuf<unsigned short>* b : {&net->a1, &net->a2, &net->g3, &net->g2, &net->g1})
      PMX_TRY(b->alloc(cells));
    PMX_TRY(net->out32.alloc(cells));
    PMX_TRY(net->part.alloc((size_t)nbc * kCnnLayer));
    PMX_TRY(net->views.alloc(nbc));
    PMX_TRY(net->loss_part.alloc(1024));
    net->nb_cap = nbc;
    net->n_cap = nc;
    net->geom_nb = -1;
  }
  if (net->geom_nb != nb || net->geom_nx != nx || net->geom_ny != ny || net->geom_nz != nz) {
    std::vector<BrickView> h(nb);
    for (int b = 0; b < nb; ++b) {
      BrickView& v = h[b];
      v.off = b * n;
      v.nx = nx;
      v.ny = ny;
      v.nz = nz;
      v.nxny = nx * ny;
      v.n = n;
      for (int d = 0; d < 6; ++d) {
        v.noff[d] = -1;
        v.next[d] = 0;
      }
    }
    PMX_HIP(hipMemcpyAsync(net->views.p, h.data(), nb * sizeof(BrickView), hipMemcpyHostToDevice,
                           net->stream));
    PMX_HIP(hipStreamSynchronize(net->stream));
    net->geom_nb = nb;
    net->geom_nx = nx;
    net->geom_ny = ny;
    net->geom_nz = nz;
  }
  return PMX_SUCCESS;
}

static int cnn_convert(pmx_cnn* net) {
  hipLaunchKernelGGL(k_cnn_weights, dim3((3 * kCnnLayerW + 255) / 256), dim3(256), 0, net->stream,
                     net->params.p, net->wf.p, net->wt.p);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

static int cnn_conv(hipStream_t st, const BrickView* views, int nb, size_t lds,
                    const unsigned short* in, const unsigned short* W, const float* bias, int relu,
                    const unsigned short* mask, unsigned short* out, float* out32) {
  hipLaunchKernelGGL(k_cnn_conv, dim3(nb), dim3(256), lds, st, views, in, W, bias, relu, mask, out,
                     out32);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

// Forward pass: a1, a2 (BF16) and out32 (FP32) from the input activations.
static int cnn_forward(pmx_cnn* net, hipStream_t st, const BrickView* views, int nb, size_t lds,
                       const unsigned short* in, unsigned short* a1, unsigned short* a2,
                       float* out32) {
  const float* P = net->params.p;
  PMX_TRY(cnn_conv(st, views, nb, lds, in, net->wf.p, P + kCnnLayerW, 1, nullptr, a1, nullptr));
  PMX_TRY(cnn_conv(st, views, nb, lds, a1, net->wf.p + kCnnLayerW, P + kCnnLayer + kCnnLayerW, 1,
                   nullptr, a2, nullptr));
  PMX_TRY(cnn_conv(st, views, nb, lds, a2, net->wf.p + 2 * kCnnLayerW,
                   P + 2 * kCnnLayer + kCnnLayerW, 0, nullptr, nullptr, out32));
  return PMX_SUCCESS;
}

// Warm start of the level-2 eigensolver on the context's bricks: x0 rows
// [cell][mb] (brick order) from the network.
int cnn_warm_start(pmx_ctx* ctx, int mb, double* x0) {
  pmx_cnn* net = ctx->cnn;
  if (!net) {
    set_error("-warm_start cnn needs a network (pmx_set_warm_start)");
    return PMX_ERR_ORDER;
  }
  int mx = 0, my = 0;
  for (const Brick& b : ctx->h_bricks) {
    PMX_TRY(cnn_check_shape(b.nx, b.ny, b.nz));
    mx = std::max(mx, b.nx);
    my = std::max(my, b.ny);
  }
  const size_t cells = (size_t)ctx->ncell * kCnnCh;
  DBuf<unsigned short> a0, a1, a2;
  DBuf<float> out32;
  PMX_TRY(a0.alloc(cells));
  PMX_TRY(a1.alloc(cells));
  PMX_TRY(a2.alloc(cells));
  PMX_TRY(out32.alloc(cells));
  hipLaunchKernelGGL(k_cnn_input, dim3(ctx->nsub), dim3(256), 0, ctx->stream, ctx->geo.views,
                     ctx->kappa[0].p, ctx->kappa[1].p, ctx->kappa[2].p, a0.p);
  PMX_HIP(hipGetLastError());
  // the network's weights live on its own stream's allocations; order the
  // conversion before the convolutions on the context stream
  hipLaunchKernelGGL(k_cnn_weights, dim3((3 * kCnnLayerW + 255) / 256), dim3(256), 0, ctx->stream,
                     net->params.p, net->wf.p, net->wt.p);
  PMX_HIP(hipGetLastError());
  PMX_TRY(cnn_forward(net, ctx->stream, ctx->geo.views, ctx->nsub, cnn_tile_bytes(mx, my), a0.p,
                      a1.p, a2.p, out32.p));
  const int cout = std::min(net->cout, mb);
  hipLaunchKernelGGL(k_cnn_x0, dim3((ctx->ncell + 255) / 256), dim3(256), 0, ctx->stream, out32.p,
                     (size_t)ctx->ncell, cout, mb, ctx->seed ^ 0xC0FFEEull, x0);
  PMX_HIP(hipGetLastError());
  PMX_HIP(hipStreamSynchronize(ctx->stream));
  return PMX_SUCCESS;
}

}  // namespace pmx

extern "C" {

int pmx_cnn_create(pmx_cnn** out, int cout, const float* params, unsigned long long seed) {
  if (!out) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  *out = nullptr;
  if (cout < 1 || cout > kCnnCh) {
    set_error("CNN: cout=%d must be in 1..%d", cout, kCnnCh);
    return PMX_ERR_ARG_OUTOFRANGE;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    set_error("no HIP device available");
    return PMX_ERR_LIB;
  }
  pmx_cnn* net = new pmx_cnn();
  net->cout = cout;
  auto fail = [&](int rc) {
    delete net;
    return rc;
  };
  if (hipStreamCreate(&net->stream) != hipSuccess) {
    set_error("hipStreamCreate failed");
    return fail(PMX_ERR_LIB);
  }
  const int np = 3 * kCnnLayer;
  std::vector<float> h(np, 0.0f);
  if (params) {
    std::copy(params, params + np, h.begin());
  } else {
    // He-uniform init of the live weights; padding channels stay zero.
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
    auto uni = [&]() {
      st ^= st << 13;
      st ^= st >> 7;
      st ^= st << 17;
      return (double)(st >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    const int cin[3] = {4, kCnnCh, kCnnCh}, cot[3] = {kCnnCh, kCnnCh, cout};
    for (int l = 0; l < 3; ++l) {
      const double bound = std::sqrt(6.0 / (cin[l] * kCnnTaps));
      for (int co = 0; co < cot[l]; ++co)
        for (int t = 0; t < kCnnTaps; ++t)
          for (int ci = 0; ci < cin[l]; ++ci)
            h[(size_t)l * kCnnLayer + (co * kCnnTaps + t) * kCnnCh + ci] = (float)(bound * uni());
    }
  }
  // zero the padding so it stays zero
  for (int l = 0; l < 3; ++l)
    for (int co = 0; co < kCnnCh; ++co)
      for (int t = 0; t < kCnnTaps; ++t)
        for (int ci = 0; ci < kCnnCh; ++ci) {
          const bool live = (l == 0 ? ci < 4 : true) && (l == 2 ? co < cout : true);
          if (!live) h[(size_t)l * kCnnLayer + (co * kCnnTaps + t) * kCnnCh + ci] = 0.0f;
        }
  for (int co = cout; co < kCnnCh; ++co) h[2 * kCnnLayer + kCnnLayerW + co] = 0.0f;
  int rc = net->params.alloc(np);
  if (!rc) rc = net->grad.alloc(np);
  if (!rc) rc = net->m.alloc(np);
  if (!rc) rc = net->s.alloc(np);
  if (!rc) rc = net->wf.alloc(3 * kCnnLayerW);
  if (!rc) rc = net->wt.alloc(3 * kCnnLayerW);
  if (rc) return fail(rc);
  if (hipMemcpy(net->params.p, h.data(), np * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(net->m.p, 0, np * sizeof(float)) != hipSuccess ||
      hipMemset(net->s.p, 0, np * sizeof(float)) != hipSuccess ||
      hipMemset(net->grad.p, 0, np * sizeof(float)) != hipSuccess) {
    set_error("CNN: device copy failed");
    return fail(PMX_ERR_LIB);
  }
  *out = net;
  return PMX_SUCCESS;
}

int pmx_cnn_num_params(void) { return 3 * kCnnLayer; }

int pmx_cnn_get_params(const pmx_cnn* net, float* params, float* grad) {
  if (!net || (!params && !grad)) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  const size_t bytes = 3 * kCnnLayer * sizeof(float);
  PMX_HIP(hipStreamSynchronize(net->stream));
  if (params) PMX_HIP(hipMemcpy(params, net->params.p, bytes, hipMemcpyDeviceToHost));
  if (grad) PMX_HIP(hipMemcpy(grad, net->grad.p, bytes, hipMemcpyDeviceToHost));
  return PMX_SUCCESS;
}

int pmx_cnn_set_params(pmx_cnn* net, const float* params) {
  if (!net || !params) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  PMX_HIP(hipStreamSynchronize(net->stream));
  PMX_HIP(hipMemcpy(net->params.p, params, 3 * kCnnLayer * sizeof(float), hipMemcpyHostToDevice));
  return PMX_SUCCESS;
}

int pmx_cnn_predict(pmx_cnn* net, int nb, int nx, int ny, int nz, const unsigned short* d_in,
                    float* d_out) {
  if (!net || !d_in || !d_out) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  PMX_TRY(cnn_geometry(net, nb, nx, ny, nz));
  PMX_TRY(cnn_convert(net));
  PMX_TRY(cnn_forward(net, net->stream, net->views.p, nb, cnn_tile_bytes(nx, ny), d_in, net->a1.p,
                      net->a2.p, d_out));
  PMX_HIP(hipStreamSynchronize(net->stream));
  return PMX_SUCCESS;
}

int pmx_cnn_loss_grad(pmx_cnn* net, int nb, int nx, int ny, int nz, const unsigned short* d_in,
                      const float* d_target, int loss_kind, double* loss) {
  if (!net || !d_in || !d_target) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  PMX_TRY(cnn_geometry(net, nb, nx, ny, nz));
  PMX_TRY(cnn_convert(net));
  hipStream_t st = net->stream;
  const size_t lds = cnn_tile_bytes(nx, ny);
  const BrickView* views = net->views.p;
  PMX_TRY(cnn_forward(net, st, views, nb, lds, d_in, net->a1.p, net->a2.p, net->out32.p));
  const size_t ncell = (size_t)nb * nx * ny * nz;
  int mb = 512;
  if (loss_kind == 0) {
    const float scale = 1.0f / (float)(ncell * net->cout);
    hipLaunchKernelGGL(k_cnn_mse, dim3(mb), dim3(256), 0, st, net->out32.p, d_target, ncell,
                       net->cout, scale, net->g3.p, net->loss_part.p);
  } else if (loss_kind == 1) {
    if (nb > 1024) {
      set_error("CNN: at most 1024 bricks per subspace-loss batch");
      return PMX_ERR_ARG_OUTOFRANGE;
    }
    mb = nb;
    hipLaunchKernelGGL(k_cnn_subspace, dim3(nb), dim3(256), 0, st, views, net->out32.p, d_target,
                       net->cout, 1.0f / (float)(nb * net->cout), net->g3.p, net->loss_part.p);
  } else {
    set_error("CNN: loss_kind must be 0 (MSE) or 1 (subspace)");
    return PMX_ERR_ARG_WRONG;
  }
  PMX_HIP(hipGetLastError());
  // layer 3
  const size_t wl = std::max(lds, (size_t)kCnnLayerW * sizeof(float));
  float* G = net->grad.p;
  const int rb = (kCnnLayer + 255) / 256;
  hipLaunchKernelGGL(k_cnn_wgrad, dim3(nb), dim3(256), wl, st, views, net->a2.p, net->g3.p,
                     net->part.p);
  hipLaunchKernelGGL(k_cnn_reduce, dim3(rb), dim3(256), 0, st, net->part.p, nb,
                     G + 2 * kCnnLayer, 0);
  PMX_TRY(cnn_conv(st, views, nb, lds, net->g3.p, net->wt.p + 2 * kCnnLayerW, nullptr, 0,
                   net->a2.p, net->g2.p, nullptr));
  // layer 2
  hipLaunchKernelGGL(k_cnn_wgrad, dim3(nb), dim3(256), wl, st, views, net->a1.p, net->g2.p,
                     net->part.p);
  hipLaunchKernelGGL(k_cnn_reduce, dim3(rb), dim3(256), 0, st, net->part.p, nb, G + kCnnLayer, 0);
  PMX_TRY(cnn_conv(st, views, nb, lds, net->g2.p, net->wt.p + kCnnLayerW, nullptr, 0, net->a1.p,
                   net->g1.p, nullptr));
  // layer 1
  hipLaunchKernelGGL(k_cnn_wgrad, dim3(nb), dim3(256), wl, st, views, d_in, net->g1.p, net->part.p);
  hipLaunchKernelGGL(k_cnn_reduce, dim3(rb), dim3(256), 0, st, net->part.p, nb, G, 0);
  PMX_HIP(hipGetLastError());
  if (loss) {
    std::vector<double> lp(mb);
    PMX_HIP(hipMemcpyAsync(lp.data(), net->loss_part.p, mb * sizeof(double),
                           hipMemcpyDeviceToHost, st));
    PMX_HIP(hipStreamSynchronize(st));
    double s = 0.0;
    for (double v : lp) s += v;
    *loss = s;
  } else {
    PMX_HIP(hipStreamSynchronize(st));
  }
  return PMX_SUCCESS;
}

int pmx_cnn_adam_step(pmx_cnn* net, float lr, float beta1, float beta2, float eps) {
  if (!net) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  net->adam_t += 1;
  const float c1 = 1.0f - std::pow(beta1, (float)net->adam_t);
  const float c2 = 1.0f - std::pow(beta2, (float)net->adam_t);
  const int n = 3 * kCnnLayer;
  hipLaunchKernelGGL(k_cnn_adam, dim3((n + 255) / 256), dim3(256), 0, net->stream, net->params.p,
                     net->grad.p, net->m.p, net->s.p, n, lr, beta1, beta2, eps, c1, c2);
  PMX_HIP(hipGetLastError());
  return PMX_SUCCESS;
}

int pmx_cnn_destroy(pmx_cnn* net) {
  if (!net) return PMX_SUCCESS;
  if (net->stream) {
    (void)hipStreamSynchronize(net->stream);
    (void)hipStreamDestroy(net->stream);
  }
  delete net;
  return PMX_SUCCESS;
}

int pmx_cnn_inputs(pmx_ctx* ctx, unsigned short* d_in) {
  if (!ctx || !d_in) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  if (!ctx->have_kappa) {
    set_error("pmx_cnn_inputs before pmx_set_kappa");
    return PMX_ERR_ORDER;
  }
  hipLaunchKernelGGL(k_cnn_input, dim3(ctx->nsub), dim3(256), 0, ctx->stream, ctx->geo.views,
                     ctx->kappa[0].p, ctx->kappa[1].p, ctx->kappa[2].p, d_in);
  PMX_HIP(hipGetLastError());
  PMX_HIP(hipStreamSynchronize(ctx->stream));
  return PMX_SUCCESS;
}

int pmx_cnn_targets(pmx_ctx* ctx, int cout, float* d_target) {
  if (!ctx || !d_target) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  if (!ctx->setup_done) {
    set_error("pmx_cnn_targets before pmx_setup");
    return PMX_ERR_ORDER;
  }
  if (cout < 1 || cout > ctx->en_lv1 - 1 || cout > kCnnCh) {
    set_error("CNN: cout=%d must be in 1..en_lv1-1", cout);
    return PMX_ERR_ARG_OUTOFRANGE;
  }
  hipLaunchKernelGGL(k_cnn_targets, dim3(ctx->nsub), dim3(256), 0, ctx->stream, ctx->geo.views,
                     ctx->phi.p, ctx->en_lv1, cout, d_target);
  PMX_HIP(hipGetLastError());
  PMX_HIP(hipStreamSynchronize(ctx->stream));
  return PMX_SUCCESS;
}

int pmx_set_warm_start(pmx_ctx* ctx, pmx_cnn* net) {
  if (!ctx) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  ctx->cnn = net;
  return PMX_SUCCESS;
}

}  // extern "C"
