// Internal declarations of the MI355X implementation of PreMixFEM's
// three-level GMsFEM preconditioner.  The public C ABI is include/premixfem_gpu.h.
//
// Device data layout ("brick order"): cells are stored subdomain-major, each
// subdomain (brick) contiguous in its own local x-fastest order
// (ez-startz)*ny*nx + (ey-starty)*nx + (ex-startx), i.e. the local vector
// order of src/PreMixFEM_3D.c:104-107, bricks in x-fastest subdomain order
// (src/PreMixFEM_3D.c:23-27).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "premixfem_gpu.h"

namespace pmx {

constexpr int kBlock = 256;        // threads per workgroup for cell kernels
constexpr int kStreamBlock = 1024; // largest workgroup of the streaming per-brick kernels
constexpr int kMaxPlanes = 768;    // hyperplanes of a brick (3 * 256 - 2 at most)
constexpr int kMaxShapes = 8;      // uneven partitions give at most 2 lengths per axis
constexpr int kMaxEig = 16;        // largest en_lv1 / en_lv2 supported on device
constexpr int kMaxCC = 64;         // largest level-3 space (all process boxes together)
constexpr double kNil = 1.0e-12;   // NIL of src/PreMixFEM_3D.h:14
constexpr int kBandAutoNc = 12288; // coarse size above which level 2 uses the band factor

void set_error(const char* fmt, ...);

#define PMX_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      ::pmx::set_error("%s:%d: %s: %s", __FILE__, __LINE__, #call,                 \
                       hipGetErrorString(e_));                                     \
      return PMX_ERR_LIB;                                                          \
    }                                                                              \
  } while (0)

#define PMX_TRY(call)                                                              \
  do {                                                                             \
    int rc_ = (call);                                                              \
    if (rc_ != PMX_SUCCESS) return rc_;                                            \
  } while (0)

// One subdomain on the device.
struct Brick {
  int off;          // first cell (brick order)
  int nx, ny, nz;   // local extents
  int gx, gy, gz;   // global start
  int shape;        // index into the shape tables
  int nb[6];        // neighbour bricks: -x, +x, -y, +y, -z, +z (-1 = domain boundary)
  int dof;          // first coarse dof (prefix sum of eigen counts)
  int neig;         // eigen count of this subdomain
  int plane_off;    // offset of this shape in the plane-position tables
  int nplanes;      // nx+ny+nz-2
  int lid;          // subdomain index inside its problem (seeds its random LOBPCG start)
  int box;          // process box (x-fastest rank of the -procs_virtual decomposition)
};

// Per-brick extents and neighbour offsets, precomputed on the host so a
// kernel's prologue is one (scalar) load.
struct BrickView {
  int off, nx, ny, nz, nxny, n;
  int noff[6];   // first cell of the neighbour brick, -1: domain boundary
  int next[6];   // the neighbour's extent along the shared axis
};

// Geometry passed by value to kernels.
struct Geo {
  int M, N, P;
  int sdx, sdy, sdz;
  int nsub;
  int ncell;
  const Brick* bricks;
  const BrickView* views;
  const int* order;       // brick handled by workgroup b of a per-brick launch
  const int* ax_sub[3];   // per axis: subdomain index of each global coordinate
  const int* ax_loc[3];   // per axis: local coordinate
  const int* plane_start; // per shape: nplanes+1 prefix counts (at shape's plane_off + shape*? see below)
  const int* plane_cell;  // per shape: packed (lx | ly<<10 | lz<<20) in plane order
  const int* shape_pstart;// per shape: offset into plane_start
  const int* shape_coff;  // per shape: offset into plane_cell
  int pin;                // brick index of the pinned cell, -1 if none
  double tau;             // pin strength
  const double* stop;     // PCG "done" flag: kernels return at once when set (nullptr: none)
  const double* wd;       // per-cell diagonal (well) term, brick order; nullptr: none
  const float* wd32;      // its FP32 copy (FP32 preconditioner)
  // Exact level-1 factors (-use_full_Cholesky_lv1): per brick the band
  // Cholesky factor by columns (l1L[off + j*(bw+1) + d] = L(j+d, j)) and by
  // rows (l1U[off + j*(bw+1) + d] = L(j, j-d)), bw = nx*ny; nullptr: none.
  const double* l1L;
  const double* l1U;
  const long long* l1off;
};

// The well term of the operator diagonal in the precision of the face
// coefficients TT (nullptr when there is none).
template <typename TT>
__device__ __forceinline__ const TT* well_term(const Geo& g);
template <>
__device__ __forceinline__ const double* well_term<double>(const Geo& g) { return g.wd; }
template <>
__device__ __forceinline__ const float* well_term<float>(const Geo& g) { return g.wd32; }

// What a captured PCG chunk depends on besides the context's buffers.
struct CgGraphKey {
  int setup_gen, chunk, max_it;
  double abstol, dtol;
  bool operator==(const CgGraphKey& o) const {
    return setup_gen == o.setup_gen && chunk == o.chunk && max_it == o.max_it &&
           abstol == o.abstol && dtol == o.dtol;
  }
};

// Device-resident PCG scalars (indices into ctx->scal).
enum ScalIdx {
  kSBeta = 0, kSBetaOld = 1, kSDpi = 2, kSDpiOld = 3, kSRnorm2 = 4, kSDone = 6, kSIts = 7,
  kSTtol = 8, kSRnorm0 = 9, kSReason = 12, kSIter = 13, kSNumScal = 16
};

#define PMX_STOP_CHECK(ptr) \
  if ((ptr) != nullptr && *(ptr) != 0.0) return;

// Neighbour index in brick order, or -1.
__device__ __forceinline__ int nbr_index(const Geo& g, const Brick& b, int li, int lx, int ly,
                                         int lz, int dir) {
  switch (dir) {
    case 0:
      if (lx > 0) return li - 1;
      if (b.nb[0] >= 0) {
        const Brick& o = g.bricks[b.nb[0]];
        return o.off + (lz * o.ny + ly) * o.nx + (o.nx - 1);
      }
      return -1;
    case 1:
      if (lx < b.nx - 1) return li + 1;
      if (b.nb[1] >= 0) {
        const Brick& o = g.bricks[b.nb[1]];
        return o.off + (lz * o.ny + ly) * o.nx;
      }
      return -1;
    case 2:
      if (ly > 0) return li - b.nx;
      if (b.nb[2] >= 0) {
        const Brick& o = g.bricks[b.nb[2]];
        return o.off + (lz * o.ny + (o.ny - 1)) * o.nx + lx;
      }
      return -1;
    case 3:
      if (ly < b.ny - 1) return li + b.nx;
      if (b.nb[3] >= 0) {
        const Brick& o = g.bricks[b.nb[3]];
        return o.off + (lz * o.ny) * o.nx + lx;
      }
      return -1;
    case 4:
      if (lz > 0) return li - b.nx * b.ny;
      if (b.nb[4] >= 0) {
        const Brick& o = g.bricks[b.nb[4]];
        return o.off + ((o.nz - 1) * o.ny + ly) * o.nx + lx;
      }
      return -1;
    default:
      if (lz < b.nz - 1) return li + b.nx * b.ny;
      if (b.nb[5] >= 0) {
        const Brick& o = g.bricks[b.nb[5]];
        return o.off + ly * o.nx + lx;
      }
      return -1;
  }
}

// One brick and the offsets/extents of its six neighbour bricks, held in
// registers by every thread of the workgroup that owns the brick.  An x
// neighbour shares ny and nz, a y neighbour nx and nz, a z neighbour nx and ny.
__device__ __forceinline__ BrickView load_view(const Geo& g, int s) { return g.views[s]; }

// Workgroups are dealt to the 8 XCDs round-robin (b mod 8), and each XCD has
// its own L2.  order[] gives XCD x a spatially compact run of bricks (a
// chunk of the Morton order), so the halo reads of neighbouring bricks hit
// the same L2.
__device__ __forceinline__ int block_brick(const Geo& g) { return g.order[blockIdx.x]; }

// Brick-order indices of the -x,+x,-y,+y,-z,+z neighbours of cell c at local
// (lx,ly,lz); -1 on the domain boundary.
__device__ __forceinline__ void nbr6(const BrickView& v, int c, int lx, int ly, int lz, int nb[6]) {
  nb[0] = lx > 0 ? c - 1
                 : (v.noff[0] >= 0 ? v.noff[0] + (lz * v.ny + ly) * v.next[0] + v.next[0] - 1 : -1);
  nb[1] = lx < v.nx - 1 ? c + 1 : (v.noff[1] >= 0 ? v.noff[1] + (lz * v.ny + ly) * v.next[1] : -1);
  nb[2] = ly > 0 ? c - v.nx
                 : (v.noff[2] >= 0 ? v.noff[2] + (lz * v.next[2] + v.next[2] - 1) * v.nx + lx : -1);
  nb[3] = ly < v.ny - 1 ? c + v.nx : (v.noff[3] >= 0 ? v.noff[3] + lz * v.next[3] * v.nx + lx : -1);
  nb[4] = lz > 0 ? c - v.nxny
                 : (v.noff[4] >= 0 ? v.noff[4] + ((v.next[4] - 1) * v.ny + ly) * v.nx + lx : -1);
  nb[5] = lz < v.nz - 1 ? c + v.nxny : (v.noff[5] >= 0 ? v.noff[5] + ly * v.nx + lx : -1);
}

// Line-based walk over a brick without divisions in the loop: thread t owns
// x-coordinate t % nx of lines t/nx, t/nx + lpi, ... (lpi = blockDim/nx).
// Requires nx <= blockDim.x.
template <typename F>
__device__ __forceinline__ void for_cells(const BrickView& v, F&& f) {
  const int lpi = (int)blockDim.x / v.nx;
  const int lx = (int)threadIdx.x % v.nx;
  const int l0 = (int)threadIdx.x / v.nx;
  if (l0 >= lpi) return;
  int ly = l0 % v.ny, lz = l0 / v.ny;
  while (lz < v.nz) {
    f(v.off + (lz * v.ny + ly) * v.nx + lx, lx, ly, lz);
    ly += lpi;
    while (ly >= v.ny) {
      ly -= v.ny;
      ++lz;
    }
  }
}

// Block-wide deterministic sum of one double (up to 16 waves; sh holds 16).
__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
    sh[0] = r;
  }
  __syncthreads();
  r = sh[0];
  __syncthreads();
  return r;
}

// Device buffer helper.
template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  // Owns its allocation: freed on scope exit (hipFree waits for the device,
  // so kernels still reading a local buffer finish first).  Never freed
  // inside a stream capture: only setup-time code holds local buffers.
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DBuf() { release(); }
  int alloc(size_t count) {
    if (count == n && p) return PMX_SUCCESS;
    release();
    if (count == 0) return PMX_SUCCESS;
    PMX_HIP(hipMalloc(&p, count * sizeof(T)));
    n = count;
    return PMX_SUCCESS;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

}  // namespace pmx

// Full context behind the opaque pmx_ctx handle.
struct pmx_cnn;

struct pmx_ctx {
  // ---- options (src/PreMixFEM_3D.c:1662-1705) ----
  double dom[3];
  int mesh[3];
  int sd[3];                        // subdomains per axis over the whole grid
  int pv[3] = {1, 1, 1};            // process boxes per axis (-procs_virtual)
  int sd_box[3] = {2, 2, 2};        // subdomains per axis inside one box
  int nbox = 1;
  int si_lv1 = 1, si_lv2 = 1;
  int en_lv1 = 4, en_lv2 = 4;
  int eg_op = PMX_EIG_OP_MOD;
  double eb_lv1 = 1.0 / pmx::kNil, eb_lv2 = 1.0 / pmx::kNil;
  int use_W_cycle = 0, no_shift_A_cc = 0, use_full_Cholesky_lv1 = 0, use_2level = 0;
  int pc_precision = 64;
  int pin[3] = {-1, -1, -1};
  int warm_start = 0;           // 0 random, 1 CNN prediction, 2 caller basis
  pmx_cnn* cnn = nullptr;       // warm-start network (owned by the caller)
  double eig_tol = 1.0e-8;      // LOBPCG relative residual tolerance
  int eig_maxit = 400;
  int eig_guard = 3;            // extra LOBPCG block vectors beyond en_lv1-1
  int quad_ok = 1;              // quad-cell stream kernels where the bricks allow (-quad 0: off)
  int eig_autotune = 0;         // time the LOBPCG block sizes on a sample and keep the fastest
  double t_tune = 0.0;          // seconds the autotuner spent in the last setup (0: cached)
  int eig_block = 0;            // block size of the last LOBPCG solve
  double eig_sigma = 1.0e-6;    // shift of the LOBPCG preconditioner A^N + sigma M
  int log_stages = 0;           // synchronise and time every apply stage
  int kernel_timing = 0;        // HIP-event timing of the solve-phase kernels and the LOBPCG launch
  double t_eig_kernel = 0.0;    // seconds of the last LOBPCG launch (-kernel_timing only)
  int eig_debug = -1;           // trace the LOBPCG iterations of one subdomain
  int bicc_prefetch = 1;        // block ICC(0) level-2 solve with row data loaded a round ahead
  int eig_kernel = 1;           // LOBPCG kernel: 1 matrix-core passes, 0 VALU
  int eig_ortho = 1;            // matrix-core LOBPCG: W orthogonalised in coefficient space (1)
                                // or by two explicit projection passes (0)
  int eig_shape = 0;            // matrix-core LOBPCG on large subdomains: 0 one per CU
                                // (512 threads), 1 two per CU (256 threads)
  int eig_order = 1;            // launch LOBPCG hardest-first from the previous setup's counts
  std::vector<int> h_lob_its;   // LOBPCG iterations per subdomain of the last setup
  pmx::DBuf<int> d_lob_order;
  int batch_threads = 8;        // pmx_setup_batch: host threads for the members' coarse levels
  uint64_t seed = 12345;

  // ---- derived geometry ----
  double Hx, Hy, Hz, meas_elem;
  double coef[3];               // |F|^2/|E| per direction
  int ncell = 0, nsub = 0;
  std::vector<pmx::Brick> h_bricks;
  std::vector<int> h_ax_sub[3], h_ax_loc[3];
  std::vector<int> h_starts[3], h_lens[3];
  std::vector<int> h_plane_start, h_plane_cell, h_shape_pstart, h_shape_coff;
  std::vector<int> h_nat2brick;     // natural index -> brick index
  int nshape = 0;

  pmx::DBuf<pmx::Brick> d_bricks;
  pmx::DBuf<pmx::BrickView> d_views;
  pmx::DBuf<int> d_order;
  pmx::DBuf<int> d_ax[6];
  pmx::DBuf<int> d_plane_start, d_plane_cell, d_shape_pstart, d_shape_coff;
  pmx::Geo geo;

  // ---- stream ----
  hipStream_t stream = nullptr;

  // ---- fields (brick order) ----
  int have_kappa = 0;
  pmx::DBuf<double> kappa[3];
  pmx::DBuf<double> T[3];       // lower-face coefficient of each cell per direction
  pmx::DBuf<float> T32[3];
  double tau = 0.0;
  int pin_brick = -1;
  // Per-cell diagonal term of the operator (the producer wells of
  // src/spe10.c:780-786), brick order; empty = none.
  int have_wd = 0;
  pmx::DBuf<double> wd;
  pmx::DBuf<float> wd32;

  // ---- level 1 ----
  pmx::DBuf<double> fac_f, fac_b;   // plane-major {dinv, cx, cy, cz} lower / upper couplings
  pmx::DBuf<double> l1L, l1U;       // exact level-1 band factors (-use_full_Cholesky_lv1)
  pmx::DBuf<long long> l1off;       // per brick offset into l1L / l1U
  pmx::DBuf<float> fac32_f, fac32_b;

  // ---- level 2 ----
  pmx::DBuf<double> mass;           // eigen mass diagonal
  pmx::DBuf<double> phi;            // [cell][en_lv1] basis rows
  pmx::DBuf<float> phi32;
  std::vector<int> h_eigen_num;     // per subdomain
  std::vector<double> h_eig_lv1;    // [nsub][en_lv1] eigenvalues
  std::vector<double> h_eigen_min_lv1, h_eigen_max_lv1;
  std::vector<int> h_dof;           // prefix sums
  int nc = 0;                        // coarse size
  pmx::DBuf<double> Ac;             // dense nc x nc Galerkin matrix
  pmx::DBuf<double> Acb;            // block-sparse rows: [nsub][7][kmax][kmax]
  pmx::DBuf<double> Ainv;           // dense inverse of the deflated coarse matrix B
  pmx::DBuf<double> c_z, c_g, c_h, c_cs;  // near-null vector, pin row, B^{-1} g, Woodbury scalars
  double coarse_gamma = 0.0;
  int coarse_deflate = 1;           // invert A_c deflated + Woodbury (1) or directly (0)
  int coarse_composite = 1;         // dense path: one composed coarse operator (1) or the
                                    // coarse stages one by one with their residual updates (0)
  pmx::DBuf<double> Fop;            // composite coarse operator of the V-cycle (nc x nc)
  int use_fop = 0;                  // the apply uses Fop (dense path, more than one stage)
  std::vector<int> coarse_prog;     // stage list: 2 = level 2, 3 = level 3
  // One process with an exact level-2 factor: the coarse stages after the
  // first level-2 stage vanish in exact arithmetic, and the program is cut
  // to that stage (-coarse_collapse 1, default); 0 keeps every stage.
  int coarse_collapse = 1;
  int coarse_collapsed = 0;         // the last setup cut the program
  // Level-2 factor of the spe10 fork (src/spe10.c:509-516): ICC(0) of A_c,
  // i.e. a block incomplete Cholesky over the subdomains (-lv2_icc).
  int lv2_icc = 0;
  // Exact level-2 factor in band storage for large coarse spaces (-coarse_band
  // 1, or automatically above kBandAutoNc unknowns): tiles of the Cholesky
  // factor of A_c and the inverses of its diagonal tiles.
  int coarse_band = -1;             // -1 auto, 0 dense, 1 band
  int use_band = 0;                 // chosen at setup
  pmx::DBuf<double> band, band_dinv, band_tmp;
  int band_nt = 0, band_w = 0;
  // The same exact factor by z-planes of subdomains (default on this path,
  // -coarse_planes 0: the band factor): the inverse Schur complements G_z
  // of the block-tridiagonal A_c, plane after plane, row-major.
  int coarse_planes = 1;
  int use_planes = 0;               // chosen at setup
  pmx::DBuf<double> pl_G, pl_w;
  std::vector<int> pl_dof;          // first dof of every plane (+ nc)
  std::vector<size_t> pl_goff;      // offset of G_z in pl_G (+ total)
  size_t pl_bytes = 0;
  pmx::DBuf<double> bicc_dinv;      // [nsub][kmax][kmax] inverse pivot blocks
  pmx::DBuf<int> bicc_plane;        // coarse hyperplane prefix counts (nplanes+1)
  pmx::DBuf<int> bicc_elem;         // subdomains in hyperplane order
  pmx::DBuf<int> bicc_slot;         // per-position sweep descriptors (hyperplane order)
  int bicc_nplanes = 0, bicc_maxcnt = 0;
  int max_plane = 0;                // most cells in one hyperplane of any brick shape
  int lds_max = 65536;              // LDS bytes per workgroup of the device
  int eig_iters_max = 0, eig_iters_sum = 0, eig_fallbacks = 0, eig_stagnated = 0;
  double eig_stag = 100.0;      // stagnation acceptance factor (x eig_tol)
  pmx::DBuf<double> lob_buf;        // LOBPCG blocks X, AX, W, AW, P, AP (kept for warm starts)

  // ---- level 3 ----
  int eigen_num_lv2 = 0;
  double eigen_min_lv2 = 0.0, eigen_max_lv2 = 0.0;
  std::vector<double> h_eig_lv2;
  std::vector<int> h_eigen_num_lv2_box;  // level-3 count of every process box
  pmx::DBuf<double> V2;             // [nc][en_lv2]
  pmx::DBuf<double> Acc_lu;         // k2 x k2 LU (row-major) + pivots packed after
  int lv3_iters = 0;
  int lv3_warm = 1;                 // -lv2_icc: start level 3 from the previous eigenvectors
  int lv3_prev_n = 0;               // coarse size of the previous level-3 solve
  pmx::DBuf<double> lv3_rand;       // level-3 random start block (seeded), kept
  unsigned long long lv3_rand_seed = 0;  // for the setups that follow
  double lv3_floor = 0.0;           // attained residual when the level-3 solver stagnated

  // ---- work vectors ----
  pmx::DBuf<double> w_x, w_y, w_r, w_b, w_p, w_z, w_q, w_tmp;   // brick order FP64
  pmx::DBuf<float> w_y32, w_y32b;
  pmx::DBuf<double> w_yb;
  pmx::DBuf<double> c_a, c_b, c_c, c_d, c_e;                       // coarse vectors
  pmx::DBuf<double> partial;                                       // per-block partial sums
  pmx::DBuf<double> scal;                                          // device scalars
  // What the PCG kernels use for the partial sums (6 nsub) and the device
  // scalars: partial / scal, or the library's fixed device array under
  // pmx_solve_device_noalloc.
  double* cg_part = nullptr;
  double* cg_scal = nullptr;
  double* scratch_dev = nullptr;    // the fixed array's address on this context's device
  pmx::DBuf<double> nat_buf;                                       // natural-order staging

  // ---- state / stats ----
  int setup_done = 0;
  double t_stages[PMX_MAX_LOG_STAGES] = {0};
  int last_its = 0;
  double last_rnorm = 0.0, last_rnorm0 = 0.0;
  int last_reason = 0;
  double t_solve = 0.0;
  double cg_rtol = 1e-8, cg_abstol = 1e-50, cg_dtol = 1e5;
  int cg_max_it = 10000;
  int cg_chunk = 8;                 // PCG iterations enqueued between host checks
  int cg_graph = 1;                 // replay the chunk as a captured HIP graph
  int setup_gen = 0;                // bumped by every setup (invalidates the graph)
  pmx::CgGraphKey cg_key{};
  hipGraph_t cg_graph_h = nullptr;
  hipGraphExec_t cg_exec = nullptr;
  double h_scal[16];
  std::vector<hipEvent_t> kt_events;
  std::vector<int> kt_pending;
  double kt_ms[PMX_NUM_TIMED_KERNELS] = {0};
  int kt_count[PMX_NUM_TIMED_KERNELS] = {0};
  std::vector<int> h_eigen_num_given;  // caller-fixed eigen counts (setup from a basis)
};
