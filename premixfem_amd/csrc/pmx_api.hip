// Host side of the C ABI (include/premixfem_gpu.h): options, geometry and
// brick tables, setup orchestration (src/PreMixFEM_3D.c:_PC_setup), the
// V-cycle (PC_apply_vec) and the GPU-resident PCG (PETSc KSPSolve_CG).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <functional>
#include <cstdarg>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <thread>
#include <tuple>

#include "pmx_kernels.h"

namespace pmx {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

// Fixed device array of pmx_solve_device_noalloc: the PCG's per-subdomain
// partial sums (6 nsub doubles) and its device scalars (16 doubles).  Each
// GPU holds its own copy until the process ends; host code never uses the
// name `scratch` for its memory, only the address hipGetSymbolAddress gives
// for the current device (scratch_address).
constexpr int kScratchDoubles = PMX_NOALLOC_SCRATCH_DOUBLES;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

// The array's address on device dev (the current device), looked up once.
static int scratch_address(int dev, double** out) {
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device index %d beyond the %d devices this library tracks", dev, kMaxDevices);
    return PMX_ERR_SUP;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("hipGetSymbolAddress of the solve scratch array: %s", hipGetErrorString(e));
      return PMX_ERR_LIB;
    }
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return PMX_SUCCESS;
}

static double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Per axis: len_i = n/sd + (n%sd > i), start by prefix (src/PreMixFEM_3D.c:62-72).
static void partition(int n, int sd, std::vector<int>& st, std::vector<int>& ln) {
  st.assign(sd, 0);
  ln.assign(sd, 0);
  for (int i = 0; i < sd; ++i) ln[i] = n / sd + ((n % sd) > i ? 1 : 0);
  for (int i = 1; i < sd; ++i) st[i] = st[i - 1] + ln[i - 1];
}

// With p process boxes along an axis: PETSc's default DMDA ownership
// (l_j = n/p + (n%p > j)) and, inside every box, the partition above from
// the box's own corner (src/PreMixFEM_3D.c:58-72), concatenated.
static void partition_boxes(int n, int p, int sd, std::vector<int>& st, std::vector<int>& ln) {
  std::vector<int> bst, bln, s1, l1;
  partition(n, p, bst, bln);
  st.clear();
  ln.clear();
  for (int j = 0; j < p; ++j) {
    partition(bln[j], sd, s1, l1);
    for (int i = 0; i < sd; ++i) {
      st.push_back(bst[j] + s1[i]);
      ln.push_back(l1[i]);
    }
  }
}

// "-procs_virtual px,py,pz" (or one count for all axes).
static int parse_procs(const std::string& v, int* pv) {
  std::vector<std::string> items;
  size_t pos = 0;
  for (;;) {
    const size_t comma = v.find(',', pos);
    items.push_back(v.substr(pos, comma == std::string::npos ? std::string::npos : comma - pos));
    if (comma == std::string::npos) break;
    pos = comma + 1;
  }
  if (items.size() != 1 && items.size() != 3) return PMX_ERR_ARG_WRONG;
  int vals[3];
  for (size_t q = 0; q < items.size(); ++q) {
    char* end = nullptr;
    const long x = strtol(items[q].c_str(), &end, 10);
    if (items[q].empty() || *end) return PMX_ERR_ARG_WRONG;
    vals[q] = (int)x;
  }
  for (int a = 0; a < 3; ++a) pv[a] = items.size() == 1 ? vals[0] : vals[a];
  return PMX_SUCCESS;
}

static int parse_options(pmx_ctx* c, const char* options) {
  if (!options) return PMX_SUCCESS;
  std::istringstream is(options);
  std::vector<std::string> tok;
  std::string t;
  while (is >> t) tok.push_back(t);
  auto need = [&](size_t i) -> bool { return i + 1 < tok.size(); };
  for (size_t i = 0; i < tok.size(); ++i) {
    const std::string& k = tok[i];
    auto geti = [&](int& dst) -> int {
      if (!need(i)) {
        set_error("option %s needs a value", k.c_str());
        return PMX_ERR_ARG_WRONG;
      }
      char* end = nullptr;
      long v = strtol(tok[++i].c_str(), &end, 10);
      if (*end) {
        set_error("bad integer for %s: %s", k.c_str(), tok[i].c_str());
        return PMX_ERR_ARG_WRONG;
      }
      dst = (int)v;
      return PMX_SUCCESS;
    };
    auto getd = [&](double& dst) -> int {
      if (!need(i)) {
        set_error("option %s needs a value", k.c_str());
        return PMX_ERR_ARG_WRONG;
      }
      char* end = nullptr;
      double v = strtod(tok[++i].c_str(), &end);
      if (*end) {
        set_error("bad number for %s: %s", k.c_str(), tok[i].c_str());
        return PMX_ERR_ARG_WRONG;
      }
      dst = v;
      return PMX_SUCCESS;
    };
    if (k == "-sd") {
      int v;
      PMX_TRY(geti(v));
      c->sd[0] = c->sd[1] = c->sd[2] = v;
    } else if (k == "-sdx") {
      PMX_TRY(geti(c->sd[0]));
    } else if (k == "-sdy") {
      PMX_TRY(geti(c->sd[1]));
    } else if (k == "-sdz") {
      PMX_TRY(geti(c->sd[2]));
    } else if (k == "-si_lv1") {
      PMX_TRY(geti(c->si_lv1));
    } else if (k == "-si_lv2") {
      PMX_TRY(geti(c->si_lv2));
    } else if (k == "-en_lv1") {
      PMX_TRY(geti(c->en_lv1));
    } else if (k == "-en_lv2") {
      PMX_TRY(geti(c->en_lv2));
    } else if (k == "-eg_op") {
      PMX_TRY(geti(c->eg_op));
    } else if (k == "-eb_lv1") {
      PMX_TRY(getd(c->eb_lv1));
    } else if (k == "-eb_lv2") {
      PMX_TRY(getd(c->eb_lv2));
    } else if (k == "-use_W_cycle") {
      c->use_W_cycle = 1;
    } else if (k == "-no_shift_A_cc") {
      c->no_shift_A_cc = 1;
    } else if (k == "-use_full_Cholesky_lv1") {
      c->use_full_Cholesky_lv1 = 1;
    } else if (k == "-use_2level") {
      c->use_2level = 1;
    } else if (k == "-batch_threads") {
      PMX_TRY(geti(c->batch_threads));
    } else if (k == "-lv2_icc") {
      c->lv2_icc = 1;
    } else if (k == "-procs_virtual") {
      if (!need(i) || parse_procs(tok[i + 1], c->pv)) {
        set_error("-procs_virtual needs px,py,pz");
        return PMX_ERR_ARG_WRONG;
      }
      ++i;
    } else if (k == "-eig_autotune") {
      PMX_TRY(geti(c->eig_autotune));
    } else if (k == "-quad") {
      PMX_TRY(geti(c->quad_ok));
    } else if (k == "-coarse_composite") {
      PMX_TRY(geti(c->coarse_composite));
    } else if (k == "-coarse_collapse") {
      PMX_TRY(geti(c->coarse_collapse));
    } else if (k == "-coarse_planes") {
      PMX_TRY(geti(c->coarse_planes));
    } else if (k == "-coarse_band") {
      PMX_TRY(geti(c->coarse_band));
    } else if (k == "-coarse_deflate") {
      PMX_TRY(geti(c->coarse_deflate));
    } else if (k == "-lv3_warm") {
      PMX_TRY(geti(c->lv3_warm));
    } else if (k == "-pc_precision") {
      PMX_TRY(geti(c->pc_precision));
    } else if (k == "-eig_tol") {
      PMX_TRY(getd(c->eig_tol));
    } else if (k == "-eig_maxit") {
      PMX_TRY(geti(c->eig_maxit));
    } else if (k == "-eig_sigma") {
      PMX_TRY(getd(c->eig_sigma));
    } else if (k == "-eig_guard") {
      PMX_TRY(geti(c->eig_guard));
    } else if (k == "-bicc_prefetch") {
      PMX_TRY(geti(c->bicc_prefetch));
    } else if (k == "-eig_kernel") {
      PMX_TRY(geti(c->eig_kernel));
      if (c->eig_kernel < 0 || c->eig_kernel > 1) {
        set_error("-eig_kernel %d: 0 (VALU) or 1 (matrix cores)", c->eig_kernel);
        return PMX_ERR_ARG_OUTOFRANGE;
      }
    } else if (k == "-eig_ortho") {
      PMX_TRY(geti(c->eig_ortho));
      if (c->eig_ortho < 0 || c->eig_ortho > 1) {
        set_error("-eig_ortho %d: 0 (explicit projections) or 1 (coefficient space)",
                  c->eig_ortho);
        return PMX_ERR_ARG_OUTOFRANGE;
      }
    } else if (k == "-eig_shape") {
      PMX_TRY(geti(c->eig_shape));
      if (c->eig_shape < 0 || c->eig_shape > 1) {
        set_error("-eig_shape %d: 0 (one subdomain per CU) or 1 (two per CU)", c->eig_shape);
        return PMX_ERR_ARG_OUTOFRANGE;
      }
    } else if (k == "-eig_order") {
      PMX_TRY(geti(c->eig_order));
    } else if (k == "-eig_debug") {
      PMX_TRY(geti(c->eig_debug));
    } else if (k == "-cg_chunk") {
      PMX_TRY(geti(c->cg_chunk));
      c->cg_chunk = std::max(1, c->cg_chunk);
    } else if (k == "-eig_stag") {
      PMX_TRY(getd(c->eig_stag));
    } else if (k == "-cg_graph") {
      PMX_TRY(geti(c->cg_graph));
    } else if (k == "-kernel_timing") {
      c->kernel_timing = 1;
    } else if (k == "-log_stages") {
      c->log_stages = 1;
    } else if (k == "-seed") {
      int v;
      PMX_TRY(geti(v));
      c->seed = (uint64_t)v;
    } else if (k == "-pin") {
      if (!need(i)) {
        set_error("-pin needs x,y,z");
        return PMX_ERR_ARG_WRONG;
      }
      if (sscanf(tok[++i].c_str(), "%d,%d,%d", &c->pin[0], &c->pin[1], &c->pin[2]) != 3) {
        set_error("-pin needs x,y,z");
        return PMX_ERR_ARG_WRONG;
      }
    } else if (k == "-warm_start") {
      if (!need(i)) {
        set_error("-warm_start needs none|cnn|basis");
        return PMX_ERR_ARG_WRONG;
      }
      const std::string v = tok[++i];
      c->warm_start = v == "cnn" ? 1 : (v == "basis" ? 2 : 0);
    } else {
      set_error("unknown option %s", k.c_str());
      return PMX_ERR_ARG_WRONG;
    }
  }
  return PMX_SUCCESS;
}

template <typename T>
static int upload(DBuf<T>& d, const std::vector<T>& h) {
  PMX_TRY(d.alloc(h.size()));
  PMX_HIP(hipMemcpy(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return PMX_SUCCESS;
}

static int build_geometry(pmx_ctx* c) {
  const int M = c->mesh[0], N = c->mesh[1], P = c->mesh[2];
  c->ncell = M * N * P;
  for (int a = 0; a < 3; ++a) {
    partition_boxes(c->mesh[a], c->pv[a], c->sd_box[a], c->h_starts[a], c->h_lens[a]);
    c->h_ax_sub[a].assign(c->mesh[a], 0);
    c->h_ax_loc[a].assign(c->mesh[a], 0);
    for (int i = 0; i < c->sd[a]; ++i)
      for (int l = 0; l < c->h_lens[a][i]; ++l) {
        c->h_ax_sub[a][c->h_starts[a][i] + l] = i;
        c->h_ax_loc[a][c->h_starts[a][i] + l] = l;
      }
  }
  const int sdx = c->sd[0], sdy = c->sd[1], sdz = c->sd[2];
  c->nsub = sdx * sdy * sdz;
  c->h_bricks.assign(c->nsub, Brick{});
  std::map<std::tuple<int, int, int>, int> shapes;
  c->h_plane_start.clear();
  c->max_plane = 0;
  c->h_plane_cell.clear();
  c->h_shape_pstart.clear();
  c->h_shape_coff.clear();
  int off = 0;
  for (int sz = 0; sz < sdz; ++sz)
    for (int sy = 0; sy < sdy; ++sy)
      for (int sx = 0; sx < sdx; ++sx) {
        const int s = sx + sdx * (sy + sdy * sz);
        Brick& b = c->h_bricks[s];
        b.off = off;
        b.nx = c->h_lens[0][sx];
        b.ny = c->h_lens[1][sy];
        b.nz = c->h_lens[2][sz];
        b.gx = c->h_starts[0][sx];
        b.gy = c->h_starts[1][sy];
        b.gz = c->h_starts[2][sz];
        b.nb[0] = sx > 0 ? s - 1 : -1;
        b.nb[1] = sx < sdx - 1 ? s + 1 : -1;
        b.nb[2] = sy > 0 ? s - sdx : -1;
        b.nb[3] = sy < sdy - 1 ? s + sdx : -1;
        b.nb[4] = sz > 0 ? s - sdx * sdy : -1;
        b.nb[5] = sz < sdz - 1 ? s + sdx * sdy : -1;
        b.nplanes = b.nx + b.ny + b.nz - 2;
        b.lid = s;
        b.box = sx / c->sd_box[0] + c->pv[0] * (sy / c->sd_box[1] + c->pv[1] * (sz / c->sd_box[2]));
        b.dof = 0;
        b.neig = 0;
        const auto key = std::make_tuple(b.nx, b.ny, b.nz);
        auto it = shapes.find(key);
        if (it == shapes.end()) {
          const int id = (int)shapes.size();
          shapes[key] = id;
          c->h_shape_pstart.push_back((int)c->h_plane_start.size());
          c->h_shape_coff.push_back((int)c->h_plane_cell.size());
          int cnt = 0;
          for (int p = 0; p < b.nplanes; ++p) {
            c->h_plane_start.push_back(cnt);
            const int cnt0 = cnt;
            for (int lz = 0; lz < b.nz; ++lz)
              for (int ly = 0; ly < b.ny; ++ly) {
                const int lx = p - ly - lz;
                if (lx < 0 || lx >= b.nx) continue;
                c->h_plane_cell.push_back(lx | (ly << 10) | (lz << 20));
                ++cnt;
              }
            c->max_plane = std::max(c->max_plane, cnt - cnt0);
          }
          c->h_plane_start.push_back(cnt);
          b.shape = id;
        } else {
          b.shape = it->second;
        }
        b.plane_off = c->h_shape_coff[b.shape];
        off += b.nx * b.ny * b.nz;
      }
  c->nshape = (int)shapes.size();
  for (const Brick& b : c->h_bricks)
    if (b.nx > kBlock || b.ny > 1023 || b.nz > 1023 || b.nx > 1023) {
      set_error("subdomain of %dx%dx%d cells: x extent above %d is not supported", b.nx, b.ny, b.nz,
                kBlock);
      return PMX_ERR_SUP;
    }
  c->h_nat2brick.assign(c->ncell, 0);
  for (int z = 0; z < P; ++z)
    for (int y = 0; y < N; ++y)
      for (int x = 0; x < M; ++x) {
        const int s = c->h_ax_sub[0][x] + sdx * (c->h_ax_sub[1][y] + sdy * c->h_ax_sub[2][z]);
        const Brick& b = c->h_bricks[s];
        c->h_nat2brick[(size_t)(z * N + y) * M + x] =
            b.off + (c->h_ax_loc[2][z] * b.ny + c->h_ax_loc[1][y]) * b.nx + c->h_ax_loc[0][x];
      }
  PMX_TRY(upload(c->d_bricks, c->h_bricks));
  {
    std::vector<BrickView> views(c->nsub);
    for (int s = 0; s < c->nsub; ++s) {
      const Brick& b = c->h_bricks[s];
      BrickView& v = views[s];
      v.off = b.off;
      v.nx = b.nx;
      v.ny = b.ny;
      v.nz = b.nz;
      v.nxny = b.nx * b.ny;
      v.n = v.nxny * b.nz;
      for (int d = 0; d < 6; ++d) {
        const int sn = b.nb[d];
        const Brick* o = sn >= 0 ? &c->h_bricks[sn] : nullptr;
        v.noff[d] = o ? o->off : -1;
        const int ax = d >> 1;
        v.next[d] = o ? (ax == 0 ? o->nx : (ax == 1 ? o->ny : o->nz)) : 0;
      }
    }
    PMX_TRY(upload(c->d_views, views));
    // XCD-aware workgroup -> brick order: Morton order of the brick grid,
    // cut into 8 contiguous runs, run x served by workgroups b = x (mod 8).
    std::vector<std::pair<uint64_t, int>> key(c->nsub);
    for (int s = 0; s < c->nsub; ++s) {
      const int bx = s % c->sd[0], by = (s / c->sd[0]) % c->sd[1], bz = s / (c->sd[0] * c->sd[1]);
      uint64_t m = 0;
      for (int bit = 0; bit < 10; ++bit)
        m |= (uint64_t)((bx >> bit) & 1) << (3 * bit) | (uint64_t)((by >> bit) & 1) << (3 * bit + 1) |
             (uint64_t)((bz >> bit) & 1) << (3 * bit + 2);
      key[s] = {m, s};
    }
    std::sort(key.begin(), key.end());
    constexpr int kXcd = 8;
    std::vector<int> start(kXcd + 1, 0), order(c->nsub);
    for (int x = 0; x < kXcd; ++x) start[x + 1] = start[x] + (c->nsub - x + kXcd - 1) / kXcd;
    for (int b = 0; b < c->nsub; ++b) order[b] = key[start[b % kXcd] + b / kXcd].second;
    PMX_TRY(upload(c->d_order, order));
  }
  for (int a = 0; a < 3; ++a) {
    PMX_TRY(upload(c->d_ax[a], c->h_ax_sub[a]));
    PMX_TRY(upload(c->d_ax[3 + a], c->h_ax_loc[a]));
  }
  PMX_TRY(upload(c->d_plane_start, c->h_plane_start));
  PMX_TRY(upload(c->d_plane_cell, c->h_plane_cell));
  PMX_TRY(upload(c->d_shape_pstart, c->h_shape_pstart));
  PMX_TRY(upload(c->d_shape_coff, c->h_shape_coff));
  Geo& g = c->geo;
  g.M = M;
  g.N = N;
  g.P = P;
  g.sdx = sdx;
  g.sdy = sdy;
  g.sdz = sdz;
  g.nsub = c->nsub;
  g.ncell = c->ncell;
  g.bricks = c->d_bricks.p;
  g.views = c->d_views.p;
  g.order = c->d_order.p;
  for (int a = 0; a < 3; ++a) {
    g.ax_sub[a] = c->d_ax[a].p;
    g.ax_loc[a] = c->d_ax[3 + a].p;
  }
  g.plane_start = c->d_plane_start.p;
  g.plane_cell = c->d_plane_cell.p;
  g.shape_pstart = c->d_shape_pstart.p;
  g.shape_coff = c->d_shape_coff.p;
  g.pin = -1;
  g.tau = 0.0;
  return PMX_SUCCESS;
}

static int alloc_work(pmx_ctx* c) {
  const size_t n = c->ncell;
  for (DBuf<double>* b : {&c->w_x, &c->w_y, &c->w_yb, &c->w_r, &c->w_b, &c->w_p, &c->w_z, &c->w_q,
                          &c->w_tmp, &c->nat_buf})
    PMX_TRY(b->alloc(n));
  if (c->pc_precision == 32) {
    PMX_TRY(c->w_y32.alloc(n));
    PMX_TRY(c->w_y32b.alloc(n));
  }
  PMX_TRY(c->partial.alloc(std::max<size_t>(6 * (size_t)c->nsub, 4096)));
  PMX_TRY(c->scal.alloc(16));
  PMX_HIP(hipMemset(c->scal.p, 0, 16 * sizeof(double)));
  c->cg_part = c->partial.p;
  c->cg_scal = c->scal.p;
  return PMX_SUCCESS;
}

// Level-2 and level-3 setup from the basis rows in ctx->phi.
static int setup_coarse(pmx_ctx* c) {
  double t0 = now();
  // Eigen counts against eb_lv1 (src/PreMixFEM_3D.c:368-384) and coarse dofs.
  c->h_eigen_num.assign(c->nsub, c->en_lv1);
  c->h_eigen_min_lv1.assign(c->nsub, 0.0);
  c->h_eigen_max_lv1.assign(c->nsub, 0.0);
  c->h_dof.assign(c->nsub + 1, 0);
  for (int s = 0; s < c->nsub; ++s) {
    int num = c->en_lv1;
    if (!c->h_eigen_num_given.empty()) {
      num = std::max(1, std::min(c->en_lv1, c->h_eigen_num_given[s]));
    } else {
      for (int j = 0; j < c->en_lv1; ++j)
        if (c->h_eig_lv1[(size_t)s * c->en_lv1 + j] >= c->eb_lv1) {
          num = j + 1;
          break;
        }
    }
    c->h_eigen_num[s] = num;
    c->h_eigen_min_lv1[s] = c->h_eig_lv1[(size_t)s * c->en_lv1];
    c->h_eigen_max_lv1[s] = c->h_eig_lv1[(size_t)s * c->en_lv1 + num - 1];
    c->h_dof[s + 1] = c->h_dof[s] + num;
    c->h_bricks[s].dof = c->h_dof[s];
    c->h_bricks[s].neig = num;
  }
  c->nc = c->h_dof[c->nsub];
  PMX_HIP(hipMemcpy(c->d_bricks.p, c->h_bricks.data(), c->h_bricks.size() * sizeof(Brick),
                    hipMemcpyHostToDevice));
  // Galerkin coarse matrix (block rows; dense too unless -lv2_icc) and its
  // level-2 factor: the dense inverse, or the block ICC(0) of the spe10 fork.
  PMX_TRY(c->Acb.alloc((size_t)c->nsub * 7 * c->en_lv1 * c->en_lv1));
  for (DBuf<double>* b : {&c->c_a, &c->c_b, &c->c_c, &c->c_d, &c->c_e}) PMX_TRY(b->alloc(c->nc));
  c->use_band = !c->lv2_icc && !c->use_2level &&
                 (c->coarse_band == 1 || (c->coarse_band < 0 && c->nc > kBandAutoNc));
  if (c->nbox > 1 && c->use_band && !c->coarse_planes) {
    set_error("-procs_virtual: the band Cholesky coarse factor is not built for process boxes");
    return PMX_ERR_SUP;
  }
  // With process boxes A_c keeps only the couplings inside a box, and every
  // box block is nonsingular through its faces to the other boxes: no
  // near-null direction to deflate.  (The 2-level variant's A_c keeps every
  // coupling, as on one process.)
  if (c->nbox > 1 && !c->use_2level) c->coarse_deflate = 0;
  if (c->lv2_icc || c->use_band) {
    for (DBuf<double>* b : {&c->Ac, &c->Ainv, &c->Fop}) b->release();
    PMX_TRY(launch_galerkin(c));
    c->use_planes = c->use_band && c->coarse_planes;
    if (c->use_planes) {
      c->band.release();
      c->band_dinv.release();
    } else {
      c->pl_G.release();
    }
    PMX_TRY(c->lv2_icc ? bicc_setup(c) : (c->use_planes ? planes_setup(c) : band_setup(c)));
  } else {
    PMX_TRY(c->Ac.alloc((size_t)c->nc * c->nc));
    PMX_TRY(c->Ainv.alloc((size_t)c->nc * c->nc));
    PMX_TRY(launch_galerkin(c));
    const double shift = (c->use_2level && !c->no_shift_A_cc) ? kNil : 0.0;
    PMX_TRY(coarse_solve_setup(c, shift));
  }
  // The coarse stages of PC_apply_vec (src/PreMixFEM_3D.c:1857-1929):
  // 2 = level 2, 3 = level 3.
  std::vector<int>& prog = c->coarse_prog;
  prog.clear();
  if (c->use_2level) {
    prog.push_back(2);
  } else if (c->use_W_cycle) {
    for (int r = 0; r < 2; ++r) {
      for (int i = 0; i < c->si_lv2; ++i) prog.push_back(2);
      prog.push_back(3);
    }
    for (int i = 0; i < c->si_lv2; ++i) prog.push_back(2);
  } else {
    for (int i = 0; i < c->si_lv2; ++i) prog.push_back(2);
    prog.push_back(3);
    for (int i = 0; i < c->si_lv2; ++i) prog.push_back(2);
  }
  // One process, exact level-2 factor (dense inverse or band Cholesky of the
  // full Galerkin matrix A_c = Phi^T A Phi): after the first level-2 stage
  // the coarse residual Phi^T (x - A y) = t - A_c A_c^{-1} t is zero, so the
  // level-3 stage and every later level-2 stage of the V- or W-cycle
  // (src/PreMixFEM_3D.c:1857-1929) add exactly zero.  (Unpinned, A_c is
  // singular with the coarse constant as null vector; the pseudo-inverse
  // leaves t's component along it, which is 1^T r = 0 for the consistent
  // residuals of a zero-sum right-hand side.)  Computed in floating point,
  // those stages add only rounding, amplified by up to 1/lambda_min(A_c)
  // (~1e11 with one pinned cell), so the program is cut to the first stage.
  // Not with process boxes (A_c keeps only the couplings inside a box) nor
  // with the inexact ICC(0) level-2 factor.
  c->coarse_collapsed = 0;
  if (c->coarse_collapse && c->nbox == 1 && !c->lv2_icc && !c->use_2level && prog.size() > 1) {
    prog.assign(1, 2);
    c->coarse_collapsed = 1;
  }
  if (c->pc_precision == 32) {
    const size_t nphi = (size_t)c->ncell * c->en_lv1;
    PMX_TRY(c->phi32.alloc(nphi));
    PMX_TRY(launch_to_f32(c, c->phi.p, c->phi32.p, nphi));
  }
  PMX_HIP(hipStreamSynchronize(c->stream));
  double t1 = now();
  c->t_stages[PMX_STAGE_SU_LV2_2] += t1 - t0;
  if (!c->use_2level) {
    PMX_TRY(c->nbox > 1 ? level3_setup_boxes(c)
                        : ((c->lv2_icc || c->use_band) ? level3_setup_icc(c) : level3_setup(c)));
    double t2 = now();
    c->t_stages[PMX_STAGE_SU_LV3_1] += t2 - t1;
  }
  const double t3 = now();
  // The dense path composes the coarse stages into one nc x nc operator; the
  // ICC(0) path runs them per apply in one workgroup (no nc^2 storage).
  // A single stage needs no composition: the GEMV with the deflated inverse
  // and its separate rank-2 correction is the same operator, and keeps the
  // large near-null term (~1/lambda_min) out of the dense entries, where its
  // rounding would reach every component of the correction.
  c->use_fop = !c->lv2_icc && !c->use_band && c->coarse_composite && c->coarse_prog.size() > 1;
  if (c->use_fop) PMX_TRY(coarse_operator_setup(c, c->coarse_prog));
  else c->Fop.release();
  c->t_stages[PMX_STAGE_SU_LV3_2] += now() - t3;
  return PMX_SUCCESS;
}

// Operator and level 1.
static int setup_fine(pmx_ctx* c) {
  if (!c->have_kappa) {
    set_error("kappa has not been set (pmx_set_kappa)");
    return PMX_ERR_ORDER;
  }
  double t0 = now();
  PMX_TRY(launch_trans(c));
  c->geo.wd = nullptr;
  c->geo.wd32 = nullptr;
  if (c->have_wd) {
    c->geo.wd = c->wd.p;
    if (c->pc_precision == 32) {
      PMX_TRY(c->wd32.alloc(c->ncell));
      PMX_TRY(launch_to_f32(c, c->wd.p, c->wd32.p, c->ncell));
      c->geo.wd32 = c->wd32.p;
    }
  }
  if (c->pin[0] >= 0) {
    const int x = c->pin[0], y = c->pin[1], z = c->pin[2];
    const int bi = c->h_nat2brick[(size_t)(z * c->mesh[1] + y) * c->mesh[0] + x];
    double tau = 0.0;
    const int nat = (z * c->mesh[1] + y) * c->mesh[0] + x;
    const int nbh[3] = {x + 1 < c->mesh[0] ? nat + 1 : -1,
                        y + 1 < c->mesh[1] ? nat + c->mesh[0] : -1,
                        z + 1 < c->mesh[2] ? nat + c->mesh[0] * c->mesh[1] : -1};
    for (int d = 0; d < 3; ++d) {
      double v = 0.0;
      PMX_HIP(hipMemcpy(&v, c->T[d].p + bi, sizeof(double), hipMemcpyDeviceToHost));
      tau = std::max(tau, v);
      if (nbh[d] >= 0) {
        PMX_HIP(hipMemcpy(&v, c->T[d].p + c->h_nat2brick[nbh[d]], sizeof(double),
                          hipMemcpyDeviceToHost));
        tau = std::max(tau, v);
      }
    }
    c->tau = tau;
    c->pin_brick = bi;
    c->geo.pin = bi;
    c->geo.tau = tau;
  } else {
    c->geo.pin = -1;
    c->geo.tau = 0.0;
  }
  // The FP64 factor serves the FP64 preconditioner and the exact level-1
  // blocks; the FP32 preconditioner sweeps its own FP32 factor only.
  if (c->pc_precision != 32 || c->use_full_Cholesky_lv1) {
    PMX_TRY(c->fac_f.alloc((size_t)c->ncell * 4));
    PMX_TRY(c->fac_b.alloc((size_t)c->ncell * 4));
    PMX_TRY(launch_dic_factor(c, 0, 0.0, c->fac_f.p, c->fac_b.p, false));
  }
  c->geo.l1L = c->geo.l1U = nullptr;
  c->geo.l1off = nullptr;
  if (c->use_full_Cholesky_lv1) PMX_TRY(chol_lv1_setup(c));
  if (c->pc_precision == 32) {
    PMX_TRY(c->fac32_f.alloc((size_t)c->ncell * 4));
    PMX_TRY(c->fac32_b.alloc((size_t)c->ncell * 4));
    PMX_TRY(launch_dic_factor(c, 0, 0.0, c->fac32_f.p, c->fac32_b.p, true));
  }
  PMX_TRY(c->mass.alloc(c->ncell));
  PMX_TRY(launch_mass(c));
  PMX_TRY(c->phi.alloc((size_t)c->ncell * c->en_lv1));
  PMX_HIP(hipStreamSynchronize(c->stream));
  c->t_stages[PMX_STAGE_SU_LV1] += now() - t0;
  return PMX_SUCCESS;
}

// Per-kernel timing with HIP events on the context stream (-kernel_timing).
static int kt_begin(pmx_ctx* c, int id) {
  if (!c->kernel_timing) return -1;
  const int slot = (int)c->kt_pending.size();
  while ((int)c->kt_events.size() < 2 * (slot + 1)) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return -1;
    c->kt_events.push_back(e);
  }
  (void)hipEventRecord(c->kt_events[2 * slot], c->stream);
  c->kt_pending.push_back(id);
  return slot;
}

static void kt_end(pmx_ctx* c, int slot) {
  if (slot < 0) return;
  (void)hipEventRecord(c->kt_events[2 * slot + 1], c->stream);
}

static int kt_collect(pmx_ctx* c) {
  for (size_t i = 0; i < c->kt_pending.size(); ++i) {
    float ms = 0.0f;
    PMX_HIP(hipEventElapsedTime(&ms, c->kt_events[2 * i], c->kt_events[2 * i + 1]));
    c->kt_ms[c->kt_pending[i]] += ms;
    c->kt_count[c->kt_pending[i]] += 1;
  }
  c->kt_pending.clear();
  return PMX_SUCCESS;
}

// V-cycle of PC_apply_vec (src/PreMixFEM_3D.c:1840-1936) on brick-order
// device vectors.  The fine-grid stages are level 1 (before and after); all
// level-2/level-3 stages in between act on the coarse residual
// t = Phi^T (x - A y): a stage with correction d updates t -= A_c d (since
// A_c = Phi^T A Phi) and accumulates u += d, and y += Phi u is applied once.
// z receives the output; partial (if not null) the per-brick z.x sums.
// cg_partial: the last kernel writes the z'r partials of the PCG.  cg_upd:
// the first level-1 kernel also performs the x/r update of the current PCG
// iteration (x is then the PCG residual, xs/p/w the CG vectors) and the
// prolongation advances the device iteration counter.
static int apply_device(pmx_ctx* c, double* x, double* z, bool cg_partial, bool cg_upd = false,
                        double* xs = nullptr, const double* p = nullptr,
                        const double* w = nullptr) {
  const bool f32 = c->pc_precision == 32;
  auto tick = [&](int stage, double& t) -> int {
    if (!c->log_stages) return PMX_SUCCESS;
    PMX_HIP(hipStreamSynchronize(c->stream));
    const double tn = now();
    c->t_stages[stage] += tn - t;
    t = tn;
    return PMX_SUCCESS;
  };
  double t = now();
  const double t_start = t;
  // Pre-smoothing: y = B1 x, then si_lv1-1 residual-corrected sweeps.
  int ev = kt_begin(c, 0);
  PMX_TRY(launch_lv1_first(c, x, cg_upd, xs, p, w));
  kt_end(c, ev);
  for (int i = 1; i < c->si_lv1; ++i) {
    PMX_TRY(launch_lv1_last(c, x, nullptr, true, false));
  }
  PMX_TRY(tick(PMX_STAGE_AV_LV1, t));
  ev = kt_begin(c, 2);
  PMX_TRY(launch_resid_restrict(c, x, false, c->c_a.p));
  kt_end(c, ev);
  double* tv = c->c_a.p;
  double* u = c->c_c.p;
  double* d = c->c_b.p;
  if (!c->use_fop) {
    PMX_TRY(launch_coarse_program(c, c->coarse_prog, tv, u, c->c_e.p, d, 6, 7, kt_begin, kt_end));
  } else {
    ev = kt_begin(c, 6);
    PMX_TRY(launch_gemv(c, c->Fop.p, tv, u, c->nc));
    kt_end(c, ev);
  }
  PMX_TRY(tick(PMX_STAGE_AV_LV2, t));
  ev = kt_begin(c, 3);
  PMX_TRY(launch_prolong(c, u, cg_upd));
  kt_end(c, ev);
  // Post-smoothing.
  for (int i = 1; i < c->si_lv1; ++i) PMX_TRY(launch_lv1_last(c, x, nullptr, true, false));
  ev = kt_begin(c, 1);
  PMX_TRY(launch_lv1_last(c, x, z, false, cg_partial));
  kt_end(c, ev);
  PMX_TRY(tick(PMX_STAGE_AV_LV1, t));
  (void)f32;
  if (c->log_stages) c->t_stages[PMX_STAGE_AV] += now() - t_start;
  return PMX_SUCCESS;
}

}  // namespace pmx

using namespace pmx;

extern "C" {

const char* pmx_last_error(void) { return g_err.c_str(); }

int pmx_init(pmx_ctx** out, const double* dom, const int* mesh, const char* options) {
  if (!out || !dom || !mesh) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  *out = nullptr;
  if (!(dom[0] > 0.0 && dom[1] > 0.0 && dom[2] > 0.0)) {
    set_error("Errors in dom(L, W, H)=[%.5f, %.5f, %.5f].", dom[0], dom[1], dom[2]);
    return PMX_ERR_ARG_WRONG;
  }
  if (!(mesh[0] > 0 && mesh[1] > 0 && mesh[2] > 0)) {
    set_error("Errors in mesh(M, N, P)=[%d, %d, %d].", mesh[0], mesh[1], mesh[2]);
    return PMX_ERR_ARG_WRONG;
  }
  pmx_ctx* c = new pmx_ctx();
  for (int a = 0; a < 3; ++a) {
    c->dom[a] = dom[a];
    c->mesh[a] = mesh[a];
    c->sd[a] = 2;
  }
  int rc = parse_options(c, options);
  auto fail = [&](int code) {
    delete c;
    return code;
  };
  if (rc) return fail(rc);
  for (int a = 0; a < 3; ++a) {
    if (c->sd[a] < 1) {
      set_error("Error in sub_domains=%d.", c->sd[a]);
      return fail(PMX_ERR_ARG_WRONG);
    }
    if (c->sd[a] == 1) c->sd[a] = 2;  // src/PreMixFEM_3D.c:1718-1721
    if (c->pv[a] < 1 || c->pv[a] > c->mesh[a]) {
      set_error("-procs_virtual: %d process boxes on the %d cells of axis %d", c->pv[a],
                c->mesh[a], a);
      return fail(PMX_ERR_ARG_WRONG);
    }
    // sub_domains counts per process box; the smallest box has n/p cells.
    if (c->sd[a] > c->mesh[a] / c->pv[a]) {
      set_error("sub_domains=%d exceeds the %d cells of axis %d", c->sd[a],
                c->mesh[a] / c->pv[a], a);
      return fail(PMX_ERR_ARG_WRONG);
    }
    c->sd_box[a] = c->sd[a];
    c->sd[a] *= c->pv[a];
  }
  c->nbox = c->pv[0] * c->pv[1] * c->pv[2];
  if (c->nbox > 1 && c->coarse_band == 1 && !c->coarse_planes) {
    set_error("-procs_virtual: the band Cholesky coarse factor (-coarse_planes 0) is not built "
              "for process boxes; the plane factor is");
    return fail(PMX_ERR_SUP);
  }
  if (c->en_lv1 < 1) {
    set_error("Error in max_eigen_num_lv1=%d for the level-1 problem.", c->en_lv1);
    return fail(PMX_ERR_ARG_WRONG);
  }
  if (c->en_lv2 < 1) {
    set_error("Error in max_eigen_num_lv2=%d for the level-2 problem.", c->en_lv2);
    return fail(PMX_ERR_ARG_WRONG);
  }
  if (c->en_lv1 > kMaxEig || c->en_lv2 > kMaxEig) {
    set_error("eigen counts above %d are not supported", kMaxEig);
    return fail(PMX_ERR_SUP);
  }
  if (c->eg_op != PMX_EIG_OP_MOD && c->eg_op != PMX_EIG_OP_MEAN && c->eg_op != PMX_EIG_OP_DIAG) {
    set_error("Error in lv2_eigen_op=%d.", c->eg_op);
    return fail(PMX_ERR_ARG_WRONG);
  }
  if (c->pc_precision != 32 && c->pc_precision != 64) {
    set_error("-pc_precision must be 32 or 64");
    return fail(PMX_ERR_ARG_WRONG);
  }
  if (c->use_full_Cholesky_lv1 && c->pc_precision != 64) {
    set_error("-use_full_Cholesky_lv1 needs -pc_precision 64");
    return fail(PMX_ERR_SUP);
  }
  c->si_lv1 = std::max(1, c->si_lv1);
  c->si_lv2 = std::max(1, c->si_lv2);
  if (c->eb_lv1 < 0.0) c->eb_lv1 = 1.0 / kNil;
  if (c->eb_lv2 < 0.0) c->eb_lv2 = 1.0 / kNil;
  if (c->pin[0] >= 0 &&
      (c->pin[0] >= mesh[0] || c->pin[1] < 0 || c->pin[1] >= mesh[1] || c->pin[2] < 0 ||
       c->pin[2] >= mesh[2])) {
    set_error("-pin cell outside the mesh");
    return fail(PMX_ERR_ARG_OUTOFRANGE);
  }
  c->Hx = dom[0] / (double)mesh[0];
  c->Hy = dom[1] / (double)mesh[1];
  c->Hz = dom[2] / (double)mesh[2];
  c->meas_elem = c->Hx * c->Hy * c->Hz;
  const double fyz = c->Hy * c->Hz, fzx = c->Hz * c->Hx, fxy = c->Hx * c->Hy;
  c->coef[0] = fyz * fyz / c->meas_elem;
  c->coef[1] = fzx * fzx / c->meas_elem;
  c->coef[2] = fxy * fxy / c->meas_elem;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    set_error("no HIP device available");
    return fail(PMX_ERR_LIB);
  }
  if (hipDeviceGetAttribute(&c->lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
      hipSuccess) {
    set_error("hipDeviceGetAttribute failed");
    return fail(PMX_ERR_LIB);
  }
  if (hipStreamCreate(&c->stream) != hipSuccess) {
    set_error("hipStreamCreate failed");
    return fail(PMX_ERR_LIB);
  }
  rc = scratch_address(dev, &c->scratch_dev);
  if (!rc) rc = build_geometry(c);
  if (!rc) {
    // The level-1 kernels keep a brick in LDS (the factorisation in FP64):
    // bound the subdomain size by the device's LDS.
    int mx = 0;
    for (const Brick& b : c->h_bricks) mx = std::max(mx, b.nx * b.ny * b.nz);
    const size_t need = (size_t)mx * sizeof(double);
    if (need > (size_t)c->lds_max) {
      set_error("subdomains of %d cells need %zu bytes of LDS per workgroup (device: %d); use "
                "more subdomains (-sd)", mx, need, c->lds_max);
      rc = PMX_ERR_SUP;
    }
  }
  if (!rc) rc = alloc_work(c);
  if (rc) {
    pmx_final(c);
    return rc;
  }
  *out = c;
  return PMX_SUCCESS;
}

int pmx_print_info(const pmx_ctx* c) {
  if (!c) return PMX_ERR_ARG_NULL;
  printf("DIM=%d, M=%d, N=%d, P=%d.\n", 3, c->mesh[0], c->mesh[1], c->mesh[2]);
  printf("GPU subdomain bricks: X=%d, Y=%d, Z=%d.\n", c->sd[0], c->sd[1], c->sd[2]);
  printf("smoothing_iters_lv1=%d, smoothing_iters_lv2=%d, sub_domains=%d, lv2_eigen_op=%d\n",
         c->si_lv1, c->si_lv2, c->sd[0], c->eg_op);
  printf("max_eigen_num_lv1=%d, eigen_bd_lv1=%.5E\n", c->en_lv1, c->eb_lv1);
  printf("max_eigen_num_lv2=%d, eigen_bd_lv2=%.5E\n", c->en_lv2, c->eb_lv2);
  if (c->use_2level) printf("Only use the 2-level version!\n");
  printf("pc_precision=%d\n", c->pc_precision);
  return PMX_SUCCESS;
}

// Pushes natural-order permeability (host or device arrays) into brick order.
static int set_kappa_impl(pmx_ctx* c, const double* const src[3], bool device) {
  for (int a = 0; a < 3; ++a) {
    if (!device) {
      for (int i = 0; i < c->ncell; ++i)
        if (!(src[a][i] > 0.0) || !std::isfinite(src[a][i])) {
          set_error("kappa[%d][%d]=%g is not positive and finite", a, i, src[a][i]);
          return PMX_ERR_ARG_WRONG;
        }
    } else {
      int bad = -1;
      PMX_TRY(launch_check_field(c, src[a], c->ncell, true, &bad));
      if (bad >= 0) {
        set_error("kappa[%d][%d] is not positive and finite", a, bad);
        return PMX_ERR_ARG_WRONG;
      }
    }
    PMX_TRY(c->kappa[a].alloc(c->ncell));
    const double* nat = src[a];
    if (!device) {
      PMX_HIP(hipMemcpyAsync(c->nat_buf.p, src[a], c->ncell * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
      nat = c->nat_buf.p;
    }
    PMX_TRY(launch_nat2brick(c, nat, c->kappa[a].p, 1, c->ncell, c->ncell));
    PMX_TRY(c->T[a].alloc(c->ncell));
    if (c->pc_precision == 32) PMX_TRY(c->T32[a].alloc(c->ncell));
  }
  PMX_HIP(hipStreamSynchronize(c->stream));
  c->have_kappa = 1;
  c->setup_done = 0;
  return PMX_SUCCESS;
}

int pmx_set_kappa(pmx_ctx* c, const double* kx, const double* ky, const double* kz) {
  if (!c || !kx || !ky || !kz) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  const double* src[3] = {kx, ky, kz};
  return set_kappa_impl(c, src, false);
}

int pmx_set_kappa_device(pmx_ctx* c, const double* d_kx, const double* d_ky,
                         const double* d_kz) {
  if (!c || !d_kx || !d_ky || !d_kz) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  const double* src[3] = {d_kx, d_ky, d_kz};
  return set_kappa_impl(c, src, true);
}

// The per-cell diagonal (well) term; null clears it.
static int set_well_impl(pmx_ctx* c, const double* w, bool device) {
  c->setup_done = 0;
  if (!w) {
    c->have_wd = 0;
    c->wd.release();
    c->wd32.release();
    c->geo.wd = nullptr;
    c->geo.wd32 = nullptr;
    return PMX_SUCCESS;
  }
  if (!device) {
    for (int i = 0; i < c->ncell; ++i)
      if (!(w[i] >= 0.0) || !std::isfinite(w[i])) {
        set_error("well term [%d]=%g is not non-negative and finite", i, w[i]);
        return PMX_ERR_ARG_WRONG;
      }
  } else {
    int bad = -1;
    PMX_TRY(launch_check_field(c, w, c->ncell, false, &bad));
    if (bad >= 0) {
      set_error("well term [%d] is not non-negative and finite", bad);
      return PMX_ERR_ARG_WRONG;
    }
  }
  PMX_TRY(c->wd.alloc(c->ncell));
  const double* nat = w;
  if (!device) {
    PMX_HIP(hipMemcpyAsync(c->nat_buf.p, w, c->ncell * sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
    nat = c->nat_buf.p;
  }
  PMX_TRY(launch_nat2brick(c, nat, c->wd.p, 1, c->ncell, c->ncell));
  PMX_HIP(hipStreamSynchronize(c->stream));
  c->have_wd = 1;
  c->geo.wd = c->wd.p;
  c->geo.wd32 = nullptr;
  return PMX_SUCCESS;
}

int pmx_set_well_term(pmx_ctx* c, const double* wd) {
  if (!c) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  return set_well_impl(c, wd, false);
}

int pmx_set_well_term_device(pmx_ctx* c, const double* d_wd) {
  if (!c) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  return set_well_impl(c, d_wd, true);
}

// Face fluxes from a pressure field (natural order, device arrays).
static int velocity_impl(pmx_ctx* c, const double* d_p, double* d_fx, double* d_fy,
                         double* d_fz) {
  if (!c->have_kappa) {
    set_error("kappa has not been set (pmx_set_kappa)");
    return PMX_ERR_ORDER;
  }
  if (!c->setup_done) PMX_TRY(launch_trans(c));
  PMX_TRY(launch_nat2brick(c, d_p, c->w_tmp.p, 1, c->ncell, c->ncell));
  PMX_TRY(launch_velocity(c, c->w_tmp.p, d_fx, d_fy, d_fz));
  PMX_HIP(hipStreamSynchronize(c->stream));
  return PMX_SUCCESS;
}

int pmx_velocity(pmx_ctx* c, const double* p, double* fx, double* fy, double* fz) {
  if (!c || !p || !fx || !fy || !fz) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  const size_t nb = c->ncell * sizeof(double);
  DBuf<double> f;
  PMX_TRY(f.alloc(3 * (size_t)c->ncell));
  PMX_HIP(hipMemcpyAsync(c->nat_buf.p, p, nb, hipMemcpyHostToDevice, c->stream));
  PMX_TRY(velocity_impl(c, c->nat_buf.p, f.p, f.p + c->ncell, f.p + 2 * (size_t)c->ncell));
  double* out[3] = {fx, fy, fz};
  for (int a = 0; a < 3; ++a)
    PMX_HIP(hipMemcpy(out[a], f.p + (size_t)a * c->ncell, nb, hipMemcpyDeviceToHost));
  return PMX_SUCCESS;
}

int pmx_velocity_device(pmx_ctx* c, const double* d_p, double* d_fx, double* d_fy, double* d_fz) {
  if (!c || !d_p || !d_fx || !d_fy || !d_fz) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  return velocity_impl(c, d_p, d_fx, d_fy, d_fz);
}

int pmx_create_A(pmx_ctx* c, double* diag, double* off) {
  if (!c || !diag || !off) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  if (!c->have_kappa) {
    set_error("kappa has not been set (pmx_set_kappa)");
    return PMX_ERR_ORDER;
  }
  if (!c->setup_done) PMX_TRY(launch_trans(c));
  const size_t n = c->ncell;
  std::vector<double> T[3];
  for (int a = 0; a < 3; ++a) {
    T[a].resize(n);
    PMX_HIP(hipMemcpyAsync(T[a].data(), c->T[a].p, n * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  }
  std::vector<double> wdb;
  if (c->have_wd) {
    wdb.resize(n);
    PMX_HIP(hipMemcpyAsync(wdb.data(), c->wd.p, n * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  }
  PMX_HIP(hipStreamSynchronize(c->stream));
  const int M = c->mesh[0], N = c->mesh[1], P = c->mesh[2];
  const long st[3] = {1, M, (long)M * N};
  // Accumulation order of PC_create_A's face loop for one cell: its lower
  // faces (x, y, z) from its own iteration, then its upper faces (x, y, z).
  for (int z = 0; z < P; ++z)
    for (int y = 0; y < N; ++y)
      for (int x = 0; x < M; ++x) {
        const size_t cn = (size_t)(z * N + y) * M + x;
        const int co[3] = {x, y, z};
        const int ext[3] = {M, N, P};
        double lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
          lo[a] = co[a] > 0 ? T[a][c->h_nat2brick[cn]] : 0.0;
          hi[a] = co[a] < ext[a] - 1 ? T[a][c->h_nat2brick[cn + st[a]]] : 0.0;
          off[6 * cn + 2 * a] = -lo[a];
          off[6 * cn + 2 * a + 1] = -hi[a];
        }
        double dsum = lo[0];
        dsum += lo[1];
        dsum += lo[2];
        dsum += hi[0];
        dsum += hi[1];
        dsum += hi[2];
        if (c->pin[0] == x && c->pin[1] == y && c->pin[2] == z) {
          double tau = 0.0;
          for (int a = 0; a < 3; ++a) tau = std::max(tau, std::max(lo[a], hi[a]));
          dsum += tau;
        }
        if (c->have_wd) dsum += wdb[c->h_nat2brick[cn]];
        diag[cn] = dsum;
      }
  return PMX_SUCCESS;
}

static int setup_common(pmx_ctx* c, const double* phi_nat, const int* eigen_num,
                        const double* eigvals) {
  std::fill(c->t_stages, c->t_stages + PMX_MAX_LOG_STAGES, 0.0);
  ++c->setup_gen;  // buffers may move: a captured PCG chunk is stale
  const double t0 = now();
  PMX_TRY(setup_fine(c));
  const double t1 = now();
  if (phi_nat) {
    // Caller basis (natural-order columns) -> brick rows.
    PMX_HIP(hipMemsetAsync(c->phi.p, 0, (size_t)c->ncell * c->en_lv1 * sizeof(double), c->stream));
    DBuf<double> tmp;
    PMX_TRY(tmp.alloc((size_t)c->ncell * c->en_lv1));
    PMX_HIP(hipMemcpyAsync(tmp.p, phi_nat, (size_t)c->ncell * c->en_lv1 * sizeof(double),
                           hipMemcpyHostToDevice, c->stream));
    PMX_TRY(launch_basis_nat2rows(c, tmp.p, c->phi.p, c->en_lv1));
    PMX_HIP(hipStreamSynchronize(c->stream));
    c->h_eig_lv1.assign((size_t)c->nsub * c->en_lv1, 0.0);
    if (eigvals) memcpy(c->h_eig_lv1.data(), eigvals, c->h_eig_lv1.size() * sizeof(double));
    if (eigen_num)
      c->h_eigen_num_given.assign(eigen_num, eigen_num + c->nsub);
    else
      c->h_eigen_num_given.assign(c->nsub, c->en_lv1);
    c->eig_iters_max = c->eig_iters_sum = 0;
  } else {
    c->h_eigen_num_given.clear();
    c->t_tune = 0.0;
    if (c->eig_autotune) {
      const double tt = now();
      PMX_TRY(lobpcg_autotune(c));
      c->t_tune = now() - tt;
    }
    if (c->warm_start == 1) {
      // learned warm start: the network's block seeds LOBPCG
      const int mb = lobpcg_block_size(c->en_lv1, c->eig_guard);
      if (mb == 0) {
        set_error("en_lv1=%d exceeds the LOBPCG block limit", c->en_lv1);
        return PMX_ERR_SUP;
      }
      DBuf<double> x0;
      PMX_TRY(x0.alloc((size_t)c->ncell * mb));
      PMX_TRY(cnn_warm_start(c, mb, x0.p));
      PMX_TRY(lobpcg_setup(c, x0.p));
    } else {
      PMX_TRY(lobpcg_setup(c, nullptr));
    }
  }
  c->t_stages[PMX_STAGE_SU_LV2_1] += now() - t1;
  PMX_TRY(setup_coarse(c));
  c->t_stages[PMX_STAGE_SU] += now() - t0;
  c->setup_done = 1;
  return PMX_SUCCESS;
}

int pmx_setup(pmx_ctx* c) {
  if (!c) return PMX_ERR_ARG_NULL;
  return setup_common(c, nullptr, nullptr, nullptr);
}

int pmx_setup_batch(pmx_ctx* const* ctxs, int nprob) {
  if (!ctxs || nprob <= 0) {
    set_error("pmx_setup_batch needs at least one context");
    return PMX_ERR_ARG_NULL;
  }
  std::vector<pmx_ctx*> cs(ctxs, ctxs + nprob);
  for (int i = 0; i < nprob; ++i) {
    if (!cs[i]) {
      set_error("pmx_setup_batch: context %d is null", i);
      return PMX_ERR_ARG_NULL;
    }
    for (int j = 0; j < i; ++j)
      if (cs[j] == cs[i]) {
        set_error("pmx_setup_batch: context %d is listed twice", i);
        return PMX_ERR_ARG_WRONG;
      }
  }
  std::vector<double> t0(nprob);
  for (pmx_ctx* c : cs) {
    std::fill(c->t_stages, c->t_stages + PMX_MAX_LOG_STAGES, 0.0);
    ++c->setup_gen;
    c->setup_done = 0;
    c->h_eigen_num_given.clear();
  }
  for (int i = 0; i < nprob; ++i) {
    t0[i] = now();
    PMX_TRY(setup_fine(cs[i]));
  }
  const double t1 = now();
  // One launch has one block size: under -eig_autotune it is tuned on the
  // problem with the most subdomains and used for all of them.
  int tuner = -1;
  for (int i = 0; i < nprob; ++i)
    if (cs[i]->eig_autotune && (tuner < 0 || cs[i]->nsub > cs[tuner]->nsub)) tuner = i;
  if (tuner >= 0) {
    PMX_TRY(lobpcg_autotune(cs[tuner]));
    for (pmx_ctx* c : cs) {
      c->eig_guard = cs[tuner]->eig_guard;
      c->eig_shape = cs[tuner]->eig_shape;
    }
  }
  PMX_TRY(lobpcg_batch(cs.data(), nprob));
  const double t_eig = now() - t1;  // one shared launch: charged to every problem
  // The coarse levels of the members are independent: host threads take the
  // members largest first, each working on its own context's stream, so the
  // one-workgroup kernels of one member (pivot inverses, small
  // eigenproblems) overlap the trailing updates of another.
  std::vector<int> order(nprob);
  for (int i = 0; i < nprob; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return cs[x]->ncell > cs[y]->ncell; });
  std::vector<int> rc(nprob, PMX_SUCCESS);
  std::vector<std::string> msg(nprob);
  std::atomic<int> next{0};
  int dev = 0;
  PMX_HIP(hipGetDevice(&dev));
  auto worker = [&]() {
    if (hipSetDevice(dev) != hipSuccess) return;
    for (;;) {
      const int k = next.fetch_add(1);
      if (k >= nprob) break;
      pmx_ctx* c = cs[order[k]];
      const double tc = now();
      c->t_stages[PMX_STAGE_SU_LV2_1] += t_eig;
      try {
        rc[order[k]] = setup_coarse(c);
      } catch (const std::exception& e) {  // host allocation inside a worker thread
        set_error("coarse setup: %s", e.what());
        rc[order[k]] = PMX_ERR_MEM;
      }
      if (rc[order[k]]) {
        msg[order[k]] = g_err;
        continue;
      }
      c->t_stages[PMX_STAGE_SU] += (t1 - t0[order[k]]) + t_eig + (now() - tc);
      c->setup_done = 1;
    }
  };
  const int nth = std::max(1, std::min(cs[0]->batch_threads, nprob));
  if (nth == 1) {
    worker();
  } else {
    std::vector<std::thread> pool;
    for (int t = 0; t < nth; ++t) pool.emplace_back(worker);
    for (std::thread& th : pool) th.join();
  }
  if (next.load() < nprob) {
    set_error("pmx_setup_batch: hipSetDevice failed in a worker thread");
    return PMX_ERR_LIB;
  }
  for (int i = 0; i < nprob; ++i)
    if (rc[i]) {
      set_error("problem %d: %s", i, msg[i].c_str());
      return rc[i];
    }
  return PMX_SUCCESS;
}

int pmx_setup_from_basis(pmx_ctx* c, const double* phi, const int* eigen_num,
                         const double* eigvals) {
  if (!c || !phi) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  return setup_common(c, phi, eigen_num, eigvals);
}

int pmx_apply(pmx_ctx* c, const double* x, double* y) {
  if (!c || !x || !y) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  if (!c->setup_done) {
    set_error("PC_apply_vec before PC_setup");
    return PMX_ERR_ORDER;
  }
  PMX_HIP(hipMemcpyAsync(c->nat_buf.p, x, c->ncell * sizeof(double), hipMemcpyHostToDevice,
                         c->stream));
  PMX_TRY(launch_nat2brick(c, c->nat_buf.p, c->w_x.p, 1, c->ncell, c->ncell));
  PMX_TRY(apply_device(c, c->w_x.p, c->w_z.p, false));
  PMX_TRY(launch_brick2nat(c, c->w_z.p, c->nat_buf.p, 1, c->ncell, c->ncell));
  PMX_HIP(hipMemcpyAsync(y, c->nat_buf.p, c->ncell * sizeof(double), hipMemcpyDeviceToHost,
                         c->stream));
  PMX_HIP(hipStreamSynchronize(c->stream));
  return PMX_SUCCESS;
}

// PETSc KSPSolve_CG (standard variant, KSP_NORM_UNPRECONDITIONED, default
// convergence test) on the brick-order vectors w_b (right-hand side) and w_x
// (initial guess if x0_nonzero, solution on return).
static void cg_graph_release(pmx_ctx* c) {
  if (c->cg_exec) (void)hipGraphExecDestroy(c->cg_exec);
  if (c->cg_graph_h) (void)hipGraphDestroy(c->cg_graph_h);
  c->cg_exec = nullptr;
  c->cg_graph_h = nullptr;
}

// Captures `chunk` PCG iterations on the context stream into a HIP graph
// (once per setup and PCG parameter set) and instantiates it.
static int cg_graph_prepare(pmx_ctx* c, int chunk,
                            const std::function<int(int)>& enqueue_chunk) {
  const CgGraphKey key{c->setup_gen, chunk, c->cg_max_it, c->cg_abstol, c->cg_dtol};
  if (c->cg_exec && key == c->cg_key) return PMX_SUCCESS;
  cg_graph_release(c);
  PMX_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
  const int rc = enqueue_chunk(chunk);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) {
    set_error("PCG: stream capture failed (%s)", hipGetErrorString(e));
    return PMX_ERR_LIB;
  }
  c->cg_graph_h = g;
  PMX_HIP(hipGraphInstantiate(&c->cg_exec, g, nullptr, nullptr, 0));
  c->cg_key = key;
  return PMX_SUCCESS;
}

static int solve_core(pmx_ctx* c, int x0_nonzero, double rtol, int max_it, int* its,
                      double* rnorm) {
  c->cg_rtol = rtol;
  c->cg_abstol = 1e-50;
  c->cg_dtol = 1e5;
  c->cg_max_it = max_it;
  double* xb = c->w_x.p;
  double* bb = c->w_b.p;
  double* r = c->w_r.p;
  double* z = c->w_z.p;
  double* p0 = c->w_p.p;
  double* p1 = c->w_tmp.p;
  double* w = c->w_q.p;
  if (x0_nonzero) {
    PMX_TRY(launch_residual(c, bb, xb, r));
  } else {
    PMX_HIP(hipMemsetAsync(xb, 0, c->ncell * sizeof(double), c->stream));
    PMX_TRY(launch_residual(c, bb, nullptr, r));
  }
  double* hs = c->h_scal;
  PMX_HIP(hipMemcpyAsync(hs, c->cg_scal, kSNumScal * sizeof(double), hipMemcpyDeviceToHost,
                         c->stream));
  PMX_HIP(hipStreamSynchronize(c->stream));
  const double t_solve0 = now();
  // Iterations are enqueued in chunks; the convergence test runs on the
  // device, which marks the solve done so the remaining kernels of the chunk
  // return at once.  The host reads the scalars once per chunk.
  c->geo.stop = c->cg_scal + kSDone;
  int rc = PMX_SUCCESS;
  if (hs[kSDone] == 0.0) {
    // z = B r0; its z'r partials go to the slot of iteration 0.
    rc = apply_device(c, r, z, true);
    const bool graph = c->cg_graph && !c->kernel_timing && !c->log_stages;
    auto enqueue_chunk = [&](int n) -> int {
      for (int q = 0; q < n; ++q) {
        int ev = kt_begin(c, 4);
        PMX_TRY(launch_cg_p_spmv(c, z, p0, p1, w));
        kt_end(c, ev);
        PMX_TRY(apply_device(c, r, z, true, true, xb, p1, w));
        std::swap(p0, p1);
      }
      return PMX_SUCCESS;
    };
    // An even chunk leaves p0/p1 (and the level-1 buffers) where they started,
    // so one captured chunk can be replayed.
    const int chunk = graph ? (c->cg_chunk + 1) / 2 * 2 : c->cg_chunk;
    if (!rc && graph) rc = cg_graph_prepare(c, chunk, enqueue_chunk);
    while (!rc) {
      if (graph) {
        if (hipGraphLaunch(c->cg_exec, c->stream) != hipSuccess) {
          set_error("PCG: hipGraphLaunch failed");
          rc = PMX_ERR_LIB;
          break;
        }
      } else {
        rc = enqueue_chunk(chunk);
        if (rc) break;
      }
      if (hipMemcpyAsync(hs, c->cg_scal, kSNumScal * sizeof(double), hipMemcpyDeviceToHost,
                         c->stream) != hipSuccess ||
          hipStreamSynchronize(c->stream) != hipSuccess) {
        set_error("PCG: reading device scalars failed");
        rc = PMX_ERR_LIB;
        break;
      }
      if (hs[kSDone] != 0.0) break;
    }
  }
  c->geo.stop = nullptr;
  PMX_TRY(rc);
  PMX_HIP(hipStreamSynchronize(c->stream));
  PMX_TRY(kt_collect(c));
  const int k = (int)hs[kSIts];
  const int reason = (int)hs[kSReason];
  const double dp = std::sqrt(hs[kSRnorm2]);
  c->last_its = k;
  c->last_rnorm = dp;
  c->last_rnorm0 = std::sqrt(hs[kSRnorm0]);
  c->last_reason = reason;
  c->t_solve = now() - t_solve0;
  if (its) *its = k;
  if (rnorm) *rnorm = dp;
  return reason > 0 ? PMX_SUCCESS : PMX_ERR_NOT_CONVERGED;
}

static int solve_check(pmx_ctx* c, const void* b, const void* x) {
  if (!c || !b || !x) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  if (!c->setup_done) {
    set_error("KSPSolve before PC_setup");
    return PMX_ERR_ORDER;
  }
  return PMX_SUCCESS;
}

int pmx_solve(pmx_ctx* c, const double* b, double* x, int x0_nonzero, double rtol, int max_it,
              int* its, double* rnorm) {
  PMX_TRY(solve_check(c, b, x));
  const size_t nb = c->ncell * sizeof(double);
  PMX_HIP(hipMemsetAsync(c->cg_scal, 0, 16 * sizeof(double), c->stream));
  PMX_HIP(hipMemcpyAsync(c->nat_buf.p, b, nb, hipMemcpyHostToDevice, c->stream));
  PMX_TRY(launch_nat2brick(c, c->nat_buf.p, c->w_b.p, 1, c->ncell, c->ncell));
  if (x0_nonzero) {
    PMX_HIP(hipMemcpyAsync(c->nat_buf.p, x, nb, hipMemcpyHostToDevice, c->stream));
    PMX_TRY(launch_nat2brick(c, c->nat_buf.p, c->w_x.p, 1, c->ncell, c->ncell));
  }
  const int rc = solve_core(c, x0_nonzero, rtol, max_it, its, rnorm);
  if (rc != PMX_SUCCESS && rc != PMX_ERR_NOT_CONVERGED) return rc;
  PMX_TRY(launch_brick2nat(c, c->w_x.p, c->nat_buf.p, 1, c->ncell, c->ncell));
  PMX_HIP(hipMemcpyAsync(x, c->nat_buf.p, nb, hipMemcpyDeviceToHost, c->stream));
  PMX_HIP(hipStreamSynchronize(c->stream));
  return rc;
}

int pmx_solve_device(pmx_ctx* c, const double* d_b, double* d_x, int x0_nonzero, double rtol,
                     int max_it, int* its, double* rnorm) {
  PMX_TRY(solve_check(c, d_b, d_x));
  PMX_HIP(hipMemsetAsync(c->cg_scal, 0, 16 * sizeof(double), c->stream));
  PMX_TRY(launch_nat2brick(c, d_b, c->w_b.p, 1, c->ncell, c->ncell));
  if (x0_nonzero) PMX_TRY(launch_nat2brick(c, d_x, c->w_x.p, 1, c->ncell, c->ncell));
  const int rc = solve_core(c, x0_nonzero, rtol, max_it, its, rnorm);
  if (rc != PMX_SUCCESS && rc != PMX_ERR_NOT_CONVERGED) return rc;
  PMX_TRY(launch_brick2nat(c, c->w_x.p, d_x, 1, c->ncell, c->ncell));
  PMX_HIP(hipStreamSynchronize(c->stream));
  return rc;
}

int pmx_solve_device_noalloc(pmx_ctx* c, const double* d_b, double* d_x, int x0_nonzero,
                             double rtol, int max_it, int* its, double* rnorm) {
  if (!c) {
    set_error("null argument");
    return PMX_ERR_ARG_NULL;
  }
  const size_t need = 6 * (size_t)c->nsub + kSNumScal;
  if (need > (size_t)kScratchDoubles) {
    set_error("pmx_solve_device_noalloc: %d subdomains need %zu doubles of solve scratch, the "
              "fixed array holds %d (at most %d subdomains)", c->nsub, need, kScratchDoubles,
              (kScratchDoubles - kSNumScal) / 6);
    return PMX_ERR_ARG_OUTOFRANGE;
  }
  PMX_TRY(solve_check(c, d_b, d_x));
  if (!c->scratch_dev) {
    set_error("pmx_solve_device_noalloc: no solve scratch on this context's device");
    return PMX_ERR_LIB;
  }
  // The PCG kernels take the partial sums and scalars from the fixed array
  // for this call, launched eagerly: no graph (its instantiation may
  // allocate), no timing events.  Restored on every return.
  struct Restore {
    pmx_ctx* c;
    double *part, *scal;
    int graph, timing, stages;
    ~Restore() {
      c->cg_part = part;
      c->cg_scal = scal;
      c->cg_graph = graph;
      c->kernel_timing = timing;
      c->log_stages = stages;
    }
  } restore{c, c->cg_part, c->cg_scal, c->cg_graph, c->kernel_timing, c->log_stages};
  c->cg_scal = c->scratch_dev;
  c->cg_part = c->scratch_dev + kSNumScal;
  c->cg_graph = 0;
  c->kernel_timing = 0;
  c->log_stages = 0;
  // The scalars start at zero (the array keeps the previous call's values).
  PMX_HIP(hipMemsetAsync(c->cg_scal, 0, kSNumScal * sizeof(double), c->stream));
  PMX_TRY(launch_nat2brick(c, d_b, c->w_b.p, 1, c->ncell, c->ncell));
  if (x0_nonzero) PMX_TRY(launch_nat2brick(c, d_x, c->w_x.p, 1, c->ncell, c->ncell));
  const int rc = solve_core(c, x0_nonzero, rtol, max_it, its, rnorm);
  if (rc != PMX_SUCCESS && rc != PMX_ERR_NOT_CONVERGED) return rc;
  PMX_TRY(launch_brick2nat(c, c->w_x.p, d_x, 1, c->ncell, c->ncell));
  PMX_HIP(hipStreamSynchronize(c->stream));
  return rc;
}

int pmx_get_kernel_times(pmx_ctx* c, double* ms, int* count, int n, int reset) {
  if (!c || !ms || !count) return PMX_ERR_ARG_NULL;
  for (int i = 0; i < n && i < PMX_NUM_TIMED_KERNELS; ++i) {
    ms[i] = c->kt_ms[i];
    count[i] = c->kt_count[i];
  }
  if (reset)
    for (int i = 0; i < PMX_NUM_TIMED_KERNELS; ++i) {
      c->kt_ms[i] = 0.0;
      c->kt_count[i] = 0;
    }
  return PMX_SUCCESS;
}

int pmx_get_stats(const pmx_ctx* c, pmx_stats* st) {
  if (!c || !st) return PMX_ERR_ARG_NULL;
  memset(st, 0, sizeof(*st));
  st->M = c->mesh[0];
  st->N = c->mesh[1];
  st->P = c->mesh[2];
  for (int a = 0; a < 3; ++a) st->sd[a] = c->sd[a];
  st->nsub = c->nsub;
  st->size_coarse = c->nc;
  st->size_mat_lv3 = c->eigen_num_lv2;
  if (!c->h_eigen_num.empty()) {
    st->eigen_num_lv1_range[0] = st->eigen_num_lv1_range[1] = c->h_eigen_num[0];
    st->eigen_min_lv1_range[0] = st->eigen_min_lv1_range[1] = c->h_eigen_min_lv1[0];
    st->eigen_max_lv1_range[0] = st->eigen_max_lv1_range[1] = c->h_eigen_max_lv1[0];
    for (int s = 1; s < c->nsub; ++s) {
      st->eigen_num_lv1_range[0] = std::min(st->eigen_num_lv1_range[0], c->h_eigen_num[s]);
      st->eigen_num_lv1_range[1] = std::max(st->eigen_num_lv1_range[1], c->h_eigen_num[s]);
      st->eigen_min_lv1_range[0] = std::min(st->eigen_min_lv1_range[0], c->h_eigen_min_lv1[s]);
      st->eigen_min_lv1_range[1] = std::max(st->eigen_min_lv1_range[1], c->h_eigen_min_lv1[s]);
      st->eigen_max_lv1_range[0] = std::min(st->eigen_max_lv1_range[0], c->h_eigen_max_lv1[s]);
      st->eigen_max_lv1_range[1] = std::max(st->eigen_max_lv1_range[1], c->h_eigen_max_lv1[s]);
    }
  }
  st->eigen_num_lv2 = c->eigen_num_lv2;
  st->eigen_min_lv2 = c->eigen_min_lv2;
  st->eigen_max_lv2 = c->eigen_max_lv2;
  for (int i = 0; i < PMX_MAX_LOG_STAGES; ++i) st->t_stages[i] = c->t_stages[i];
  st->eig_iters_max = c->eig_iters_max;
  st->eig_iters_mean = c->nsub ? (double)c->eig_iters_sum / c->nsub : 0.0;
  st->eig_fallbacks = c->eig_fallbacks;
  st->eig_stagnated = c->eig_stagnated;
  st->eig_block = c->eig_block;
  st->coarse_mode = c->lv2_icc ? 3
                    : (c->use_band ? (c->use_planes ? 4 : 2) : (c->use_fop ? 0 : 1));
  st->coarse_stages = (int)c->coarse_prog.size();
  st->lv3_floor = c->lv3_floor;
  st->t_eig_tune = c->t_tune;
  st->eig_shape = c->eig_shape;
  st->t_eig_kernel = c->t_eig_kernel;
  {
    const double nc2 = (double)c->nc * c->nc * sizeof(double);
    const double kk = (double)c->nsub * 7 * c->en_lv1 * c->en_lv1 * sizeof(double);
    if (c->lv2_icc) {
      // block rows and the inverse pivot blocks, read by both sweeps
      st->coarse_factor_bytes = kk + (double)c->nsub * c->en_lv1 * c->en_lv1 * sizeof(double);
      st->coarse_solve_bytes = 2.0 * st->coarse_factor_bytes;
    } else if (c->use_band && c->use_planes) {
      const int Z = (int)c->pl_dof.size() - 1;
      double last = 0.0;
      if (Z > 0) last = (double)(c->pl_goff[Z] - c->pl_goff[Z - 1]) * sizeof(double);
      st->coarse_factor_bytes = (double)c->pl_bytes;
      st->coarse_solve_bytes = 2.0 * c->pl_bytes - last;  // G_z twice, the last plane once
    } else if (c->use_band) {
      st->coarse_factor_bytes = (double)(c->band.n + c->band_dinv.n) * sizeof(double);
      st->coarse_solve_bytes = 2.0 * st->coarse_factor_bytes;
    } else {
      st->coarse_factor_bytes = nc2;
      st->coarse_solve_bytes = nc2;
    }
  }
  st->lv3_iters = c->lv3_iters;
  st->pc_precision = c->pc_precision;
  st->its = c->last_its;
  st->rnorm = c->last_rnorm;
  st->rnorm0 = c->last_rnorm0;
  st->reason = c->last_reason;
  st->t_solve = c->t_solve;
  return PMX_SUCCESS;
}

int pmx_print_stat(const pmx_ctx* c) {
  pmx_stats s;
  PMX_TRY(pmx_get_stats(c, &s));
  printf("eigen_max_lv1_range=[%.5E, %.5E], \teigen_min_lv1_range=[%.5E, %.5E], "
         "\teigen_num_lv1_range=[%d, %d].\n",
         s.eigen_max_lv1_range[0], s.eigen_max_lv1_range[1], s.eigen_min_lv1_range[0],
         s.eigen_min_lv1_range[1], s.eigen_num_lv1_range[0], s.eigen_num_lv1_range[1]);
  // With process boxes the count is per box (PC_get_range over the ranks).
  int num_lo = s.eigen_num_lv2, num_hi = s.eigen_num_lv2;
  if (!c->h_eigen_num_lv2_box.empty() && c->nbox > 1) {
    const auto mm = std::minmax_element(c->h_eigen_num_lv2_box.begin(),
                                        c->h_eigen_num_lv2_box.end());
    num_lo = *mm.first;
    num_hi = *mm.second;
  }
  printf("eigen_max_lv2_range=[%.5E, %.5E], \teigen_min_lv2_range=[%.5E, %.5E], "
         "\teigen_num_lv2_range=[%d, %d], \tsize_mat_lv3=%d.\n",
         s.eigen_max_lv2, s.eigen_max_lv2, s.eigen_min_lv2, s.eigen_min_lv2, num_lo, num_hi,
         s.size_mat_lv3);
  const double* t = s.t_stages;
  printf("su_lv1=[%.5f, %.5f], \tsu_lv2_1=[%.5f, %.5f], \tsu_lv2_2=[%.5f, %.5f], "
         "\tsu_lv3_1=[%.5f, %.5f], \tsu_lv3_2=[%.5f, %.5f], \tsu=[%.5f, %.5f].\n",
         t[0], t[0], t[1], t[1], t[2], t[2], t[3], t[3], t[4], t[4], t[8], t[8]);
  printf("av_lv1=[%.5f, %.5f], \tav_lv2=[%.5f, %.5f], \tav_lv3=[%.5f, %.5f], \tav=[%.5f, %.5f].\n",
         t[5], t[5], t[6], t[6], t[7], t[7], t[9], t[9]);
  return PMX_SUCCESS;
}

int pmx_get_basis(const pmx_ctx* cc, int level, double* out, double* eigvals) {
  pmx_ctx* c = const_cast<pmx_ctx*>(cc);
  if (!c || !out) return PMX_ERR_ARG_NULL;
  if (!c->setup_done) {
    set_error("basis requested before setup");
    return PMX_ERR_ORDER;
  }
  if (level == 2) {
    DBuf<double> tmp;
    PMX_TRY(tmp.alloc((size_t)c->ncell * c->en_lv1));
    PMX_TRY(launch_basis_rows2nat(c, c->phi.p, tmp.p, c->en_lv1));
    PMX_HIP(hipMemcpyAsync(out, tmp.p, (size_t)c->ncell * c->en_lv1 * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
    PMX_HIP(hipStreamSynchronize(c->stream));
    if (eigvals) memcpy(eigvals, c->h_eig_lv1.data(), c->h_eig_lv1.size() * sizeof(double));
    return PMX_SUCCESS;
  }
  if (level == 3) {
    if (c->use_2level) {
      set_error("no level-3 basis in the 2-level variant");
      return PMX_ERR_ORDER;
    }
    // Phi_cc = Phi V on the host.
    const int k2 = c->eigen_num_lv2, kmax = c->en_lv1;
    std::vector<double> V((size_t)c->nc * k2), rows((size_t)c->ncell * kmax);
    PMX_HIP(hipMemcpy(V.data(), c->V2.p, V.size() * sizeof(double), hipMemcpyDeviceToHost));
    PMX_HIP(hipMemcpy(rows.data(), c->phi.p, rows.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int n = 0; n < c->ncell; ++n) {
      const int bi = c->h_nat2brick[n];
      int s = 0;
      // find the brick of bi
      int lo = 0, hi = c->nsub - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (c->h_bricks[mid].off <= bi)
          lo = mid;
        else
          hi = mid - 1;
      }
      s = lo;
      const Brick& b = c->h_bricks[s];
      for (int j = 0; j < k2; ++j) {
        double v = 0.0;
        for (int i = 0; i < b.neig; ++i) v += rows[(size_t)bi * kmax + i] * V[(size_t)j * c->nc + b.dof + i];
        out[(size_t)j * c->ncell + n] = v;
      }
    }
    if (eigvals) memcpy(eigvals, c->h_eig_lv2.data(), c->h_eig_lv2.size() * sizeof(double));
    return PMX_SUCCESS;
  }
  set_error("level must be 2 or 3");
  return PMX_ERR_ARG_WRONG;
}

int pmx_get_coarse_matrix(const pmx_ctx* c, double* out) {
  if (!c || !out) return PMX_ERR_ARG_NULL;
  if (!c->setup_done) {
    set_error("coarse matrix requested before setup");
    return PMX_ERR_ORDER;
  }
  if (c->Ac.p) {
    PMX_HIP(hipMemcpy(out, c->Ac.p, (size_t)c->nc * c->nc * sizeof(double),
                      hipMemcpyDeviceToHost));
    return PMX_SUCCESS;
  }
  // -lv2_icc and the plane factor keep block rows only: expand them (with
  // process boxes, the level-2 matrix keeps only the blocks inside a box).
  const int k = c->en_lv1;
  std::vector<double> blk(c->Acb.n);
  PMX_HIP(hipMemcpy(blk.data(), c->Acb.p, blk.size() * sizeof(double), hipMemcpyDeviceToHost));
  std::fill(out, out + (size_t)c->nc * c->nc, 0.0);
  for (int s = 0; s < c->nsub; ++s) {
    const Brick& b = c->h_bricks[s];
    for (int slot = 0; slot < 7; ++slot) {
      const int o = slot == 0 ? s : b.nb[slot - 1];
      if (o < 0) continue;
      const Brick& ob = c->h_bricks[o];
      if (ob.box != b.box) continue;
      const double* B = blk.data() + ((size_t)s * 7 + slot) * k * k;
      for (int i = 0; i < b.neig; ++i)
        for (int j = 0; j < ob.neig; ++j) out[(size_t)(b.dof + i) * c->nc + ob.dof + j] = B[i * k + j];
    }
  }
  return PMX_SUCCESS;
}

int pmx_final(pmx_ctx* c) {
  if (!c) return PMX_SUCCESS;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  cg_graph_release(c);
  c->d_bricks.release();
  c->d_views.release();
  c->d_order.release();
  for (auto& d : c->d_ax) d.release();
  c->d_plane_start.release();
  c->d_plane_cell.release();
  c->d_shape_pstart.release();
  c->d_shape_coff.release();
  for (int a = 0; a < 3; ++a) {
    c->kappa[a].release();
    c->T[a].release();
    c->T32[a].release();
  }
  for (DBuf<double>* b : {&c->c_z, &c->c_g, &c->c_h, &c->c_cs, &c->Fop, &c->lob_buf, &c->bicc_dinv,
                          &c->wd})
    b->release();
  c->wd32.release();
  c->bicc_plane.release();
  c->bicc_elem.release();
  c->bicc_slot.release();
  c->band.release();
  c->band_dinv.release();
  c->band_tmp.release();
  c->pl_G.release();
  c->pl_w.release();
  for (DBuf<double>* b : {&c->fac_f, &c->fac_b, &c->mass, &c->phi, &c->Ac, &c->Acb, &c->Ainv,
                          &c->V2, &c->Acc_lu, &c->w_x, &c->w_y, &c->w_yb, &c->w_r, &c->w_b, &c->w_p,
                          &c->w_z, &c->w_q, &c->w_tmp, &c->c_a, &c->c_b, &c->c_c, &c->c_d, &c->c_e,
                          &c->partial, &c->scal, &c->nat_buf})
    b->release();
  for (DBuf<float>* b : {&c->fac32_f, &c->fac32_b, &c->phi32, &c->w_y32, &c->w_y32b}) b->release();
  for (hipEvent_t e : c->kt_events) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PMX_SUCCESS;
}

}  // extern "C"
