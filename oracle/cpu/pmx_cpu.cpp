// FP64 CPU restatement (C++ / OpenMP) of PreMixFEM's three-level GMsFEM
// preconditioner and the PETSc CG that drives it, on one process.
//
// TEST INFRASTRUCTURE ONLY: the CPU baseline of the benchmark (timed on the
// host cores beside the GPU) and a second, compiled checker of the numpy
// oracle (oracle/premixfem_oracle.py), which it must agree with.  The
// product path (premixfem_amd) never loads it.
//
// What it restates, in the reference's own algorithms:
// * TPFA operator (PC_create_A, src/PreMixFEM_3D.c:1757-1824), pin term as
//   the workloads define it;
// * level 1: block ICC(0) of A[Omega_i, Omega_i] in natural order
//   (_PC_setup_lv1_J, :80-222; on the 7-point graph ICC(0) keeps no fill);
// * level 2: per-subdomain GHEP A_i^N phi = lambda M_i phi with the MOD mass
//   and its x-direction accumulation (:224-421) by shift-invert Lanczos at
//   target -1e-12 on an exact band Cholesky factor of A_i^N + 1e-12 M_i (the
//   SLEPc Krylov-Schur + STSINVERT + MUMPS set-up of :354-365), the constant
//   first pair deflated; Galerkin A_c in the energy form (:737-890) and its
//   exact band Cholesky factor (MUMPS, :894-906);
// * level 3: the smallest en_lv2 eigenpairs of A_c by shift-invert Lanczos on
//   that factor (:918-1175), Phi_cc = Phi V, A_cc = Phi_cc^T A Phi_cc + NIL I
//   and its LU (:1177-1390);
// * PC_apply_vec's V-cycle with its five residuals r = x - A y (:1840-1936);
// * KSPSolve_CG with the unpreconditioned norm (src/Expe_scal.c:101-119).
// Arrays are in the natural order arr[ez][ey][ex] (x fastest).
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

constexpr double kNil = 1.0e-12;
constexpr double kTarget = -1.0e-12;   // SLEPc target of the shift-invert (:354-365)

struct Part {
  std::vector<int> st, ln;   // per axis: starts, lengths
};

void partition(int n, int sd, std::vector<int>& st, std::vector<int>& ln) {
  st.assign(sd, 0);
  ln.assign(sd, 0);
  for (int i = 0; i < sd; ++i) ln[i] = n / sd + (n % sd > i ? 1 : 0);
  for (int i = 1; i < sd; ++i) st[i] = st[i - 1] + ln[i - 1];
}

// Lower band matrix in row storage: entry (i, j), i - bw <= j <= i, at
// a[i * (bw + 1) + (j - i + bw)].
struct Band {
  int n = 0, bw = 0;
  std::vector<double> a;
  void init(int n_, int bw_) {
    n = n_;
    bw = bw_;
    a.assign((size_t)n * (bw + 1), 0.0);
  }
  double& at(int i, int j) { return a[(size_t)i * (bw + 1) + (j - i + bw)]; }
  double at(int i, int j) const { return a[(size_t)i * (bw + 1) + (j - i + bw)]; }
  const double* row(int i) const { return &a[(size_t)i * (bw + 1) + bw - i]; }  // row(i)[j]
  double* row(int i) { return &a[(size_t)i * (bw + 1) + bw - i]; }
};

// Sequential band Cholesky (left-looking, row dot products); false on a
// non-positive pivot.
bool band_chol_seq(Band& A) {
  const int n = A.n, bw = A.bw;
  for (int i = 0; i < n; ++i) {
    double* ri = A.row(i);
    const int j0 = std::max(0, i - bw);
    for (int j = j0; j <= i; ++j) {
      const double* rj = A.row(j);
      const int p0 = std::max(j0, j - bw);
      double s = ri[j];
      for (int p = p0; p < j; ++p) s -= ri[p] * rj[p];
      if (j == i) {
        if (!(s > 0.0)) return false;
        ri[i] = std::sqrt(s);
      } else {
        ri[j] = s / rj[j];
      }
    }
  }
  return true;
}

// Blocked right-looking band Cholesky with OpenMP (panels of NB columns).
bool band_chol_par(Band& A) {
  const int n = A.n, bw = A.bw;
  constexpr int NB = 64;
  int fail = 0;
  for (int J0 = 0; J0 < n && !fail; J0 += NB) {
    const int J1 = std::min(n, J0 + NB);
    const int R1 = std::min(n, J1 + bw);   // rows with entries in the panel
    // Panel: the trailing updates of earlier panels are applied already.
    for (int j = J0; j < J1; ++j) {
      double* rj = A.row(j);
      double s = rj[j];
      for (int p = std::max(J0, j - bw); p < j; ++p) s -= rj[p] * rj[p];
      if (!(s > 0.0)) {
        fail = 1;
        break;
      }
      const double d = std::sqrt(s);
      rj[j] = d;
      const int iend = std::min(n - 1, j + bw);
#pragma omp parallel for schedule(static) if (iend - j > 256)
      for (int i = j + 1; i <= iend; ++i) {
        double* ri = A.row(i);
        double t = ri[j];
        for (int p = std::max(J0, i - bw); p < j; ++p) t -= ri[p] * rj[p];
        ri[j] = t / d;
      }
    }
    if (fail) break;
    // Trailing update of rows J1..R1-1 with the panel's columns.
#pragma omp parallel for schedule(dynamic, 8)
    for (int i = J1; i < R1; ++i) {
      double* ri = A.row(i);
      const int pi0 = std::max(J0, i - bw);
      if (pi0 >= J1) continue;
      for (int l = std::max(J1, i - bw); l <= i; ++l) {
        const double* rl = A.row(l);
        const int p0 = std::max(pi0, l - bw);
        double s = 0.0;
        for (int p = p0; p < J1; ++p) s += ri[p] * rl[p];
        ri[l] -= s;
      }
    }
  }
  return !fail;
}

// x <- L^{-T} L^{-1} x (one right-hand side), sequential.
void band_solve_seq(const Band& L, double* x) {
  const int n = L.n, bw = L.bw;
  for (int i = 0; i < n; ++i) {
    const double* ri = L.row(i);
    double s = x[i];
    for (int p = std::max(0, i - bw); p < i; ++p) s -= ri[p] * x[p];
    x[i] = s / ri[i];
  }
  for (int i = n - 1; i >= 0; --i) {
    const double* ri = L.row(i);
    const double z = x[i] / ri[i];
    x[i] = z;
    for (int p = std::max(0, i - bw); p < i; ++p) x[p] -= ri[p] * z;
  }
}

// The same with OpenMP over blocks of 64 rows (the large coarse factor).
void band_solve_par(const Band& L, double* x) {
  const int n = L.n, bw = L.bw;
  constexpr int RB = 64;
  for (int I0 = 0; I0 < n; I0 += RB) {
    const int I1 = std::min(n, I0 + RB);
#pragma omp parallel for schedule(static)
    for (int i = I0; i < I1; ++i) {
      const double* ri = L.row(i);
      double s = 0.0;
      for (int p = std::max(0, i - bw); p < I0; ++p) s += ri[p] * x[p];
      x[i] -= s;
    }
    for (int i = I0; i < I1; ++i) {
      const double* ri = L.row(i);
      double s = x[i];
      for (int p = std::max(I0, i - bw); p < i; ++p) s -= ri[p] * x[p];
      x[i] = s / ri[i];
    }
  }
  for (int I1 = n; I1 > 0; I1 -= RB) {
    const int I0 = std::max(0, I1 - RB);
    for (int i = I1 - 1; i >= I0; --i) {
      const double* ri = L.row(i);
      const double z = x[i] / ri[i];
      x[i] = z;
      for (int p = std::max(I0, i - bw); p < i; ++p) x[p] -= ri[p] * z;
    }
    const int P0 = std::max(0, I0 - bw);
#pragma omp parallel for schedule(static)
    for (int p = P0; p < I0; ++p) {
      double s = 0.0;
      for (int i = I0; i < I1; ++i)
        if (p >= i - bw) s += L.row(i)[p] * x[i];
      x[p] -= s;
    }
  }
}

// Symmetric eigenproblem of a small dense matrix (row-major, n x n) by cyclic
// Jacobi; eigenvalues ascending in w, eigenvectors in the columns of V.
void jacobi_eig(std::vector<double> A, int n, std::vector<double>& w, std::vector<double>& V) {
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const double v = A[(size_t)i * n + j] * A[(size_t)i * n + j];
        tot += v;
        if (i != j) off += v;
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[(size_t)p * n + q];
        if (apq == 0.0) continue;
        const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
        const double theta = 0.5 * (aqq - app) / apq;
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        const double c = 1.0 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
          A[(size_t)k * n + p] = c * akp - s * akq;
          A[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
          A[(size_t)p * n + k] = c * apk - s * aqk;
          A[(size_t)q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
          V[(size_t)k * n + p] = c * vkp - s * vkq;
          V[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> ord(n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  std::sort(ord.begin(), ord.end(),
            [&](int a, int b) { return A[(size_t)a * n + a] < A[(size_t)b * n + b]; });
  std::vector<double> V2((size_t)n * n);
  w.assign(n, 0.0);
  for (int c = 0; c < n; ++c) {
    w[c] = A[(size_t)ord[c] * n + ord[c]];
    for (int r = 0; r < n; ++r) V2[(size_t)r * n + c] = V[(size_t)r * n + ord[c]];
  }
  V.swap(V2);
}

uint64_t mix(uint64_t a) {
  a += 0x9E3779B97F4A7C15ull;
  a = (a ^ (a >> 30)) * 0xBF58476D1CE4E5B9ull;
  a = (a ^ (a >> 27)) * 0x94D049BB133111EBull;
  return a ^ (a >> 31);
}

// Shift-invert Lanczos for the largest eigenvalues theta of Op = K^{-1} B
// (B-inner product; B = diag(bd), or the identity when bd is null) on the
// B-orthogonal complement of the columns of D (nd of them, B-orthonormal).
// Returns the nwant largest theta (descending) and their B-orthonormal Ritz
// vectors; converged when |beta_m s_mi| <= tol theta_i for all of them.
struct LanczosOut {
  std::vector<double> theta, X;   // X: nwant columns of length n
  int steps = 0, converged = 0;
};

template <typename OpF>
LanczosOut lanczos_si(int n, const double* bd, OpF op, const double* D, int nd, int nwant,
                      double tol, int mmax, uint64_t seed) {
  LanczosOut out;
  mmax = std::min(mmax, n - nd);
  if (nwant <= 0 || mmax <= 0) return out;
  std::vector<double> Vb((size_t)(mmax + 1) * n), w(n), alpha, beta;
  auto dotB = [&](const double* a, const double* b) {
    double s = 0.0;
    if (bd)
      for (int i = 0; i < n; ++i) s += a[i] * bd[i] * b[i];
    else
      for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
  };
  auto orth = [&](double* v, int upto) {
    for (int pass = 0; pass < 2; ++pass) {
      for (int c = 0; c < nd; ++c) {
        const double h = dotB(D + (size_t)c * n, v);
        for (int i = 0; i < n; ++i) v[i] -= h * D[(size_t)c * n + i];
      }
      for (int c = 0; c < upto; ++c) {
        const double* q = &Vb[(size_t)c * n];
        const double h = dotB(q, v);
        for (int i = 0; i < n; ++i) v[i] -= h * q[i];
      }
    }
  };
  double* v0 = &Vb[0];
  for (int i = 0; i < n; ++i) v0[i] = (double)(mix(seed ^ (uint64_t)i) >> 11) * (1.0 / 9007199254740992.0) - 0.5;
  orth(v0, 0);
  double nv = std::sqrt(dotB(v0, v0));
  for (int i = 0; i < n; ++i) v0[i] /= nv;
  std::vector<double> wT, S;
  int m = 0;
  for (int j = 0; j < mmax; ++j) {
    op(&Vb[(size_t)j * n], w.data());
    const double a = dotB(&Vb[(size_t)j * n], w.data());
    alpha.push_back(a);
    orth(w.data(), j + 1);
    const double b = std::sqrt(dotB(w.data(), w.data()));
    beta.push_back(b);
    m = j + 1;
    const bool last = (m == mmax) || !(b > 1e-300);
    if (m >= nwant && (m % 4 == 0 || last)) {
      std::vector<double> T((size_t)m * m, 0.0);
      for (int i = 0; i < m; ++i) {
        T[(size_t)i * m + i] = alpha[i];
        if (i + 1 < m) T[(size_t)i * m + i + 1] = T[(size_t)(i + 1) * m + i] = beta[i];
      }
      jacobi_eig(T, m, wT, S);
      int conv = 1;
      for (int c = 0; c < nwant; ++c) {
        const int col = m - 1 - c;
        if (!(std::fabs(b * S[(size_t)(m - 1) * m + col]) <= tol * std::fabs(wT[col]))) conv = 0;
      }
      if (conv || last) {
        out.converged = conv;
        break;
      }
    }
    if (!(b > 1e-300)) break;
    double* vn = &Vb[(size_t)(j + 1) * n];
    for (int i = 0; i < n; ++i) vn[i] = w[i] / b;
  }
  out.steps = m;
  if ((int)wT.size() != m && m > 0) {   // no Ritz check ran (m < nwant)
    std::vector<double> T((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) {
      T[(size_t)i * m + i] = alpha[i];
      if (i + 1 < m) T[(size_t)i * m + i + 1] = T[(size_t)(i + 1) * m + i] = beta[i];
    }
    jacobi_eig(T, m, wT, S);
  }
  out.theta.assign(nwant, 0.0);
  out.X.assign((size_t)nwant * n, 0.0);
  for (int c = 0; c < nwant && c < m; ++c) {
    const int col = m - 1 - c;
    out.theta[c] = wT[col];
    double* x = &out.X[(size_t)c * n];
    for (int q = 0; q < m; ++q) {
      const double s = S[(size_t)q * m + col];
      const double* vq = &Vb[(size_t)q * n];
      for (int i = 0; i < n; ++i) x[i] += s * vq[i];
    }
    const double nx = std::sqrt(dotB(x, x));
    for (int i = 0; i < n; ++i) x[i] /= nx;
  }
  return out;
}

}  // namespace

extern "C" {

struct pmxc_problem {
  int M, N, P;
  double dom[3];
  const double* kx;
  const double* ky;
  const double* kz;
  int sd[3];
  int en_lv1, en_lv2;
  int pin[3];          // pinned cell, pin[0] < 0: none
  const double* rhs;
  double rtol;
  int max_it;
  int threads;         // OpenMP threads (<= 0: the runtime's default)
  double eig_tol;      // relative tolerance of the eigensolvers (SLEPc's default 1e-8)
};

struct pmxc_result {
  double* x;           // output pressure (natural order), M*N*P doubles
  int its, reason;     // PCG iterations, 2 = converged (rtol), -3 = max_it
  double rnorm, rnorm0;
  double t[7];         // seconds: operator+level 1, level-2 eigen, Galerkin, coarse factor,
                       // level 3, PCG, total
  int nc, lanczos_steps_max, lanczos_unconverged, threads;
  double eig_min_lv1, eig_max_lv1, eig_lv2[16];
  double* eig_lv1;     // optional output [nsub][en_lv1] (may be null)
};

int pmxc_setup_solve(const pmxc_problem* pr, pmxc_result* res) {
  const double t_all = now();
  if (pr->threads > 0) omp_set_num_threads(pr->threads);
  res->threads = omp_get_max_threads();
  const int M = pr->M, N = pr->N, P = pr->P, k = pr->en_lv1;
  const long long n = (long long)M * N * P;
  const size_t MN = (size_t)M * N;
  const double Hx = pr->dom[0] / M, Hy = pr->dom[1] / N, Hz = pr->dom[2] / P;
  const double E = Hx * Hy * Hz;
  const double cx = (Hy * Hz) * (Hy * Hz) / E, cy = (Hz * Hx) * (Hz * Hx) / E,
               cz = (Hx * Hy) * (Hx * Hy) / E;
  const double* kap[3] = {pr->kx, pr->ky, pr->kz};
  // ---- operator: face coefficients Tx[c] of the face (c, c+x) etc. ----
  double t0 = now();
  std::vector<double> Tf[3];
  for (int d = 0; d < 3; ++d) Tf[d].assign(n, 0.0);
#pragma omp parallel for schedule(static)
  for (int ez = 0; ez < P; ++ez)
    for (int ey = 0; ey < N; ++ey)
      for (int ex = 0; ex < M; ++ex) {
        const size_t c = ez * MN + (size_t)ey * M + ex;
        if (ex + 1 < M) Tf[0][c] = cx * (2.0 / (1.0 / kap[0][c] + 1.0 / kap[0][c + 1]));
        if (ey + 1 < N) Tf[1][c] = cy * (2.0 / (1.0 / kap[1][c] + 1.0 / kap[1][c + M]));
        if (ez + 1 < P) Tf[2][c] = cz * (2.0 / (1.0 / kap[2][c] + 1.0 / kap[2][c + MN]));
      }
  const double *Tx = Tf[0].data(), *Ty = Tf[1].data(), *Tz = Tf[2].data();
  long long pin = -1;
  double tau = 0.0;
  if (pr->pin[0] >= 0) {
    const int x = pr->pin[0], y = pr->pin[1], z = pr->pin[2];
    pin = z * MN + (size_t)y * M + x;
    tau = std::max({x > 0 ? Tx[pin - 1] : 0.0, x + 1 < M ? Tx[pin] : 0.0,
                    y > 0 ? Ty[pin - M] : 0.0, y + 1 < N ? Ty[pin] : 0.0,
                    z > 0 ? Tz[pin - MN] : 0.0, z + 1 < P ? Tz[pin] : 0.0});
  }
  auto spmv = [&](const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int ez = 0; ez < P; ++ez)
      for (int ey = 0; ey < N; ++ey) {
        const size_t r = ez * MN + (size_t)ey * M;
        for (int ex = 0; ex < M; ++ex) {
          const size_t c = r + ex;
          const double xc = x[c];
          double s = 0.0;
          if (ex > 0) s += Tx[c - 1] * (xc - x[c - 1]);
          if (ex + 1 < M) s += Tx[c] * (xc - x[c + 1]);
          if (ey > 0) s += Ty[c - M] * (xc - x[c - M]);
          if (ey + 1 < N) s += Ty[c] * (xc - x[c + M]);
          if (ez > 0) s += Tz[c - MN] * (xc - x[c - MN]);
          if (ez + 1 < P) s += Tz[c] * (xc - x[c + MN]);
          if ((long long)c == pin) s += tau * xc;
          y[c] = s;
        }
      }
  };
  // ---- partition ----
  std::vector<int> st[3], ln[3];
  partition(M, pr->sd[0], st[0], ln[0]);
  partition(N, pr->sd[1], st[1], ln[1]);
  partition(P, pr->sd[2], st[2], ln[2]);
  const int sdx = pr->sd[0], sdy = pr->sd[1], sdz = pr->sd[2];
  const int nsub = sdx * sdy * sdz;
  std::vector<int> subx(M), suby(N), subz(P);
  for (int a = 0; a < sdx; ++a)
    for (int i = 0; i < ln[0][a]; ++i) subx[st[0][a] + i] = a;
  for (int a = 0; a < sdy; ++a)
    for (int i = 0; i < ln[1][a]; ++i) suby[st[1][a] + i] = a;
  for (int a = 0; a < sdz; ++a)
    for (int i = 0; i < ln[2][a]; ++i) subz[st[2][a] + i] = a;
  auto corners = [&](int s, int& x0, int& nx, int& y0, int& ny, int& z0, int& nz) {
    const int a = s % sdx, b = (s / sdx) % sdy, c = s / (sdx * sdy);
    x0 = st[0][a];
    nx = ln[0][a];
    y0 = st[1][b];
    ny = ln[1][b];
    z0 = st[2][c];
    nz = ln[2][c];
  };
  // ---- level 1: block ICC(0), one inverse pivot per cell ----
  std::vector<double> dinv(n);
#pragma omp parallel for schedule(dynamic, 1)
  for (int s = 0; s < nsub; ++s) {
    int x0, nx, y0, ny, z0, nz;
    corners(s, x0, nx, y0, ny, z0, nz);
    for (int lz = 0; lz < nz; ++lz)
      for (int ly = 0; ly < ny; ++ly)
        for (int lx = 0; lx < nx; ++lx) {
          const int ex = x0 + lx, ey = y0 + ly, ez = z0 + lz;
          const size_t c = ez * MN + (size_t)ey * M + ex;
          double a = (ex > 0 ? Tx[c - 1] : 0.0) + (ex + 1 < M ? Tx[c] : 0.0) +
                     (ey > 0 ? Ty[c - M] : 0.0) + (ey + 1 < N ? Ty[c] : 0.0) +
                     (ez > 0 ? Tz[c - MN] : 0.0) + (ez + 1 < P ? Tz[c] : 0.0);
          if ((long long)c == pin) a += tau;
          if (lx > 0) a -= Tx[c - 1] * Tx[c - 1] * dinv[c - 1];
          if (ly > 0) a -= Ty[c - M] * Ty[c - M] * dinv[c - M];
          if (lz > 0) a -= Tz[c - MN] * Tz[c - MN] * dinv[c - MN];
          dinv[c] = 1.0 / a;
        }
  }
  auto lv1_solve = [&](const double* r, double* y) {   // y = B1 r
#pragma omp parallel for schedule(dynamic, 1)
    for (int s = 0; s < nsub; ++s) {
      int x0, nx, y0, ny, z0, nz;
      corners(s, x0, nx, y0, ny, z0, nz);
      for (int lz = 0; lz < nz; ++lz)
        for (int ly = 0; ly < ny; ++ly)
          for (int lx = 0; lx < nx; ++lx) {
            const size_t c = (z0 + lz) * MN + (size_t)(y0 + ly) * M + x0 + lx;
            double v = r[c];
            if (lx > 0) v += Tx[c - 1] * y[c - 1];
            if (ly > 0) v += Ty[c - M] * y[c - M];
            if (lz > 0) v += Tz[c - MN] * y[c - MN];
            y[c] = v * dinv[c];
          }
      for (int lz = nz - 1; lz >= 0; --lz)
        for (int ly = ny - 1; ly >= 0; --ly)
          for (int lx = nx - 1; lx >= 0; --lx) {
            const size_t c = (z0 + lz) * MN + (size_t)(y0 + ly) * M + x0 + lx;
            double v = 0.0;
            if (lx + 1 < nx) v += Tx[c] * y[c + 1];
            if (ly + 1 < ny) v += Ty[c] * y[c + M];
            if (lz + 1 < nz) v += Tz[c] * y[c + MN];
            y[c] += v * dinv[c];
          }
    }
  };
  // ---- eigen mass (MOD, with the reference's x accumulation) ----
  std::vector<double> mass(n);
#pragma omp parallel for schedule(static)
  for (int ez = 0; ez < P; ++ez)
    for (int ey = 0; ey < N; ++ey)
      for (int ex = 0; ex < M; ++ex) {
        const size_t c = ez * MN + (size_t)ey * M + ex;
        auto hm = [](double a, double b) { return 2.0 / (1.0 / a + 1.0 / b); };
        const double hx = ex + 1 < M ? 2.0 * hm(kap[0][c], kap[0][c + 1]) / Hx / Hx : 0.0;
        const double lx_ = ex > 0 ? 2.0 * hm(kap[0][c - 1], kap[0][c]) / Hx / Hx : 0.0;
        const double hy = ey + 1 < N ? 2.0 * hm(kap[1][c], kap[1][c + M]) / Hy / Hy : 0.0;
        const double ly_ = ey > 0 ? 2.0 * hm(kap[1][c - M], kap[1][c]) / Hy / Hy : 0.0;
        const double hz = ez + 1 < P ? 2.0 * hm(kap[2][c], kap[2][c + MN]) / Hz / Hz : 0.0;
        const double lz_ = ez > 0 ? 2.0 * hm(kap[2][c - MN], kap[2][c]) / Hz / Hz : 0.0;
        const bool first = ex == st[0][subx[ex]];
        const bool next_same = ex + 1 < M && subx[ex + 1] == subx[ex];
        double m = first ? lx_ : 0.0;
        m = m + hx;
        m = m + ly_;
        m = m + hy;
        m = m + lz_;
        m = m + hz;
        m = m + (next_same ? hx : 0.0);
        mass[c] = m * E;
      }
  res->t[0] = now() - t0;
  // ---- level 2: subdomain GHEPs ----
  t0 = now();
  std::vector<double> phi((size_t)n * k);      // rows [cell][k]
  std::vector<double> eig((size_t)nsub * k);
  int steps_max = 0, unconv = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(max : steps_max) reduction(+ : unconv)
  for (int s = 0; s < nsub; ++s) {
    int x0, nx, y0, ny, z0, nz;
    corners(s, x0, nx, y0, ny, z0, nz);
    const int ns = nx * ny * nz, bw = nx * ny;
    std::vector<size_t> gc(ns);
    std::vector<double> m(ns);
    for (int lz = 0; lz < nz; ++lz)
      for (int ly = 0; ly < ny; ++ly)
        for (int lx = 0; lx < nx; ++lx) {
          const int li = (lz * ny + ly) * nx + lx;
          gc[li] = (z0 + lz) * MN + (size_t)(y0 + ly) * M + x0 + lx;
          m[li] = mass[gc[li]];
        }
    // K = A_i^N - target M (interior faces only), exact band Cholesky
    Band K;
    K.init(ns, bw);
    std::vector<double> an_diag(ns, 0.0), tx(ns, 0.0), ty(ns, 0.0), tz(ns, 0.0);
    for (int lz = 0; lz < nz; ++lz)
      for (int ly = 0; ly < ny; ++ly)
        for (int lx = 0; lx < nx; ++lx) {
          const int li = (lz * ny + ly) * nx + lx;
          const size_t c = gc[li];
          if (lx + 1 < nx) {
            tx[li] = Tx[c];
            an_diag[li] += Tx[c];
            an_diag[li + 1] += Tx[c];
            K.at(li + 1, li) = -Tx[c];
          }
          if (ly + 1 < ny) {
            ty[li] = Ty[c];
            an_diag[li] += Ty[c];
            an_diag[li + nx] += Ty[c];
            K.at(li + nx, li) = -Ty[c];
          }
          if (lz + 1 < nz) {
            tz[li] = Tz[c];
            an_diag[li] += Tz[c];
            an_diag[li + bw] += Tz[c];
            K.at(li + bw, li) = -Tz[c];
          }
        }
    for (int i = 0; i < ns; ++i) K.at(i, i) = an_diag[i] - kTarget * m[i];
    const bool ok = band_chol_seq(K);
    double msum = 0.0;
    for (int i = 0; i < ns; ++i) msum += m[i];
    std::vector<double> one(ns, 1.0 / std::sqrt(msum));   // M-normalised constant
    std::vector<double> tmp(ns);
    auto op = [&](const double* v, double* w) {
      for (int i = 0; i < ns; ++i) w[i] = m[i] * v[i];
      band_solve_seq(K, w);
    };
    LanczosOut lo;
    if (ok && k > 1)
      lo = lanczos_si(ns, m.data(), op, one.data(), 1, k - 1, pr->eig_tol,
                      std::min(ns - 1, std::max(4 * k + 24, 48)), 0x5eedull + (uint64_t)s * 7919);
    steps_max = std::max(steps_max, lo.steps);
    if (!ok || (k > 1 && !lo.converged)) unconv += 1;
    // outputs: the constant pair, then the Lanczos pairs by ascending lambda
    for (int i = 0; i < ns; ++i) phi[gc[i] * k] = one[i];
    eig[(size_t)s * k] = 0.0;
    for (int j = 1; j < k; ++j) {
      const double* x = j - 1 < (int)lo.theta.size() ? &lo.X[(size_t)(j - 1) * ns] : nullptr;
      // Rayleigh quotient on the original pencil
      double num = 0.0, den = 0.0;
      if (x) {
        for (int li = 0; li < ns; ++li) {
          double ax = an_diag[li] * x[li];
          const int lx = li % nx, ly = (li / nx) % ny, lz = li / bw;
          if (lx > 0) ax -= tx[li - 1] * x[li - 1];
          if (lx + 1 < nx) ax -= tx[li] * x[li + 1];
          if (ly > 0) ax -= ty[li - nx] * x[li - nx];
          if (ly + 1 < ny) ax -= ty[li] * x[li + nx];
          if (lz > 0) ax -= tz[li - bw] * x[li - bw];
          if (lz + 1 < nz) ax -= tz[li] * x[li + bw];
          num += x[li] * ax;
          den += x[li] * m[li] * x[li];
        }
      }
      eig[(size_t)s * k + j] = x ? num / den : 0.0;
      for (int i = 0; i < ns; ++i) phi[gc[i] * k + j] = x ? x[i] : 0.0;
    }
  }
  res->t[1] = now() - t0;
  res->lanczos_steps_max = steps_max;
  res->lanczos_unconverged = unconv;
  res->eig_min_lv1 = 1e300;
  res->eig_max_lv1 = -1e300;
  for (int s = 0; s < nsub; ++s) {
    res->eig_min_lv1 = std::min(res->eig_min_lv1, eig[(size_t)s * k + k - 1]);
    res->eig_max_lv1 = std::max(res->eig_max_lv1, eig[(size_t)s * k + k - 1]);
  }
  if (res->eig_lv1) std::copy(eig.begin(), eig.end(), res->eig_lv1);
  // ---- level 2: Galerkin A_c (energy form) in lower band storage ----
  t0 = now();
  const int nc = nsub * k;
  res->nc = nc;
  const int bwc = sdx * sdy * k + k - 1;
  Band Ac;
  Ac.init(nc, std::min(bwc, nc - 1));
  auto sub_of = [&](size_t c) {
    const int ex = (int)(c % M), ey = (int)((c / M) % N), ez = (int)(c / MN);
    return subx[ex] + sdx * (suby[ey] + sdy * subz[ez]);
  };
#pragma omp parallel for schedule(dynamic, 1)
  for (int s = 0; s < nsub; ++s) {
    int x0, nx, y0, ny, z0, nz;
    corners(s, x0, nx, y0, ny, z0, nz);
    std::vector<double> D((size_t)k * k, 0.0), Ox((size_t)k * k, 0.0), Oy((size_t)k * k, 0.0),
        Oz((size_t)k * k, 0.0);
    int sx_n = -1, sy_n = -1, sz_n = -1;
    for (int lz = 0; lz < nz; ++lz)
      for (int ly = 0; ly < ny; ++ly)
        for (int lx = 0; lx < nx; ++lx) {
          const size_t c = (z0 + lz) * MN + (size_t)(y0 + ly) * M + x0 + lx;
          const double* pc = &phi[c * k];
          auto face = [&](size_t c2, double T, bool inside, std::vector<double>& O, int& sn) {
            const double* p2 = &phi[c2 * k];
            if (inside) {
              for (int i = 0; i < k; ++i)
                for (int j = 0; j < k; ++j) D[i * k + j] += T * (pc[i] - p2[i]) * (pc[j] - p2[j]);
            } else {
              // this cell's side of a face to the next subdomain, and the
              // coupling block (rows of the neighbour, columns of s)
              for (int i = 0; i < k; ++i)
                for (int j = 0; j < k; ++j) {
                  D[i * k + j] += T * pc[i] * pc[j];
                  O[i * k + j] -= T * p2[i] * pc[j];
                }
              sn = sub_of(c2);
            }
          };
          if (x0 + lx + 1 < M) face(c + 1, Tx[c], lx + 1 < nx, Ox, sx_n);
          if (y0 + ly + 1 < N) face(c + M, Ty[c], ly + 1 < ny, Oy, sy_n);
          if (z0 + lz + 1 < P) face(c + MN, Tz[c], lz + 1 < nz, Oz, sz_n);
          // lower faces to the previous subdomains: this cell's side
          if (lx == 0 && x0 > 0) {
            const double T = Tx[c - 1];
            for (int i = 0; i < k; ++i)
              for (int j = 0; j < k; ++j) D[i * k + j] += T * pc[i] * pc[j];
          }
          if (ly == 0 && y0 > 0) {
            const double T = Ty[c - M];
            for (int i = 0; i < k; ++i)
              for (int j = 0; j < k; ++j) D[i * k + j] += T * pc[i] * pc[j];
          }
          if (lz == 0 && z0 > 0) {
            const double T = Tz[c - MN];
            for (int i = 0; i < k; ++i)
              for (int j = 0; j < k; ++j) D[i * k + j] += T * pc[i] * pc[j];
          }
          if ((long long)c == pin)
            for (int i = 0; i < k; ++i)
              for (int j = 0; j < k; ++j) D[i * k + j] += tau * pc[i] * pc[j];
        }
    const int r0 = s * k;
    for (int i = 0; i < k; ++i)
      for (int j = 0; j <= i; ++j) Ac.at(r0 + i, r0 + j) = 0.5 * (D[i * k + j] + D[j * k + i]);
    const std::vector<double>* Os[3] = {&Ox, &Oy, &Oz};
    const int sn[3] = {sx_n, sy_n, sz_n};
    for (int d = 0; d < 3; ++d) {
      if (sn[d] < 0) continue;
      const int q0 = sn[d] * k;
      for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) Ac.at(q0 + i, r0 + j) = (*Os[d])[i * k + j];
    }
  }
  res->t[2] = now() - t0;
  // ---- exact coarse factor ----
  t0 = now();
  Band Lc = Ac;
  if (!band_chol_par(Lc)) return 91;
  res->t[3] = now() - t0;
  // ---- level 3 ----
  t0 = now();
  const int k2 = std::min(pr->en_lv2, nc);
  auto cop = [&](const double* v, double* w) {
    std::copy(v, v + nc, w);
    band_solve_par(Lc, w);
  };
  LanczosOut l3 = lanczos_si(nc, nullptr, cop, nullptr, 0, k2, pr->eig_tol,
                             std::min(nc, std::max(6 * k2 + 24, 60)), 0x1a3ull);
  // Phi_cc = Phi V (natural order, [cell][k2]); A_cc = Phi_cc^T A Phi_cc + NIL I
  std::vector<double> pcc((size_t)n * k2);
#pragma omp parallel for schedule(static)
  for (long long c = 0; c < n; ++c) {
    const int s = sub_of(c);
    for (int j = 0; j < k2; ++j) {
      double v = 0.0;
      for (int a = 0; a < k; ++a) v += phi[c * k + a] * l3.X[(size_t)j * nc + s * k + a];
      pcc[c * k2 + j] = v;
    }
  }
  for (int j = 0; j < k2; ++j) res->eig_lv2[j] = j < (int)l3.theta.size() ? 1.0 / l3.theta[j] : 0.0;
  std::vector<double> Acc((size_t)k2 * k2, 0.0), colv(n), acol(n);
  for (int j = 0; j < k2; ++j) {
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < n; ++c) colv[c] = pcc[c * k2 + j];
    spmv(colv.data(), acol.data());
    for (int i = 0; i < k2; ++i) {
      double sum = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : sum)
      for (long long c = 0; c < n; ++c) sum += pcc[c * k2 + i] * acol[c];
      Acc[(size_t)i * k2 + j] = sum;
    }
  }
  for (int i = 0; i < k2; ++i) Acc[(size_t)i * k2 + i] += kNil;
  // LU with partial pivoting
  std::vector<int> piv(k2);
  for (int c = 0; c < k2; ++c) {
    int p = c;
    for (int r = c + 1; r < k2; ++r)
      if (std::fabs(Acc[(size_t)r * k2 + c]) > std::fabs(Acc[(size_t)p * k2 + c])) p = r;
    piv[c] = p;
    if (p != c)
      for (int q = 0; q < k2; ++q) std::swap(Acc[(size_t)c * k2 + q], Acc[(size_t)p * k2 + q]);
    for (int r = c + 1; r < k2; ++r) {
      Acc[(size_t)r * k2 + c] /= Acc[(size_t)c * k2 + c];
      for (int q = c + 1; q < k2; ++q)
        Acc[(size_t)r * k2 + q] -= Acc[(size_t)r * k2 + c] * Acc[(size_t)c * k2 + q];
    }
  }
  res->t[4] = now() - t0;
  // ---- PCG with the V-cycle of PC_apply_vec ----
  t0 = now();
  std::vector<double> cvec(nc), ycc(k2);
  auto lv2 = [&](const double* r, double* y) {   // y += Phi A_c^{-1} Phi^T r
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nc; ++i) cvec[i] = 0.0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int s = 0; s < nsub; ++s) {
      int x0, nx, y0, ny, z0, nz;
      corners(s, x0, nx, y0, ny, z0, nz);
      double acc[64] = {0.0};
      for (int lz = 0; lz < nz; ++lz)
        for (int ly = 0; ly < ny; ++ly) {
          const size_t c0 = (z0 + lz) * MN + (size_t)(y0 + ly) * M + x0;
          for (int lx = 0; lx < nx; ++lx)
            for (int a = 0; a < k; ++a) acc[a] += phi[(c0 + lx) * k + a] * r[c0 + lx];
        }
      for (int a = 0; a < k; ++a) cvec[s * k + a] = acc[a];
    }
    band_solve_par(Lc, cvec.data());
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < n; ++c) {
      const int s = sub_of(c);
      double v = 0.0;
      for (int a = 0; a < k; ++a) v += phi[c * k + a] * cvec[s * k + a];
      y[c] += v;
    }
  };
  auto lv3 = [&](const double* r, double* y) {   // y += Phi_cc A_cc^{-1} Phi_cc^T r
    for (int j = 0; j < k2; ++j) {
      double sum = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : sum)
      for (long long c = 0; c < n; ++c) sum += pcc[c * k2 + j] * r[c];
      ycc[j] = sum;
    }
    for (int c = 0; c < k2; ++c) std::swap(ycc[c], ycc[piv[c]]);
    for (int r2 = 0; r2 < k2; ++r2)
      for (int q = 0; q < r2; ++q) ycc[r2] -= Acc[(size_t)r2 * k2 + q] * ycc[q];
    for (int r2 = k2 - 1; r2 >= 0; --r2) {
      for (int q = r2 + 1; q < k2; ++q) ycc[r2] -= Acc[(size_t)r2 * k2 + q] * ycc[q];
      ycc[r2] /= Acc[(size_t)r2 * k2 + r2];
    }
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < n; ++c) {
      double v = 0.0;
      for (int j = 0; j < k2; ++j) v += pcc[c * k2 + j] * ycc[j];
      y[c] += v;
    }
  };
  std::vector<double> rr(n), tt(n);
  auto apply = [&](const double* x, double* y) {   // y = B x
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < n; ++c) y[c] = 0.0;
    auto resid = [&]() {
      spmv(y, rr.data());
#pragma omp parallel for schedule(static)
      for (long long c = 0; c < n; ++c) rr[c] = x[c] - rr[c];
    };
    resid();
    lv1_solve(rr.data(), tt.data());
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < n; ++c) y[c] += tt[c];
    resid();
    lv2(rr.data(), y);
    resid();
    lv3(rr.data(), y);
    resid();
    lv2(rr.data(), y);
    resid();
    lv1_solve(rr.data(), tt.data());
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < n; ++c) y[c] += tt[c];
  };
  std::vector<double> r(pr->rhs, pr->rhs + n), z(n), p(n), w(n);
  double* x = res->x;
  std::fill(x, x + n, 0.0);
  auto dot = [&](const double* a, const double* b) {
    double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
    for (long long c = 0; c < n; ++c) s += a[c] * b[c];
    return s;
  };
  double dp = std::sqrt(dot(r.data(), r.data()));
  res->rnorm0 = dp;
  const double ttol = std::max(pr->rtol * dp, 1e-50);
  int its = 0, reason = 0;
  if (dp <= ttol) reason = 2;
  double beta = 0.0, betaold = 0.0;
  if (!reason) {
    apply(r.data(), z.data());
    beta = dot(z.data(), r.data());
  }
  while (!reason) {
    ++its;
    if (its == 1) {
      std::copy(z.begin(), z.end(), p.begin());
    } else {
      const double bb = beta / betaold;
#pragma omp parallel for schedule(static)
      for (long long c = 0; c < n; ++c) p[c] = z[c] + bb * p[c];
    }
    spmv(p.data(), w.data());
    const double dpi = dot(p.data(), w.data());
    betaold = beta;
    const double a = beta / dpi;
    double r2 = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : r2)
    for (long long c = 0; c < n; ++c) {
      x[c] += a * p[c];
      r[c] -= a * w[c];
      r2 += r[c] * r[c];
    }
    dp = std::sqrt(r2);
    if (dp <= ttol) {
      reason = 2;
      break;
    }
    if (its >= pr->max_it) {
      reason = -3;
      break;
    }
    apply(r.data(), z.data());
    beta = dot(z.data(), r.data());
  }
  res->its = its;
  res->reason = reason;
  res->rnorm = dp;
  res->t[5] = now() - t0;
  res->t[6] = now() - t_all;
  return 0;
}

}  // extern "C"
