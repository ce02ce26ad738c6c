// Kernel launch wrappers shared between the translation units.
#pragma once

#include <algorithm>

#include "pmx_internal.h"

namespace pmx {

size_t brick_lds_bytes(const pmx_ctx* ctx, size_t elem);

// fine grid (pmx_fine.hip)
int launch_nat2brick(pmx_ctx* ctx, const double* src, double* dst, int nvec, int sstride,
                     int dstride);
int launch_brick2nat(pmx_ctx* ctx, const double* src, double* dst, int nvec, int sstride,
                     int dstride);
int launch_basis_nat2rows(pmx_ctx* ctx, const double* src, double* rows, int k);
int launch_basis_rows2nat(pmx_ctx* ctx, const double* rows, double* dst, int k);
int launch_trans(pmx_ctx* ctx);
int launch_mass(pmx_ctx* ctx);
int launch_dic_factor(pmx_ctx* ctx, int mode, double sigma, void* fac_f, void* fac_b, bool f32);
int launch_lv1_first(pmx_ctx* ctx, double* x, bool upd, double* xs, const double* p,
                     const double* w);
int stream_threads(const pmx_ctx* ctx);
int quad_threads(const pmx_ctx* ctx);
int chol_lv1_setup(pmx_ctx* ctx);
int launch_lv1_last(pmx_ctx* ctx, const double* x, double* z, bool keep_y, bool cg_partial);
int launch_resid_restrict(pmx_ctx* ctx, const double* x, bool y_zero, double* cvec);
int launch_prolong(pmx_ctx* ctx, const double* cvec, bool count_iter);
int launch_cg_p_spmv(pmx_ctx* ctx, const double* z, const double* pold, double* pnew, double* w);
int launch_residual(pmx_ctx* ctx, const double* b, const double* x, double* r);
int launch_spmv(pmx_ctx* ctx, const double* x, double* y);
int launch_to_f32(pmx_ctx* ctx, const double* src, float* dst, size_t n);
int launch_velocity(pmx_ctx* ctx, const double* p, double* fx, double* fy, double* fz);
int launch_check_field(pmx_ctx* ctx, const double* a, int n, bool strict, int* bad);

// coarse levels (pmx_coarse.hip)
int launch_galerkin(pmx_ctx* ctx);
int coarse_invert(pmx_ctx* ctx, const double* A, double* Ainv, double shift);
// Inverse of a dense row-major SPD n x n matrix (+ shift I), on stream.
int spd_invert(hipStream_t stream, int n, const double* A, double* Ainv, double shift);
int coarse_solve_setup(pmx_ctx* ctx, double delta);
int launch_coarse_solve(pmx_ctx* ctx, const double* t, double* d, double* ytmp, int ncol);
int launch_coarse_fix(pmx_ctx* ctx, const double* y, const double* t, double* d);
int coarse_operator_setup(pmx_ctx* ctx, const std::vector<int>& prog);
int launch_coarse_program(pmx_ctx* ctx, const std::vector<int>& prog, double* t, double* u,
                          double* y, double* d, int kt_gemv, int kt_small,
                          int (*kt_begin)(pmx_ctx*, int), void (*kt_end)(pmx_ctx*, int));
int launch_resid_restrict_f64(pmx_ctx* ctx, const double* x, double* cvec);
int level3_setup(pmx_ctx* ctx);
int level3_setup_boxes(pmx_ctx* ctx);
int launch_gemv(pmx_ctx* ctx, const double* A, const double* x, double* y, int n);
int launch_coarse_update(pmx_ctx* ctx, const double* d, double* t, double* u, int first,
                         int update_t);
int launch_lv3_apply(pmx_ctx* ctx, const double* t, double* d);
// level 2 of the spe10 fork: block ICC(0) of A_c
int bicc_setup(pmx_ctx* ctx);
int launch_bicc_solve(pmx_ctx* ctx, const double* t, double* d, int ncol);
int level3_setup_icc(pmx_ctx* ctx);
// exact level-2 factor in band storage (pmx_band.hip)
int band_setup(pmx_ctx* ctx);
int launch_band_solve(pmx_ctx* ctx, const double* t, double* d, int ncol);
// exact level-2 factor by z-planes of subdomains (pmx_planes.hip)
int planes_setup(pmx_ctx* ctx);
int launch_planes_solve(pmx_ctx* ctx, const double* t, double* d, int ncol);
// level-2 solve of the sparse coarse paths: ICC(0) sweeps, the band or the plane factor
int launch_coarse_prec(pmx_ctx* ctx, const double* t, double* d, int ncol);

// level-2 eigensolver (pmx_lobpcg.hip)
int lobpcg_setup(pmx_ctx* ctx, const double* x0_rows);
int lobpcg_block_size(int en_lv1, int guard);
int lobpcg_autotune(pmx_ctx* ctx);
int lobpcg_batch(pmx_ctx** ctxs, int nprob);

// learned warm start (pmx_cnn.hip): x0 rows [cell][mb] in brick order
int cnn_warm_start(pmx_ctx* ctx, int mb, double* x0);

}  // namespace pmx
