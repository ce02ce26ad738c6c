/*
 * premixfem_gpu.h -- C ABI of the MI355X implementation of PreMixFEM's
 * three-level GMsFEM preconditioner for the 3D mixed-FEM (RT0/P0,
 * velocity-eliminated) Darcy pressure system.
 *
 * Every entry point replaces one function of the reference's
 * src/PreMixFEM_3D.h (or the PETSc call it relies on); the citation is given
 * above each declaration.  Signatures use plain pointers and sizes only.
 * All entry points return 0 on success or one of the PMX_ERR_* codes, whose
 * values equal the PETSc error codes the reference raises in the same
 * situation (PETSC_ERR_ARG_WRONG for bad options, PETSC_ERR_USER when the
 * eigensolver cannot deliver enough eigenvectors, ...).
 *
 * Host arrays are in the reference's natural DMDA order arr[ez][ey][ex]
 * (x fastest), length M*N*P.  One context drives one GPU (the current HIP
 * device when pmx_init is called); calls on one context must not overlap.
 */
#ifndef PREMIXFEM_GPU_H
#define PREMIXFEM_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes: the PETSc values the reference's PetscCheck/PetscCall use. */
#define PMX_SUCCESS 0
#define PMX_ERR_MEM 55            /* PETSC_ERR_MEM */
#define PMX_ERR_SUP 56            /* PETSC_ERR_SUP: option not implemented */
#define PMX_ERR_ARG_WRONG 62      /* PETSC_ERR_ARG_WRONG (src/PreMixFEM_3D.c:1653-1690) */
#define PMX_ERR_ARG_OUTOFRANGE 63 /* PETSC_ERR_ARG_OUTOFRANGE */
#define PMX_ERR_ORDER 73          /* PETSC_ERR_ORDER: call sequence violated */
#define PMX_ERR_LIB 76            /* PETSC_ERR_LIB: HIP runtime failure */
#define PMX_ERR_USER 83           /* PETSC_ERR_USER: eigensolver shortfall (:367, :1098) */
#define PMX_ERR_ARG_NULL 85       /* PETSC_ERR_ARG_NULL */
#define PMX_ERR_NOT_CONVERGED 91  /* PETSC_ERR_NOT_CONVERGED */

/* Eigen mass operators (src/PreMixFEM_3D.h:28-30). */
#define PMX_EIG_OP_MOD 0
#define PMX_EIG_OP_MEAN 1
#define PMX_EIG_OP_DIAG 2

/* Stage timer ids (src/PreMixFEM_3D.h:16-26). */
#define PMX_STAGE_SU_LV1 0
#define PMX_STAGE_SU_LV2_1 1
#define PMX_STAGE_SU_LV2_2 2
#define PMX_STAGE_SU_LV3_1 3
#define PMX_STAGE_SU_LV3_2 4
#define PMX_STAGE_AV_LV1 5
#define PMX_STAGE_AV_LV2 6
#define PMX_STAGE_AV_LV3 7
#define PMX_STAGE_SU 8
#define PMX_STAGE_AV 9
#define PMX_MAX_LOG_STAGES 10

/* PCG convergence reasons (PETSc KSPConvergedReason values). */
#define PMX_CONVERGED_RTOL 2
#define PMX_CONVERGED_ATOL 3
#define PMX_DIVERGED_ITS (-3)
#define PMX_DIVERGED_DTOL (-4)
#define PMX_DIVERGED_INDEFINITE_PC (-8)
#define PMX_DIVERGED_NANORINF (-9)
#define PMX_DIVERGED_INDEFINITE_MAT (-10)

typedef struct pmx_ctx pmx_ctx;

/* Statistics PC_print_stat reports (src/PreMixFEM_3D.c:1938-1999), plus the
 * eigensolver and outer-solver counters of this implementation. */
typedef struct {
  int M, N, P;
  int sd[3];
  int nsub;
  int size_coarse;           /* level-2 coarse size (sum of eigen counts) */
  int size_mat_lv3;          /* level-3 size (eigen_num_lv2) */
  int eigen_num_lv1_range[2];
  double eigen_min_lv1_range[2];
  double eigen_max_lv1_range[2];
  int eigen_num_lv2;
  double eigen_min_lv2, eigen_max_lv2;
  double t_stages[PMX_MAX_LOG_STAGES]; /* seconds, same ids as the reference */
  int eig_iters_max;         /* LOBPCG iterations, worst subdomain */
  double eig_iters_mean;
  int eig_fallbacks;         /* subdomains restarted from a random block */
  int lv3_iters;
  int pc_precision;          /* 64 or 32 */
  int its;                   /* last pmx_solve iteration count */
  double rnorm, rnorm0;      /* last pmx_solve final / initial residual norms */
  int reason;                /* last pmx_solve convergence reason */
  double t_solve;            /* last pmx_solve wall seconds (PCG iterations only) */
  int eig_stagnated;         /* subdomains accepted at the eigensolver's stagnation test
                                (worst residual within -eig_stag x -eig_tol) */
  int eig_block;             /* LOBPCG block size used (chosen by -eig_autotune 1) */
  int coarse_mode;           /* level-2/3 path of the apply: 0 dense composite operator
                                (one GEMV), 1 dense stage by stage, 2 band Cholesky,
                                3 ICC(0) */
  int coarse_stages;         /* coarse stages run per apply: the level-2/3 stages of the
                                cycle, or 1 when -coarse_collapse cut them to the first
                                level-2 stage (one process, exact level-2 factor) */
  double lv3_floor;          /* level-3 eigensolver: the residual it stagnated at and
                                accepted (0 when it met -eig_tol) */
  double coarse_factor_bytes; /* stored bytes of the level-2 factor (dense inverse or
                                 composite operator, band tiles, plane inverses, ICC(0)) */
  double coarse_solve_bytes;  /* bytes one level-2 solve reads from that factor */
  double t_eig_tune;          /* seconds of the last setup spent timing LOBPCG block sizes
                                 (-eig_autotune 1; 0 when the process had them cached) */
  int eig_shape;              /* matrix-core LOBPCG workgroup shape used (-eig_shape, or
                                 chosen by -eig_autotune 1): 0 one subdomain per CU,
                                 1 two per CU */
  double t_eig_kernel;        /* seconds of the last setup's subdomain LOBPCG launch, from
                                 HIP events on the library stream (-kernel_timing; else 0) */
} pmx_stats;

/* PC_init (src/PreMixFEM_3D.c:1650-1741): domain lengths dom[3] (L, W, H),
 * cells mesh[3] (M, N, P) and an options string with the reference's option
 * names (-sd -en_lv1 -en_lv2 -si_lv1 -si_lv2 -eg_op -eb_lv1 -eb_lv2
 * -use_W_cycle -no_shift_A_cc -use_full_Cholesky_lv1 -use_2level), plus
 * -sdx/-sdy/-sdz (per-direction subdomains), -pin x,y,z, -pc_precision 32|64,
 * -eig_tol, -eig_maxit, -eig_stag, -warm_start none|cnn|basis, -seed, -cg_graph,
 * -cg_chunk, -kernel_timing, and -lv2_icc (the spe10 fork's level-2 factor:
 * ICC(0) of the coarse matrix, src/spe10.c:509-516; its level-3 eigenvectors
 * come from ICC(0)-preconditioned block LOBPCG, started from the previous
 * setup's unless -lv3_warm 0), -coarse_band 0|1 (exact level-2 Cholesky
 * factor in band storage; automatic above 12288 coarse unknowns),
 * -eig_guard g (LOBPCG guard vectors), -eig_autotune 0|1 (time the LOBPCG
 * block sizes, and the workgroup shapes of the matrix-core kernel, on a
 * sample of subdomains and keep the fastest, cached per subdomain size and
 * eigen count for the process), -eig_kernel 0|1 (the
 * subdomain LOBPCG kernel: 1, default, Grams and block transforms on the FP64
 * matrix cores with fused passes; 0, the VALU kernel; same eigenpairs to the
 * eigensolver tolerance), -eig_shape 0|1 (matrix-core kernel: one
 * 512-thread workgroup per CU on large subdomains, default, or two 256-thread
 * ones per CU; same eigenpairs), -eig_ortho 0|1 (matrix-core kernel: W
 * made M-orthogonal to [X P] and M-orthonormal from one Gram of [X P W] in
 * coefficient space, default, or by two explicit projection passes, not
 * available for subdomains of at most 512 cells or in pmx_setup_batch; same
 * eigenpairs to the eigensolver tolerance),
 * -coarse_collapse 0|1 (one process with an exact level-2 factor: apply only
 * the first level-2 stage of the coarse cycle, since the later level-3 and
 * level-2 stages add exactly zero in exact arithmetic; default 1),
 * -coarse_composite 0|1 (dense path: compose the coarse stages into one
 * operator, default, or apply them one by one), -coarse_deflate 0|1 (dense
 * path: invert A_c deflated with a Woodbury correction, default, or
 * directly), -quad 0|1 (quad-cell stream kernels where every brick's
 * x-extent is a multiple of 4, default 1), -procs_virtual px,py,pz (the
 * preconditioner of px*py*pz PETSc processes on one GPU: process boxes by
 * PETSc's default DMDA ownership (src/PreMixFEM_3D.c:1711), -sd subdomains
 * per axis inside every box (:58-72), level 2 coupled inside a box only
 * (:806-888), en_lv2 level-3 vectors per box from its box-local Neumann
 * coarse matrix (:918-1175), A_cc over all boxes (:1177-1390); dense coarse
 * path, at most 64 level-3 vectors in total), -eig_order 0|1 (launch the
 * subdomain eigenproblems of a repeated setup hardest first by the previous
 * setup's iteration counts, default 1; results do not depend on it),
 * -batch_threads t (pmx_setup_batch host threads, default 8).  Allocates the
 * context and selects the current HIP device. */
int pmx_init(pmx_ctx **ctx, const double *dom, const int *mesh, const char *options);

/* PC_print_info (src/PreMixFEM_3D.c:1743-1755). */
int pmx_print_info(const pmx_ctx *ctx);

/* Caller fills s_ctx.kappa[0..2] (src/Expe_scal.c:11-42): the three
 * permeability components, host arrays of M*N*P doubles. */
int pmx_set_kappa(pmx_ctx *ctx, const double *kx, const double *ky, const double *kz);

/* pmx_set_kappa with device arrays (natural order, M*N*P doubles each), as
 * the spe10 fork's INTER_upd_kappa (src/spe10.c:620-648) refreshes kappa
 * before every pressure solve. */
int pmx_set_kappa_device(pmx_ctx *ctx, const double *d_kx, const double *d_ky,
                         const double *d_kz);

/* The producer-well term of the spe10 fork's pressure system: a per-cell
 * diagonal addition (|E| perm_x lambda_t WI at the producer columns,
 * src/spe10.c:780-786) that the operator, the level-1 blocks (:921-927),
 * the eigen stiffness and the DIAG eigen mass (:1005-1012) and, through A,
 * the coarse matrices (:1128-1131) all include.  wd: M*N*P non-negative
 * doubles, natural order; NULL removes the term.  Takes effect at the next
 * pmx_setup. */
int pmx_set_well_term(pmx_ctx *ctx, const double *wd);
int pmx_set_well_term_device(pmx_ctx *ctx, const double *d_wd);

/* Velocity recovery: the RT0 face fluxes F = T (p_L - p_R) of a pressure
 * field p (the face coefficient T of PC_create_A, src/PreMixFEM_3D.c:1785-1810,
 * used by the spe10 transport at src/spe10.c:1977-2034).  fx[c] is the flux
 * from the -x neighbour into cell c across its lower x face (0 on the domain
 * boundary), likewise fy, fz.  All arrays natural order, M*N*P doubles.  The
 * outflow (fx[c+x]-fx[c]) + (fy[c+y]-fy[c]) + (fz[c+z]-fz[c]) of every cell
 * equals (A p)_c without the pin and well terms. */
int pmx_velocity(pmx_ctx *ctx, const double *p, double *fx, double *fy, double *fz);
int pmx_velocity_device(pmx_ctx *ctx, const double *d_p, double *d_fx, double *d_fy,
                        double *d_fz);

/* PC_create_A (src/PreMixFEM_3D.c:1757-1824): the 7-point operator as seven
 * diagonals in natural order: diag[c] and the couplings to the -x,+x,-y,+y,
 * -z,+z neighbours (off[6*c + d], zero where there is no neighbour).  The
 * pin term, if requested, is included in diag. */
int pmx_create_A(pmx_ctx *ctx, double *diag, double *off);

/* PC_setup (src/PreMixFEM_3D.c:1826-1838 -> _PC_setup :1392-1427): builds
 * level 1 (block ICC(0)), level 2 (per-subdomain eigenproblems by batched
 * LOBPCG, Galerkin coarse matrix and its FP64 factorisation) and level 3. */
int pmx_setup(pmx_ctx *ctx);

/* PC_setup of nprob independent problems (distinct contexts, each with its
 * own mesh, kappa and options) in one call: level 1 and the coarse levels
 * are built per context as pmx_setup does, while the subdomain eigenproblems
 * of ALL problems are solved by ONE batched LOBPCG launch (one workgroup per
 * subdomain of any problem).  A brick's random start depends only on its
 * index inside its own problem, so every context ends up with the basis its
 * own pmx_setup computes.  Requires equal en_lv1, -eig_* options and
 * warm-start mode; under -eig_autotune the block size is tuned on the
 * problem with the most subdomains and used for all.  The coarse levels of
 * the problems are then built concurrently by host threads (the first
 * context's -batch_threads, default 8), each on its own context's stream.
 * The reference sets up one problem per PC_setup call
 * (src/PreMixFEM_3D.c:1826-1838); a caller with many problems calls it once
 * per problem. */
int pmx_setup_batch(pmx_ctx *const *ctxs, int nprob);

/* PC_setup with the level-2 eigenvectors supplied by the caller (host,
 * natural order, en_lv1 columns of M*N*P doubles each, column j of every
 * subdomain in vector j; eigen_num per subdomain, may be NULL = en_lv1).
 * Replaces _PC_setup_lv2_eigen (:224-421) by the given basis; everything
 * else as pmx_setup.  Used for restarts from a previous basis. */
int pmx_setup_from_basis(pmx_ctx *ctx, const double *phi, const int *eigen_num,
                         const double *eigvals);

/* PC_apply_vec (src/PreMixFEM_3D.c:1840-1936): y = B x, host arrays. */
int pmx_apply(pmx_ctx *ctx, const double *x, double *y);

/* PETSc KSPSolve with -ksp_type cg, KSP_NORM_UNPRECONDITIONED and the
 * PCSHELL above (src/Expe_scal.c:101-119), run entirely on the GPU.
 * b, x: host arrays.  x0_nonzero selects the spe10-style initial guess
 * (src/spe10.c:191).  On return *its and *rnorm hold the iteration count and
 * the final unpreconditioned residual norm. */
int pmx_solve(pmx_ctx *ctx, const double *b, double *x, int x0_nonzero, double rtol,
              int max_it, int *its, double *rnorm);

/* pmx_solve with b and x in device memory (natural order, length M*N*P):
 * the same PCG without the host<->device copies of the vectors. */
int pmx_solve_device(pmx_ctx *ctx, const double *d_b, double *d_x, int x0_nonzero, double rtol,
                     int max_it, int *its, double *rnorm);

/* pmx_solve_device without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The PCG's per-subdomain partial sums and device scalars live in a fixed
 * device array of the library (PMX_NOALLOC_SCRATCH_DOUBLES doubles, 4 MiB,
 * held on every GPU the library runs on until the process ends) instead of
 * the context's buffers; the iterations are launched eagerly (no HIP graph,
 * no -kernel_timing events).  The context must have been set up by
 * pmx_setup, which allocates the preconditioner and the PCG vectors; this
 * call then calls no hipMalloc / hipFree.  Limit: 6 * nsub + 16 doubles of
 * scratch, i.e. at most 87378 subdomains (PMX_ERR_ARG_OUTOFRANGE beyond).
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (from several threads or streams) race.  Results are those of
 * pmx_solve_device. */
#define PMX_NOALLOC_SCRATCH_DOUBLES 524288
int pmx_solve_device_noalloc(pmx_ctx *ctx, const double *d_b, double *d_x, int x0_nonzero,
                             double rtol, int max_it, int *its, double *rnorm);

/* Per-kernel GPU time of the solve phase, measured with HIP events on the
 * context's stream when the context was created with -kernel_timing.
 * Kernels, in order: 0 level-1 first sweep, 1 level-1 last sweep (with the
 * residual), 2 residual + restriction, 3 prolongation, 4 PCG p-update +
 * SpMV, 5 PCG x/r update, 6 coarse dense solve (GEMV), 7 coarse small ops.
 * ms[i] receives the summed milliseconds, count[i] the launch count, for
 * i < n (n <= PMX_NUM_TIMED_KERNELS).  reset clears the counters after
 * reading. */
#define PMX_NUM_TIMED_KERNELS 8
int pmx_get_kernel_times(pmx_ctx *ctx, double *ms, int *count, int n, int reset);

/* PC_print_stat / PC_get_range (src/PreMixFEM_3D.c:1938-1999). */
int pmx_get_stats(const pmx_ctx *ctx, pmx_stats *st);
int pmx_print_stat(const pmx_ctx *ctx);

/* The ms_bases_c / ms_bases_cc vectors of PCCtx (src/PreMixFEM_3D.h:34):
 * level 2 (level=2): en_lv1 vectors of M*N*P doubles (zero outside each
 * subdomain; vector j holds eigenvector j of every subdomain); level 3
 * (level=3): eigen_num_lv2 vectors.  eigvals (optional) receives the
 * eigenvalues ([nsub][en_lv1] for level 2, [eigen_num_lv2] for level 3). */
int pmx_get_basis(const pmx_ctx *ctx, int level, double *out, double *eigvals);

/* The level-2 Galerkin matrix Phi^T A Phi (dense, size_coarse^2, row-major). */
int pmx_get_coarse_matrix(const pmx_ctx *ctx, double *out);

/* PC_final / PC_final_default (src/PreMixFEM_3D.c:2001-2046). */
int pmx_final(pmx_ctx *ctx);

/* ---- Learned warm start of the level-2 eigensolver ----------------------
 * No counterpart in the reference: a 3-layer 3D CNN (3x3x3 convolutions,
 * 4 -> 16 -> 16 -> cout channels, BF16 weights, FP32 accumulation on the
 * matrix cores) maps a subdomain's log-permeability patch and cell
 * coordinates to an initial block for its eigenproblem (the eigensolver of
 * src/PreMixFEM_3D.c:224-421); LOBPCG refines it to the same tolerance and
 * restarts from a random block where it fails, so the basis is unchanged.
 * Parameters: pmx_cnn_num_params() FP32 values, per layer W[16][27][16]
 * (co, tap = (dz+1)*9+(dy+1)*3+(dx+1), ci) then bias[16].  Activations are
 * device arrays [cell][16] of BF16 (uint16) in brick order, uniform batches
 * of nb bricks of nx*ny*nz cells (brick b starts at cell b*nx*ny*nz). */
typedef struct pmx_cnn pmx_cnn;
int pmx_cnn_num_params(void);
/* params (host) may be NULL: seeded He-uniform initialisation. */
int pmx_cnn_create(pmx_cnn **net, int cout, const float *params, unsigned long long seed);
int pmx_cnn_get_params(const pmx_cnn *net, float *params, float *grad);
int pmx_cnn_set_params(pmx_cnn *net, const float *params);
/* d_out: [cell][16] FP32 device array (channels >= cout are 0). */
int pmx_cnn_predict(pmx_cnn *net, int nb, int nx, int ny, int nz, const unsigned short *d_in,
                    float *d_out);
/* Forward + backward against d_target ([cell][16] FP32); the gradient is
 * kept in the network (pmx_cnn_get_params), the loss returned in *loss.
 * loss_kind 0: MSE sum_{c<cout} (out - target)^2 / (ncell*cout);
 * loss_kind 1: subspace distance per brick, sum_b (cout - tr(C^T (G + dI)^{-1}
 * C)) / (nb*cout) with G = Y^T Y, C = Y^T U/sqrt(n) (Y = the first cout
 * output channels, U = targets with orthogonal unit-RMS columns, d = 1e-6 n):
 * invariant to sign and rotation of the eigenvectors. */
int pmx_cnn_loss_grad(pmx_cnn *net, int nb, int nx, int ny, int nz, const unsigned short *d_in,
                      const float *d_target, int loss_kind, double *loss);
int pmx_cnn_adam_step(pmx_cnn *net, float lr, float beta1, float beta2, float eps);
int pmx_cnn_destroy(pmx_cnn *net);
/* Training data from a context: the network input of every brick (after
 * pmx_set_kappa) and the normalised level-2 eigenvectors 1..cout (after
 * pmx_setup), [ncell][16] device arrays in brick order. */
int pmx_cnn_inputs(pmx_ctx *ctx, unsigned short *d_in);
int pmx_cnn_targets(pmx_ctx *ctx, int cout, float *d_target);
/* Network used by pmx_setup under -warm_start cnn (not owned; keep it alive
 * while the context uses it). */
int pmx_set_warm_start(pmx_ctx *ctx, pmx_cnn *net);

/* ---- The spe10 fork's reservoir loop (src/spe10.c) ----------------------
 * Two-phase IMPES on an M x N x P grid of Hx x Hy x Hz cells: producers at
 * the four corner columns, the injector at the four centre columns, Corey
 * relative permeabilities (exponent 2), mu_o / mu_w = 10 (SPE10_init,
 * :345-405).  The preconditioner context used with it must have the same
 * mesh and dom = (M Hx, N Hy, P Hz); the reference's settings are
 * "-eg_op 2 -no_shift_A_cc -lv2_icc" (DIAG mass, no NIL shift, ICC(0) level-2
 * factor, :509-530). */
typedef struct pmx_spe10 pmx_spe10;

/* One row of the P_solver.csv log (src/spe10.c:206-252) plus timings. */
typedef struct {
  int step;            /* simu_time_ind */
  double simu_time;
  int stage;           /* stage_ind */
  double DS_max;
  int S_w_steps;
  int its;             /* PCG iterations of the pressure solve */
  int reason;          /* its convergence reason */
  double S_w_min, S_w_max;
  double delta_t;      /* time advanced by the transport substeps */
  double t_assemble, t_setup, t_solve, t_transport;  /* wall seconds */
} pmx_spe10_log;

/* SPE10_init + INTER_init + get_modified_poro + get_init_S_w (:345-405,
 * :569-590, :2065-2101): perm_{x,y,z} (the reference's units after its
 * 10.7639 scaling) and poro, host arrays in natural order; S_w = S_wi,
 * P = 0. */
int pmx_spe10_create(pmx_spe10 **sp, const int *mesh, const double *H, const double *perm_x,
                     const double *perm_y, const double *perm_z, const double *poro);
/* INTER_upd_kappa + get_linear_system (:620-648, :723-800): kappa and the
 * well term go into ctx (device to device), the rhs stays in sp. */
int pmx_spe10_assemble(pmx_spe10 *sp, pmx_ctx *ctx);
/* The assembled fields as host copies (any pointer may be NULL). */
int pmx_spe10_get_system(pmx_spe10 *sp, double *kx, double *ky, double *kz, double *wd,
                         double *rhs);
/* KSPSetUp (PC_setup) + KSPSolve with the previous pressure as the initial
 * guess (KSPSetInitialGuessNonzero, :191), all on the GPU. */
int pmx_spe10_solve(pmx_spe10 *sp, pmx_ctx *ctx, double rtol, int max_it, int *its,
                    double *rnorm);
/* get_next_S_w_multisteps (:1942-2063): `steps` explicit upwind substeps of
 * dt = DS_max / max(rate); *delta_t receives their sum. */
int pmx_spe10_transport(pmx_spe10 *sp, int steps, double DS_max, double *delta_t);
/* VecMin / VecMax of S_w (:248-249). */
int pmx_spe10_saturation_range(pmx_spe10 *sp, double *smin, double *smax);
int pmx_spe10_get_state(const pmx_spe10 *sp, double *S_w, double *P);
int pmx_spe10_set_state(pmx_spe10 *sp, const double *S_w, const double *P);
/* The IMPES loop of main (:197-274): up to max_steps of assemble, setup,
 * solve, S_w bounds check (PMX_ERR_USER outside [0.2, 0.8], :250) and
 * transport, with the stage tables simu_time_stages[7], DS_max_stages[7],
 * steps_stages[7] (:117-119).  log (may be NULL) receives one row per step;
 * *steps_done the number of steps run. */
int pmx_spe10_run(pmx_spe10 *sp, pmx_ctx *ctx, int max_steps, const double *simu_time_stages,
                  const double *DS_max_stages, const int *steps_stages, double rtol, int max_it,
                  pmx_spe10_log *log, int *steps_done);
/* Per-step output hook of pmx_spe10_run: called after the pressure solve
 * and the bounds check, before the transport of each step (where the
 * reference writes its log row and the S_w / P snapshots, :246-271), with
 * the step's log row.  A nonzero return stops the loop with that code.
 * cb == NULL removes the hook. */
typedef int (*pmx_spe10_step_fn)(void *user, const pmx_spe10_log *row);
int pmx_spe10_set_step_hook(pmx_spe10 *sp, pmx_spe10_step_fn cb, void *user);
int pmx_spe10_destroy(pmx_spe10 *sp);

/* Last error message of this thread (empty if none). */
const char *pmx_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* PREMIXFEM_GPU_H */
