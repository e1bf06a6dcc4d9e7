/*
 * gputo.h -- C ABI of libgputo.so, the MI355X-native top3D125 MGCG
 * elasticity solver (matrix-free hex8 operator, geometric multigrid
 * preconditioned CG, topology-optimisation driver kernels).
 *
 * Plain C types only: host pointers are double / int arrays laid out as the
 * reference's MEX receives them (column vectors, DOF = 3*node + d, node =
 * (nely+1)*(nelz+1)*i + (nely+1)*k + j, 0-based).  Entry points whose name
 * ends in _d take device pointers (hipMalloc'd by the caller) instead.
 *
 * Every entry point returns 0 on success or a negative gputo_status.
 * A handle is not thread-safe; independent handles may be used from
 * different threads.  Each handle owns one HIP stream.
 */
#ifndef GPUTO_H
#define GPUTO_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gputo_handle gputo_handle;

/* Status codes.  The reference reports only nit = -1 (workspace shortfall,
 * Cantilever/omp/mgcg1.c:91-97) and otherwise calls exit(); here every
 * failure is a return code and nit < 0 carries the same code. */
enum gputo_status {
    GPUTO_OK = 0,
    GPUTO_ERR_WORKSPACE = -1,    /* device memory / coarsest size limit */
    GPUTO_ERR_NOT_SPD = -2,      /* coarsest matrix not positive definite */
    GPUTO_ERR_NONFINITE = -3,    /* NaN/Inf in an input or in the iteration */
    GPUTO_ERR_DEVICE = -4,       /* HIP runtime error */
    GPUTO_ERR_ARGUMENT = -5      /* invalid sizes or arguments */
};

/* Multigrid presets.
 *  GPUTO_PRESET_REF_C: constants of the reference C/CUDA MEX
 *    (pre-relaxation U = 0.7 F/AD, 3 pre and 3 post damped-Jacobi sweeps
 *    U = 0.52 U + 0.48 (R + AD U)/AD, level 1 rediscretised from child
 *    densities, levels >= 2 element Galerkin, coarsest 1e10 diagonal
 *    penalty with the reference's node table;
 *    Cantilever/omp/mgcg1.c:110,308,353,363-364, mgcg1.cu:188,1196,1205).
 *  GPUTO_PRESET_MATLAB_CFG0: the MATLAB top3D125mgcg smoother (omega 0.6,
 *    one sweep down from zero, one sweep up; Cantilever/TOP3D125MGCG/
 *    top3D125mgcg.m:190-246) on element-Galerkin operators at every
 *    coarse level.
 *  GPUTO_PRESET_MGCG9: the mgcg9 engine of the CrossWalk and QNC drivers
 *    (Homogenization/QNC/mgcg9.cu): every level matrix-free with the unit
 *    Ke and element scales CX_l = 2 * mean of the children (formCXc,
 *    :753-789); the smoother's defect skips elements with CX < 1e-8
 *    (defect0c, :939-996) while the CG matvec keeps them; pre-relaxation
 *    U = 0.7 F/AD and `nswp` damped-Jacobi sweeps (omega 0.33) each way
 *    (:416-635, Jacobc :929-937); a coarsest level below 20000 DOFs is
 *    factored with its fixed rows and columns zeroed and 1e10 on their
 *    diagonal (CorrCsAQc, :1005-1028), a larger one gets 10*nswp sweeps. */
enum gputo_preset { GPUTO_PRESET_REF_C = 0, GPUTO_PRESET_MATLAB_CFG0 = 1, GPUTO_PRESET_MGCG9 = 2 };

/* Arithmetic of the solve.
 *  GPUTO_PREC_F64: everything in FP64 (the reference's arithmetic).
 *  GPUTO_PREC_MIXED_F32: the PCG recurrences, the level-0 matvec, dot
 *    products, alpha/beta and the coarsest Cholesky stay in FP64; the
 *    preconditioner (the whole V-cycle) runs on FP32 copies of the hierarchy
 *    with FP32 accumulation, so rounding changes only the iteration count.
 *  GPUTO_PREC_MIXED_BF16: as MIXED_F32 with the V-cycle's fine-level
 *    Ke x u contraction on BF16 MFMA (split-BF16 operands, FP32
 *    accumulation). */
enum gputo_precision { GPUTO_PREC_F64 = 0, GPUTO_PREC_MIXED_F32 = 1, GPUTO_PREC_MIXED_BF16 = 2 };

typedef struct gputo_params {
    int preset;        /* enum gputo_preset */
    int precision;     /* enum gputo_precision */
    int learned;       /* learned level-2 correction inside flexible CG: 0 none; 1 applied in
                          every V-cycle once the network is trained; 2 the same with a
                          solve-level fallback: successive solves are timed, and the stage
                          is kept only while a solve with it is faster than the last one
                          without it (failed trials back off 4 .. 64 solves; a kept stage
                          is re-checked every 8 solves) */
    int print_level;   /* 0 silent; >=1 summary on stderr; >=2 per-iteration */
    int l0_kernel;     /* level-0 operator of the FP32 V-cycle: 0 by precision, 1 FP32
                          node-gather, 2 FP32 MFMA, 3 split-BF16 MFMA (hi+lo),
                          4 split-BF16 MFMA (hi+mid+lo), 5 single-product
                          BF16 MFMA (smoother-grade accuracy only) */
    int nswp;          /* GPUTO_PRESET_MGCG9: Jacobi sweeps each way (0: 2, the MEX default) */
    int l1_kernel;     /* level-1 operator in the V-cycle: 0 by preset (matrix-free
                          where level 1 is rediscretised from the child densities,
                          stored element matrices otherwise), 1 stored element
                          matrices, 2 matrix-free (rediscretised level 1 only) */
    int coarse_lag;    /* coarsest factorisation: 0 exact (the reference algorithm),
                          1 lagged: each setup factors its coarsest matrix on a
                          side stream while the solve uses the factor of the
                          previous setup (the first setup, and any setup after a
                          parameter change, waits for its own factor).  Lagging
                          hides the factorisation but costs iterations when the
                          design changes fast (early design iterations).
                          n >= 2: the first n V-cycles of a solve use the previous
                          factor, the later ones wait for this setup's (flexible CG;
                          the switch is at a fixed V-cycle, so results do not
                          depend on timing) */
    int periodic;      /* 1: periodic unit cell.  The mesh wraps in x, y and z: the
                          nodes of the faces i = nelx, j = nely, k = nelz are the
                          images of those at 0.  Vectors keep the
                          (nelx+1)(nely+1)(nelz+1)-node layout; a right-hand side
                          may be split between a node and its images (they are
                          summed), a solution holds equal values on them.  Every
                          level needs at least 2 elements per direction; the
                          learned stage is not available */
    int mixed_npre0;   /* FP32 V-cycle of the mixed precisions: Jacobi sweeps on level 0
                          before / after the coarse correction (0: the preset's; -1: none,
                          the level is pre-smoothed by the prerelaxed iterate
                          prerelax*F/AD alone).  The preconditioner changes, not the
                          solution: the FP64 CG still stops at the FP64 tolerance.  With
                          the prerelaxation counted as the first pre-sweep the total must
                          be even (-5 otherwise) */
    int mixed_npost0;
    int mixed_npre1;   /* the same for level 1 */
    int mixed_npost1;
    int mixed_npre2;   /* the same for the levels from 2 down to the last above the coarsest
                          (from level 3 on: mixed_npre3 / mixed_npost3 when set) */
    int mixed_npost2;
    int homog_sequential; /* gputo_homogenize: 0 (default) solves the six load cases
                             together, each operator launch applied to all six
                             vectors; 1 solves them one after the other */
    int graph_vcycles;    /* mixed precisions: 1 replays the FP32 V-cycle of the PCG
                             loop from one recorded HIP graph where its launches
                             do not change (one host launch instead of ~45);
                             0 (default) launches every kernel from the host,
                             measured faster on MI355X.  The results are the
                             same bit for bit */
    int mixed_npre3;      /* FP32 V-cycle: Jacobi sweeps before / after the coarse
                             correction on the levels from 3 down to the last above the
                             coarsest (0: those of mixed_npre2 / mixed_npost2; the same
                             parity rule) */
    int mixed_npost3;
} gputo_params;

/* Fill *p with the defaults (REF_C, F64, no learned stage, operator
 * kernels chosen by precision and preset). */
void gputo_default_params(gputo_params *p);

/* Create a solver for an nelx x nely x nelz hex mesh with nl multigrid
 * levels (each size divisible by 2^(nl-1)).  Replaces the per-call setup of
 * mgcg1 (Cantilever/TOP3D125MGCGCUDA/mgcg1.cu:184-233, estWorkSpace
 * 293-322): the device, workspace and stream persist across solves. */
int gputo_create(int nelx, int nely, int nelz, int nl, const gputo_params *params, gputo_handle **out);
int gputo_destroy(gputo_handle *h);
int gputo_set_params(gputo_handle *h, const gputo_params *params);

/* The mgcg1 MEX solve (mgcg1.cu:120-182, mgcg1.c:51-105): K(CX) U = F with
 * zero initial guess, Dirichlet DOFs `fixed` (1-based doubles, nfd of them),
 * stop when ||r||/||F|| < tol or after maxit iterations.  U (ndof doubles)
 * is overwritten with the solution in the MATLAB frame (y components
 * negated, mgcg1.cu:282-283).  *relres gets the recursively updated
 * relative residual, *nit the 0-based exit index (mgcg1.cu:285-286).
 * Ke0 is the 24x24 element matrix as MATLAB passes it (column-major,
 * symmetric). */
int gputo_mgcg1(gputo_handle *h, const double *Ke0, const double *F, double *U, const double *CX,
                const double *fixed, int nfd, double tol, int maxit, double *relres, int *nit);

/* The same solve on device-resident data (no host copies, no y flip):
 * d_F, d_U (ndof doubles), d_CX (nel doubles), d_fixed (nfd int32,
 * 0-based DOF indices).  Ke0 is a host array.  Results as above. */
int gputo_solve_d(gputo_handle *h, const double *Ke0, const double *d_F, double *d_U, const double *d_CX,
                  const int *d_fixed, int nfd, double tol, int maxit, double *relres, int *nit);

/* Workspace estimate in doubles, with the reference's formula
 * (mgcg1.cu:293-322) -- kept so that callers that size memory from it keep
 * working; this library's own footprint is gputo_device_bytes(). */
long long gputo_est_workspace(int nfd, int nl, int nelx, int nely, int nelz);
long long gputo_device_bytes(const gputo_handle *h);

/* ---- kernel-level entry points (parity tests, microbenchmarks) ------- */
/* Host arrays in/out; each call runs on the handle's stream and waits.  */

/* y = K(CX) u with y[fixed] = 0 (multAV0c + ZeroBCc, mgcg1.cu:941-997,1326). */
int gputo_matvec_f64(gputo_handle *h, const double *Ke0, const double *CX, const double *u,
                     const double *fixed, int nfd, double *y);
/* The same product with a reduced-precision level-0 kernel of the
 * mixed-precision solve, FP64 in and out: kernel 1 FP32 node-gather, 2 FP32
 * MFMA, 3 split-BF16 MFMA (hi+lo), 4 split-BF16 MFMA (hi+mid+lo), 5 single
 * BF16 MFMA product bf16(Ke) bf16(u). */
int gputo_matvec_mixed(gputo_handle *h, int kernel, const double *Ke0, const double *CX, const double *u,
                       const double *fixed, int nfd, double *y);
/* y = K(CX) u with y[fixed] = 0 on device arrays, allocating nothing: the
 * same result, bit for bit, as gputo_matvec_f64.  No counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops).  No handle: nelx x nely x nelz mesh;
 * d_CX (nel doubles), d_u and d_y (ndof doubles) and d_fixed (nfd int32,
 * 0-based DOF indices; entries outside [0, ndof) are skipped) are device
 * pointers; Ke0 is a host 24x24 array with the unit-H8 entry structure of
 * the top3D125 Ke0 (GPUTO_ERR_ARGUMENT otherwise).  The node mask lives in
 * a static device buffer of the library, 8 MiB per GPU (meshes up to
 * 8,388,608 nodes; GPUTO_ERR_WORKSPACE beyond), shared by every call on that
 * GPU: calls that overlap on one GPU race.  The work is queued on `stream`
 * (a hipStream_t; NULL: the null stream) and not waited for. */
int gputo_matvec_f64_noalloc(int nelx, int nely, int nelz, const double *Ke0, const double *d_CX, const double *d_u,
                             const int *d_fixed, int nfd, double *d_y, void *stream);
/* Lane-layout check of the BF16 and FP32 MFMA instructions the level-0
 * kernels use: *max_err = max |D - A B| on small exact integer products
 * (0 expected). */
int gputo_selftest_mfma(double *max_err);
/* r = F - K(CX) u, r[fixed] = 0 (defect0c + ZeroBCc, mgcg1.cu:1208-1264). */
int gputo_defect_f64(gputo_handle *h, const double *Ke0, const double *CX, const double *F,
                     const double *u, const double *fixed, int nfd, double *r);
/* Diagonal of K(CX) (prepAD0c, mgcg1.cu:1447-1487). */
int gputo_diag_f64(gputo_handle *h, const double *Ke0, const double *CX, double *ad);
/* One V-cycle z = V(r) on the hierarchy built for (Ke0, CX, fixed)
 * (vcycle, mgcg1.cu:468-651). */
int gputo_vcycle_f64(gputo_handle *h, const double *Ke0, const double *CX, const double *fixed, int nfd,
                     const double *r, double *z);
/* Level data of the hierarchy built by the last solve / vcycle / setup:
 * element matrices of level l (nel_l x 24 x 24 row-major, l >= 1),
 * diagonal (ndof_l) and fixed list (1-based doubles, returned count). */
int gputo_setup(gputo_handle *h, const double *Ke0, const double *CX, const double *fixed, int nfd);
int gputo_level_info(const gputo_handle *h, int level, int *nelx, int *nely, int *nelz, int *nfd);
int gputo_level_matrices(gputo_handle *h, int level, double *A, double *ad, double *fixed);
/* Element scales of level `level` of a GPUTO_PRESET_MGCG9 hierarchy
 * (formCXc, Homogenization/QNC/mgcg9.cu:753-789; level 0: the input CX). */
int gputo_level_cx(gputo_handle *h, int level, double *cx);
/* Grid transfers on level l (fine) <-> l+1 (coarse) of the current
 * hierarchy: restriction F_c = P^T r_f with F_c[fixed_c] = 0 (rstrFc,
 * mgcg1.cu:792-831) and prolongation u_f += P u_c (interpc, 1587-1641). */
int gputo_restrict_f64(gputo_handle *h, int level, const double *r_fine, double *f_coarse);
int gputo_interp_f64(gputo_handle *h, int level, const double *u_coarse, double *u_fine);
/* Mean duration (ms, HIP events on the handle's stream) of `reps`
 * back-to-back launches of the level-0 operator y = K(CX) u on the
 * hierarchy of the last solve / setup (microbenchmark of multAV0c,
 * mgcg1.cu:941-997): kernel 0 FP64 node-gather, 1-5 as gputo_matvec_mixed. */
int gputo_bench_operator(gputo_handle *h, int kernel, int reps, double *avg_ms);
/* One application of the operator of level `level` (0 <= level < nl-1) of
 * the hierarchy of the last solve / setup: y = A u (mode 0, masked; multAV,
 * Cantilever/omp/mgcg1.c:667-715), y = f - A u (mode 1; defectc,
 * Cantilever/TOP3D125MGCGCUDA/mgcg1.cu:1266-1324), one damped-Jacobi
 * sweep from u with the preset's omega (mode 2; defectc then Jacobc,
 * mgcg1.cu:1199-1206), the same sweep from the prerelaxed iterate
 * prerelax*f/AD instead of u (mode 3; preRelc then one sweep,
 * mgcg1.cu:1190-1206) or the defect f - A (prerelax*f/AD) (mode 4; u is
 * ignored in modes 3 and 4).  op 1: stored element matrices or node
 * stencils (level 0: the level-0 kernel of the precision, for fp32 the
 * handle's l0_kernel); op 2: matrix-free level-1 operator (rediscretised
 * level 1 only).  fp32 != 0 runs the FP32 kernels of the mixed path.
 * Host vectors. */
int gputo_level_apply(gputo_handle *h, int level, int op, int mode, int fp32, const double *u, const double *f,
                      double *y);
/* Mean duration (ms, HIP events on the handle's stream) of `reps`
 * back-to-back matvecs of gputo_level_apply's operator. */
int gputo_bench_level(gputo_handle *h, int level, int op, int fp32, int reps, double *avg_ms);
/* ---- learned level-2 correction (params.learned = 1, mixed precision,
 * nl >= 4).  No counterpart in the reference: an extra stage of the FP32
 * V-cycle after the level-2 post-smoothing.  A 3-D convolutional
 * encoder-decoder (4 -> 16 -> 32 -> 64 -> 32 -> 16 -> 3 channels, 3x3x3
 * kernels, ReLU, strides 1/2/2, nearest upsampling with skip additions) maps
 * the Jacobi-scaled level-2 defect d and a stiffness channel to a correction
 * e; U += gamma e with gamma = <d,e>/<e,Ae> when gamma > 0, else the plain
 * V-cycle result stands (the fallback).  The PCG then uses the flexible
 * (Polak-Ribiere) beta.  Every solve logs level-2 defects as training
 * samples; gputo_learn_train runs Adam on the level-2 energy
 * (1/2 e^T A e - e^T d)/sigma^2 of those samples on the device.
 * Parameters are a flat FP32 array: per layer W[cout][cin][27] then b[cout].
 * Activations/inputs are channel-major on the level-2 node grid
 * n0 x n1 x n2 = (nnx, nnz, nny), position (i*n1 + k)*n2 + j. */
int gputo_learn_info(gputo_handle *h, long long *nparam, int *n0, int *n1, int *n2, int *nsamples);
/* set != 0: load params from buf (marks the network trained); 0: read them. */
int gputo_learn_params(gputo_handle *h, int set, double *buf);
int gputo_learn_reset(gputo_handle *h, unsigned seed);
/* on != 0 (the default): each solve logs the level-2 defects of its first 64
 * V-cycles as training samples; 0: no logging (a trained network in use
 * then costs nothing when the fallback has switched it off). */
int gputo_learn_logging(gputo_handle *h, int on);
/* y[3][n] = network(x[4][n]) */
int gputo_learn_forward(gputo_handle *h, const double *x, double *y);
/* g = d<gy, network(x)>/dparams */
int gputo_learn_grad(gputo_handle *h, const double *x, const double *gy, double *g);
/* `steps` Adam steps (learning rate lr) on the samples of the last solve. */
int gputo_learn_train(gputo_handle *h, int steps, double lr, double *loss_first, double *loss_last);

/* Coarsest-level solve x = A_c^{-1} b (cholsl1, mgcg1.cu:2175-2232). */
int gputo_coarse_solve_f64(gputo_handle *h, const double *b, double *x);

/* ---- index maps (bit-exact parity with the reference) ---------------- */
/* MATLAB connectivity cMat, 1-based DOFs, MATLAB local order
 * (top3D125mod.m:37-41); nel x 24 int32 row-major. */
int gputo_edofmat(gputo_handle *h, int *cmat);
/* Fixed-DOF list of level l as the reference's rstrFD builds it
 * (mgcg1.c:532-562): 1-based doubles in fine-list order. */
int gputo_fixed_list(const gputo_handle *h, int level, double *fixed, int *nfd);

/* ---- topology-optimisation driver (top3D125mod.m:105-165) ------------ */
typedef struct gputo_to_config {
    double volfrac, penal, rmin, eta, beta, move, Emin, E0;
    int ft;            /* 1: density filter; 3: filter + volume-preserving projection */
    double cg_tol;
    int cg_maxit;
} gputo_to_config;

typedef struct gputo_to_stats {
    double compliance, volume, change, eta, lmid, relres;
    int cg_iters, status;
    double seconds_total, seconds_solve;
    int inner_iters, outer_iters, learned_accepted, learned_rejected;
} gputo_to_stats;

/* Start a design loop on this handle: x0 (nel doubles, design variables),
 * F (ndof doubles, C frame), fixed (1-based doubles).  Device-resident
 * state persists until gputo_to_end. */
int gputo_to_begin(gputo_handle *h, const double *Ke0, const double *x0, const double *F, const double *fixed,
                   int nfd, const gputo_to_config *cfg);
/* One design iteration: filter, projection, CX, MGCG solve, sensitivity,
 * filtered sensitivities, OC bisection (top3D125mod.m:105-165). */
int gputo_to_iterate(gputo_handle *h, gputo_to_stats *st);
/* Copy state back: x, xPhys (nel doubles, may be NULL), U (ndof, C frame,
 * may be NULL). */
int gputo_to_get(gputo_handle *h, double *x, double *xPhys, double *U);
int gputo_to_end(gputo_handle *h);

/* Device-side driver kernels on host arrays (parity tests). */
int gputo_filter_f64(gputo_handle *h, const double *in, double rmin, double *out);
int gputo_sens_f64(gputo_handle *h, const double *Ke0, const double *U_cframe, double *ce);

/* Periodic homogenisation of a unit cell (no counterpart in the reference,
 * which has no periodic solve; QNC_V1_mod1m.m is a supported compliance
 * problem).  The handle must have params.periodic = 1.  The six unit
 * macroscopic strains (Voigt order xx, yy, zz, xy, yz, zx, engineering
 * shears; x along i, y along j, z along k, the frame of Ke0's local node
 * order) give six load cases f_q = sum_e CX_e Ke0 u0_q, u0_q = eps_q . x;
 * each is solved with the handle's MGCG (one hierarchy, set up once, node 0
 * and its images pinned) and
 *   CH[6a+b] = 1/nel * sum_e CX_e (u0_a - chi_a)_e^T Ke0 (u0_b - chi_b)_e,
 * evaluated as (sum CX) Ke0-energy of the unit strains minus the cross
 * terms chi_a.f_b + chi_b.f_a plus chi_a.K chi_b.  CH is 36 doubles
 * (row-major, symmetric).  relres and nit get 6 entries each (may be NULL);
 * chi (6 x ndof doubles, may be NULL) gets the periodic fluctuation fields
 * on the full node layout, equal on image nodes. */
int gputo_homogenize(gputo_handle *h, const double *Ke0, const double *CX, double tol, int maxit, double *CH,
                     double *relres, int *nit, double *chi);

/* Library version string. */
const char *gputo_version(void);
/* Wait for all work on the current device (benchmark timing brackets). */
int gputo_device_synchronize(void);

#ifdef __cplusplus
}
#endif
#endif /* GPUTO_H */
