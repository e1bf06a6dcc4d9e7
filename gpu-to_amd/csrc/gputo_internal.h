// Internal declarations of libgputo: grid geometry, tiling of the node-gather
// kernels, device state of one multigrid level and of a solver handle.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/gputo.h"

namespace gputo {

// Largest halo / element box of one node-gather tile (LDS arrays are sized
// from these; the tile chooser keeps every tile inside them).
constexpr int kMaxHalo = 864;
constexpr int kMaxElemBox = 640;
constexpr int kBlock = 256;          // threads per block of the gather kernels
constexpr int kStageRows = 4;        // halo rows per wave loaded before any is stored
constexpr int kStNodes = kBlock / 8; // nodes per block of the stored-matrix operator
constexpr int kRedBlocks = 1024;     // blocks of the two-stage reductions
constexpr int kMaxLevels = 12;
constexpr int kMaxCoarseDofs = 20000;  // dense coarsest solve (mgcg9.cu:74, maxDofs0)
constexpr int kKeMaxClasses = 16;    // Ke0 entry classes passed by value to the level-0 kernels

// One structured hex grid: elements nx*ny*nz, nodes (nx+1)(ny+1)(nz+1);
// node id = nny*nnz*i + nny*k + j, element id = ny*nz*i + ny*k + j.
struct Grid {
    int nx, ny, nz;
    int nnx, nny, nnz;
    long nel, nnod, ndof;
};

Grid make_grid(int nx, int ny, int nz);

// Node tile of a gather kernel: TJ x TK x TI nodes per block.
struct Tile {
    int TJ, TK, TI;
    int ntj, ntk, nti;
    __host__ __device__ long blocks() const { return (long)ntj * ntk * nti; }
};

// XCD-aware block order.  Consecutive blocks are dealt round-robin over the
// 8 XCDs, each with its own L2; launching 8*ceil(n/8) blocks and giving XCD x
// the contiguous tile range [x*m, (x+1)*m), m = ceil(n/8), keeps
// neighbouring tiles, which share halo nodes, on one L2.
constexpr int kXcds = 8;
inline long xcd_launch(long ntiles) { return (ntiles + kXcds - 1) / kXcds * kXcds; }
__device__ __forceinline__ long xcd_tile(long b, long ntiles)
{
    const long m = (ntiles + kXcds - 1) / kXcds;
    return (b % kXcds) * m + b / kXcds;
}

Tile choose_tile(const Grid &g);   // level-0 node-gather bricks
Tile stored_tile(const Grid &g);   // stored-matrix operator (at most kStNodes nodes)

// Arguments of the node-gather operator kernels (T = double or float).
template <class T>
struct OpArgsT {
    Grid g;
    Tile t;
    const T *Ke;           // 24x24 row-major (level 0)
    const T *CX;           // element scale (level 0)
    const T *A;            // packed 300 x nel entry-major element matrices (levels >= 1)
    const T *S;            // assembled node stencils [27][3][3][nnod] of the same operator, or null
    const T *U;            // input vector
    const T *F;            // right-hand side (defect / Jacobi)
    const T *AD;           // diagonal (Jacobi)
    const uint8_t *mask;   // per node: bit d set when DOF 3n+d is fixed
    T *out;                // result vector
    T *out2;               // second result (prerelax+defect writes U and R)
    double *partials;      // per-block partial sums (fused dot), may be null
    const float *Uf;       // OP_MATVEC_UF: the input vector in FP32
    const int *ctl;        // ctl[0] != 0: skip (converged)
    int nb;                // right-hand sides of a batched launch (0 or 1: one)
    double omega, prerelax;
    double cx_floor;       // > 0: element scales below it count as 0 (mgcg9 smoother defect)
    int kc;                // Ke0 has the unit-H8 class structure: use kv
    T kv[kKeMaxClasses];   // value of each Ke0 entry class (kKeClass)
};
using OpArgs = OpArgsT<double>;

enum OpMode {
    OP_MATVEC = 0,      // out = K U, masked; partials = sum U.out
    OP_DEFECT = 1,      // out = F - K U, masked
    OP_JACOBI = 2,      // out = (1-w) U + w (r + AD U)/AD with r = F - K U masked; partials = sum F.out
    OP_PRE_JACOBI = 3,  // as OP_JACOBI with U = prerelax*F/AD formed on the fly
    OP_PRE_DEFECT = 4,  // U = prerelax*F/AD written to out2, out = F - K U masked
    OP_MATVEC_UF = 5    // as OP_MATVEC with the input Uf in FP32 (level-0 node-gather, FP64)
};

// Scalars of one device-resident PCG (all on the device, one struct).
struct CgScalars {
    double rho, rho_p, alpha, beta, res0, relres, pq, rr;
    int done, nit, it, bad;
    // mixed precision: residual replacement state
    double rmax;     // largest recursive residual norm since the last replacement
    int noru;        // 0: replace the recursive residual by b - A x this iteration
    int nru;         // replacements done
};

// Batched launches: several right-hand sides on one hierarchy, one per
// blockIdx.z.  The vectors of a batch follow each other with the level's DOF
// count (or the vector length n) as stride, each right-hand side has its own
// CgScalars, and a producer of per-block partial sums with G blocks writes
// right-hand side z's at partials + z G.  The convergence flag of right-hand
// side z is ctl[z * kCtlStride] (ctl points at CgScalars::done).  A launch
// with one right-hand side has blockIdx.z = 0 and is unchanged.
static_assert(sizeof(CgScalars) % sizeof(int) == 0, "CgScalars stride in ints");
constexpr int kCtlStride = (int)(sizeof(CgScalars) / sizeof(int));
constexpr int kMaxBatch = 6;  // the six load cases of a homogenisation

template <class P>
__device__ __forceinline__ P *batch_ptr(P *p, long stride)
{
    return p ? p + (long)blockIdx.z * stride : p;
}
__device__ __forceinline__ const int *batch_ctl(const int *ctl) { return batch_ptr(ctl, kCtlStride); }
template <class T>
__device__ __forceinline__ void batch_shift(OpArgsT<T> &a)
{
    const long st = a.g.ndof;
    a.U = batch_ptr(a.U, st);
    a.F = batch_ptr(a.F, st);
    a.out = batch_ptr(a.out, st);
    a.out2 = batch_ptr(a.out2, st);
    a.partials = batch_ptr(a.partials, (long)gridDim.x);
    a.ctl = batch_ctl(a.ctl);
}
inline unsigned batch_dim(int nb) { return nb > 1 ? (unsigned)nb : 1u; }

struct Level {
    Grid g;
    Tile t;
    Tile tmf;                     // tiling of the matrix-free level-1 operator (FP32)
    Tile tmf64;                   // its FP64 tiling (half the elements per block)
    Tile tg;                      // node-gather tiling of this level (matrix-free levels)
    double *CXl = nullptr;        // element scales of a matrix-free coarse level (mgcg9)
    double *U = nullptr, *F = nullptr, *R = nullptr, *AD = nullptr, *T = nullptr;
    double *A = nullptr;          // packed element matrices (levels >= 1)
    double *S = nullptr;          // the level's operator as node stencils (stored levels below the coarsest)
    uint8_t *mask = nullptr;      // fixed-DOF node mask
    int *fixed_d = nullptr;       // fixed list (0-based DOF) on the device
    std::vector<int> fixed;       // fixed list (0-based DOF), reference order
};

struct Coarse {
    int n = 0;
    bool direct = true;      // factored (else smoothed: mgcg9 above 20000 DOFs)
    int bw = 0;              // lower bandwidth of the assembled matrix (node-major DOFs)
    double *AQ = nullptr;       // dense matrix, then its Cholesky factor (lower)
    double *Linv = nullptr;     // inverse of the factor used by the solves (lower, row-major)
    double *LinvT = nullptr;    // its transpose (row-major)
    double *Lbuf[2] = {}, *LTbuf[2] = {};  // two inverse buffers (lagged factorisation)
    int cur = 0;                // buffer of Linv/LinvT
    int target = 0;             // buffer the factorisation in flight writes
    bool have = false;          // buffers[cur] hold a factor of an earlier setup
    bool invalid = false;       // set by a parameter or operator change: the next setup
                                // drops a lagged factor still in flight and factors exactly
    bool solve_waits = true;    // solves wait for the factorisation in flight
    double *y = nullptr;     // work vector
    int *flag = nullptr;     // SPD failure flag
    // the factorisation runs on a side stream, overlapped with the start of
    // the solve: ev_in orders it after the assembly, ev_done orders every
    // coarsest solve and the next assembly after it
    hipStream_t side = nullptr;
    hipEvent_t ev_in = nullptr, ev_done = nullptr;
    bool pending = false;  // factorisation launched, SPD flag not read yet
    mutable bool inv_pending = false;  // its inverse not enqueued yet (coarse_finish)
    mutable bool waited = false;       // the solve stream already waits for it (stream order holds after)
};

struct MixedState;
struct LearnState;

// Buffers of a batched PCG (several right-hand sides on one hierarchy):
// every vector holds nb vectors one after another, every level of the
// V-cycle has its own batched work vectors (U, F, R, Tmp; FP64 for an FP64
// solve, FP32 for the mixed one).
struct BatchState {
    int nb = 0;
    int precision = -1;
    double *R = nullptr, *P = nullptr, *Q = nullptr, *Z = nullptr;
    double *partials = nullptr;  // nb x (the most per-block partials of a launch)
    CgScalars *sc = nullptr, *sc_host = nullptr;
    double *v64[kMaxLevels][4] = {};
    float *v32[kMaxLevels][4] = {};
    float *r32 = nullptr, *z32 = nullptr;  // the FP32 V-cycle's right-hand side and result
};

}  // namespace gputo

struct gputo_handle {
    gputo_params prm{};
    int nl = 0;
    gputo::Level L[gputo::kMaxLevels];
    gputo::Coarse C;
    hipStream_t stream = nullptr;
    int device = 0;
    double *Ke = nullptr;         // fine element matrix (row-major, symmetric)
    int ke_classes = 0;           // Ke has the unit-H8 class structure (values in kv)
    double kv[gputo::kKeMaxClasses] = {};
    double *Aelm = nullptr;       // level-1 Gauss-point matrices, packed [300][8]
    double *CX = nullptr;         // element scales of level 0
    double *Fv = nullptr, *Uv = nullptr, *R = nullptr, *P = nullptr, *Q = nullptr, *Z = nullptr;
    double *partials = nullptr;   // reduction partials
    gputo::CgScalars *sc = nullptr;
    gputo::CgScalars *sc_host = nullptr;  // pinned mirror
    int *ctl_zero = nullptr;      // an always-zero control word
    long long bytes = 0;
    bool hier_ready = false;
    bool masks_ready = false;     // level fixed lists and masks match L[0].fixed
    std::vector<void *> allocs;
    struct ToState *to = nullptr;
    gputo::MixedState *mx = nullptr;
    gputo::LearnState *ln = nullptr;  // learned level-2 correction (params.learned)
    // per-fine-element Galerkin tables of levels 1..gt_max (device, packed
    // [8^l][300]; null where the level is built from the level above)
    double *gt[gputo::kMaxLevels] = {};
    int gt_max = 0;
    int gt_preset = -1;          // preset the tables were built for (-1: none yet)
    std::vector<double> gt_ke;  // Ke they were built from (Galerkin level 1)
    double *hom = nullptr;      // periodic homogenisation: 6 load vectors and 6 fluctuations
    gputo::BatchState *bt = nullptr;  // batched PCG buffers (allocated on first use)
};

namespace gputo {

// ---- launch helpers (kernels.hip) ------------------------------------
template <class T> void launch_op0(hipStream_t s, int mode, const OpArgsT<T> &a);
void launch_op0_uf(hipStream_t s, const OpArgs &a);  // OP_MATVEC_UF
template <class T> void launch_opA(hipStream_t s, int mode, const OpArgsT<T> &a);
void launch_diag0(hipStream_t s, const Grid &g, const double *Ke, const double *CX, double *AD);
void launch_diagA(hipStream_t s, const Grid &g, const double *A, double *AD);
// S = the operator of the packed element matrices A as node stencils
void launch_stencil_build(hipStream_t s, const Grid &g, const double *A, double *S);
constexpr int kStencilEntries = 126;  // the own and 13 forward 3 x 3 blocks per node (the rest by symmetry)
template <class T>
void launch_restrict(hipStream_t s, const Grid &gc, const T *Rf, T *Fc, const uint8_t *maskc, const int *ctl,
                     int nb = 1);
template <class T>
void launch_interp(hipStream_t s, const Grid &gf, const T *Uc, T *Uf, const uint8_t *maskf, int zero_fixed,
                   const int *ctl, int nb = 1);
template <class T>  // Uout = Uin + P Uc (Uout may equal Uin)
void launch_interp_to(hipStream_t s, const Grid &gf, const T *Uc, const T *Uin, T *Uout, const uint8_t *maskf,
                      int zero_fixed, const int *ctl, int nb = 1);
template <class T> void launch_fill(hipStream_t s, T *v, long n, double val);
template <class Ta, class Tb>
void launch_dot_partial(hipStream_t s, const Ta *a, const Tb *b, long n, double *partials, const int *ctl,
                        int nb = 1);
template <class Ti, class To>
void launch_copy(hipStream_t s, const Ti *src, To *dst, long n, const int *ctl, int nb = 1);
template <class T>  // U = prerelax F / AD
void launch_prerelax(hipStream_t s, const T *F, const T *AD, T *U, long n, double prerelax, const int *ctl,
                     int nb = 1);
void launch_coarse_l1(hipStream_t s, const Grid &gc, const double *Aelm, const double *CXf, double *A);
void launch_coarse_galerkin(hipStream_t s, const Grid &gc, const double *Af, const double *Ke, const double *CXf,
                            double *Ac);
void launch_galerkin_table(hipStream_t s, const Grid &gc, int depth, const double *T, const double *CX0, double *A);
std::vector<std::vector<double>> galerkin_tables(const double *Ke, const double *base1, int maxdepth);
constexpr int kGtMaxDepth = 4;  // deepest level built from tables (8^4 x 300 doubles = 9.8 MB)
// A table product costs 300 FMA per level-0 element at every level, spread
// over nel_l/16 blocks that each stream 8^l table rows; below this many
// elements the child-by-child Galerkin of the level above (about 1.7e5 FMA
// per element, one block each) is faster.  Measured at 192x96x96: level 3
// (3456 elements) takes 260 us from its table and 452 us child by child;
// level 4 (432 elements) 95 us child by child.
constexpr long kGtMinElems = 2048;
void launch_corrcsa(hipStream_t s, const Grid &g, const int *fixed_d, int nfd, double *A, int mirrored);
// mgcg9 coarse levels: CX_c = 2 mean(children), element matrices CX Ke, and
// the zeroing of fixed rows/columns of the dense coarsest matrix
void launch_coarse_cx(hipStream_t s, const Grid &gc, const double *CXf, double *CXc);
void launch_ke_elements(hipStream_t s, const Grid &g, const double *CX, const double *Ke, double *A);
void launch_zero_rowcol(hipStream_t s, int N, const int *fixed_d, int nfd, double *AQ);
void launch_assemble(hipStream_t s, const Grid &g, const double *A, double *AQ);
// Factor C.AQ on C.side after the work on s.  lag: when an earlier factor
// exists, write the new one into the other buffer and let the solves keep
// the earlier one (coarse_adopt switches at the next setup).
void coarse_factor(hipStream_t s, Coarse &C, bool lag);
// enqueues the inverse of a factorisation launched by coarse_factor and
// records ev_done; every wait on ev_done comes after it
void coarse_finish(const Coarse &C);
int coarse_check(Coarse &C);  // waits for an exact factorisation; GPUTO_ERR_NOT_SPD on a bad pivot
void coarse_adopt(Coarse &C);  // before a setup: take a finished lagged factor
// after a parameter or operator change: no factor of an earlier setup
// (current or still in flight) is used again
void coarse_invalidate(Coarse &C);
// mid-solve: make the solves on s wait for the factorisation in flight and
// use it from now on (no-op without a lagged factorisation)
void coarse_switch(Coarse &C, hipStream_t s);
template <class T>
void launch_coarse_solve(hipStream_t s, const Coarse &C, const T *b, T *x, const uint8_t *mask, const int *ctl,
                         int nb = 1);

// PCG scalar steps (single-block kernels reading the reduction partials).
void launch_cg_init(hipStream_t s, CgScalars *sc, const double *partials, int nparts, int nb = 1);
void launch_cg_rho0(hipStream_t s, CgScalars *sc, const double *partials, int nparts, int nb = 1);
void launch_cg_alpha(hipStream_t s, CgScalars *sc, const double *partials, int nparts, int nb = 1);
// U += alpha P, R -= alpha Q, partials of R.R; R32 (if not null) gets R in FP32
template <class T>
void launch_cg_update_ur(hipStream_t s, T *U, T *R, const T *P, const T *Q, long n, CgScalars *sc, double *partials,
                         float *R32 = nullptr, int nb = 1);
void launch_cg_check(hipStream_t s, CgScalars *sc, const double *partials, int nparts, double tol, int maxit,
                     int nb = 1);
void launch_cg_beta(hipStream_t s, CgScalars *sc, const double *partials, int nparts, int maxit, int nb = 1);
template <class T> void launch_cg_update_p(hipStream_t s, T *P, const T *Z, long n, CgScalars *sc, int nb = 1);
// P = Z + beta P with the FP32 output of the mixed V-cycle (widened exactly)
void launch_cg_update_p(hipStream_t s, double *P, const float *Z, long n, CgScalars *sc, int nb = 1);
// the same with P kept in FP32 (Pout = Z + beta Pin, formed in FP64 and
// rounded once; Pout may be Pin)
void launch_cg_update_p32(hipStream_t s, float *Pout, const float *Pin, const float *Z, long n, CgScalars *sc);
// PCG with deferred u: the iterations' alphas are logged (0 once converged)
// and u += sum alpha_t p_t is applied every kCgLog iterations from the
// kept FP32 directions
constexpr int kCgLog = 16;
void launch_cg_alpha_log(hipStream_t s, CgScalars *sc, const double *partials, int nparts, double *alog);
// X0 (optional): also writes X0 = prerelax R32 / AD32, the prerelaxed
// level-0 iterate of the next FP32 V-cycle (mixed_prerelax_target)
void launch_cg_update_r(hipStream_t s, double *R, const double *Q, long n, CgScalars *sc, double *partials,
                        float *R32, const float *AD32 = nullptr, float *X0 = nullptr, double prerelax = 0.0);
void launch_cg_flush_u(hipStream_t s, double *U, const float *P, long n, const double *alog, int m);
// flexible beta (partials needs 2*kRedBlocks entries); z FP64, or the
// mixed V-cycle's FP32 output (products in FP64, the same values)
template <class TZ>
void launch_cg_beta_flex(hipStream_t s, const double *r, const TZ *z, const double *q, long n, CgScalars *sc,
                         double *partials, int maxit, int nb = 1);

void check_hip(hipError_t e, const char *what);  // throws on an error
#define HIPCHK(x) ::gputo::check_hip((x), #x)
void set_ke(gputo_handle *h, const double *Ke0);  // upload Ke0 and detect its class structure

// Deterministic sum over a kBlock-thread block: fixed butterfly per wave,
// then the wave sums in order.  Every thread of the block must call it.
__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double ws[kBlock / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) ws[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < kBlock / 64; ++i) s += ws[i];
    return s;
}


void to_release(gputo_handle *h);  // frees the design-loop state (driver.hip)

// ---- periodic unit cell (periodic.hip) --------------------------------
// A periodic vector lives on the full node grid; the nodes with i = nx,
// j = ny or k = nz are images of the master node (i mod nx, j mod ny,
// k mod nz).  Primal vectors (iterates) hold equal values on a master and
// its images; dual vectors (right-hand sides, residuals) are kept with the
// whole value on the master and zeros on the images, so that dot products
// over the full grid count every periodic DOF once.  The non-periodic
// operators applied to a primal vector give each master part of its row
// and the rest on its images: summing the images into the master (fold)
// completes the periodic operator.
// fold: master += images; images = bcast ? master : 0
template <class T> void launch_per_fold(hipStream_t s, const Grid &g, T *v, int bcast, const int *ctl, int nb = 1);
// images = master
template <class T> void launch_per_bcast(hipStream_t s, const Grid &g, T *v, const int *ctl, int nb = 1);
// pre: U = prerelax F/AD (0 on fixed DOFs); else U += omega fold(R)/AD;
// the new master value is written to the images too
template <class T>
void launch_per_jacobi(hipStream_t s, const Grid &g, T *U, const T *R, const T *F, const T *AD, const uint8_t *mask,
                       double omega, double prerelax, int pre, const int *ctl, int nb = 1);
// dense coarsest matrix of the periodic operator in the full numbering:
// master rows and columns hold the wrapped assembly, image rows are unit
void launch_assemble_per(hipStream_t s, const Grid &g, const double *A, double *AQ);

// unit-strain loads: fe[q] = Ke0 u0_q on the unit element, Q0[q][p] = u0_q.Ke0 u0_p
void homog_loads(const double *Ke0, double fe[6][24], double Q0[6][6]);
// the six load vectors (6 x ndof, master form) of the element scales CX
void launch_homog_rhs(hipStream_t s, const Grid &g, const double *CX, const double fe[6][24], const uint8_t *mask,
                      double *F);

int setup_hierarchy(gputo_handle *h, const std::vector<int> &fixed0);
int run_pcg(gputo_handle *h, const double *F, double *U, double tol, int maxit, double *relres, int *nit);
// nb <= kMaxBatch right-hand sides together (F, U: nb vectors one after another)
int run_pcg_batch(gputo_handle *h, const double *F, double *U, int nb, double tol, int maxit, double *relres,
                  int *nit);

}  // namespace gputo
