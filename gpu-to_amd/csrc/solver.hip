// Host side of libgputo: handle and workspace, multigrid hierarchy set-up,
// the V-cycle and the device-resident FP64 PCG, and the C-ABI entry points
// that replace the mgcg1 MEX solve (Cantilever/TOP3D125MGCGCUDA/mgcg1.cu).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "gputo_internal.h"
#include "ke_classes.h"
#include "learn.h"
#include "mixed.h"
#include "vcycle.h"

namespace gputo {

struct HipError : std::runtime_error {
    explicit HipError(const std::string &s) : std::runtime_error(s) {}
};

void check_hip(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}


static_assert(kKeClasses <= kKeMaxClasses, "Ke class table larger than the kernel argument array");

// Upload Ke0 (read row-major, as the reference reads MATLAB's column-major
// symmetric array, mgcg1.cu:131-133) and record whether it has the unit-H8
// class structure bit for bit: then the level-0 kernels take its class
// values as arguments instead of reading the 576 entries.
void set_ke(gputo_handle *h, const double *Ke0)
{
    HIPCHK(hipMemcpyAsync(h->Ke, Ke0, 576 * sizeof(double), hipMemcpyHostToDevice, h->stream));
    for (int k = 0; k < kKeClasses; ++k) h->kv[k] = Ke0[kKeFirst[k]];
    int ok = 1;
    for (int q = 0; q < 576 && ok; ++q) ok = std::memcmp(&Ke0[q], &h->kv[kKeClass[q]], sizeof(double)) == 0;
    h->ke_classes = ok;
}

Grid make_grid(int nx, int ny, int nz)
{
    Grid g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.nnx = nx + 1;
    g.nny = ny + 1;
    g.nnz = nz + 1;
    g.nel = (long)nx * ny * nz;
    g.nnod = (long)g.nnx * g.nny * g.nnz;
    g.ndof = 3 * g.nnod;
    return g;
}

static int balance(int n, int tmax)
{
    // smallest tile >= n/ceil(n/tmax) so that ceil(n/T) tiles cover n evenly
    if (tmax < 1) tmax = 1;
    if (tmax >= n) return n;
    const int nt = (n + tmax - 1) / tmax;
    return (n + nt - 1) / nt;
}

Tile choose_tile(const Grid &g)
{
    Tile t;
    t.TJ = balance(g.nny, 32);
    t.TK = balance(g.nnz, std::max(1, kBlock / (t.TJ * 2)));
    t.TI = balance(g.nnx, std::max(1, kBlock / (t.TJ * t.TK)));
    while ((t.TJ + 2) * (t.TK + 2) * (t.TI + 2) > kMaxHalo || (t.TJ + 1) * (t.TK + 1) * (t.TI + 1) > kMaxElemBox) {
        if (t.TI > 1)
            t.TI--;
        else if (t.TK > 1)
            t.TK--;
        else
            t.TJ--;
    }
    t.ntj = (g.nny + t.TJ - 1) / t.TJ;
    t.ntk = (g.nnz + t.TK - 1) / t.TK;
    t.nti = (g.nnx + t.TI - 1) / t.TI;
    return t;
}

Tile stored_tile(const Grid &g)
{
    Tile t;
    t.TJ = balance(g.nny, kStNodes);
    t.TK = balance(g.nnz, std::max(1, kStNodes / t.TJ));
    t.TI = balance(g.nnx, std::max(1, kStNodes / (t.TJ * t.TK)));
    t.ntj = (g.nny + t.TJ - 1) / t.TJ;
    t.ntk = (g.nnz + t.TK - 1) / t.TK;
    t.nti = (g.nnx + t.TI - 1) / t.TI;
    return t;
}

// ---- Gauss-point element stiffness (h = 2, level-1 rediscretisation) ----
// B^T D B |J| at one 2x2x2 Gauss point of a cube of side h in the element
// order ie/je/ke; rows of B: exx, eyy, ezz, gxy, gyz, gzx (the formulation
// of Kem3/defD/Bem3, Cantilever/omp/mgcg1.c:1356-1537).
static void gauss_point_matrix(double nu, double h, double s, double e, double t, double *K)
{
    static const int ie[8] = {0, 1, 1, 0, 0, 1, 1, 0}, je[8] = {0, 0, 1, 1, 0, 0, 1, 1}, ke[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    double B[6][24] = {};
    for (int a = 0; a < 8; ++a) {
        const double sx = ie[a] ? 1 : -1, sy = je[a] ? 1 : -1, sz = ke[a] ? 1 : -1;
        const double dx = 0.25 * sx * (1 + sy * e) * (1 + sz * t) / h;
        const double dy = 0.25 * sy * (1 + sx * s) * (1 + sz * t) / h;
        const double dz = 0.25 * sz * (1 + sx * s) * (1 + sy * e) / h;
        B[0][3 * a] = dx;
        B[1][3 * a + 1] = dy;
        B[2][3 * a + 2] = dz;
        B[3][3 * a] = dy;
        B[3][3 * a + 1] = dx;
        B[4][3 * a + 1] = dz;
        B[4][3 * a + 2] = dy;
        B[5][3 * a] = dz;
        B[5][3 * a + 2] = dx;
    }
    const double c1 = 1.0 / ((1 + nu) * (1 - 2 * nu));
    double D[6][6] = {};
    for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b) D[a][b] = c1 * (a == b ? 1 - nu : nu);
        D[3 + a][3 + a] = c1 * (1 - 2 * nu) / 2;
    }
    const double w = h * h * h / 8.0;
    for (int m = 0; m < 24; ++m)
        for (int n = 0; n < 24; ++n) {
            double acc = 0.0;
            for (int p = 0; p < 6; ++p)
                for (int q = 0; q < 6; ++q) acc += B[p][m] * D[p][q] * B[q][n];
            K[24 * m + n] = w * acc;
        }
}

// Packed [300][8] table: Aelm[8 p + na], na = 4 iP + 2 kP + jP with the
// Gauss point (g[iP], g[jP], g[kP]) (prepAelm, mgcg1.cu:1998-2034).
static std::vector<double> level1_gauss_table(double nu, double h)
{
    const double g[2] = {-1.0 / std::sqrt(3.0), 1.0 / std::sqrt(3.0)};
    std::vector<double> tab(300 * 8);
    double K[576];
    for (int iP = 0; iP < 2; ++iP)
        for (int kP = 0; kP < 2; ++kP)
            for (int jP = 0; jP < 2; ++jP) {
                gauss_point_matrix(nu, h, g[iP], g[jP], g[kP], K);
                const int na = 4 * iP + 2 * kP + jP;
                for (int m = 0; m < 24; ++m)
                    for (int n = 0; n <= m; ++n) tab[8 * (m * (m + 1) / 2 + n) + na] = K[24 * m + n];
            }
    return tab;
}

// rstrFD (Cantilever/omp/mgcg1.c:532-562): a fine fixed DOF survives on the
// next level iff its node has even (i, j, k); order follows the fine list.
static std::vector<int> restrict_fixed(const std::vector<int> &fine, const Grid &gc)
{
    const int fny = 2 * gc.ny + 1, fnz = 2 * gc.nz + 1;
    std::vector<int> out;
    out.reserve(fine.size() / 2 + 1);
    for (int dof : fine) {
        const int idof = dof % 3, nodP = dof / 3;
        const int iP = nodP / (fnz * fny);
        if (iP % 2) continue;
        const int kP = (nodP - fnz * fny * iP) / fny;
        if (kP % 2) continue;
        const int jP = nodP - fnz * fny * iP - fny * kP;
        if (jP % 2) continue;
        const long nod = (long)gc.nny * gc.nnz * (iP / 2) + (long)gc.nny * (kP / 2) + jP / 2;
        out.push_back((int)(3 * nod + idof));
    }
    return out;
}

static std::vector<uint8_t> node_mask(const std::vector<int> &fixed, const Grid &g)
{
    std::vector<uint8_t> m(g.nnod, 0);
    for (int dof : fixed) m[dof / 3] |= (uint8_t)(1u << (dof % 3));
    return m;
}

}  // namespace gputo

using namespace gputo;

namespace {

template <class T>
T *dalloc(gputo_handle *h, size_t n)
{
    void *p = nullptr;
    if (n == 0) n = 1;
    check_hip(hipMalloc(&p, n * sizeof(T)), "hipMalloc");
    h->allocs.push_back(p);
    h->bytes += (long long)(n * sizeof(T));
    return static_cast<T *>(p);
}

int map_exception()
{
    try {
        throw;
    } catch (const HipError &e) {
        std::fprintf(stderr, "gputo: %s\n", e.what());
        return GPUTO_ERR_DEVICE;
    } catch (const std::bad_alloc &) {
        return GPUTO_ERR_WORKSPACE;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "gputo: %s\n", e.what());
        return GPUTO_ERR_DEVICE;
    }
}

HierView<double> f64_view(const gputo_handle *h)
{
    HierView<double> H;
    for (int l = 0; l < h->nl; ++l) {
        const Level &v = h->L[l];
        H.L[l].U = v.U;
        H.L[l].F = v.F;
        H.L[l].R = v.R;
        H.L[l].AD = v.AD;
        H.L[l].Tmp = v.T;
        H.L[l].A = v.A;
        H.L[l].S = v.S;
        H.L[l].CX = l == 0 ? h->CX : v.CXl;
    }
    H.Ke = h->Ke;
    H.CX = h->CX;
    H.l1_mf = use_l1_mf(h->prm, h->nl);
    H.mg9 = preset_of(h->prm).mg9;
    H.per = h->prm.periodic != 0;
    return H;
}

OpArgs op_args(const gputo_handle *h, int l) { return view_args(h, f64_view(h), l, &h->sc->done); }

void vcycle(gputo_handle *h, const double *F0, double *U0) { vcycle_t(h, f64_view(h), F0, U0, &h->sc->done); }

int validate_fixed(const std::vector<int> &f, long ndof)
{
    for (int d : f)
        if (d < 0 || d >= ndof) return GPUTO_ERR_ARGUMENT;
    return 0;
}

bool all_finite(const double *v, long n)
{
    for (long i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

}  // namespace

namespace gputo {

// (Re)build the Galerkin tables of the current preset: once for a
// rediscretised level 1 (they depend only on its Gauss-point matrices),
// and whenever Ke changes when level 1 is itself Galerkin.
static void ensure_tables(gputo_handle *h)
{
    if (preset_of(h->prm).mg9) return;  // no Galerkin levels
    const bool gal1 = preset_of(h->prm).l1_galerkin;
    if (h->gt_max < (gal1 ? 1 : 2)) return;
    std::vector<double> ke;
    if (gal1) {
        ke.resize(576);
        HIPCHK(hipMemcpyAsync(ke.data(), h->Ke, 576 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        if (h->gt_preset == h->prm.preset && ke == h->gt_ke) return;
    } else if (h->gt_preset == h->prm.preset) {
        return;
    }
    std::vector<std::vector<double>> t;
    if (gal1) {
        t = galerkin_tables(ke.data(), nullptr, h->gt_max);
    } else {
        std::vector<double> base(8 * 576);
        const double g[2] = {-1.0 / std::sqrt(3.0), 1.0 / std::sqrt(3.0)};
        for (int iP = 0; iP < 2; ++iP)
            for (int kP = 0; kP < 2; ++kP)
                for (int jP = 0; jP < 2; ++jP)
                    gauss_point_matrix(kL1Nu, 2.0, g[iP], g[jP], g[kP], &base[576 * (4 * iP + 2 * kP + jP)]);
        t = galerkin_tables(nullptr, base.data(), h->gt_max);
    }
    for (int l = 1; l <= h->gt_max; ++l)
        if (!t[l].empty())
            HIPCHK(hipMemcpyAsync(h->gt[l], t[l].data(), t[l].size() * sizeof(double), hipMemcpyHostToDevice,
                                  h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));  // t is a temporary
    h->gt_preset = h->prm.preset;
    h->gt_ke = ke;
}

// Build the hierarchy for the current Ke, CX (device) and fixed list
// (initVcycle, mgcg1.cu:326-466).
static bool coarse_lagged(const gputo_params &p) { return p.coarse_lag >= 1; }

int setup_hierarchy(gputo_handle *h, const std::vector<int> &fixed0)
{
    const Preset p = preset_of(h->prm);
    hipStream_t s = h->stream;
    Level &L0 = h->L[0];
    if (int rc = validate_fixed(fixed0, L0.g.ndof)) return rc;
    // a factorisation still running on the side stream reads the coarsest
    // matrix this setup overwrites; a finished lagged one becomes current
    coarse_finish(h->C);
    if (h->C.ev_done) HIPCHK(hipStreamWaitEvent(s, h->C.ev_done, 0));
    coarse_adopt(h->C);
    ensure_tables(h);
    // the fixed lists and masks of every level are rebuilt only when the
    // fixed DOFs change (a design loop keeps them)
    if (!h->masks_ready || fixed0 != L0.fixed) {
        L0.fixed = fixed0;
        for (int l = 1; l < h->nl; ++l) h->L[l].fixed = restrict_fixed(h->L[l - 1].fixed, h->L[l].g);
        for (int l = 0; l < h->nl; ++l) {
            Level &v = h->L[l];
            std::vector<uint8_t> m = node_mask(v.fixed, v.g);
            HIPCHK(hipMemcpyAsync(v.mask, m.data(), m.size(), hipMemcpyHostToDevice, s));
            if (!v.fixed.empty())
                HIPCHK(hipMemcpyAsync(v.fixed_d, v.fixed.data(), v.fixed.size() * sizeof(int),
                                      hipMemcpyHostToDevice, s));
            HIPCHK(hipStreamSynchronize(s));  // host vectors above are temporaries
        }
        h->masks_ready = true;
    }
    const bool per = h->prm.periodic != 0;
    launch_diag0(s, L0.g, h->Ke, h->CX, L0.AD);
    if (per) launch_per_fold<double>(s, L0.g, L0.AD, 1, nullptr);
    if (p.mg9) {
        // mgcg9 (initVcycle, mgcg9.cu:342-414): coarse element scales and
        // diagonals from the unit Ke; the coarsest matrix, when factored,
        // is CX*Ke assembled with its fixed rows and columns decoupled
        for (int l = 1; l < h->nl; ++l)
            launch_coarse_cx(s, h->L[l].g, l == 1 ? h->CX : h->L[l - 1].CXl, h->L[l].CXl);
        Level &vc = h->L[h->nl - 1];
        h->C.direct = vc.g.ndof < kMaxCoarseDofs;
        // a smoothed coarsest level starts each solve from zero and then
        // keeps its iterate from one V-cycle to the next (mgcg9.cu:553-563)
        if (!h->C.direct) launch_fill<double>(s, vc.U, vc.g.ndof, 0.0);
        if (h->C.direct) {
            launch_ke_elements(s, vc.g, vc.CXl, h->Ke, vc.A);
            if (per)
                launch_assemble_per(s, vc.g, vc.A, h->C.AQ);
            else
                launch_assemble(s, vc.g, vc.A, h->C.AQ);
            launch_zero_rowcol(s, (h->C.n + 31) / 32 * 32, vc.fixed_d, (int)vc.fixed.size(), h->C.AQ);
            HIPCHK(hipGetLastError());
            coarse_factor(s, h->C, coarse_lagged(h->prm));
        }
        // the diagonals are not needed by the factorisation: launched after
        // it, they run while it does
        for (int l = 1; l < h->nl; ++l) {
            Level &v = h->L[l];
            launch_diag0(s, v.g, h->Ke, v.CXl, v.AD);
            if (per) launch_per_fold<double>(s, v.g, v.AD, 1, nullptr);
        }
        HIPCHK(hipGetLastError());
        h->hier_ready = true;
        return 0;
    }
    h->C.direct = true;
    for (int l = 1; l < h->nl; ++l) {
        Level &v = h->L[l];
        if (l == 1 && !p.l1_galerkin)
            launch_coarse_l1(s, v.g, h->Aelm, h->CX, v.A);
        else if (l <= h->gt_max && (v.g.nel >= kGtMinElems || l == 1))
            launch_galerkin_table(s, v.g, l, h->gt[l], h->CX, v.A);
        else
            launch_coarse_galerkin(s, v.g, l == 1 ? nullptr : h->L[l - 1].A, h->Ke, h->CX, v.A);
        if (l == h->nl - 1) {
            if (per) {
                // the wrapped matrix with the fixed rows and columns decoupled
                launch_assemble_per(s, v.g, v.A, h->C.AQ);
                launch_zero_rowcol(s, (h->C.n + 31) / 32 * 32, v.fixed_d, (int)v.fixed.size(), h->C.AQ);
            } else {
                launch_corrcsa(s, v.g, v.fixed_d, (int)v.fixed.size(), v.A, p.corr_mirrored);
                launch_assemble(s, v.g, v.A, h->C.AQ);
            }
        }
    }
    HIPCHK(hipGetLastError());
    // overlaps the solve (lagged) or the first V-cycle's descent (exact);
    // errors surface in coarse_check / coarse_adopt
    coarse_factor(s, h->C, coarse_lagged(h->prm));
    // diagonals and node stencils, which the factorisation does not read,
    // run on the main stream while it does
    for (int l = 1; l < h->nl; ++l) {
        Level &v = h->L[l];
        launch_diagA(s, v.g, v.A, v.AD);
        if (per) launch_per_fold<double>(s, v.g, v.AD, 1, nullptr);
        if (v.S) launch_stencil_build(s, v.g, v.A, v.S);
    }
    HIPCHK(hipGetLastError());
    h->hier_ready = true;
    return 0;
}

// A level-0 or level-1 schedule of the FP32 V-cycle must take an even number of
// smoothing steps (the prerelaxation is the first pre-step): damped
// Jacobi's error factor 1 - omega*lambda is negative on the highest modes,
// so an odd count makes the preconditioner indefinite (measured: 1 + 2 and
// 2 + 3 sweeps stall the CG at 48x24x24, 1 + 1, 2 + 2 and 1 + 3 do not).
static bool schedule_ok(const gputo_params &p)
{
    const Preset ps = preset_of(p);
    for (int l = 0; l < 4; ++l) {
        const int a = mixed_sweeps(p, l, false), b = mixed_sweeps(p, l, true);
        if (a < -1 || b < -1) return false;
        if (a == 0 && b == 0) continue;
        // -1: no sweep (0: the preset's count)
        const int pre = a > 0 ? a : (a < 0 ? 0 : ps.npre), post = b > 0 ? b : (b < 0 ? 0 : ps.npost);
        if ((std::max(pre, 1) + post) % 2) return false;
    }
    return true;
}

static bool custom_schedule(const gputo_params &p)
{
    return p.precision != GPUTO_PREC_F64 &&
           (p.mixed_npre0 != 0 || p.mixed_npost0 != 0 || p.mixed_npre1 != 0 || p.mixed_npost1 != 0 ||
            p.mixed_npre2 != 0 || p.mixed_npost2 != 0 || p.mixed_npre3 != 0 || p.mixed_npost3 != 0);
}

// Device-resident PCG (mgcg1.cu:235-280); U, F on the device.
// z = M r: the FP64 V-cycle, or the FP32 V-cycle of the mixed precisions
static void precond(gputo_handle *h, const double *r, double *z, bool r32_ready = false, bool pre0_ready = false)
{
    if (h->prm.precision == GPUTO_PREC_F64)
        vcycle(h, r, z);
    else
        mixed_vcycle(h, r, z, &h->sc->done, r32_ready, pre0_ready);
}

int run_pcg(gputo_handle *h, const double *F, double *U, double tol, int maxit, double *relres, int *nit)
{
    hipStream_t s = h->stream;
    const auto t_begin = std::chrono::steady_clock::now();
    if (h->prm.precision != GPUTO_PREC_F64) mixed_begin(h);
    if (h->ln) learn_begin_solve(h);
    // a trained learned stage makes the preconditioner vary, and a level-0
    // sweep schedule other than the preset's makes it less symmetric:
    // flexible beta
    // a factor switched in mid-solve (coarse_lag >= 2) changes it too
    const bool flex = (h->ln && h->ln->trained) || custom_schedule(h->prm) || h->prm.coarse_lag >= 2;
    // coarse_lag = n >= 2: the first n V-cycles use the previous setup's
    // coarsest factor, later ones the factor of this setup
    int vcycles = 0;
    // the residual update also writes the FP32 right-hand side of the mixed V-cycle
    float *r32 = h->prm.precision != GPUTO_PREC_F64 ? mixed_r32(h) : nullptr;
    // ... whose FP32 result the CG then reads directly (not for a periodic
    // cell, whose preconditioned residual is folded in FP64)
    const float *z32 = (r32 && !h->prm.periodic) ? mixed_z32(h) : nullptr;
    // graph_vcycles: the loop's V-cycles of that form replay one recorded HIP
    // graph (their launches are the same each time)
    const bool graphs = z32 && h->prm.graph_vcycles;
    // ... and then also forms the V-cycle's prerelaxed level-0 iterate
    // (prerelax*r/AD), in place of the V-cycle's first elementwise pass
    const float *ad32 = nullptr;
    double prx = 0.0;
    float *x0 = z32 ? mixed_prerelax_target(h, &ad32, &prx) : nullptr;
    auto precond_n = [&](const double *r, double *z, bool r32_ready) {
        if (h->prm.coarse_lag >= 2 && vcycles == h->prm.coarse_lag) {
            coarse_switch(h->C, s);
            if (graphs) mixed_graph_drop(h);  // it holds the previous factor
        }
        const bool pre0 = x0 && r32_ready && !z;  // the loop's V-cycles after the residual update
        if (graphs && r32_ready && !z)
            mixed_vcycle_graph(h, &h->sc->done, pre0);
        else
            precond(h, r, z, r32_ready, pre0);
        ++vcycles;
    };
    // ... and then keeps the search direction in FP32: p = z + beta p formed
    // in FP64 and rounded once; u, r and the matvec take exactly the stored
    // values, so the recurrences stay consistent (the FP64 residual decides
    // convergence), at half the traffic of p.  u += alpha p is deferred:
    // the last kCgLog directions stay in a ring with their alphas, and u
    // takes them every kCgLog iterations in one pass (the same fused
    // multiply-adds in the same order), so u is not read and written by
    // every iteration.
    float *p32 = z32 ? mixed_p32(h) : nullptr;
    double *alog = z32 ? mixed_alog(h) : nullptr;
    int folded = 0;  // iterations whose alpha p is in U
    Level &L0 = h->L[0];
    const long n = L0.g.ndof;
    const int *ctl = &h->sc->done;
    const int nmv = (int)xcd_launch(L0.t.blocks());  // partials of the level-0 matvec
    const bool per = h->prm.periodic != 0;
    launch_fill(s, U, n, 0.0);
    launch_copy(s, F, h->R, n, nullptr);
    if (per) launch_per_fold<double>(s, L0.g, h->R, 0, nullptr);  // images' share to the masters
    launch_dot_partial(s, h->R, h->R, n, h->partials, nullptr);
    launch_cg_init(s, h->sc, h->partials, kRedBlocks);
    precond_n(h->R, h->P, false);
    launch_copy(s, h->P, h->Z, n, ctl);
    if (p32) launch_copy<double, float>(s, h->P, p32, n, ctl);  // exact: z is the V-cycle's FP32 output
    launch_dot_partial(s, h->R, h->Z, n, h->partials, ctl);
    launch_cg_rho0(s, h->sc, h->partials, kRedBlocks);
    const int chunk = 4;
    int enq = 0;
    for (;;) {
        for (int c = 0; c < chunk && enq < maxit; ++c, ++enq) {
            const int sl = enq % kCgLog;  // ring slot of this iteration's direction
            float *pk = p32 ? p32 + (long)sl * n : nullptr;
            OpArgs a = op_args(h, 0);
            a.cx_floor = 0.0;  // the CG matvec keeps every element (multAV0c, mgcg9.cu:791-848)
            a.U = h->P;
            a.out = h->Q;
            a.partials = h->partials;
            if (p32) {
                a.Uf = pk;
                launch_op0_uf(s, a);
            } else {
                launch_op0(s, OP_MATVEC, a);
            }
            // p.q over the full grid is the periodic p.Kp before the fold
            // (p is equal on images); the fold completes q for the update
            if (per) launch_per_fold<double>(s, L0.g, h->Q, 0, ctl);
            if (p32) {
                launch_cg_alpha_log(s, h->sc, h->partials, nmv, alog + sl);
                launch_cg_update_r(s, h->R, h->Q, n, h->sc, h->partials, r32, ad32, x0, prx);
                if (sl == kCgLog - 1) {  // before slot 0 is overwritten by this iteration's p update
                    launch_cg_flush_u(s, U, p32, n, alog, kCgLog);
                    folded = enq + 1;
                }
            } else {
                launch_cg_alpha(s, h->sc, h->partials, nmv);
                launch_cg_update_ur(s, U, h->R, h->P, h->Q, n, h->sc, h->partials, r32);
            }
            launch_cg_check(s, h->sc, h->partials, kRedBlocks, tol, maxit);
            if (z32) {
                // the FP32 V-cycle's result feeds the FP64 dots and the P
                // update directly (widened exactly, so nothing changes but
                // the traffic of an FP64 copy of z)
                precond_n(h->R, nullptr, true);
                if (flex) {
                    launch_cg_beta_flex<float>(s, h->R, z32, h->Q, n, h->sc, h->partials, maxit);
                } else {
                    launch_dot_partial<double, float>(s, h->R, z32, n, h->partials, ctl);
                    launch_cg_beta(s, h->sc, h->partials, kRedBlocks, maxit);
                }
                launch_cg_update_p32(s, p32 + (long)((sl + 1) % kCgLog) * n, pk, z32, n, h->sc);
                continue;
            }
            precond_n(h->R, h->Z, r32 != nullptr);
            if (flex) {
                launch_cg_beta_flex<double>(s, h->R, h->Z, h->Q, n, h->sc, h->partials, maxit);
            } else {
                launch_dot_partial(s, h->R, h->Z, n, h->partials, ctl);
                launch_cg_beta(s, h->sc, h->partials, kRedBlocks, maxit);
            }
            launch_cg_update_p(s, h->P, h->Z, n, h->sc);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(h->sc_host, h->sc, sizeof(CgScalars), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (h->sc_host->done || enq >= maxit) {
            if (p32) {  // the directions since the last pass (converged ones logged alpha 0)
                launch_cg_flush_u(s, U, p32, n, alog, enq - folded);
                HIPCHK(hipStreamSynchronize(s));
            }
            break;
        }
        // the learned stage's fallback watches a trial's progress
        if (h->ln)
            learn_poll(h, h->sc_host->it,
                       std::chrono::duration<double>(std::chrono::steady_clock::now() - t_begin).count(),
                       h->sc_host->relres);
    }
    const CgScalars &r = *h->sc_host;
    *relres = r.relres;
    *nit = r.done ? r.nit : maxit;
    if (maxit <= 0) *nit = 0;
    if (int frc = coarse_check(h->C)) {  // the solve used a failed factorisation
        h->hier_ready = false;
        coarse_invalidate(h->C);
        *nit = frc;
        if (h->prm.precision != GPUTO_PREC_F64) mixed_end(h, 0, 0);
        if (h->ln) {
            int acc = 0, rej = 0;
            learn_end_solve(h, &acc, &rej);
        }
        return frc;
    }
    if (h->prm.precision != GPUTO_PREC_F64) mixed_end(h, *nit, r.done);
    if (h->ln) {
        int acc = 0, rej = 0;
        learn_end_solve(h, &acc, &rej,
                        std::chrono::duration<double>(std::chrono::steady_clock::now() - t_begin).count());
        if (h->mx) {
            h->mx->last_learned_ok = acc;
            h->mx->last_learned_bad = rej;
        }
    }
    if (h->prm.print_level >= 1)
        std::fprintf(stderr, "gputo: MGCG nit=%d relres=%.6g res0=%.6g\n", *nit, *relres, r.res0);
    if (r.bad) {
        *nit = GPUTO_ERR_NONFINITE;
        return GPUTO_ERR_NONFINITE;
    }
    return 0;
}

// Buffers of a batched PCG with nb right-hand sides for the handle's
// precision (kept for later calls with the same nb and precision).
static BatchState &batch_state(gputo_handle *h, int nb)
{
    BatchState *&b = h->bt;
    const int prec = h->prm.precision;
    if (b && b->nb == nb && b->precision == prec) return *b;
    if (b) {
        // a different shape: the device buffers stay with the handle (freed
        // at destroy), the pinned scalars are replaced
        if (b->sc_host) hipHostFree(b->sc_host);
        delete b;
        b = nullptr;
    }
    b = new BatchState();
    b->nb = nb;
    b->precision = prec;
    const long n = h->L[0].g.ndof;
    b->R = dalloc<double>(h, nb * n);
    b->P = dalloc<double>(h, nb * n);
    b->Q = dalloc<double>(h, nb * n);
    b->Z = dalloc<double>(h, nb * n);
    const long nmv = xcd_launch(h->L[0].t.blocks());
    b->partials = dalloc<double>(h, nb * std::max<long>(nmv, 2 * kRedBlocks));
    b->sc = dalloc<CgScalars>(h, nb);
    HIPCHK(hipHostMalloc((void **)&b->sc_host, nb * sizeof(CgScalars), hipHostMallocDefault));
    std::memset(b->sc_host, 0, nb * sizeof(CgScalars));
    HIPCHK(hipMemcpy(b->sc, b->sc_host, nb * sizeof(CgScalars), hipMemcpyHostToDevice));
    const bool mixed = prec != GPUTO_PREC_F64;
    for (int l = 0; l < h->nl; ++l) {
        const long nl_dof = h->L[l].g.ndof;
        for (int k = 0; k < 4; ++k) {
            if (l == 0 && k < 2) continue;  // level 0's U and F are the V-cycle's output and input
            if (mixed)
                b->v32[l][k] = dalloc<float>(h, nb * nl_dof);
            else
                b->v64[l][k] = dalloc<double>(h, nb * nl_dof);
        }
    }
    if (mixed) {
        b->r32 = dalloc<float>(h, nb * n);
        b->z32 = dalloc<float>(h, nb * n);
    }
    return *b;
}

template <class T>
static void batch_view(HierView<T> &H, T *const (&v)[kMaxLevels][4], int nl, int nb)
{
    for (int l = 0; l < nl; ++l) {
        H.L[l].U = v[l][0];
        H.L[l].F = v[l][1];
        H.L[l].R = v[l][2];
        H.L[l].Tmp = v[l][3];
    }
    H.nb = nb;
}

// The PCG of run_pcg for nb right-hand sides at once on one hierarchy
// (F and U hold nb vectors one after another): every operator, transfer,
// smoother, coarsest solve and vector launch applies to all nb vectors, each
// with its own CG scalars, and a right-hand side that has converged skips
// the rest (its control word).  The FP32 V-cycle of the mixed path applies
// level 0 with the FP32 node-gather kernel, whose launches take a batch.
int run_pcg_batch(gputo_handle *h, const double *F, double *U, int nb, double tol, int maxit, double *relres,
                  int *nit)
{
    if (nb < 1 || nb > kMaxBatch || h->ln) return GPUTO_ERR_ARGUMENT;
    hipStream_t s = h->stream;
    const bool mixed = h->prm.precision != GPUTO_PREC_F64;
    if (mixed) mixed_begin(h);
    BatchState &b = batch_state(h, nb);
    const bool flex = custom_schedule(h->prm);
    Level &L0 = h->L[0];
    const long n = L0.g.ndof;
    const int *ctl = &b.sc->done;
    const int nmv = (int)xcd_launch(L0.t.blocks());
    const bool per = h->prm.periodic != 0;
    HierView<double> H64 = f64_view(h);
    HierView<float> H32;
    if (mixed) {
        H32 = h->mx->H;
        H32.l0 = L0_GATHER;
        H32.learn = false;
        batch_view<float>(H32, b.v32, h->nl, nb);
    } else {
        batch_view<double>(H64, b.v64, h->nl, nb);
    }
    float *r32 = mixed ? b.r32 : nullptr;
    auto precond = [&](const double *r, double *z, bool r32_ready) {
        if (!mixed) {
            vcycle_t<double>(h, H64, r, z, ctl);
            return;
        }
        if (!r32_ready) launch_copy<double, float>(s, r, b.r32, n, ctl, nb);
        vcycle_t<float>(h, H32, b.r32, b.z32, ctl);
        launch_copy<float, double>(s, b.z32, z, n, ctl, nb);
    };
    launch_fill(s, U, nb * n, 0.0);
    launch_copy(s, F, b.R, n, nullptr, nb);
    if (per) launch_per_fold<double>(s, L0.g, b.R, 0, nullptr, nb);
    launch_dot_partial(s, b.R, b.R, n, b.partials, nullptr, nb);
    launch_cg_init(s, b.sc, b.partials, kRedBlocks, nb);
    precond(b.R, b.P, false);
    launch_copy(s, b.P, b.Z, n, ctl, nb);
    launch_dot_partial(s, b.R, b.Z, n, b.partials, ctl, nb);
    launch_cg_rho0(s, b.sc, b.partials, kRedBlocks, nb);
    const int chunk = 4;
    int enq = 0;
    for (;;) {
        for (int c = 0; c < chunk && enq < maxit; ++c, ++enq) {
            OpArgs a = op_args(h, 0);
            a.cx_floor = 0.0;
            a.U = b.P;
            a.out = b.Q;
            a.partials = b.partials;
            a.ctl = ctl;
            a.nb = nb;
            launch_op0(s, OP_MATVEC, a);
            if (per) launch_per_fold<double>(s, L0.g, b.Q, 0, ctl, nb);
            launch_cg_alpha(s, b.sc, b.partials, nmv, nb);
            launch_cg_update_ur(s, U, b.R, b.P, b.Q, n, b.sc, b.partials, r32, nb);
            launch_cg_check(s, b.sc, b.partials, kRedBlocks, tol, maxit, nb);
            precond(b.R, b.Z, r32 != nullptr);
            if (flex) {
                launch_cg_beta_flex<double>(s, b.R, b.Z, b.Q, n, b.sc, b.partials, maxit, nb);
            } else {
                launch_dot_partial(s, b.R, b.Z, n, b.partials, ctl, nb);
                launch_cg_beta(s, b.sc, b.partials, kRedBlocks, maxit, nb);
            }
            launch_cg_update_p(s, b.P, b.Z, n, b.sc, nb);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(b.sc_host, b.sc, nb * sizeof(CgScalars), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        bool all = true;
        for (int q = 0; q < nb; ++q) all = all && b.sc_host[q].done;
        if (all || enq >= maxit) break;
    }
    if (int frc = coarse_check(h->C)) {  // the solve used a failed factorisation
        h->hier_ready = false;
        coarse_invalidate(h->C);
        for (int q = 0; q < nb; ++q) nit[q] = frc;
        return frc;
    }
    int rc = 0;
    for (int q = 0; q < nb; ++q) {
        const CgScalars &r = b.sc_host[q];
        relres[q] = r.relres;
        nit[q] = r.done ? r.nit : maxit;
        if (maxit <= 0) nit[q] = 0;
        if (r.bad) {
            nit[q] = GPUTO_ERR_NONFINITE;
            rc = GPUTO_ERR_NONFINITE;
        }
    }
    if (mixed) mixed_end(h, nit[0], b.sc_host[0].done);
    if (h->prm.print_level >= 1)
        for (int q = 0; q < nb; ++q)
            std::fprintf(stderr, "gputo: MGCG rhs %d nit=%d relres=%.6g\n", q, nit[q], relres[q]);
    return rc;
}

}  // namespace gputo

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

void gputo_default_params(gputo_params *p)
{
    std::memset(p, 0, sizeof(*p));
    p->preset = GPUTO_PRESET_REF_C;
    p->precision = GPUTO_PREC_F64;
    p->learned = 0;
    p->l0_kernel = 0;
    p->l1_kernel = 0;
    p->print_level = 0;
}

const char *gputo_version(void) { return "gputo-mi355x 0.1 (gfx950)"; }

int gputo_device_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : GPUTO_ERR_DEVICE; }

long long gputo_est_workspace(int nfd, int nl, int nelx, int nely, int nelz)
{
    // estWorkSpace (mgcg1.cu:293-322), in doubles
    long long nnel0 = (long long)nelx * nely * nelz;
    long long ndof0 = 3LL * (nelx + 1) * (nely + 1) * (nelz + 1);
    long long lW = ndof0 * 6 + nnel0 + nfd;
    for (int lev = 0; lev < nl; ++lev) {
        long long ex = nelx >> lev, ey = nely >> lev, ez = nelz >> lev;
        long long ndof = 3 * (ex + 1) * (ey + 1) * (ez + 1);
        lW += lev == 0 ? 2 * ndof : 4 * ndof + 300 * ex * ey * ez + nfd;
        if (lev == nl - 1) lW += 3 * ndof + ndof * ndof;
    }
    return lW;
}

long long gputo_device_bytes(const gputo_handle *h) { return h ? h->bytes : 0; }


int gputo_create(int nelx, int nely, int nelz, int nl, const gputo_params *params, gputo_handle **out)
{
    if (!out) return GPUTO_ERR_ARGUMENT;
    *out = nullptr;
    if (nelx < 1 || nely < 1 || nelz < 1 || nl < 2 || nl > kMaxLevels) return GPUTO_ERR_ARGUMENT;
    const int f = 1 << (nl - 1);
    if (nelx % f || nely % f || nelz % f) return GPUTO_ERR_ARGUMENT;
    const Grid gc = make_grid(nelx / f, nely / f, nelz / f);
    // a coarsest level too large for the dense solve is only allowed for the
    // mgcg9 engine, which smooths it instead
    const bool dense_ok = gc.ndof <= kMaxCoarseDofs;
    if (!dense_ok && !(params && params->preset == GPUTO_PRESET_MGCG9)) return GPUTO_ERR_WORKSPACE;
    // the learned stage works on level 2 of the FP32 V-cycle
    if (params && params->learned && (params->precision == GPUTO_PREC_F64 || nl < 4)) return GPUTO_ERR_ARGUMENT;
    if (params && !schedule_ok(*params)) return GPUTO_ERR_ARGUMENT;
    // a periodic cell keeps a master distinct from its images on every
    // level; the learned stage is trained on non-periodic defects only
    if (params && params->periodic && (gc.nx < 2 || gc.ny < 2 || gc.nz < 2 || params->learned))
        return GPUTO_ERR_ARGUMENT;
    gputo_handle *h = new gputo_handle();
    if (params)
        h->prm = *params;
    else
        gputo_default_params(&h->prm);
    h->nl = nl;
    try {
        HIPCHK(hipGetDevice(&h->device));
        HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        HIPCHK(hipStreamCreateWithFlags(&h->C.side, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&h->C.ev_in, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&h->C.ev_done, hipEventDisableTiming));
        int maxblocks = 2 * kRedBlocks;  // two reductions share the partials (flexible beta)
        for (int l = 0; l < nl; ++l) {
            Level &v = h->L[l];
            v.g = make_grid(nelx >> l, nely >> l, nelz >> l);
            v.t = l == 0 ? choose_tile(v.g) : stored_tile(v.g);
            v.tg = choose_tile(v.g);
            maxblocks = std::max<long>(maxblocks, xcd_launch(v.t.blocks()));
            maxblocks = std::max<long>(maxblocks, xcd_launch(v.tg.blocks()));
            if (l == 1) {
                v.tmf = l1_tile(v.g);
                v.tmf64 = l1_tile64(v.g);
                maxblocks = std::max<long>(maxblocks, std::max(v.tmf.blocks(), v.tmf64.blocks()));
            }
            v.R = dalloc<double>(h, v.g.ndof);
            v.AD = dalloc<double>(h, v.g.ndof);
            v.T = dalloc<double>(h, v.g.ndof);
            v.mask = dalloc<uint8_t>(h, v.g.nnod);
            v.fixed_d = dalloc<int>(h, v.g.ndof);
            if (l > 0) {
                v.U = dalloc<double>(h, v.g.ndof);
                v.F = dalloc<double>(h, v.g.ndof);
                v.A = dalloc<double>(h, 300 * v.g.nel);
                v.CXl = dalloc<double>(h, v.g.nel);
                // node stencils of the stored levels the V-cycle smooths
                if (l < nl - 1 && !(l == 1 && use_l1_mf(h->prm, nl)))
                    v.S = dalloc<double>(h, (long)kStencilEntries * v.g.nnod);
            }
        }
        const Grid &g0 = h->L[0].g;
        maxblocks = (int)std::max<long>(maxblocks, mfma_tile_blocks(g0));  // partials of the MFMA level-0 kernel
        h->Ke = dalloc<double>(h, 576);
        h->Aelm = dalloc<double>(h, 2400);
        h->CX = dalloc<double>(h, g0.nel);
        h->Fv = dalloc<double>(h, g0.ndof);
        h->Uv = dalloc<double>(h, g0.ndof);
        h->R = dalloc<double>(h, g0.ndof);
        h->P = dalloc<double>(h, g0.ndof);
        h->Q = dalloc<double>(h, g0.ndof);
        h->Z = dalloc<double>(h, g0.ndof);
        h->partials = dalloc<double>(h, maxblocks);
        h->sc = dalloc<CgScalars>(h, 1);
        HIPCHK(hipHostMalloc((void **)&h->sc_host, sizeof(CgScalars), hipHostMallocDefault));
        std::memset(h->sc_host, 0, sizeof(CgScalars));
        HIPCHK(hipMemcpy(h->sc, h->sc_host, sizeof(CgScalars), hipMemcpyHostToDevice));
        Coarse &C = h->C;
        C.n = (int)h->L[nl - 1].g.ndof;
        {
            // DOFs of one element are at most 3(nny*nnz + nny + 1) + 2 apart
            const Grid &gc = h->L[nl - 1].g;
            C.bw = 3 * (gc.nny * gc.nnz + gc.nny + 1) + 2;
            if (h->prm.periodic) C.bw = C.n;  // wrapped couplings: no band
        }
        const long N = (C.n + 31) / 32 * 32;
        if (dense_ok) {
            C.AQ = dalloc<double>(h, N * N);
            for (int b = 0; b < 2; ++b) {
                C.Lbuf[b] = dalloc<double>(h, N * N);
                C.LTbuf[b] = dalloc<double>(h, N * N);
            }
            C.Linv = C.Lbuf[0];
            C.LinvT = C.LTbuf[0];
            C.y = dalloc<double>(h, (long)kMaxBatch * N);  // one per right-hand side of a batch
        }
        C.direct = dense_ok;
        C.flag = dalloc<int>(h, 1);
        std::vector<double> tab = level1_gauss_table(0.3, 2.0);
        HIPCHK(hipMemcpy(h->Aelm, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
        h->gt_max = std::min(nl - 1, kGtMaxDepth);
        for (int l = 1; l <= h->gt_max; ++l) h->gt[l] = dalloc<double>(h, (size_t)300 << (3 * l));
        if (h->prm.precision != GPUTO_PREC_F64) mixed_create(h);
        if (h->prm.learned) learn_create(h);
    } catch (...) {
        int rc = map_exception();
        gputo_destroy(h);
        return rc;
    }
    *out = h;
    return 0;
}

int gputo_destroy(gputo_handle *h)
{
    if (!h) return 0;
    to_release(h);
    learn_destroy(h);
    mixed_destroy(h);
    if (h->stream) hipStreamSynchronize(h->stream);
    if (h->C.side) hipStreamSynchronize(h->C.side);
    for (void *p : h->allocs) hipFree(p);
    if (h->sc_host) hipHostFree(h->sc_host);
    if (h->bt) {
        if (h->bt->sc_host) hipHostFree(h->bt->sc_host);
        delete h->bt;
    }
    if (h->stream) hipStreamDestroy(h->stream);
    if (h->C.side) hipStreamDestroy(h->C.side);
    if (h->C.ev_in) hipEventDestroy(h->C.ev_in);
    if (h->C.ev_done) hipEventDestroy(h->C.ev_done);
    delete h;
    return 0;
}

int gputo_set_params(gputo_handle *h, const gputo_params *params)
{
    if (!h || !params) return GPUTO_ERR_ARGUMENT;
    if (!!params->learned != !!h->prm.learned) return GPUTO_ERR_ARGUMENT;  // fixed at creation
    if (!!params->periodic != !!h->prm.periodic) return GPUTO_ERR_ARGUMENT;
    if (!schedule_ok(*params)) return GPUTO_ERR_ARGUMENT;
    if (params->precision != h->prm.precision && params->precision == GPUTO_PREC_F64 && h->ln)
        return GPUTO_ERR_ARGUMENT;  // the learned stage runs in the FP32 V-cycle
    if (!h->C.AQ && params->preset != GPUTO_PRESET_MGCG9) return GPUTO_ERR_WORKSPACE;  // no dense coarsest
    if (params->preset != h->prm.preset) h->hier_ready = false;
    coarse_invalidate(h->C);  // the next setup factors exactly
    if (h->ln) learn_policy_reset(h->ln);  // the fallback re-times the stage under the new parameters
    h->prm = *params;
    try {
        if (h->prm.precision != GPUTO_PREC_F64) mixed_create(h);
    } catch (...) {
        return map_exception();
    }
    return 0;
}

static int upload_problem(gputo_handle *h, const double *Ke0, const double *CX, const double *fixed, int nfd,
                          std::vector<int> &fx)
{
    const Grid &g0 = h->L[0].g;
    if (!Ke0 || !CX || (nfd > 0 && !fixed) || nfd < 0 || nfd > g0.ndof) return GPUTO_ERR_ARGUMENT;
    if (!all_finite(Ke0, 576) || !all_finite(CX, g0.nel)) return GPUTO_ERR_NONFINITE;
    fx.resize(nfd);
    for (int n = 0; n < nfd; ++n) fx[n] = (int)fixed[n] - 1;  // (int)FDofs[n]-1, mgcg1.c:663
    set_ke(h, Ke0);
    HIPCHK(hipMemcpyAsync(h->CX, CX, g0.nel * sizeof(double), hipMemcpyHostToDevice, h->stream));
    return 0;
}

int gputo_setup(gputo_handle *h, const double *Ke0, const double *CX, const double *fixed, int nfd)
{
    if (!h) return GPUTO_ERR_ARGUMENT;
    try {
        std::vector<int> fx;
        if (int rc = upload_problem(h, Ke0, CX, fixed, nfd, fx)) return rc;
        if (int rc = setup_hierarchy(h, fx)) return rc;
        if (int rc = coarse_check(h->C)) {
            h->hier_ready = false;
            coarse_invalidate(h->C);
            return rc;
        }
        return 0;
    } catch (...) {
        return map_exception();
    }
}

int gputo_mgcg1(gputo_handle *h, const double *Ke0, const double *F, double *U, const double *CX, const double *fixed,
                int nfd, double tol, int maxit, double *relres, int *nit)
{
    if (!h || !F || !U || !relres || !nit) return GPUTO_ERR_ARGUMENT;
    *relres = 0.0;
    *nit = GPUTO_ERR_ARGUMENT;
    try {
        const Grid &g0 = h->L[0].g;
        std::vector<int> fx;
        int rc = upload_problem(h, Ke0, CX, fixed, nfd, fx);
        if (rc == 0 && !all_finite(F, g0.ndof)) rc = GPUTO_ERR_NONFINITE;
        if (rc) {
            *nit = rc;
            return rc;
        }
        HIPCHK(hipMemcpyAsync(h->Fv, F, g0.ndof * sizeof(double), hipMemcpyHostToDevice, h->stream));
        rc = setup_hierarchy(h, fx);
        if (rc == 0) {
            rc = run_pcg(h, h->Fv, h->Uv, tol, maxit, relres, nit);
        }
        if (rc) {
            *nit = rc;
            return rc;
        }
        HIPCHK(hipMemcpyAsync(U, h->Uv, g0.ndof * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        for (long n = 1; n < g0.ndof; n += 3) U[n] = -U[n];  // MATLAB frame (mgcg1.cu:282-283)
        return 0;
    } catch (...) {
        int rc = map_exception();
        *nit = rc;
        return rc;
    }
}

int gputo_solve_d(gputo_handle *h, const double *Ke0, const double *d_F, double *d_U, const double *d_CX,
                  const int *d_fixed, int nfd, double tol, int maxit, double *relres, int *nit)
{
    if (!h || !Ke0 || !d_F || !d_U || !d_CX || !relres || !nit || nfd < 0) return GPUTO_ERR_ARGUMENT;
    try {
        const Grid &g0 = h->L[0].g;
        std::vector<int> fx(nfd);
        if (nfd) HIPCHK(hipMemcpyAsync(fx.data(), d_fixed, nfd * sizeof(int), hipMemcpyDeviceToHost, h->stream));
        set_ke(h, Ke0);
        HIPCHK(hipMemcpyAsync(h->CX, d_CX, g0.nel * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        int rc = setup_hierarchy(h, fx);
        if (rc == 0) {
            rc = run_pcg(h, d_F, d_U, tol, maxit, relres, nit);
        }
        if (rc) *nit = rc;
        return rc;
    } catch (...) {
        int rc = map_exception();
        *nit = rc;
        return rc;
    }
}

// ---- kernel-level entry points ----------------------------------------------

static int host_op(gputo_handle *h, int mode, const double *Ke0, const double *CX, const double *F, const double *u,
                   const double *fixed, int nfd, double *y)
{
    if (!h || !y || (mode != 2 && !u)) return GPUTO_ERR_ARGUMENT;
    try {
        const Grid &g0 = h->L[0].g;
        std::vector<int> fx;
        if (int rc = upload_problem(h, Ke0, CX, fixed, nfd, fx)) return rc;
        if (int rc = validate_fixed(fx, g0.ndof)) return rc;
        std::vector<uint8_t> m = node_mask(fx, g0);
        hipStream_t s = h->stream;
        HIPCHK(hipMemcpyAsync(h->L[0].mask, m.data(), m.size(), hipMemcpyHostToDevice, s));
        h->masks_ready = false;  // the level-0 mask no longer matches L[0].fixed
        if (mode == 2) {
            launch_diag0(s, g0, h->Ke, h->CX, h->Q);
        } else {
            HIPCHK(hipMemcpyAsync(h->P, u, g0.ndof * sizeof(double), hipMemcpyHostToDevice, s));
            if (F) HIPCHK(hipMemcpyAsync(h->Fv, F, g0.ndof * sizeof(double), hipMemcpyHostToDevice, s));
            OpArgs a = op_args(h, 0);
            a.ctl = nullptr;
            a.cx_floor = 0.0;  // multAV0c / defect0c of mgcg1 keep every element
            a.U = h->P;
            a.F = h->Fv;
            a.out = h->Q;
            launch_op0(s, mode == 0 ? OP_MATVEC : OP_DEFECT, a);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(y, h->Q, g0.ndof * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        h->hier_ready = false;  // the level-0 mask was replaced
        coarse_invalidate(h->C);
        return 0;
    } catch (...) {
        return map_exception();
    }
}

int gputo_matvec_f64(gputo_handle *h, const double *Ke0, const double *CX, const double *u, const double *fixed, int nfd,
                     double *y)
{
    return host_op(h, 0, Ke0, CX, nullptr, u, fixed, nfd, y);
}

int gputo_matvec_mixed(gputo_handle *h, int kernel, const double *Ke0, const double *CX, const double *u,
                       const double *fixed, int nfd, double *y)
{
    if (!h || !u || !y || kernel < 1 || kernel > 5) return GPUTO_ERR_ARGUMENT;
    try {
        const Grid &g0 = h->L[0].g;
        std::vector<int> fx;
        if (int rc = upload_problem(h, Ke0, CX, fixed, nfd, fx)) return rc;
        if (int rc = validate_fixed(fx, g0.ndof)) return rc;
        std::vector<uint8_t> m = node_mask(fx, g0);
        hipStream_t s = h->stream;
        HIPCHK(hipMemcpyAsync(h->L[0].mask, m.data(), m.size(), hipMemcpyHostToDevice, s));
        h->masks_ready = false;  // the level-0 mask no longer matches L[0].fixed
        HIPCHK(hipMemcpyAsync(h->P, u, g0.ndof * sizeof(double), hipMemcpyHostToDevice, s));
        mixed_apply_l0(h, kernel, h->P, h->Q);
        HIPCHK(hipMemcpyAsync(y, h->Q, g0.ndof * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        h->hier_ready = false;
        coarse_invalidate(h->C);
        return 0;
    } catch (...) {
        return map_exception();
    }
}

int gputo_defect_f64(gputo_handle *h, const double *Ke0, const double *CX, const double *F, const double *u,
                     const double *fixed, int nfd, double *r)
{
    if (!F) return GPUTO_ERR_ARGUMENT;
    return host_op(h, 1, Ke0, CX, F, u, fixed, nfd, r);
}

int gputo_diag_f64(gputo_handle *h, const double *Ke0, const double *CX, double *ad)
{
    return host_op(h, 2, Ke0, CX, nullptr, nullptr, nullptr, 0, ad);
}

int gputo_vcycle_f64(gputo_handle *h, const double *Ke0, const double *CX, const double *fixed, int nfd, const double *r,
                     double *z)
{
    if (!h || !r || !z) return GPUTO_ERR_ARGUMENT;
    try {
        int rc = gputo_setup(h, Ke0, CX, fixed, nfd);
        if (rc) return rc;
        const Grid &g0 = h->L[0].g;
        hipStream_t s = h->stream;
        std::memset(h->sc_host, 0, sizeof(CgScalars));
        HIPCHK(hipMemcpyAsync(h->sc, h->sc_host, sizeof(CgScalars), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(h->R, r, g0.ndof * sizeof(double), hipMemcpyHostToDevice, s));
        vcycle(h, h->R, h->Z);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(z, h->Z, g0.ndof * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        return 0;
    } catch (...) {
        return map_exception();
    }
}

int gputo_level_info(const gputo_handle *h, int level, int *nelx, int *nely, int *nelz, int *nfd)
{
    if (!h || level < 0 || level >= h->nl) return GPUTO_ERR_ARGUMENT;
    const Level &v = h->L[level];
    if (nelx) *nelx = v.g.nx;
    if (nely) *nely = v.g.ny;
    if (nelz) *nelz = v.g.nz;
    if (nfd) *nfd = (int)v.fixed.size();
    return 0;
}

int gputo_level_matrices(gputo_handle *h, int level, double *A, double *ad, double *fixed)
{
    if (!h || level < 0 || level >= h->nl || !h->hier_ready) return GPUTO_ERR_ARGUMENT;
    try {
        const Level &v = h->L[level];
        hipStream_t s = h->stream;
        if (A && level > 0) {
            std::vector<double> packed(300 * v.g.nel);
            HIPCHK(hipMemcpyAsync(packed.data(), v.A, packed.size() * sizeof(double), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            for (long e = 0; e < v.g.nel; ++e)
                for (int m = 0; m < 24; ++m)
                    for (int n = 0; n < 24; ++n) {
                        const int p = n <= m ? m * (m + 1) / 2 + n : n * (n + 1) / 2 + m;
                        A[576 * e + 24 * m + n] = packed[(long)p * v.g.nel + e];
                    }
        }
        if (ad) HIPCHK(hipMemcpyAsync(ad, v.AD, v.g.ndof * sizeof(double), hipMemcpyDeviceToHost, s));
        if (fixed)
            for (size_t n = 0; n < v.fixed.size(); ++n) fixed[n] = v.fixed[n] + 1.0;
        HIPCHK(hipStreamSynchronize(s));
        return 0;
    } catch (...) {
        return map_exception();
    }
}

int gputo_level_cx(gputo_handle *h, int level, double *cx)
{
    if (!h || level < 0 || level >= h->nl || !h->hier_ready || !cx || !preset_of(h->prm).mg9)
        return GPUTO_ERR_ARGUMENT;
    try {
        const Level &v = h->L[level];
        HIPCHK(hipMemcpyAsync(cx, level == 0 ? h->CX : v.CXl, v.g.nel * sizeof(double), hipMemcpyDeviceToHost,
                              h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        return 0;
    } catch (...) {
        return map_exception();
    }
}

int gputo_fixed_list(const gputo_handle *h, int level, double *fixed, int *nfd)
{
    if (!h || level < 0 || level >= h->nl) return GPUTO_ERR_ARGUMENT;
    const Level &v = h->L[level];
    if (nfd) *nfd = (int)v.fixed.size();
    if (fixed)
        for (size_t n = 0; n < v.fixed.size(); ++n) fixed[n] = v.fixed[n] + 1.0;
    return 0;
}

int gputo_restrict_f64(gputo_handle *h, int level, const double *r_fine, double *f_coarse)
{
    if (!h || level < 0 || level + 1 >= h->nl || !h->hier_ready || !r_fine || !f_coarse) return GPUTO_ERR_ARGUMENT;
    try {
        Level &vf = h->L[level], &vc = h->L[level + 1];
        hipStream_t s = h->stream;
        HIPCHK(hipMemcpyAsync(vf.T, r_fine, vf.g.ndof * sizeof(double), hipMemcpyHostToDevice, s));
        launch_restrict(s, vc.g, vf.T, vc.T, vc.mask, nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(f_coarse, vc.T, vc.g.ndof * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        return 0;
    } catch (...) {
        return map_exception();
    }
}

int gputo_interp_f64(gputo_handle *h, int level, const double *u_coarse, double *u_fine)
{
    if (!h || level < 0 || level + 1 >= h->nl || !h->hier_ready || !u_coarse || !u_fine) return GPUTO_ERR_ARGUMENT;
    try {
        Level &vf = h->L[level], &vc = h->L[level + 1];
        hipStream_t s = h->stream;
        HIPCHK(hipMemcpyAsync(vc.T, u_coarse, vc.g.ndof * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(vf.T, u_fine, vf.g.ndof * sizeof(double), hipMemcpyHostToDevice, s));
        launch_interp(s, vf.g, vc.T, vf.T, vf.mask, 0, nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(u_fine, vf.T, vf.g.ndof * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        return 0;
    } catch (...) {
        return map_exception();
    }
}

int gputo_bench_operator(gputo_handle *h, int kernel, int reps, double *avg_ms)
{
    if (!h || !h->hier_ready || reps < 1 || !avg_ms || kernel < 0 || kernel > 5) return GPUTO_ERR_ARGUMENT;
    try {
        hipStream_t s = h->stream;
        const Grid &g0 = h->L[0].g;
        launch_fill(s, h->P, g0.ndof, 1.0);
        if (kernel != 0) return mixed_bench_operator(h, kernel, reps, avg_ms);
        OpArgs a = op_args(h, 0);
        a.ctl = nullptr;
        a.U = h->P;
        a.out = h->Q;
        launch_op0(s, OP_MATVEC, a);  // warm-up
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) launch_op0(s, OP_MATVEC, a);
        HIPCHK(hipEventRecord(e1, s));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        HIPCHK(hipGetLastError());
        *avg_ms = ms / reps;
        return 0;
    } catch (...) {
        return map_exception();
    }
}

}  // extern "C"

namespace {

// Launch one level-`level` operator application of mode `mode` (0 matvec,
// 1 defect, 2 damped Jacobi) with operator `op` on device vectors.
template <class T>
void level_apply_t(gputo_handle *h, HierView<T> H, int level, int op, int mode, const T *u, const T *f, T *y,
                   T *tmp = nullptr)
{
    H.l1_mf = (level == 1 && op == 2);
    OpArgsT<T> a = view_args(h, H, level, nullptr);
    a.U = u;
    a.F = f;
    a.out = y;
    a.out2 = tmp;
    static const int kModes[5] = {OP_MATVEC, OP_DEFECT, OP_JACOBI, OP_PRE_JACOBI, OP_PRE_DEFECT};
    level_op(h, H, level, kModes[mode], a);
}

int level_args_ok(const gputo_handle *h, int level, int op, int mode)
{
    if (!h || !h->hier_ready || level < 0 || level >= h->nl - 1 || mode < 0 || mode > 4) return 0;
    if (op == 2) return level == 1 && !preset_of(h->prm).l1_galerkin && !preset_of(h->prm).mg9;
    return op == 1;
}

}  // namespace

extern "C" {

int gputo_level_apply(gputo_handle *h, int level, int op, int mode, int fp32, const double *u, const double *f,
                      double *y)
{
    if (!level_args_ok(h, level, op, mode) || !u || !y || (mode > 0 && !f)) return GPUTO_ERR_ARGUMENT;
    try {
        hipStream_t s = h->stream;
        const long n = h->L[level].g.ndof;
        HIPCHK(hipMemcpyAsync(h->P, u, n * sizeof(double), hipMemcpyHostToDevice, s));
        if (f) HIPCHK(hipMemcpyAsync(h->Z, f, n * sizeof(double), hipMemcpyHostToDevice, s));
        if (fp32) {
            HierView<float> H = mixed_level_view(h);
            MixedState *m = h->mx;
            launch_copy<double, float>(s, h->P, m->P, n, nullptr);
            launch_copy<double, float>(s, h->Z, m->X, n, nullptr);
            level_apply_t<float>(h, H, level, op, mode, m->P, m->X, m->Q, m->Z);
            launch_copy<float, double>(s, m->Q, h->Q, n, nullptr);
        } else {
            level_apply_t<double>(h, f64_view(h), level, op, mode, h->P, h->Z, h->Q, h->R);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(y, h->Q, n * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        return 0;
    } catch (...) {
        return map_exception();
    }
}

int gputo_bench_level(gputo_handle *h, int level, int op, int fp32, int reps, double *avg_ms)
{
    if (!level_args_ok(h, level, op, 0) || reps < 1 || !avg_ms) return GPUTO_ERR_ARGUMENT;
    try {
        hipStream_t s = h->stream;
        const long n = h->L[level].g.ndof;
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        if (fp32) {
            HierView<float> H = mixed_level_view(h);
            MixedState *m = h->mx;
            launch_fill<float>(s, m->P, n, 1.0);
            level_apply_t<float>(h, H, level, op, 0, m->P, nullptr, m->Q);  // warm-up
            HIPCHK(hipEventRecord(e0, s));
            for (int r = 0; r < reps; ++r) level_apply_t<float>(h, H, level, op, 0, m->P, nullptr, m->Q);
        } else {
            const HierView<double> H = f64_view(h);
            launch_fill<double>(s, h->P, n, 1.0);
            level_apply_t<double>(h, H, level, op, 0, h->P, nullptr, h->Q);
            HIPCHK(hipEventRecord(e0, s));
            for (int r = 0; r < reps; ++r) level_apply_t<double>(h, H, level, op, 0, h->P, nullptr, h->Q);
        }
        HIPCHK(hipEventRecord(e1, s));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        HIPCHK(hipGetLastError());
        *avg_ms = ms / reps;
        return 0;
    } catch (...) {
        return map_exception();
    }
}

int gputo_coarse_solve_f64(gputo_handle *h, const double *b, double *x)
{
    if (!h || !h->hier_ready || !b || !x) return GPUTO_ERR_ARGUMENT;
    try {
        Level &v = h->L[h->nl - 1];
        hipStream_t s = h->stream;
        HIPCHK(hipMemcpyAsync(v.F, b, v.g.ndof * sizeof(double), hipMemcpyHostToDevice, s));
        launch_coarse_solve(s, h->C, v.F, v.T, v.mask, nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(x, v.T, v.g.ndof * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        return 0;
    } catch (...) {
        return map_exception();
    }
}

}  // extern "C"

namespace {

// a.b over the full grid, partials summed on the host in block order
double dot_host(gputo_handle *h, const double *a, const double *b, long n)
{
    launch_dot_partial(h->stream, a, b, n, h->partials, nullptr);
    std::vector<double> p(kRedBlocks);
    HIPCHK(hipMemcpyAsync(p.data(), h->partials, kRedBlocks * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    double acc = 0.0;
    for (double v : p) acc += v;
    return acc;
}

}  // namespace

extern "C" {

int gputo_homogenize(gputo_handle *h, const double *Ke0, const double *CX, double tol, int maxit, double *CH,
                     double *relres, int *nit, double *chi)
{
    if (!h || !Ke0 || !CX || !CH || !h->prm.periodic) return GPUTO_ERR_ARGUMENT;
    try {
        const Grid &g0 = h->L[0].g;
        const long n = g0.ndof;
        if (!all_finite(Ke0, 576) || !all_finite(CX, g0.nel)) return GPUTO_ERR_NONFINITE;
        hipStream_t s = h->stream;
        set_ke(h, Ke0);
        HIPCHK(hipMemcpyAsync(h->CX, CX, g0.nel * sizeof(double), hipMemcpyHostToDevice, s));
        // node 0 and its 7 images are pinned: the periodic operator's null
        // space is the three rigid translations
        std::vector<int> fixed;
        for (int a = 0; a < 2; ++a)
            for (int c = 0; c < 2; ++c)
                for (int b = 0; b < 2; ++b) {
                    const long nod = (long)g0.nny * g0.nnz * (a * g0.nx) + (long)g0.nny * (c * g0.nz) + b * g0.ny;
                    for (int d = 0; d < 3; ++d) fixed.push_back((int)(3 * nod + d));
                }
        if (int rc = setup_hierarchy(h, fixed)) return rc;
        if (!h->hom) h->hom = dalloc<double>(h, 12 * (size_t)n);
        double *Fq = h->hom, *Xq = h->hom + 6 * n;
        double fe[6][24], Q0[6][6];
        homog_loads(Ke0, fe, Q0);  // Ke0 read row-major, as the operators read it
        launch_homog_rhs(s, g0, h->CX, fe, h->L[0].mask, Fq);
        HIPCHK(hipGetLastError());
        if (h->prm.homog_sequential) {
            for (int q = 0; q < 6; ++q) {
                double rr = 0.0;
                int ni = 0;
                int rc = run_pcg(h, Fq + q * n, Xq + q * n, tol, maxit, &rr, &ni);
                if (relres) relres[q] = rr;
                if (nit) nit[q] = rc ? rc : ni;
                if (rc) return rc;
            }
        } else {
            double rr[6];
            int ni[6];
            const int rc = run_pcg_batch(h, Fq, Xq, 6, tol, maxit, rr, ni);
            for (int q = 0; q < 6; ++q) {
                if (relres) relres[q] = rr[q];
                if (nit) nit[q] = ni[q];
            }
            if (rc) return rc;
        }
        // CH_ab |Y| = sum_e CX_e (u0_a - chi_a)^T Ke0 (u0_b - chi_b)
        //           = Q0_ab sum CX - chi_a.f_b - chi_b.f_a + chi_a.K chi_b
        double sumcx = 0.0;
        for (long e = 0; e < g0.nel; ++e) sumcx += CX[e];
        double XF[6][6], XKX[6][6];
        for (int a = 0; a < 6; ++a)
            for (int b = 0; b < 6; ++b) XF[a][b] = dot_host(h, Xq + a * n, Fq + b * n, n);
        for (int b = 0; b < 6; ++b) {
            OpArgs o = op_args(h, 0);
            o.ctl = nullptr;
            o.cx_floor = 0.0;
            o.U = Xq + b * n;
            o.out = h->Q;
            o.partials = nullptr;
            launch_op0(s, OP_MATVEC, o);  // unfolded: chi_a is equal on images
            for (int a = 0; a <= b; ++a) XKX[a][b] = XKX[b][a] = dot_host(h, Xq + a * n, h->Q, n);
        }
        const double vol = (double)g0.nel;
        for (int a = 0; a < 6; ++a)
            for (int b = 0; b < 6; ++b)
                CH[6 * a + b] = (Q0[a][b] * sumcx - XF[a][b] - XF[b][a] + XKX[a][b]) / vol;
        if (chi) {
            HIPCHK(hipMemcpyAsync(chi, Xq, 6 * n * sizeof(double), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
        }
        return 0;
    } catch (...) {
        return map_exception();
    }
}

}  // extern "C"
