// Mixed-precision solve: FP64 iterative refinement around an inner PCG
// whose vectors, matvec and smoother are FP32 (FP32 accumulation).  The
// outer residual f - K u, every dot product (accumulated in FP64), alpha and
// beta, and the coarsest-level Cholesky solve stay in FP64.  The level-0
// operator of the inner solve is either the FP32 node-gather kernel or the
// split-BF16 MFMA element kernel (mfma.hip).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "mixed.h"
#include "vcycle.h"

namespace gputo {

int setup_hierarchy(gputo_handle *h, const std::vector<int> &fixed0);

template <class T>
static T *malloc_dev(MixedState *m, size_t n)
{
    void *p = nullptr;
    check_hip(hipMalloc(&p, (n ? n : 1) * sizeof(T)), "hipMalloc");
    m->allocs.push_back(p);
    m->bytes += (long long)(n * sizeof(T));
    return static_cast<T *>(p);
}

void mixed_create(gputo_handle *h)
{
    if (h->mx) return;
    MixedState *m = new MixedState();
    h->mx = m;
    for (int l = 0; l < h->nl; ++l) {
        const Grid &g = h->L[l].g;
        LevelView<float> &v = m->H.L[l];
        v.U = malloc_dev<float>(m, g.ndof);
        v.F = malloc_dev<float>(m, g.ndof);
        v.R = malloc_dev<float>(m, g.ndof);
        v.AD = malloc_dev<float>(m, g.ndof);
        v.Tmp = malloc_dev<float>(m, g.ndof);
        if (l > 0) {
            float *A = malloc_dev<float>(m, 300 * g.nel);
            m->A32[l] = A;
            v.A = A;
        }
    }
    const Grid &g0 = h->L[0].g;
    m->Ke32 = malloc_dev<float>(m, 576);
    m->CX32 = malloc_dev<float>(m, g0.nel);
    m->H.Ke = m->Ke32;
    m->H.CX = m->CX32;
    m->X = malloc_dev<float>(m, g0.ndof);
    m->R = malloc_dev<float>(m, g0.ndof);
    m->P = malloc_dev<float>(m, g0.ndof);
    m->Q = malloc_dev<float>(m, g0.ndof);
    m->Z = malloc_dev<float>(m, g0.ndof);
    m->sc = malloc_dev<CgScalars>(m, 1);
    m->norm = malloc_dev<double>(m, 2);
    check_hip(hipHostMalloc((void **)&m->sc_host, sizeof(CgScalars), hipHostMallocDefault), "hipHostMalloc");
    std::memset(m->sc_host, 0, sizeof(CgScalars));
    mfma_create(h, m);
    h->bytes += m->bytes;
}

void mixed_destroy(gputo_handle *h)
{
    MixedState *m = h->mx;
    if (!m) return;
    mfma_destroy(h, m);
    for (void *p : m->allocs) (void)hipFree(p);
    if (m->sc_host) (void)hipHostFree(m->sc_host);
    delete m;
    h->mx = nullptr;
}

void mixed_stats(gputo_handle *h, gputo_to_stats *st)
{
    if (!h->mx) return;
    st->inner_iters = h->mx->last_inner;
    st->outer_iters = h->mx->last_outer;
    st->learned_accepted = h->mx->last_learned_ok;
    st->learned_rejected = h->mx->last_learned_bad;
}

// Level-0 kernel of the inner solve: explicit choice, else MFMA FP32 for
// MIXED_F32 and the three-way split-BF16 MFMA for MIXED_BF16.
static int l0_kind(const gputo_params &p, int precision)
{
    if (p.l0_kernel >= L0_GATHER && p.l0_kernel <= L0_MFMA_BF16X3) return p.l0_kernel;
    return precision == GPUTO_PREC_MIXED_BF16 ? L0_MFMA_BF16X3 : L0_MFMA_F32;
}

// FP32 copies of the FP64 hierarchy (operators, diagonals) built by
// setup_hierarchy.
static void mixed_prepare(gputo_handle *h)
{
    MixedState *m = h->mx;
    hipStream_t s = h->stream;
    const Grid &g0 = h->L[0].g;
    launch_copy<double, float>(s, h->Ke, m->Ke32, 576, nullptr);
    launch_copy<double, float>(s, h->CX, m->CX32, g0.nel, nullptr);
    m->H.l1_mf = use_l1_mf(h->prm, h->nl);
    for (int l = 0; l < h->nl; ++l) {
        const Level &v = h->L[l];
        launch_copy<double, float>(s, v.AD, m->H.L[l].AD, v.g.ndof, nullptr);
        if (l > 0 && !(l == 1 && m->H.l1_mf)) launch_copy<double, float>(s, v.A, m->A32[l], 300 * v.g.nel, nullptr);
    }
    m->H.l0 = l0_kind(h->prm, h->prm.precision);
    m->H.l0_ctx = m;
    if (m->H.l0 == L0_MFMA_BF16) mfma_prepare(h, m);
}

// r32 = r / scale; d32 = 0
__global__ void k_scale_down(const double *__restrict__ r, float *__restrict__ r32, float *__restrict__ x32, long n,
                             const double *scale)
{
    const double inv = 1.0 / scale[0];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        r32[i] = (float)(r[i] * inv);
        x32[i] = 0.0f;
    }
}

// u += scale * d32
__global__ void k_scale_up(double *__restrict__ u, const float *__restrict__ d32, long n, const double *scale)
{
    const double sc = scale[0];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        u[i] += sc * (double)d32[i];
}

// ||r|| from the partials into scale[0] (single block, fixed order)
__global__ __launch_bounds__(kBlock) void k_norm(const double *p, int n, double *scale)
{
    __shared__ double ws[kBlock / 64];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += kBlock) s += p[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < kBlock / 64; ++i) t += ws[i];
        scale[0] = sqrt(t);
    }
}

static unsigned nblk(long n)
{
    long nb = (n + kBlock - 1) / kBlock;
    return (unsigned)std::min<long>(std::max<long>(nb, 1), 4096);
}

// Inner PCG in FP32: K d = r32 to relative tolerance itol (same device-side
// scalar machinery as the FP64 loop).  Returns the iterations taken.
static int inner_pcg(gputo_handle *h, double itol, int maxit)
{
    MixedState *m = h->mx;
    hipStream_t s = h->stream;
    const long n = h->L[0].g.ndof;
    const int *ctl = &m->sc->done;
    const HierView<float> &H = m->H;
    launch_dot_partial<float, float>(s, m->R, m->R, n, h->partials, nullptr);
    launch_cg_init(s, m->sc, h->partials, kRedBlocks);
    vcycle_t<float>(h, H, m->R, m->P, ctl);
    launch_copy<float, float>(s, m->P, m->Z, n, ctl);
    launch_dot_partial<float, float>(s, m->R, m->Z, n, h->partials, ctl);
    launch_cg_rho0(s, m->sc, h->partials, kRedBlocks);
    const int chunk = 4;
    int enq = 0;
    const int nmv = (int)h->L[0].t.blocks();
    for (;;) {
        for (int c = 0; c < chunk && enq < maxit; ++c, ++enq) {
            OpArgsT<float> a = view_args(h, H, 0, ctl);
            a.U = m->P;
            a.out = m->Q;
            a.partials = h->partials;
            level_op(h, H, 0, OP_MATVEC, a);
            launch_cg_alpha(s, m->sc, h->partials, H.l0 >= L0_MFMA_F32 ? mfma_blocks(h) : nmv);
            launch_cg_update_ur<float>(s, m->X, m->R, m->P, m->Q, n, m->sc, h->partials);
            launch_cg_check(s, m->sc, h->partials, kRedBlocks, itol, maxit);
            vcycle_t<float>(h, H, m->R, m->Z, ctl);
            launch_dot_partial<float, float>(s, m->R, m->Z, n, h->partials, ctl);
            launch_cg_beta(s, m->sc, h->partials, kRedBlocks, maxit);
            launch_cg_update_p<float>(s, m->P, m->Z, n, m->sc);
        }
        check_hip(hipGetLastError(), "launch");
        check_hip(hipMemcpyAsync(m->sc_host, m->sc, sizeof(CgScalars), hipMemcpyDeviceToHost, s), "copy");
        check_hip(hipStreamSynchronize(s), "sync");
        if (m->sc_host->done || enq >= maxit) break;
    }
    const CgScalars &r = *m->sc_host;
    if (r.bad) return -1;
    return r.done ? r.nit + 1 : maxit;
}

int mixed_solve(gputo_handle *h, const double *F, double *U, double tol, int maxit, double *relres, int *nit)
{
    if (!h->mx) mixed_create(h);
    MixedState *m = h->mx;
    hipStream_t s = h->stream;
    const Grid &g0 = h->L[0].g;
    const long n = g0.ndof;
    mixed_prepare(h);
    // outer state in FP64: U = 0, R = F, res0 = ||F||
    launch_fill<double>(s, U, n, 0.0);
    launch_copy<double, double>(s, F, h->R, n, nullptr);
    launch_dot_partial<double, double>(s, F, F, n, h->partials, nullptr);
    k_norm<<<1, kBlock, 0, s>>>(h->partials, kRedBlocks, m->norm);
    check_hip(hipMemcpyAsync(&m->norm_host[0], m->norm, sizeof(double), hipMemcpyDeviceToHost, s), "copy");
    check_hip(hipStreamSynchronize(s), "sync");
    const double res0 = m->norm_host[0];
    int total = 0, outer = 0;
    double rel = 0.0;
    if (!(res0 < INFINITY)) return GPUTO_ERR_NONFINITE;
    if (res0 == 0.0) {
        *relres = 0.0;
        *nit = 0;
        return 0;
    }
    double rn = res0;
    while (total < maxit) {
        // inner solve of K d = r / ||r|| in FP32
        k_scale_down<<<nblk(n), kBlock, 0, s>>>(h->R, m->R, m->X, n, m->norm);
        const double itol = std::max(h->prm.inner_tol, 0.5 * tol * res0 / rn);
        const int k = inner_pcg(h, itol, maxit - total);
        if (k < 0) return GPUTO_ERR_NONFINITE;
        total += k;
        ++outer;
        // u += ||r|| d;  r = F - K u (FP64, masked)
        k_scale_up<<<nblk(n), kBlock, 0, s>>>(U, m->X, n, m->norm);
        OpArgs a{};
        a.g = g0;
        a.t = h->L[0].t;
        a.Ke = h->Ke;
        a.CX = h->CX;
        a.mask = h->L[0].mask;
        a.U = U;
        a.F = F;
        a.out = h->R;
        launch_op0<double>(s, OP_DEFECT, a);
        launch_dot_partial<double, double>(s, h->R, h->R, n, h->partials, nullptr);
        k_norm<<<1, kBlock, 0, s>>>(h->partials, kRedBlocks, m->norm);
        check_hip(hipMemcpyAsync(&m->norm_host[0], m->norm, sizeof(double), hipMemcpyDeviceToHost, s), "copy");
        check_hip(hipStreamSynchronize(s), "sync");
        rn = m->norm_host[0];
        rel = rn / res0;
        if (h->prm.print_level >= 2)
            std::fprintf(stderr, "gputo: outer %d inner %d total %d relres %.3e\n", outer, k, total, rel);
        if (!(rel < INFINITY)) return GPUTO_ERR_NONFINITE;
        if (rel < tol) break;
    }
    m->last_inner = total;
    m->last_outer = outer;
    *relres = rel;
    *nit = total - 1;  // 0-based like the FP64 loop: iterations taken minus one
    if (h->prm.print_level >= 1)
        std::fprintf(stderr, "gputo: mixed MGCG inner=%d outer=%d relres=%.6g\n", total, outer, rel);
    return 0;
}

// y = K(CX) u with the reduced-precision level-0 kernel `kind` (an L0Kind),
// u and y FP64 device vectors; Ke, CX and the level-0 mask must be current
// on the handle.
int mixed_apply_l0(gputo_handle *h, int kind, const double *u, double *y)
{
    if (!h->mx) mixed_create(h);
    MixedState *m = h->mx;
    hipStream_t s = h->stream;
    const Grid &g0 = h->L[0].g;
    launch_copy<double, float>(s, h->Ke, m->Ke32, 576, nullptr);
    launch_copy<double, float>(s, h->CX, m->CX32, g0.nel, nullptr);
    m->H.l0 = kind;
    m->H.l0_ctx = m;
    launch_copy<double, float>(s, u, m->P, g0.ndof, nullptr);
    OpArgsT<float> a = view_args(h, m->H, 0, nullptr);
    a.U = m->P;
    a.out = m->Q;
    level_op(h, m->H, 0, OP_MATVEC, a);
    launch_copy<float, double>(s, m->Q, y, g0.ndof, nullptr);
    check_hip(hipGetLastError(), "launch");
    return 0;
}

// FP32 view of the current hierarchy with every level's stored matrices
// cast (level-operator checks and benchmarks); vectors X/P/Q are free.
HierView<float> mixed_level_view(gputo_handle *h)
{
    if (!h->mx) mixed_create(h);
    MixedState *m = h->mx;
    mixed_prepare(h);
    if (m->H.l1_mf && h->nl > 1)
        launch_copy<double, float>(h->stream, h->L[1].A, m->A32[1], 300 * h->L[1].g.nel, nullptr);
    return m->H;
}

// mean launch time of the reduced-precision level-0 kernel `kind`
int mixed_bench_operator(gputo_handle *h, int kind, int reps, double *avg_ms)
{
    if (!h->mx) mixed_create(h);
    MixedState *m = h->mx;
    hipStream_t s = h->stream;
    mixed_prepare(h);
    m->H.l0 = kind;
    launch_fill<float>(s, m->P, h->L[0].g.ndof, 1.0);
    OpArgsT<float> a = view_args(h, m->H, 0, nullptr);
    a.U = m->P;
    a.out = m->Q;
    level_op(h, m->H, 0, OP_MATVEC, a);  // warm-up
    hipEvent_t e0, e1;
    check_hip(hipEventCreate(&e0), "event");
    check_hip(hipEventCreate(&e1), "event");
    check_hip(hipEventRecord(e0, s), "event");
    for (int r = 0; r < reps; ++r) level_op(h, m->H, 0, OP_MATVEC, a);
    check_hip(hipEventRecord(e1, s), "event");
    check_hip(hipEventSynchronize(e1), "event");
    float ms = 0.f;
    check_hip(hipEventElapsedTime(&ms, e0, e1), "event");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    check_hip(hipGetLastError(), "launch");
    *avg_ms = ms / reps;
    return 0;
}

}  // namespace gputo
