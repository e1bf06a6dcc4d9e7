// Mixed precision: the PCG recurrences, the level-0 matvec, every dot
// product, alpha/beta and the coarsest Cholesky solve stay FP64; the
// preconditioner -- the whole V-cycle, i.e. nearly all of the work -- runs on
// FP32 copies of the hierarchy with FP32 accumulation inside the operators.
// Rounding in the preconditioner changes only how fast CG converges, never
// what it converges to.  (An FP32 Krylov loop does not reach 1e-8: the FP32
// operator applied to a smooth field has a relative error of about
// 6e-8 (L/h)^2, 2e-3 at 192 elements, so its true residual stalls.)  The
// level-0 FP32 operator is the node-gather kernel or one of the MFMA element
// kernels (mfma.hip).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "learn.h"
#include "mixed.h"
#include "vcycle.h"

namespace gputo {

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
            if (h->L[l].S) v.S = m->S32[l] = malloc_dev<float>(m, (long)kStencilEntries * g.nnod);
            v.CX = m->CXl32[l] = malloc_dev<float>(m, g.nel);
        }
    }
    const Grid &g0 = h->L[0].g;
    m->Ke32 = malloc_dev<float>(m, 576);
    m->CX32 = malloc_dev<float>(m, g0.nel);
    m->H.Ke = m->Ke32;
    m->H.CX = m->CX32;
    m->H.L[0].CX = m->CX32;
    m->X = malloc_dev<float>(m, g0.ndof);
    m->R = malloc_dev<float>(m, g0.ndof);
    m->P = malloc_dev<float>(m, g0.ndof);
    m->Q = malloc_dev<float>(m, g0.ndof);
    m->Z = malloc_dev<float>(m, g0.ndof);
    m->Pcg = malloc_dev<float>(m, (long)kCgLog * g0.ndof);
    launch_fill<float>(h->stream, m->Pcg, (long)kCgLog * g0.ndof, 0.0);  // finite: unused slots add 0 * p
    m->alog = malloc_dev<double>(m, kCgLog);
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
    mixed_graph_drop(h);
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

// Level-0 kernel of the FP32 operators: explicit choice, else the FP32
// node-gather for MIXED_F32 and the three-way split-BF16 MFMA for MIXED_BF16.
static int l0_kind(const gputo_params &p, int precision)
{
    if (p.l0_kernel >= L0_GATHER && p.l0_kernel <= L0_MFMA_BF16X1) return p.l0_kernel;
    return precision == GPUTO_PREC_MIXED_BF16 ? L0_MFMA_BF16X3 : L0_GATHER;
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
    m->H.mg9 = preset_of(h->prm).mg9;
    m->H.per = h->prm.periodic != 0;
    if (m->H.mg9 && !h->C.direct) launch_fill<float>(s, m->H.L[h->nl - 1].U, h->L[h->nl - 1].g.ndof, 0.0);
    for (int l = 0; l < h->nl; ++l) {
        const Level &v = h->L[l];
        launch_copy<double, float>(s, v.AD, m->H.L[l].AD, v.g.ndof, nullptr);
        if (m->H.mg9) {
            if (l > 0) launch_copy<double, float>(s, v.CXl, m->CXl32[l], v.g.nel, nullptr);
        } else if (l > 0 && !(l == 1 && m->H.l1_mf)) {
            // the stencil replaces the element matrices where it exists
            if (m->S32[l])
                launch_copy<double, float>(s, v.S, m->S32[l], (long)kStencilEntries * v.g.nnod, nullptr);
            else
                launch_copy<double, float>(s, v.A, m->A32[l], 300 * v.g.nel, nullptr);
        }
    }
    m->H.l0 = l0_kind(h->prm, h->prm.precision);
    m->H.l0_ctx = m;
    m->H.learn = h->ln != nullptr;
}


// z = V(r) with the FP32 V-cycle: r (FP64) is rounded to FP32, the whole
// cycle runs on the FP32 copies of the hierarchy (FP64 coarsest Cholesky
// inside), and z is widened back to FP64 for the FP64 Krylov recurrences.
void mixed_vcycle(gputo_handle *h, const double *r, double *z, const int *ctl, bool r32_ready, bool pre0_ready)
{
    MixedState *m = h->mx;
    hipStream_t s = h->stream;
    const long n = h->L[0].g.ndof;
    if (!r32_ready) launch_copy<double, float>(s, r, m->R, n, ctl);
    m->H.pre0_ready = pre0_ready;
    vcycle_t<float>(h, m->H, m->R, m->Z, ctl);
    m->H.pre0_ready = false;
    if (z) launch_copy<float, double>(s, m->Z, z, n, ctl);
}

float *mixed_prerelax_target(gputo_handle *h, const float **ad, double *prerelax)
{
    MixedState *m = h->mx;
    const HierView<float> &H = m->H;
    // the condition under which the V-cycle's level-0 pre-smoothing starts
    // with the elementwise prerelaxation (vcycle_t, sweeps_t)
    const int npre = sweeps_of<float>(h, 0, false);
    if (h->nl < 2 || H.per || H.mg9 || H.l0 < L0_MFMA_F32 || npre <= 0) return nullptr;
    *ad = H.L[0].AD;
    *prerelax = preset_of(h->prm).prerelax;
    return npre % 2 == 1 ? H.L[0].Tmp : m->Z;
}

void mixed_graph_drop(gputo_handle *h)
{
    MixedState *m = h->mx;
    if (!m || !m->vc_exec) return;
    (void)hipGraphExecDestroy(m->vc_exec);
    m->vc_exec = nullptr;
}

void mixed_vcycle_graph(gputo_handle *h, const int *ctl, bool pre0_ready)
{
    MixedState *m = h->mx;
    hipStream_t s = h->stream;
    const int kind = m->H.learn ? learn_stage_kind(h) : 0;
    if (kind < 0) {
        mixed_vcycle(h, nullptr, nullptr, ctl, true, pre0_ready);
        return;
    }
    if (!m->vc_exec || kind != m->vc_kind || ctl != m->vc_ctl || pre0_ready != m->vc_pre0) {
        mixed_graph_drop(h);
        // recorded, not run: the host side (learn_stage's count included)
        // runs as for an eager V-cycle
        hipGraph_t g = nullptr;
        check_hip(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "graph capture");
        try {
            mixed_vcycle(h, nullptr, nullptr, ctl, true, pre0_ready);
        } catch (...) {
            m->H.pre0_ready = false;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        check_hip(hipStreamEndCapture(s, &g), "graph capture");
        const hipError_t e = hipGraphInstantiate(&m->vc_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) m->vc_exec = nullptr;
        check_hip(e, "graph instantiate");
        m->vc_kind = kind;
        m->vc_ctl = ctl;
        m->vc_pre0 = pre0_ready;
    } else if (m->H.learn) {
        learn_stage_count(h);
    }
    check_hip(hipGraphLaunch(m->vc_exec, s), "graph launch");
}

float *mixed_r32(gputo_handle *h) { return h->mx->R; }
const float *mixed_z32(gputo_handle *h) { return h->mx->Z; }
float *mixed_p32(gputo_handle *h) { return h->mx->Pcg; }
double *mixed_alog(gputo_handle *h) { return h->mx->alog; }

// Start of a mixed-precision solve: FP32 copies of the current hierarchy.
void mixed_begin(gputo_handle *h)
{
    if (!h->mx) mixed_create(h);
    mixed_graph_drop(h);
    mixed_prepare(h);
    h->mx->H.pre0_ready = false;
}

void mixed_end(gputo_handle *h, int nit, int converged)
{
    if (!h->mx) return;
    mixed_graph_drop(h);
    h->mx->last_inner = converged ? nit + 1 : nit;
    h->mx->last_outer = 1;
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
