// Topology-optimisation driver on the device: the design loop of
// Cantilever/omp/top3D125mod.m:105-165 (density filter with symmetric
// padding, volume-preserving Heaviside projection, SIMP interpolation,
// MGCG solve, compliance sensitivities, OC bisection), with every
// element-sized array resident in HBM.  Reductions are two-stage with a
// fixed summation order.
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "gputo_internal.h"
#include "mixed.h"

using namespace gputo;

struct ToState {
    gputo_to_config cfg;
    long nel = 0, ndof = 0;
    double *x = nullptr, *xPhys = nullptr, *xOld = nullptr, *Hs = nullptr, *dHs = nullptr, *dsK = nullptr;
    double *dc = nullptr, *dv = nullptr, *tmp = nullptr, *ocP = nullptr, *CX = nullptr, *F = nullptr, *U = nullptr;
    double *w = nullptr;     // filter weights
    double *part = nullptr;  // reduction partials [32][kRedBlocks]
    double *red = nullptr;   // finalised sums [32]
    double *red_host = nullptr;
    int *fixed_d = nullptr;
    std::vector<int> fixed;
    std::vector<double> Ke;
    int fw = 0;              // filter half width
    int loop = 0;
    double penal = 3, beta = 2, eta = 0.4;
    std::vector<void *> allocs;
};

namespace {

constexpr int kCand = 31;  // OC candidates per pass: a depth-5 bisection tree

template <class T>
T *talloc(ToState *t, size_t n)
{
    void *p = nullptr;
    check_hip(hipMalloc(&p, (n ? n : 1) * sizeof(T)), "hipMalloc");
    t->allocs.push_back(p);
    return static_cast<T *>(p);
}

__device__ __forceinline__ int mirror_idx(int a, int n)
{
    // 'symmetric' padding of imfilter: -1 -> 0, -2 -> 1, n -> n-1
    while (a < 0 || a >= n) a = a < 0 ? -a - 1 : 2 * n - a - 1;
    return a;
}

__device__ __forceinline__ double wsum256(double v)
{
    __shared__ double ws[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    return ws[0] + ws[1] + ws[2] + ws[3];
}

// imfilter(in, h, 'symmetric') on the (nely, nelz, nelx) array = element
// order (top3D125mod.m:96,108,140-141); optional pre-division by `div`.
__global__ __launch_bounds__(256) void k_filter(Grid g, const double *__restrict__ in, const double *__restrict__ div,
                                               const double *__restrict__ w, int fw, double *__restrict__ out)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= g.nel) return;
    const int j = (int)(e % g.ny), k = (int)((e / g.ny) % g.nz), i = (int)(e / ((long)g.ny * g.nz));
    double acc = 0.0;
    int q = 0;
    for (int dx = -fw; dx <= fw; ++dx) {
        const int ii = mirror_idx(i + dx, g.nx);
        for (int dz = -fw; dz <= fw; ++dz) {
            const int kk = mirror_idx(k + dz, g.nz);
            for (int dy = -fw; dy <= fw; ++dy, ++q) {
                const double wq = w[q];
                if (wq == 0.0) continue;
                const long f = (long)g.ny * g.nz * ii + (long)g.ny * kk + mirror_idx(j + dy, g.ny);
                acc += wq * (div ? in[f] / div[f] : in[f]);
            }
        }
    }
    out[e] = acc;
}

__device__ __forceinline__ double prj_d(double v, double eta, double beta)
{
    return (tanh(beta * eta) + tanh(beta * (v - eta))) / (tanh(beta * eta) + tanh(beta * (1 - eta)));
}
__device__ __forceinline__ double deta_d(double v, double eta, double beta)
{
    const double se = 1.0 / cosh(beta * (v - eta));
    return -beta / sinh(beta) * se * se * sinh(v * beta) * sinh((1 - v) * beta);
}
__device__ __forceinline__ double dprj_d(double v, double eta, double beta)
{
    const double t = tanh(beta * (v - eta));
    return beta * (1 - t * t) / (tanh(beta * eta) + tanh(beta * (1 - eta)));
}

// xTilde = filtered / Hs (top3D125mod.m:108-109)
__global__ void k_div(double *__restrict__ x, const double *__restrict__ d, long n)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e < n) x[e] /= d[e];
}

// partial sums of prj and deta for the eta Newton step (top3D125mod.m:110-116)
__global__ __launch_bounds__(256) void k_eta_sums(const double *__restrict__ xp, long n, double eta, double beta,
                                                 double *part)
{
    double s0 = 0.0, s1 = 0.0;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)kRedBlocks * 256) {
        s0 += prj_d(xp[e], eta, beta);
        s1 += deta_d(xp[e], eta, beta);
    }
    s0 = wsum256(s0);
    s1 = wsum256(s1);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s0;
        part[kRedBlocks + blockIdx.x] = s1;
    }
}

// one block per sum: fixed-order reduction of kRedBlocks partials
__global__ __launch_bounds__(256) void k_finalise(const double *part, int nsums, double *out)
{
    const int q = blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < kRedBlocks; i += 256) s += part[(long)q * kRedBlocks + i];
    s = wsum256(s);
    if (threadIdx.x == 0) out[q] = s;
}

// projection, SIMP and change (top3D125mod.m:117-129): dHs = Hs/dprj,
// xPhys = prj, ch partials, xOld = xPhys, CX and dsK.
__global__ __launch_bounds__(256) void k_physics(double *__restrict__ xPhys, double *__restrict__ xOld,
                                                const double *__restrict__ Hs, double *__restrict__ dHs,
                                                double *__restrict__ CX, double *__restrict__ dsK, long n, int project,
                                                double eta, double beta, double penal, double Emin, double E0,
                                                double *part)
{
    double s = 0.0, vs = 0.0;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)kRedBlocks * 256) {
        double xp = xPhys[e];
        if (project) {
            dHs[e] = Hs[e] / dprj_d(xp, eta, beta);
            xp = prj_d(xp, eta, beta);
            xPhys[e] = xp;
        } else {
            dHs[e] = Hs[e];
        }
        const double d = xp - xOld[e];
        s += d * d;
        vs += xp;
        xOld[e] = xp;
        CX[e] = Emin + pow(xp, penal) * (E0 - Emin);
        dsK[e] = -penal * (E0 - Emin) * pow(xp, penal - 1);
    }
    s = wsum256(s);
    vs = wsum256(vs);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s;
        part[kRedBlocks + blockIdx.x] = vs;
    }
}

// Compliance density ce_e = u_e^T Ke u_e in the C frame and C local order,
// scaled into dc = dsK ce / dHs, and dv = (1/(nel volfrac)) / dHs
// (top3D125mod.m:139-141 before filtering).  The quadratic form is the same
// in the MATLAB frame with cMat because Ke is invariant under the y-mirror.
__global__ __launch_bounds__(256) void k_sens(Grid g, const double *__restrict__ Ke, const double *__restrict__ U,
                                             const double *__restrict__ dsK, const double *__restrict__ dHs,
                                             double dv0, double *__restrict__ dc, double *__restrict__ dv,
                                             double *__restrict__ ce_out)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= g.nel) return;
    const int j = (int)(e % g.ny), k = (int)((e / g.ny) % g.nz), i = (int)(e / ((long)g.ny * g.nz));
    const int ie[8] = {0, 1, 1, 0, 0, 1, 1, 0}, je[8] = {0, 0, 1, 1, 0, 0, 1, 1}, ke[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    double u[24];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const long n = (long)g.nny * g.nnz * (i + ie[a]) + (long)g.nny * (k + ke[a]) + (j + je[a]);
        u[3 * a] = U[3 * n];
        u[3 * a + 1] = U[3 * n + 1];
        u[3 * a + 2] = U[3 * n + 2];
    }
    double ce = 0.0;
#pragma unroll
    for (int r = 0; r < 24; ++r) {
        double t = 0.0;
#pragma unroll
        for (int c = 0; c < 24; ++c) t += Ke[24 * r + c] * u[c];
        ce += u[r] * t;
    }
    if (ce_out) {
        ce_out[e] = ce;
        return;
    }
    dc[e] = dsK[e] * ce / dHs[e];
    dv[e] = dv0 / dHs[e];
}

// OC: ocP = x sqrt(-dc/dv) (top3D125mod.m:145)
__global__ void k_ocp(const double *__restrict__ x, const double *__restrict__ dc, const double *__restrict__ dv,
                      double *__restrict__ ocP, long n, double *part)
{
    double s = 0.0;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)kRedBlocks * 256) {
        const double v = x[e] * sqrt(-dc[e] / dv[e]);
        ocP[e] = v;
        s += v;
    }
    s = wsum256(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__device__ __forceinline__ double oc_x(double ocp, double xt, double lmid, double move)
{
    // max(max(min(min(ocP/lmid, xU), 1), xL), 0) (top3D125mod.m:149)
    double v = ocp / lmid;
    v = fmin(v, xt + move);
    v = fmin(v, 1.0);
    v = fmax(v, xt - move);
    v = fmax(v, 0.0);
    return v;
}

// Volume for kCand candidate multipliers at once (the nodes of the next
// five bisection levels), read x once.
struct Cands {
    double l[kCand];
};

__global__ __launch_bounds__(256) void k_oc_volumes(const double *__restrict__ ocP, const double *__restrict__ x, long n,
                                                   Cands c, int nc, double move, double *part)
{
    double s[kCand];
#pragma unroll
    for (int q = 0; q < kCand; ++q) s[q] = 0.0;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)kRedBlocks * 256) {
        const double o = ocP[e], xt = x[e];
#pragma unroll
        for (int q = 0; q < kCand; ++q)
            if (q < nc) s[q] += oc_x(o, xt, c.l[q], move);
    }
#pragma unroll
    for (int q = 0; q < kCand; ++q) {
        const double t = wsum256(s[q]);
        if (threadIdx.x == 0 && q < nc) part[(long)q * kRedBlocks + blockIdx.x] = t;
    }
}

__global__ void k_oc_apply(const double *__restrict__ ocP, double *__restrict__ x, long n, double lmid, double move)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e < n) x[e] = oc_x(ocP[e], x[e], lmid, move);
}

__global__ void k_edofmat(Grid g, int *cmat)
{
    // cMat (top3D125mod.m:37-41): cVec = 3*nodeNrs+1 (1-based nodes) plus
    // 24 offsets, in MATLAB local order
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= g.nel) return;
    const int j = (int)(e % g.ny), k = (int)((e / g.ny) % g.nz), i = (int)(e / ((long)g.ny * g.nz));
    const long s1 = 3L * g.nny * g.nnz, s2 = 3L * g.nny * (g.nnz + 1), s3 = 3L * g.nny;
    const long off[24] = {0,      1,      2,      s1,     s1 + 1, s1 + 2, s1 - 3, s1 - 2,
                          s1 - 1, -3,     -2,     -1,     s3,     s3 + 1, s3 + 2, s2,
                          s2 + 1, s2 + 2, s2 - 3, s2 - 2, s2 - 1, s3 - 3, s3 - 2, s3 - 1};
    const long nodeNr = (long)g.nny * g.nnz * i + (long)g.nny * k + j + 1;
    const long cvec = 3 * nodeNr + 1;
#pragma unroll
    for (int q = 0; q < 24; ++q) cmat[24 * e + q] = (int)(cvec + off[q]);
}

unsigned nb256(long n) { return (unsigned)((n + 255) / 256); }

std::vector<double> filter_weights(double rmin, int *fw)
{
    // h = max(0, rmin - dist) on offsets -ceil(rmin)+1 .. ceil(rmin)-1
    // (top3D125mod.m:92-95), ordered [dx][dz][dy]
    const int w = (int)std::ceil(rmin) - 1;
    std::vector<double> h;
    for (int dx = -w; dx <= w; ++dx)
        for (int dz = -w; dz <= w; ++dz)
            for (int dy = -w; dy <= w; ++dy) {
                const double v = rmin - std::sqrt((double)(dx * dx + dy * dy + dz * dz));
                h.push_back(v > 0 ? v : 0.0);
            }
    *fw = w;
    return h;
}

double reduce_host(ToState *t, hipStream_t s, int nsums, double *out)
{
    k_finalise<<<nsums, 256, 0, s>>>(t->part, nsums, t->red);
    check_hip(hipMemcpyAsync(t->red_host, t->red, nsums * sizeof(double), hipMemcpyDeviceToHost, s), "copy");
    check_hip(hipStreamSynchronize(s), "sync");
    for (int q = 0; q < nsums; ++q) out[q] = t->red_host[q];
    return out[0];
}

int map_exc()
{
    try {
        throw;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "gputo: %s\n", e.what());
        return GPUTO_ERR_DEVICE;
    }
}

}  // namespace

void gputo::to_release(gputo_handle *h)
{
    if (!h->to) return;
    for (void *p : h->to->allocs) hipFree(p);
    if (h->to->red_host) hipHostFree(h->to->red_host);
    delete h->to;
    h->to = nullptr;
}

extern "C" {

int gputo_edofmat(gputo_handle *h, int *cmat)
{
    if (!h || !cmat) return GPUTO_ERR_ARGUMENT;
    try {
        const Grid &g = h->L[0].g;
        int *d = nullptr;
        check_hip(hipMalloc(&d, 24 * g.nel * sizeof(int)), "hipMalloc");
        k_edofmat<<<nb256(g.nel), 256, 0, h->stream>>>(g, d);
        hipError_t e1 = hipMemcpyAsync(cmat, d, 24 * g.nel * sizeof(int), hipMemcpyDeviceToHost, h->stream);
        hipError_t e2 = hipStreamSynchronize(h->stream);
        hipFree(d);
        check_hip(e1, "copy");
        check_hip(e2, "sync");
        return 0;
    } catch (...) {
        return map_exc();
    }
}

int gputo_filter_f64(gputo_handle *h, const double *in, double rmin, double *out)
{
    if (!h || !in || !out || !(rmin > 0) || rmin > 8) return GPUTO_ERR_ARGUMENT;
    try {
        const Grid &g = h->L[0].g;
        int fw;
        std::vector<double> w = filter_weights(rmin, &fw);
        double *dw, *din, *dout;
        check_hip(hipMalloc(&dw, w.size() * sizeof(double)), "hipMalloc");
        check_hip(hipMalloc(&din, g.nel * sizeof(double)), "hipMalloc");
        check_hip(hipMalloc(&dout, g.nel * sizeof(double)), "hipMalloc");
        hipStream_t s = h->stream;
        hipMemcpyAsync(dw, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice, s);
        hipMemcpyAsync(din, in, g.nel * sizeof(double), hipMemcpyHostToDevice, s);
        k_filter<<<nb256(g.nel), 256, 0, s>>>(g, din, nullptr, dw, fw, dout);
        hipMemcpyAsync(out, dout, g.nel * sizeof(double), hipMemcpyDeviceToHost, s);
        hipError_t e = hipStreamSynchronize(s);
        hipFree(dw);
        hipFree(din);
        hipFree(dout);
        check_hip(e, "filter");
        return 0;
    } catch (...) {
        return map_exc();
    }
}

int gputo_sens_f64(gputo_handle *h, const double *Ke0, const double *U_cframe, double *ce)
{
    if (!h || !Ke0 || !U_cframe || !ce) return GPUTO_ERR_ARGUMENT;
    try {
        const Grid &g = h->L[0].g;
        hipStream_t s = h->stream;
        set_ke(h, Ke0);
        check_hip(hipMemcpyAsync(h->P, U_cframe, g.ndof * sizeof(double), hipMemcpyHostToDevice, s), "copy");
        k_sens<<<nb256(g.nel), 256, 0, s>>>(g, h->Ke, h->P, nullptr, nullptr, 0.0, nullptr, nullptr, h->Q);
        check_hip(hipMemcpyAsync(ce, h->Q, g.nel * sizeof(double), hipMemcpyDeviceToHost, s), "copy");
        check_hip(hipStreamSynchronize(s), "sync");
        return 0;
    } catch (...) {
        return map_exc();
    }
}

int gputo_to_begin(gputo_handle *h, const double *Ke0, const double *x0, const double *F, const double *fixed, int nfd,
                   const gputo_to_config *cfg)
{
    if (!h || !Ke0 || !x0 || !F || !cfg || nfd < 0 || (nfd && !fixed)) return GPUTO_ERR_ARGUMENT;
    if (!(cfg->rmin > 0) || cfg->rmin > 8 || (cfg->ft != 1 && cfg->ft != 3)) return GPUTO_ERR_ARGUMENT;
    to_release(h);
    // the fixed list is validated (finite, 1 <= v <= ndof) before any state
    // is published: a rejected call leaves no design loop on the handle
    std::vector<int> fx((size_t)nfd);
    {
        const double ndof = (double)h->L[0].g.ndof;
        for (int n = 0; n < nfd; ++n) {
            const double v = fixed[n];
            if (!(v >= 1.0 && v <= ndof)) return GPUTO_ERR_ARGUMENT;  // NaN fails too
            fx[n] = (int)v - 1;
        }
    }
    try {
        const Grid &g = h->L[0].g;
        ToState *t = new ToState();
        h->to = t;
        t->cfg = *cfg;
        t->nel = g.nel;
        t->ndof = g.ndof;
        t->penal = cfg->penal;
        t->beta = cfg->beta;
        t->eta = cfg->eta;
        t->Ke.assign(Ke0, Ke0 + 576);
        t->fixed = std::move(fx);
        const long ne = g.nel;
        t->x = talloc<double>(t, ne);
        t->xPhys = talloc<double>(t, ne);
        t->xOld = talloc<double>(t, ne);
        t->Hs = talloc<double>(t, ne);
        t->dHs = talloc<double>(t, ne);
        t->dsK = talloc<double>(t, ne);
        t->dc = talloc<double>(t, ne);
        t->dv = talloc<double>(t, ne);
        t->tmp = talloc<double>(t, ne);
        t->ocP = talloc<double>(t, ne);
        t->CX = talloc<double>(t, ne);
        t->F = talloc<double>(t, g.ndof);
        t->U = talloc<double>(t, g.ndof);
        t->part = talloc<double>(t, (size_t)kCand * kRedBlocks);
        t->red = talloc<double>(t, kCand);
        t->fixed_d = talloc<int>(t, nfd);
        check_hip(hipHostMalloc((void **)&t->red_host, kCand * sizeof(double), hipHostMallocDefault), "hostmalloc");
        std::vector<double> w = filter_weights(cfg->rmin, &t->fw);
        t->w = talloc<double>(t, w.size());
        hipStream_t s = h->stream;
        check_hip(hipMemcpyAsync(t->w, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice, s), "copy");
        check_hip(hipMemcpyAsync(t->x, x0, ne * sizeof(double), hipMemcpyHostToDevice, s), "copy");
        check_hip(hipMemcpyAsync(t->xPhys, x0, ne * sizeof(double), hipMemcpyHostToDevice, s), "copy");
        check_hip(hipMemcpyAsync(t->F, F, g.ndof * sizeof(double), hipMemcpyHostToDevice, s), "copy");
        if (nfd)
            check_hip(hipMemcpyAsync(t->fixed_d, t->fixed.data(), nfd * sizeof(int), hipMemcpyHostToDevice, s), "copy");
        launch_fill(s, t->xOld, ne, 1.0);  // xOld = 1 (top3D125mod.m:102)
        launch_fill(s, t->U, g.ndof, 0.0);
        launch_fill(s, t->tmp, ne, 1.0);
        k_filter<<<nb256(ne), 256, 0, s>>>(g, t->tmp, nullptr, t->w, t->fw, t->Hs);  // Hs = imfilter(ones)
        check_hip(hipGetLastError(), "launch");
        check_hip(hipStreamSynchronize(s), "sync");
        t->loop = 0;
        return 0;
    } catch (...) {
        int rc = map_exc();
        to_release(h);
        return rc;
    }
}

int gputo_to_iterate(gputo_handle *h, gputo_to_stats *st)
{
    if (!h || !h->to || !st) return GPUTO_ERR_ARGUMENT;
    std::memset(st, 0, sizeof(*st));
    ToState *t = h->to;
    const gputo_to_config &c = t->cfg;
    const Grid &g = h->L[0].g;
    const long ne = g.nel;
    hipStream_t s = h->stream;
    try {
        const auto t0 = std::chrono::steady_clock::now();
        t->loop += 1;
        // RL 1: filtered field and projection (top3D125mod.m:108-121)
        k_filter<<<nb256(ne), 256, 0, s>>>(g, t->x, nullptr, t->w, t->fw, t->xPhys);
        k_div<<<nb256(ne), 256, 0, s>>>(t->xPhys, t->Hs, ne);
        double sums[kCand];
        if (c.ft == 3) {
            k_eta_sums<<<kRedBlocks, 256, 0, s>>>(t->xPhys, ne, t->eta, t->beta, t->part);
            reduce_host(t, s, 2, sums);
            double f = sums[0] / ne - c.volfrac;
            int it = 0;
            while (std::fabs(f) > 1e-6 && it < 100) {
                t->eta = t->eta - f / (sums[1] / ne);
                k_eta_sums<<<kRedBlocks, 256, 0, s>>>(t->xPhys, ne, t->eta, t->beta, t->part);
                reduce_host(t, s, 2, sums);
                f = sums[0] / ne - c.volfrac;
                ++it;
            }
        }
        k_physics<<<kRedBlocks, 256, 0, s>>>(t->xPhys, t->xOld, t->Hs, t->dHs, t->CX, t->dsK, ne, c.ft == 3 ? 1 : 0,
                                             t->eta, t->beta, t->penal, c.Emin, c.E0, t->part);
        reduce_host(t, s, 2, sums);
        st->change = std::sqrt(sums[0]) / ne;
        st->volume = sums[1] / ne;
        // RL 2: solve (top3D125mod.m:130-134)
        hipEvent_t e0, e1;
        hipEventCreate(&e0);
        hipEventCreate(&e1);
        hipEventRecord(e0, s);
        double relres = 0.0;
        int nit = 0;
        set_ke(h, t->Ke.data());
        check_hip(hipMemcpyAsync(h->CX, t->CX, ne * sizeof(double), hipMemcpyDeviceToDevice, s), "copy");
        int rc = setup_hierarchy(h, t->fixed);
        if (rc == 0) {
            rc = run_pcg(h, t->F, t->U, c.cg_tol, c.cg_maxit, &relres, &nit);
        }
        hipEventRecord(e1, s);
        hipEventSynchronize(e1);
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        st->seconds_solve = ms * 1e-3;
        st->relres = relres;
        st->cg_iters = nit;
        if (rc) {
            st->status = rc;
            return rc;
        }
        // RL 3: sensitivities (top3D125mod.m:139-141)
        k_sens<<<nb256(ne), 256, 0, s>>>(g, h->Ke, t->U, t->dsK, t->dHs, 1.0 / ne / c.volfrac, t->tmp, t->ocP,
                                          nullptr);
        k_filter<<<nb256(ne), 256, 0, s>>>(g, t->tmp, nullptr, t->w, t->fw, t->dc);
        k_filter<<<nb256(ne), 256, 0, s>>>(g, t->ocP, nullptr, t->w, t->fw, t->dv);
        // RL 4: OC bisection (top3D125mod.m:143-151)
        k_ocp<<<kRedBlocks, 256, 0, s>>>(t->x, t->dc, t->dv, t->ocP, ne, t->part);
        reduce_host(t, s, 1, sums);
        double l1 = 0.0, l2 = sums[0] / ne / c.volfrac, lmid = 0.0;
        while ((l2 - l1) / (l2 + l1) > 1e-4) {
            // candidates: the midpoints of the next 5 bisection levels, in
            // heap order (node q has children 2q+1 = "volume > volfrac")
            Cands cd;
            double lo[kCand], hi[kCand];
            lo[0] = l1;
            hi[0] = l2;
            for (int q = 0; q < kCand; ++q) {
                cd.l[q] = 0.5 * (lo[q] + hi[q]);
                if (2 * q + 2 < kCand) {
                    lo[2 * q + 1] = cd.l[q];  // volume > volfrac: l1 = lmid
                    hi[2 * q + 1] = hi[q];
                    lo[2 * q + 2] = lo[q];    // else: l2 = lmid
                    hi[2 * q + 2] = cd.l[q];
                }
            }
            k_oc_volumes<<<kRedBlocks, 256, 0, s>>>(t->ocP, t->x, ne, cd, kCand, c.move, t->part);
            reduce_host(t, s, kCand, sums);
            int q = 0;
            for (;;) {
                lmid = cd.l[q];
                if (sums[q] / ne > c.volfrac)
                    l1 = lmid;
                else
                    l2 = lmid;
                const int nxt = (sums[q] / ne > c.volfrac) ? 2 * q + 1 : 2 * q + 2;
                if (!((l2 - l1) / (l2 + l1) > 1e-4) || nxt >= kCand) break;
                q = nxt;
            }
        }
        k_oc_apply<<<nb256(ne), 256, 0, s>>>(t->ocP, t->x, ne, lmid, c.move);
        // continuation (top3D125mod.m:32-33,152): cnt(v,{1,1,25,d},loop)
        const int L = t->loop;
        if (L >= 1 && t->penal < 1 && L % 25 == 0) t->penal += 0.25;
        if (L >= 1 && t->beta < 1 && L % 25 == 0) t->beta += 2;
        // RL 5: compliance F'U (top3D125mod.m:155)
        launch_dot_partial(s, t->F, t->U, g.ndof, t->part, nullptr);
        reduce_host(t, s, 1, sums);
        check_hip(hipGetLastError(), "launch");
        st->compliance = sums[0];
        st->eta = t->eta;
        st->lmid = lmid;
        st->status = 0;
        st->seconds_total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (h->mx) mixed_stats(h, st);
        return 0;
    } catch (...) {
        st->status = map_exc();
        return st->status;
    }
}

int gputo_to_get(gputo_handle *h, double *x, double *xPhys, double *U)
{
    if (!h || !h->to) return GPUTO_ERR_ARGUMENT;
    try {
        ToState *t = h->to;
        hipStream_t s = h->stream;
        if (x) check_hip(hipMemcpyAsync(x, t->x, t->nel * sizeof(double), hipMemcpyDeviceToHost, s), "copy");
        if (xPhys) check_hip(hipMemcpyAsync(xPhys, t->xPhys, t->nel * sizeof(double), hipMemcpyDeviceToHost, s), "copy");
        if (U) check_hip(hipMemcpyAsync(U, t->U, t->ndof * sizeof(double), hipMemcpyDeviceToHost, s), "copy");
        check_hip(hipStreamSynchronize(s), "sync");
        return 0;
    } catch (...) {
        return map_exc();
    }
}

int gputo_to_end(gputo_handle *h)
{
    if (!h) return GPUTO_ERR_ARGUMENT;
    if (h->stream) hipStreamSynchronize(h->stream);
    to_release(h);
    return 0;
}

}  // extern "C"
