// Learned level-2 correction of the FP32 V-cycle.
//
// After the level-2 post-smoothing of each V-cycle the stage computes the
// level-2 defect d = F - A U, maps it (Jacobi-scaled, with a stiffness
// channel) through a 3-D convolutional encoder-decoder to a correction e,
// and adds gamma e with gamma = <d,e>/<e,Ae>, the step that minimises the
// level-2 energy 1/2 e^T A e - e^T d along e.  A step that would not lower
// that energy (gamma <= 0 or not finite) is dropped: the plain V-cycle is the
// fallback, decided on the device.  The network is trained on the defects
// the solve itself logs, with the same energy as the loss,
//   L(theta) = (1/2 e^T A e - e^T d) / sigma^2,   e = e(d; theta),
// whose minimiser is e = A^{-1} d; its gradient needs only level-2
// operator products, so no reference errors have to be solved for.
//
// Every kernel here is this library's own: direct 3x3x3 convolutions
// (forward, input gradient, weight gradient), nearest upsampling with skip
// additions, Adam.  FP32 throughout with deterministic reductions.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <stdexcept>

#include "learn.h"
#include "mixed.h"

namespace gputo {

namespace {

// position p = (i*n1 + k)*n2 + j of a grid (all sizes here are < 2^31:
// 32-bit divisions)
__device__ __forceinline__ void pos3(int p, const Dims3 &d, int &i, int &k, int &j)
{
    const int r = p / d.n2;
    j = p - r * d.n2;
    i = r / d.n1;
    k = r - i * d.n1;
}

// Copy n floats, element q from src[idx(q)], into LDS with 8 loads in flight
// per thread (indices clamped to 0 past the end, those values dropped).
template <class Idx>
__device__ __forceinline__ void stage_lds(float *__restrict__ dst, int n, const float *__restrict__ src, Idx idx)
{
    for (int q0 = 0; q0 < n; q0 += 8 * 256) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int q = q0 + u * 256 + (int)threadIdx.x;
            v[u] = src[q < n ? idx(q) : 0];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int q = q0 + u * 256 + (int)threadIdx.x;
            if (q < n) dst[q] = v[u];
        }
    }
}

// out[co][p] = act(b[co] + sum_ci sum_t w[co][ci][t] in[ci][p*s + t - 1]).
// A block owns 256/KS output positions and CT output channels; its
// CT x cin x 27 weights are staged in LDS as [ci][t][c] (one vector read
// per tap feeds CT FMAs) and the 27 tap offsets of a position are computed
// once.  KS > 1 splits the input channels over KS threads per position
// (small grids have too few positions to fill the device); the KS partial
// sums are added in a fixed order.
template <int CT, int KS, bool RELU>
__global__ __launch_bounds__(256) void k_conv_fwd(const float *__restrict__ in, int cin, Dims3 din,
                                                 const float *__restrict__ w, const float *__restrict__ b,
                                                 float *__restrict__ out, Dims3 dout, int stride, const int *ctl)
{
    if (ctl && ctl[0]) return;
    extern __shared__ float wl[];  // [cin*27][CT], then KS partials [KS][256/KS][CT]
    constexpr int NP = 256 / KS;
    const int co0 = blockIdx.y * CT;
    stage_lds(wl, cin * 27 * CT, w, [=](int q) {
        const int c = q % CT, k = q / CT;  // k = ci*27 + t
        return (co0 + c) * cin * 27 + k;
    });
    __syncthreads();
    const int pl = threadIdx.x % NP, ks = threadIdx.x / NP;
    const int p = blockIdx.x * NP + pl;
    const int nout = (int)dout.size(), nin = (int)din.size();
    float acc[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[c] = 0.0f;
    if (p < nout) {
        int oi, ok, oj;
        pos3(p, dout, oi, ok, oj);
        int off[27];
        unsigned valid = 0;
#pragma unroll
        for (int t = 0; t < 27; ++t) {
            const int ii = oi * stride + t / 9 - 1, ik = ok * stride + (t / 3) % 3 - 1, ij = oj * stride + t % 3 - 1;
            const bool v = ii >= 0 && ii < din.n0 && ik >= 0 && ik < din.n1 && ij >= 0 && ij < din.n2;
            valid |= (unsigned)v << t;
            off[t] = v ? (ii * din.n1 + ik) * din.n2 + ij : 0;
        }
        if (KS == 1) {
#pragma unroll
            for (int c = 0; c < CT; ++c) acc[c] = b[co0 + c];
        }
        for (int ci = ks; ci < cin; ci += KS) {
            const float *src = in + ci * nin;
            const float *wc = wl + ci * 27 * CT;
            // unconditional loads (off = 0 outside the grid), masked after:
            // all 27 loads of a channel are in flight together
            float v[27];
#pragma unroll
            for (int t = 0; t < 27; ++t) v[t] = src[off[t]];
#pragma unroll
            for (int t = 0; t < 27; ++t) {
                const float vt = ((valid >> t) & 1u) ? v[t] : 0.0f;
#pragma unroll
                for (int c = 0; c < CT; ++c) acc[c] = fmaf(wc[t * CT + c], vt, acc[c]);
            }
        }
    }
    if (KS > 1) {
        float *part = wl + cin * 27 * CT;
        __syncthreads();
#pragma unroll
        for (int c = 0; c < CT; ++c) part[(ks * NP + pl) * CT + c] = acc[c];
        __syncthreads();
        if (ks != 0 || p >= nout) return;
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            float sm = b[co0 + c];
            for (int k = 0; k < KS; ++k) sm += part[(k * NP + pl) * CT + c];
            acc[c] = sm;
        }
    } else if (p >= nout) {
        return;
    }
#pragma unroll
    for (int c = 0; c < CT; ++c) out[(long)(co0 + c) * nout + p] = RELU ? fmaxf(acc[c], 0.0f) : acc[c];
}

// gin[ci][q] = sum_co sum_t w[co][ci][t] gz[co][p] over the outputs p that
// read q through tap t (q = p*s + t - 1), gz = gout * relu'(act).  The
// block's CT input channels of every output channel's weights are staged
// in LDS as [co][t][c].
template <int CT>
__global__ __launch_bounds__(256) void k_conv_bwd_data(const float *__restrict__ gout, const float *__restrict__ act,
                                                      int cout, Dims3 dout, const float *__restrict__ w, int cin,
                                                      Dims3 din, int stride, float *__restrict__ gin)
{
    extern __shared__ float wl[];  // [cout*27][CT]
    const int ci0 = blockIdx.y * CT;
    stage_lds(wl, cout * 27 * CT, w, [=](int q) {
        const int c = q % CT, k = q / CT;  // k = co*27 + t
        const int co = k / 27, t = k - 27 * co;
        return (co * cin + ci0 + c) * 27 + t;
    });
    __syncthreads();
    const int q = blockIdx.x * 256 + threadIdx.x;
    const int nin = (int)din.size(), nout = (int)dout.size();
    if (q >= nin) return;
    int qi, qk, qj;
    pos3(q, din, qi, qk, qj);
    int off[27];
    unsigned valid = 0;
#pragma unroll
    for (int t = 0; t < 27; ++t) {
        const int xi = qi + 1 - t / 9, xk = qk + 1 - (t / 3) % 3, xj = qj + 1 - t % 3;
        int pi = xi, pk = xk, pj = xj;
        bool ok = xi >= 0 && xk >= 0 && xj >= 0;
        if (stride == 2) {
            ok = ok && !(xi & 1) && !(xk & 1) && !(xj & 1);
            pi = xi >> 1;
            pk = xk >> 1;
            pj = xj >> 1;
        }
        ok = ok && pi < dout.n0 && pk < dout.n1 && pj < dout.n2;
        valid |= (unsigned)ok << t;
        off[t] = ok ? (pi * dout.n1 + pk) * dout.n2 + pj : 0;
    }
    float acc[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[c] = 0.0f;
    for (int co = 0; co < cout; ++co) {
        const float *g = gout + co * nout;
        const float *a = act ? act + co * nout : nullptr;
        const float *wc = wl + co * 27 * CT;
        float gv[27], av[27];
#pragma unroll
        for (int t = 0; t < 27; ++t) {
            gv[t] = g[off[t]];  // off = 0 outside the grid: masked below
            av[t] = a ? a[off[t]] : 1.0f;
        }
#pragma unroll
        for (int t = 0; t < 27; ++t) {
            const float gz = (((valid >> t) & 1u) && av[t] > 0.0f) ? gv[t] : 0.0f;
#pragma unroll
            for (int c = 0; c < CT; ++c) acc[c] = fmaf(wc[t * CT + c], gz, acc[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < CT; ++c) gin[(ci0 + c) * nin + q] = acc[c];
}

// Weight gradient, split over chunks of kWChunk output positions:
// part[ch][co*cin + ci][t] = sum_{p in chunk ch} gz[co][p] in[ci][p*s + t - 1]
// and, at t = 27, sum_p gz[co][p].  One block per (co, ci, chunk), fixed-
// order block reduction; k_conv_bwd_w_sum adds the chunks in order.
constexpr int kWChunk = 2048;

__global__ __launch_bounds__(256) void k_conv_bwd_w(const float *__restrict__ gout, const float *__restrict__ act,
                                                   Dims3 dout, const float *__restrict__ in, int cin, Dims3 din,
                                                   int stride, float *__restrict__ part)
{
    __shared__ float red[28][257];
    const int pair = blockIdx.x, co = pair / cin, ci = pair % cin;
    const int nout = (int)dout.size(), nin = (int)din.size();
    const float *g = gout + co * nout;
    const float *a = act ? act + co * nout : nullptr;
    const float *src = in + ci * nin;
    float acc[28];
#pragma unroll
    for (int t = 0; t < 28; ++t) acc[t] = 0.0f;
    const int p1 = min(nout, (int)(blockIdx.y + 1) * kWChunk);
    for (int p = blockIdx.y * kWChunk + threadIdx.x; p < p1; p += 256) {
        float gz = g[p];
        if (a && !(a[p] > 0.0f)) gz = 0.0f;
        if (gz == 0.0f) continue;
        int oi, ok, oj;
        pos3(p, dout, oi, ok, oj);
        acc[27] += gz;
        float v[27];
#pragma unroll
        for (int t = 0; t < 27; ++t) {
            const int ii = oi * stride + t / 9 - 1, ik = ok * stride + (t / 3) % 3 - 1, ij = oj * stride + t % 3 - 1;
            const bool in_grid = ii >= 0 && ii < din.n0 && ik >= 0 && ik < din.n1 && ij >= 0 && ij < din.n2;
            v[t] = src[in_grid ? (ii * din.n1 + ik) * din.n2 + ij : 0];
            v[t] = in_grid ? v[t] : 0.0f;
        }
#pragma unroll
        for (int t = 0; t < 27; ++t) acc[t] = fmaf(gz, v[t], acc[t]);
    }
#pragma unroll
    for (int t = 0; t < 28; ++t) red[t][threadIdx.x] = acc[t];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int t = 0; t < 28; ++t) red[t][threadIdx.x] += red[t][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 28) part[((long)blockIdx.y * gridDim.x + pair) * 28 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_conv_bwd_w_sum(const float *__restrict__ part, int nch, int npair, int cin, float *__restrict__ gw,
                                 float *__restrict__ gb)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= npair * 28) return;
    const int pair = q / 28, t = q - 28 * pair;
    float sm = 0.0f;
    for (int ch = 0; ch < nch; ++ch) sm += part[((long)ch * npair + pair) * 28 + t];
    if (t < 27)
        gw[(long)pair * 27 + t] += sm;
    else if (pair % cin == 0)
        gb[pair / cin] += sm;
}

// out[c][p] = d[c][parent(p)] + e[c][p]  (nearest upsampling by 2 plus skip)
__global__ void k_up_add(const float *__restrict__ dcoarse, Dims3 dc, const float *__restrict__ efine, Dims3 df,
                         int C, float *__restrict__ out, const int *ctl)
{
    if (ctl && ctl[0]) return;
    const int n = (int)df.size(), total = n * C, ncs = (int)dc.size();
    for (int q = blockIdx.x * 256 + threadIdx.x; q < total; q += gridDim.x * 256) {
        const int c = q / n;
        const int p = q - c * n;
        int i, k, j;
        pos3(p, df, i, k, j);
        out[q] = dcoarse[c * ncs + ((i >> 1) * dc.n1 + (k >> 1)) * dc.n2 + (j >> 1)] + efine[q];
    }
}

// gd[c][pc] = sum over the (up to 8) fine children of pc of gu[c][p]
__global__ void k_up_bwd(const float *__restrict__ gu, Dims3 df, int C, Dims3 dc, float *__restrict__ gd)
{
    const int nc = (int)dc.size(), total = nc * C, nf = (int)df.size();
    for (int q = blockIdx.x * 256 + threadIdx.x; q < total; q += gridDim.x * 256) {
        const int c = q / nc;
        const int pc = q - c * nc;
        int i, k, j;
        pos3(pc, dc, i, k, j);
        float s = 0.0f;
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int e = 0; e < 2; ++e) {
                    const int fi = 2 * i + a, fk = 2 * k + b, fj = 2 * j + e;
                    if (fi < df.n0 && fk < df.n1 && fj < df.n2) s += gu[c * nf + (fi * df.n1 + fk) * df.n2 + fj];
                }
        gd[q] = s;
    }
}

__global__ void k_add_inplace(float *__restrict__ dst, const float *__restrict__ src, long n)
{
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long)gridDim.x * 256) dst[q] += src[q];
}

__global__ void k_zero(float *__restrict__ v, long n)
{
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long)gridDim.x * 256) v[q] = 0.0f;
}

// ---- the level-2 stage: features, output scaling, loss gradient, step ----
// sc[0] = sigma (RMS of the Jacobi-scaled defect), sc[1] = gamma,
// sc[2] = <d,e>, sc[3] = <e,Ae>, sc[4] = loss
__device__ double block_total(const double *p, int n)
{
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += kBlock) s += p[i];
    return block_sum(s);
}

// partials of sum_d d^2 / ad
__global__ __launch_bounds__(kBlock) void k_scaled_norm(const float *__restrict__ d, const float *__restrict__ ad,
                                                       long n, double *partials, const int *ctl)
{
    if (ctl && ctl[0]) return;
    double s = 0.0;
    for (long q = (long)blockIdx.x * kBlock + threadIdx.x; q < n; q += (long)kRedBlocks * kBlock)
        s += (double)d[q] * (double)d[q] / (double)ad[q];
    s = block_sum(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_set_sigma(double *sc, const double *p, long n, const int *ctl)
{
    if (ctl && ctl[0]) return;
    const double s = block_total(p, kRedBlocks);
    if (threadIdx.x == 0) {
        const double sig = sqrt(s / (double)n);
        sc[0] = sig > 1e-300 ? sig : 1.0;
    }
}

// x[c][p] = d[3p+c] / (sqrt(ad[3p+c]) sigma), x[3][p] = log10(mean_c ad)/9
__global__ void k_features(const float *__restrict__ d, const float *__restrict__ ad, long nnod,
                           const double *sc, float *__restrict__ x, const int *ctl)
{
    if (ctl && ctl[0]) return;
    const float inv = (float)(1.0 / sc[0]);
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < nnod; p += (long)gridDim.x * 256) {
        float m = 0.0f;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float a = ad[3 * p + c];
            x[c * nnod + p] = d[3 * p + c] * inv / sqrtf(a);
            m += a;
        }
        x[3 * nnod + p] = log10f(fmaxf(m / 3.0f, 1e-30f)) / 9.0f;
    }
}

// e[3p+c] = fixed ? 0 : sigma y[c][p] / sqrt(ad[3p+c])
__global__ void k_correction(const float *__restrict__ y, const float *__restrict__ ad,
                             const uint8_t *__restrict__ mask, long nnod, const double *sc, float *__restrict__ e,
                             const int *ctl)
{
    if (ctl && ctl[0]) return;
    const float sig = (float)sc[0];
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < nnod; p += (long)gridDim.x * 256) {
        const unsigned m = mask[p];
#pragma unroll
        for (int c = 0; c < 3; ++c)
            e[3 * p + c] = ((m >> c) & 1u) ? 0.0f : sig * y[c * nnod + p] / sqrtf(ad[3 * p + c]);
    }
}

// gy[c][p] += w (Ae - d)[3p+c] / (sqrt(ad) sigma) at free DOFs
__global__ void k_loss_grad(const float *__restrict__ Ae, const float *__restrict__ d, const float *__restrict__ ad,
                            const uint8_t *__restrict__ mask, long nnod, const double *sc, float w,
                            float *__restrict__ gy)
{
    const float inv = (float)(1.0 / sc[0]);
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < nnod; p += (long)gridDim.x * 256) {
        const unsigned m = mask[p];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const long q = 3 * p + c;
            gy[c * nnod + p] = ((m >> c) & 1u) ? 0.0f : w * (Ae[q] - d[q]) * inv / sqrtf(ad[q]);
        }
    }
}

// gamma = <d,e>/<e,Ae> if positive and finite, else 0 (fallback); counters
__global__ __launch_bounds__(kBlock) void k_gamma(double *sc, const double *p_de, const double *p_eAe, int *count,
                                                 const int *ctl)
{
    if (ctl && ctl[0]) return;
    const double de = block_total(p_de, kRedBlocks);
    const double eAe = block_total(p_eAe, kRedBlocks);
    if (threadIdx.x == 0) {
        double g = de / eAe;
        const bool ok = de > 0.0 && eAe > 0.0 && g > 0.0 && g < INFINITY;
        sc[1] = ok ? g : 0.0;
        sc[2] = de;
        sc[3] = eAe;
        count[ok ? 0 : 1] += 1;
    }
}

__global__ void k_axpy_gamma(float *__restrict__ U, const float *__restrict__ e, long n, const double *sc,
                             const int *ctl)
{
    if (ctl && ctl[0]) return;
    const float g = (float)sc[1];
    if (g == 0.0f) return;
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long)gridDim.x * 256) U[q] += g * e[q];
}

// energy loss of one sample: (1/2 <e,Ae> - <e,d>) / sigma^2 accumulated in sc[4]
__global__ __launch_bounds__(kBlock) void k_loss_acc(double *sc, const double *p_eAe, const double *p_ed)
{
    const double eAe = block_total(p_eAe, kRedBlocks);
    const double ed = block_total(p_ed, kRedBlocks);
    if (threadIdx.x == 0) sc[4] += (0.5 * eAe - ed) / (sc[0] * sc[0]);
}

__global__ void k_log_sample(const float *__restrict__ d, float *__restrict__ samples, long n, int *count,
                             int nmax, const int *ctl)
{
    if (ctl && ctl[0]) return;
    const int s = count[2];
    if (s >= nmax) return;
    float *dst = samples + (long)s * n;
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long)gridDim.x * 256) dst[q] = d[q];
}

__global__ void k_log_count(int *count, int nmax, const int *ctl)
{
    if (ctl && ctl[0]) return;
    if (threadIdx.x == 0 && blockIdx.x == 0 && count[2] < nmax) count[2] += 1;
}

__global__ void k_adam(float *__restrict__ P, const float *__restrict__ G, float *__restrict__ M,
                       float *__restrict__ V, long n, float lr, float b1, float b2, float c1, float c2, float scale)
{
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long)gridDim.x * 256) {
        const float g = G[q] * scale;
        const float m = b1 * M[q] + (1.0f - b1) * g;
        const float v = b2 * V[q] + (1.0f - b2) * g * g;
        M[q] = m;
        V[q] = v;
        P[q] -= lr * (m / c1) / (sqrtf(v / c2) + 1e-8f);
    }
}

unsigned grid256(long n) { return (unsigned)std::min<long>(std::max<long>((n + 255) / 256, 1), 8192); }

template <class T>
T *ldalloc(LearnState *L, size_t n)
{
    void *p = nullptr;
    check_hip(hipMalloc(&p, (n ? n : 1) * sizeof(T)), "hipMalloc");
    L->allocs.push_back(p);
    L->bytes += (long long)(n * sizeof(T));
    return static_cast<T *>(p);
}

Dims3 half(const Dims3 &d) { return {(d.n0 - 1) / 2 + 1, (d.n1 - 1) / 2 + 1, (d.n2 - 1) / 2 + 1}; }

template <int CT, int KS>
void conv_fwd_t(hipStream_t s, const NetLayer &ly, const float *w, const float *b, const float *in, Dims3 din,
                float *out, Dims3 dout, const int *ctl)
{
    const dim3 grid((unsigned)((dout.size() + 256 / KS - 1) / (256 / KS)), ly.cout / CT);
    const size_t lds = ((size_t)ly.cin * 27 * CT + (KS > 1 ? 256 * CT : 0)) * sizeof(float);
    if (ly.relu)
        k_conv_fwd<CT, KS, true><<<grid, 256, lds, s>>>(in, ly.cin, din, w, b, out, dout, ly.stride, ctl);
    else
        k_conv_fwd<CT, KS, false><<<grid, 256, lds, s>>>(in, ly.cin, din, w, b, out, dout, ly.stride, ctl);
}

// Input-channel split: the smallest KS (a power of two, at most cin) that
// gives at least kConvBlocks blocks, so the small grids fill the device.
constexpr long kConvBlocks = 1024;

template <int CT>
void conv_fwd_ct(hipStream_t s, const NetLayer &ly, const float *w, const float *b, const float *in, Dims3 din,
                 float *out, Dims3 dout, const int *ctl)
{
    int ks = 1;
    while (ks < 32 && 2 * ks <= ly.cin &&
           ((dout.size() + 256 / ks - 1) / (256 / ks)) * (ly.cout / CT) < kConvBlocks)
        ks *= 2;
    switch (ks) {
    case 1: conv_fwd_t<CT, 1>(s, ly, w, b, in, din, out, dout, ctl); break;
    case 2: conv_fwd_t<CT, 2>(s, ly, w, b, in, din, out, dout, ctl); break;
    case 4: conv_fwd_t<CT, 4>(s, ly, w, b, in, din, out, dout, ctl); break;
    case 8: conv_fwd_t<CT, 8>(s, ly, w, b, in, din, out, dout, ctl); break;
    case 16: conv_fwd_t<CT, 16>(s, ly, w, b, in, din, out, dout, ctl); break;
    default: conv_fwd_t<CT, 32>(s, ly, w, b, in, din, out, dout, ctl); break;
    }
}

void conv_fwd(hipStream_t s, const NetLayer &ly, const float *P, const float *in, Dims3 din, float *out, Dims3 dout,
              const int *ctl)
{
    const float *w = P + ly.w_off, *b = P + ly.b_off;
    if (ly.cout % 4 == 0)
        conv_fwd_ct<4>(s, ly, w, b, in, din, out, dout, ctl);
    else if (ly.cout == 3)
        conv_fwd_ct<3>(s, ly, w, b, in, din, out, dout, ctl);
    else
        conv_fwd_ct<1>(s, ly, w, b, in, din, out, dout, ctl);
}

// gradients of one layer: weights/bias accumulate into G; gin (if non-null)
// is overwritten with the input gradient
void conv_bwd(hipStream_t s, const NetLayer &ly, const float *P, float *G, float *wpart, const float *in, Dims3 din,
              const float *out, Dims3 dout, const float *gout, float *gin)
{
    const float *act = ly.relu ? out : nullptr;
    const int npair = ly.cout * ly.cin, nch = (int)((dout.size() + kWChunk - 1) / kWChunk);
    k_conv_bwd_w<<<dim3(npair, nch), 256, 0, s>>>(gout, act, dout, in, ly.cin, din, ly.stride, wpart);
    k_conv_bwd_w_sum<<<(npair * 28 + 255) / 256, 256, 0, s>>>(wpart, nch, npair, ly.cin, G + ly.w_off, G + ly.b_off);
    if (gin) {
        const unsigned gx = (unsigned)((din.size() + 255) / 256);
        if (ly.cin % 4 == 0)
            k_conv_bwd_data<4><<<dim3(gx, ly.cin / 4), 256, (size_t)ly.cout * 27 * 4 * sizeof(float), s>>>(
                gout, act, ly.cout, dout, P + ly.w_off, ly.cin, din, ly.stride, gin);
        else
            k_conv_bwd_data<1><<<dim3(gx, ly.cin), 256, (size_t)ly.cout * 27 * sizeof(float), s>>>(
                gout, act, ly.cout, dout, P + ly.w_off, ly.cin, din, ly.stride, gin);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
void learn_create(gputo_handle *h)
{
    if (h->ln) return;
    if (h->nl < 4) throw std::runtime_error("learned correction needs at least 4 levels");
    LearnState *L = new LearnState();
    h->ln = L;
    const Grid &g2 = h->L[2].g;
    L->d0 = {g2.nnx, g2.nnz, g2.nny};
    L->d1 = half(L->d0);
    L->d2 = half(L->d1);
    const int spec[kNetLayers][4] = {{4, 16, 1, 1}, {16, 32, 2, 1}, {32, 64, 2, 1},
                                     {64, 32, 1, 1}, {32, 16, 1, 1}, {16, 3, 1, 0}};
    long off = 0;
    for (int i = 0; i < kNetLayers; ++i) {
        NetLayer &ly = L->L[i];
        ly.cin = spec[i][0];
        ly.cout = spec[i][1];
        ly.stride = spec[i][2];
        ly.relu = spec[i][3];
        ly.w_off = off;
        off += (long)ly.cout * ly.cin * 27;
        ly.b_off = off;
        off += ly.cout;
    }
    L->nparam = off;
    L->P = ldalloc<float>(L, off);
    L->G = ldalloc<float>(L, off);
    L->M = ldalloc<float>(L, off);
    L->V = ldalloc<float>(L, off);
    const long n0 = L->d0.size(), n1 = L->d1.size(), n2 = L->d2.size();
    L->x0 = ldalloc<float>(L, 4 * n0);
    L->e1 = ldalloc<float>(L, 16 * n0);
    L->e2 = ldalloc<float>(L, 32 * n1);
    L->e3 = ldalloc<float>(L, 64 * n2);
    L->a2 = ldalloc<float>(L, 32 * n2);
    L->u2 = ldalloc<float>(L, 32 * n1);
    L->a1 = ldalloc<float>(L, 16 * n1);
    L->u1 = ldalloc<float>(L, 16 * n0);
    L->y = ldalloc<float>(L, 3 * n0);
    L->gy = ldalloc<float>(L, 3 * n0);
    L->gu1 = ldalloc<float>(L, 16 * n0);
    L->ga1 = ldalloc<float>(L, 16 * n1);
    L->gu2 = ldalloc<float>(L, 32 * n1);
    L->ga2 = ldalloc<float>(L, 32 * n2);
    L->ge3 = ldalloc<float>(L, 64 * n2);
    L->ge2 = ldalloc<float>(L, 32 * n1);
    L->ge1 = ldalloc<float>(L, 16 * n0);
    const long nd = g2.ndof;
    L->d = ldalloc<float>(L, nd);
    L->e = ldalloc<float>(L, nd);
    L->Ae = ldalloc<float>(L, nd);
    L->sc = ldalloc<double>(L, 8);
    L->count = ldalloc<int>(L, 4);
    L->partials = ldalloc<double>(L, 2 * kRedBlocks);
    {
        // weight-gradient partials: the largest (cout*cin) x chunks x 28 over the layers
        const Dims3 dims[kNetLayers] = {L->d0, L->d1, L->d2, L->d2, L->d1, L->d0};
        long mx = 1;
        for (int i = 0; i < kNetLayers; ++i)
            mx = std::max(mx, (long)L->L[i].cout * L->L[i].cin * 28 * ((dims[i].size() + kWChunk - 1) / kWChunk));
        L->wpart = ldalloc<float>(L, mx);
    }
    L->nsamp_max = 64;
    L->samples = ldalloc<float>(L, (size_t)L->nsamp_max * nd);
    check_hip(hipMemset(L->count, 0, 4 * sizeof(int)), "memset");
    check_hip(hipMemset(L->sc, 0, 8 * sizeof(double)), "memset");
    learn_init_params(h, 1234u);
    h->bytes += L->bytes;
}

void learn_destroy(gputo_handle *h)
{
    LearnState *L = h->ln;
    if (!L) return;
    for (void *p : L->allocs) (void)hipFree(p);
    delete L;
    h->ln = nullptr;
}

// He-uniform weights from a fixed seed (ReLU layers), small output layer,
// zero biases; Adam state cleared.
void learn_init_params(gputo_handle *h, unsigned seed)
{
    LearnState *L = h->ln;
    std::vector<float> p(L->nparam, 0.0f);
    std::mt19937 rng(seed);
    for (int i = 0; i < kNetLayers; ++i) {
        const NetLayer &ly = L->L[i];
        const float a = std::sqrt(6.0f / (ly.cin * 27.0f)) * (ly.relu ? 1.0f : 0.1f);
        std::uniform_real_distribution<float> U(-a, a);
        for (long q = 0; q < (long)ly.cout * ly.cin * 27; ++q) p[ly.w_off + q] = U(rng);
    }
    check_hip(hipMemcpy(L->P, p.data(), p.size() * sizeof(float), hipMemcpyHostToDevice), "copy");
    check_hip(hipMemset(L->M, 0, L->nparam * sizeof(float)), "memset");
    check_hip(hipMemset(L->V, 0, L->nparam * sizeof(float)), "memset");
    L->adam_t = 0;
    L->trained = false;
}

// y = N(x0) with the forward activations cached
void learn_forward(gputo_handle *h, const int *ctl)
{
    LearnState *L = h->ln;
    hipStream_t s = h->stream;
    const Dims3 d0 = L->d0, d1 = L->d1, d2 = L->d2;
    conv_fwd(s, L->L[0], L->P, L->x0, d0, L->e1, d0, ctl);
    conv_fwd(s, L->L[1], L->P, L->e1, d0, L->e2, d1, ctl);
    conv_fwd(s, L->L[2], L->P, L->e2, d1, L->e3, d2, ctl);
    conv_fwd(s, L->L[3], L->P, L->e3, d2, L->a2, d2, ctl);
    k_up_add<<<grid256(32 * d1.size()), 256, 0, s>>>(L->a2, d2, L->e2, d1, 32, L->u2, ctl);
    conv_fwd(s, L->L[4], L->P, L->u2, d1, L->a1, d1, ctl);
    k_up_add<<<grid256(16 * d0.size()), 256, 0, s>>>(L->a1, d1, L->e1, d0, 16, L->u1, ctl);
    conv_fwd(s, L->L[5], L->P, L->u1, d0, L->y, d0, ctl);
}

// parameter gradients of <gy, y> accumulated into G (forward cache current)
void learn_backward(gputo_handle *h)
{
    LearnState *L = h->ln;
    hipStream_t s = h->stream;
    const Dims3 d0 = L->d0, d1 = L->d1, d2 = L->d2;
    const float *P = L->P;
    float *G = L->G;
    conv_bwd(s, L->L[5], P, G, L->wpart, L->u1, d0, L->y, d0, L->gy, L->gu1);
    // u1 = up(a1) + e1
    k_up_bwd<<<grid256(16 * d1.size()), 256, 0, s>>>(L->gu1, d0, 16, d1, L->ga1);
    conv_bwd(s, L->L[4], P, G, L->wpart, L->u2, d1, L->a1, d1, L->ga1, L->gu2);
    // u2 = up(a2) + e2
    k_up_bwd<<<grid256(32 * d2.size()), 256, 0, s>>>(L->gu2, d1, 32, d2, L->ga2);
    conv_bwd(s, L->L[3], P, G, L->wpart, L->e3, d2, L->a2, d2, L->ga2, L->ge3);
    conv_bwd(s, L->L[2], P, G, L->wpart, L->e2, d1, L->e3, d2, L->ge3, L->ge2);
    k_add_inplace<<<grid256(32 * d1.size()), 256, 0, s>>>(L->ge2, L->gu2, 32 * d1.size());
    conv_bwd(s, L->L[1], P, G, L->wpart, L->e1, d0, L->e2, d1, L->ge2, L->ge1);
    k_add_inplace<<<grid256(16 * d0.size()), 256, 0, s>>>(L->ge1, L->gu1, 16 * d0.size());
    conv_bwd(s, L->L[0], P, G, L->wpart, L->x0, d0, L->e1, d0, L->ge1, nullptr);
}

// e = correction for the defect in L->d (features, forward, output scaling)
static void correction(gputo_handle *h, const HierView<float> &H, int l, const int *ctl)
{
    LearnState *L = h->ln;
    hipStream_t s = h->stream;
    const Grid &g = h->L[l].g;
    k_scaled_norm<<<kRedBlocks, kBlock, 0, s>>>(L->d, H.L[l].AD, g.ndof, L->partials, ctl);
    k_set_sigma<<<1, kBlock, 0, s>>>(L->sc, L->partials, g.ndof, ctl);
    k_features<<<grid256(g.nnod), 256, 0, s>>>(L->d, H.L[l].AD, g.nnod, L->sc, L->x0, ctl);
    learn_forward(h, ctl);
    k_correction<<<grid256(g.nnod), 256, 0, s>>>(L->y, H.L[l].AD, h->L[l].mask, g.nnod, L->sc, L->e, ctl);
}

void learn_stage(gputo_handle *h, const HierView<float> &H, int l, const float *F, float *U, const int *ctl)
{
    LearnState *L = h->ln;
    hipStream_t s = h->stream;
    const Grid &g = h->L[l].g;
    OpArgsT<float> a = view_args(h, H, l, ctl);
    a.U = U;
    a.F = F;
    a.out = L->d;
    level_op(h, H, l, OP_DEFECT, a);
    if (L->vc_count++ % L->log_every == 0) {
        k_log_sample<<<grid256(g.ndof), 256, 0, s>>>(L->d, L->samples, g.ndof, L->count, L->nsamp_max, ctl);
        k_log_count<<<1, 64, 0, s>>>(L->count, L->nsamp_max, ctl);
    }
    if (!L->trained) return;
    correction(h, H, l, ctl);
    OpArgsT<float> b = view_args(h, H, l, ctl);
    b.U = L->e;
    b.out = L->Ae;
    level_op(h, H, l, OP_MATVEC, b);
    launch_dot_partial<float, float>(s, L->d, L->e, g.ndof, L->partials, ctl);
    launch_dot_partial<float, float>(s, L->e, L->Ae, g.ndof, L->partials + kRedBlocks, ctl);
    k_gamma<<<1, kBlock, 0, s>>>(L->sc, L->partials, L->partials + kRedBlocks, L->count, ctl);
    k_axpy_gamma<<<grid256(g.ndof), 256, 0, s>>>(U, L->e, g.ndof, L->sc, ctl);
}

void learn_begin_solve(gputo_handle *h)
{
    LearnState *L = h->ln;
    check_hip(hipMemsetAsync(L->count, 0, 4 * sizeof(int), h->stream), "memset");
    L->vc_count = 0;
}

void learn_end_solve(gputo_handle *h, int *accepted, int *rejected)
{
    LearnState *L = h->ln;
    int c[4];
    check_hip(hipMemcpyAsync(c, L->count, sizeof(c), hipMemcpyDeviceToHost, h->stream), "copy");
    check_hip(hipStreamSynchronize(h->stream), "sync");
    *accepted = c[0];
    *rejected = c[1];
    L->nsamp = c[2];
}

// Adam steps on the logged defects of the last solve (one sample per step,
// cycling), loss = level-2 energy of the correction.  Needs the FP32
// hierarchy of that solve (mixed state current).
int learn_train(gputo_handle *h, int steps, double lr, double *loss_first, double *loss_last)
{
    LearnState *L = h->ln;
    if (!L || !h->mx || L->nsamp <= 0 || steps <= 0) return GPUTO_ERR_ARGUMENT;
    hipStream_t s = h->stream;
    const int l = 2;
    const HierView<float> &H = h->mx->H;
    const Grid &g = h->L[l].g;
    double lf = 0.0, ll = 0.0;
    for (int it = 0; it < steps; ++it) {
        const int smp = it % L->nsamp;
        launch_copy<float, float>(s, L->samples + (long)smp * g.ndof, L->d, g.ndof, nullptr);
        k_zero<<<grid256(L->nparam), 256, 0, s>>>(L->G, L->nparam);
        correction(h, H, l, nullptr);
        OpArgsT<float> b = view_args(h, H, l, nullptr);
        b.U = L->e;
        b.out = L->Ae;
        level_op(h, H, l, OP_MATVEC, b);
        // loss for monitoring
        check_hip(hipMemsetAsync(L->sc + 4, 0, sizeof(double), s), "memset");
        launch_dot_partial<float, float>(s, L->e, L->Ae, g.ndof, L->partials, nullptr);
        launch_dot_partial<float, float>(s, L->e, L->d, g.ndof, L->partials + kRedBlocks, nullptr);
        k_loss_acc<<<1, kBlock, 0, s>>>(L->sc, L->partials, L->partials + kRedBlocks);
        k_loss_grad<<<grid256(g.nnod), 256, 0, s>>>(L->Ae, L->d, H.L[l].AD, h->L[l].mask, g.nnod, L->sc, 1.0f,
                                                    L->gy);
        learn_backward(h);
        L->adam_t += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - std::pow(b1, (float)L->adam_t), c2 = 1.0f - std::pow(b2, (float)L->adam_t);
        k_adam<<<grid256(L->nparam), 256, 0, s>>>(L->P, L->G, L->M, L->V, L->nparam, (float)lr, b1, b2, c1, c2,
                                                   1.0f);
        if (it == 0 || it == steps - 1) {
            double v = 0.0;
            check_hip(hipMemcpyAsync(&v, L->sc + 4, sizeof(double), hipMemcpyDeviceToHost, s), "copy");
            check_hip(hipStreamSynchronize(s), "sync");
            if (it == 0) lf = v;
            ll = v;
        }
    }
    check_hip(hipGetLastError(), "launch");
    L->trained = true;
    if (loss_first) *loss_first = lf;
    if (loss_last) *loss_last = ll;
    return 0;
}

}  // namespace gputo

// ============================================================================
// C ABI of the learned stage (include/gputo.h)
// ============================================================================
using namespace gputo;

namespace {

int learn_guard(gputo_handle *h) { return (h && h->ln) ? 0 : GPUTO_ERR_ARGUMENT; }

template <class F>
int learn_call(F &&f)
{
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "gputo: %s\n", e.what());
        return GPUTO_ERR_DEVICE;
    }
}

void to_float(hipStream_t s, const double *host, float *dev, long n)
{
    std::vector<float> t(host, host + n);
    check_hip(hipMemcpyAsync(dev, t.data(), n * sizeof(float), hipMemcpyHostToDevice, s), "copy");
    check_hip(hipStreamSynchronize(s), "sync");
}

void to_double(hipStream_t s, const float *dev, double *host, long n)
{
    std::vector<float> t(n);
    check_hip(hipMemcpyAsync(t.data(), dev, n * sizeof(float), hipMemcpyDeviceToHost, s), "copy");
    check_hip(hipStreamSynchronize(s), "sync");
    std::copy(t.begin(), t.end(), host);
}

}  // namespace

extern "C" {

int gputo_learn_info(gputo_handle *h, long long *nparam, int *n0, int *n1, int *n2, int *nsamples)
{
    if (int rc = learn_guard(h)) return rc;
    const LearnState *L = h->ln;
    if (nparam) *nparam = L->nparam;
    if (n0) *n0 = L->d0.n0;
    if (n1) *n1 = L->d0.n1;
    if (n2) *n2 = L->d0.n2;
    if (nsamples) *nsamples = L->nsamp;
    return 0;
}

int gputo_learn_params(gputo_handle *h, int set, double *buf)
{
    if (int rc = learn_guard(h)) return rc;
    if (!buf) return GPUTO_ERR_ARGUMENT;
    return learn_call([&] {
        if (set) {
            to_float(h->stream, buf, h->ln->P, h->ln->nparam);
            h->ln->trained = true;
        } else {
            to_double(h->stream, h->ln->P, buf, h->ln->nparam);
        }
    });
}

int gputo_learn_reset(gputo_handle *h, unsigned seed)
{
    if (int rc = learn_guard(h)) return rc;
    return learn_call([&] { learn_init_params(h, seed); });
}

int gputo_learn_forward(gputo_handle *h, const double *x, double *y)
{
    if (int rc = learn_guard(h)) return rc;
    if (!x || !y) return GPUTO_ERR_ARGUMENT;
    return learn_call([&] {
        LearnState *L = h->ln;
        const long n0 = L->d0.size();
        to_float(h->stream, x, L->x0, 4 * n0);
        learn_forward(h, nullptr);
        to_double(h->stream, L->y, y, 3 * n0);
    });
}

int gputo_learn_grad(gputo_handle *h, const double *x, const double *gy, double *g)
{
    if (int rc = learn_guard(h)) return rc;
    if (!x || !gy || !g) return GPUTO_ERR_ARGUMENT;
    return learn_call([&] {
        LearnState *L = h->ln;
        const long n0 = L->d0.size();
        hipStream_t s = h->stream;
        to_float(s, x, L->x0, 4 * n0);
        learn_forward(h, nullptr);
        to_float(s, gy, L->gy, 3 * n0);
        k_zero<<<grid256(L->nparam), 256, 0, s>>>(L->G, L->nparam);
        learn_backward(h);
        check_hip(hipGetLastError(), "launch");
        to_double(s, L->G, g, L->nparam);
    });
}

int gputo_learn_train(gputo_handle *h, int steps, double lr, double *loss_first, double *loss_last)
{
    if (int rc = learn_guard(h)) return rc;
    if (!h->hier_ready || h->ln->nsamp <= 0 || steps < 1 || !(lr > 0.0)) return GPUTO_ERR_ARGUMENT;
    int rc = 0;
    const int crc = learn_call([&] { rc = learn_train(h, steps, lr, loss_first, loss_last); });
    return crc ? crc : rc;
}

}  // extern "C"
