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
// A block of KS waves owns 64 output positions (one per lane) and CT
// output channels; wave v takes the input channels ci = v, v + KS, ...
// The block's CT x cin x 27 weights are staged in LDS once as [ci][t][c]:
// one broadcast read of 4 consecutive c feeds 4 FMAs, and each vector load
// of an input value feeds CT FMAs.  The grids are small, so a layer has
// about one block per CU and the staging is paid about once per CU.  The
// KS partial sums meet in LDS (over the weights, after a barrier) and are
// added in wave order; wave v then finishes the channels c = v, v + KS, ...
// (bias, ReLU, epilogue).
//
// Options carried in ConvFwd:
//  * norm (conv 1 of the correction): input channels 0-2 are scaled by
//    1/sigma, sigma = sqrt(sum(norm[0..norm_n)) / norm_count) from the
//    partials of k_features; every wave sums them in the same order, and
//    block (0, 0) stores sigma in sc[0];
//  * up (convs 4 and 5): the output is also upsampled to the finer grid
//    dup and added to the skip activation there, up[c][q] = out[c][p] +
//    skip[c][q] for the up to 8 fine nodes q with parent p;
//  * corr (conv 6): the output is also turned into the level-2 correction
//    corr[3p+c] = fixed ? 0 : sigma y[c][p] / sqrt(ad[3p+c]).
// input channels whose gathers a thread issues together (one memory
// latency per group instead of per channel)
constexpr int kConvGroup = 4;

struct ConvFwd {
    const float *in;
    Dims3 din;
    int cin, cout, stride, relu;
    const float *w, *b;
    float *out;
    Dims3 dout;
    const double *norm = nullptr;
    int norm_n = 0;
    double norm_count = 1.0;
    double *sc = nullptr;
    const float *skip = nullptr;
    float *up = nullptr;
    Dims3 dup{};
    const float *ad = nullptr;
    const uint8_t *mask = nullptr;
    float *corr = nullptr;
    const int *ctl = nullptr;
};

// 1/sigma from the feature partials (every wave sums them in the same
// order); block (0, 0) stores sigma in sc[0].  Called by whole waves.
__device__ __forceinline__ float conv_inv_sigma(const ConvFwd &a)
{
    double s = 0.0;
    for (int i = (int)(threadIdx.x & 63); i < a.norm_n; i += 64) s += a.norm[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    double sig = sqrt(s / a.norm_count);
    sig = sig > 1e-300 ? sig : 1.0;
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) a.sc[0] = sig;
    return (float)(1.0 / sig);
}

// bias-added sum r of output channel co at position p = (oi, ok, oj):
// activation, store, the corr epilogue and (with_up) the up epilogue
__device__ __forceinline__ void conv_finish(const ConvFwd &a, int co, int p, int nout, int oi, int ok, int oj,
                                            float r, bool with_up)
{
    if (a.relu) r = fmaxf(r, 0.0f);
    a.out[(long)co * nout + p] = r;
    if (with_up && a.up) {
        const int nup = (int)a.dup.size();
        for (int di = 0; di < 2; ++di)
            for (int dk = 0; dk < 2; ++dk)
                for (int dj = 0; dj < 2; ++dj) {
                    const int fi = 2 * oi + di, fk = 2 * ok + dk, fj = 2 * oj + dj;
                    if (fi < a.dup.n0 && fk < a.dup.n1 && fj < a.dup.n2) {
                        const long q = (long)co * nup + (fi * a.dup.n1 + fk) * a.dup.n2 + fj;
                        a.up[q] = r + a.skip[q];
                    }
                }
    }
    if (a.corr) {
        const float sig = (float)a.sc[0];
        const long q = 3L * p + co;
        a.corr[q] = ((a.mask[p] >> co) & 1u) ? 0.0f : sig * r / sqrtf(a.ad[q]);
    }
}

template <int CT, int KS>
__global__ __launch_bounds__(64 * KS) void k_conv_fwd(ConvFwd a)
{
    if (a.ctl && a.ctl[0]) return;
    extern __shared__ __align__(16) float lds[];  // weights [cin][27][CT]; then partials [KS][CT][64]
    constexpr int NT = 64 * KS;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int co0 = blockIdx.y * CT;
    const int nout = (int)a.dout.size(), nin = (int)a.din.size();
    const int p = blockIdx.x * 64 + lane;
    const bool live = p < nout;
    const int nw = a.cin * 27 * CT;
    for (int q0 = 0; q0 < nw; q0 += 8 * NT) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int q = q0 + u * NT + (int)threadIdx.x;
            v[u] = a.w[q < nw ? (co0 + q % CT) * a.cin * 27 + q / CT : 0];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int q = q0 + u * NT + (int)threadIdx.x;
            if (q < nw) lds[q] = v[u];
        }
    }
    const float inv = a.norm ? conv_inv_sigma(a) : 1.0f;
    int oi, ok, oj;
    pos3(live ? p : 0, a.dout, oi, ok, oj);
    int off[27];
    unsigned valid = 0;
#pragma unroll
    for (int t = 0; t < 27; ++t) {
        const int ii = oi * a.stride + t / 9 - 1, ik = ok * a.stride + (t / 3) % 3 - 1,
                  ij = oj * a.stride + t % 3 - 1;
        const bool v = live && ii >= 0 && ii < a.din.n0 && ik >= 0 && ik < a.din.n1 && ij >= 0 && ij < a.din.n2;
        valid |= (unsigned)v << t;
        off[t] = v ? (ii * a.din.n1 + ik) * a.din.n2 + ij : 0;
    }
    __syncthreads();
    float acc[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[c] = 0.0f;
    for (int cg = wv; cg < a.cin; cg += kConvGroup * KS) {
        // unconditional loads (off = 0 outside the grid), masked after: the
        // 27 loads of up to kConvGroup channels are in flight together
        float vg[kConvGroup][27];
#pragma unroll
        for (int u = 0; u < kConvGroup; ++u) {
            const int ci = cg + u * KS;
            const float *src = a.in + (long)(ci < a.cin ? ci : 0) * nin;
#pragma unroll
            for (int t = 0; t < 27; ++t) vg[u][t] = src[off[t]];
        }
#pragma unroll
        for (int u = 0; u < kConvGroup; ++u) {
        const int ci = cg + u * KS;
        if (ci >= a.cin) break;
        const float *v = vg[u];
        const float sc = ci < 3 ? inv : 1.0f;
        const float *wc = lds + ci * 27 * CT;
#pragma unroll
        for (int t = 0; t < 27; ++t) {
            const float vt = ((valid >> t) & 1u) ? v[t] * sc : 0.0f;
            if constexpr (CT % 4 == 0) {
#pragma unroll
                for (int c4 = 0; c4 < CT; c4 += 4) {
                    const float4 w4 = *reinterpret_cast<const float4 *>(wc + t * CT + c4);
                    acc[c4] = fmaf(w4.x, vt, acc[c4]);
                    acc[c4 + 1] = fmaf(w4.y, vt, acc[c4 + 1]);
                    acc[c4 + 2] = fmaf(w4.z, vt, acc[c4 + 2]);
                    acc[c4 + 3] = fmaf(w4.w, vt, acc[c4 + 3]);
                }
            } else {
#pragma unroll
                for (int c = 0; c < CT; ++c) acc[c] = fmaf(wc[t * CT + c], vt, acc[c]);
            }
        }
        }
    }
    float *part = lds;
    if (KS > 1) {
        __syncthreads();  // every wave is done with the weights
#pragma unroll
        for (int c = 0; c < CT; ++c) part[(wv * CT + c) * 64 + lane] = acc[c];
        __syncthreads();
    }
    if (!live) return;
    auto finish = [&](int c, float r) { conv_finish(a, co0 + c, p, nout, oi, ok, oj, r, true); };
    if constexpr (KS == 1) {
#pragma unroll
        for (int c = 0; c < CT; ++c) finish(c, a.b[co0 + c] + acc[c]);
    } else {
        for (int c = wv; c < CT; c += KS) {
            float r = a.b[co0 + c];
            for (int k = 0; k < KS; ++k) r += part[(k * CT + c) * 64 + lane];
            finish(c, r);
        }
    }
}

// The same convolution with the input channels split over the lanes: a
// 256-thread block owns NP = 256/KS positions and CT output channels,
// thread (ks, pl) takes position pl and the channels ci = ks, ks + KS, ...
// The coarse grids (a few hundred positions) use it: their blocks are many
// and small, each input value gathered serves CT outputs, and the KS
// partials of a position are added in ks order.
template <int CT, int KS>
__global__ __launch_bounds__(256) void k_conv_fwd_lanes(ConvFwd a)
{
    if (a.ctl && a.ctl[0]) return;
    extern __shared__ float wl[];  // [cin*27][CT], then KS partials [KS][256/KS][CT]
    constexpr int NP = 256 / KS;
    const int co0 = blockIdx.y * CT;
    const int cin = a.cin;
    stage_lds(wl, cin * 27 * CT, a.w, [=](int q) {
        const int c = q % CT, k = q / CT;  // k = ci*27 + t
        return (co0 + c) * cin * 27 + k;
    });
    const float inv = a.norm ? conv_inv_sigma(a) : 1.0f;
    __syncthreads();
    const int pl = threadIdx.x % NP, ks = threadIdx.x / NP;
    const int p = blockIdx.x * NP + pl;
    const int nout = (int)a.dout.size(), nin = (int)a.din.size();
    float acc[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[c] = 0.0f;
    int oi = 0, ok = 0, oj = 0;
    if (p < nout) {
        pos3(p, a.dout, oi, ok, oj);
        int off[27];
        unsigned valid = 0;
#pragma unroll
        for (int t = 0; t < 27; ++t) {
            const int ii = oi * a.stride + t / 9 - 1, ik = ok * a.stride + (t / 3) % 3 - 1,
                      ij = oj * a.stride + t % 3 - 1;
            const bool v = ii >= 0 && ii < a.din.n0 && ik >= 0 && ik < a.din.n1 && ij >= 0 && ij < a.din.n2;
            valid |= (unsigned)v << t;
            off[t] = v ? (ii * a.din.n1 + ik) * a.din.n2 + ij : 0;
        }
        for (int cg = ks; cg < cin; cg += kConvGroup * KS) {
            // the 27 loads of up to kConvGroup channels in flight together
            float vg[kConvGroup][27];
#pragma unroll
            for (int u = 0; u < kConvGroup; ++u) {
                const int ci = cg + u * KS;
                const float *src = a.in + (long)(ci < cin ? ci : 0) * nin;
#pragma unroll
                for (int t = 0; t < 27; ++t) vg[u][t] = src[off[t]];
            }
#pragma unroll
            for (int u = 0; u < kConvGroup; ++u) {
                const int ci = cg + u * KS;
                if (ci >= cin) break;
                const float *wc = wl + ci * 27 * CT;
                const float sc = ci < 3 ? inv : 1.0f;
#pragma unroll
                for (int t = 0; t < 27; ++t) {
                    const float vt = ((valid >> t) & 1u) ? vg[u][t] * sc : 0.0f;
#pragma unroll
                    for (int c = 0; c < CT; ++c) acc[c] = fmaf(wc[t * CT + c], vt, acc[c]);
                }
            }
        }
    }
    const bool own = ks == 0 && p < nout;  // the thread that finishes position p
    if (KS > 1) {
        float *part = wl + cin * 27 * CT;
        __syncthreads();
#pragma unroll
        for (int c = 0; c < CT; ++c) part[(ks * NP + pl) * CT + c] = acc[c];
        __syncthreads();
        if (own) {
#pragma unroll
            for (int c = 0; c < CT; ++c) {
                float sm = a.b[co0 + c];
                for (int k = 0; k < KS; ++k) sm += part[(k * NP + pl) * CT + c];
                acc[c] = sm;
            }
        }
    } else if (own) {
#pragma unroll
        for (int c = 0; c < CT; ++c) acc[c] += a.b[co0 + c];
    }
    if (own) {
#pragma unroll
        for (int c = 0; c < CT; ++c) conv_finish(a, co0 + c, p, nout, oi, ok, oj, acc[c], false);
    }
    if (a.up) {
        // the up epilogue over the whole block: item (position, channel,
        // child) per thread, from the activated outputs in LDS
        float *res = wl;
        __syncthreads();  // every thread is done with the weights and partials
        if (own) {
#pragma unroll
            for (int c = 0; c < CT; ++c) res[pl * CT + c] = a.relu ? fmaxf(acc[c], 0.0f) : acc[c];
        }
        __syncthreads();
        const int nup = (int)a.dup.size();
        for (int it = threadIdx.x; it < NP * CT * 8; it += 256) {
            const int ch = it & 7, c = (it >> 3) % CT, ql = it / (8 * CT);
            const int pp = blockIdx.x * NP + ql;
            if (pp >= nout) continue;
            int pi, pk, pj;
            pos3(pp, a.dout, pi, pk, pj);
            const int fi = 2 * pi + (ch >> 2), fk = 2 * pk + ((ch >> 1) & 1), fj = 2 * pj + (ch & 1);
            if (fi < a.dup.n0 && fk < a.dup.n1 && fj < a.dup.n2) {
                const long q = (long)(co0 + c) * nup + (fi * a.dup.n1 + fk) * a.dup.n2 + fj;
                a.up[q] = res[ql * CT + c] + a.skip[q];
            }
        }
    }
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

// x[c][p] = d[3p+c] / sqrt(ad[3p+c]) (c < 3), x[3][p] = log10(mean_c ad)/9,
// and partials[block] = the block's sum of d^2/ad in double: conv 1 turns
// the kFeatBlocks partials into sigma, the RMS of the Jacobi-scaled defect,
// and scales channels 0-2 by 1/sigma as it reads them.
constexpr int kFeatBlocks = 128;

__global__ __launch_bounds__(kBlock) void k_features(const float *__restrict__ d, const float *__restrict__ ad,
                                                    long nnod, float *__restrict__ x, double *partials,
                                                    const int *ctl)
{
    if (ctl && ctl[0]) return;
    double s = 0.0;
    for (long p = (long)blockIdx.x * kBlock + threadIdx.x; p < nnod; p += (long)kFeatBlocks * kBlock) {
        float m = 0.0f;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float av = ad[3 * p + c], dv = d[3 * p + c];
            x[c * nnod + p] = dv / sqrtf(av);
            s += (double)dv * (double)dv / (double)av;
            m += av;
        }
        x[3 * nnod + p] = log10f(fmaxf(m / 3.0f, 1e-30f)) / 9.0f;
    }
    s = block_sum(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// x[c][p] *= 1/sigma for c < 3: the input conv 1 saw, for the backward pass
__global__ void k_scale_features(float *__restrict__ x, long nnod, const double *sc)
{
    const float inv = (float)(1.0 / sc[0]);
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < 3 * nnod; q += (long)gridDim.x * 256) x[q] *= inv;
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

// partials of <d,e> (first kFeatBlocks) and <e,Ae> (next kFeatBlocks)
__global__ __launch_bounds__(kBlock) void k_dot2f(const float *__restrict__ d, const float *__restrict__ e,
                                                 const float *__restrict__ Ae, long n, double *part, const int *ctl)
{
    if (ctl && ctl[0]) return;
    double s0 = 0.0, s1 = 0.0;
    for (long q = (long)blockIdx.x * kBlock + threadIdx.x; q < n; q += (long)kFeatBlocks * kBlock) {
        const double ev = e[q];
        s0 += (double)d[q] * ev;
        s1 += ev * (double)Ae[q];
    }
    s0 = block_sum(s0);
    s1 = block_sum(s1);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s0;
        part[kFeatBlocks + blockIdx.x] = s1;
    }
}

// gamma = <d,e>/<e,Ae> from the k_dot2f partials; U += gamma e if gamma is
// positive and finite, else the step is dropped (the plain V-cycle is the
// fallback).  Every block sums the partials in the same order; block 0
// records gamma and the dots (sc[1..3]), the accept/reject counters and,
// when logged > 0, the number of logged samples.
__global__ __launch_bounds__(kBlock) void k_step(float *__restrict__ U, const float *__restrict__ e, long n,
                                                const double *part, double *sc, int *count, int logged,
                                                const int *ctl)
{
    if (ctl && ctl[0]) return;
    const double de = block_total(part, kFeatBlocks);
    const double eAe = block_total(part + kFeatBlocks, kFeatBlocks);
    const double g = de / eAe;
    const bool ok = de > 0.0 && eAe > 0.0 && g > 0.0 && g < INFINITY;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc[1] = ok ? g : 0.0;
        sc[2] = de;
        sc[3] = eAe;
        count[ok ? 0 : 1] += 1;
        if (logged > 0) count[2] = logged;
    }
    if (!ok) return;
    const float gf = (float)g;
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long)gridDim.x * 256) U[q] += gf * e[q];
}

// energy loss of one sample: (1/2 <e,Ae> - <e,d>) / sigma^2 accumulated in sc[4]
__global__ __launch_bounds__(kBlock) void k_loss_acc(double *sc, const double *p_eAe, const double *p_ed)
{
    const double eAe = block_total(p_eAe, kRedBlocks);
    const double ed = block_total(p_ed, kRedBlocks);
    if (threadIdx.x == 0) sc[4] += (0.5 * eAe - ed) / (sc[0] * sc[0]);
}

// number of logged samples of the solve (the defect went to its slot)
__global__ void k_log_count(int *count, int logged, const int *ctl)
{
    if (ctl && ctl[0]) return;
    if (threadIdx.x == 0 && blockIdx.x == 0) count[2] = logged;
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

// Choice of kernel per layer.  A grid with at least kWaveTiles blocks of
// 64 positions x 16 output channels (3 for the output layer) goes to the
// wave kernel with KS = min(8, cin) waves; its blocks stage their weights
// once and each gather feeds 16 FMAs.  Coarser grids go to the lane kernel
// with CT = 4 (3) and the smallest KS (a power of two, at most 32 and at
// most cin) that gives kLaneBlocks blocks, so that they fill the device.
constexpr long kWaveTiles = 128;
constexpr long kLaneBlocks = 1024;

template <int CT>
void conv_wave(hipStream_t s, const ConvFwd &a, int ks)
{
    const dim3 grid((unsigned)((a.dout.size() + 63) / 64), (unsigned)(a.cout / CT));
    const size_t lds = std::max<size_t>((size_t)a.cin * 27 * CT, ks > 1 ? (size_t)ks * CT * 64 : 0) * sizeof(float);
    switch (ks) {
    case 1: k_conv_fwd<CT, 1><<<grid, 64, lds, s>>>(a); break;
    case 2: k_conv_fwd<CT, 2><<<grid, 128, lds, s>>>(a); break;
    case 4: k_conv_fwd<CT, 4><<<grid, 256, lds, s>>>(a); break;
    default: k_conv_fwd<CT, 8><<<grid, 512, lds, s>>>(a); break;
    }
}

template <int CT, int KS>
void conv_lanes_ks(hipStream_t s, const ConvFwd &a)
{
    const dim3 grid((unsigned)((a.dout.size() + 256 / KS - 1) / (256 / KS)), (unsigned)(a.cout / CT));
    // weights and partials; the up epilogue reuses the front for NP x CT outputs
    const size_t lds = std::max<size_t>((size_t)a.cin * 27 * CT + (KS > 1 ? 256 * CT : 0),
                                        a.up ? (size_t)(256 / KS) * CT : 0) * sizeof(float);
    k_conv_fwd_lanes<CT, KS><<<grid, 256, lds, s>>>(a);
}

template <int CT>
void conv_lanes(hipStream_t s, const ConvFwd &a)
{
    int ks = 1;
    while (ks < 32 && 2 * ks <= a.cin &&
           ((a.dout.size() + 256 / ks - 1) / (256 / ks)) * (a.cout / CT) < kLaneBlocks)
        ks *= 2;
    switch (ks) {
    case 1: conv_lanes_ks<CT, 1>(s, a); break;
    case 2: conv_lanes_ks<CT, 2>(s, a); break;
    case 4: conv_lanes_ks<CT, 4>(s, a); break;
    case 8: conv_lanes_ks<CT, 8>(s, a); break;
    case 16: conv_lanes_ks<CT, 16>(s, a); break;
    default: conv_lanes_ks<CT, 32>(s, a); break;
    }
}

void conv_fwd(hipStream_t s, const ConvFwd &a)
{
    const long tiles = (a.dout.size() + 63) / 64;
    int ks = 1;
    while (ks < 8 && 2 * ks <= a.cin) ks *= 2;
    if (a.cout == 3) {
        if (tiles >= kWaveTiles)
            conv_wave<3>(s, a, ks);
        else
            conv_lanes<3>(s, a);
        return;
    }
    if (a.cout % 16 == 0 && a.cin * 27 * 16 <= 16384 && tiles * (a.cout / 16) >= kWaveTiles)
        conv_wave<16>(s, a, ks);
    else if (a.cout % 4 == 0)
        conv_lanes<4>(s, a);
    else
        throw std::runtime_error("conv: output channels not 3 or a multiple of 4");
}

ConvFwd conv_args(const NetLayer &ly, const float *P, const float *in, Dims3 din, float *out, Dims3 dout,
                  const int *ctl)
{
    ConvFwd a;
    a.in = in;
    a.din = din;
    a.cin = ly.cin;
    a.cout = ly.cout;
    a.stride = ly.stride;
    a.relu = ly.relu;
    a.w = P + ly.w_off;
    a.b = P + ly.b_off;
    a.out = out;
    a.dout = dout;
    a.ctl = ctl;
    return a;
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
    learn_policy_reset(L);
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

// y = N(x0) with the forward activations cached.  With io (the correction
// stage) conv 1 scales the defect channels by 1/sigma from the feature
// partials and stores sigma, and conv 6 also writes the correction.
void learn_forward(gputo_handle *h, const int *ctl, const CorrIO *io)
{
    LearnState *L = h->ln;
    hipStream_t s = h->stream;
    const Dims3 d0 = L->d0, d1 = L->d1, d2 = L->d2;
    ConvFwd c1 = conv_args(L->L[0], L->P, L->x0, d0, L->e1, d0, ctl);
    if (io) {
        c1.norm = io->norm;
        c1.norm_n = io->norm_n;
        c1.norm_count = io->norm_count;
        c1.sc = L->sc;
    }
    conv_fwd(s, c1);
    conv_fwd(s, conv_args(L->L[1], L->P, L->e1, d0, L->e2, d1, ctl));
    conv_fwd(s, conv_args(L->L[2], L->P, L->e2, d1, L->e3, d2, ctl));
    ConvFwd c4 = conv_args(L->L[3], L->P, L->e3, d2, L->a2, d2, ctl);  // u2 = up(a2) + e2
    c4.skip = L->e2;
    c4.up = L->u2;
    c4.dup = d1;
    conv_fwd(s, c4);
    ConvFwd c5 = conv_args(L->L[4], L->P, L->u2, d1, L->a1, d1, ctl);  // u1 = up(a1) + e1
    c5.skip = L->e1;
    c5.up = L->u1;
    c5.dup = d0;
    conv_fwd(s, c5);
    ConvFwd c6 = conv_args(L->L[5], L->P, L->u1, d0, L->y, d0, ctl);
    if (io) {
        c6.sc = L->sc;
        c6.ad = io->ad;
        c6.mask = io->mask;
        c6.corr = io->e;
    }
    conv_fwd(s, c6);
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

// e = correction for the defect d (features, forward, output scaling)
static void correction(gputo_handle *h, const HierView<float> &H, int l, const float *d, const int *ctl)
{
    LearnState *L = h->ln;
    const Grid &g = h->L[l].g;
    k_features<<<kFeatBlocks, kBlock, 0, h->stream>>>(d, H.L[l].AD, g.nnod, L->x0, L->partials, ctl);
    const CorrIO io{L->partials, kFeatBlocks, (double)g.ndof, H.L[l].AD, h->L[l].mask, L->e};
    learn_forward(h, ctl, &io);
}

// The first nsamp_max V-cycles of a solve log their defect as training
// samples: the defect is computed straight into its slot.
void learn_stage(gputo_handle *h, const HierView<float> &H, int l, const float *F, float *U, const int *ctl)
{
    LearnState *L = h->ln;
    hipStream_t s = h->stream;
    const Grid &g = h->L[l].g;
    const int slot = L->vc_count++;
    const bool logged = L->logging && slot < L->nsamp_max;
    const bool apply = L->trained && L->use;
    if (!apply && !logged) return;
    float *d = logged ? L->samples + (long)slot * g.ndof : L->d;
    OpArgsT<float> a = view_args(h, H, l, ctl);
    a.U = U;
    a.F = F;
    a.out = d;
    level_op(h, H, l, OP_DEFECT, a);
    if (!apply) {
        k_log_count<<<1, 64, 0, s>>>(L->count, slot + 1, ctl);
        return;
    }
    correction(h, H, l, d, ctl);
    OpArgsT<float> b = view_args(h, H, l, ctl);
    b.U = L->e;
    b.out = L->Ae;
    level_op(h, H, l, OP_MATVEC, b);
    k_dot2f<<<kFeatBlocks, kBlock, 0, s>>>(d, L->e, L->Ae, g.ndof, L->partials, ctl);
    k_step<<<grid256(g.ndof), 256, 0, s>>>(U, L->e, g.ndof, L->partials, L->sc, L->count, logged ? slot + 1 : 0,
                                           ctl);
}

int learn_stage_kind(const gputo_handle *h)
{
    const LearnState *L = h->ln;
    if (L->logging && L->vc_count < L->nsamp_max) return -1;
    return L->trained && L->use ? 1 : 0;
}

void learn_stage_count(gputo_handle *h) { ++h->ln->vc_count; }

// Solve-level fallback of params.learned = 2.  The stage is worth its cost
// only if it shortens the whole solve, which a per-V-cycle test cannot
// see, so successive solves of a design loop are compared by wall time:
//   REF    a solve without the stage; its time is the reference -> TRIAL
//   TRIAL  a solve with the stage: faster than the reference -> ON, else OFF
//   ON     solves with the stage; after kPolRecheck of them -> REF
//   OFF    solves without the stage (each refreshing the reference) for a
//          backoff that doubles after every failed trial (4 .. 64) -> TRIAL
// Neighbouring designs of a loop differ a little, so a trial compares two
// slightly different problems; a harder next design counts against the
// stage (the bias is towards the plain V-cycle).  A new training (or new
// parameters) restarts at REF.
// A trial is also watched while it runs: at every host poll from CG
// iteration kPolAbortIt on, its wall time is compared with the time the
// reference solve took to reach the same relative residual (log-linear
// between the reference's polls).  More than kPolAbortSlack times that, and
// the stage is switched off for the rest of the solve (the flexible beta
// absorbs the change) and the trial counts as failed: a losing trial costs
// a few V-cycles instead of a whole solve.
enum LearnPolicy { POL_REF = 0, POL_TRIAL = 1, POL_ON = 2, POL_OFF = 3 };
constexpr int kPolRecheck = 8, kPolBackoff0 = 4, kPolBackoffMax = 64;
constexpr int kPolAbortIt = 12;
constexpr double kPolAbortSlack = 1.1;

void learn_policy_reset(LearnState *L)
{
    L->pol_state = POL_REF;
    L->pol_off_left = 0;
    L->pol_backoff = kPolBackoff0;
    L->pol_on_count = 0;
    L->pol_t_plain = 0.0;
    L->pol_ref.clear();
}

void learn_begin_solve(gputo_handle *h)
{
    LearnState *L = h->ln;
    check_hip(hipMemsetAsync(L->count, 0, 4 * sizeof(int), h->stream), "memset");
    L->vc_count = 0;
    L->pol_cur.clear();
    L->pol_abort = false;
    if (h->prm.learned == 2 && L->trained)
        L->use = L->pol_state == POL_TRIAL || L->pol_state == POL_ON;
    else
        L->use = true;
}

void learn_end_solve(gputo_handle *h, int *accepted, int *rejected, double seconds)
{
    LearnState *L = h->ln;
    int c[4];
    check_hip(hipMemcpyAsync(c, L->count, sizeof(c), hipMemcpyDeviceToHost, h->stream), "copy");
    check_hip(hipStreamSynchronize(h->stream), "sync");
    *accepted = c[0];
    *rejected = c[1];
    L->nsamp = c[2];
    if (h->prm.learned != 2 || !L->trained || seconds < 0.0) return;
    switch (L->pol_state) {
    case POL_REF:
        L->pol_t_plain = seconds;
        L->pol_ref.swap(L->pol_cur);
        L->pol_state = POL_TRIAL;
        break;
    case POL_TRIAL:
        if (!L->pol_abort && seconds < L->pol_t_plain) {
            L->pol_state = POL_ON;
            L->pol_on_count = 0;
            L->pol_backoff = kPolBackoff0;
        } else {
            L->pol_state = POL_OFF;
            L->pol_off_left = L->pol_backoff;
            L->pol_backoff = std::min(2 * L->pol_backoff, kPolBackoffMax);
        }
        break;
    case POL_ON:
        if (++L->pol_on_count >= kPolRecheck) L->pol_state = POL_REF;
        break;
    default:  // POL_OFF
        L->pol_t_plain = seconds;
        L->pol_ref.swap(L->pol_cur);
        if (--L->pol_off_left <= 0) L->pol_state = POL_TRIAL;
        break;
    }
}

void learn_poll(gputo_handle *h, int it, double seconds, double relres)
{
    LearnState *L = h->ln;
    if (h->prm.learned != 2 || !L->trained) return;
    L->pol_cur.emplace_back(seconds, relres);
    if (L->pol_state != POL_TRIAL || !L->use || it < kPolAbortIt || !(relres > 0.0) || L->pol_ref.empty()) return;
    // the reference's time to reach relres, from (0 s, 1) through its polls
    double t0 = 0.0, r0 = 1.0, t_ref = -1.0;
    for (const auto &pt : L->pol_ref) {
        if (pt.second <= relres && pt.second > 0.0) {
            const double span = std::log(r0) - std::log(pt.second);
            const double f = span > 0.0 ? (std::log(r0) - std::log(relres)) / span : 1.0;
            t_ref = t0 + std::min(1.0, std::max(0.0, f)) * (pt.first - t0);
            break;
        }
        t0 = pt.first;
        r0 = pt.second;
    }
    if (t_ref <= 0.0) return;  // the reference never got this far: the trial is ahead
    if (seconds > kPolAbortSlack * t_ref) {
        L->use = false;
        L->pol_abort = true;
    }
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
        correction(h, H, l, L->d, nullptr);
        k_scale_features<<<grid256(3 * g.nnod), 256, 0, s>>>(L->x0, g.nnod, L->sc);
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
    learn_policy_reset(L);
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
            learn_policy_reset(h->ln);
        } else {
            to_double(h->stream, h->ln->P, buf, h->ln->nparam);
        }
    });
}

int gputo_learn_logging(gputo_handle *h, int on)
{
    if (int rc = learn_guard(h)) return rc;
    h->ln->logging = on != 0;
    return 0;
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
        learn_forward(h, nullptr, nullptr);
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
        learn_forward(h, nullptr, nullptr);
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
