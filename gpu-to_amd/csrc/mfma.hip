// Level-0 operator of the FP32 V-cycle on the matrix cores: the shared
// 24x24 element matrix runs against 24 x 16 tiles of element vectors (16
// elements per MFMA column tile, Ke rows padded 24 -> 32) with FP32
// accumulation.  Flavours of the same kernel:
//   MF_F32    v_mfma_f32_16x16x4_f32, exact FP32 products (K = 24 in 6 steps)
//   MF_BF16X1 v_mfma_f32_16x16x32_bf16 on bf16(Ke) and bf16(u), one product
//   MF_BF16   split operands x = hi + lo, products hi*hi + hi*lo + lo*hi
//   MF_BF16X3 three-way split x = hi + mid + lo (exact for FP32 inputs) and
//             the six products whose order is >= 2^-24
//
// One block owns a TJ x TK x TI brick of nodes (bricks dealt in XCD order).
// It stages the nodal values of the brick plus a one-node halo and the
// element scales of its (TJ+1)(TK+1)(TI+1) element box in LDS, one j-row per
// wave.  Each wave multiplies 16 elements of one colour at a time (element
// coordinates from the colour index by reciprocal multiplication), scales
// each column by E_e and adds the 24 rows into LDS node accumulators with
// ds_add_f32.  Elements of one colour share no node, so a colour pass never
// adds twice to one address, and every node receives its 8 element
// contributions in the fixed colour order: no global atomics, bitwise
// deterministic.
#include <algorithm>
#include <cmath>
#include <vector>

#include "mixed.h"

namespace gputo {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

enum MfFlavor { MF_F32 = 0, MF_BF16 = 1, MF_BF16X3 = 2, MF_BF16X1 = 3 };

constexpr int kMfHalo = 2048;  // halo nodes per block
constexpr int kMfElem = 1280;  // element box per block
constexpr int kMfOwn = 1024;   // owned nodes per block

struct MfmaState {
    Tile t;  // node brick of the MFMA kernel on level 0
};

__device__ constexpr int mIE[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__device__ constexpr int mJE[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__device__ constexpr int mKE[8] = {0, 0, 0, 0, 1, 1, 1, 1};


// halo-relative offset (in floats) of element DOF k for an element whose
// origin node has halo index 0
__device__ __forceinline__ int dof_off(int k, int HJ, int HK)
{
    const int ln = k / 3, comp = k - 3 * (k / 3);
    return 3 * ((mIE[ln] * HK + mKE[ln]) * HJ + mJE[ln]) + comp;
}

__device__ __forceinline__ void split2(float v, __bf16 &h, __bf16 &l)
{
    h = (__bf16)v;
    l = (__bf16)(v - (float)h);
}

__device__ __forceinline__ void split3(float v, __bf16 &h, __bf16 &m, __bf16 &l)
{
    h = (__bf16)v;
    const float r = v - (float)h;
    m = (__bf16)r;
    l = (__bf16)(r - (float)m);
}

// q / d for 0 <= q < 2^22 with inv = 1/d (float quotient corrected both ways)
__device__ __forceinline__ int mf_udiv(int q, int d, float inv)
{
    int r = (int)((float)q * inv);
    r -= (r * d > q) ? 1 : 0;
    r += ((r + 1) * d <= q) ? 1 : 0;
    return r;
}

template <int FL, int MODE>
__global__ __launch_bounds__(256) void k_mfma_op(OpArgsT<float> a, Tile t)
{
    if (a.ctl && a.ctl[0]) return;
    extern __shared__ float smem[];
    const Grid g = a.g;
    const long tb = xcd_tile(blockIdx.x, t.blocks());
    if (tb >= t.blocks()) {  // padding block of the XCD order
        if (a.partials && threadIdx.x == 0) a.partials[blockIdx.x] = 0.0;
        return;
    }
    int b = (int)tb;
    const int bj = b % t.ntj;
    b /= t.ntj;
    const int bk = b % t.ntk;
    const int bi = b / t.ntk;
    const int j0 = bj * t.TJ, k0 = bk * t.TK, i0 = bi * t.TI;
    const int HJ = t.TJ + 2, HK = t.TK + 2, HI = t.TI + 2;
    const int EJ = t.TJ + 1, EK = t.TK + 1, EI = t.TI + 1;
    const int NH = HJ * HK * HI;
    float *su = smem;          // [3*NH] nodal values, AoS
    float *acc = su + 3 * NH;  // [3*NH] node accumulators
    float *cE = acc + 3 * NH;  // [EI*EK*EJ] element scales

    // stage rows of the halo brick and of the element box, one j-row per
    // wave (coalesced, no per-lane integer division)
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int row = wave; row < HK * HI; row += 4) {
        const int hi = row / HK, hk = row - hi * HK;
        const int gi = i0 - 1 + hi, gk = k0 - 1 + hk;
        const bool rowok = gi >= 0 && gi < g.nnx && gk >= 0 && gk < g.nnz;
        const long base = 3 * ((long)g.nny * g.nnz * gi + (long)g.nny * gk + (j0 - 1));
        for (int col = lane; col < 3 * HJ; col += 64) {
            const int gj = j0 - 1 + col / 3;
            float v = 0.0f;
            if (rowok && gj >= 0 && gj < g.nny) {
                const long dof = base + col;
                if (MODE == OP_PRE_JACOBI || MODE == OP_PRE_DEFECT)
                    v = (float)a.prerelax * a.F[dof] / a.AD[dof];
                else
                    v = a.U[dof];
            }
            su[3 * row * HJ + col] = v;
            acc[3 * row * HJ + col] = 0.0f;
        }
    }
    for (int row = wave; row < EK * EI; row += 4) {
        const int ei = row / EK, ek = row - ei * EK;
        const int gi = i0 - 1 + ei, gk = k0 - 1 + ek;
        const bool rowok = gi >= 0 && gi < g.nx && gk >= 0 && gk < g.nz;
        for (int ej = lane; ej < EJ; ej += 64) {
            const int gj = j0 - 1 + ej;
            float E = 0.0f;
            if (rowok && gj >= 0 && gj < g.ny) E = a.CX[(long)g.ny * g.nz * gi + (long)g.ny * gk + gj];
            cE[row * EJ + ej] = E;
        }
    }

    const int lr = lane & 15, lg = lane >> 4;
    // A fragments (Ke rows 16*tile + lr) and per-lane DOF offsets
    float Af[2][6];
    bf16x8 Ah[2], Am[2], Al[2];
    int koff[8];
    if (FL == MF_F32) {
#pragma unroll
        for (int tt = 0; tt < 2; ++tt)
#pragma unroll
            for (int s = 0; s < 6; ++s) {
                const int row = 16 * tt + lr;
                Af[tt][s] = row < 24 ? a.Ke[24 * row + 4 * s + lg] : 0.0f;
            }
#pragma unroll
        for (int s = 0; s < 6; ++s) koff[s] = dof_off(4 * s + lg, HJ, HK);
    } else {
#pragma unroll
        for (int tt = 0; tt < 2; ++tt) {
            const int row = 16 * tt + lr;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int k = 8 * lg + i;
                const float v = (row < 24 && k < 24) ? a.Ke[24 * row + k] : 0.0f;
                __bf16 h, m, l;
                if (FL == MF_BF16X3) {
                    split3(v, h, m, l);
                    Am[tt][i] = m;
                } else {
                    split2(v, h, l);
                }
                Ah[tt][i] = h;
                Al[tt][i] = l;
            }
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) koff[i] = (8 * lg + i < 24) ? dof_off(8 * lg + i, HJ, HK) : -1;
    }
    int ooff[2][4];
#pragma unroll
    for (int tt = 0; tt < 2; ++tt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int m = 16 * tt + 4 * lg + r;
            ooff[tt][r] = m < 24 ? dof_off(m, HJ, HK) : -1;
        }
    __syncthreads();

    // colour c = (ej&1) + 2(ek&1) + 4(ei&1): elements of one colour share no
    // node, so a pass adds at most once to any accumulator
    for (int c = 0; c < 8; ++c) {
        const int pj = c & 1, pk = (c >> 1) & 1, pi = (c >> 2) & 1;
        const int nj = (EJ - pj + 1) / 2, nk = (EK - pk + 1) / 2, ni = (EI - pi + 1) / 2;
        const int cnt = nj * nk * ni;
        const float inv_nj = 1.0f / (float)nj, inv_nk = 1.0f / (float)nk;
        const int nb = (cnt + 15) / 16;
        for (int bb = wave; bb < nb; bb += 4) {
            const int q = 16 * bb + lr;
            const bool valid = q < cnt;
            const int qq = valid ? q : 0;
            const int r1 = mf_udiv(qq, nj, inv_nj);
            const int r2 = mf_udiv(r1, nk, inv_nk);
            const int ej = pj + 2 * (qq - r1 * nj), ek = pk + 2 * (r1 - r2 * nk), ei = pi + 2 * r2;
            const int base = 3 * ((ei * HK + ek) * HJ + ej);
            const float E = valid ? cE[(ei * EK + ek) * EJ + ej] : 0.0f;
            f32x4 d0 = {0.f, 0.f, 0.f, 0.f}, d1 = {0.f, 0.f, 0.f, 0.f};
            if (FL == MF_F32) {
#pragma unroll
                for (int s = 0; s < 6; ++s) {
                    const float bv = valid ? su[base + koff[s]] : 0.0f;
                    d0 = __builtin_amdgcn_mfma_f32_16x16x4f32(Af[0][s], bv, d0, 0, 0, 0);
                    d1 = __builtin_amdgcn_mfma_f32_16x16x4f32(Af[1][s], bv, d1, 0, 0, 0);
                }
            } else {
                bf16x8 Bh, Bm, Bl;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const float v = (valid && koff[i] >= 0) ? su[base + koff[i]] : 0.0f;
                    __bf16 h, m, l;
                    if (FL == MF_BF16X3) {
                        split3(v, h, m, l);
                        Bm[i] = m;
                    } else {
                        split2(v, h, l);
                    }
                    Bh[i] = h;
                    Bl[i] = l;
                }
#pragma unroll
                for (int tt = 0; tt < 2; ++tt) {
                    f32x4 d = {0.f, 0.f, 0.f, 0.f};
                    // smallest terms first
                    if (FL == MF_BF16X3) {
                        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al[tt], Bh, d, 0, 0, 0);
                        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Am[tt], Bm, d, 0, 0, 0);
                        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah[tt], Bl, d, 0, 0, 0);
                        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Am[tt], Bh, d, 0, 0, 0);
                        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah[tt], Bm, d, 0, 0, 0);
                    } else if (FL == MF_BF16) {
                        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Al[tt], Bh, d, 0, 0, 0);
                        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah[tt], Bl, d, 0, 0, 0);
                    }
                    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ah[tt], Bh, d, 0, 0, 0);
                    if (tt == 0)
                        d0 = d;
                    else
                        d1 = d;
                }
            }
            if (valid) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    atomicAdd(&acc[base + ooff[0][r]], E * d0[r]);
                    if (ooff[1][r] >= 0) atomicAdd(&acc[base + ooff[1][r]], E * d1[r]);
                }
            }
        }
        __syncthreads();
    }

    double part = 0.0;
    const float om = (float)a.omega;
    const int nown = t.TJ * t.TK * t.TI;
    const float inv_tj = 1.0f / (float)t.TJ, inv_tk = 1.0f / (float)t.TK;
    for (int own = threadIdx.x; own < nown; own += 256) {
        const int r = mf_udiv(own, t.TJ, inv_tj);
        const int tj = own - r * t.TJ;
        const int ti = mf_udiv(r, t.TK, inv_tk);
        const int tk = r - ti * t.TK;
        const int gi = i0 + ti, gj = j0 + tj, gk = k0 + tk;
        if (gi >= g.nnx || gj >= g.nny || gk >= g.nnz) continue;
        const long nod = (long)g.nny * g.nnz * gi + (long)g.nny * gk + gj;
        const long dof = 3 * nod;
        const unsigned msk = a.mask[nod];
        const int hc = 3 * (((ti + 1) * HK + (tk + 1)) * HJ + (tj + 1));
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            const bool fixed = (msk >> cc) & 1u;
            const float y = acc[hc + cc];
            const float u = su[hc + cc];
            if (MODE == OP_MATVEC) {
                const float v = fixed ? 0.0f : y;
                a.out[dof + cc] = v;
                part += (double)u * (double)v;
            } else if (MODE == OP_DEFECT || MODE == OP_PRE_DEFECT) {
                const float rv = fixed ? 0.0f : a.F[dof + cc] - y;
                a.out[dof + cc] = rv;
                if (MODE == OP_PRE_DEFECT) a.out2[dof + cc] = u;
            } else {
                const float f = a.F[dof + cc];
                const float rv = fixed ? 0.0f : f - y;
                const float ad = a.AD[dof + cc];
                const float un = (1.0f - om) * u + om * (rv + ad * u) / ad;
                a.out[dof + cc] = un;
                part += (double)f * (double)un;
            }
        }
    }
    if (a.partials) {
        const double s = block_sum(part);
        if (threadIdx.x == 0) a.partials[blockIdx.x] = s;
    }
}

static int mf_balance(int n, int tmax)
{
    if (tmax >= n) return n;
    const int nt = (n + tmax - 1) / tmax;
    return (n + nt - 1) / nt;
}

static Tile mfma_tile(const Grid &g)
{
    Tile t;
    t.TJ = mf_balance(g.nny, 32);
    t.TK = mf_balance(g.nnz, 8);
    t.TI = mf_balance(g.nnx, 4);
    while ((t.TJ + 2) * (t.TK + 2) * (t.TI + 2) > kMfHalo || (t.TJ + 1) * (t.TK + 1) * (t.TI + 1) > kMfElem ||
           t.TJ * t.TK * t.TI > kMfOwn) {
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

static size_t mf_lds_bytes(const Tile &t)
{
    const size_t NH = (size_t)(t.TJ + 2) * (t.TK + 2) * (t.TI + 2);
    const size_t NE = (size_t)(t.TJ + 1) * (t.TK + 1) * (t.TI + 1);
    return (6 * NH + NE) * sizeof(float);
}

template <int FL>
static void launch_flavor(hipStream_t s, int mode, const OpArgsT<float> &a, const Tile &t)
{
    const dim3 grid((unsigned)xcd_launch(t.blocks())), block(256);
    const size_t lds = mf_lds_bytes(t);
    switch (mode) {
    case OP_MATVEC: k_mfma_op<FL, OP_MATVEC><<<grid, block, lds, s>>>(a, t); break;
    case OP_DEFECT: k_mfma_op<FL, OP_DEFECT><<<grid, block, lds, s>>>(a, t); break;
    case OP_JACOBI: k_mfma_op<FL, OP_JACOBI><<<grid, block, lds, s>>>(a, t); break;
    case OP_PRE_JACOBI: k_mfma_op<FL, OP_PRE_JACOBI><<<grid, block, lds, s>>>(a, t); break;
    case OP_PRE_DEFECT: k_mfma_op<FL, OP_PRE_DEFECT><<<grid, block, lds, s>>>(a, t); break;
    default: break;
    }
}

void launch_op0_mfma(hipStream_t s, int mode, const OpArgsT<float> &a, const void *ctx, int flavor)
{
    const MixedState *m = static_cast<const MixedState *>(ctx);
    const Tile &t = m->mf->t;
    if (flavor == MF_F32)
        launch_flavor<MF_F32>(s, mode, a, t);
    else if (flavor == MF_BF16X3)
        launch_flavor<MF_BF16X3>(s, mode, a, t);
    else if (flavor == MF_BF16X1)
        launch_flavor<MF_BF16X1>(s, mode, a, t);
    else
        launch_flavor<MF_BF16>(s, mode, a, t);
}

long mfma_tile_blocks(const Grid &g) { return xcd_launch(mfma_tile(g).blocks()); }

void mfma_create(gputo_handle *h, MixedState *m)
{
    m->mf = new MfmaState();
    m->mf->t = mfma_tile(h->L[0].g);
}

void mfma_destroy(gputo_handle *, MixedState *m)
{
    delete m->mf;
    m->mf = nullptr;
}

int mfma_blocks(const gputo_handle *h) { return (int)xcd_launch(h->mx->mf->t.blocks()); }

// ---------------------------------------------------------------------------
// Lane-layout self-checks of the two MFMA instructions the level-0 kernel
// uses: D = A B for small integer matrices (exact in BF16 and FP32) with the
// fragment layouts the operator kernel assumes.  One wave, all accesses in
// bounds.
// ---------------------------------------------------------------------------
__global__ void k_mfma_selftest(const float *A, const float *B, float *D, const float *A4, const float *B4,
                                float *D4)
{
    const int lane = threadIdx.x;
    const int lr = lane & 15, lg = lane >> 4;
    bf16x8 fa, fb;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        fa[i] = (__bf16)A[32 * lr + 8 * lg + i];  // A: 16 x 32
        fb[i] = (__bf16)B[16 * (8 * lg + i) + lr];  // B: 32 x 16
    }
    f32x4 d = {0.f, 0.f, 0.f, 0.f};
    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa, fb, d, 0, 0, 0);
    f32x4 d4 = {0.f, 0.f, 0.f, 0.f};
    d4 = __builtin_amdgcn_mfma_f32_16x16x4f32(A4[4 * lr + lg], B4[16 * lg + lr], d4, 0, 0, 0);  // 16x4 * 4x16
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        D[16 * (4 * lg + r) + lr] = d[r];
        D4[16 * (4 * lg + r) + lr] = d4[r];
    }
}

}  // namespace gputo

extern "C" int gputo_selftest_mfma(double *max_err)
{
    using namespace gputo;
    if (!max_err) return GPUTO_ERR_ARGUMENT;
    std::vector<float> A(16 * 32), B(32 * 16), R(256, 0.f), A4(16 * 4), B4(4 * 16), R4(256, 0.f);
    for (int i = 0; i < 16 * 32; ++i) A[i] = (float)((i * 7 + 3) % 11 - 5);
    for (int i = 0; i < 32 * 16; ++i) B[i] = (float)((i * 5 + 1) % 9 - 4);
    for (int i = 0; i < 64; ++i) {
        A4[i] = (float)((i * 3 + 2) % 7 - 3);
        B4[i] = (float)((i * 11 + 5) % 13 - 6);
    }
    for (int r = 0; r < 16; ++r)
        for (int c = 0; c < 16; ++c) {
            for (int k = 0; k < 32; ++k) R[16 * r + c] += A[32 * r + k] * B[16 * k + c];
            for (int k = 0; k < 4; ++k) R4[16 * r + c] += A4[4 * r + k] * B4[16 * k + c];
        }
    float *dev = nullptr;
    const size_t n = A.size() + B.size() + 256 + A4.size() + B4.size() + 256;
    if (hipMalloc(&dev, n * sizeof(float)) != hipSuccess) return GPUTO_ERR_DEVICE;
    float *dA = dev, *dB = dA + A.size(), *dD = dB + B.size(), *dA4 = dD + 256, *dB4 = dA4 + A4.size(),
          *dD4 = dB4 + B4.size();
    hipError_t e = hipMemcpy(dA, A.data(), A.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dB, B.data(), B.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dA4, A4.data(), A4.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dB4, B4.data(), B4.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        k_mfma_selftest<<<1, 64>>>(dA, dB, dD, dA4, dB4, dD4);
        e = hipGetLastError();
    }
    std::vector<float> D(256), D4(256);
    if (e == hipSuccess) e = hipMemcpy(D.data(), dD, 256 * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(D4.data(), dD4, 256 * 4, hipMemcpyDeviceToHost);
    (void)hipFree(dev);
    if (e != hipSuccess) return GPUTO_ERR_DEVICE;
    double err = 0.0;
    for (int i = 0; i < 256; ++i) {
        err = std::max(err, (double)std::fabs(D[i] - R[i]));
        err = std::max(err, (double)std::fabs(D4[i] - R4[i]));
    }
    *max_err = err;
    return 0;
}
