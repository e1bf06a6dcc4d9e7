// Level-0 operator of the mixed-precision inner solve on MFMA: the shared
// 24x24 element matrix runs against 24 x 16 tiles of element vectors with
// v_mfma_f32_16x16x32_bf16 (gfx950), split-BF16 operands (hi + lo, three
// products hi*hi + hi*lo + lo*hi) and FP32 accumulation.
//
// One block owns a TJ x TK x TI brick of nodes.  It stages the nodal values
// of the brick plus a one-node halo and the element scales of the
// (TJ+1)(TK+1)(TI+1) element box in LDS, multiplies every element of the box
// (16 elements per MFMA column tile, rows of Ke padded 24 -> 32, K padded
// 24 -> 32), scales each column by E_e, and adds the rows that belong to
// owned nodes into LDS accumulators in 8-colour element order: elements of
// one colour share no node, so a colour pass is conflict-free and every node
// receives its 8 contributions in a fixed order.  No atomics; deterministic.
#include <algorithm>
#include <cmath>
#include <vector>

#include "mixed.h"

namespace gputo {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMfHalo = 2048;  // halo nodes per block
constexpr int kMfElem = 1280;  // element box per block
constexpr int kMfOwn = 1024;   // owned nodes per block

struct MfmaState {
    Tile t;  // node brick of the MFMA kernel on level 0
};

__device__ constexpr int mIE[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__device__ constexpr int mJE[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__device__ constexpr int mKE[8] = {0, 0, 0, 0, 1, 1, 1, 1};

__device__ __forceinline__ double mf_block_sum(double v)
{
    __shared__ double ws[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    return ws[0] + ws[1] + ws[2] + ws[3];
}

template <int MODE>
__global__ __launch_bounds__(256) void k_mfma_op(OpArgsT<float> a, Tile t)
{
    if (a.ctl && a.ctl[0]) return;
    __shared__ float su[3 * kMfHalo];
    __shared__ float sE[kMfElem];
    __shared__ float acc[3 * kMfOwn];
    const Grid g = a.g;
    int b = blockIdx.x;
    const int bj = b % t.ntj;
    b /= t.ntj;
    const int bk = b % t.ntk;
    const int bi = b / t.ntk;
    const int j0 = bj * t.TJ, k0 = bk * t.TK, i0 = bi * t.TI;
    const int HJ = t.TJ + 2, HK = t.TK + 2, HI = t.TI + 2;
    const int EJ = t.TJ + 1, EK = t.TK + 1, EI = t.TI + 1;
    const int nown = t.TJ * t.TK * t.TI;

    // stage nodal values (AoS, node h -> su[3h + c]) and element scales
    const int rowlen = 3 * HJ;
    for (int q = threadIdx.x; q < rowlen * HK * HI; q += 256) {
        const int row = q / rowlen;
        const int col = q - row * rowlen;
        const int hk = row % HK, hi = row / HK;
        const int hj = col / 3, c = col - 3 * hj;
        const int gi = i0 - 1 + hi, gj = j0 - 1 + hj, gk = k0 - 1 + hk;
        float v = 0.0f;
        if (gi >= 0 && gi < g.nnx && gj >= 0 && gj < g.nny && gk >= 0 && gk < g.nnz) {
            const long dof = 3 * ((long)g.nny * g.nnz * gi + (long)g.nny * gk + gj) + c;
            if (MODE == OP_PRE_JACOBI || MODE == OP_PRE_DEFECT)
                v = (float)a.prerelax * a.F[dof] / a.AD[dof];
            else
                v = a.U[dof];
        }
        su[3 * ((hi * HK + hk) * HJ + hj) + c] = v;
    }
    for (int q = threadIdx.x; q < EJ * EK * EI; q += 256) {
        const int ej = q % EJ;
        const int r = q / EJ;
        const int ek = r % EK, ei = r / EK;
        const int gi = i0 - 1 + ei, gj = j0 - 1 + ej, gk = k0 - 1 + ek;
        float v = 0.0f;
        if (gi >= 0 && gi < g.nx && gj >= 0 && gj < g.ny && gk >= 0 && gk < g.nz)
            v = a.CX[(long)g.ny * g.nz * gi + (long)g.ny * gk + gj];
        sE[q] = v;
    }
    for (int q = threadIdx.x; q < 3 * nown; q += 256) acc[q] = 0.0f;

    // Ke fragments: lane holds rows 16*tile + (lane % 16), K slice 8*(lane/16) .. +7
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lg = lane >> 4;
    bf16x8 Ahi[2], Alo[2];
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
        const int row = 16 * tt + lr;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int k = 8 * lg + i;
            const float v = (row < 24 && k < 24) ? a.Ke[24 * row + k] : 0.0f;
            const __bf16 hi = (__bf16)v;
            Ahi[tt][i] = hi;
            Alo[tt][i] = (__bf16)(v - (float)hi);
        }
    }
    __syncthreads();

    for (int c = 0; c < 8; ++c) {
        const int pj = c & 1, pk = (c >> 1) & 1, pi = (c >> 2) & 1;
        const int nj = (EJ - pj + 1) / 2, nk = (EK - pk + 1) / 2, ni = (EI - pi + 1) / 2;
        const int cnt = nj * nk * ni;
        const int nb = (cnt + 15) / 16;
        for (int bb = wave; bb < nb; bb += 4) {
            const int idx = 16 * bb + lr;
            const bool valid = idx < cnt;
            int ej = 0, ek = 0, ei = 0;
            if (valid) {
                ej = pj + 2 * (idx % nj);
                const int r = idx / nj;
                ek = pk + 2 * (r % nk);
                ei = pi + 2 * (r / nk);
            }
            bf16x8 Bhi, Blo;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int k = 8 * lg + i;
                float v = 0.0f;
                if (valid && k < 24) {
                    const int ln = k / 3, comp = k - 3 * (k / 3);
                    const int h = ((ei + mIE[ln]) * HK + (ek + mKE[ln])) * HJ + (ej + mJE[ln]);
                    v = su[3 * h + comp];
                }
                const __bf16 hi = (__bf16)v;
                Bhi[i] = hi;
                Blo[i] = (__bf16)(v - (float)hi);
            }
            f32x4 d0 = {0.f, 0.f, 0.f, 0.f}, d1 = {0.f, 0.f, 0.f, 0.f};
            d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ahi[0], Bhi, d0, 0, 0, 0);
            d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ahi[0], Blo, d0, 0, 0, 0);
            d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Alo[0], Bhi, d0, 0, 0, 0);
            d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ahi[1], Bhi, d1, 0, 0, 0);
            d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ahi[1], Blo, d1, 0, 0, 0);
            d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Alo[1], Bhi, d1, 0, 0, 0);
            if (valid) {
                const float E = sE[(ei * EK + ek) * EJ + ej];
#pragma unroll
                for (int tt = 0; tt < 2; ++tt) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int m = 16 * tt + 4 * lg + r;  // element DOF of this output row
                        if (m >= 24) continue;
                        const int ln = m / 3, comp = m - 3 * (m / 3);
                        const int hj = ej + mJE[ln], hk = ek + mKE[ln], hi = ei + mIE[ln];
                        if (hj < 1 || hj > t.TJ || hk < 1 || hk > t.TK || hi < 1 || hi > t.TI) continue;
                        const int own = ((hi - 1) * t.TK + (hk - 1)) * t.TJ + (hj - 1);
                        acc[3 * own + comp] += E * (tt ? d1[r] : d0[r]);
                    }
                }
            }
        }
        __syncthreads();
    }

    double part = 0.0;
    const float om = (float)a.omega;
    for (int own = threadIdx.x; own < nown; own += 256) {
        const int tj = own % t.TJ;
        const int r = own / t.TJ;
        const int tk = r % t.TK, ti = r / t.TK;
        const int gi = i0 + ti, gj = j0 + tj, gk = k0 + tk;
        if (gi >= g.nnx || gj >= g.nny || gk >= g.nnz) continue;
        const long nod = (long)g.nny * g.nnz * gi + (long)g.nny * gk + gj;
        const long dof = 3 * nod;
        const unsigned msk = a.mask[nod];
        const int hc = ((ti + 1) * HK + (tk + 1)) * HJ + (tj + 1);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const bool fixed = (msk >> c) & 1u;
            const float y = acc[3 * own + c];
            const float u = su[3 * hc + c];
            if (MODE == OP_MATVEC) {
                const float v = fixed ? 0.0f : y;
                a.out[dof + c] = v;
                part += (double)u * (double)v;
            } else if (MODE == OP_DEFECT || MODE == OP_PRE_DEFECT) {
                const float rv = fixed ? 0.0f : a.F[dof + c] - y;
                a.out[dof + c] = rv;
                if (MODE == OP_PRE_DEFECT) a.out2[dof + c] = u;
            } else {
                const float f = a.F[dof + c];
                const float rv = fixed ? 0.0f : f - y;
                const float ad = a.AD[dof + c];
                const float un = (1.0f - om) * u + om * (rv + ad * u) / ad;
                a.out[dof + c] = un;
                part += (double)f * (double)un;
            }
        }
    }
    if (a.partials) {
        const double s = mf_block_sum(part);
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

void launch_op0_mfma(hipStream_t s, int mode, const OpArgsT<float> &a, const void *ctx)
{
    const MixedState *m = static_cast<const MixedState *>(ctx);
    const Tile &t = m->mf->t;
    const dim3 grid((unsigned)t.blocks()), block(256);
    switch (mode) {
    case OP_MATVEC: k_mfma_op<OP_MATVEC><<<grid, block, 0, s>>>(a, t); break;
    case OP_DEFECT: k_mfma_op<OP_DEFECT><<<grid, block, 0, s>>>(a, t); break;
    case OP_JACOBI: k_mfma_op<OP_JACOBI><<<grid, block, 0, s>>>(a, t); break;
    case OP_PRE_JACOBI: k_mfma_op<OP_PRE_JACOBI><<<grid, block, 0, s>>>(a, t); break;
    case OP_PRE_DEFECT: k_mfma_op<OP_PRE_DEFECT><<<grid, block, 0, s>>>(a, t); break;
    default: break;
    }
}

long mfma_tile_blocks(const Grid &g) { return mfma_tile(g).blocks(); }

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

void mfma_prepare(gputo_handle *, MixedState *) {}

int mfma_blocks(const gputo_handle *h) { return (int)h->mx->mf->t.blocks(); }

// ---------------------------------------------------------------------------
// Lane-layout self-check of the BF16 MFMA: D = A B for a 16x32 A and a
// 32x16 B of small integers (exact in BF16), with the fragment layout the
// operator kernel assumes.  One wave, all accesses in bounds.
// ---------------------------------------------------------------------------
__global__ void k_mfma_selftest(const float *A, const float *B, float *D)
{
    const int lane = threadIdx.x;
    const int lr = lane & 15, lg = lane >> 4;
    bf16x8 fa, fb;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        fa[i] = (__bf16)A[32 * lr + 8 * lg + i];
        fb[i] = (__bf16)B[16 * (8 * lg + i) + lr];
    }
    f32x4 d = {0.f, 0.f, 0.f, 0.f};
    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa, fb, d, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) D[16 * (4 * lg + r) + lr] = d[r];
}

}  // namespace gputo

extern "C" int gputo_selftest_mfma(double *max_err)
{
    using namespace gputo;
    if (!max_err) return GPUTO_ERR_ARGUMENT;
    std::vector<float> A(16 * 32), B(32 * 16), D(256), R(256, 0.f);
    for (int i = 0; i < 16 * 32; ++i) A[i] = (float)((i * 7 + 3) % 11 - 5);
    for (int i = 0; i < 32 * 16; ++i) B[i] = (float)((i * 5 + 1) % 9 - 4);
    for (int r = 0; r < 16; ++r)
        for (int c = 0; c < 16; ++c)
            for (int k = 0; k < 32; ++k) R[16 * r + c] += A[32 * r + k] * B[16 * k + c];
    float *dA = nullptr, *dB = nullptr, *dD = nullptr;
    if (hipMalloc(&dA, A.size() * 4) != hipSuccess || hipMalloc(&dB, B.size() * 4) != hipSuccess ||
        hipMalloc(&dD, D.size() * 4) != hipSuccess)
        return GPUTO_ERR_DEVICE;
    hipError_t e = hipMemcpy(dA, A.data(), A.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dB, B.data(), B.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        k_mfma_selftest<<<1, 64>>>(dA, dB, dD);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(D.data(), dD, D.size() * 4, hipMemcpyDeviceToHost);
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dD);
    if (e != hipSuccess) return GPUTO_ERR_DEVICE;
    double err = 0.0;
    for (int i = 0; i < 256; ++i) err = std::max(err, (double)std::fabs(D[i] - R[i]));
    *max_err = err;
    return 0;
}
