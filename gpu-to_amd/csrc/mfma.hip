// Level-0 operator of the FP32 V-cycle on the matrix cores.
//
// Element form: the shared element matrix Ke (24 x 24) multiplies the 24
// nodal values of 32 elements at a time with v_mfma_f32_32x32x16_bf16 (Ke
// rows and columns padded to 32, FP32 accumulation), each result column is
// scaled by its element's E, and every node then sums the contributions of
// its 8 elements in a fixed order.  FP32 operands are split into BF16
// parts x = hi + mid + lo; the flavours differ in the products they keep:
//   MF_F32    all nine products (hi/mid/lo x hi/mid/lo)
//   MF_BF16X3 the six products of order >= 2^-16 (FP32-level accuracy)
//   MF_BF16   hi*hi + hi*lo + lo*hi on a two-way split
//   MF_BF16X1 hi*hi only
//
// One block owns a TJ x TK column of nodes and streams TI node layers along
// i.  LDS holds a ring of three halo node planes (each node as BF16 parts,
// 4 per node with a zero pad, so a pair of j-neighbours is one 16-byte MFMA
// operand fragment) and a ring of two element layers of results (32 floats
// per element: 8 nodes x (3 components + pad)).  Per node layer: the waves
// multiply the next element layer, a barrier, each thread sums the 8 element
// results of one node and applies the operator epilogue, the next node plane
// is staged, a barrier.  No atomics, no colouring, one fixed summation
// order: bitwise deterministic.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "mixed.h"

namespace gputo {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

enum MfFlavor { MF_F32 = 0, MF_BF16 = 1, MF_BF16X3 = 2, MF_BF16X1 = 3 };

struct MfmaState {
    Tile t;  // node brick of the MFMA kernel on level 0
};

// q / d for 0 <= q < 2^22 with inv = 1/d (float quotient corrected both ways)
__device__ __forceinline__ int mf_udiv(int q, int d, float inv)
{
    int r = (int)((float)q * inv);
    r -= (r * d > q) ? 1 : 0;
    r += ((r + 1) * d <= q) ? 1 : 0;
    return r;
}

// local node number of offset (x, y, z) = (i, j, k) in {0,1}^3 in the element
// order ie={0,1,1,0,0,1,1,0}, je={0,0,1,1,0,0,1,1}, ke={0,0,0,0,1,1,1,1}
__host__ __device__ constexpr int mf_lidx(int x, int y, int z) { return 4 * z + (y ? (x ? 2 : 3) : (x ? 1 : 0)); }

// slot of local node ln in element q's 8 result slots (16 bytes each, 128
// bytes per element): permuted by q so that 8 consecutive elements touch
// the 8 different 16-byte bank groups of 128 bytes that a 16-byte LDS access
// of 8 lanes is served from (the counters show a third fewer bank-conflict
// cycles than with groups of 256 bytes assumed)
__device__ __forceinline__ int mf_swz(int q, int ln) { return ln ^ (q & 7); }

// number of BF16 parts per operand for a flavour
__host__ __device__ constexpr int mf_parts(int fl) { return fl == MF_BF16X1 ? 1 : (fl == MF_BF16 ? 2 : 3); }

struct MfGeom {
    int HJ, HK, NPL;  // halo plane (TK+2) x (TJ+2) nodes
    int EJ, EK, NEL;  // element layer (TK+1) x (TJ+1)
};

__device__ __forceinline__ MfGeom mf_geom(const Tile &t)
{
    MfGeom m;
    m.HJ = t.TJ + 2;
    m.HK = t.TK + 2;
    m.NPL = m.HJ * m.HK;
    m.EJ = t.TJ + 1;
    m.EK = t.TK + 1;
    m.NEL = m.EJ * m.EK;
    return m;
}

__device__ __forceinline__ uint32_t bf16_bits(__bf16 v) { return (uint32_t) * reinterpret_cast<unsigned short *>(&v); }

// the NP BF16 parts of v: part 0 = bf16(v), part 1 = bf16(v - part0), ...
template <int NP>
__device__ __forceinline__ void mf_split(float v, float out[NP])
{
    float r = v;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const __bf16 h = (__bf16)r;
        out[p] = (float)h;
        r -= out[p];
    }
}

// Stage node plane gi (halo coordinates: j from j0-1, k from k0-1) as BF16
// parts: slot layout [part][node] of uint2 = (c0 | c1 << 16, c2 | 0 << 16).
// Values are first loaded into registers (vals) so the load can be issued
// early and the LDS write done later.
template <int MODE>
__device__ __forceinline__ void mf_plane_load(const OpArgsT<float> &a, const MfGeom &m, int gi, int j0, int k0,
                                             float vals[2][3])
{
    const Grid &g = a.g;
#pragma unroll
    for (int rr = 0; rr < 2; ++rr) {
        const int node = threadIdx.x + 256 * rr;
        vals[rr][0] = vals[rr][1] = vals[rr][2] = 0.0f;
        if (node >= m.NPL) continue;
        const int hk = node / m.HJ, hj = node - hk * m.HJ;
        const int gj = j0 - 1 + hj, gk = k0 - 1 + hk;
        if (gi < 0 || gi >= g.nnx || gj < 0 || gj >= g.nny || gk < 0 || gk >= g.nnz) continue;
        const long dof = 3 * ((long)g.nny * g.nnz * gi + (long)g.nny * gk + gj);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (MODE == OP_PRE_JACOBI || MODE == OP_PRE_DEFECT)
                vals[rr][c] = (float)a.prerelax * a.F[dof + c] / a.AD[dof + c];
            else
                vals[rr][c] = a.U[dof + c];
        }
    }
}

template <int NP>
__device__ __forceinline__ void mf_plane_store(uint2 *su_slot, const MfGeom &m, const float vals[2][3])
{
#pragma unroll
    for (int rr = 0; rr < 2; ++rr) {
        const int node = threadIdx.x + 256 * rr;
        if (node >= m.NPL) continue;
        float p0[NP], p1[NP], p2[NP];
        mf_split<NP>(vals[rr][0], p0);
        mf_split<NP>(vals[rr][1], p1);
        mf_split<NP>(vals[rr][2], p2);
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            uint2 w;
            w.x = bf16_bits((__bf16)p0[p]) | (bf16_bits((__bf16)p1[p]) << 16);
            w.y = bf16_bits((__bf16)p2[p]);
            su_slot[p * m.NPL + node] = w;
        }
    }
}

constexpr int kMfBatchesPerWave = 3;  // element layers of at most 12 batches of 32

// Element coordinates of the batches a lane works on (batch b = wave + 4 kb,
// element q = 32 b + lane % 32), the same in every element layer: first-node
// position in the halo plane (np), result slot (qd: NEL for the padding lanes
// of the last batch) and the element's (k, j) offset in the mesh (eo, -1 when
// outside it or a padding lane); np = -1 for batches past the last.
struct MfBatch {
    int np[kMfBatchesPerWave], qd[kMfBatchesPerWave], eo[kMfBatchesPerWave];
};

__device__ __forceinline__ void mf_batches(const Grid &g, const MfGeom &m, int j0, int k0, MfBatch &B)
{
    const int n = threadIdx.x & 31, wave = threadIdx.x >> 6;
    const int nb = (m.NEL + 31) / 32;
    const float inv_ej = 1.0f / (float)m.EJ;
#pragma unroll
    for (int kb = 0; kb < kMfBatchesPerWave; ++kb) {
        const int b = wave + 4 * kb;
        const int q = 32 * b + n;
        const bool valid = q < m.NEL;
        const int qq = valid ? q : 0;
        const int ek = mf_udiv(qq, m.EJ, inv_ej), ej = qq - ek * m.EJ;
        const int gj = j0 - 1 + ej, gk = k0 - 1 + ek;
        B.np[kb] = b < nb ? ek * m.HJ + ej : -1;
        B.qd[kb] = valid ? q : m.NEL;
        B.eo[kb] = (valid && gj >= 0 && gj < g.ny && gk >= 0 && gk < g.nz) ? g.ny * gk + gj : -1;
    }
}

// Element scales of element layer ge_i for the lane's batches (zero outside
// the mesh and for padding lanes).
__device__ __forceinline__ void mf_load_E(const OpArgsT<float> &a, int ge_i, const MfBatch &B,
                                          float (&E)[kMfBatchesPerWave])
{
    const Grid &g = a.g;
    const bool in = ge_i >= 0 && ge_i < g.nx;
#pragma unroll
    for (int kb = 0; kb < kMfBatchesPerWave; ++kb)
        E[kb] = (in && B.eo[kb] >= 0) ? a.CX[(long)g.ny * g.nz * ge_i + B.eo[kb]] : 0.0f;
}

// Multiply one element layer: element (ej, ek) of the layer has its first
// node at halo plane position ek*HJ + ej of planes sx0 (x = 0) and sx1
// (x = 1).  Results E_e * (Ke u_e) go to T[q] (q = ek*EJ + ej), rows in
// local-node order with a pad: T[q][slot(ln)][0..3].
template <int FL>
__device__ __forceinline__ void mf_layer(const MfGeom &m, const uint2 *__restrict__ sx0,
                                         const uint2 *__restrict__ sx1, float4 *__restrict__ T,
                                         const bf16x8 (&Ka)[3][2], const MfBatch &B,
                                         const float (&E)[kMfBatchesPerWave])
{
    constexpr int NP = mf_parts(FL);
    const int lane = threadIdx.x & 63;
    const int h = lane >> 5;
#pragma unroll
    for (int kb = 0; kb < kMfBatchesPerWave; ++kb) {
        const int np = B.np[kb];
        if (np < 0) break;  // wave-uniform
        // operand fragments: K step s covers node-pair groups 2s and 2s+1,
        // lane half h takes group 2s + h = (x, z) = (h, s)
        bf16x8 Bf[NP][2];
#pragma unroll
        for (int st = 0; st < 2; ++st) {
            const uint2 *src = (h ? sx1 : sx0) + np + st * m.HJ;
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                const uint2 w0 = src[p * m.NPL], w1 = src[p * m.NPL + 1];
                union {
                    uint32_t u[4];
                    bf16x8 v;
                } cv;
                cv.u[0] = w0.x;
                cv.u[1] = w0.y;
                cv.u[2] = w1.x;
                cv.u[3] = w1.y;
                Bf[p][st] = cv.v;
            }
        }
        // one accumulation chain, smallest products first
        f32x16 d = {};
#pragma unroll
        for (int st = 0; st < 2; ++st) {
            if (FL == MF_F32) {
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[2][st], Bf[2][st], d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[2][st], Bf[1][st], d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[1][st], Bf[2][st], d, 0, 0, 0);
            }
            if (FL == MF_F32 || FL == MF_BF16X3) {
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[2][st], Bf[0][st], d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[1][st], Bf[1][st], d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[0][st], Bf[2][st], d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[1][st], Bf[0][st], d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[0][st], Bf[1][st], d, 0, 0, 0);
            }
            if (FL == MF_BF16) {
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[1][st], Bf[0][st], d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[0][st], Bf[1][st], d, 0, 0, 0);
            }
        }
#pragma unroll
        for (int st = 0; st < 2; ++st) d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[0][st], Bf[0][st], d, 0, 0, 0);
        // rows 8j + 4h + r: node pair group j = (x, z) = (j & 1, j >> 1), the
        // pair member y = h, component r (r = 3: pad row, zero)
        const float e = E[kb];
        const int qd = B.qd[kb];
        float4 *dst = T + (long)qd * 8;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int ln = h ? mf_lidx(j & 1, 1, j >> 1) : mf_lidx(j & 1, 0, j >> 1);
            dst[mf_swz(qd, ln)] = make_float4(e * d[4 * j], e * d[4 * j + 1], e * d[4 * j + 2], e * d[4 * j + 3]);
        }
    }
}

// Operands of the epilogue of one owned node (loaded a layer ahead).
struct MfNodeIn {
    float u[3], f[3], ad[3];
    unsigned msk;
    bool ok;
};

template <int MODE>
__device__ __forceinline__ void mf_load_node(const OpArgsT<float> &a, const Tile &t, int gi, int j0, int k0,
                                             MfNodeIn &in)
{
    const Grid &g = a.g;
    const int own = threadIdx.x;
    in.ok = false;
    if (own >= t.TJ * t.TK || gi >= g.nnx) return;
    const int tk = own / t.TJ, tj = own - tk * t.TJ;
    const int gj = j0 + tj, gk = k0 + tk;
    if (gj >= g.nny || gk >= g.nnz) return;
    in.ok = true;
    const long nod = (long)g.nny * g.nnz * gi + (long)g.nny * gk + gj;
    const long dof = 3 * nod;
    in.msk = a.mask[nod];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (MODE != OP_MATVEC) in.f[c] = a.F[dof + c];
        if (MODE == OP_JACOBI || MODE == OP_PRE_JACOBI || MODE == OP_PRE_DEFECT) in.ad[c] = a.AD[dof + c];
        if (MODE == OP_PRE_JACOBI || MODE == OP_PRE_DEFECT)
            in.u[c] = (float)a.prerelax * in.f[c] / in.ad[c];
        else if (MODE != OP_DEFECT)
            in.u[c] = a.U[dof + c];
    }
}

template <int FL, int MODE>
__global__ __launch_bounds__(256, 3) void k_mfma_op(OpArgsT<float> a, Tile t)
{
    if (a.ctl && a.ctl[0]) return;
    constexpr int NP = mf_parts(FL);
    extern __shared__ __align__(16) unsigned char smem[];
    const Grid &g = a.g;
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
    const MfGeom m = mf_geom(t);
    float4 *T = reinterpret_cast<float4 *>(smem);                     // [2][NEL+1][8]
    uint2 *su = reinterpret_cast<uint2 *>(T + 2 * (m.NEL + 1) * 8);  // [3][NP][NPL]
    const int lane = threadIdx.x & 63;
    const int n = lane & 31, h = lane >> 5;

    // A fragments: lane row r = n (output row 8j + 4hm + cm), K step s,
    // columns 16s + 8h + 4yy + c -> Ke column 3*lidx(h, yy, s) + c
    bf16x8 Ka[3][2];
    {
        const int jm = n >> 3, hm = (n >> 2) & 1, cm = n & 3;
        const int lnr = mf_lidx(jm & 1, hm, jm >> 1);
#pragma unroll
        for (int st = 0; st < 2; ++st)
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int yy = i >> 2, c = i & 3;
                const int lnc = mf_lidx(h, yy, st);
                const float v = (cm < 3 && c < 3) ? a.Ke[24 * (3 * lnr + cm) + 3 * lnc + c] : 0.0f;
                float pv[3] = {0.0f, 0.0f, 0.0f};
                mf_split<NP>(v, pv);
#pragma unroll
                for (int p = 0; p < 3; ++p) Ka[p][st][i] = (__bf16)pv[p];
            }
    }
    const int nlay = min(t.TI, g.nnx - i0);
    MfBatch B;
    mf_batches(g, m, j0, k0, B);

    // prologue: planes i0-1, i0, i0+1 into slots 0, 1, 2; element layer i0-1.
    // Software pipeline of the layer loop: plane q = tt + 3 is loaded two
    // layers ahead (at layer tt - 1) and stored at the end of layer tt; the
    // epilogue operands of node layer tt and the element scales of element
    // layer tt are loaded one layer ahead.
    float En[kMfBatchesPerWave];
    float pv[2][3];  // plane loaded one layer earlier, stored at the end of this layer
    MfNodeIn in;
    {
        float E0[kMfBatchesPerWave];
        mf_load_E(a, i0 - 1, B, E0);
        float vals[2][3];
#pragma unroll 1
        for (int q = 0; q < 3; ++q) {
            mf_plane_load<MODE>(a, m, i0 - 1 + q, j0, k0, vals);
            mf_plane_store<NP>(su + q * NP * m.NPL, m, vals);
        }
        mf_load_E(a, i0, B, En);
        mf_plane_load<MODE>(a, m, i0 + 2, j0, k0, pv);
        mf_load_node<MODE>(a, t, i0, j0, k0, in);
        __syncthreads();
        mf_layer<FL>(m, su, su + NP * m.NPL, T, Ka, B, E0);
    }

    double part = 0.0;
    const float om = (float)a.omega;
    for (int tt = 0; tt < nlay; ++tt) {
        const int gi = i0 + tt;
        // loads for the next layer first
        float pn[2][3];
        mf_plane_load<MODE>(a, m, gi + 3, j0, k0, pn);
        MfNodeIn inn;
        mf_load_node<MODE>(a, t, gi + 1, j0, k0, inn);
        float Enn[kMfBatchesPerWave];
        mf_load_E(a, gi + 1, B, Enn);
        mf_layer<FL>(m, su + ((tt + 1) % 3) * NP * m.NPL, su + ((tt + 2) % 3) * NP * m.NPL,
                     T + ((tt + 1) & 1) * (m.NEL + 1) * 8, Ka, B, En);
        __syncthreads();
        // node layer gi: element layers gi-1 (node at x = 1) and gi (x = 0)
        if (in.ok) {
            const int own = threadIdx.x;
            const int tk = own / t.TJ, tj = own - tk * t.TJ;
            float y0 = 0.0f, y1 = 0.0f, y2 = 0.0f;
#pragma unroll
            for (int x = 1; x >= 0; --x) {
                const float4 *Tl = T + ((x ? tt : tt + 1) & 1) * (m.NEL + 1) * 8;
#pragma unroll
                for (int z = 1; z >= 0; --z)
#pragma unroll
                    for (int y = 1; y >= 0; --y) {
                        // element with first node at halo (tj + 1 - y, tk + 1 - z)
                        const int q = (tk + 1 - z) * m.EJ + (tj + 1 - y);
                        const float4 v = Tl[q * 8 + mf_swz(q, mf_lidx(x, y, z))];
                        y0 += v.x;
                        y1 += v.y;
                        y2 += v.z;
                    }
            }
            const long dof = 3 * ((long)g.nny * g.nnz * gi + (long)g.nny * (k0 + tk) + (j0 + tj));
            const float yv[3] = {y0, y1, y2};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const bool fixed = (in.msk >> c) & 1u;
                if (MODE == OP_MATVEC) {
                    const float v = fixed ? 0.0f : yv[c];
                    a.out[dof + c] = v;
                    part += (double)in.u[c] * (double)v;
                } else if (MODE == OP_DEFECT || MODE == OP_PRE_DEFECT) {
                    const float rv = fixed ? 0.0f : in.f[c] - yv[c];
                    a.out[dof + c] = rv;
                    if (MODE == OP_PRE_DEFECT) a.out2[dof + c] = in.u[c];
                } else {
                    const float f = in.f[c], u = in.u[c], ad = in.ad[c];
                    const float rv = fixed ? 0.0f : f - yv[c];
                    // U = (1-w) U + w (R + AD U)/AD  (mgcg1.cu:1199-1206)
                    const float un = (1.0f - om) * u + om * (rv + ad * u) / ad;
                    a.out[dof + c] = un;
                    part += (double)f * (double)un;
                }
            }
        }
        mf_plane_store<NP>(su + (tt % 3) * NP * m.NPL, m, pv);  // plane gi + 2
#pragma unroll
        for (int kb = 0; kb < kMfBatchesPerWave; ++kb) En[kb] = Enn[kb];
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) pv[r][c] = pn[r][c];
        in = inn;
        __syncthreads();
    }
    if (a.partials) {
        const double sm = block_sum(part);
        if (threadIdx.x == 0) a.partials[blockIdx.x] = sm;
    }
}

static int mf_balance(int n, int tmax)
{
    if (tmax >= n) return n;
    const int nt = (n + tmax - 1) / tmax;
    return (n + nt - 1) / nt;
}

template <int NP>
static size_t mf_lds_bytes_t(const Tile &t)
{
    const size_t NEL = (size_t)(t.TJ + 1) * (t.TK + 1), NPL = (size_t)(t.TJ + 2) * (t.TK + 2);
    return 2 * (NEL + 1) * 8 * sizeof(float4) + 3 * NP * NPL * sizeof(uint2);
}

constexpr size_t kMfLdsMax = 52 * 1024;  // three blocks per CU

static Tile mfma_tile(const Grid &g)
{
    Tile t;
    // 20 x 5 node columns: with two BF16 parts the block needs 39.9 KB of
    // LDS, so four blocks fit a CU (measured at 192x96x96: 55 us against
    // 62 us for 20 x 6 at three blocks per CU)
    t.TJ = mf_balance(g.nny, 20);
    t.TK = mf_balance(g.nnz, 5);
    t.TI = 8;
    // LDS, one owned node per thread, at most kMfBatchesPerWave batches per wave
    while (mf_lds_bytes_t<3>(t) > kMfLdsMax || t.TJ * t.TK > 256 ||
           ((t.TJ + 1) * (t.TK + 1) + 31) / 32 > 4 * kMfBatchesPerWave) {
        if (t.TK > 1)
            t.TK--;
        else
            t.TJ--;
    }
    t.ntj = (g.nny + t.TJ - 1) / t.TJ;
    t.ntk = (g.nnz + t.TK - 1) / t.TK;
    t.nti = (g.nnx + t.TI - 1) / t.TI;
    // a halo plane is staged two nodes per thread
    if ((t.TJ + 2) * (t.TK + 2) > 512) throw std::runtime_error("MFMA level-0 tile: halo plane too large");
    return t;
}

template <int FL>
static void launch_flavor(hipStream_t s, int mode, const OpArgsT<float> &a, const Tile &t)
{
    const dim3 grid((unsigned)xcd_launch(t.blocks())), block(256);
    const size_t lds = mf_lds_bytes_t<mf_parts(FL)>(t);
    static bool attr = false;
    if (!attr) {  // dynamic LDS above 64 KB
        (void)hipFuncSetAttribute((const void *)k_mfma_op<FL, OP_MATVEC>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMfLdsMax);
        (void)hipFuncSetAttribute((const void *)k_mfma_op<FL, OP_DEFECT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMfLdsMax);
        (void)hipFuncSetAttribute((const void *)k_mfma_op<FL, OP_JACOBI>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMfLdsMax);
        (void)hipFuncSetAttribute((const void *)k_mfma_op<FL, OP_PRE_JACOBI>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMfLdsMax);
        (void)hipFuncSetAttribute((const void *)k_mfma_op<FL, OP_PRE_DEFECT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMfLdsMax);
        attr = true;
    }
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
                                float *D4, const float *A32, const float *B32, float *D32)
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
    // 32x32x16: lane holds A[l%32][8(l/32)+i], B[8(l/32)+i][l%32] and
    // D[8j + 4(l/32) + r][l%32] in accumulator 4j + r
    const int l32 = lane & 31, h = lane >> 5;
    bf16x8 ga, gb;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        ga[i] = (__bf16)A32[16 * l32 + 8 * h + i];   // A: 32 x 16
        gb[i] = (__bf16)B32[32 * (8 * h + i) + l32];  // B: 16 x 32
    }
    f32x16 d32 = {};
    d32 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ga, gb, d32, 0, 0, 0);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) D32[32 * (8 * j + 4 * h + r) + l32] = d32[4 * j + r];
}

}  // namespace gputo

extern "C" int gputo_selftest_mfma(double *max_err)
{
    using namespace gputo;
    if (!max_err) return GPUTO_ERR_ARGUMENT;
    std::vector<float> A(16 * 32), B(32 * 16), R(256, 0.f), A4(16 * 4), B4(4 * 16), R4(256, 0.f);
    std::vector<float> A32(32 * 16), B32(16 * 32), R32(1024, 0.f);
    for (int i = 0; i < 16 * 32; ++i) A[i] = (float)((i * 7 + 3) % 11 - 5);
    for (int i = 0; i < 32 * 16; ++i) B[i] = (float)((i * 5 + 1) % 9 - 4);
    for (int i = 0; i < 64; ++i) {
        A4[i] = (float)((i * 3 + 2) % 7 - 3);
        B4[i] = (float)((i * 11 + 5) % 13 - 6);
    }
    for (int i = 0; i < 512; ++i) {
        A32[i] = (float)((i * 13 + 4) % 15 - 7);
        B32[i] = (float)((i * 9 + 2) % 17 - 8);
    }
    for (int r = 0; r < 16; ++r)
        for (int c = 0; c < 16; ++c) {
            for (int k = 0; k < 32; ++k) R[16 * r + c] += A[32 * r + k] * B[16 * k + c];
            for (int k = 0; k < 4; ++k) R4[16 * r + c] += A4[4 * r + k] * B4[16 * k + c];
        }
    for (int r = 0; r < 32; ++r)
        for (int c = 0; c < 32; ++c)
            for (int k = 0; k < 16; ++k) R32[32 * r + c] += A32[16 * r + k] * B32[32 * k + c];
    float *dev = nullptr;
    const size_t n = A.size() + B.size() + 256 + A4.size() + B4.size() + 256 + A32.size() + B32.size() + 1024;
    if (hipMalloc(&dev, n * sizeof(float)) != hipSuccess) return GPUTO_ERR_DEVICE;
    float *dA = dev, *dB = dA + A.size(), *dD = dB + B.size(), *dA4 = dD + 256, *dB4 = dA4 + A4.size(),
          *dD4 = dB4 + B4.size(), *dA32 = dD4 + 256, *dB32 = dA32 + A32.size(), *dD32 = dB32 + B32.size();
    hipError_t e = hipMemcpy(dA, A.data(), A.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dB, B.data(), B.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dA4, A4.data(), A4.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dB4, B4.data(), B4.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dA32, A32.data(), A32.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dB32, B32.data(), B32.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        k_mfma_selftest<<<1, 64>>>(dA, dB, dD, dA4, dB4, dD4, dA32, dB32, dD32);
        e = hipGetLastError();
    }
    std::vector<float> D(256), D4(256), D32(1024);
    if (e == hipSuccess) e = hipMemcpy(D.data(), dD, 256 * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(D4.data(), dD4, 256 * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(D32.data(), dD32, 1024 * 4, hipMemcpyDeviceToHost);
    (void)hipFree(dev);
    if (e != hipSuccess) return GPUTO_ERR_DEVICE;
    double err = 0.0;
    for (int i = 0; i < 256; ++i) {
        err = std::max(err, (double)std::fabs(D[i] - R[i]));
        err = std::max(err, (double)std::fabs(D4[i] - R4[i]));
    }
    for (int i = 0; i < 1024; ++i) err = std::max(err, (double)std::fabs(D32[i] - R32[i]));
    *max_err = err;
    return 0;
}
