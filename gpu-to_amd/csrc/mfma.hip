// Level-0 operator of the FP32 V-cycle on the matrix cores.
//
// Element form: the shared element matrix Ke (24 x 24, rows and columns
// padded to 32) multiplies the 24 nodal values of 32 elements at a time
// with v_mfma_f32_32x32x16_bf16 (FP32 accumulation); each result column is
// scaled by its element's E and added into the sums of the element's 8
// nodes.  FP32 operands are split into BF16 parts x = hi + mid + lo; the
// flavours differ in the products they keep:
//   MF_F32    all nine products (hi/mid/lo x hi/mid/lo)
//   MF_BF16X3 the six products of order >= 2^-16 (FP32-level accuracy)
//   MF_BF16   hi*hi + hi*lo + lo*hi on a two-way split
//   MF_BF16X1 hi*hi only
//
// Work of one wave: a strip of 31 nodes along j (element columns j0-1 ..
// j0+30 on the 32 MFMA columns), KR node rows along k and TI node layers
// along i, streamed layer by layer.  Lane (h, c), h = lane / 32, holds for
// element column c the results of the element's four nodes with y offset
// h: the (x, z) offsets.  The z offsets of consecutive element rows and the
// x offsets of consecutive element layers meet in the lane's registers; the
// two y halves of a node meet by one lane permutation per value.  The
// operands come straight from global memory (a node is read by two lanes
// per layer, the repeats hit L1/L2), so the kernel has no LDS staging and
// no barriers, and the waves of a block are independent.  No atomics; every
// node's sum has one fixed order, so results are bitwise deterministic.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <stdexcept>
#include <vector>

#include "mixed.h"

namespace gputo {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

enum MfFlavor { MF_F32 = 0, MF_BF16 = 1, MF_BF16X3 = 2, MF_BF16X1 = 3 };

constexpr int kMfStrip = 31;  // nodes per wave along j (32 element columns)
constexpr int kMfWaves = 4;   // independent waves per block

// Work decomposition: strips of kMfStrip nodes along j, KR node rows along
// k, TI node layers along i; one unit per wave, strips fastest.
struct MfWork {
    int KR, TI;
    int ns, nkb, nib;
    __host__ __device__ long units() const { return (long)ns * nkb * nib; }
    __host__ __device__ long blocks() const { return (units() + kMfWaves - 1) / kMfWaves; }
};

struct MfmaState {
    MfWork w;
};

// local node number of offset (x, y, z) = (i, j, k) in {0,1}^3 in the element
// order ie={0,1,1,0,0,1,1,0}, je={0,0,1,1,0,0,1,1}, ke={0,0,0,0,1,1,1,1}
__host__ __device__ constexpr int mf_lidx(int x, int y, int z) { return 4 * z + (y ? (x ? 2 : 3) : (x ? 1 : 0)); }

// number of BF16 parts per operand for a flavour
__host__ __device__ constexpr int mf_parts(int fl) { return fl == MF_BF16X1 ? 1 : (fl == MF_BF16 ? 2 : 3); }

// the NP BF16 parts of v: part 0 = bf16(v), part 1 = bf16(v - part0), ...
template <int NP>
__device__ __forceinline__ void mf_split(float v, float out[NP])
{
    float r = v;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const __bf16 hb = (__bf16)r;
        out[p] = (float)hb;
        r -= out[p];
    }
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// (a, b) rounded to BF16 and packed, a in the low half
__device__ __forceinline__ unsigned mf_pk(float a, float b)
{
    return __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){a, b}, bf16x2));
}
__device__ __forceinline__ float mf_lo16(unsigned w) { return __builtin_bit_cast(float, w << 16); }
__device__ __forceinline__ float mf_hi16(unsigned w) { return __builtin_bit_cast(float, w & 0xffff0000u); }

// B operand fragments of one node row: the 6 values of the lane's node pair
// (y = 0: v[0..2], y = 1: v[3..5]) as NP BF16 parts, slots (3 components,
// zero) per node; part p + 1 splits what parts 0..p leave.
template <int NP>
__device__ __forceinline__ void mf_frag(const float (&v)[6], bf16x8 (&f)[NP])
{
    float r0 = v[0], r1 = v[1], r2 = v[2], r3 = v[3], r4 = v[4], r5 = v[5];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const unsigned w0 = mf_pk(r0, r1), w1 = mf_pk(r2, 0.0f), w2 = mf_pk(r3, r4), w3 = mf_pk(r5, 0.0f);
        const u32x4 w = {w0, w1, w2, w3};
        f[p] = __builtin_bit_cast(bf16x8, w);
        if (p + 1 < NP) {
            r0 -= mf_lo16(w0);
            r1 -= mf_hi16(w0);
            r2 -= mf_lo16(w1);
            r3 -= mf_lo16(w2);
            r4 -= mf_hi16(w2);
            r5 -= mf_lo16(w3);
        }
    }
}

// Buffer resources: loads past the end (and, with the offset taken as
// unsigned, before the start) return 0, stores there are dropped.
typedef __amdgpu_buffer_rsrc_t MfBuf;
__device__ __forceinline__ MfBuf mf_buf(const void *p, long bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float mf_ld(MfBuf b, int off, int soff = 0)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(b, off, soff, 0));
}
// Three consecutive floats as three dword loads.  (The vector-valued load
// builtins are read through a bit cast of the whole vector: a bit cast of
// one component of their result came out wrong in this toolchain.)
__device__ __forceinline__ void mf_ld3(MfBuf b, int off, float *v)
{
#pragma unroll
    for (int i = 0; i < 3; ++i) v[i] = mf_ld(b, off, 4 * i);
}
// Six consecutive floats: one 16-byte and one 8-byte load.
__device__ __forceinline__ void mf_ld6(MfBuf b, int off, float (&v)[6])
{
    const f32x4 x = __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(b, off, 0, 0));
    const f32x2 y = __builtin_bit_cast(f32x2, __builtin_amdgcn_raw_buffer_load_b64(b, off + 16, 0, 0));
    v[0] = x[0];
    v[1] = x[1];
    v[2] = x[2];
    v[3] = x[3];
    v[4] = y[0];
    v[5] = y[1];
}
constexpr int kMfNoStore = 0x7ffffff0;  // a store offset past every buffer: dropped

// Buffers of one launch.
struct MfBufs {
    MfBuf U, F, AD, CX, M, O, O2;
};

// The node pair (je, je+1) of node row kr in node plane pi as a lane's 6
// operand values: U, or prerelax*F/AD in the prerelaxed modes (preRelc,
// mgcg1.cu:1190-1197).  Pairs outside the mesh only feed elements outside
// it, whose scale is 0: the loads need no bounds test (they read other
// nodes' values, or 0 past the ends of the vector).
template <int MODE>
__device__ __forceinline__ void mf_row(const OpArgsT<float> &a, const MfBufs &B, int pi, int kr, int je, float (&v)[6])
{
    const Grid &g = a.g;
    const int off = 12 * (g.nny * g.nnz * pi + g.nny * kr + je);
    if (MODE == OP_PRE_JACOBI || MODE == OP_PRE_DEFECT) {
        float f[6], ad[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            f[i] = mf_ld(B.F, off, 4 * i);
            ad[i] = mf_ld(B.AD, off, 4 * i);
        }
#pragma unroll
        for (int i = 0; i < 6; ++i) v[i] = ad[i] != 0.0f ? (float)a.prerelax * f[i] / ad[i] : 0.0f;
    } else {
        mf_ld6(B.U, off, v);
    }
}

// scale * E of element (ie, je, ke), zero outside the mesh (jin: je inside)
__device__ __forceinline__ float mf_E(const OpArgsT<float> &a, const MfBufs &B, int ie, int ke, int je, bool jin,
                                      float scale)
{
    const Grid &g = a.g;
    const bool in = jin && ie >= 0 && ie < g.nx && ke >= 0 && ke < g.nz;
    const float v = mf_ld(B.CX, 4 * (g.ny * g.nz * ie + g.ny * ke + je));
    return in ? v * scale : 0.0f;
}

// Operands of the epilogue of one output node.
struct MfOut {
    float u[3], f[3], ad[3];
    unsigned msk;
    int off;  // byte offset of the node's first DOF, or kMfNoStore
};

// (ni is inside the mesh; `want` says whether the lane owns column nj)
template <int MODE>
__device__ __forceinline__ void mf_out_load(const OpArgsT<float> &a, const MfBufs &B, int ni, int nk, int nj,
                                            bool want, MfOut &o)
{
    const Grid &g = a.g;
    const bool ok = want && nk < g.nnz && nj < g.nny;
    const int node = g.nny * g.nnz * ni + g.nny * nk + nj;
    const int off = 12 * node;
    o.off = ok ? off : kMfNoStore;
    o.msk = __builtin_amdgcn_raw_buffer_load_b8(B.M, node, 0, 0);
#pragma unroll
    for (int c = 0; c < 3; ++c) o.u[c] = o.f[c] = o.ad[c] = 0.0f;
    if (MODE != OP_MATVEC) mf_ld3(B.F, off, o.f);
    if (MODE == OP_JACOBI || MODE == OP_PRE_JACOBI || MODE == OP_PRE_DEFECT) mf_ld3(B.AD, off, o.ad);
    if (MODE == OP_PRE_JACOBI || MODE == OP_PRE_DEFECT) {
#pragma unroll
        for (int c = 0; c < 3; ++c) o.u[c] = o.ad[c] != 0.0f ? (float)a.prerelax * o.f[c] / o.ad[c] : 0.0f;
    } else if (MODE == OP_JACOBI) {
        mf_ld3(B.U, off, o.u);
    }
}

// The operator epilogue of one node from y = (K u) at the node
// (multAV0c / defect0c / Jacobc, mgcg1.cu:941-997,1199-1264; ZeroBC fused).
// Lanes that own no node store past the end of the buffer (dropped).
template <int MODE>
__device__ __forceinline__ void mf_out_store(const OpArgsT<float> &a, const MfBufs &B, const MfOut &o,
                                             const float (&y)[3], float om)
{
    float r[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const bool fixed = (o.msk >> c) & 1u;
        if (MODE == OP_MATVEC) {
            r[c] = fixed ? 0.0f : y[c];
        } else if (MODE == OP_DEFECT || MODE == OP_PRE_DEFECT) {
            r[c] = fixed ? 0.0f : o.f[c] - y[c];
        } else {
            const float f = o.f[c], u = o.u[c], ad = o.ad[c];
            const float rv = fixed ? 0.0f : f - y[c];
            // U = (1-w) U + w (R + AD U)/AD  (mgcg1.cu:1199-1206)
            r[c] = (1.0f - om) * u + om * (rv + ad * u) * __builtin_amdgcn_rcpf(ad);
        }
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, r[c]), B.O, o.off, 4 * c, 0);
        if (MODE == OP_PRE_DEFECT)
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, o.u[c]), B.O2, o.off, 4 * c, 0);
    }
}

// A operand parts: the BF16 split of Ke, or with KX (Ke = S Kint, Kint
// integer and exact in BF16) Kint alone.
template <int FL, bool KX>
struct MfA {
    static constexpr int n = KX ? 1 : mf_parts(FL);
};

// D = Ke u_e (KX: Kint u_e) for the 32 elements of one element row: K step
// 0 is the z = 0 node row (fragments f0), step 1 the z = 1 row (f1); one
// accumulation chain, smallest products first.
template <int FL, bool KX>
__device__ __forceinline__ f32x16 mf_chain(const bf16x8 (&Ka)[MfA<FL, KX>::n][2], const bf16x8 (&f0)[mf_parts(FL)],
                                           const bf16x8 (&f1)[mf_parts(FL)])
{
    constexpr int NP = mf_parts(FL);
    f32x16 d = {};
    if constexpr (KX) {
        // exact A: every product of the split of u, lo parts first
#pragma unroll
        for (int p = NP - 1; p >= 0; --p)
#pragma unroll
            for (int st = 0; st < 2; ++st)
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[0][st], st ? f1[p] : f0[p], d, 0, 0, 0);
        return d;
    } else {
#pragma unroll
        for (int st = 0; st < 2; ++st) {
            const bf16x8(&B)[NP] = st ? f1 : f0;
            if constexpr (FL == MF_F32) {
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[2][st], B[2], d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[2][st], B[1], d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[1][st], B[2], d, 0, 0, 0);
            }
            if constexpr (FL == MF_F32 || FL == MF_BF16X3) {
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[2][st], B[0], d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[1][st], B[1], d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[0][st], B[2], d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[1][st], B[0], d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[0][st], B[1], d, 0, 0, 0);
            }
            if constexpr (FL == MF_BF16) {
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[1][st], B[0], d, 0, 0, 0);
                d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[0][st], B[1], d, 0, 0, 0);
            }
        }
#pragma unroll
        for (int st = 0; st < 2; ++st)
            d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ka[0][st], st ? f1[0] : f0[0], d, 0, 0, 0);
        return d;
    }
}

// Operands of one element layer: the lane's node rows k0-1 .. k0+KR in node
// plane layer + h, and the (scaled) element scales of element rows
// k0-1 .. k0+KR-1.
template <int KR>
struct MfLayer {
    float u[KR + 2][6];
    float e[KR + 1];
};

template <int MODE, int KR>
__device__ __forceinline__ void mf_load_layer(const OpArgsT<float> &a, const MfBufs &B, int ie, int k0, int je,
                                              bool jin, int h, float scale, MfLayer<KR> &L)
{
#pragma unroll
    for (int t = 0; t < KR + 2; ++t) mf_row<MODE>(a, B, ie + h, k0 - 1 + t, je, L.u[t]);
#pragma unroll
    for (int q = 0; q <= KR; ++q) L.e[q] = mf_E(a, B, ie, k0 - 1 + q, je, jin, scale);
}

// Element layer ie: products of its KR+1 element rows added into the node
// sums of node layer ie (cur: x = 0 offsets; it already holds the x = 1
// offsets of layer ie-1) and started for node layer ie+1 (nxt: x = 1
// offsets; each row's first product is assigned, so nxt needs no clearing).
// Accumulator 4j + r holds node group j = (x, z) = (j & 1, j >> 1),
// component r (r = 3 is the zero pad row).
template <int FL, int KR, bool KX>
__device__ __forceinline__ void mf_products(const bf16x8 (&Ka)[MfA<FL, KX>::n][2], const MfLayer<KR> &L,
                                            float (&cur)[KR][3], float (&nxt)[KR][3])
{
    constexpr int NP = mf_parts(FL);
    bf16x8 f0[NP], f1[NP];
    mf_frag<NP>(L.u[0], f0);
#pragma unroll
    for (int q = 0; q <= KR; ++q) {
        mf_frag<NP>(L.u[q + 1], f1);
        const f32x16 d = mf_chain<FL, KX>(Ka, f0, f1);
        const float e = L.e[q];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (q >= 1) {  // z = 0 offsets: node row k0-1+q
                cur[q - 1][c] += e * d[c];
                nxt[q - 1][c] += e * d[4 + c];
            }
            if (q < KR) {  // z = 1 offsets: node row k0+q
                cur[q][c] += e * d[8 + c];
                nxt[q][c] = e * d[12 + c];
            }
        }
#pragma unroll
        for (int p = 0; p < NP; ++p) f0[p] = f1[p];
    }
}

template <int FL, int MODE, int KR, bool KX>
__global__ __launch_bounds__(256) void k_mfma_op(OpArgsT<float> a, MfWork w, float kscale)
{
    constexpr int NA = MfA<FL, KX>::n;
    constexpr int KH = KR / 2;  // output rows per lane half
    const Grid &g = a.g;
    const long nb = w.blocks();
    const long tb = xcd_tile(blockIdx.x, nb);
    const int lane = threadIdx.x & 63, c = lane & 31, h = lane >> 5;
    const long unit = tb * kMfWaves + (threadIdx.x >> 6);
    if (tb < nb && unit < w.units()) {  // wave-uniform
        // 32-bit: far fewer than 2^32 wave units
        unsigned un = (unsigned)unit;
        const int s = (int)(un % (unsigned)w.ns);
        un /= (unsigned)w.ns;
        const int kb = (int)(un % (unsigned)w.nkb);
        const int ib = (int)(un / (unsigned)w.nkb);
        const int j0 = kMfStrip * s, k0 = KR * kb, i0 = w.TI * ib;
        const int je = j0 - 1 + c;  // element column of this lane
        const bool jin = je >= 0 && je < g.ny;
        const float scale = KX ? kscale : 1.0f;
        const float kinv = KX ? 1.0f / kscale : 1.0f;
        const long vb = 4 * g.ndof;
        MfBufs B;
        B.U = mf_buf(a.U, a.U ? vb : 0);
        B.F = mf_buf(a.F, a.F ? vb : 0);
        B.AD = mf_buf(a.AD, a.AD ? vb : 0);
        B.CX = mf_buf(a.CX, 4 * g.nel);
        B.M = mf_buf(a.mask, g.nnod);
        B.O = mf_buf(a.out, vb);
        B.O2 = mf_buf(a.out2, a.out2 ? vb : 0);
        // the first element layer's operands are in flight while Ke is read
        // and split, and while the skip flag is read
        MfLayer<KR> LA, LB;
        mf_load_layer<MODE, KR>(a, B, i0 - 1, k0, je, jin, h, scale, LA);

        // A fragments: lane row r = c (output row 8j + 4hm + cm), K step st,
        // columns 16 st + 8 h + 4 yy + cc -> Ke column 3 lidx(h, yy, st) + cc
        bf16x8 Ka[NA][2];
        {
            const int jm = c >> 3, hm = (c >> 2) & 1, cm = c & 3;
            const int lnr = mf_lidx(jm & 1, hm, jm >> 1);
            float kev[2][8];
#pragma unroll
            for (int st = 0; st < 2; ++st)
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int yy = i >> 2, cc = i & 3;
                    const int lnc = mf_lidx(h, yy, st);
                    const float v = a.Ke[24 * (3 * lnr + min(cm, 2)) + 3 * lnc + min(cc, 2)];
                    kev[st][i] = (cm < 3 && cc < 3) ? v : 0.0f;
                }
            if (a.ctl && a.ctl[0]) return;  // converged: read with the loads above
#pragma unroll
            for (int st = 0; st < 2; ++st)
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const float v = kev[st][i];
                    if constexpr (KX) {
                        // v / kscale is within rounding of an integer of at most 8 bits, so
                        // the product with the rounded reciprocal rounds to the same one
                        Ka[0][st][i] = (__bf16)rintf(v * kinv);
                    } else {
                        float pv[NA];
                        mf_split<NA>(v, pv);
#pragma unroll
                        for (int p = 0; p < NA; ++p) Ka[p][st][i] = (__bf16)pv[p];
                    }
                }
        }

        // output node of the lane: the low half owns nodes j0 + c - 1 of node
        // rows k0 .. k0+KH-1, the high half nodes j0 + c of rows k0+KH ..
        const int jo = h ? j0 + c : j0 + c - 1;
        const bool jok = h ? (c < kMfStrip) : (c >= 1);
        const int src = h ? lane - 31 : lane + 31;  // the lane with the other y half
        const float om = (float)a.omega;
        const int nlay = min(w.TI, g.nnx - i0);

        float accA[KR][3], accB[KR][3];
#pragma unroll
        for (int r = 0; r < KR; ++r)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) accA[r][cc] = accB[r][cc] = 0.0f;

        // element layer ie = i0 - 1 + tt; after it node layer ie is complete.
        // Software pipeline: the operands of element layer ie + 1 are loaded
        // while layer ie computes (two register sets, swapped every layer);
        // the epilogue operands of node layer ie are loaded at the start of
        // layer ie, whose products cover their latency.
        auto step = [&](int tt, float (&cur)[KR][3], float (&nxt)[KR][3], const MfLayer<KR> &L, MfLayer<KR> &N) {
            const int ie = i0 - 1 + tt;
            // issue order matters: loads complete in order for the wait
            // counter, so the epilogue operands go before the prefetch
            MfOut O[KH];
            if (tt >= 1) {
#pragma unroll
                for (int rr = 0; rr < KH; ++rr) mf_out_load<MODE>(a, B, ie, k0 + rr + h * KH, jo, jok, O[rr]);
            }
            if (tt < nlay) mf_load_layer<MODE, KR>(a, B, ie + 1, k0, je, jin, h, scale, N);
            mf_products<FL, KR, KX>(Ka, L, cur, nxt);
            if (tt >= 1) {
                float y[KH][3];
#pragma unroll
                for (int r = 0; r < KR; ++r)
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) {
                        const float full = cur[r][cc] + __shfl(cur[r][cc], src, 64);
                        if (r < KH) {
                            if (!h) y[r][cc] = full;
                        } else {
                            if (h) y[r - KH][cc] = full;
                        }
                    }
#pragma unroll
                for (int rr = 0; rr < KH; ++rr) mf_out_store<MODE>(a, B, O[rr], y[rr], om);
            }
        };
        for (int tt = 0; tt <= nlay; tt += 2) {
            step(tt, accA, accB, LA, LB);
            if (tt + 1 <= nlay) step(tt + 1, accB, accA, LB, LA);
        }
    }
}

// S with Ke = S Kint, Kint integer with |entries| <= 256 (exact in BF16),
// from the entry classes of a unit-H8 Ke0; 0 when there is none.  For the
// top3D125 Ke0 (nu = 0.3), Kint = 1440 (1+nu)(2nu-1) Ke0 has the entries
// -176 .. 80 (top3D125mod.m:50-76 in integer form).
static float ke_int_scale(const OpArgsT<float> &a)
{
    if (!a.kc) return 0.0f;
    float vmin = 0.0f;
    for (int k = 0; k < kKeMaxClasses; ++k)
        if (a.kv[k] != 0.0f && (vmin == 0.0f || std::fabs(a.kv[k]) < vmin)) vmin = std::fabs(a.kv[k]);
    if (vmin == 0.0f) return 0.0f;
    for (int m = 1; m <= 256; ++m) {
        const double S = (double)vmin / m;
        bool ok = true;
        for (int k = 0; k < kKeMaxClasses && ok; ++k) {
            const double q = a.kv[k] / S, r = std::nearbyint(q);
            ok = std::fabs(r) <= 256.0 && std::fabs(q - r) <= 1e-4;
        }
        if (ok) return (float)S;
    }
    return 0.0f;
}

// Tiling: KR node rows per wave (a template argument) and TI node layers.
// A wave streams its TI layers one after another, so the grid is sized to
// be resident at once: TI is the smallest layer count whose units fit the
// waves the device holds of the kernel (its occupancy times the CUs; one
// round, no tail).  GPUTO_MF_TILE="KR,TI" overrides it (KR 4 or 6).
static MfWork mfma_work(const Grid &g)
{
    MfWork w;
    w.KR = 4;
    w.TI = 0;
    if (const char *e = std::getenv("GPUTO_MF_TILE")) {
        int kr = 0, ti = 0;
        if (std::sscanf(e, "%d,%d", &kr, &ti) == 2 && (kr == 4 || kr == 6) && ti >= 1) {
            w.KR = kr;
            w.TI = ti;
        }
    }
    w.ns = (g.nny + kMfStrip - 1) / kMfStrip;
    w.nkb = (g.nnz + w.KR - 1) / w.KR;
    if (w.TI == 0) {
        int dev = 0, ncu = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
            ncu = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mfma_op<MF_BF16, OP_JACOBI, 4, true>,
                                                         kMfWaves * 64, 0) != hipSuccess || per_cu < 1)
            per_cu = 2;
        const long waves = (long)per_cu * ncu * kMfWaves;
        const long nib = std::max(1L, waves / ((long)w.ns * w.nkb));
        w.TI = (int)((g.nnx + nib - 1) / nib);
    }
    w.nib = (g.nnx + w.TI - 1) / w.TI;
    return w;
}

template <int FL, int KR, bool KX>
static void launch_kr(hipStream_t s, int mode, const OpArgsT<float> &a, const MfWork &w, float S)
{
    const dim3 grid((unsigned)xcd_launch(w.blocks())), block(kMfWaves * 64);
    switch (mode) {
    case OP_MATVEC: k_mfma_op<FL, OP_MATVEC, KR, KX><<<grid, block, 0, s>>>(a, w, S); break;
    case OP_DEFECT: k_mfma_op<FL, OP_DEFECT, KR, KX><<<grid, block, 0, s>>>(a, w, S); break;
    case OP_JACOBI: k_mfma_op<FL, OP_JACOBI, KR, KX><<<grid, block, 0, s>>>(a, w, S); break;
    case OP_PRE_JACOBI: k_mfma_op<FL, OP_PRE_JACOBI, KR, KX><<<grid, block, 0, s>>>(a, w, S); break;
    case OP_PRE_DEFECT: k_mfma_op<FL, OP_PRE_DEFECT, KR, KX><<<grid, block, 0, s>>>(a, w, S); break;
    default: break;
    }
}

template <int FL, bool KX>
static void launch_fx(hipStream_t s, int mode, const OpArgsT<float> &a, const MfWork &w, float S)
{
    if (w.KR == 6)
        launch_kr<FL, 6, KX>(s, mode, a, w, S);
    else
        launch_kr<FL, 4, KX>(s, mode, a, w, S);
}

template <int FL>
static void launch_flavor(hipStream_t s, int mode, const OpArgsT<float> &a, const MfWork &w)
{
    const float S = ke_int_scale(a);
    if (S > 0.0f)
        launch_fx<FL, true>(s, mode, a, w, S);
    else
        launch_fx<FL, false>(s, mode, a, w, 1.0f);
}

void launch_op0_mfma(hipStream_t s, int mode, const OpArgsT<float> &a, const void *ctx, int flavor)
{
    if (a.partials) throw std::runtime_error("MFMA level-0 operator: no fused dot product");
    if (a.nb > 1) throw std::runtime_error("MFMA level-0 operator: one right-hand side per launch");
    const MixedState *m = static_cast<const MixedState *>(ctx);
    const MfWork &w = m->mf->w;
    if (flavor == MF_F32)
        launch_flavor<MF_F32>(s, mode, a, w);
    else if (flavor == MF_BF16X3)
        launch_flavor<MF_BF16X3>(s, mode, a, w);
    else if (flavor == MF_BF16X1)
        launch_flavor<MF_BF16X1>(s, mode, a, w);
    else
        launch_flavor<MF_BF16>(s, mode, a, w);
}

long mfma_tile_blocks(const Grid &g) { return xcd_launch(mfma_work(g).blocks()); }

void mfma_create(gputo_handle *h, MixedState *m)
{
    m->mf = new MfmaState();
    m->mf->w = mfma_work(h->L[0].g);
}

void mfma_destroy(gputo_handle *, MixedState *m)
{
    delete m->mf;
    m->mf = nullptr;
}

int mfma_blocks(const gputo_handle *h) { return (int)xcd_launch(h->mx->mf->w.blocks()); }

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
