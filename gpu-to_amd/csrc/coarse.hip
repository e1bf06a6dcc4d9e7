// Coarse levels of the MGCG hierarchy: level-1 rediscretisation from child
// densities, element-Galerkin operators, the coarsest-level penalty,
// dense assembly, blocked FP64 Cholesky, the explicit inverse of the
// factor and the two triangular products that apply it in every V-cycle.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "gputo_internal.h"

namespace gputo {

typedef double f64x4 __attribute__((ext_vector_type(4)));

__host__ __device__ constexpr int lidx2(int x, int y, int z) { return 4 * z + (y ? (x ? 2 : 3) : (x ? 1 : 0)); }
__host__ __device__ constexpr int pidx2(int m, int n) { return n <= m ? m * (m + 1) / 2 + n : n * (n + 1) / 2 + m; }
__device__ constexpr int kIE[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__device__ constexpr int kJE[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__device__ constexpr int kKE[8] = {0, 0, 0, 0, 1, 1, 1, 1};

// ---------------------------------------------------------------------------
// Level-1 rediscretisation (coarseAe2c, mgcg1.cu:2042-2070): the packed
// coarse element matrix is sum_na CX[child(na)] * Aelm[:, na] where Aelm
// holds the h=2 Gauss-point stiffness matrices and child na = 4 iP + 2 kP
// + jP.  One thread per coarse element; Aelm is read through the scalar
// cache (uniform addresses).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_coarse_l1(Grid gc, const double *__restrict__ Aelm,
                                                  const double *__restrict__ CXf, double *__restrict__ A)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= gc.nel) return;
    const int j = (int)(e % gc.ny), k = (int)((e / gc.ny) % gc.nz), i = (int)(e / ((long)gc.ny * gc.nz));
    const int fny = 2 * gc.ny, fnz = 2 * gc.nz;
    double cx[8];
#pragma unroll
    for (int iP = 0; iP < 2; ++iP)
#pragma unroll
        for (int kP = 0; kP < 2; ++kP)
#pragma unroll
            for (int jP = 0; jP < 2; ++jP)
                cx[4 * iP + 2 * kP + jP] = CXf[(long)fny * fnz * (2 * i + iP) + (long)fny * (2 * k + kP) + 2 * j + jP];
    for (int m = 0; m < 300; ++m) {
        double s = 0.0;
#pragma unroll
        for (int na = 0; na < 8; ++na) s += cx[na] * Aelm[8 * m + na];
        A[(long)m * gc.nel + e] = s;
    }
}

void launch_coarse_l1(hipStream_t s, const Grid &gc, const double *Aelm, const double *CXf, double *A)
{
    k_coarse_l1<<<(unsigned)((gc.nel + 255) / 256), 256, 0, s>>>(gc, Aelm, CXf, A);
}

// ---------------------------------------------------------------------------
// Element Galerkin (coarseAe3c, mgcg1.cu:1678-1743): A_c = P (sum of the 8
// children on the 27-node patch) P^T.  P is the trilinear interpolation of
// the coarse element's 8 nodes onto its 27 fine nodes, so the (a,b) node
// block of A_c is
//   sum_ch sum_{p,q} w(a, ch+p) w(b, ch+q) A_ch[p q block],
// with w(a, x) = prod_axis (1 - |2 a_axis - x_axis| / 2).  One 64-thread
// block per coarse element stages its 8 children (full 24x24 each) in LDS;
// thread (a >= b) forms one 3x3 block.  Children come from the packed
// matrices of the finer level, or are CX*Ke on level 0.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double trilin_w(int a, int x0, int x1, int x2)
{
    // coarse local node a (offsets kIE/kJE/kKE), fine patch node (x0,x1,x2) in {0,1,2}^3
    const double w0 = 1.0 - 0.5 * abs(2 * kIE[a] - x0);
    const double w1 = 1.0 - 0.5 * abs(2 * kJE[a] - x1);
    const double w2 = 1.0 - 0.5 * abs(2 * kKE[a] - x2);
    return w0 * w1 * w2;
}

__global__ __launch_bounds__(64) void k_galerkin(Grid gc, const double *__restrict__ Af, const double *__restrict__ Ke,
                                                const double *__restrict__ CXf, double *__restrict__ Ac)
{
    __shared__ double S[8][576];
    const long e = blockIdx.x;
    const int j = (int)(e % gc.ny), k = (int)((e / gc.ny) % gc.nz), i = (int)(e / ((long)gc.ny * gc.nz));
    const int fny = 2 * gc.ny, fnz = 2 * gc.nz;
    const long nelf = 8 * gc.nel;
    for (int ch = 0; ch < 8; ++ch) {
        const int ci = kIE[ch], cj = kJE[ch], ck = kKE[ch];
        const long eP = (long)fny * fnz * (2 * i + ci) + (long)fny * (2 * k + ck) + 2 * j + cj;
        const double cx = Af ? 0.0 : CXf[eP];
        for (int m = threadIdx.x; m < 576; m += 64) {
            const int r = m / 24, c = m - 24 * r;
            S[ch][m] = Af ? Af[(long)pidx2(r, c) * nelf + eP] : cx * Ke[m];
        }
    }
    __syncthreads();
    const int q = threadIdx.x;
    if (q >= 36) return;
    int a = 0;
    while ((a + 1) * (a + 2) / 2 <= q) ++a;
    const int b = q - a * (a + 1) / 2;
    double blk[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int ch = 0; ch < 8; ++ch) {
        const int ci = kIE[ch], cj = kJE[ch], ck = kKE[ch];
        for (int p = 0; p < 8; ++p) {
            const double wa = trilin_w(a, ci + kIE[p], cj + kJE[p], ck + kKE[p]);
            if (wa == 0.0) continue;
            for (int r = 0; r < 8; ++r) {
                const double wb = trilin_w(b, ci + kIE[r], cj + kJE[r], ck + kKE[r]);
                if (wb == 0.0) continue;
                const double w = wa * wb;
                const double *s = &S[ch][24 * (3 * p) + 3 * r];
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int d = 0; d < 3; ++d) blk[c][d] += w * s[24 * c + d];
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int r = 3 * a + c, s = 3 * b + d;
            if (r >= s) Ac[(long)pidx2(r, s) * gc.nel + e] = blk[c][d];
        }
}

void launch_coarse_galerkin(hipStream_t s, const Grid &gc, const double *Af, const double *Ke, const double *CXf,
                            double *Ac)
{
    k_galerkin<<<(unsigned)gc.nel, 64, 0, s>>>(gc, Af, Ke, CXf, Ac);
}

// ---------------------------------------------------------------------------
// Element-Galerkin operators from per-fine-element tables.  The Galerkin
// product of level l (coarseAe3c, mgcg1.cu:1653-1743: A_c = P^T (sum of the 8
// children) P with the trilinear P of mgcg1.cu:1653-1676) is linear in the
// level-0 element scales, so the level-l element matrix is
//   A_e = sum_f CX[f] T_l[f],   f over the 8^l level-0 elements inside e,
// with T_l[f] = P_chain(f)^T B(f) P_chain(f) fixed per fine position: B = Ke
// when level 1 is itself Galerkin, B = the Gauss-point matrix of f's position
// in its level-1 parent when level 1 is rediscretised (coarseAe2c,
// mgcg1.cu:2042-2070).  The tables are built once on the host; the device
// work per level is one [nel_l x 8^l] x [8^l x 300] product.  Same operator
// as the child-by-child product, summed in a different order.
// ---------------------------------------------------------------------------
// The product on the matrix cores: D[p][e] = sum_f T[f][p] CX0[fine(e, f)]
// as v_mfma_f64_16x16x4f64 tiles (operand layout at dgemm_tile64), p (the
// packed entry) along the rows and e along the columns, so the stores of a
// tile row are contiguous in A (p-major).  One wave per 32 x 32 tile of
// (p, e), four waves per workgroup; the operands come straight from L2 (the
// table is at most 1.2 MB, CX0 is read once per p tile), kGtAhead K steps
// of loads in flight.
constexpr int kGtAhead = 8;

__global__ __launch_bounds__(256) void k_galerkin_table(Grid gc, int depth, const double *__restrict__ T,
                                                       const double *__restrict__ CX0, double *__restrict__ A)
{
    const int lane = threadIdx.x & 63;
    const long wt = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const long etiles = (gc.nel + 31) / 32;
    const int p0 = (int)(wt / etiles) * 32;
    const long e0 = (wt % etiles) * 32;
    if (p0 >= 300) return;  // whole wave
    const int S = 1 << depth, nf = S * S * S;
    const long ny0 = (long)gc.ny * S, nz0 = (long)gc.nz * S;
    const int i16 = lane & 15, kq = lane >> 4;
    // B operand: element e0 + 16 b + i16 of this lane, its first fine element
    long eb[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        const long e = e0 + 16 * b + i16;
        eb[b] = -1;
        if (e < gc.nel) {
            const long i = e / ((long)gc.ny * gc.nz), k = (e / gc.ny) % gc.nz, j = e % gc.ny;
            eb[b] = (i * S) * nz0 * ny0 + (k * S) * ny0 + j * S;
        }
    }
    f64x4 d[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) d[a][b] = f64x4{0.0, 0.0, 0.0, 0.0};
    for (int f0 = 0; f0 < nf; f0 += 4 * kGtAhead) {
        double fa[kGtAhead][2], fb[kGtAhead][2];
#pragma unroll
        for (int u = 0; u < kGtAhead; ++u) {
            const int f = f0 + 4 * u + kq;  // this lane's K index
            const bool fin = f < nf;
            const long fo = (long)(f / (S * S)) * nz0 * ny0 + (long)((f / S) % S) * ny0 + f % S;
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const int p = p0 + 16 * a + i16;
                fa[u][a] = (fin && p < 300) ? T[(long)f * 300 + p] : 0.0;
            }
#pragma unroll
            for (int b = 0; b < 2; ++b) fb[u][b] = (fin && eb[b] >= 0) ? CX0[eb[b] + fo] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kGtAhead; ++u)
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    d[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[u][a], fb[u][b], d[a][b], 0, 0, 0);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const long e = e0 + 16 * b + i16;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int p = p0 + 16 * a + 4 * v + kq;
                if (p < 300 && e < gc.nel) A[(long)p * gc.nel + e] = d[a][b][v];
            }
        }
}

void launch_galerkin_table(hipStream_t s, const Grid &gc, int depth, const double *T, const double *CX0, double *A)
{
    const long waves = 10 * ((gc.nel + 31) / 32);  // ceil(300 / 32) p tiles
    k_galerkin_table<<<(unsigned)((waves + 3) / 4), 256, 0, s>>>(gc, depth, T, CX0, A);
}

// Host: the tables T_l for l = 1..maxdepth, packed [8^l][300] with the
// fine position f = (fi*S + fk)*S + fj, S = 2^l.  base1: the 8 full matrices
// of depth 1 (index 4 fi + 2 fk + fj), or nullptr to start from Ke at depth 0.
static void child_prolongation(int cx, int cy, int cz, double P[24][24])
{
    // P[3a+d][3b+d] = trilinear weight of parent node b at child node a
    static const int ie[8] = {0, 1, 1, 0, 0, 1, 1, 0}, je[8] = {0, 0, 1, 1, 0, 0, 1, 1}, ke[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    for (int r = 0; r < 24; ++r)
        for (int c = 0; c < 24; ++c) P[r][c] = 0.0;
    for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b) {
            const double w = (1.0 - 0.5 * std::abs(cx + ie[a] - 2 * ie[b])) * (1.0 - 0.5 * std::abs(cy + je[a] - 2 * je[b])) *
                             (1.0 - 0.5 * std::abs(cz + ke[a] - 2 * ke[b]));
            for (int d = 0; d < 3; ++d) P[3 * a + d][3 * b + d] = w;
        }
}

std::vector<std::vector<double>> galerkin_tables(const double *Ke, const double *base1, int maxdepth)
{
    std::vector<std::vector<double>> out(maxdepth + 1);
    std::vector<double> cur;  // full matrices of the current depth, [S^3][576]
    int d = 0;
    if (base1) {
        cur.assign(8 * 576, 0.0);
        for (int fi = 0; fi < 2; ++fi)
            for (int fk = 0; fk < 2; ++fk)
                for (int fj = 0; fj < 2; ++fj)
                    std::copy(base1 + 576 * (4 * fi + 2 * fk + fj), base1 + 576 * (4 * fi + 2 * fk + fj) + 576,
                              cur.begin() + 576 * ((fi * 2 + fk) * 2 + fj));
        d = 1;
    } else {
        cur.assign(Ke, Ke + 576);
    }
    auto pack = [](const std::vector<double> &full, int n) {
        std::vector<double> t((size_t)n * 300);
        for (int f = 0; f < n; ++f)
            for (int m = 0; m < 24; ++m)
                for (int c = 0; c <= m; ++c) t[(size_t)f * 300 + m * (m + 1) / 2 + c] = full[(size_t)f * 576 + 24 * m + c];
        return t;
    };
    if (d >= 1 && d <= maxdepth) out[d] = pack(cur, 8);
    double P[8][24][24];
    for (int c = 0; c < 8; ++c) child_prolongation(c & 1, (c >> 1) & 1, c >> 2, P[c]);
    while (d < maxdepth) {
        const int S = 1 << d, S2 = 2 * S;
        std::vector<double> nxt((size_t)S2 * S2 * S2 * 576);
        double tmp[24][24];
        for (int fi = 0; fi < S2; ++fi)
            for (int fk = 0; fk < S2; ++fk)
                for (int fj = 0; fj < S2; ++fj) {
                    const int cx = fi / S, cy = fj / S, cz = fk / S;
                    const double(*Pc)[24] = P[cx + 2 * cy + 4 * cz];
                    const double *B = &cur[(size_t)(((fi % S) * S + (fk % S)) * S + (fj % S)) * 576];
                    // tmp = B P, then out = P^T tmp
                    for (int r = 0; r < 24; ++r)
                        for (int c = 0; c < 24; ++c) {
                            double acc = 0.0;
                            for (int q = 0; q < 24; ++q) acc += B[24 * r + q] * Pc[q][c];
                            tmp[r][c] = acc;
                        }
                    double *o = &nxt[(size_t)((fi * S2 + fk) * S2 + fj) * 576];
                    for (int r = 0; r < 24; ++r)
                        for (int c = 0; c < 24; ++c) {
                            double acc = 0.0;
                            for (int q = 0; q < 24; ++q) acc += Pc[q][r] * tmp[q][c];
                            o[24 * r + c] = acc;
                        }
                }
        cur.swap(nxt);
        ++d;
        out[d] = pack(cur, S2 * S2 * S2);
    }
    return out;
}

// ---------------------------------------------------------------------------
// mgcg9 coarse levels (Homogenization/QNC/mgcg9.cu).
// formCXc (:753-789): CX_c = 2 * mean of the 8 children, summed in the
// reference's child order.
__global__ void k_coarse_cx(Grid gc, const double *__restrict__ CXf, double *__restrict__ CXc)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= gc.nel) return;
    const int j = (int)(e % gc.ny), k = (int)((e / gc.ny) % gc.nz), i = (int)(e / ((long)gc.ny * gc.nz));
    const long fny = 2L * gc.ny, fnz = 2L * gc.nz;
    double s = 0.0;
    for (int iP = 0; iP < 2; ++iP)
        for (int kP = 0; kP < 2; ++kP)
            for (int jP = 0; jP < 2; ++jP) s += CXf[fnz * fny * (2 * i + iP) + fny * (2 * k + kP) + 2 * j + jP];
    CXc[e] = 2.0 * (s / 8);
}

void launch_coarse_cx(hipStream_t s, const Grid &gc, const double *CXf, double *CXc)
{
    k_coarse_cx<<<(unsigned)((gc.nel + 255) / 256), 256, 0, s>>>(gc, CXf, CXc);
}

// packed element matrices CX_e Ke (assembAQc's element products, :1747-...)
__global__ void k_ke_elements(Grid g, const double *__restrict__ CX, const double *__restrict__ Ke,
                              double *__restrict__ A)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 300 * g.nel) return;
    const int p = (int)(q / g.nel);
    const long e = q % g.nel;
    int m = 0;
    while ((m + 1) * (m + 2) / 2 <= p) ++m;
    const int n = p - m * (m + 1) / 2;
    A[q] = CX[e] * Ke[24 * m + n];
}

void launch_ke_elements(hipStream_t s, const Grid &g, const double *CX, const double *Ke, double *A)
{
    k_ke_elements<<<(unsigned)((300 * g.nel + 255) / 256), 256, 0, s>>>(g, CX, Ke, A);
}

// CorrCsAQc (:1005-1028): zero the rows and columns of the fixed DOFs of
// the dense coarsest matrix (N x N, row-major), then 1e10 on their diagonal
__global__ void k_zero_rowcol(int N, const int *__restrict__ fixed, int nfd, double *__restrict__ AQ)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long)nfd * N) return;
    const long d = fixed[q / N], i = q % N;
    AQ[d * N + i] = 0.0;
    AQ[i * N + d] = 0.0;
}

__global__ void k_fixed_diag(int N, const int *__restrict__ fixed, int nfd, double *__restrict__ AQ)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nfd) AQ[(long)fixed[f] * N + fixed[f]] = 1e10;
}

void launch_zero_rowcol(hipStream_t s, int N, const int *fixed_d, int nfd, double *AQ)
{
    if (nfd <= 0) return;
    k_zero_rowcol<<<(unsigned)(((long)nfd * N + 255) / 256), 256, 0, s>>>(N, fixed_d, nfd, AQ);
    k_fixed_diag<<<(unsigned)((nfd + 255) / 256), 256, 0, s>>>(N, fixed_d, nfd, AQ);
}

// ---------------------------------------------------------------------------
// Coarsest-level penalty (CorrCsAc, mgcg1.cu:1368-1402): the diagonal of
// every incident element matrix is set to 1e10 for each fixed DOF.  The
// reference's table {5,6,1,2,4,7,0,3} is the y-mirror of the node's own
// local index {6,5,2,1,7,4,3,0}; `mirrored` selects it.  Threads that hit
// the same entry store the same value.
// ---------------------------------------------------------------------------
__global__ void k_corrcsa(Grid g, const int *__restrict__ fixed, int nfd, double *__restrict__ A, int mirrored)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nfd) return;
    const int tmir[8] = {5, 6, 1, 2, 4, 7, 0, 3};
    const int town[8] = {6, 5, 2, 1, 7, 4, 3, 0};
    const int dof = fixed[n];
    const int idof = dof % 3, nod = dof / 3;
    const int i = nod / (g.nnz * g.nny);
    const int k = (nod - g.nnz * g.nny * i) / g.nny;
    const int j = nod - g.nnz * g.nny * i - g.nny * k;
    for (int di = -1; di <= 0; di++)
        for (int dk = -1; dk <= 0; dk++)
            for (int dj = -1; dj <= 0; dj++) {
                const int iP = i + di, jP = j + dj, kP = k + dk;
                if (iP < 0 || iP > g.nx - 1 || jP < 0 || jP > g.ny - 1 || kP < 0 || kP > g.nz - 1) continue;
                const long e = (long)g.ny * g.nz * iP + (long)g.ny * kP + jP;
                const int t = 4 * (di + 1) + 2 * (dk + 1) + dj + 1;
                const int n1 = mirrored ? tmir[t] : town[t];
                const int m = 3 * n1 + idof;
                A[(long)pidx2(m, m) * g.nel + e] = 1e10;
            }
}

void launch_corrcsa(hipStream_t s, const Grid &g, const int *fixed_d, int nfd, double *A, int mirrored)
{
    if (nfd > 0) k_corrcsa<<<(nfd + 255) / 256, 256, 0, s>>>(g, fixed_d, nfd, A, mirrored);
}

// ---------------------------------------------------------------------------
// Dense coarsest assembly (assembAc, mgcg1.cu:1747-1806): one thread per
// matrix row gathers the 8 incident element matrices.  The matrix is padded
// to N = roundup(n, 32) with an identity block so the blocked kernels below
// work on whole 32x32 tiles.
// ---------------------------------------------------------------------------
constexpr int NB = 32;

__global__ void k_assemble(Grid g, const double *__restrict__ A, double *__restrict__ AQ, int N)
{
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    if (r >= g.ndof) {
        AQ[r * N + r] = 1.0;
        return;
    }
    const long nod = r / 3;
    const int c = (int)(r % 3);
    const int j = (int)(nod % g.nny), k = (int)((nod / g.nny) % g.nnz), i = (int)(nod / ((long)g.nny * g.nnz));
    for (int ei = 0; ei < 2; ++ei)
        for (int ek = 0; ek < 2; ++ek)
            for (int ej = 0; ej < 2; ++ej) {
                const int ex = i - 1 + ei, ey = j - 1 + ej, ez = k - 1 + ek;
                if (ex < 0 || ex >= g.nx || ey < 0 || ey >= g.ny || ez < 0 || ez >= g.nz) continue;
                const long e = (long)g.ny * g.nz * ex + (long)g.ny * ez + ey;
                const int la = lidx2(1 - ei, 1 - ej, 1 - ek);
                for (int lb = 0; lb < 8; ++lb) {
                    const long nb = (long)g.nny * g.nnz * (ex + kIE[lb]) + (long)g.nny * (ez + kKE[lb]) + (ey + kJE[lb]);
                    for (int d = 0; d < 3; ++d) AQ[r * N + 3 * nb + d] += A[(long)pidx2(3 * la + c, 3 * lb + d) * g.nel + e];
                }
            }
}

void launch_assemble(hipStream_t s, const Grid &g, const double *A, double *AQ)
{
    const int N = (int)((g.ndof + NB - 1) / NB * NB);
    launch_fill(s, AQ, (long)N * N, 0.0);
    k_assemble<<<(N + 255) / 256, 256, 0, s>>>(g, A, AQ, N);
}

// ---------------------------------------------------------------------------
// Blocked right-looking Cholesky A = L L^T on the N x N row-major matrix
// (lower triangle used), NB = 32: factor the diagonal tile in LDS, solve
// the panel below it, update the trailing lower triangle.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Columns J0..J1-1 of the tile factorisation in one wave, lane r < NB
// holding row r in L: at step jj the pivot comes from lane jj and each
// multiplier L[c][jj] from lane c by v_readlane (scalar operands of the
// updates), so no memory round trip is on the chain of pivots.  Only
// columns below J1 are updated.  Entries right of the diagonal (c > r) take
// garbage that is never read: lj = 0 on rows above the column, and only
// c <= r is stored.
template <int J0, int J1>
__device__ __forceinline__ void potrf_cols(double (&L)[NB], int r, bool &bad)
{
#pragma unroll
    for (int jj = J0; jj < J1; ++jj) {
        double d = readlane_f64(L[jj], jj);  // updated A[jj][jj]
        if (!(d > 0.0)) {  // not positive definite (or NaN)
            bad = true;
            d = 1.0;
        }
        // y = 1/sqrt(d): v_rsq_f64 (2^-29 relative) and one Newton step;
        // the pivot is d y and the column is scaled by y, so the chain of a
        // column has no square root or division in it
        double y = __builtin_amdgcn_rsq(d);
        y = y * fma(-0.5 * d * y, y, 1.5);
        const double lj = r == jj ? d * y : (r > jj ? L[jj] * y : 0.0);
        L[jj] = lj;
#pragma unroll
        for (int c = jj + 1; c < J1; ++c) L[c] -= lj * readlane_f64(lj, c);  // L[c][jj]
    }
}

__device__ __forceinline__ void wave_lds_sync()
{
    // LDS operations of one wave complete in order; the fences keep the
    // compiler from moving them across
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One full wave factors the 32x32 diagonal tile at T in LDS (row stride
// ld, lower triangle) in two halves of 16 columns.  The first half's
// contribution to the second, T11 -= L10 L10^T (16x16, K = 16), runs as
// four v_mfma_f64_16x16x4f64 with L10 staged in LDS (operand layout at
// dgemm_tile64); the product is passed back through the tile's unused
// upper-right block.  Returns true on a non-positive pivot.
__device__ __forceinline__ bool potrf_tile(double *T, int ld)
{
    const int lane = threadIdx.x & 63;
    const int r = lane < NB ? lane : NB - 1;  // lanes >= NB shadow row NB-1, results dropped
    double L[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) L[c] = c <= r ? T[r * ld + c] : 0.0;
    bool bad = false;
    potrf_cols<0, 16>(L, r, bad);
    if (lane < NB) {
#pragma unroll
        for (int c = 0; c < 16; ++c) T[r * ld + c] = L[c];
    }
    wave_lds_sync();
    f64x4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kk = 0; kk < 16; kk += 4) {
        const double a = T[(16 + (lane & 15)) * ld + kk + (lane >> 4)];  // A = B = L10
        d = __builtin_amdgcn_mfma_f64_16x16x4f64(a, a, d, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) T[(4 * v + (lane >> 4)) * ld + 16 + (lane & 15)] = d[v];
    wave_lds_sync();
    if (r >= 16) {
#pragma unroll
        for (int c = 16; c < NB; ++c) L[c] -= T[(r - 16) * ld + c];
    }
    potrf_cols<16, NB>(L, r, bad);
    wave_lds_sync();  // every lane has read the product before the tile is overwritten
    if (lane < NB) {
#pragma unroll
        for (int c = 0; c < NB; ++c)
            if (c <= r) T[r * ld + c] = L[c];
    }
    return bad;
}

__global__ __launch_bounds__(64) void k_potrf_diag(double *A, int N, int k0, int *flag)
{
    __shared__ double T[NB][NB + 1];
    for (int q = threadIdx.x; q < NB * NB; q += 64) T[q / NB][q % NB] = A[(long)(k0 + q / NB) * N + k0 + q % NB];
    __syncthreads();
    const bool bad = potrf_tile(&T[0][0], NB + 1);
    __syncthreads();
    for (int q = threadIdx.x; q < NB * NB; q += 64)
        if (q % NB <= q / NB) A[(long)(k0 + q / NB) * N + k0 + q % NB] = T[q / NB][q % NB];
    if (bad && threadIdx.x == 0) flag[0] = 1;
}


// Panel rows solve x L^T = p: x[jj] = (p[jj] - sum_{m<jj} x[m] L[jj][m]) rd[jj]
// with rd[jj] = 1/L[jj][jj] computed once per tile.
__global__ __launch_bounds__(256) void k_trsm_panel(double *A, int N, int k0, int rows)
{
    __shared__ double L[NB][NB + 1];
    __shared__ double rd[NB];
    for (int q = threadIdx.x; q < NB * NB; q += 256) {
        const int r = q / NB, c = q % NB;
        L[r][c] = c <= r ? A[(long)(k0 + r) * N + k0 + c] : 0.0;
    }
    __syncthreads();
    if (threadIdx.x < NB) rd[threadIdx.x] = 1.0 / L[threadIdx.x][threadIdx.x];
    __syncthreads();
    const long row = (long)k0 + NB + (long)blockIdx.x * 256 + threadIdx.x;
    if (row >= (long)k0 + NB + rows) return;
    double x[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) x[c] = A[row * N + k0 + c];
#pragma unroll
    for (int jj = 0; jj < NB; ++jj) {
        double s = x[jj];
#pragma unroll
        for (int m = 0; m < jj; ++m) s -= x[m] * L[jj][m];
        x[jj] = s * rd[jj];
    }
#pragma unroll
    for (int c = 0; c < NB; ++c) A[row * N + k0 + c] = x[c];
}

__device__ __forceinline__ void tri_index(int b, int &ti, int &tj)
{
    // b-th tile of a lower-triangular tile set: (ti, tj) with tj <= ti
    int t = (int)((sqrt(8.0 * b + 1.0) - 1.0) * 0.5);
    while ((t + 1) * (t + 2) / 2 <= b) ++t;
    while (t * (t + 1) / 2 > b) --t;
    ti = t;
    tj = b - t * (t + 1) / 2;
}

// the same for the few tiles of a band step, by integer steps (the index is
// wave-uniform, so this runs on the scalar unit)
__device__ __forceinline__ void tri_index_small(int b, int &ti, int &tj)
{
    int t = 0;
    while ((t + 1) * (t + 2) / 2 <= b) ++t;
    ti = t;
    tj = b - t * (t + 1) / 2;
}

__global__ __launch_bounds__(256) void k_syrk_update(double *A, int N, int k0)
{
    __shared__ double Pi[NB][NB + 1], Pj[NB][NB + 1];
    int ti, tj;
    tri_index(blockIdx.x, ti, tj);
    const int base = k0 + NB;
    const int r0 = base + ti * NB, c0 = base + tj * NB;
    for (int q = threadIdx.x; q < NB * NB; q += 256) {
        const int r = q / NB, c = q % NB;
        Pi[r][c] = A[(long)(r0 + r) * N + k0 + c];
        Pj[r][c] = A[(long)(c0 + r) * N + k0 + c];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < NB * NB; q += 256) {
        const int r = q / NB, c = q % NB;
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < NB; ++m) s += Pi[r][m] * Pj[c][m];
        A[(long)(r0 + r) * N + c0 + c] -= s;
    }
}

// Inverse of each 32x32 diagonal factor tile, written into Linv.
__global__ __launch_bounds__(64) void k_trtri_diag(const double *A, double *Linv, int N)
{
    __shared__ double L[NB][NB + 1];
    const int k0 = blockIdx.x * NB;
    for (int q = threadIdx.x; q < NB * NB; q += 64) {
        const int r = q / NB, c = q % NB;
        L[r][c] = c <= r ? A[(long)(k0 + r) * N + k0 + c] : 0.0;
    }
    __syncthreads();
    const int col = threadIdx.x;
    if (col >= NB) return;
    double x[NB];
#pragma unroll
    for (int r = 0; r < NB; ++r) {
        double s = (r == col) ? 1.0 : 0.0;
#pragma unroll
        for (int m = 0; m < r; ++m) s -= L[r][m] * x[m];
        x[r] = (r < col) ? 0.0 : s / L[r][r];
    }
#pragma unroll
    for (int r = 0; r < NB; ++r) Linv[(long)(k0 + r) * N + k0 + col] = x[r];
}

__global__ void k_transpose(const double *__restrict__ In, double *__restrict__ Out, int N)
{
    __shared__ double T[NB][NB + 1];
    const int bi = blockIdx.y * NB, bj = blockIdx.x * NB;
    for (int q = threadIdx.x; q < NB * NB; q += blockDim.x) T[q / NB][q % NB] = In[(long)(bi + q / NB) * N + bj + q % NB];
    __syncthreads();
    for (int q = threadIdx.x; q < NB * NB; q += blockDim.x) Out[(long)(bj + q / NB) * N + bi + q % NB] = T[q % NB][q / NB];
}

// ---------------------------------------------------------------------------
// The same factorisation in one workgroup when the band spans at most
// kBandRowsMax rows below the diagonal tile (the coarsest grids of the
// usual hierarchies), with a look-ahead.  At step k block column k (the
// diagonal tile and the panel below it) sits in LDS, its tile already
// updated by the earlier steps:
//   A. wave 0 factors the tile while waves off its SIMD apply step k-1's
//      update to the trailing band (in global memory) and to the panel rows
//      of column k (in LDS);
//   B. one thread per panel row solves the panel in two halves of 16
//      columns while the other waves load block column k+1 and store the
//      factored tile;
//   C. panel k is stored, and the diagonal tile of column k+1 receives step
//      k's update, so it is ready for step k+1's factorisation.
// Every entry receives the updates of the steps in increasing order, as in
// the tile-parallel kernels.  The trailing products run on
// v_mfma_f64_16x16x4f64 in 16x16 blocks (operand layout at dgemm_tile64).
// Blocks above the diagonal of a diagonal tile are never read and are not
// updated, and blocks whose panel rows lie past the band are exact zeros
// and are skipped: panel row x of a step (counted from the row below its
// tile) is zero for x >= bw.
// ---------------------------------------------------------------------------
constexpr int kBandRowsMax = 256;
constexpr int kCholThreads = 512;  // 2 waves per SIMD: 256 VGPRs for a panel row
constexpr int kCholWaves = kCholThreads / 64;
constexpr int kCholRows = NB + kBandRowsMax;  // rows of a block column in LDS

// sum_m P[ra + i][m] P[rb + j][m] for a 16x16 block, D layout of the MFMA;
// every operand is read from LDS before the chain starts, so the matrix
// core does not wait on each pair of reads
__device__ __forceinline__ f64x4 panel_product(const double (*P)[NB + 1], int ra, int rb, int lane)
{
    const int i = lane & 15, k = lane >> 4;
    double fa[NB / 4], fb[NB / 4];
#pragma unroll
    for (int s = 0; s < NB / 4; ++s) {
        fa[s] = P[ra + i][4 * s + k];
        fb[s] = P[rb + i][4 * s + k];
    }
    f64x4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < NB / 4; ++s) d = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[s], fb[s], d, 0, 0, 0);
    return d;
}

// two blocks at once: independent accumulation chains overlap on the
// matrix core
__device__ __forceinline__ void panel_product2(const double (*P)[NB + 1], int ra0, int rb0, int ra1, int rb1, int lane,
                                               f64x4 &d0, f64x4 &d1)
{
    const int i = lane & 15, k = lane >> 4;
    double fa0[NB / 4], fb0[NB / 4], fa1[NB / 4], fb1[NB / 4];
#pragma unroll
    for (int s = 0; s < NB / 4; ++s) {
        fa0[s] = P[ra0 + i][4 * s + k];
        fb0[s] = P[rb0 + i][4 * s + k];
        fa1[s] = P[ra1 + i][4 * s + k];
        fb1[s] = P[rb1 + i][4 * s + k];
    }
    d0 = f64x4{0.0, 0.0, 0.0, 0.0};
    d1 = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < NB / 4; ++s) {
        d0 = __builtin_amdgcn_mfma_f64_16x16x4f64(fa0[s], fb0[s], d0, 0, 0, 0);
        d1 = __builtin_amdgcn_mfma_f64_16x16x4f64(fa1[s], fb1[s], d1, 0, 0, 0);
    }
}

// Columns J0..J1-1 of one panel row's substitution x L^T = p (x[m] receives
// its updates in the same order as in k_trsm_panel), updating columns below
// J1 only; Lr: row c of the factored tile in lane c.
template <int J0, int J1>
__device__ __forceinline__ void trsm_cols(double (&x)[NB], const double (&Lr)[NB], const double *rdg)
{
#pragma unroll
    for (int jj = J0; jj < J1; ++jj) {
        x[jj] *= rdg[jj];
        // column jj of the tile is read out once x[jj] is final: otherwise
        // the compiler hoists every readlane and spills them
        double lcol = Lr[jj];
        asm volatile("" : "+v"(lcol) : "v"(x[jj]));
#pragma unroll
        for (int m = jj + 1; m < J1; ++m) x[m] -= x[jj] * readlane_f64(lcol, m);
    }
}

// rows x NB entries of the matrix at G (row stride N) into LDS, kLoads
// loads in flight per thread; thread t0 of nt
constexpr int kLoads = 24;
__device__ __forceinline__ void load_column(double (*D)[NB + 1], const double *G, long N, int rows, int t0, int nt)
{
    const int n = rows * NB;
    for (int q0 = t0; q0 < n; q0 += kLoads * nt) {
        double v[kLoads];
#pragma unroll
        for (int u = 0; u < kLoads; ++u) {
            const int q = q0 + u * nt;
            v[u] = q < n ? G[(long)(q / NB) * N + q % NB] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kLoads; ++u) {
            const int q = q0 + u * nt;
            if (q < n) D[q / NB][q % NB] = v[u];
        }
    }
}

// Step k-1's update of global 16x16 blocks (R, C) of the trailing band, R
// and C counted in 16-row blocks from row k0 + NB: A -= P_R P_C^T with P
// the panel of step k-1 in LDS (row NB + 16 R for block row R).  Items
// i0, i0 + stride, ... < items, mapped to blocks by MAP: 0 the blocks of
// tile column k+1 (C < 2), 1 those right of it (2 <= C <= R).  All loads
// of a batch are issued before its first store (the memory counter also
// waits for earlier stores).
template <int MAP>
__device__ __forceinline__ void rest_blocks(double *A, long N, long k0, const double (*P)[NB + 1], int lane, int items,
                                            int i0, int stride)
{
    constexpr int kBatch = 6;
    const long lo = (long)(lane >> 4) * N + (lane & 15);  // the lane's entry of a block
    for (int w0 = i0; w0 < items; w0 += kBatch * stride) {
        double old[kBatch][4];
        long go[kBatch];  // uniform: the block's first entry
        int ra[kBatch], rb[kBatch];
#pragma unroll
        for (int b = 0; b < kBatch; ++b) {
            const int w = w0 + b * stride;
            if (w < items) {
                int R, C;
                if (MAP == 0) {  // slot 2R + C, slot 1 (above the diagonal) skipped
                    const int sl = w + (w >= 1);
                    R = sl >> 1;
                    C = sl & 1;
                } else {
                    tri_index_small(w, R, C);
                    R += 2;
                    C += 2;
                }
                ra[b] = NB + 16 * R;
                rb[b] = NB + 16 * C;
                go[b] = (k0 + ra[b]) * N + k0 + rb[b];
#pragma unroll
                for (int v = 0; v < 4; ++v) old[b][v] = A[go[b] + lo + 4L * v * N];
            }
        }
        // every product of the batch first, then the stores: a store before
        // a product would make the next operand waits wait for it too
        f64x4 d[kBatch];
#pragma unroll
        for (int b = 0; b < kBatch; b += 2) {
            if (w0 + (b + 1) * stride < items)
                panel_product2(P, NB + ra[b], NB + rb[b], NB + ra[b + 1], NB + rb[b + 1], lane, d[b], d[b + 1]);
            else if (w0 + b * stride < items)
                d[b] = panel_product(P, NB + ra[b], NB + rb[b], lane);
        }
#pragma unroll
        for (int b = 0; b < kBatch; ++b) {
            if (w0 + b * stride < items) {
#pragma unroll
                for (int v = 0; v < 4; ++v) A[go[b] + lo + 4L * v * N] = old[b][v] - d[b][v];
            }
        }
    }
}

// Step k's update of the column buffer D (the next block column, row q of D
// = panel row q of P's step): the 16x16 blocks in slots sl = 2R + C, C < 2,
// sl = s0, s0 + stride, ... < s1, slot 1 (above the diagonal) excluded
__device__ __forceinline__ void column_blocks(double (*D)[NB + 1], const double (*P)[NB + 1], int lane, int s0, int s1,
                                              int stride)
{
    for (int sl = s0; sl < s1; sl += 2 * stride) {
        const int t0 = sl, t1 = sl + stride;
        const int ra0 = (t0 >> 1) * 16, rb0 = (t0 & 1) * 16, ra1 = (t1 >> 1) * 16, rb1 = (t1 & 1) * 16;
        if (t1 < s1) {
            f64x4 d0, d1;
            panel_product2(P, NB + ra0, NB + rb0, NB + ra1, NB + rb1, lane, d0, d1);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                D[ra0 + 4 * v + (lane >> 4)][rb0 + (lane & 15)] -= d0[v];
                D[ra1 + 4 * v + (lane >> 4)][rb1 + (lane & 15)] -= d1[v];
            }
        } else {
            const f64x4 d = panel_product(P, NB + ra0, NB + rb0, lane);
#pragma unroll
            for (int v = 0; v < 4; ++v) D[ra0 + 4 * v + (lane >> 4)][rb0 + (lane & 15)] -= d[v];
        }
    }
}

__global__ __launch_bounds__(kCholThreads) void k_chol_band(double *A, int N, int band_rows, int bw, int *flag)
{
    __shared__ double buf[2][kCholRows][NB + 1];
    __shared__ double rdg[NB];  // 1/L[jj][jj] of the current tile
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nb = N / NB;
    bool bad = false;
    load_column(buf[0], A, N, min(N, NB + band_rows), tid, kCholThreads);  // block column 0 as assembled
    __syncthreads();
    for (int k = 0; k < nb; ++k) {
        double(*Cur)[NB + 1] = buf[k & 1];
        double(*Oth)[NB + 1] = buf[(k + 1) & 1];  // panel k-1 in phase A, block column k+1 from phase B on
        const long k0 = (long)k * NB;
        const int rows = min(N - (int)k0 - NB, band_rows);   // panel rows of step k
        const int prows = k > 0 ? min(N - (int)k0, band_rows) : 0;  // panel rows of step k-1 (from row k0)
        const int rc = prows > NB ? min(2 * (prows / NB - 1), max(0, (bw - NB + 15) / 16)) : 0;  // rest block rows
        // A. factor the tile | step k-1's update of the trailing band in
        // global memory and of column k's panel rows in LDS (slots from 4:
        // below the diagonal tile), by the waves off the factoring wave's
        // SIMD (waves w and w + 4 share a SIMD)
        if (wave == 0) {
            bad |= potrf_tile(&Cur[0][0], NB + 1);
            if (lane < NB) rdg[lane] = 1.0 / Cur[lane][lane];
        } else if (wave != 4 && k > 0) {
            const int aw = wave - 1 - (wave > 4);  // 0..5
            if (rc > 0) rest_blocks<0>(A, N, k0, Oth, lane, 2 * rc - 1, aw, kCholWaves - 2);
            if (rc > 2) rest_blocks<1>(A, N, k0, Oth, lane, (rc - 2) * (rc - 1) / 2, aw, kCholWaves - 2);
            column_blocks(Cur, Oth, lane, 4 + aw, 2 * (min(prows, bw + 15) / 16), kCholWaves - 2);
        }
        __syncthreads();
        // B. panel solve, one row per thread, in two halves of 16 columns
        // (column-oriented substitution; lane c of each wave holds row c of
        // the factored tile, whose entries reach the other lanes by
        // v_readlane) | the other waves load block column k+1 into x (it
        // reaches LDS after the second half, so the loads stay in flight
        // meanwhile) and store the tile (after the loads, which would
        // otherwise wait for the stores).  Between the halves, P[:, 16:32]
        // -= P[:, 0:16] L[16:32, 0:16]^T on the matrix cores.
        const int tw = (rows + 63) & ~63;  // threads of the waves with panel rows
        double Lr[NB], x[NB];
        if (tid < tw) {
            const int lr = lane < NB ? lane : NB - 1;
#pragma unroll
            for (int c = 0; c < NB; ++c) Lr[c] = Cur[lr][c];
#pragma unroll
            for (int c = 0; c < 16; ++c) x[c] = Cur[NB + tid][c];  // rows past the panel: stale, not stored
            trsm_cols<0, 16>(x, Lr, rdg);
            if (tid < rows) {
#pragma unroll
                for (int c = 0; c < 16; ++c) Cur[NB + tid][c] = x[c];
            }
        } else {
            const int t0 = tid - tw, nt = kCholThreads - tw;
            const int n = k + 1 < nb ? min(N - (int)k0 - NB, NB + band_rows) * NB : 0;
            const double *G = A + (k0 + NB) * N + k0 + NB;
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                const int q = t0 + u * nt;
                x[u] = q < n ? G[(long)(q / NB) * N + q % NB] : 0.0;
            }
            for (int q = t0; q < NB * NB; q += nt) {
                const int r = q / NB, c = q % NB;
                if (c <= r) A[(k0 + r) * N + k0 + c] = Cur[r][c];
            }
        }
        __syncthreads();
        for (int w = wave; w < min(rows, bw + 15) / 16; w += kCholWaves) {  // panel rows past the band are zero
            f64x4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int kk = 0; kk < 16; kk += 4)
                d = __builtin_amdgcn_mfma_f64_16x16x4f64(Cur[NB + 16 * w + (lane & 15)][kk + (lane >> 4)],
                                                         Cur[16 + (lane & 15)][kk + (lane >> 4)], d, 0, 0, 0);
#pragma unroll
            for (int v = 0; v < 4; ++v) Cur[NB + 16 * w + 4 * v + (lane >> 4)][16 + (lane & 15)] -= d[v];
        }
        __syncthreads();
        if (tid < tw) {
#pragma unroll
            for (int c = 16; c < NB; ++c) x[c] = Cur[NB + tid][c];
            trsm_cols<16, NB>(x, Lr, rdg);
            if (tid < rows) {
#pragma unroll
                for (int c = 16; c < NB; ++c) Cur[NB + tid][c] = x[c];
            }
        } else {
            const int t0 = tid - tw, nt = kCholThreads - tw;
            const int n = k + 1 < nb ? min(N - (int)k0 - NB, NB + band_rows) * NB : 0;
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                const int q = t0 + u * nt;
                if (q < n) Oth[q / NB][q % NB] = x[u];
            }
            load_column(Oth, A + (k0 + NB) * N + k0 + NB, N, n / NB, t0 + NB * nt, nt);  // the rest, if any
        }
        __syncthreads();
        // C. store panel k; the diagonal tile of column k+1 -= panel k's
        // first tile x its transpose (slots 0, 2, 3)
        for (int q = tid; q < rows * NB; q += kCholThreads) A[(k0 + NB + q / NB) * N + k0 + q % NB] = Cur[NB + q / NB][q % NB];
        if (k + 1 < nb && wave < 3) {
            const int sl = wave == 0 ? 0 : wave + 1;
            const int ra = (sl >> 1) * 16, rb = (sl & 1) * 16;
            const f64x4 d = panel_product(Cur, NB + ra, NB + rb, lane);
#pragma unroll
            for (int v = 0; v < 4; ++v) Oth[ra + 4 * v + (lane >> 4)][rb + (lane & 15)] -= d[v];
        }
        __syncthreads();
    }
    if (bad && tid == 0) flag[0] = 1;
}

// ---------------------------------------------------------------------------
// L^-1 by recursive halving.  With the block rows of a group of g tiles
// split in halves, L = [L11 0; L21 L22] and inv(L) = [X11 0; X21 X22] with
//   X21 = -X22 L21 X11,
// where X11 and X22, the inverses of the halves, come from the level below
// (g/2; level 1 is the 32x32 diagonal tile inverses of k_trtri_diag).  L21
// has nonzero tiles only in its top-right nbt x nbt corner (the band), so
//   W   = L21[0:r, n1-c:n1] X11[n1-c:n1, 0:n1]      (r x n1 tiles)
//   X21 = -X22[0:n2, 0:r] W                          (n2 x n1 tiles)
// with n1 = g/2, n2 <= n1 the tiles of the second half, r = min(nbt, n2)
// and c = min(nbt, n1).  The groups of a level are independent: each level
// is two batched products (blockIdx.z = group), so the inverse takes
// 2 ceil(log2 nb) launches instead of a march down nb dependent block rows.
// ---------------------------------------------------------------------------
struct LinvLevel {
    int g, nb, nbt;  // group size in tiles, tiles of the matrix, band in tiles
    long wstride;    // doubles of W per group
};

// One 64x64 tile of C = alpha A B, with A (M x K), B (K x Nc) and C
// row-major (leading dimensions lda, ldb, ldc), M, Nc and K multiples of 32.
// Four waves, each a 32x32 quadrant as 2 x 2 v_mfma_f64_16x16x4f64 tiles;
// the operand layout of that instruction: A[i][k] in lane i + 16k, B[k][j]
// in lane j + 16k, D[4v + q][j] in element v of lane j + 16q.
__device__ __forceinline__ void dgemm_tile64(const double *__restrict__ A, long lda, const double *__restrict__ B,
                                             long ldb, double *__restrict__ C, long ldc, int M, int Nc, int K,
                                             double alpha)
{
    __shared__ double As[64][33];
    __shared__ double Bs[32][65];
    const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
    if (i0 >= M || j0 >= Nc) return;  // whole workgroup
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = w >> 1, wc = w & 1;
    f64x4 d[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) d[a][b] = f64x4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += 32) {
        // rows past M (A) and columns past Nc (B) read as zeros
        double av[8], bv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int r = (tid >> 5) + 8 * u, c = tid & 31;
            av[u] = i0 + r < M ? A[(long)(i0 + r) * lda + k0 + c] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int r = (tid >> 6) + 4 * u, c = tid & 63;
            bv[u] = j0 + c < Nc ? B[(long)(k0 + r) * ldb + j0 + c] : 0.0;
        }
        __syncthreads();  // the previous step's readers are done
#pragma unroll
        for (int u = 0; u < 8; ++u) As[(tid >> 5) + 8 * u][tid & 31] = av[u];
#pragma unroll
        for (int u = 0; u < 8; ++u) Bs[(tid >> 6) + 4 * u][tid & 63] = bv[u];
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 32; kk += 4) {
            double fa[2], fb[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) fa[a] = As[wr * 32 + a * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
            for (int b = 0; b < 2; ++b) fb[b] = Bs[kk + (lane >> 4)][wc * 32 + b * 16 + (lane & 15)];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) d[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[a], fb[b], d[a][b], 0, 0, 0);
        }
    }
    const int qi = i0 + wr * 32, qj = j0 + wc * 32;
    if (qi >= M || qj >= Nc) return;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                C[(long)(qi + a * 16 + 4 * v + (lane >> 4)) * ldc + qj + b * 16 + (lane & 15)] = alpha * d[a][b][v];
}

template <int STEP>
__global__ __launch_bounds__(256) void k_linv_level(const double *__restrict__ L, double *X, double *W, int N,
                                                   LinvLevel lv)
{
    const int z = blockIdx.z, n1 = lv.g / 2, s = z * lv.g, h = s + n1;
    if (h >= lv.nb) return;
    const int n2 = min(lv.g, lv.nb - s) - n1;
    const int r = min(lv.nbt, n2), c = min(lv.nbt, n1);
    double *Wz = W + z * lv.wstride;
    const long ldw = (long)n1 * NB;
    if (STEP == 0)  // W = L21[0:r, n1-c:n1] X11[n1-c:n1, 0:n1]
        dgemm_tile64(L + (long)h * NB * N + (long)(h - c) * NB, N, X + (long)(h - c) * NB * N + (long)s * NB, N, Wz,
                     ldw, r * NB, n1 * NB, c * NB, 1.0);
    else  // X21 = -X22[0:n2, 0:r] W
        dgemm_tile64(X + (long)h * NB * N + (long)h * NB, N, Wz, ldw, X + (long)h * NB * N + (long)s * NB, N, n2 * NB,
                     n1 * NB, r * NB, -1.0);
}

// Levels 2, 4, ... of the halving; X holds the diagonal tile inverses and
// zeros above them, W has room for wcap doubles.
static void launch_linv_levels(hipStream_t s, const double *L, double *X, double *W, long wcap, int N, int nbt)
{
    const int nb = N / NB;
    for (int g = 2; g / 2 < nb; g *= 2) {
        const int n1 = g / 2;
        const int groups = (nb - n1 + g - 1) / g;
        const int rmax = std::min(nbt, n1);
        const LinvLevel lv{g, nb, nbt, (long)rmax * NB * n1 * NB};
        if (lv.wstride * groups > wcap) throw std::runtime_error("coarsest inverse: work buffer too small");
        const unsigned tn = (unsigned)((n1 * NB + 63) / 64);
        k_linv_level<0><<<dim3(tn, (unsigned)((rmax * NB + 63) / 64), (unsigned)groups), 256, 0, s>>>(L, X, W, N, lv);
        k_linv_level<1><<<dim3(tn, tn, (unsigned)groups), 256, 0, s>>>(L, X, W, N, lv);
    }
}

void coarse_factor(hipStream_t s0, Coarse &C, bool lag)
{
    C.target = (lag && C.have) ? 1 - C.cur : C.cur;
    C.solve_waits = C.target == C.cur;
    C.Linv = C.Lbuf[C.cur];
    C.LinvT = C.LTbuf[C.cur];
    const int N = (C.n + NB - 1) / NB * NB;
    const int nb = N / NB;
    // the factor of a banded matrix keeps its band: panel rows and trailing
    // tiles beyond it are exact zeros and are skipped
    const int bw = std::max(1, std::min(C.bw, N));
    const int band_rows = (bw + NB - 1) / NB * NB;
    hipStream_t s = C.side;
    HIPCHK(hipEventRecord(C.ev_in, s0));
    HIPCHK(hipStreamWaitEvent(s, C.ev_in, 0));
    HIPCHK(hipMemsetAsync(C.flag, 0, sizeof(int), s));
    if (band_rows <= kBandRowsMax) {
        k_chol_band<<<1, kCholThreads, 0, s>>>(C.AQ, N, band_rows, bw, C.flag);
    } else {
        for (int kb = 0; kb < nb; ++kb) {
            const int k0 = kb * NB;
            k_potrf_diag<<<1, 64, 0, s>>>(C.AQ, N, k0, C.flag);
            const int rows = std::min(N - k0 - NB, band_rows);
            if (rows > 0) {
                k_trsm_panel<<<(rows + 255) / 256, 256, 0, s>>>(C.AQ, N, k0, rows);
                const int nt = rows / NB;
                k_syrk_update<<<nt * (nt + 1) / 2, 256, 0, s>>>(C.AQ, N, k0);
            }
        }
    }
    HIPCHK(hipGetLastError());
    // the inverse's launches are enqueued by coarse_finish, when the host
    // has enqueued the start of the solve on the main stream
    C.inv_pending = true;
    C.waited = false;
    C.pending = true;
}

void coarse_finish(const Coarse &C)
{
    if (!C.inv_pending) return;
    C.inv_pending = false;
    const int N = (C.n + NB - 1) / NB * NB;
    const int nb = N / NB;
    const int bw = std::max(1, std::min(C.bw, N));
    double *Li = C.Lbuf[C.target], *LiT = C.LTbuf[C.target];
    hipStream_t s = C.side;
    launch_fill(s, Li, (long)N * N, 0.0);
    k_trtri_diag<<<nb, 64, 0, s>>>(C.AQ, Li, N);
    const int nbt = std::min((bw + 2 * NB - 2) / NB, nb);
    // the transpose's buffer is free until the last step: it holds W
    launch_linv_levels(s, C.AQ, Li, LiT, (long)N * N, N, nbt);
    k_transpose<<<dim3(nb, nb), 256, 0, s>>>(Li, LiT, N);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(C.ev_done, s));
}

int coarse_check(Coarse &C)
{
    if (!C.pending || !C.solve_waits) return 0;  // a lagged factor is checked when adopted
    coarse_finish(C);
    int flag = 0;
    HIPCHK(hipMemcpyAsync(&flag, C.flag, sizeof(int), hipMemcpyDeviceToHost, C.side));
    HIPCHK(hipStreamSynchronize(C.side));
    C.pending = false;
    C.have = flag == 0;
    return flag ? GPUTO_ERR_NOT_SPD : 0;
}

void coarse_invalidate(Coarse &C)
{
    C.have = false;
    C.invalid = true;
}

void coarse_adopt(Coarse &C)
{
    const bool invalid = C.invalid;
    C.invalid = false;
    if (!C.pending) {
        if (invalid) C.have = false;
        return;
    }
    coarse_finish(C);
    int flag = 0;
    HIPCHK(hipMemcpyAsync(&flag, C.flag, sizeof(int), hipMemcpyDeviceToHost, C.side));
    HIPCHK(hipStreamSynchronize(C.side));  // launched a solve ago: normally finished
    C.pending = false;
    if (flag || invalid) {
        // a failed factor, or one of an earlier configuration: factor
        // again, exactly, at this setup
        C.have = false;
        return;
    }
    C.cur = C.target;
    C.have = true;
}

void coarse_switch(Coarse &C, hipStream_t s)
{
    if (!C.pending || C.solve_waits) return;  // no lagged factorisation in flight
    coarse_finish(C);
    HIPCHK(hipStreamWaitEvent(s, C.ev_done, 0));
    C.waited = true;
    C.cur = C.target;
    C.Linv = C.Lbuf[C.cur];
    C.LinvT = C.LTbuf[C.cur];
    C.solve_waits = true;  // coarse_check reads its SPD flag when the solve ends
}

// ---------------------------------------------------------------------------
// Coarsest solve x = L^-T L^-1 b as two row-parallel triangular products
// (one 256-thread block per row, fixed reduction order), then x[fixed] = 0
// (mgcg1.c:431-432).  Each thread issues kTrmvLoads row entries (and as
// many of b) before its first product, so a row of up to 2048 entries is
// one memory round trip; the skip flag is read while they are in flight.
// ---------------------------------------------------------------------------
constexpr int kTrmvThreads = 256, kTrmvLoads = 8;

template <bool UPPER, class Ti, class To>
__global__ __launch_bounds__(kTrmvThreads) void k_trmv(const double *__restrict__ M, const Ti *__restrict__ b,
                                                      To *__restrict__ x, int N, int n,
                                                      const uint8_t *__restrict__ mask, const int *ctl)
{
    __shared__ double red[kTrmvThreads / 64];
    ctl = batch_ctl(ctl);
    // lower: b has n entries, x has N (the padded y); upper: the reverse
    b = batch_ptr(b, UPPER ? (long)N : (long)n);
    x = batch_ptr(x, UPPER ? (long)n : (long)N);
    const int row = blockIdx.x;  // < n: the grid has n blocks
    const int lo = UPPER ? row : 0, hi = UPPER ? n - 1 : row;
    const double *__restrict__ Mr = M + (long)row * N;
    double s = 0.0;
    for (int c0 = lo + (int)threadIdx.x; c0 <= hi; c0 += kTrmvThreads * kTrmvLoads) {
        double mv[kTrmvLoads], bv[kTrmvLoads];
#pragma unroll
        for (int u = 0; u < kTrmvLoads; ++u) {
            const int c = c0 + kTrmvThreads * u;
            const int cc = c <= hi ? c : c0;  // in the row's band (c0 <= hi)
            const double m = Mr[cc];
            mv[u] = c <= hi ? m : 0.0;  // entries past the triangle are never used
            bv[u] = (double)b[cc];
        }
        if (c0 == lo + (int)threadIdx.x && ctl && ctl[0]) return;
#pragma unroll
        for (int u = 0; u < kTrmvLoads; ++u) s += mv[u] * bv[u];
    }
    if (ctl && ctl[0]) return;  // the threads with no entries
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        s = (red[0] + red[1]) + (red[2] + red[3]);
        if (UPPER && mask && ((mask[row / 3] >> (row % 3)) & 1u)) s = 0.0;
        x[row] = (To)s;
    }
}

// The solve is FP64 for both vector types: the right-hand side is widened,
// the result rounded once on output.
template <class T>
void launch_coarse_solve(hipStream_t s, const Coarse &C, const T *b, T *x, const uint8_t *mask, const int *ctl, int nb)
{
    const int N = (C.n + NB - 1) / NB * NB;
    coarse_finish(C);  // also a lagged factorisation's, which runs meanwhile
    if (C.solve_waits && !C.waited) {  // the factorisation of this hierarchy, once per stream order
        HIPCHK(hipStreamWaitEvent(s, C.ev_done, 0));
        C.waited = true;
    }
    // C.y has room for kMaxBatch right-hand sides
    k_trmv<false, T, double>
        <<<dim3(C.n, 1, batch_dim(nb)), kTrmvThreads, 0, s>>>(C.Linv, b, C.y, N, C.n, nullptr, ctl);
    k_trmv<true, double, T><<<dim3(C.n, 1, batch_dim(nb)), kTrmvThreads, 0, s>>>(C.LinvT, C.y, x, N, C.n, mask, ctl);
}
template void launch_coarse_solve<double>(hipStream_t, const Coarse &, const double *, double *, const uint8_t *,
                                          const int *, int);
template void launch_coarse_solve<float>(hipStream_t, const Coarse &, const float *, float *, const uint8_t *,
                                         const int *, int);

}  // namespace gputo
