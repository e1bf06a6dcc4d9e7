// Matrix-free level-1 operator of the REF_C hierarchy.  The reference stores
// the rediscretised level-1 element matrices A_e = sum_g CX[child_g] Aelm_g
// (coarseAe2c, Cantilever/TOP3D125MGCGCUDA/mgcg1.cu:2042-2070; 300 doubles
// per element, 531 MB per apply at 192x96x96) and applies them with defectc
// (mgcg1.cu:1266-1324).  Here the same operator is applied from the child
// densities alone, Gauss point by Gauss point:
//   A_e u = sum_g CX_g |J| B_g^T D B_g u,     h = 2, |J| w_g = 1,
// with B_g the strain-displacement matrix of the coarse hex at Gauss point
// g = 4 iP + 2 kP + jP, (s,e,t) = (g_iP, g_jP, g_kP), g_0 = -1/sqrt3,
// g_1 = +1/sqrt3 (the Gauss-point / child pairing of coarseAe2,
// Cantilever/omp/mgcg1.c:1539-1584).
//
// One thread per coarse element evaluates the 8 Gauss points (k_l1_ew
// below); the nodal forces meet in LDS.
#include <type_traits>

#include "vcycle.h"

namespace gputo {

constexpr int kL1Own = 512;  // owned nodes per block

template <class T>
__device__ __forceinline__ T lane_w(int x, T xi)
{
    // linear shape factor in one direction: (1 -/+ xi)/2
    return x ? T(0.5) * (T(1) + xi) : T(0.5) * (T(1) - xi);
}

// A_e u of one element, summed over its 2x2x2 Gauss points by sum
// factorisation (the FP32 form): the displacement gradient at the Gauss
// points from the nodal differences interpolated one direction at a time
// (a trilinear field's x-derivative depends on y and z only, and so on),
// the forces B^T S from the stresses summed over one direction at a time.
// Node order of y as the element matrices (n: x = n in {1,2,5,6}, y = n in
// {2,3,6,7}, z = n >= 4); child / Gauss point g = 4 sx + 2 sz + sy.
template <class T>
__device__ __forceinline__ void l1_element_sf(const T (&u)[2][2][2][3], const T (&cx)[8], T (&y)[24], T lam2mu,
                                              T lam, T mu)
{
    const T g1 = T(0.57735026918962576451);  // 1/sqrt(3)
    const T ih = T(0.5);                     // 1/h, h = 2
    const T wa = T(0.5) * (T(1) + g1), wb = T(0.5) * (T(1) - g1);
    // shape function n of one direction at Gauss coordinate s (0: -g1, 1: +g1)
    auto W = [&](int n, int s) { return n == s ? wa : wb; };
    T gx[3][2][2], gy[3][2][2], gz[3][2][2];  // [c][.][.]: at (sy,sz), (sx,sz), (sx,sy)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        T t[2][2];
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int s = 0; s < 2; ++s)
                t[p][s] = W(0, s) * (u[1][p][0][c] - u[0][p][0][c]) + W(1, s) * (u[1][p][1][c] - u[0][p][1][c]);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) gx[c][a][b] = ih * (W(0, a) * t[0][b] + W(1, a) * t[1][b]);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int s = 0; s < 2; ++s)
                t[p][s] = W(0, s) * (u[p][1][0][c] - u[p][0][0][c]) + W(1, s) * (u[p][1][1][c] - u[p][0][1][c]);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) gy[c][a][b] = ih * (W(0, a) * t[0][b] + W(1, a) * t[1][b]);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int s = 0; s < 2; ++s)
                t[p][s] = W(0, s) * (u[p][0][1][c] - u[p][0][0][c]) + W(1, s) * (u[p][1][1][c] - u[p][1][0][c]);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) gz[c][a][b] = ih * (W(0, a) * t[0][b] + W(1, a) * t[1][b]);
    }
    // stresses S[c][d] (symmetric) at every Gauss point [sx][sy][sz]
    T S[3][3][2][2][2];
#pragma unroll
    for (int sx = 0; sx < 2; ++sx)
#pragma unroll
        for (int sy = 0; sy < 2; ++sy)
#pragma unroll
            for (int sz = 0; sz < 2; ++sz) {
                T G[3][3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    G[c][0] = gx[c][sy][sz];
                    G[c][1] = gy[c][sx][sz];
                    G[c][2] = gz[c][sx][sy];
                }
                const T c0 = cx[4 * sx + 2 * sz + sy];
                const T tr = G[0][0] + G[1][1] + G[2][2];
#pragma unroll
                for (int c = 0; c < 3; ++c) S[c][c][sx][sy][sz] = c0 * (lam * tr + (lam2mu - lam) * G[c][c]);
                S[0][1][sx][sy][sz] = S[1][0][sx][sy][sz] = c0 * mu * (G[0][1] + G[1][0]);
                S[1][2][sx][sy][sz] = S[2][1][sx][sy][sz] = c0 * mu * (G[1][2] + G[2][1]);
                S[0][2][sx][sy][sz] = S[2][0][sx][sy][sz] = c0 * mu * (G[2][0] + G[0][2]);
            }
    auto nidx = [](int x, int yy, int z) { return 4 * z + (yy ? (x ? 2 : 3) : (x ? 1 : 0)); };
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        // d = 0 (x-derivative of the test function): sum over sx, then z, then y
        {
            T P[2][2], Q[2][2], R[2][2];
#pragma unroll
            for (int sy = 0; sy < 2; ++sy)
#pragma unroll
                for (int sz = 0; sz < 2; ++sz) P[sy][sz] = S[c][0][0][sy][sz] + S[c][0][1][sy][sz];
#pragma unroll
            for (int sy = 0; sy < 2; ++sy)
#pragma unroll
                for (int z = 0; z < 2; ++z) Q[sy][z] = W(z, 0) * P[sy][0] + W(z, 1) * P[sy][1];
#pragma unroll
            for (int yy = 0; yy < 2; ++yy)
#pragma unroll
                for (int z = 0; z < 2; ++z) R[yy][z] = ih * (W(yy, 0) * Q[0][z] + W(yy, 1) * Q[1][z]);
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int yy = 0; yy < 2; ++yy)
#pragma unroll
                    for (int z = 0; z < 2; ++z) y[3 * nidx(x, yy, z) + c] += x ? R[yy][z] : -R[yy][z];
        }
        // d = 1: sum over sy, then z, then x
        {
            T P[2][2], Q[2][2], R[2][2];
#pragma unroll
            for (int sx = 0; sx < 2; ++sx)
#pragma unroll
                for (int sz = 0; sz < 2; ++sz) P[sx][sz] = S[c][1][sx][0][sz] + S[c][1][sx][1][sz];
#pragma unroll
            for (int sx = 0; sx < 2; ++sx)
#pragma unroll
                for (int z = 0; z < 2; ++z) Q[sx][z] = W(z, 0) * P[sx][0] + W(z, 1) * P[sx][1];
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int z = 0; z < 2; ++z) R[x][z] = ih * (W(x, 0) * Q[0][z] + W(x, 1) * Q[1][z]);
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int yy = 0; yy < 2; ++yy)
#pragma unroll
                    for (int z = 0; z < 2; ++z) y[3 * nidx(x, yy, z) + c] += yy ? R[x][z] : -R[x][z];
        }
        // d = 2: sum over sz, then y, then x
        {
            T P[2][2], Q[2][2], R[2][2];
#pragma unroll
            for (int sx = 0; sx < 2; ++sx)
#pragma unroll
                for (int sy = 0; sy < 2; ++sy) P[sx][sy] = S[c][2][sx][sy][0] + S[c][2][sx][sy][1];
#pragma unroll
            for (int sx = 0; sx < 2; ++sx)
#pragma unroll
                for (int yy = 0; yy < 2; ++yy) Q[sx][yy] = W(yy, 0) * P[sx][0] + W(yy, 1) * P[sx][1];
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int yy = 0; yy < 2; ++yy) R[x][yy] = ih * (W(x, 0) * Q[0][yy] + W(x, 1) * Q[1][yy]);
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int yy = 0; yy < 2; ++yy)
#pragma unroll
                    for (int z = 0; z < 2; ++z) y[3 * nidx(x, yy, z) + c] += z ? R[x][yy] : -R[x][yy];
        }
    }
}

// Element-per-thread form: one thread per coarse element of the tile (and
// its one-element halo) evaluates all 8 Gauss points in registers and writes
// its 24 nodal forces to LDS; after one barrier every owned node sums the
// forces of its 8 elements in a fixed order.  No colouring, one barrier,
// deterministic.  HALO nodes and NE elements per block: 1024 and 512 in
// FP32, 512 and 256 in FP64 (63.5 KB of LDS either way, two blocks per CU).
constexpr int kL1Ys = 25;  // force row stride (odd: conflict-free rows)
template <class T> struct L1Ew;
template <> struct L1Ew<float> { static constexpr int HALO = 1024, NE = 512; };
template <> struct L1Ew<double> { static constexpr int HALO = 512, NE = 256; };

template <class T, int MODE>
__global__ __launch_bounds__(256, 2) void k_l1_ew(OpArgsT<T> a, Tile t, const T *__restrict__ CXf, T lam2mu,
                                                 T lam, T mu)
{
    batch_shift(a);
    __shared__ T su[3 * L1Ew<T>::HALO];
    __shared__ T sy[L1Ew<T>::NE * kL1Ys];
    const Grid g = a.g;  // the coarse (level-1) grid
    int b = blockIdx.x;
    const int bj = b % t.ntj;
    b /= t.ntj;
    const int bk = b % t.ntk;
    const int bi = b / t.ntk;
    const int j0 = bj * t.TJ, k0 = bk * t.TK, i0 = bi * t.TI;
    const int HJ = t.TJ + 2, HK = t.TK + 2, HI = t.TI + 2;
    const int EJ = t.TJ + 1, EK = t.TK + 1, EI = t.TI + 1;
    const int nown = t.TJ * t.TK * t.TI;
    const int nel = EJ * EK * EI;

    // halo brick of u, one j-row per wave (coalesced rows, wave-uniform row
    // split); kStageRows rows per wave are loaded before any is stored, with
    // addresses outside the mesh clamped and their values masked
    // a wave's rows are wv, wv + 4, ...: their (i, k) position is stepped
    // from row to row (no division per row); node indices are 32-bit
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), ln = threadIdx.x & 63;
    const int rowlen = 3 * HJ;
    const int nrows = HK * HI;
    const unsigned nplane = (unsigned)g.nny * (unsigned)g.nnz;
    int shi = wv / HK, shk = wv - shi * HK;
    for (int r0 = wv; r0 < nrows; r0 += 4 * kStageRows) {
        T v[kStageRows][2];
#pragma unroll
        for (int u = 0; u < kStageRows; ++u) {
            const int row = r0 + 4 * u;
            const int gi = i0 - 1 + shi, gk = k0 - 1 + shk;
            for (shk += 4; shk >= HK; shk -= HK) ++shi;
            const bool rowok = row < nrows && gi >= 0 && gi < g.nnx && gk >= 0 && gk < g.nnz;
            // node (gi, gk, j0 - 1), which is -1 at the grid's first corner
            const long base = 3 * ((long)(nplane * (unsigned)gi + (unsigned)(g.nny * gk)) + (j0 - 1));
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
                const int col = ln + 64 * cc;
                const int gj = j0 - 1 + col / 3;
                const bool ok = rowok && col < rowlen && gj >= 0 && gj < g.nny;
                const long dof = ok ? base + col : 0;
                T x;
                if (MODE == OP_PRE_JACOBI || MODE == OP_PRE_DEFECT)
                    x = T(a.prerelax) * a.F[dof] / a.AD[dof];
                else
                    x = a.U[dof];
                v[u][cc] = ok ? x : T(0);
            }
        }
        // the skip flag is read while the first rows' loads are in flight
        if (r0 == wv && a.ctl && a.ctl[0]) return;
#pragma unroll
        for (int u = 0; u < kStageRows; ++u) {
            const int row = r0 + 4 * u;
            if (row >= nrows) break;
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
                const int col = ln + 64 * cc;
                if (col < rowlen) su[3 * row * HJ + col] = v[u][cc];
            }
        }
    }
    if (a.ctl && a.ctl[0]) return;  // every thread tests it before the barrier
    __syncthreads();

    // element phase
    const int fny = 2 * g.ny, fnz = 2 * g.nz;
    const T g1 = T(0.57735026918962576451);  // 1/sqrt(3)
    const T ih = T(0.5);                     // 1/h, h = 2; |J| w_g = h^3/8 = 1
#pragma unroll 1
    for (int q = threadIdx.x; q < nel; q += 256) {
        const int ej = q % EJ, r = q / EJ;
        const int ek = r % EK, ei = r / EK;
        const int ex = i0 - 1 + ei, ey = j0 - 1 + ej, ez = k0 - 1 + ek;
        T y[24];
#pragma unroll
        for (int m = 0; m < 24; ++m) y[m] = T(0);
        if (ex >= 0 && ex < g.nx && ey >= 0 && ey < g.ny && ez >= 0 && ez < g.nz && std::is_same<T, float>::value) {
            T cx[8];
#pragma unroll
            for (int gp = 0; gp < 8; ++gp) {
                const int is = gp >> 2, it = (gp >> 1) & 1, ie = gp & 1;
                cx[gp] = CXf[(long)fny * fnz * (2 * ex + is) + (long)fny * (2 * ez + it) + 2 * ey + ie];
            }
            T u[2][2][2][3];  // [x][y][z][c]
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int yy = 0; yy < 2; ++yy)
#pragma unroll
                    for (int z = 0; z < 2; ++z) {
                        const int h = 3 * (((ei + x) * HK + (ek + z)) * HJ + (ej + yy));
                        u[x][yy][z][0] = su[h];
                        u[x][yy][z][1] = su[h + 1];
                        u[x][yy][z][2] = su[h + 2];
                    }
            l1_element_sf<T>(u, cx, y, lam2mu, lam, mu);
        } else if (ex >= 0 && ex < g.nx && ey >= 0 && ey < g.ny && ez >= 0 && ez < g.nz) {
            T cx[8];
#pragma unroll
            for (int gp = 0; gp < 8; ++gp) {
                const int is = gp >> 2, it = (gp >> 1) & 1, ie = gp & 1;
                cx[gp] = CXf[(long)fny * fnz * (2 * ex + is) + (long)fny * (2 * ez + it) + 2 * ey + ie];
            }
            T u[2][2][2][3];  // [x][y][z][c]
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int yy = 0; yy < 2; ++yy)
#pragma unroll
                    for (int z = 0; z < 2; ++z) {
                        const int h = 3 * (((ei + x) * HK + (ek + z)) * HJ + (ej + yy));
                        u[x][yy][z][0] = su[h];
                        u[x][yy][z][1] = su[h + 1];
                        u[x][yy][z][2] = su[h + 2];
                    }
#pragma unroll
            for (int gp = 0; gp < 8; ++gp) {
                const int is = gp >> 2, it = (gp >> 1) & 1, ie = gp & 1;
                const T sx = is ? g1 : -g1, sy = ie ? g1 : -g1, sz = it ? g1 : -g1;
                const T wx[2] = {lane_w(0, sx), lane_w(1, sx)};
                const T wy[2] = {lane_w(0, sy), lane_w(1, sy)};
                const T wz[2] = {lane_w(0, sz), lane_w(1, sz)};
                T G[3][3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    T gx = T(0), gy = T(0), gz = T(0);
#pragma unroll
                    for (int p = 0; p < 2; ++p)
#pragma unroll
                        for (int q2 = 0; q2 < 2; ++q2) {
                            gx = fma(wy[p] * wz[q2], u[1][p][q2][c] - u[0][p][q2][c], gx);
                            gy = fma(wx[p] * wz[q2], u[p][1][q2][c] - u[p][0][q2][c], gy);
                            gz = fma(wx[p] * wy[q2], u[p][q2][1][c] - u[p][q2][0][c], gz);
                        }
                    G[c][0] = ih * gx;
                    G[c][1] = ih * gy;
                    G[c][2] = ih * gz;
                }
                const T c0 = cx[gp];
                const T tr = G[0][0] + G[1][1] + G[2][2];
                T S[3][3];
                S[0][0] = c0 * (lam * tr + (lam2mu - lam) * G[0][0]);
                S[1][1] = c0 * (lam * tr + (lam2mu - lam) * G[1][1]);
                S[2][2] = c0 * (lam * tr + (lam2mu - lam) * G[2][2]);
                S[0][1] = S[1][0] = c0 * mu * (G[0][1] + G[1][0]);
                S[1][2] = S[2][1] = c0 * mu * (G[1][2] + G[2][1]);
                S[0][2] = S[2][0] = c0 * mu * (G[2][0] + G[0][2]);
#pragma unroll
                for (int n = 0; n < 8; ++n) {
                    const int x = (n == 1 || n == 2 || n == 5 || n == 6) ? 1 : 0;
                    const int yy = (n == 2 || n == 3 || n == 6 || n == 7) ? 1 : 0;
                    const int z = n >= 4 ? 1 : 0;
                    const T dx = (x ? ih : -ih) * wy[yy] * wz[z];
                    const T dy = (yy ? ih : -ih) * wx[x] * wz[z];
                    const T dz = (z ? ih : -ih) * wx[x] * wy[yy];
#pragma unroll
                    for (int c = 0; c < 3; ++c) y[3 * n + c] += dx * S[c][0] + dy * S[c][1] + dz * S[c][2];
                }
            }
        }
#pragma unroll
        for (int m = 0; m < 24; ++m) sy[q * kL1Ys + m] = y[m];
    }
    __syncthreads();

    // node phase: the forces of the 8 elements of each owned node
    double part = 0.0;
    const T om = T(a.omega);
    for (int own = threadIdx.x; own < nown; own += 256) {
        const int tj = own % t.TJ;
        const int r = own / t.TJ;
        const int tk = r % t.TK, ti = r / t.TK;
        const int gi = i0 + ti, gj = j0 + tj, gk = k0 + tk;
        if (gi >= g.nnx || gj >= g.nny || gk >= g.nnz) continue;
        T yv[3] = {T(0), T(0), T(0)};
#pragma unroll
        for (int x = 1; x >= 0; --x)
#pragma unroll
            for (int z = 1; z >= 0; --z)
#pragma unroll
                for (int yy = 1; yy >= 0; --yy) {
                    // element with first node at halo (tj + 1 - yy, tk + 1 - z, ti + 1 - x)
                    const int q = ((ti + 1 - x) * EK + (tk + 1 - z)) * EJ + (tj + 1 - yy);
                    const int nloc = 4 * z + (yy ? (x ? 2 : 3) : (x ? 1 : 0));
#pragma unroll
                    for (int c = 0; c < 3; ++c) yv[c] += sy[q * kL1Ys + 3 * nloc + c];
                }
        const long nod = (long)g.nny * g.nnz * gi + (long)g.nny * gk + gj;
        const long dof = 3 * nod;
        const unsigned msk = a.mask[nod];
        const int hc = 3 * (((ti + 1) * HK + (tk + 1)) * HJ + (tj + 1));
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const bool fixed = (msk >> c) & 1u;
            const T u = su[hc + c];
            if (MODE == OP_MATVEC) {
                const T v = fixed ? T(0) : yv[c];
                a.out[dof + c] = v;
                part += (double)u * (double)v;
            } else if (MODE == OP_DEFECT || MODE == OP_PRE_DEFECT) {
                const T rv = fixed ? T(0) : a.F[dof + c] - yv[c];
                a.out[dof + c] = rv;
                if (MODE == OP_PRE_DEFECT) a.out2[dof + c] = u;
            } else {
                const T f = a.F[dof + c];
                const T rv = fixed ? T(0) : f - yv[c];
                const T ad = a.AD[dof + c];
                const T un = (T(1) - om) * u + om * (rv + ad * u) / ad;
                a.out[dof + c] = un;
                part += (double)f * (double)un;
            }
        }
    }
    if (a.partials) {
        const double sm = block_sum(part);
        if (threadIdx.x == 0) a.partials[blockIdx.x] = sm;
    }
}

static int l1_balance(int n, int tmax)
{
    if (tmax >= n) return n;
    const int nt = (n + tmax - 1) / tmax;
    return (n + nt - 1) / nt;
}

static Tile l1_tile_limits(const Grid &g, int halo, int own, int elems)
{
    Tile t;
    t.TJ = l1_balance(g.nny, 7);
    t.TK = l1_balance(g.nnz, 7);
    t.TI = l1_balance(g.nnx, 7);
    while ((t.TJ + 2) * (t.TK + 2) * (t.TI + 2) > halo || t.TJ * t.TK * t.TI > own ||
           (t.TJ + 1) * (t.TK + 1) * (t.TI + 1) > elems) {
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

Tile l1_tile(const Grid &g) { return l1_tile_limits(g, L1Ew<float>::HALO, kL1Own, L1Ew<float>::NE); }
Tile l1_tile64(const Grid &g) { return l1_tile_limits(g, L1Ew<double>::HALO, kL1Own, L1Ew<double>::NE); }

template <class T>
void launch_op1_mf(hipStream_t s, int mode, const OpArgsT<T> &a, const T *CXf, const Tile &t, double nu)
{
    // Lame constants of D (E = 1): D11 = lam + 2mu, D12 = lam, D44 = mu
    const double c1 = 1.0 / ((1 + nu) * (1 - 2 * nu));
    const T l2m = T(c1 * (1 - nu)), lam = T(c1 * nu), mu = T(c1 * (1 - 2 * nu) / 2);
    const dim3 grid((unsigned)t.blocks(), 1, batch_dim(a.nb)), block(256);
    switch (mode) {
    case OP_MATVEC: k_l1_ew<T, OP_MATVEC><<<grid, block, 0, s>>>(a, t, CXf, l2m, lam, mu); break;
    case OP_DEFECT: k_l1_ew<T, OP_DEFECT><<<grid, block, 0, s>>>(a, t, CXf, l2m, lam, mu); break;
    case OP_JACOBI: k_l1_ew<T, OP_JACOBI><<<grid, block, 0, s>>>(a, t, CXf, l2m, lam, mu); break;
    case OP_PRE_JACOBI: k_l1_ew<T, OP_PRE_JACOBI><<<grid, block, 0, s>>>(a, t, CXf, l2m, lam, mu); break;
    case OP_PRE_DEFECT: k_l1_ew<T, OP_PRE_DEFECT><<<grid, block, 0, s>>>(a, t, CXf, l2m, lam, mu); break;
    default: break;
    }
}
template void launch_op1_mf<double>(hipStream_t, int, const OpArgsT<double> &, const double *, const Tile &, double);
template void launch_op1_mf<float>(hipStream_t, int, const OpArgsT<float> &, const float *, const Tile &, double);

}  // namespace gputo
