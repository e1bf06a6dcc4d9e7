// Kernels of the MGCG hierarchy, instantiated for FP64 (the reference
// arithmetic) and FP32 (the inner solve of the mixed-precision path, FP32
// accumulation): node-gather operators with LDS-staged nodal bricks,
// operator diagonals, grid transfers, BLAS-1 with deterministic two-stage
// reductions (always accumulated in FP64) and the device-side PCG scalar
// steps.
//
// Layout: DOF vector [3*nnod], node id = nny*nnz*i + nny*k + j (j fastest);
// element scales [nel], element id = ny*nz*i + ny*k + j; coarse element
// matrices packed lower-triangular and entry-major: A[p*nel + e] with
// p = m(m+1)/2 + n for n <= m (the packing of Cantilever/TOP3D125MGCGCUDA/
// mgcg1.cu:1313-1316, stored entry-major so that neighbouring elements are
// neighbouring in memory).
#include "gputo_internal.h"
#include "ke_classes.h"

namespace gputo {

__host__ __device__ constexpr int lidx(int x, int y, int z)
{
    // local node number of offset (x,y,z) in {0,1}^3 in the element order
    // ie={0,1,1,0,0,1,1,0}, je={0,0,1,1,0,0,1,1}, ke={0,0,0,0,1,1,1,1}
    // (Cantilever/omp/mgcg1.c:621-624)
    return 4 * z + (y ? (x ? 2 : 3) : (x ? 1 : 0));
}

__host__ __device__ constexpr int pidx(int m, int n) { return n <= m ? m * (m + 1) / 2 + n : n * (n + 1) / 2 + m; }


// ---------------------------------------------------------------------------
// Level-0 node-gather operator (multAV0c / defect0c, Cantilever/
// TOP3D125MGCGCUDA/mgcg1.cu:941-997,1208-1264).  One block owns a
// TJ x TK x TI brick of nodes; it stages the nodal values of the brick plus a
// one-node halo in LDS (structure of arrays, one plane per component), the
// element scales of the (TJ+1)(TK+1)(TI+1) element box.  Each thread
// then forms the 3 rows of K U at its node from the 27 neighbouring nodes:
//   y_c = sum_e E_e sum_b Ke[3a_e + c][3b + d] u_{b,d}.
// No atomics: each output DOF is written by exactly one thread.
// ---------------------------------------------------------------------------
template <class T, int MODE, bool KC>
__global__ __launch_bounds__(kBlock) void k_gather(OpArgsT<T> a)
{
    batch_shift(a);
    if (a.ctl && a.ctl[0]) return;
    __shared__ T su[3][kMaxHalo];
    __shared__ T sE[kMaxElemBox];
    __shared__ T sK[KC ? 1 : 576];  // general Ke: wave-uniform (broadcast) LDS reads
    const Grid g = a.g;
    const Tile t = a.t;
    const long tb = xcd_tile(blockIdx.x, t.blocks());
    if (tb >= t.blocks()) {  // padding block of the XCD order
        if (a.partials && threadIdx.x == 0) a.partials[blockIdx.x] = 0.0;
        return;
    }
    if (!KC)
        for (int q = threadIdx.x; q < 576; q += kBlock) sK[q] = a.Ke[q];
    int b = (int)tb;
    const int bj = b % t.ntj;
    b /= t.ntj;
    const int bk = b % t.ntk;
    const int bi = b / t.ntk;
    const int j0 = bj * t.TJ, k0 = bk * t.TK, i0 = bi * t.TI;
    const int HJ = t.TJ + 2, HK = t.TK + 2, HI = t.TI + 2;

    // Stage the halo brick one j-row per wave (rows are contiguous in the DOF
    // vector, so the loads coalesce) and the element box the same way; the
    // row split is wave-uniform, so no per-lane integer division is needed.
    // kStageRows rows per wave are loaded into registers before any is
    // stored, so their loads are in flight together (addresses outside the
    // mesh are clamped to 0 and the values masked).
    // A wave's rows are wave, wave + 4, ...: their (i, k) position is
    // stepped from row to row (no division per row), node indices are
    // 32-bit (a grid has far fewer than 2^32 nodes).
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    constexpr int kStep = kBlock / 64;  // waves, and the row stride of one wave
    const int rowlen = 3 * HJ;          // <= 128: two values per lane
    const int nrows = HK * HI;
    const unsigned nplane = (unsigned)g.nny * (unsigned)g.nnz;
    int shi = wave / HK, shk = wave - shi * HK;  // position of the wave's next row
    for (int r0 = wave; r0 < nrows; r0 += kStep * kStageRows) {
        T v[kStageRows][2];
#pragma unroll
        for (int u = 0; u < kStageRows; ++u) {
            const int row = r0 + kStep * u;
            const int gi = i0 - 1 + shi, gk = k0 - 1 + shk;
            const bool rowok = row < nrows && gi >= 0 && gi < g.nnx && gk >= 0 && gk < g.nnz;
            // node (gi, gk, j0 - 1), which is -1 at the grid's first corner
            const long base = 3 * ((long)(nplane * (unsigned)gi + (unsigned)(g.nny * gk)) + (j0 - 1));
            for (shk += kStep; shk >= HK; shk -= HK) ++shi;
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
                const int col = lane + 64 * cc;
                const int gj = j0 - 1 + col / 3;
                const bool ok = rowok && col < rowlen && gj >= 0 && gj < g.nny;
                const long dof = ok ? base + col : 0;
                T x;
                if (MODE == OP_PRE_JACOBI || MODE == OP_PRE_DEFECT)
                    x = T(a.prerelax) * a.F[dof] / a.AD[dof];
                else if (MODE == OP_MATVEC_UF)
                    x = T(a.Uf[dof]);
                else
                    x = a.U[dof];
                v[u][cc] = ok ? x : T(0);
            }
        }
#pragma unroll
        for (int u = 0; u < kStageRows; ++u) {
            const int row = r0 + kStep * u;
            if (row >= nrows) break;
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
                const int col = lane + 64 * cc;
                if (col < rowlen) {
                    const int hj = col / 3, c = col - 3 * hj;
                    su[c][row * HJ + hj] = v[u][cc];
                }
            }
        }
    }
    const int EJ = t.TJ + 1, EK = t.TK + 1;
    {
        const int EI = t.TI + 1;
        const int erows = EK * EI;
        const unsigned eplane = (unsigned)g.ny * (unsigned)g.nz;
        int sei = wave / EK, sek = wave - sei * EK;
        for (int r0 = wave; r0 < erows; r0 += kStep * kStageRows) {
            T v[kStageRows];
#pragma unroll
            for (int u = 0; u < kStageRows; ++u) {
                const int row = r0 + kStep * u;
                const int gi = i0 - 1 + sei, gk = k0 - 1 + sek, gj = j0 - 1 + lane;
                for (sek += kStep; sek >= EK; sek -= EK) ++sei;
                const bool ok = row < erows && lane < EJ && gi >= 0 && gi < g.nx && gk >= 0 && gk < g.nz && gj >= 0 &&
                                gj < g.ny;
                const T x = a.CX[ok ? (long)(eplane * (unsigned)gi + (unsigned)(g.ny * gk + gj)) : 0];
                T e = ok ? x : T(0);
                if (a.cx_floor > 0.0 && e < T(a.cx_floor)) e = T(0);  // mgcg9 defect0c skip
                v[u] = e;
            }
#pragma unroll
            for (int u = 0; u < kStageRows; ++u) {
                const int row = r0 + kStep * u;
                if (row >= erows) break;
                if (lane < EJ) sE[row * EJ + lane] = v[u];
            }
        }
    }
    __syncthreads();

    double part = 0.0;
    const int tid = threadIdx.x;
    if (tid < t.TJ * t.TK * t.TI) {
        const int tj = tid % t.TJ;
        const int r = tid / t.TJ;
        const int tk = r % t.TK, ti = r / t.TK;
        const int gi = i0 + ti, gj = j0 + tj, gk = k0 + tk;
        if (gi < g.nnx && gj < g.nny && gk < g.nnz) {
            const int hb = (ti * HK + tk) * HJ + tj;  // halo index of neighbour offset (0,0,0)
            T y0, y1, y2;
            if (KC) {
                // Neighbour form: for each of the 27 neighbour nodes m (u_m read
                // from LDS once) and each incident element s holding both this
                // node and m, z_s += Ke[3a_s.., 3b..] u_m; then y = sum_s E_s z_s.
                // Every Ke index is a compile-time constant after unrolling and
                // Ke0 has the unit-H8 class structure, so each entry is one of
                // the kKeClasses values passed as kernel arguments (scalar
                // registers): LDS carries only the 81 neighbour values.
                T z[8][3];
#pragma unroll
                for (int q = 0; q < 8; ++q) z[q][0] = z[q][1] = z[q][2] = T(0);
#pragma unroll
                for (int di = -1; di <= 1; ++di)
#pragma unroll
                    for (int dk = -1; dk <= 1; ++dk)
#pragma unroll
                        for (int dj = -1; dj <= 1; ++dj) {
                            const int h = hb + ((1 + di) * HK + (1 + dk)) * HJ + (1 + dj);
                            const T u0 = su[0][h], u1 = su[1][h], u2 = su[2][h];
#pragma unroll
                            for (int si = 0; si < 2; ++si)
#pragma unroll
                                for (int sk = 0; sk < 2; ++sk)
#pragma unroll
                                    for (int sj = 0; sj < 2; ++sj) {
                                        // element s has its first node at n + (s - 1)
                                        const int bx = di - si + 1, by = dj - sj + 1, bz = dk - sk + 1;
                                        if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
                                        const int la = lidx(1 - si, 1 - sj, 1 - sk), lb = lidx(bx, by, bz);
                                        T *zs = z[4 * si + 2 * sk + sj];
#pragma unroll
                                        for (int c = 0; c < 3; ++c) {
                                            const int r = 24 * (3 * la + c) + 3 * lb;
                                            zs[c] = fma(a.kv[kKeClass[r + 0]], u0, zs[c]);
                                            zs[c] = fma(a.kv[kKeClass[r + 1]], u1, zs[c]);
                                            zs[c] = fma(a.kv[kKeClass[r + 2]], u2, zs[c]);
                                        }
                                    }
                        }
                const int eb = (ti * EK + tk) * EJ + tj;
                y0 = T(0), y1 = T(0), y2 = T(0);
#pragma unroll
                for (int si = 0; si < 2; ++si)
#pragma unroll
                    for (int sk = 0; sk < 2; ++sk)
#pragma unroll
                        for (int sj = 0; sj < 2; ++sj) {
                            const T E = sE[eb + (si * EK + sk) * EJ + sj];
                            const T *zs = z[4 * si + 2 * sk + sj];
                            y0 = fma(E, zs[0], y0);
                            y1 = fma(E, zs[1], y1);
                            y2 = fma(E, zs[2], y2);
                        }
            } else {
                y0 = T(0), y1 = T(0), y2 = T(0);
                {
                    // one incident element at a time (runtime loop keeps the live
                    // set small): t = Ke[3a..3a+2, :] u_e, y += E_e t, with the
                    // node's local index a and the Ke rows uniform across the wave
                    const int eb = (ti * EK + tk) * EJ + tj;
#pragma unroll 1
                    for (int e = 0; e < 8; ++e) {
                        const int ei = e >> 2, ek = (e >> 1) & 1, ej = e & 1;
                        const T E = sE[eb + (ei * EK + ek) * EJ + ej];
                        const T *kr = sK + 72 * (4 * (1 - ek) + ((1 - ej) ? ((1 - ei) ? 2 : 3) : ((1 - ei) ? 1 : 0)));
                        const int he = hb + (ei * HK + ek) * HJ + ej;
                        T t0 = T(0), t1 = T(0), t2 = T(0);
#pragma unroll
                        for (int lb = 0; lb < 8; ++lb) {
                            const int bx = (lb == 1 || lb == 2 || lb == 5 || lb == 6) ? 1 : 0;
                            const int by = (lb == 2 || lb == 3 || lb == 6 || lb == 7) ? 1 : 0;
                            const int bz = lb >= 4 ? 1 : 0;
                            const int h = he + (bx * HK + bz) * HJ + by;
                            const T u0 = su[0][h], u1 = su[1][h], u2 = su[2][h];
                            t0 = fma(kr[3 * lb + 0], u0, t0);
                            t0 = fma(kr[3 * lb + 1], u1, t0);
                            t0 = fma(kr[3 * lb + 2], u2, t0);
                            t1 = fma(kr[24 + 3 * lb + 0], u0, t1);
                            t1 = fma(kr[24 + 3 * lb + 1], u1, t1);
                            t1 = fma(kr[24 + 3 * lb + 2], u2, t1);
                            t2 = fma(kr[48 + 3 * lb + 0], u0, t2);
                            t2 = fma(kr[48 + 3 * lb + 1], u1, t2);
                            t2 = fma(kr[48 + 3 * lb + 2], u2, t2);
                        }
                        y0 = fma(E, t0, y0);
                        y1 = fma(E, t1, y1);
                        y2 = fma(E, t2, y2);
                    }
                }
            }
            const long nod = (long)g.nny * g.nnz * gi + (long)g.nny * gk + gj;
            const long dof = 3 * nod;
            const unsigned m = a.mask[nod];
            const T y[3] = {y0, y1, y2};
            const int hc = hb + (HK + 1) * HJ + 1;  // the node itself
            const T om = T(a.omega), one = T(1);
            // all epilogue operands loaded before use (one round trip)
            T fv[3] = {T(0), T(0), T(0)}, adv[3] = {T(1), T(1), T(1)};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (MODE != OP_MATVEC && MODE != OP_MATVEC_UF) fv[c] = a.F[dof + c];
                if (MODE == OP_JACOBI || MODE == OP_PRE_JACOBI) adv[c] = a.AD[dof + c];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const bool fixed = (m >> c) & 1u;
                if (MODE == OP_MATVEC || MODE == OP_MATVEC_UF) {
                    const T v = fixed ? T(0) : y[c];
                    a.out[dof + c] = v;
                    part += (double)su[c][hc] * (double)v;
                } else if (MODE == OP_DEFECT || MODE == OP_PRE_DEFECT) {
                    const T rv = fixed ? T(0) : fv[c] - y[c];
                    a.out[dof + c] = rv;
                    if (MODE == OP_PRE_DEFECT) a.out2[dof + c] = su[c][hc];
                } else {
                    const T f = fv[c];
                    const T rv = fixed ? T(0) : f - y[c];
                    const T u = su[c][hc];
                    const T ad = adv[c];
                    // U = (1-w) U + w (R + AD U)/AD  (mgcg1.cu:1199-1206)
                    const T un = (one - om) * u + om * (rv + ad * u) / ad;
                    a.out[dof + c] = un;
                    part += (double)f * (double)un;
                }
            }
        }
    }
    if (a.partials) {
        const double s = block_sum(part);
        if (threadIdx.x == 0) a.partials[blockIdx.x] = s;
    }
}

template <class T>
void launch_op0(hipStream_t s, int mode, const OpArgsT<T> &a)
{
    const dim3 grid((unsigned)xcd_launch(a.t.blocks()), 1, batch_dim(a.nb)), block(kBlock);
    if (a.kc) {
        switch (mode) {
        case OP_MATVEC: k_gather<T, OP_MATVEC, true><<<grid, block, 0, s>>>(a); break;
        case OP_DEFECT: k_gather<T, OP_DEFECT, true><<<grid, block, 0, s>>>(a); break;
        case OP_JACOBI: k_gather<T, OP_JACOBI, true><<<grid, block, 0, s>>>(a); break;
        case OP_PRE_JACOBI: k_gather<T, OP_PRE_JACOBI, true><<<grid, block, 0, s>>>(a); break;
        case OP_PRE_DEFECT: k_gather<T, OP_PRE_DEFECT, true><<<grid, block, 0, s>>>(a); break;
        default: break;
        }
        return;
    }
    switch (mode) {
    case OP_MATVEC: k_gather<T, OP_MATVEC, false><<<grid, block, 0, s>>>(a); break;
    case OP_DEFECT: k_gather<T, OP_DEFECT, false><<<grid, block, 0, s>>>(a); break;
    case OP_JACOBI: k_gather<T, OP_JACOBI, false><<<grid, block, 0, s>>>(a); break;
    case OP_PRE_JACOBI: k_gather<T, OP_PRE_JACOBI, false><<<grid, block, 0, s>>>(a); break;
    case OP_PRE_DEFECT: k_gather<T, OP_PRE_DEFECT, false><<<grid, block, 0, s>>>(a); break;
    default: break;
    }
}

// ---------------------------------------------------------------------------
// Operator of a level with stored element matrices (levels >= 1; the
// element-matrix products of defectc / multAV, Cantilever/TOP3D125MGCGCUDA/
// mgcg1.cu:1266-1324, Cantilever/omp/mgcg1.c:667-715).  A block holds
// kStNodes nodes x 8 element slots: thread (slot, node) forms the 3 rows of
// the node's local index in the incident element at offset slot, from that
// element's packed matrix.  The slot is uniform over each half-wave and nodes
// run along j, so every packed entry is read as coalesced runs of
// neighbouring elements; the 8 slot sums of a node are added in slot order
// (deterministic).  Eight times the threads of a node-per-thread gather keeps
// the small coarse levels, whose cost is load latency, occupied.
// ---------------------------------------------------------------------------
template <class T, int MODE>
__device__ __forceinline__ T stored_u(const OpArgsT<T> &a, long dof)
{
    if (MODE == OP_PRE_JACOBI || MODE == OP_PRE_DEFECT) return T(a.prerelax) * a.F[dof] / a.AD[dof];
    return a.U[dof];
}

template <class T, int MODE>
__global__ __launch_bounds__(kBlock) void k_stored(OpArgsT<T> a)
{
    batch_shift(a);
    if (a.ctl && a.ctl[0]) return;
    __shared__ T red[8][3][kStNodes];
    const Grid g = a.g;
    const Tile t = a.t;
    int b = blockIdx.x;
    const int bj = b % t.ntj;
    b /= t.ntj;
    const int bk = b % t.ntk;
    const int bi = b / t.ntk;
    const int slot = threadIdx.x / kStNodes, n = threadIdx.x % kStNodes;
    const int tj = n % t.TJ, tk = (n / t.TJ) % t.TK, ti = n / (t.TJ * t.TK);
    const int gi = bi * t.TI + ti, gj = bj * t.TJ + tj, gk = bk * t.TK + tk;
    const bool node_ok = n < t.TJ * t.TK * t.TI && gi < g.nnx && gj < g.nny && gk < g.nnz;
    T y0 = T(0), y1 = T(0), y2 = T(0);
    {
        const int ei = slot >> 2, ek = (slot >> 1) & 1, ej = slot & 1;
        const int ex = gi - 1 + ei, ey = gj - 1 + ej, ez = gk - 1 + ek;
        if (node_ok && ex >= 0 && ex < g.nx && ey >= 0 && ey < g.ny && ez >= 0 && ez < g.nz) {
            const long nel = g.nel;
            const T *__restrict__ Ae = a.A + ((long)g.ny * g.nz * ex + (long)g.ny * ez + ey);
            const int la = lidx(1 - ei, 1 - ej, 1 - ek);
            // all loads of a pass (the nodal values and the 3 x 3 blocks of
            // NB nodes) are issued before the products: one memory latency
            // per pass instead of one per node; FP64 takes two passes of 4
            // nodes to bound its registers
            constexpr int NB = sizeof(T) == 8 ? 4 : 8;
#pragma unroll
            for (int l0 = 0; l0 < 8; l0 += NB) {
                T uu[NB][3], av[NB][3][3];
#pragma unroll
                for (int q = 0; q < NB; ++q) {
                    const int lb = l0 + q;
                    const int bx = (lb == 1 || lb == 2 || lb == 5 || lb == 6) ? 1 : 0;
                    const int by = (lb == 2 || lb == 3 || lb == 6 || lb == 7) ? 1 : 0;
                    const int bz = lb >= 4 ? 1 : 0;
                    const long dof = 3 * ((long)g.nny * g.nnz * (ex + bx) + (long)g.nny * (ez + bz) + (ey + by));
#pragma unroll
                    for (int c = 0; c < 3; ++c) uu[q][c] = stored_u<T, MODE>(a, dof + c);
#pragma unroll
                    for (int r = 0; r < 3; ++r)
#pragma unroll
                        for (int c = 0; c < 3; ++c) av[q][r][c] = Ae[pidx(3 * la + r, 3 * lb + c) * nel];
                }
#pragma unroll
                for (int q = 0; q < NB; ++q) {
                    y0 += av[q][0][0] * uu[q][0] + av[q][0][1] * uu[q][1] + av[q][0][2] * uu[q][2];
                    y1 += av[q][1][0] * uu[q][0] + av[q][1][1] * uu[q][1] + av[q][1][2] * uu[q][2];
                    y2 += av[q][2][0] * uu[q][0] + av[q][2][1] * uu[q][1] + av[q][2][2] * uu[q][2];
                }
            }
        }
    }
    red[slot][0][n] = y0;
    red[slot][1][n] = y1;
    red[slot][2][n] = y2;
    __syncthreads();
    double part = 0.0;
    if (slot == 0 && node_ok) {
        const long nod = (long)g.nny * g.nnz * gi + (long)g.nny * gk + gj;
        const long dof = 3 * nod;
        const unsigned m = a.mask[nod];
        const T om = T(a.omega), one = T(1);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T yc = red[0][c][n];
#pragma unroll
            for (int q = 1; q < 8; ++q) yc += red[q][c][n];
            const bool fixed = (m >> c) & 1u;
            const T u = stored_u<T, MODE>(a, dof + c);
            if (MODE == OP_MATVEC) {
                const T v = fixed ? T(0) : yc;
                a.out[dof + c] = v;
                part += (double)u * (double)v;
            } else if (MODE == OP_DEFECT || MODE == OP_PRE_DEFECT) {
                a.out[dof + c] = fixed ? T(0) : a.F[dof + c] - yc;
                if (MODE == OP_PRE_DEFECT) a.out2[dof + c] = u;
            } else {
                const T f = a.F[dof + c];
                const T rv = fixed ? T(0) : f - yc;
                const T ad = a.AD[dof + c];
                const T un = (one - om) * u + om * (rv + ad * u) / ad;
                a.out[dof + c] = un;
                part += (double)f * (double)un;
            }
        }
    }
    if (a.partials) {
        const double s = block_sum(part);
        if (threadIdx.x == 0) a.partials[blockIdx.x] = s;
    }
}

// ---------------------------------------------------------------------------
// The same operator assembled per node: block b = 9 (di+1) + 3 (dk+1) +
// (dj+1) of node (i, k, j) couples it to node (i+di, k+dk, j+dj) and holds
// the sum of the 3x3 sub-blocks of every element containing both nodes.
// Both nodes of a pair sum the same elements in the same order (increasing
// element position), and the element matrices are symmetric, so block
// 26 - b of node p + d is exactly the transpose of block b of node p: only
// the node's own block and the 13 forward blocks (b >= 13) are stored,
// S[(9 (b - 13) + 3 r + c) nnod + node], and a backward block is read
// transposed at the neighbour.  Zero blocks stand for neighbours outside
// the grid.  Built once per setup from the element matrices, applied by
// k_stencil with 81 coalesced matrix loads and 27 neighbour loads per node,
// where the element form reads 576 matrix entries and 192 neighbour values
// per node.
// ---------------------------------------------------------------------------
// one thread per (node, stored block b = 13 + blockIdx.y)
__global__ void k_stencil_build(Grid g, const double *__restrict__ A, double *__restrict__ S)
{
    const long nnod = g.nnod;
    const long nod = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (nod >= nnod) return;
    const int b = 13 + blockIdx.y;
    const int j = (int)(nod % g.nny), k = (int)((nod / g.nny) % g.nnz), i = (int)(nod / ((long)g.nny * g.nnz));
    const int di = b / 9 - 1, dk = (b / 3) % 3 - 1, dj = b % 3 - 1;
    double sb[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) sb[q] = 0.0;
#pragma unroll
    for (int ei = 0; ei < 2; ++ei)
#pragma unroll
        for (int ek = 0; ek < 2; ++ek)
#pragma unroll
            for (int ej = 0; ej < 2; ++ej) {
                const int ex = i - 1 + ei, ez = k - 1 + ek, ey = j - 1 + ej;
                const int bx = di + 1 - ei, bz = dk + 1 - ek, by = dj + 1 - ej;
                if (ex < 0 || ex >= g.nx || ey < 0 || ey >= g.ny || ez < 0 || ez >= g.nz) continue;
                if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
                const long e = (long)g.ny * g.nz * ex + (long)g.ny * ez + ey;
                const int la = lidx(1 - ei, 1 - ej, 1 - ek), lb = lidx(bx, by, bz);
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c) sb[3 * r + c] += A[(long)pidx(3 * la + r, 3 * lb + c) * g.nel + e];
            }
#pragma unroll
    for (int q = 0; q < 9; ++q) S[(long)(9 * (b - 13) + q) * nnod + nod] = sb[q];
}

void launch_stencil_build(hipStream_t s, const Grid &g, const double *A, double *S)
{
    k_stencil_build<<<dim3((unsigned)((g.nnod + 255) / 256), 14), 256, 0, s>>>(g, A, S);
}

// Three waves per 64 consecutive nodes: wave w sums the 9 blocks of the
// neighbour plane di = w - 1 (FP32: all 108 loads in flight together; FP64:
// one dk row of 36 at a time), the three plane sums meet in LDS and wave 1
// (the node's own plane) adds them in plane order and finishes as k_stored
// does.  On the small levels a launch is a few memory round trips long, so
// every load is issued in the first one: wave 1 fetches its epilogue
// operands (mask, F, AD, own u) with the stencil, and the skip flag is read
// alongside and tested once they are in flight.
// Stencil products of neighbour plane di = W - 1 for one node (lane): the
// blocks b = 9W + 3 (dk+1) + (dj+1), read at the node when b >= 13 and
// transposed at the neighbour (its block 26 - b) when b < 13.  Returns false
// when the solve has converged (the flag is read once the first loads are
// in flight).
template <int W, class T, int MODE>
__device__ __forceinline__ bool stencil_plane(const OpArgsT<T> &a, long p, long pc, int j, int k, long pw, T &y0,
                                              T &y1, T &y2)
{
    const long nnod = a.g.nnod;
    constexpr int NR = sizeof(T) == 8 ? 1 : 3;  // dk rows per batch
#pragma unroll
    for (int r0 = 0; r0 < 3; r0 += NR) {
        T sv[NR][3][9], uu[NR][3][3];
#pragma unroll
        for (int rr = 0; rr < NR; ++rr) {
            const int dk = r0 + rr - 1;
#pragma unroll
            for (int c3 = 0; c3 < 3; ++c3) {
                const int dj = c3 - 1;
                const int kk = k + dk, jj = j + dj;
                const bool ok = kk >= 0 && kk < a.g.nnz && jj >= 0 && jj < a.g.nny;
                // outside the grid the block is zero: read any valid node
                const long nb = ok ? pw + (dk * a.g.nny + dj) : pc;
                const int b = 9 * W + 3 * (r0 + rr) + c3;
                if (b >= 13) {
                    const T *__restrict__ Sb = a.S + (long)(9 * (b - 13)) * nnod + p;
#pragma unroll
                    for (int q = 0; q < 9; ++q) sv[rr][c3][q] = Sb[(long)q * nnod];
                } else {
                    const T *__restrict__ Sb = a.S + (long)(9 * (13 - b)) * nnod + nb;
#pragma unroll
                    for (int r = 0; r < 3; ++r)
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            const T v = Sb[(long)(3 * c + r) * nnod];
                            sv[rr][c3][3 * r + c] = ok ? v : T(0);
                        }
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) uu[rr][c3][c] = stored_u<T, MODE>(a, 3 * nb + c);
            }
        }
        if (r0 == 0 && a.ctl && a.ctl[0]) return false;  // converged: nothing is written
#pragma unroll
        for (int rr = 0; rr < NR; ++rr)
#pragma unroll
            for (int c3 = 0; c3 < 3; ++c3) {
                const T *m = sv[rr][c3];
                const T *u = uu[rr][c3];
                y0 += m[0] * u[0] + m[1] * u[1] + m[2] * u[2];
                y1 += m[3] * u[0] + m[4] * u[1] + m[5] * u[2];
                y2 += m[6] * u[0] + m[7] * u[1] + m[8] * u[2];
            }
    }
    return true;
}

template <class T, int MODE>
__global__ __launch_bounds__(192) void k_stencil(OpArgsT<T> a)
{
    batch_shift(a);
    __shared__ T red[2][3][64];
    const Grid g = a.g;
    const long nnod = g.nnod;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long p = (long)blockIdx.x * 64 + lane;
    const bool live = p < nnod;
    const long pc = live ? p : 0;
    // 32-bit decomposition (a level has far fewer than 2^32 nodes)
    const unsigned pu = (unsigned)pc, q1 = pu / (unsigned)g.nny;
    const int j = (int)(pu - q1 * (unsigned)g.nny), k = (int)(q1 % (unsigned)g.nnz), i = (int)(q1 / (unsigned)g.nnz);
    const int ii = i + w - 1;
    const long pw = pc + (long)(w - 1) * ((long)g.nny * g.nnz);  // the node's image in plane ii
    // epilogue operands of wave 1 (in bounds for every lane: node pc)
    const long dof = 3 * pc;
    unsigned msk = 0;
    T fe[3] = {T(0), T(0), T(0)}, ade[3] = {T(1), T(1), T(1)}, ue[3] = {T(0), T(0), T(0)};
    if (w == 1) {
        msk = a.mask[pc];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (MODE != OP_MATVEC) fe[c] = a.F[dof + c];
            if (MODE != OP_MATVEC && MODE != OP_DEFECT) ade[c] = a.AD[dof + c];
            if (MODE == OP_JACOBI) ue[c] = a.U[dof + c];
        }
    }
    T y0 = T(0), y1 = T(0), y2 = T(0);
    if (live && ii >= 0 && ii < g.nnx) {
        const bool go = w == 0   ? stencil_plane<0, T, MODE>(a, p, pc, j, k, pw, y0, y1, y2)
                        : w == 1 ? stencil_plane<1, T, MODE>(a, p, pc, j, k, pw, y0, y1, y2)
                                 : stencil_plane<2, T, MODE>(a, p, pc, j, k, pw, y0, y1, y2);
        if (!go) return;
    }
    if (a.ctl && a.ctl[0]) return;
    if (w != 1) {
        red[w >> 1][0][lane] = y0;
        red[w >> 1][1][lane] = y1;
        red[w >> 1][2][lane] = y2;
    }
    __syncthreads();
    if (w != 1 || !live) return;
    const T om = T(a.omega), one = T(1);
    const T yv[3] = {y0, y1, y2};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const T yc = red[0][c][lane] + yv[c] + red[1][c][lane];
        const bool fixed = (msk >> c) & 1u;
        const T u = (MODE == OP_PRE_JACOBI || MODE == OP_PRE_DEFECT) ? T(a.prerelax) * fe[c] / ade[c] : ue[c];
        if (MODE == OP_MATVEC) {
            a.out[dof + c] = fixed ? T(0) : yc;
        } else if (MODE == OP_DEFECT || MODE == OP_PRE_DEFECT) {
            a.out[dof + c] = fixed ? T(0) : fe[c] - yc;
            if (MODE == OP_PRE_DEFECT) a.out2[dof + c] = u;
        } else {
            const T rv = fixed ? T(0) : fe[c] - yc;
            a.out[dof + c] = (one - om) * u + om * (rv + ade[c] * u) / ade[c];
        }
    }
}

template <class T>
void launch_opA(hipStream_t s, int mode, const OpArgsT<T> &a)
{
    if (a.S && !a.partials) {
        // the stencil form (its launches carry no fused dot)
        const dim3 sg((unsigned)((a.g.nnod + 63) / 64), 1, batch_dim(a.nb)), sb(192);
        switch (mode) {
        case OP_MATVEC: k_stencil<T, OP_MATVEC><<<sg, sb, 0, s>>>(a); return;
        case OP_DEFECT: k_stencil<T, OP_DEFECT><<<sg, sb, 0, s>>>(a); return;
        case OP_JACOBI: k_stencil<T, OP_JACOBI><<<sg, sb, 0, s>>>(a); return;
        case OP_PRE_JACOBI: k_stencil<T, OP_PRE_JACOBI><<<sg, sb, 0, s>>>(a); return;
        case OP_PRE_DEFECT: k_stencil<T, OP_PRE_DEFECT><<<sg, sb, 0, s>>>(a); return;
        default: return;
        }
    }
    const dim3 grid((unsigned)a.t.blocks(), 1, batch_dim(a.nb)), block(kBlock);
    switch (mode) {
    case OP_MATVEC: k_stored<T, OP_MATVEC><<<grid, block, 0, s>>>(a); break;
    case OP_DEFECT: k_stored<T, OP_DEFECT><<<grid, block, 0, s>>>(a); break;
    case OP_JACOBI: k_stored<T, OP_JACOBI><<<grid, block, 0, s>>>(a); break;
    case OP_PRE_JACOBI: k_stored<T, OP_PRE_JACOBI><<<grid, block, 0, s>>>(a); break;
    case OP_PRE_DEFECT: k_stored<T, OP_PRE_DEFECT><<<grid, block, 0, s>>>(a); break;
    default: break;
    }
}
template void launch_op0<double>(hipStream_t, int, const OpArgsT<double> &);

void launch_op0_uf(hipStream_t s, const OpArgs &a)
{
    const dim3 grid((unsigned)xcd_launch(a.t.blocks()), 1, batch_dim(a.nb)), block(kBlock);
    if (a.kc)
        k_gather<double, OP_MATVEC_UF, true><<<grid, block, 0, s>>>(a);
    else
        k_gather<double, OP_MATVEC_UF, false><<<grid, block, 0, s>>>(a);
}
template void launch_op0<float>(hipStream_t, int, const OpArgsT<float> &);
template void launch_opA<double>(hipStream_t, int, const OpArgsT<double> &);
template void launch_opA<float>(hipStream_t, int, const OpArgsT<float> &);

// ---------------------------------------------------------------------------
// Operator diagonals (prepAD0c / prepADc, mgcg1.cu:1404-1487): one thread
// per node, sum over its incident elements.
// ---------------------------------------------------------------------------
template <bool STORED>
__global__ __launch_bounds__(kBlock) void k_diag(Grid g, const double *__restrict__ Ke, const double *__restrict__ CX,
                                                const double *__restrict__ A, double *__restrict__ AD)
{
    const long nod = (long)blockIdx.x * kBlock + threadIdx.x;
    if (nod >= g.nnod) return;
    const int j = (int)(nod % g.nny);
    const int k = (int)((nod / g.nny) % g.nnz);
    const int i = (int)(nod / ((long)g.nny * g.nnz));
    double d[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int ei = 0; ei < 2; ++ei)
#pragma unroll
        for (int ek = 0; ek < 2; ++ek)
#pragma unroll
            for (int ej = 0; ej < 2; ++ej) {
                const int ex = i - 1 + ei, ey = j - 1 + ej, ez = k - 1 + ek;
                if (ex < 0 || ex >= g.nx || ey < 0 || ey >= g.ny || ez < 0 || ez >= g.nz) continue;
                const long e = (long)g.ny * g.nz * ex + (long)g.ny * ez + ey;
                const int la = lidx(1 - ei, 1 - ej, 1 - ek);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const int m = 3 * la + c;
                    d[c] += STORED ? A[pidx(m, m) * g.nel + e] : CX[e] * Ke[25 * m];
                }
            }
    AD[3 * nod + 0] = d[0];
    AD[3 * nod + 1] = d[1];
    AD[3 * nod + 2] = d[2];
}

void launch_diag0(hipStream_t s, const Grid &g, const double *Ke, const double *CX, double *AD)
{
    k_diag<false><<<(unsigned)((g.nnod + kBlock - 1) / kBlock), kBlock, 0, s>>>(g, Ke, CX, nullptr, AD);
}

void launch_diagA(hipStream_t s, const Grid &g, const double *A, double *AD)
{
    k_diag<true><<<(unsigned)((g.nnod + kBlock - 1) / kBlock), kBlock, 0, s>>>(g, nullptr, nullptr, A, AD);
}

// ---------------------------------------------------------------------------
// Restriction F_c = P^T R_f, F_c[fixed] = 0 (rstrFc + ZeroBCc,
// mgcg1.cu:792-831): one thread per coarse node gathers the 27 fine nodes
// around (2I,2J,2K) with weight 2^-(|di|+|dj|+|dk|).
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(kBlock) void k_restrict(Grid gc, const T *__restrict__ Rf, T *__restrict__ Fc,
                                                    const uint8_t *__restrict__ maskc, const int *ctl)
{
    ctl = batch_ctl(ctl);
    const long nod = (long)blockIdx.x * kBlock + threadIdx.x;
    if (nod >= gc.nnod) return;
    const unsigned nu = (unsigned)nod, q1 = nu / (unsigned)gc.nny;  // 32-bit: far fewer than 2^32 nodes
    const int J = (int)(nu - q1 * (unsigned)gc.nny);
    const int K = (int)(q1 % (unsigned)gc.nnz);
    const int I = (int)(q1 / (unsigned)gc.nnz);
    const int fny = 2 * gc.ny + 1, fnz = 2 * gc.nz + 1, fnx = 2 * gc.nx + 1;
    const long fc = (long)fny * fnz * (2 * I) + (long)fny * (2 * K) + 2 * J;  // the fine node at (2I,2K,2J)
    Rf = batch_ptr(Rf, 3L * fnx * fny * fnz);
    Fc = batch_ptr(Fc, gc.ndof);
    // all 27 fine nodes' loads are issued before the first product; nodes
    // outside the mesh read node 0 with weight 0 (no branches around loads)
    T v[27][3], w[27];
#pragma unroll
    for (int di = -1; di <= 1; ++di)
#pragma unroll
        for (int dk = -1; dk <= 1; ++dk)
#pragma unroll
            for (int dj = -1; dj <= 1; ++dj) {
                const int q = 9 * (di + 1) + 3 * (dk + 1) + (dj + 1);
                const int iP = 2 * I + di, kP = 2 * K + dk, jP = 2 * J + dj;
                const bool ok = iP >= 0 && iP < fnx && kP >= 0 && kP < fnz && jP >= 0 && jP < fny;
                const long f = ok ? 3 * (fc + ((long)di * fny * fnz + dk * fny + dj)) : 0;
                w[q] = ok ? T(1) / (T)(1 << (abs(di) + abs(dj) + abs(dk))) : T(0);
                v[q][0] = Rf[f];
                v[q][1] = Rf[f + 1];
                v[q][2] = Rf[f + 2];
            }
    const unsigned m = maskc[nod];
    if (ctl && ctl[0]) return;  // read while the loads above are in flight
    T s0 = T(0), s1 = T(0), s2 = T(0);
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        s0 += w[q] * v[q][0];
        s1 += w[q] * v[q][1];
        s2 += w[q] * v[q][2];
    }
    Fc[3 * nod + 0] = (m & 1u) ? T(0) : s0;
    Fc[3 * nod + 1] = (m & 2u) ? T(0) : s1;
    Fc[3 * nod + 2] = (m & 4u) ? T(0) : s2;
}

template <class T>
void launch_restrict(hipStream_t s, const Grid &gc, const T *Rf, T *Fc, const uint8_t *maskc, const int *ctl, int nb)
{
    k_restrict<T><<<dim3((unsigned)((gc.nnod + kBlock - 1) / kBlock), 1, batch_dim(nb)), kBlock, 0, s>>>(gc, Rf, Fc,
                                                                                                       maskc, ctl);
}
template void launch_restrict<double>(hipStream_t, const Grid &, const double *, double *, const uint8_t *, const int *,
                                      int);
template void launch_restrict<float>(hipStream_t, const Grid &, const float *, float *, const uint8_t *, const int *,
                                     int);

// ---------------------------------------------------------------------------
// Prolongation U_f += P U_c (interpc, mgcg1.cu:1587-1641) followed by
// U_f[fixed] = 0 (mgcg1.cu:631): one thread per fine node averages the 1,
// 2, 4 or 8 coarse nodes around it according to the parity of (i,j,k).
// The sum may go to another vector (Uout = Uin + P U_c; Uout == Uin is the
// in-place form).
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(kBlock) void k_interp(Grid gf, const T *__restrict__ Uc, const T *Uin, T *Uout,
                                                  const uint8_t *__restrict__ maskf, int zero_fixed, const int *ctl)
{
    ctl = batch_ctl(ctl);
    const long nod = (long)blockIdx.x * kBlock + threadIdx.x;
    if (nod >= gf.nnod) return;
    const unsigned nu = (unsigned)nod, q1 = nu / (unsigned)gf.nny;  // 32-bit: far fewer than 2^32 nodes
    const int j = (int)(nu - q1 * (unsigned)gf.nny);
    const int k = (int)(q1 % (unsigned)gf.nnz);
    const int i = (int)(q1 / (unsigned)gf.nnz);
    const int cny = gf.ny / 2 + 1, cnz = gf.nz / 2 + 1;
    Uc = batch_ptr(Uc, 3L * (gf.nx / 2 + 1) * cny * cnz);
    Uin = batch_ptr(Uin, gf.ndof);
    Uout = batch_ptr(Uout, gf.ndof);
    const int ia = i >> 1, ib = (i + 1) >> 1, ja = j >> 1, jb = (j + 1) >> 1, ka = k >> 1, kb = (k + 1) >> 1;
    const T w = T(1) / (T)(1 << ((i & 1) + (j & 1) + (k & 1)));
    // corners (ia|ib, ja|jb, ka|kb), all 8 loaded together; for an even
    // index a == b and the second corner counts 0, so each coarse node
    // enters once
    T v[8][3], on[8];
#pragma unroll
    for (int ci = 0; ci < 2; ++ci)
#pragma unroll
        for (int ck = 0; ck < 2; ++ck)
#pragma unroll
            for (int cj = 0; cj < 2; ++cj) {
                const int q = 4 * ci + 2 * ck + cj;
                const int I = ci ? ib : ia, K = ck ? kb : ka, J = cj ? jb : ja;
                const bool dup = (ci && ib == ia) || (ck && kb == ka) || (cj && jb == ja);
                const long c = 3 * ((long)cny * cnz * I + (long)cny * K + J);
                on[q] = dup ? T(0) : T(1);
                v[q][0] = Uc[c];
                v[q][1] = Uc[c + 1];
                v[q][2] = Uc[c + 2];
            }
    const long d = 3 * nod;
    const T u0 = Uin[d + 0], u1 = Uin[d + 1], u2 = Uin[d + 2];
    const unsigned m = zero_fixed ? maskf[nod] : 0u;
    if (ctl && ctl[0]) return;  // read while the loads above are in flight
    T s0 = T(0), s1 = T(0), s2 = T(0);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        s0 += on[q] * v[q][0];
        s1 += on[q] * v[q][1];
        s2 += on[q] * v[q][2];
    }
    Uout[d + 0] = (m & 1u) ? T(0) : u0 + w * s0;
    Uout[d + 1] = (m & 2u) ? T(0) : u1 + w * s1;
    Uout[d + 2] = (m & 4u) ? T(0) : u2 + w * s2;
}

template <class T>
void launch_interp_to(hipStream_t s, const Grid &gf, const T *Uc, const T *Uin, T *Uout, const uint8_t *maskf,
                      int zero_fixed, const int *ctl, int nb)
{
    k_interp<T><<<dim3((unsigned)((gf.nnod + kBlock - 1) / kBlock), 1, batch_dim(nb)), kBlock, 0, s>>>(
        gf, Uc, Uin, Uout, maskf, zero_fixed, ctl);
}
template void launch_interp_to<double>(hipStream_t, const Grid &, const double *, const double *, double *,
                                       const uint8_t *, int, const int *, int);
template void launch_interp_to<float>(hipStream_t, const Grid &, const float *, const float *, float *,
                                      const uint8_t *, int, const int *, int);

template <class T>
void launch_interp(hipStream_t s, const Grid &gf, const T *Uc, T *Uf, const uint8_t *maskf, int zero_fixed,
                   const int *ctl, int nb)
{
    launch_interp_to<T>(s, gf, Uc, Uf, Uf, maskf, zero_fixed, ctl, nb);
}
template void launch_interp<double>(hipStream_t, const Grid &, const double *, double *, const uint8_t *, int,
                                    const int *, int);
template void launch_interp<float>(hipStream_t, const Grid &, const float *, float *, const uint8_t *, int,
                                   const int *, int);

// ---------------------------------------------------------------------------
// BLAS-1 and reductions.  Dot products are two-stage with a fixed grid of
// kRedBlocks blocks and a fixed per-thread stride, so the summation order
// (and the result) does not depend on timing.  Products and sums are FP64
// for both vector types.
// ---------------------------------------------------------------------------
static unsigned grid_for(long n)
{
    long nb = (n + kBlock - 1) / kBlock;
    if (nb > 4096) nb = 4096;
    if (nb < 1) nb = 1;
    return (unsigned)nb;
}

template <class T>
__global__ void k_fill(T *v, long n, T val)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) v[i] = val;
}

template <class T>
void launch_fill(hipStream_t s, T *v, long n, double val)
{
    k_fill<T><<<grid_for(n), kBlock, 0, s>>>(v, n, T(val));
}
template void launch_fill<double>(hipStream_t, double *, long, double);
template void launch_fill<float>(hipStream_t, float *, long, double);

template <class Ti, class To>
__global__ void k_copy(const Ti *__restrict__ src, To *__restrict__ dst, long n, const int *ctl)
{
    ctl = batch_ctl(ctl);
    if (ctl && ctl[0]) return;
    src = batch_ptr(src, n);
    dst = batch_ptr(dst, n);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = (To)src[i];
}

template <class Ti, class To>
void launch_copy(hipStream_t s, const Ti *src, To *dst, long n, const int *ctl, int nb)
{
    k_copy<Ti, To><<<dim3(grid_for(n), 1, batch_dim(nb)), kBlock, 0, s>>>(src, dst, n, ctl);
}
// U = prerelax F / AD (preRelc, mgcg1.cu:1190-1197), the same expression the
// fused PRE modes form on the fly
template <class T>
__global__ void k_prerelax(const T *__restrict__ F, const T *__restrict__ AD, T *__restrict__ U, long n, T pr,
                           const int *ctl)
{
    ctl = batch_ctl(ctl);
    F = batch_ptr(F, n);
    U = batch_ptr(U, n);
    // two elements per pass, both loaded before either is stored (the
    // memory counter retires in order: a load issued after a store waits
    // for it); the skip flag is read while the first pass's loads fly
    const long stride = (long)gridDim.x * blockDim.x, i0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (long i = i0; i < n; i += 2 * stride) {
        const bool two = i + stride < n;
        const long j = two ? i + stride : i;  // in bounds; unused when !two
        const T f0 = F[i], a0 = AD[i], f1 = F[j], a1 = AD[j];
        if (i == i0 && ctl && ctl[0]) return;
        U[i] = pr * f0 / a0;
        if (two) U[j] = pr * f1 / a1;
    }
}

template <class T>
void launch_prerelax(hipStream_t s, const T *F, const T *AD, T *U, long n, double prerelax, const int *ctl, int nb)
{
    k_prerelax<T><<<dim3(grid_for(n), 1, batch_dim(nb)), kBlock, 0, s>>>(F, AD, U, n, T(prerelax), ctl);
}
template void launch_prerelax<double>(hipStream_t, const double *, const double *, double *, long, double,
                                      const int *, int);
template void launch_prerelax<float>(hipStream_t, const float *, const float *, float *, long, double, const int *,
                                     int);

template void launch_copy<double, double>(hipStream_t, const double *, double *, long, const int *, int);
template void launch_copy<float, float>(hipStream_t, const float *, float *, long, const int *, int);
template void launch_copy<double, float>(hipStream_t, const double *, float *, long, const int *, int);
template void launch_copy<float, double>(hipStream_t, const float *, double *, long, const int *, int);

template <class Ta, class Tb>
__global__ __launch_bounds__(kBlock) void k_dot_partial(const Ta *__restrict__ a, const Tb *__restrict__ b, long n,
                                                       double *partials, const int *ctl)
{
    ctl = batch_ctl(ctl);
    if (ctl && ctl[0]) return;
    a = batch_ptr(a, n);
    b = batch_ptr(b, n);
    partials = batch_ptr(partials, (long)gridDim.x);
    double s = 0.0;
#pragma unroll 4
    for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long)kRedBlocks * kBlock)
        s += (double)a[i] * (double)b[i];
    s = block_sum(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

template <class Ta, class Tb>
void launch_dot_partial(hipStream_t s, const Ta *a, const Tb *b, long n, double *partials, const int *ctl, int nb)
{
    k_dot_partial<Ta, Tb><<<dim3(kRedBlocks, 1, batch_dim(nb)), kBlock, 0, s>>>(a, b, n, partials, ctl);
}
template void launch_dot_partial<double, double>(hipStream_t, const double *, const double *, long, double *,
                                                 const int *, int);
template void launch_dot_partial<float, float>(hipStream_t, const float *, const float *, long, double *, const int *,
                                               int);
template void launch_dot_partial<double, float>(hipStream_t, const double *, const float *, long, double *,
                                                const int *, int);

// Each thread adds p[t], p[t + kBlock], p[t + 2 kBlock], ... in that order;
// the loads of a chunk are all issued before the first add, so a few
// thousand partials cost a round trip per chunk, not one per partial.
__device__ double sum_partials(const double *p, int n)
{
    constexpr int kChunk = 16;
    double s = 0.0;
    for (int i0 = threadIdx.x; i0 < n; i0 += kChunk * kBlock) {
        double v[kChunk];
#pragma unroll
        for (int k = 0; k < kChunk; ++k) v[k] = i0 + k * kBlock < n ? p[i0 + k * kBlock] : 0.0;
#pragma unroll
        for (int k = 0; k < kChunk; ++k)
            if (i0 + k * kBlock < n) s += v[k];
    }
    return block_sum(s);
}

// PCG scalar steps (the scalar arithmetic of mgcg1.cu:235-275, kept on the
// device so the loop never waits on the host).
__global__ __launch_bounds__(kBlock) void k_cg_init(CgScalars *sc, const double *p, int n)
{
    sc += blockIdx.z;
    p += (long)blockIdx.z * n;
    const double ff = sum_partials(p, n);
    if (threadIdx.x == 0) {
        sc->res0 = sqrt(ff);
        sc->done = 0;
        sc->nit = 0;
        sc->it = 0;
        sc->bad = 0;
        sc->relres = 0.0;
        if (!(sc->res0 < INFINITY)) {  // NaN or Inf in F
            sc->bad = 1;
            sc->done = 1;
        } else if (sc->res0 == 0.0) {  // F = 0: U = 0 is exact
            sc->done = 1;
        }
    }
}

// The scalar steps read their CgScalars fields and the partials first and
// test the converged flag after: one memory round trip, not three.
__global__ __launch_bounds__(kBlock) void k_cg_rho0(CgScalars *sc, const double *p, int n)
{
    sc += blockIdx.z;
    p += (long)blockIdx.z * n;
    const double rz = sum_partials(p, n);
    if (sc->done) return;
    if (threadIdx.x == 0) {
        sc->rho = rz;
        sc->rho_p = rz;
    }
}

// alog (optional): the iteration's slot of the alpha log, 0 once converged
__global__ __launch_bounds__(kBlock) void k_cg_alpha(CgScalars *sc, const double *p, int n, double *alog)
{
    sc += blockIdx.z;
    p += (long)blockIdx.z * n;
    const double rho_p = sc->rho_p;
    const double pq = sum_partials(p, n);
    if (sc->done) {
        if (alog && threadIdx.x == 0) *alog = 0.0;
        return;
    }
    if (threadIdx.x == 0) {
        sc->pq = pq;
        sc->alpha = rho_p / pq;
        if (alog) *alog = sc->alpha;
        if (!(fabs(sc->alpha) < INFINITY)) {
            sc->bad = 1;
            sc->done = 1;
            sc->nit = sc->it;
        }
    }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_update_ur(T *__restrict__ U, T *__restrict__ R, const T *__restrict__ P,
                                                        const T *__restrict__ Q, long n, const CgScalars *sc,
                                                        double *partials, float *__restrict__ R32)
{
    sc += blockIdx.z;
    U = batch_ptr(U, n);
    R = batch_ptr(R, n);
    P = batch_ptr(P, n);
    Q = batch_ptr(Q, n);
    R32 = batch_ptr(R32, n);
    partials = batch_ptr(partials, (long)gridDim.x);
    const T al = (T)sc->alpha;
    const long stride = (long)kRedBlocks * kBlock;
    const long i0 = (long)blockIdx.x * kBlock + threadIdx.x;
    double s = 0.0;
    // two elements per pass, both loaded before either is stored
    for (long i = i0; i < n; i += 2 * stride) {
        const bool two = i + stride < n;
        const long j = two ? i + stride : i;  // in bounds; unused when !two
        const T u0 = U[i], p0 = P[i], r0 = R[i], q0 = Q[i];
        const T u1 = U[j], p1 = P[j], r1 = R[j], q1 = Q[j];
        if (i == i0 && sc->done) return;  // read with alpha while the loads are in flight
        const T ra = r0 - al * q0, rb = r1 - al * q1;
        U[i] = u0 + al * p0;
        R[i] = ra;
        if (R32) R32[i] = (float)ra;  // the FP32 V-cycle's right-hand side
        s += (double)ra * (double)ra;
        if (two) {
            U[j] = u1 + al * p1;
            R[j] = rb;
            if (R32) R32[j] = (float)rb;
            s += (double)rb * (double)rb;
        }
    }
    if (sc->done) return;  // every thread tests it before the barrier
    s = block_sum(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_cg_check(CgScalars *sc, const double *p, int n, double tol, int maxit)
{
    sc += blockIdx.z;
    p += (long)blockIdx.z * n;
    const double res0 = sc->res0;
    const double rr = sum_partials(p, n);
    if (sc->done) return;
    if (threadIdx.x == 0) {
        sc->rr = rr;
        sc->relres = sqrt(rr) / res0;
        if (!(sc->relres < INFINITY)) {
            sc->bad = 1;
            sc->done = 1;
            sc->nit = sc->it;
        } else if (sc->relres < tol) {
            sc->done = 1;
            sc->nit = sc->it;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_cg_beta(CgScalars *sc, const double *p, int n, int maxit)
{
    sc += blockIdx.z;
    p += (long)blockIdx.z * n;
    const double rho_p = sc->rho_p;
    const double rz = sum_partials(p, n);
    if (sc->done) return;
    if (threadIdx.x == 0) {
        sc->rho = rz;
        sc->beta = rz / rho_p;
        sc->rho_p = rz;
        sc->it += 1;
        if (sc->it >= maxit) {
            sc->done = 1;
            sc->nit = maxit;
        }
    }
}

// Flexible (Polak-Ribiere) beta for a preconditioner that varies between
// iterations: beta = <z, r - r_old>/rho_old with r - r_old = -alpha q, so
// beta = -alpha <z,q>/rho_old; rho = <r,z> is kept for the next alpha.
template <class TZ>
__global__ __launch_bounds__(kBlock) void k_dot2_partial(const double *__restrict__ r, const TZ *__restrict__ z,
                                                       const double *__restrict__ q, long n, double *partials,
                                                       const int *ctl)
{
    ctl = batch_ctl(ctl);
    if (ctl && ctl[0]) return;
    r = batch_ptr(r, n);
    z = batch_ptr(z, n);
    q = batch_ptr(q, n);
    partials = batch_ptr(partials, 2L * kRedBlocks);
    double a = 0.0, b = 0.0;
#pragma unroll 4
    for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long)kRedBlocks * kBlock) {
        const double zi = (double)z[i];  // exact for an FP32 z
        a += r[i] * zi;
        b += q[i] * zi;
    }
    a = block_sum(a);
    b = block_sum(b);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = a;
        partials[kRedBlocks + blockIdx.x] = b;
    }
}

__global__ __launch_bounds__(kBlock) void k_cg_beta_flex(CgScalars *sc, const double *p, int maxit)
{
    sc += blockIdx.z;
    p += 2L * kRedBlocks * blockIdx.z;
    const double alpha = sc->alpha, rho_p = sc->rho_p;
    const double rz = sum_partials(p, kRedBlocks);
    const double qz = sum_partials(p + kRedBlocks, kRedBlocks);
    if (sc->done) return;
    if (threadIdx.x == 0) {
        sc->rho = rz;
        sc->beta = -alpha * qz / rho_p;
        sc->rho_p = rz;
        sc->it += 1;
        if (sc->it >= maxit) {
            sc->done = 1;
            sc->nit = maxit;
        }
    }
}

template <class TZ>
void launch_cg_beta_flex(hipStream_t s, const double *r, const TZ *z, const double *q, long n, CgScalars *sc,
                         double *partials, int maxit, int nb)
{
    k_dot2_partial<TZ><<<dim3(kRedBlocks, 1, batch_dim(nb)), kBlock, 0, s>>>(r, z, q, n, partials, &sc->done);
    k_cg_beta_flex<<<dim3(1, 1, batch_dim(nb)), kBlock, 0, s>>>(sc, partials, maxit);
}
template void launch_cg_beta_flex<double>(hipStream_t, const double *, const double *, const double *, long,
                                          CgScalars *, double *, int, int);
template void launch_cg_beta_flex<float>(hipStream_t, const double *, const float *, const double *, long,
                                         CgScalars *, double *, int, int);

template <class T, class TZ>
__global__ __launch_bounds__(kBlock) void k_cg_update_p(T *__restrict__ P, const TZ *__restrict__ Z, long n,
                                                       const CgScalars *sc)
{
    sc += blockIdx.z;
    P = batch_ptr(P, n);
    Z = batch_ptr(Z, n);
    const T be = (T)sc->beta;
    // two elements per pass, as k_prerelax; done is read with beta
    const long stride = (long)gridDim.x * kBlock, i0 = (long)blockIdx.x * kBlock + threadIdx.x;
    for (long i = i0; i < n; i += 2 * stride) {
        const bool two = i + stride < n;
        const long j = two ? i + stride : i;  // in bounds; unused when !two
        const T p0 = P[i], p1 = P[j], z0 = (T)Z[i], z1 = (T)Z[j];
        if (i == i0 && sc->done) return;
        P[i] = fma(be, p0, z0);
        if (two) P[j] = fma(be, p1, z1);
    }
}

void launch_cg_init(hipStream_t s, CgScalars *sc, const double *partials, int nparts, int nb)
{
    k_cg_init<<<dim3(1, 1, batch_dim(nb)), kBlock, 0, s>>>(sc, partials, nparts);
}
void launch_cg_rho0(hipStream_t s, CgScalars *sc, const double *partials, int nparts, int nb)
{
    k_cg_rho0<<<dim3(1, 1, batch_dim(nb)), kBlock, 0, s>>>(sc, partials, nparts);
}
void launch_cg_alpha(hipStream_t s, CgScalars *sc, const double *partials, int nparts, int nb)
{
    k_cg_alpha<<<dim3(1, 1, batch_dim(nb)), kBlock, 0, s>>>(sc, partials, nparts, nullptr);
}
void launch_cg_alpha_log(hipStream_t s, CgScalars *sc, const double *partials, int nparts, double *alog)
{
    k_cg_alpha<<<1, kBlock, 0, s>>>(sc, partials, nparts, alog);
}

// Residual update of the PCG with deferred u (u += alpha p is applied later
// from the logged alphas and directions): R -= alpha Q, R32 = R in FP32,
// partials of R.R, two elements per pass as k_cg_update_ur.  PRE: also the
// FP32 V-cycle's prerelaxed level-0 iterate X0 = pr R32 / AD32, the
// expression of k_prerelax on the same FP32 values.
template <bool PRE>
__global__ __launch_bounds__(kBlock) void k_cg_update_r(double *__restrict__ R, const double *__restrict__ Q, long n,
                                                       const CgScalars *sc, double *partials,
                                                       float *__restrict__ R32, const float *__restrict__ AD32,
                                                       float *__restrict__ X0, float pr)
{
    const double al = sc->alpha;
    const long stride = (long)kRedBlocks * kBlock;
    const long i0 = (long)blockIdx.x * kBlock + threadIdx.x;
    double s = 0.0;
    for (long i = i0; i < n; i += 2 * stride) {
        const bool two = i + stride < n;
        const long j = two ? i + stride : i;  // in bounds; unused when !two
        const double r0 = R[i], q0 = Q[i], r1 = R[j], q1 = Q[j];
        float d0 = 0.0f, d1 = 0.0f;
        if (PRE) {
            d0 = AD32[i];
            d1 = AD32[j];
        }
        if (i == i0 && sc->done) return;  // read with alpha while the loads are in flight
        const double ra = r0 - al * q0, rb = r1 - al * q1;
        const float fa = (float)ra, fb = (float)rb;
        R[i] = ra;
        R32[i] = fa;
        if (PRE) X0[i] = pr * fa / d0;
        s += ra * ra;
        if (two) {
            R[j] = rb;
            R32[j] = fb;
            if (PRE) X0[j] = pr * fb / d1;
            s += rb * rb;
        }
    }
    if (sc->done) return;  // every thread tests it before the barrier
    s = block_sum(s);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}
void launch_cg_update_r(hipStream_t s, double *R, const double *Q, long n, CgScalars *sc, double *partials,
                        float *R32, const float *AD32, float *X0, double prerelax)
{
    if (X0)
        k_cg_update_r<true><<<kRedBlocks, kBlock, 0, s>>>(R, Q, n, sc, partials, R32, AD32, X0, (float)prerelax);
    else
        k_cg_update_r<false><<<kRedBlocks, kBlock, 0, s>>>(R, Q, n, sc, partials, R32, nullptr, nullptr, 0.0f);
}

// U += alpha_0 P_0 + ... + alpha_{m-1} P_{m-1}, term by term in that order
// (the same fused multiply-adds as one update per iteration).  Converged
// iterations logged alpha 0; their directions are finite, so they add 0.
// One element per thread, its m + 1 loads all in flight together.
__global__ __launch_bounds__(kBlock) void k_cg_flush_u(double *__restrict__ U, const float *__restrict__ P, long n,
                                                      const double *__restrict__ alog, int m)
{
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float pv[kCgLog];
#pragma unroll
    for (int t = 0; t < kCgLog; ++t) pv[t] = t < m ? P[(long)t * n + i] : 0.0f;
    double u = U[i];
#pragma unroll
    for (int t = 0; t < kCgLog; ++t)
        if (t < m) u = fma(alog[t], (double)pv[t], u);
    U[i] = u;
}
void launch_cg_flush_u(hipStream_t s, double *U, const float *P, long n, const double *alog, int m)
{
    if (m > 0) k_cg_flush_u<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(U, P, n, alog, m);
}
template <class T>
void launch_cg_update_ur(hipStream_t s, T *U, T *R, const T *P, const T *Q, long n, CgScalars *sc, double *partials,
                         float *R32, int nb)
{
    k_cg_update_ur<T><<<dim3(kRedBlocks, 1, batch_dim(nb)), kBlock, 0, s>>>(U, R, P, Q, n, sc, partials, R32);
}
template void launch_cg_update_ur<double>(hipStream_t, double *, double *, const double *, const double *, long,
                                          CgScalars *, double *, float *, int);
template void launch_cg_update_ur<float>(hipStream_t, float *, float *, const float *, const float *, long,
                                         CgScalars *, double *, float *, int);
void launch_cg_check(hipStream_t s, CgScalars *sc, const double *partials, int nparts, double tol, int maxit, int nb)
{
    k_cg_check<<<dim3(1, 1, batch_dim(nb)), kBlock, 0, s>>>(sc, partials, nparts, tol, maxit);
}
void launch_cg_beta(hipStream_t s, CgScalars *sc, const double *partials, int nparts, int maxit, int nb)
{
    k_cg_beta<<<dim3(1, 1, batch_dim(nb)), kBlock, 0, s>>>(sc, partials, nparts, maxit);
}
template <class T>
void launch_cg_update_p(hipStream_t s, T *P, const T *Z, long n, CgScalars *sc, int nb)
{
    k_cg_update_p<T, T><<<dim3(grid_for(n), 1, batch_dim(nb)), kBlock, 0, s>>>(P, Z, n, sc);
}
template void launch_cg_update_p<double>(hipStream_t, double *, const double *, long, CgScalars *, int);
template void launch_cg_update_p<float>(hipStream_t, float *, const float *, long, CgScalars *, int);
// P = Z + beta P with P kept in FP32: formed in FP64 from the FP32 values,
// rounded once (two elements per pass, as k_cg_update_p)
// (Pout = Z + beta Pin; Pout may be Pin)
__global__ __launch_bounds__(kBlock) void k_cg_update_p32(float *Pout, const float *Pin, const float *__restrict__ Z,
                                                         long n, const CgScalars *sc)
{
    const double be = sc->beta;
    const long stride = (long)gridDim.x * kBlock, i0 = (long)blockIdx.x * kBlock + threadIdx.x;
    for (long i = i0; i < n; i += 2 * stride) {
        const bool two = i + stride < n;
        const long j = two ? i + stride : i;  // in bounds; unused when !two
        const float p0 = Pin[i], p1 = Pin[j], z0 = Z[i], z1 = Z[j];
        if (i == i0 && sc->done) return;
        Pout[i] = (float)fma(be, (double)p0, (double)z0);
        if (two) Pout[j] = (float)fma(be, (double)p1, (double)z1);
    }
}
void launch_cg_update_p32(hipStream_t s, float *Pout, const float *Pin, const float *Z, long n, CgScalars *sc)
{
    k_cg_update_p32<<<grid_for(n), kBlock, 0, s>>>(Pout, Pin, Z, n, sc);
}

void launch_cg_update_p(hipStream_t s, double *P, const float *Z, long n, CgScalars *sc, int nb)
{
    k_cg_update_p<double, float><<<dim3(grid_for(n), 1, batch_dim(nb)), kBlock, 0, s>>>(P, Z, n, sc);
}

}  // namespace gputo
