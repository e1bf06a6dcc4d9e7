// Periodic unit cell: the fold, image and Jacobi kernels that turn the
// non-periodic operators into the periodic ones, the wrapped assembly of
// the dense coarsest matrix, and the load vectors of the six unit
// macroscopic strains.  The reference has no periodic solve; these follow
// the standard periodic homogenisation of a hex-meshed cell (six load cases
// from the unit strains, fluctuations with one node pinned, the energy
// average of the total strain fields).
//
// Layout: a periodic field keeps the non-periodic (nx+1)(ny+1)(nz+1) node
// grid.  Node (i, j, k) with i < nx, j < ny, k < nz is a master; a node
// with i = nx, j = ny or k = nz is an image of (i mod nx, j mod ny,
// k mod nz).  A master on the face i = 0 (j = 0, k = 0) has images at
// i = nx (j = ny, k = nz) and their combinations: at most 7.
#include <hip/hip_runtime.h>

#include "gputo_internal.h"

namespace gputo {

namespace {

constexpr int kPerBlock = 256;

__device__ __forceinline__ long per_node(const Grid &g, int i, int j, int k)
{
    return (long)g.nny * g.nnz * i + (long)g.nny * k + j;
}

// Number of masters with i == 0, j == 0 or k == 0 (the masters that have
// images), and the t-th of them: first the face i = 0, then j = 0 with
// i >= 1, then k = 0 with i, j >= 1.
__host__ __device__ __forceinline__ long face_masters(const Grid &g)
{
    return (long)g.ny * g.nz + (long)(g.nx - 1) * g.nz + (long)(g.nx - 1) * (g.ny - 1);
}

__device__ __forceinline__ void face_master(const Grid &g, long t, int &i, int &j, int &k)
{
    const long a = (long)g.ny * g.nz, b = (long)(g.nx - 1) * g.nz;
    if (t < a) {
        i = 0;
        j = (int)(t % g.ny);
        k = (int)(t / g.ny);
    } else if (t < a + b) {
        t -= a;
        j = 0;
        i = 1 + (int)(t / g.nz);
        k = (int)(t % g.nz);
    } else {
        t -= a + b;
        k = 0;
        i = 1 + (int)(t / (g.ny - 1));
        j = 1 + (int)(t % (g.ny - 1));
    }
}

// Images of master (i, j, k): up to 7 node indices, in a fixed order.
__device__ __forceinline__ int images_of(const Grid &g, int i, int j, int k, long *img)
{
    int n = 0;
    for (int a = 0; a <= (i == 0); ++a)
        for (int b = 0; b <= (j == 0); ++b)
            for (int c = 0; c <= (k == 0); ++c)
                if (a | b | c) img[n++] = per_node(g, i + a * g.nx, j + b * g.ny, k + c * g.nz);
    return n;
}

template <class T>
__global__ __launch_bounds__(kPerBlock) void k_per_fold(Grid g, T *__restrict__ v, int bcast, const int *ctl)
{
    ctl = batch_ctl(ctl);
    if (ctl && ctl[0]) return;
    v = batch_ptr(v, g.ndof);
    const long t = (long)blockIdx.x * kPerBlock + threadIdx.x;
    if (t >= face_masters(g)) return;
    int i, j, k;
    face_master(g, t, i, j, k);
    long img[7];
    const int n = images_of(g, i, j, k, img);
    const long m = per_node(g, i, j, k);
    for (int d = 0; d < 3; ++d) {
        T s = v[3 * m + d];
        for (int q = 0; q < n; ++q) s += v[3 * img[q] + d];
        v[3 * m + d] = s;
        for (int q = 0; q < n; ++q) v[3 * img[q] + d] = bcast ? s : T(0);
    }
}

template <class T>
__global__ __launch_bounds__(kPerBlock) void k_per_bcast(Grid g, T *__restrict__ v, const int *ctl)
{
    ctl = batch_ctl(ctl);
    if (ctl && ctl[0]) return;
    v = batch_ptr(v, g.ndof);
    const long t = (long)blockIdx.x * kPerBlock + threadIdx.x;
    if (t >= face_masters(g)) return;
    int i, j, k;
    face_master(g, t, i, j, k);
    long img[7];
    const int n = images_of(g, i, j, k, img);
    const long m = per_node(g, i, j, k);
    for (int d = 0; d < 3; ++d) {
        const T s = v[3 * m + d];
        for (int q = 0; q < n; ++q) v[3 * img[q] + d] = s;
    }
}

// One thread per master node (j fastest).  R is the unfolded defect
// F - K U of the non-periodic operator, so its images hold the rest of the
// master's row.
template <class T>
__global__ __launch_bounds__(kPerBlock) void k_per_jacobi(Grid g, T *__restrict__ U, const T *__restrict__ R,
                                                          const T *__restrict__ F, const T *__restrict__ AD,
                                                          const uint8_t *__restrict__ mask, T omega, T prerelax,
                                                          int pre, const int *ctl)
{
    ctl = batch_ctl(ctl);
    if (ctl && ctl[0]) return;
    U = batch_ptr(U, g.ndof);
    R = batch_ptr(R, g.ndof);
    F = batch_ptr(F, g.ndof);
    const long t = (long)blockIdx.x * kPerBlock + threadIdx.x;
    if (t >= (long)g.nx * g.ny * g.nz) return;
    const int j = (int)(t % g.ny);
    const int k = (int)((t / g.ny) % g.nz);
    const int i = (int)(t / ((long)g.ny * g.nz));
    long img[7];
    const int n = (i == 0 || j == 0 || k == 0) ? images_of(g, i, j, k, img) : 0;
    const long m = per_node(g, i, j, k);
    const unsigned fx = mask[m];
    for (int d = 0; d < 3; ++d) {
        const long q0 = 3 * m + d;
        T u;
        if (pre) {
            u = ((fx >> d) & 1u) ? T(0) : prerelax * F[q0] / AD[q0];
        } else {
            T r = R[q0];
            for (int q = 0; q < n; ++q) r += R[3 * img[q] + d];
            u = U[q0] + omega * r / AD[q0];
        }
        U[q0] = u;
        for (int q = 0; q < n; ++q) U[3 * img[q] + d] = u;
    }
}

// Row-owner assembly of the periodic coarsest matrix in the full numbering
// (element indices and column nodes wrap; image rows are unit rows).
// offsets of local node a: x {0,1,1,0,0,1,1,0}, y {0,0,1,1,0,0,1,1}, z {0,0,0,0,1,1,1,1}
__host__ __device__ constexpr int off_x(int a) { return ((a + 1) >> 1) & 1; }
__host__ __device__ constexpr int off_y(int a) { return (a >> 1) & 1; }
__host__ __device__ constexpr int off_z(int a) { return a >> 2; }

__device__ __forceinline__ int local_of(int x, int y, int z) { return 4 * z + (y ? (x ? 2 : 3) : (x ? 1 : 0)); }
__device__ __forceinline__ int packed_of(int m, int n) { return n <= m ? m * (m + 1) / 2 + n : n * (n + 1) / 2 + m; }

__global__ void k_assemble_per(Grid g, const double *__restrict__ A, double *__restrict__ AQ, int N)
{
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    const long nod = r / 3;
    const int c = (int)(r % 3);
    const int j = (int)(nod % g.nny), k = (int)((nod / g.nny) % g.nnz), i = (int)(nod / ((long)g.nny * g.nnz));
    if (r >= g.ndof || i == g.nx || j == g.ny || k == g.nz) {
        AQ[r * N + r] = 1.0;
        return;
    }
    for (int ei = 0; ei < 2; ++ei)
        for (int ek = 0; ek < 2; ++ek)
            for (int ej = 0; ej < 2; ++ej) {
                const int ex = (i - 1 + ei + g.nx) % g.nx;
                const int ey = (j - 1 + ej + g.ny) % g.ny;
                const int ez = (k - 1 + ek + g.nz) % g.nz;
                const long e = (long)g.ny * g.nz * ex + (long)g.ny * ez + ey;
                const int la = local_of(1 - ei, 1 - ej, 1 - ek);
                for (int lb = 0; lb < 8; ++lb) {
                    const long nb = per_node(g, (ex + off_x(lb)) % g.nx, (ey + off_y(lb)) % g.ny,
                                             (ez + off_z(lb)) % g.nz);
                    for (int d = 0; d < 3; ++d)
                        AQ[r * N + 3 * nb + d] += A[(long)packed_of(3 * la + c, 3 * lb + d) * g.nel + e];
                }
            }
}

// Load vectors of the unit strains: F_q at a master node is the sum over
// its 8 (wrapped) elements of CX_e fe[q][3 la + d]; images and fixed DOFs
// get 0.  One thread per node of the full grid.
struct HomLoads {
    double fe[6][24];
};

__global__ __launch_bounds__(kPerBlock) void k_homog_rhs(Grid g, const double *__restrict__ CX, HomLoads L,
                                                         const uint8_t *__restrict__ mask, double *__restrict__ F)
{
    const long nod = (long)blockIdx.x * kPerBlock + threadIdx.x;
    if (nod >= g.nnod) return;
    const int j = (int)(nod % g.nny), k = (int)((nod / g.nny) % g.nnz), i = (int)(nod / ((long)g.nny * g.nnz));
    double s[6][3] = {};
    if (i < g.nx && j < g.ny && k < g.nz) {
        for (int ei = 0; ei < 2; ++ei)
            for (int ek = 0; ek < 2; ++ek)
                for (int ej = 0; ej < 2; ++ej) {
                    const int ex = (i - 1 + ei + g.nx) % g.nx;
                    const int ey = (j - 1 + ej + g.ny) % g.ny;
                    const int ez = (k - 1 + ek + g.nz) % g.nz;
                    const double cx = CX[(long)g.ny * g.nz * ex + (long)g.ny * ez + ey];
                    const int la = local_of(1 - ei, 1 - ej, 1 - ek);
                    for (int q = 0; q < 6; ++q)
                        for (int d = 0; d < 3; ++d) s[q][d] += cx * L.fe[q][3 * la + d];
                }
    }
    const unsigned fx = mask[nod];
    for (int q = 0; q < 6; ++q)
        for (int d = 0; d < 3; ++d) F[q * g.ndof + 3 * nod + d] = ((fx >> d) & 1u) ? 0.0 : s[q][d];
}

}  // namespace

template <class T>
void launch_per_fold(hipStream_t s, const Grid &g, T *v, int bcast, const int *ctl, int nb)
{
    const long n = face_masters(g);
    k_per_fold<T><<<dim3((unsigned)((n + kPerBlock - 1) / kPerBlock), 1, batch_dim(nb)), kPerBlock, 0, s>>>(g, v, bcast,
                                                                                                          ctl);
}
template void launch_per_fold<double>(hipStream_t, const Grid &, double *, int, const int *, int);
template void launch_per_fold<float>(hipStream_t, const Grid &, float *, int, const int *, int);

template <class T>
void launch_per_bcast(hipStream_t s, const Grid &g, T *v, const int *ctl, int nb)
{
    const long n = face_masters(g);
    k_per_bcast<T><<<dim3((unsigned)((n + kPerBlock - 1) / kPerBlock), 1, batch_dim(nb)), kPerBlock, 0, s>>>(g, v, ctl);
}
template void launch_per_bcast<double>(hipStream_t, const Grid &, double *, const int *, int);
template void launch_per_bcast<float>(hipStream_t, const Grid &, float *, const int *, int);

template <class T>
void launch_per_jacobi(hipStream_t s, const Grid &g, T *U, const T *R, const T *F, const T *AD, const uint8_t *mask,
                       double omega, double prerelax, int pre, const int *ctl, int nb)
{
    const long n = (long)g.nx * g.ny * g.nz;
    k_per_jacobi<T><<<dim3((unsigned)((n + kPerBlock - 1) / kPerBlock), 1, batch_dim(nb)), kPerBlock, 0, s>>>(
        g, U, R, F, AD, mask, (T)omega, (T)prerelax, pre, ctl);
}
template void launch_per_jacobi<double>(hipStream_t, const Grid &, double *, const double *, const double *,
                                        const double *, const uint8_t *, double, double, int, const int *, int);
template void launch_per_jacobi<float>(hipStream_t, const Grid &, float *, const float *, const float *,
                                       const float *, const uint8_t *, double, double, int, const int *, int);

void launch_assemble_per(hipStream_t s, const Grid &g, const double *A, double *AQ)
{
    const int N = (int)((g.ndof + 31) / 32 * 32);
    launch_fill(s, AQ, (long)N * N, 0.0);
    k_assemble_per<<<(N + 255) / 256, 256, 0, s>>>(g, A, AQ, N);
}

// fe[q] = Ke0 u0_q for the unit element (x = IE, y = JE, z = KE of each
// local node), Q0[q][p] = u0_q . Ke0 u0_p.
void homog_loads(const double *Ke0, double fe[6][24], double Q0[6][6])
{
    double u0[6][24] = {};
    for (int a = 0; a < 8; ++a) {
        const double x = off_x(a), y = off_y(a), z = off_z(a);
        double *u;
        u = u0[0] + 3 * a; u[0] = x;                       // exx
        u = u0[1] + 3 * a; u[1] = y;                       // eyy
        u = u0[2] + 3 * a; u[2] = z;                       // ezz
        u = u0[3] + 3 * a; u[0] = 0.5 * y; u[1] = 0.5 * x;  // gxy
        u = u0[4] + 3 * a; u[1] = 0.5 * z; u[2] = 0.5 * y;  // gyz
        u = u0[5] + 3 * a; u[0] = 0.5 * z; u[2] = 0.5 * x;  // gzx
    }
    for (int q = 0; q < 6; ++q)
        for (int m = 0; m < 24; ++m) {
            double acc = 0.0;
            for (int n = 0; n < 24; ++n) acc += Ke0[24 * m + n] * u0[q][n];
            fe[q][m] = acc;
        }
    for (int q = 0; q < 6; ++q)
        for (int p = 0; p < 6; ++p) {
            double acc = 0.0;
            for (int m = 0; m < 24; ++m) acc += u0[q][m] * fe[p][m];
            Q0[q][p] = acc;
        }
}

void launch_homog_rhs(hipStream_t s, const Grid &g, const double *CX, const double fe[6][24], const uint8_t *mask,
                      double *F)
{
    HomLoads L;
    for (int q = 0; q < 6; ++q)
        for (int m = 0; m < 24; ++m) L.fe[q][m] = fe[q][m];
    k_homog_rhs<<<(unsigned)((g.nnod + kPerBlock - 1) / kPerBlock), kPerBlock, 0, s>>>(g, CX, L, mask, F);
}

}  // namespace gputo
