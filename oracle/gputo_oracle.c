/*
 * gputo_oracle.c -- CPU FP64 oracle for the top3D125 MGCG elasticity solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C99 + OpenMP, the
 * algorithm of the reference's OpenMP MEX solver (Cantilever/omp/mgcg1.c) with
 * the GPU deltas of Cantilever/TOP3D125MGCGCUDA/mgcg1.cu, and the driver maths
 * of Cantilever/omp/top3D125mod.m.  Only tests/, the smoke check and the
 * cpu_baseline leg of bench.py load it, as the checker and as the CPU
 * baseline; the product (libgputo.so) never links or calls it.
 *
 * Conventions (all 0-based, C frame):
 *   element id  e   = nely*nelz*i + nely*k + j          (mgcg1.c:644)
 *   node id     n   = (nely+1)*(nelz+1)*i + (nely+1)*k + j (mgcg1.c:645)
 *   dof             = 3*n + d, d = 0,1,2 for x,y,z
 *   element local node order ie/je/ke                     (mgcg1.c:621-624)
 *   fixed DOF lists are 1-based doubles, as MATLAB passes them (mgcg1.c:663).
 *
 * Parity pinning of every function is described at the top of each test that
 * exercises it (tests/test_oracle_*.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const int IE[8] = {0, 1, 1, 0, 0, 1, 1, 0};
static const int JE[8] = {0, 0, 1, 1, 0, 0, 1, 1};
static const int KE[8] = {0, 0, 0, 0, 1, 1, 1, 1};

#define NODE(i, j, k) ((long)(nely + 1) * (nelz + 1) * (i) + (long)(nely + 1) * (k) + (j))
#define ELEM(i, j, k) ((long)nely * nelz * (i) + (long)nely * (k) + (j))

/* ------------------------------------------------------------------ */
/* Element stiffness: defD / Bem3 / Kem3 (mgcg1.c:1356-1537)           */
/* ------------------------------------------------------------------ */
static void defD(double D[6][6], double v)
{
    /* mgcg1.c:1395-1420 */
    double v1 = 1.0 - v, v2 = (1.0 - 2.0 * v) / 2.0, c1 = 0.5 / (1 + v) / v2;
    memset(D, 0, 36 * sizeof(double));
    for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b) D[a][b] = c1 * (a == b ? v1 : v);
        D[3 + a][3 + a] = c1 * v2;
    }
}

static void bem3(double B[6][24], double s, double e, double t, double hx, double hy, double hz)
{
    /* Strain-displacement matrix of the trilinear hex at (s,e,t) in the
     * ie/je/ke local order; rows exx, eyy, ezz, gxy, gyz, gzx
     * (mgcg1.c:1422-1537). */
    memset(B, 0, 144 * sizeof(double));
    for (int a = 0; a < 8; ++a) {
        double sx = IE[a] ? 1.0 : -1.0, sy = JE[a] ? 1.0 : -1.0, sz = KE[a] ? 1.0 : -1.0;
        double nx = 0.25 * sx * (1 + sy * e) * (1 + sz * t) / hx;
        double ny = 0.25 * sy * (1 + sx * s) * (1 + sz * t) / hy;
        double nz = 0.25 * sz * (1 + sx * s) * (1 + sy * e) / hz;
        B[0][3 * a] = nx;
        B[1][3 * a + 1] = ny;
        B[2][3 * a + 2] = nz;
        B[3][3 * a] = ny;
        B[3][3 * a + 1] = nx;
        B[4][3 * a + 1] = nz;
        B[4][3 * a + 2] = ny;
        B[5][3 * a] = nz;
        B[5][3 * a + 2] = nx;
    }
}

/* Gauss-point stiffness B^T D B * w * |J| at one 2x2x2 point, added into K. */
static void gauss_point_add(double K[576], double nu, double s, double e, double t,
                            double hx, double hy, double hz)
{
    double D[6][6], B[6][24], DB[6][24];
    defD(D, nu);
    bem3(B, s, e, t, hx, hy, hz);
    for (int m = 0; m < 6; ++m)
        for (int n = 0; n < 24; ++n) {
            double acc = 0.0;
            for (int l = 0; l < 6; ++l) acc += D[m][l] * B[l][n];
            DB[m][n] = acc;
        }
    for (int m = 0; m < 24; ++m)
        for (int n = 0; n < 24; ++n) {
            double acc = 0.0;
            for (int l = 0; l < 6; ++l) acc += 0.125 * hx * hy * hz * B[l][m] * DB[l][n];
            K[24 * m + n] += acc;
        }
}

/* Kem3 (mgcg1.c:1356-1393): full 2x2x2 Gauss H8 stiffness, C local order. */
void or_kem3(double nu, double hx, double hy, double hz, double *Ke)
{
    const double g[2] = {-1.0 / sqrt(3.0), 1.0 / sqrt(3.0)};
    memset(Ke, 0, 576 * sizeof(double));
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int c = 0; c < 2; ++c) gauss_point_add(Ke, nu, g[a], g[b], g[c], hx, hy, hz);
}

/* Per-Gauss-point matrices of the coarse element used by the level-1
 * rediscretisation: Ael[na] with na = 4*iP + 2*kP + jP, Gauss point
 * (s,e,t) = (g[iP], g[jP], g[kP])  (coarseAe2, mgcg1.c:1539-1584;
 * prepAelm, mgcg1.cu:1998-2034). */
void or_aelm(double nu, double h, double *Ael /* [8][576] */)
{
    const double g[2] = {-1.0 / sqrt(3.0), 1.0 / sqrt(3.0)};
    memset(Ael, 0, 8 * 576 * sizeof(double));
    for (int iP = 0; iP < 2; ++iP)
        for (int kP = 0; kP < 2; ++kP)
            for (int jP = 0; jP < 2; ++jP)
                gauss_point_add(Ael + 576 * (4 * iP + 2 * kP + jP), nu, g[iP], g[jP], g[kP], h, h, h);
}

/* ------------------------------------------------------------------ */
/* Index maps                                                          */
/* ------------------------------------------------------------------ */

/* MATLAB connectivity cMat (1-based DOFs, MATLAB local order)
 * (top3D125mod.m:37-41): cVec = 3*nodeNrs(1:nely,1:nelz,1:nelx)+1 with
 * nodeNrs 1-based, then 24 fixed offsets. */
void or_edofmat_matlab(int nelx, int nely, int nelz, int32_t *cmat /* [nel][24] */)
{
    long s1 = 3L * (nely + 1) * (nelz + 1), s2 = 3L * (nely + 1) * (nelz + 2), s3 = 3L * (nely + 1);
    long off[24] = {0, 1, 2,
                    s1 + 0, s1 + 1, s1 + 2, s1 - 3, s1 - 2, s1 - 1,
                    -3, -2, -1,
                    s3 + 0, s3 + 1, s3 + 2,
                    s2 + 0, s2 + 1, s2 + 2, s2 - 3, s2 - 2, s2 - 1,
                    s3 - 3, s3 - 2, s3 - 1};
    for (int i = 0; i < nelx; ++i)
        for (int k = 0; k < nelz; ++k)
            for (int j = 0; j < nely; ++j) {
                long e = ELEM(i, j, k);
                long nodeNr = NODE(i, j, k) + 1; /* 1-based */
                long cvec = 3 * nodeNr + 1;
                for (int q = 0; q < 24; ++q) cmat[24 * e + q] = (int32_t)(cvec + off[q]);
            }
}

/* C-order element -> node table (mgcg1.c:630-631, 644-647). */
void or_elem_nodes_c(int nelx, int nely, int nelz, int32_t *en /* [nel][8] */)
{
    for (int i = 0; i < nelx; ++i)
        for (int k = 0; k < nelz; ++k)
            for (int j = 0; j < nely; ++j) {
                long e = ELEM(i, j, k);
                for (int a = 0; a < 8; ++a) en[8 * e + a] = (int32_t)NODE(i + IE[a], j + JE[a], k + KE[a]);
            }
}

/* rstrFD (mgcg1.c:532-562): fixed DOFs of the next-coarser level; kept iff
 * the fine node has all-even indices; order follows the fine list. */
int or_rstrfd(const double *FDofsP, double *FDofs, int nelx, int nely, int nelz, int nFDP)
{
    int nelyP = 2 * nely, nelzP = 2 * nelz, cnt = 0;
    for (int n = 0; n < nFDP; ++n) {
        int g = (int)FDofsP[n] - 1;
        int idof = g % 3, nodP = g / 3;
        int iP = nodP / ((nelzP + 1) * (nelyP + 1));
        if (iP % 2) continue;
        int kP = (nodP - (nelzP + 1) * (nelyP + 1) * iP) / (nelyP + 1);
        if (kP % 2) continue;
        int jP = nodP - (nelzP + 1) * (nelyP + 1) * iP - (nelyP + 1) * kP;
        if (jP % 2) continue;
        long nod = NODE(iP / 2, jP / 2, kP / 2);
        FDofs[cnt++] = (double)(3 * nod + idof + 1);
    }
    return cnt;
}

static void zero_bc(double *V, const double *FDofs, int nFD)
{
    for (int n = 0; n < nFD; ++n) V[(long)FDofs[n] - 1] = 0.0;
}

/* ------------------------------------------------------------------ */
/* Operators                                                           */
/* ------------------------------------------------------------------ */

/* Accumulate s * Ae_e * U_e over all elements into R with the 8-colour
 * element sweep of multAV0/defect0 (mgcg1.c:617-665, 775-825).  Ae is the
 * shared 24x24 (row-major) scaled by CX[e] when A == NULL, else A[e]. */
static void accumulate(const double *Ke, const double *A, const double *CX, const double *U,
                       double *R, double sgn, int nelx, int nely, int nelz)
{
    long ne[8];
    for (int n = 0; n < 8; ++n) ne[n] = NODE(IE[n], JE[n], KE[n]);
    for (int k0 = 0; k0 < 2; ++k0)
        for (int j0 = 0; j0 < 2; ++j0)
            for (int i0 = 0; i0 < 2; ++i0) {
#pragma omp parallel for collapse(2) schedule(static)
                for (int k = k0; k < nelz; k += 2)
                    for (int j = j0; j < nely; j += 2)
                        for (int i = i0; i < nelx; i += 2) {
                            long e = ELEM(i, j, k), nod0 = NODE(i, j, k);
                            const double *Ae = A ? A + 576 * e : Ke;
                            double s = A ? sgn : sgn * CX[e];
                            for (int n1 = 0; n1 < 8; ++n1) {
                                long nod1 = nod0 + ne[n1];
                                for (int n2 = 0; n2 < 8; ++n2) {
                                    long nod2 = nod0 + ne[n2];
                                    for (int d = 0; d < 3; ++d) {
                                        double u = U[3 * nod2 + d];
                                        R[3 * nod1 + 0] += s * Ae[24 * (3 * n1 + 0) + 3 * n2 + d] * u;
                                        R[3 * nod1 + 1] += s * Ae[24 * (3 * n1 + 1) + 3 * n2 + d] * u;
                                        R[3 * nod1 + 2] += s * Ae[24 * (3 * n1 + 2) + 3 * n2 + d] * u;
                                    }
                                }
                            }
                        }
            }
}

/* multAV0 (mgcg1.c:617-665): R = K(CX) U, then R[fixed] = 0. */
void or_matvec0(const double *Ke, const double *CX, const double *U, double *R,
                const double *FDofs, int nFD, int nelx, int nely, int nelz)
{
    long ndof = 3L * (nelx + 1) * (nely + 1) * (nelz + 1);
    memset(R, 0, ndof * sizeof(double));
    accumulate(Ke, NULL, CX, U, R, 1.0, nelx, nely, nelz);
    zero_bc(R, FDofs, nFD);
}

/* defect0 (mgcg1.c:775-825): R = F - K(CX) U, then R[fixed] = 0. */
void or_defect0(const double *Ke, const double *F, const double *U, double *R, const double *CX,
                const double *FDofs, int nFD, int nelx, int nely, int nelz)
{
    long ndof = 3L * (nelx + 1) * (nely + 1) * (nelz + 1);
    memcpy(R, F, ndof * sizeof(double));
    accumulate(Ke, NULL, CX, U, R, -1.0, nelx, nely, nelz);
    zero_bc(R, FDofs, nFD);
}

/* multAV / defect with stored element matrices (mgcg1.c:667-715, 827-877). */
void or_matvec_el(const double *A, const double *U, double *R, const double *FDofs, int nFD,
                  int nelx, int nely, int nelz)
{
    long ndof = 3L * (nelx + 1) * (nely + 1) * (nelz + 1);
    memset(R, 0, ndof * sizeof(double));
    accumulate(NULL, A, NULL, U, R, 1.0, nelx, nely, nelz);
    zero_bc(R, FDofs, nFD);
}

void or_defect_el(const double *A, const double *F, const double *U, double *R, const double *FDofs,
                  int nFD, int nelx, int nely, int nelz)
{
    long ndof = 3L * (nelx + 1) * (nely + 1) * (nelz + 1);
    memcpy(R, F, ndof * sizeof(double));
    accumulate(NULL, A, NULL, U, R, -1.0, nelx, nely, nelz);
    zero_bc(R, FDofs, nFD);
}

/* prepAD0 / prepAD (mgcg1.c:879-916, 955-992): operator diagonal. */
void or_prepad(const double *Ke, const double *A, const double *CX, double *AD, int nelx, int nely, int nelz)
{
    long ndof = 3L * (nelx + 1) * (nely + 1) * (nelz + 1);
    memset(AD, 0, ndof * sizeof(double));
    long ne[8];
    for (int n = 0; n < 8; ++n) ne[n] = NODE(IE[n], JE[n], KE[n]);
    for (int k0 = 0; k0 < 2; ++k0)
        for (int j0 = 0; j0 < 2; ++j0)
            for (int i0 = 0; i0 < 2; ++i0) {
#pragma omp parallel for collapse(2) schedule(static)
                for (int k = k0; k < nelz; k += 2)
                    for (int j = j0; j < nely; j += 2)
                        for (int i = i0; i < nelx; i += 2) {
                            long e = ELEM(i, j, k), nod0 = NODE(i, j, k);
                            for (int n = 0; n < 8; ++n) {
                                long nod1 = nod0 + ne[n];
                                for (int d = 0; d < 3; ++d) {
                                    int m = 3 * n + d;
                                    AD[3 * nod1 + d] += A ? A[576 * e + 25 * m] : CX[e] * Ke[25 * m];
                                }
                            }
                        }
            }
}

/* rstrF (mgcg1.c:564-607): F_c = P^T R_f with trilinear full weighting,
 * then F_c[fixed_c] = 0.  (nelx,nely,nelz) are the coarse sizes. */
void or_restrict(const double *RP, double *F, const double *FDofs, int nFD, int nelx, int nely, int nelz)
{
    int nelxP = 2 * nelx, nelyP = 2 * nely, nelzP = 2 * nelz;
    long ndof = 3L * (nelx + 1) * (nely + 1) * (nelz + 1);
    memset(F, 0, ndof * sizeof(double));
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nelz + 1; k++)
        for (int j = 0; j < nely + 1; j++)
            for (int i = 0; i < nelx + 1; i++) {
                long nod = NODE(i, j, k);
                for (int dk = -1; dk <= 1; dk++) {
                    int kP = 2 * k + dk;
                    if (kP < 0 || kP > nelzP) continue;
                    for (int dj = -1; dj <= 1; dj++) {
                        int jP = 2 * j + dj;
                        if (jP < 0 || jP > nelyP) continue;
                        for (int di = -1; di <= 1; di++) {
                            int iP = 2 * i + di;
                            if (iP < 0 || iP > nelxP) continue;
                            double w = 1.0 / (double)(1 << (abs(di) + abs(dj) + abs(dk)));
                            long nodP = (long)(nelzP + 1) * (nelyP + 1) * iP + (long)(nelyP + 1) * kP + jP;
                            for (int d = 0; d < 3; d++) F[3 * nod + d] += w * RP[3 * nodP + d];
                        }
                    }
                }
            }
    zero_bc(F, FDofs, nFD);
}

/* interp (mgcg1.c:1011-1068): U_f += P U_c, 8-colour scatter from the
 * coarse nodes.  (nelx,nely,nelz) are the fine sizes. */
void or_interp(const double *UP, double *U, int nelx, int nely, int nelz)
{
    int nelxP = nelx / 2, nelyP = nely / 2, nelzP = nelz / 2;
    for (int k1 = 0; k1 < 2; ++k1)
        for (int j1 = 0; j1 < 2; ++j1)
            for (int i1 = 0; i1 < 2; ++i1) {
#pragma omp parallel for collapse(2) schedule(static)
                for (int k = k1; k < nelzP + 1; k += 2)
                    for (int j = j1; j < nelyP + 1; j += 2)
                        for (int i = i1; i < nelxP + 1; i += 2) {
                            long nodP = (long)(nelzP + 1) * (nelyP + 1) * i + (long)(nelyP + 1) * k + j;
                            for (int di = -1; di <= 1; di++) {
                                int i0 = 2 * i + di;
                                if (i0 < 0 || i0 > nelx) continue;
                                for (int dk = -1; dk <= 1; dk++) {
                                    int k0 = 2 * k + dk;
                                    if (k0 < 0 || k0 > nelz) continue;
                                    for (int dj = -1; dj <= 1; dj++) {
                                        int j0 = 2 * j + dj;
                                        if (j0 < 0 || j0 > nely) continue;
                                        long nod = NODE(i0, j0, k0);
                                        double w = 1.0 / (double)(1 << (abs(di) + abs(dj) + abs(dk)));
                                        for (int d = 0; d < 3; d++) U[3 * nod + d] += w * UP[3 * nodP + d];
                                    }
                                }
                            }
                        }
            }
}

/* Level-1 rediscretised operator (coarseAe2, mgcg1.c:1539-1610): the
 * coarse element matrix is sum over its 8 children of CXP[child] times the
 * coarse-element Gauss-point stiffness at the matching Gauss point.
 * (nelx,nely,nelz) are the coarse sizes; h the coarse element size. */
void or_coarse_l1(double *A, const double *CXP, double nu, double h, int nelx, int nely, int nelz)
{
    double *Ael = (double *)malloc(8 * 576 * sizeof(double));
    or_aelm(nu, h, Ael);
    int nelyP = 2 * nely, nelzP = 2 * nelz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nelz; k++)
        for (int j = 0; j < nely; j++)
            for (int i = 0; i < nelx; i++) {
                long e = ELEM(i, j, k);
                double *Ae = A + 576 * e;
                memset(Ae, 0, 576 * sizeof(double));
                for (int iP = 0; iP < 2; ++iP)
                    for (int kP = 0; kP < 2; ++kP)
                        for (int jP = 0; jP < 2; ++jP) {
                            int na = 4 * iP + 2 * kP + jP;
                            long eP = (long)nelzP * nelyP * (2 * i + iP) + (long)nelyP * (2 * k + kP) + 2 * j + jP;
                            for (int m = 0; m < 576; ++m) Ae[m] += CXP[eP] * Ael[576 * na + m];
                        }
            }
    free(Ael);
}

/* Trilinear prolongation of one coarse element onto its 27 fine nodes:
 * P[24][81] with fine local index 9*i + 3*k + j (coarseAe3,
 * mgcg1.c:1168-1201; table P1[8][27] at mgcg1.c:1185-1192, P at
 * mgcg1.cu:1653-1676).  Weight = prod over axes of (1 - |2c - f|/2). */
void or_prolong_local(double *P /* [24][81] */)
{
    memset(P, 0, 24 * 81 * sizeof(double));
    for (int a = 0; a < 8; ++a)
        for (int fi = 0; fi < 3; ++fi)
            for (int fk = 0; fk < 3; ++fk)
                for (int fj = 0; fj < 3; ++fj) {
                    double w = (1.0 - 0.5 * abs(2 * IE[a] - fi)) * (1.0 - 0.5 * abs(2 * JE[a] - fj)) *
                               (1.0 - 0.5 * abs(2 * KE[a] - fk));
                    if (w <= 0) continue;
                    int f = 9 * fi + 3 * fk + fj;
                    for (int d = 0; d < 3; ++d) P[81 * (3 * a + d) + 3 * f + d] = w;
                }
}

/* Element Galerkin (coarseAe3, mgcg1.c:1168-1253): A_c = P (sum of the 8
 * children's matrices on the 27-node patch) P^T.  When AP == NULL the
 * children are CX[child]*Ke (the fine level-0 operator).  (nelx,...) are
 * the coarse sizes. */
void or_coarse_galerkin(const double *AP, const double *Ke, const double *CXP, double *A,
                        int nelx, int nely, int nelz)
{
    double P[24 * 81];
    or_prolong_local(P);
    int nelyP = 2 * nely, nelzP = 2 * nelz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nelz; k++)
        for (int j = 0; j < nely; j++)
            for (int i = 0; i < nelx; i++) {
                double *AL = (double *)calloc(81 * 81, sizeof(double));
                double *ALP = (double *)malloc(81 * 24 * sizeof(double));
                long e = ELEM(i, j, k);
                for (int iP = 0; iP < 2; ++iP)
                    for (int kP = 0; kP < 2; ++kP)
                        for (int jP = 0; jP < 2; ++jP) {
                            long eP = (long)nelzP * nelyP * (2 * i + iP) + (long)nelyP * (2 * k + kP) + 2 * j + jP;
                            int nod0 = 9 * iP + 3 * kP + jP;
                            for (int n1 = 0; n1 < 8; ++n1) {
                                int nod1 = nod0 + 9 * IE[n1] + 3 * KE[n1] + JE[n1];
                                for (int n2 = 0; n2 < 8; ++n2) {
                                    int nod2 = nod0 + 9 * IE[n2] + 3 * KE[n2] + JE[n2];
                                    for (int d1 = 0; d1 < 3; ++d1)
                                        for (int d2 = 0; d2 < 3; ++d2) {
                                            int m = 24 * (3 * n1 + d1) + 3 * n2 + d2;
                                            double v = AP ? AP[576 * eP + m] : CXP[eP] * Ke[m];
                                            AL[81 * (3 * nod1 + d1) + 3 * nod2 + d2] += v;
                                        }
                                }
                            }
                        }
                for (int iA = 0; iA < 81; iA++)
                    for (int jA = 0; jA < 24; jA++) {
                        double acc = 0.0;
                        for (int l = 0; l < 81; l++) acc += AL[81 * iA + l] * P[81 * jA + l];
                        ALP[24 * iA + jA] = acc;
                    }
                for (int iA = 0; iA < 24; iA++)
                    for (int jA = 0; jA < 24; jA++) {
                        double acc = 0.0;
                        for (int l = 0; l < 81; l++) acc += ALP[24 * l + jA] * P[81 * iA + l];
                        A[576 * e + 24 * iA + jA] = acc;
                    }
                free(AL);
                free(ALP);
            }
}

/* CorrCsA (mgcg1.c:918-953; mgcg1.cu:1368-1402): diagonal penalty 1e10 on
 * the coarsest level.  mirrored=1 uses the reference's table
 * {5,6,1,2,4,7,0,3}, which puts the penalty on the y-neighbour node
 * (i, j-/+1, k); mirrored=0 uses the node's own local index
 * {6,5,2,1,7,4,3,0} (mgcg1.cu:952). */
void or_corrcsa(double *A, const double *FDofs, int nFD, int nelx, int nely, int nelz, int mirrored)
{
    static const int tmir[8] = {5, 6, 1, 2, 4, 7, 0, 3};
    static const int town[8] = {6, 5, 2, 1, 7, 4, 3, 0};
    const int *tab = mirrored ? tmir : town;
    for (int n = 0; n < nFD; n++) {
        int g = (int)FDofs[n] - 1;
        int idof = g % 3, nod = g / 3;
        int i = nod / ((nelz + 1) * (nely + 1));
        int k = (nod - (nelz + 1) * (nely + 1) * i) / (nely + 1);
        int j = nod - (nelz + 1) * (nely + 1) * i - (nely + 1) * k;
        for (int di = -1; di <= 0; di++)
            for (int dk = -1; dk <= 0; dk++)
                for (int dj = -1; dj <= 0; dj++) {
                    int iP = i + di, jP = j + dj, kP = k + dk;
                    if (iP < 0 || iP > nelx - 1 || jP < 0 || jP > nely - 1 || kP < 0 || kP > nelz - 1) continue;
                    long e = ELEM(iP, jP, kP);
                    int n1 = tab[4 * (di + 1) + 2 * (dk + 1) + dj + 1];
                    A[576 * e + 25 * (3 * n1 + idof)] = 1e10;
                }
    }
}

/* assembA (mgcg1.c:1255-1306): dense row-major coarsest matrix. */
void or_assemble(const double *A, double *AQ, int nelx, int nely, int nelz)
{
    long ndof = 3L * (nelx + 1) * (nely + 1) * (nelz + 1);
    memset(AQ, 0, ndof * ndof * sizeof(double));
    for (int k = 0; k < nelz; ++k)
        for (int j = 0; j < nely; ++j)
            for (int i = 0; i < nelx; ++i) {
                long e = ELEM(i, j, k);
                for (int n1 = 0; n1 < 8; ++n1) {
                    long nod1 = NODE(i + IE[n1], j + JE[n1], k + KE[n1]);
                    for (int n2 = 0; n2 < 8; ++n2) {
                        long nod2 = NODE(i + IE[n2], j + JE[n2], k + KE[n2]);
                        for (int d1 = 0; d1 < 3; ++d1)
                            for (int d2 = 0; d2 < 3; ++d2)
                                AQ[(3 * nod1 + d1) * ndof + 3 * nod2 + d2] += A[576 * e + 24 * (3 * n1 + d1) + 3 * n2 + d2];
                    }
                }
            }
}

/* choldc (mgcg1.c:1694-1717): L in the strict lower triangle, diag in p.
 * Returns the first failing row + 1, or 0 on success. */
int or_choldc(int n, double *a, double *p)
{
    int fail = 0;
    for (int i = 0; i < n; i++)
        for (int j = i; j < n; j++) {
            double sum = a[(long)i * n + j];
            for (int k = i - 1; k >= 0; k--) sum -= a[(long)i * n + k] * a[(long)j * n + k];
            if (i == j) {
                if (sum <= 0.0 && !fail) fail = i + 1;
                p[i] = sqrt(sum);
            } else
                a[(long)j * n + i] = sum / p[i];
        }
    return fail;
}

/* cholsl (mgcg1.c:1719-1737). */
void or_cholsl(int n, const double *a, const double *p, const double *b, double *x)
{
    for (int i = 0; i < n; i++) {
        double sum = b[i];
        for (int k = i - 1; k >= 0; k--) sum -= a[(long)i * n + k] * x[k];
        x[i] = sum / p[i];
    }
    for (int i = n - 1; i >= 0; i--) {
        double sum = x[i];
        for (int k = i + 1; k < n; k++) sum -= a[(long)k * n + i] * x[k];
        x[i] = sum / p[i];
    }
}

/* ------------------------------------------------------------------ */
/* MGCG                                                                */
/* ------------------------------------------------------------------ */

/* formCXc (Homogenization/QNC/mgcg9.cu:753-789): CX_c = 2 * mean of the 8
 * children (nelx.. are the coarse sizes). */
void or_coarse_cx(const double *CXP, double *CX, int nelx, int nely, int nelz)
{
    int nelyP = 2 * nely, nelzP = 2 * nelz;
    for (int i = 0; i < nelx; ++i)
        for (int k = 0; k < nelz; ++k)
            for (int j = 0; j < nely; ++j) {
                double s = 0.0;
                for (int iP = 0; iP < 2; ++iP)
                    for (int kP = 0; kP < 2; ++kP)
                        for (int jP = 0; jP < 2; ++jP)
                            s += CXP[(long)nelzP * nelyP * (2 * i + iP) + (long)nelyP * (2 * k + kP) + 2 * j + jP];
                CX[(long)nelz * nely * i + (long)nely * k + j] = 2.0 * (s / 8);
            }
}

/* CorrCsAQc (Homogenization/QNC/mgcg9.cu:1005-1028): rows and columns of
 * the fixed DOFs zeroed, 1e10 on their diagonal. */
void or_zero_rowcol(double *AQ, long n, const double *FDofs, int nFD)
{
    for (int f = 0; f < nFD; ++f) {
        long d = (long)FDofs[f] - 1;
        for (long i = 0; i < n; ++i) {
            AQ[d * n + i] = 0.0;
            AQ[i * n + d] = 0.0;
        }
    }
    for (int f = 0; f < nFD; ++f) {
        long d = (long)FDofs[f] - 1;
        AQ[d * n + d] = 1e10;
    }
}

/* Smoother / hierarchy preset.  ref_c follows mgcg1.c + mgcg1.cu
 * (mgcg1.c:110,308,353,363-364; mgcg1.cu:188,1196,1205,631);
 * matlab_cfg0 follows top3D125mgcg.m:190-246 (omega 0.6, one sweep,
 * start from zero) on element-Galerkin coarse operators. */
typedef struct {
    int npre, npost;      /* Jacobi sweeps after the pre-relaxation / after interp */
    double omega;         /* U <- (1-omega) U + omega (R + AD U)/AD */
    double prerelax;      /* U <- prerelax * F / AD on the way down */
    int l1_galerkin;      /* 0: level-1 rediscretised; 1: element Galerkin */
    int zero_after_interp;
    int corr_mirrored;    /* CorrCsA table choice */
    double nu;            /* Poisson ratio of the level-1 rediscretisation */
    int mg9;              /* mgcg9 engine: every level matrix-free with a coarse CX */
} or_params;

/* nswp of the mgcg9 preset (its MEX argument, mgcg9.cu:153) */
static int s_mg9_nswp = 2;
void or_set_nswp(int n) { s_mg9_nswp = n; }

void or_params_preset(int preset, or_params *p)
{
    p->mg9 = 0;
    if (preset == 2) { /* mgcg9 (Homogenization/QNC/mgcg9.cu): omega 0.33 (Jacobc :929-937),
                          preRelc 0.7 (:920-927), nswp sweeps each way (set by the caller) */
        p->npre = s_mg9_nswp; p->npost = s_mg9_nswp; p->omega = 0.33; p->prerelax = 0.7;
        p->l1_galerkin = 0; p->zero_after_interp = 1; p->corr_mirrored = 0; p->mg9 = 1;
        p->nu = 0.3;
        return;
    }
    if (preset == 1) { /* matlab_cfg0 */
        p->npre = 0; p->npost = 1; p->omega = 0.6; p->prerelax = 0.6;
        p->l1_galerkin = 1; p->zero_after_interp = 1; p->corr_mirrored = 0;
    } else { /* ref_c */
        p->npre = 3; p->npost = 3; p->omega = 0.48; p->prerelax = 0.7;
        p->l1_galerkin = 0; p->zero_after_interp = 1; p->corr_mirrored = 1;
    }
    p->nu = 0.3;
}

typedef struct {
    int nelx, nely, nelz;
    long ndof, nel;
    double *U, *F, *R, *AD, *A, *FD;
    int nFD;
    double *AQ, *pdiag; /* coarsest only */
    double *CX, *CXd;   /* mgcg9: level element scales, and the same with
                           scales below 1e-8 dropped (its defect operator) */
    int direct;         /* coarsest level solved by Cholesky */
} or_level;

typedef struct {
    int nl;
    or_level L[16];
    const double *Ke, *CX;
    or_params prm;
} or_hier;

static double dot(const double *a, const double *b, long n)
{
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static void lev_defect(or_hier *H, int l, const double *F, const double *U, double *R)
{
    or_level *v = &H->L[l];
    if (H->prm.mg9)
        or_defect0(H->Ke, F, U, R, v->CXd, v->FD, v->nFD, v->nelx, v->nely, v->nelz);
    else if (l == 0)
        or_defect0(H->Ke, F, U, R, H->CX, v->FD, v->nFD, v->nelx, v->nely, v->nelz);
    else
        or_defect_el(v->A, F, U, R, v->FD, v->nFD, v->nelx, v->nely, v->nelz);
}

static void jacobi(double *U, const double *R, const double *AD, double om, long n)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) U[i] = (1.0 - om) * U[i] + om * (R[i] + AD[i] * U[i]) / AD[i];
}

/* mgcg9 hierarchy (initVcycle, Homogenization/QNC/mgcg9.cu:342-414): every
 * level keeps element scales, CX_l = 2 * mean of the 8 children (formCXc,
 * :753-789, the factor 2 standing for Ke(2h) = 2 Ke(h)), the unit Ke and
 * the diagonal from them (formADc, :1030-1070); the defect skips elements
 * with CX < 1e-8 (defect0c, :939-996) while the CG matvec does not
 * (multAV0c, :791-848).  The coarsest level (not level 0) is factored when
 * it has fewer than 20000 DOFs (maxDofs0, :74): CX*Ke assembled
 * (assembAQc), the rows and columns of its fixed DOFs zeroed and 1e10 put
 * on their diagonal (CorrCsAQc, :1005-1028). */
static int hier_init_mg9(or_hier *H, const double *Ke, const double *CX, const double *FDofs, int nFD,
                         int nelx, int nely, int nelz, int nl)
{
    for (int l = 0; l < nl; ++l) {
        or_level *v = &H->L[l];
        v->nelx = nelx >> l;
        v->nely = nely >> l;
        v->nelz = nelz >> l;
        v->nel = (long)v->nelx * v->nely * v->nelz;
        v->ndof = 3L * (v->nelx + 1) * (v->nely + 1) * (v->nelz + 1);
        v->R = (double *)calloc(v->ndof, sizeof(double));
        v->AD = (double *)calloc(v->ndof, sizeof(double));
        v->CX = (double *)malloc(v->nel * sizeof(double));
        v->CXd = (double *)malloc(v->nel * sizeof(double));
        if (l == 0) {
            memcpy(v->CX, CX, v->nel * sizeof(double));
            v->FD = (double *)malloc((nFD + 1) * sizeof(double));
            memcpy(v->FD, FDofs, nFD * sizeof(double));
            v->nFD = nFD;
        } else {
            or_level *p = &H->L[l - 1];
            v->U = (double *)calloc(v->ndof, sizeof(double));
            v->F = (double *)calloc(v->ndof, sizeof(double));
            v->FD = (double *)malloc((p->nFD + 1) * sizeof(double));
            v->nFD = or_rstrfd(p->FD, v->FD, v->nelx, v->nely, v->nelz, p->nFD);
            or_coarse_cx(p->CX, v->CX, v->nelx, v->nely, v->nelz);
        }
        for (long e = 0; e < v->nel; ++e) v->CXd[e] = v->CX[e] < 1e-8 ? 0.0 : v->CX[e];
        or_prepad(Ke, NULL, v->CX, v->AD, v->nelx, v->nely, v->nelz);
        if (l == nl - 1 && l > 0 && v->ndof < 20000) {
            v->A = (double *)malloc(576 * v->nel * sizeof(double));
            for (long e = 0; e < v->nel; ++e)
                for (int m = 0; m < 576; ++m) v->A[576 * e + m] = v->CX[e] * Ke[m];
            v->AQ = (double *)malloc(v->ndof * v->ndof * sizeof(double));
            v->pdiag = (double *)malloc(v->ndof * sizeof(double));
            or_assemble(v->A, v->AQ, v->nelx, v->nely, v->nelz);
            or_zero_rowcol(v->AQ, v->ndof, v->FD, v->nFD);
            if (or_choldc((int)v->ndof, v->AQ, v->pdiag)) return -2;
            v->direct = 1;
        }
    }
    return 0;
}

/* initVcycle (mgcg1.c:209-281; mgcg1.cu:326-466).  Returns 0, or -2 when
 * the coarsest matrix is not positive definite, -1 when the coarsest
 * dense matrix is too large. */
static int hier_init(or_hier *H, const double *Ke, const double *CX, const double *FDofs, int nFD,
                     int nelx, int nely, int nelz, int nl, const or_params *prm)
{
    memset(H, 0, sizeof(*H));
    H->nl = nl;
    H->Ke = Ke;
    H->CX = CX;
    H->prm = *prm;
    if (prm->mg9) return hier_init_mg9(H, Ke, CX, FDofs, nFD, nelx, nely, nelz, nl);
    for (int l = 0; l < nl; ++l) {
        or_level *v = &H->L[l];
        v->nelx = nelx >> l;
        v->nely = nely >> l;
        v->nelz = nelz >> l;
        v->nel = (long)v->nelx * v->nely * v->nelz;
        v->ndof = 3L * (v->nelx + 1) * (v->nely + 1) * (v->nelz + 1);
        v->R = (double *)calloc(v->ndof, sizeof(double));
        v->AD = (double *)calloc(v->ndof, sizeof(double));
        if (l == 0) {
            v->FD = (double *)malloc((nFD + 1) * sizeof(double));
            memcpy(v->FD, FDofs, nFD * sizeof(double));
            v->nFD = nFD;
            or_prepad(Ke, NULL, CX, v->AD, v->nelx, v->nely, v->nelz);
            if (nl > 1) continue;
        } else {
            or_level *p = &H->L[l - 1];
            v->U = (double *)calloc(v->ndof, sizeof(double));
            v->F = (double *)calloc(v->ndof, sizeof(double));
            v->A = (double *)malloc(576 * v->nel * sizeof(double));
            v->FD = (double *)malloc((p->nFD + 1) * sizeof(double));
            v->nFD = or_rstrfd(p->FD, v->FD, v->nelx, v->nely, v->nelz, p->nFD);
            if (l == 1 && !prm->l1_galerkin)
                or_coarse_l1(v->A, CX, prm->nu, 2.0, v->nelx, v->nely, v->nelz);
            else
                or_coarse_galerkin(l == 1 ? NULL : p->A, Ke, CX, v->A, v->nelx, v->nely, v->nelz);
        }
        if (l == nl - 1) {
            if (v->ndof > 20000) return -1;
            double *Ac = v->A;
            if (l == 0) { /* single level: element matrices CX*Ke */
                Ac = (double *)malloc(576 * v->nel * sizeof(double));
                for (long e = 0; e < v->nel; ++e)
                    for (int m = 0; m < 576; ++m) Ac[576 * e + m] = CX[e] * Ke[m];
                v->A = Ac;
            }
            or_corrcsa(v->A, v->FD, v->nFD, v->nelx, v->nely, v->nelz, prm->corr_mirrored);
            v->AQ = (double *)malloc(v->ndof * v->ndof * sizeof(double));
            v->pdiag = (double *)malloc(v->ndof * sizeof(double));
            or_assemble(v->A, v->AQ, v->nelx, v->nely, v->nelz);
            if (or_choldc((int)v->ndof, v->AQ, v->pdiag)) return -2;
        }
        if (l > 0) or_prepad(NULL, v->A, NULL, v->AD, v->nelx, v->nely, v->nelz);
    }
    return 0;
}

static void hier_free(or_hier *H)
{
    for (int l = 0; l < H->nl; ++l) {
        or_level *v = &H->L[l];
        free(v->R); free(v->AD); free(v->FD);
        if (l > 0) { free(v->U); free(v->F); }
        free(v->A); free(v->AQ); free(v->pdiag);
        free(v->CX); free(v->CXd);
    }
}

/* vcycle (mgcg1.c:283-447; mgcg1.cu:468-651): U0 = V(F0). */
static void vcycle(or_hier *H, const double *F0, double *U0)
{
    const or_params *p = &H->prm;
    int nl = H->nl;
    for (int l = 0; l < nl; ++l) {
        or_level *v = &H->L[l];
        const double *F = l == 0 ? F0 : v->F;
        double *U = l == 0 ? U0 : v->U;
        if (l > 0) or_restrict(H->L[l - 1].R, v->F, v->FD, v->nFD, v->nelx, v->nely, v->nelz);
        if (p->mg9 && l == nl - 1 && l > 0 && !v->direct) {
            /* coarsest level too large to factor: 10 nswp sweeps from the
               level's previous iterate (mgcg9.cu:553-563) */
            for (int it = 0; it < 10 * p->npre; ++it) {
                lev_defect(H, l, F, U, v->R);
                jacobi(U, v->R, v->AD, p->omega, v->ndof);
            }
        } else if (l < nl - 1 || (p->mg9 && l == 0)) {
#pragma omp parallel for schedule(static)
            for (long n = 0; n < v->ndof; ++n) U[n] = p->prerelax * F[n] / v->AD[n];
            for (int it = 0; it < p->npre; ++it) {
                lev_defect(H, l, F, U, v->R);
                jacobi(U, v->R, v->AD, p->omega, v->ndof);
            }
            lev_defect(H, l, F, U, v->R);
        } else {
            or_cholsl((int)v->ndof, v->AQ, v->pdiag, F, U);
            if (!p->mg9) zero_bc(U, v->FD, v->nFD);
        }
    }
    for (int l = nl - 2; l >= 0; --l) {
        or_level *v = &H->L[l];
        const double *F = l == 0 ? F0 : v->F;
        double *U = l == 0 ? U0 : v->U;
        or_interp(H->L[l + 1].U, U, v->nelx, v->nely, v->nelz);
        if (p->zero_after_interp) zero_bc(U, v->FD, v->nFD);
        for (int it = 0; it < p->npost; ++it) {
            lev_defect(H, l, F, U, v->R);
            jacobi(U, v->R, v->AD, p->omega, v->ndof);
        }
    }
}

/* mgcg1 (mgcg1.c:107-179; mgcg1.cu:184-291).  U is overwritten (zero
 * initial guess), then its y-components are negated (MATLAB frame).
 * Returns 0 or a negative error code; *nit is the 0-based exit index. */
int or_mgcg(int preset, const double *Ke, const double *F, double *U, const double *CX,
            const double *FDofs, int nFD, int nelx, int nely, int nelz, int nl, double tol,
            int maxit, double *res, int *nit, int flip_y)
{
    or_params prm;
    or_params_preset(preset, &prm);
    long ndof = 3L * (nelx + 1) * (nely + 1) * (nelz + 1);
    or_hier H;
    int rc = hier_init(&H, Ke, CX, FDofs, nFD, nelx, nely, nelz, nl, &prm);
    if (rc) {
        hier_free(&H);
        *res = 0.0;
        *nit = rc;
        return rc;
    }
    H.L[0].U = NULL;
    double *R = (double *)malloc(ndof * sizeof(double));
    double *P = (double *)calloc(ndof, sizeof(double));
    double *Q = (double *)malloc(ndof * sizeof(double));
    double *Z = (double *)calloc(ndof, sizeof(double));
    memset(U, 0, ndof * sizeof(double));
    memcpy(R, F, ndof * sizeof(double));
    double res0 = sqrt(dot(F, F, ndof)), relres = 0.0, rho, rho_p;
    vcycle(&H, R, P);
    memcpy(Z, P, ndof * sizeof(double));
    rho = dot(R, Z, ndof);
    rho_p = rho;
    int itcg;
    for (itcg = 0; itcg < maxit; ++itcg) {
        or_matvec0(Ke, CX, P, Q, FDofs, nFD, nelx, nely, nelz);
        double alpha = rho_p / dot(P, Q, ndof);
#pragma omp parallel for schedule(static)
        for (long n = 0; n < ndof; ++n) {
            U[n] += alpha * P[n];
            R[n] -= alpha * Q[n];
        }
        vcycle(&H, R, Z);
        rho = dot(R, Z, ndof);
        double beta = rho / rho_p;
#pragma omp parallel for schedule(static)
        for (long n = 0; n < ndof; ++n) P[n] = Z[n] + beta * P[n];
        rho_p = rho;
        *res = sqrt(dot(R, R, ndof));
        relres = *res / res0;
        if (relres < tol || itcg >= maxit) break;
    }
    if (flip_y)
        for (long n = 1; n < ndof; n += 3) U[n] = -U[n];
    *res = relres;
    *nit = itcg;
    free(R); free(P); free(Q); free(Z);
    hier_free(&H);
    return 0;
}

/* One V-cycle z = V(r) on a fresh hierarchy (test entry). */
int or_vcycle_once(int preset, const double *Ke, const double *CX, const double *FDofs, int nFD,
                   int nelx, int nely, int nelz, int nl, const double *r, double *z)
{
    or_params prm;
    or_params_preset(preset, &prm);
    or_hier H;
    int rc = hier_init(&H, Ke, CX, FDofs, nFD, nelx, nely, nelz, nl, &prm);
    if (!rc) vcycle(&H, r, z);
    hier_free(&H);
    return rc;
}

/* Level operators of a fresh hierarchy (test entry): copies level l's
 * element matrices (l>=1, with the coarsest penalty when l==nl-1), its
 * diagonal and its fixed list; returns nFD of level l. */
int or_level_data(int preset, const double *Ke, const double *CX, const double *FDofs, int nFD,
                  int nelx, int nely, int nelz, int nl, int l, double *A, double *AD, double *FD)
{
    or_params prm;
    or_params_preset(preset, &prm);
    or_hier H;
    int rc = hier_init(&H, Ke, CX, FDofs, nFD, nelx, nely, nelz, nl, &prm);
    if (rc) { hier_free(&H); return rc; }
    or_level *v = &H.L[l];
    if (A && v->A) memcpy(A, v->A, 576 * v->nel * sizeof(double));
    if (AD) memcpy(AD, v->AD, v->ndof * sizeof(double));
    if (FD) memcpy(FD, v->FD, v->nFD * sizeof(double));
    int n = v->nFD;
    hier_free(&H);
    return n;
}

/* Element scales of level l of a fresh mgcg9 hierarchy (test entry). */
int or_level_cx(const double *Ke, const double *CX, const double *FDofs, int nFD, int nelx, int nely, int nelz,
                int nl, int l, double *CXl)
{
    or_params prm;
    or_params_preset(2, &prm);
    or_hier H;
    int rc = hier_init(&H, Ke, CX, FDofs, nFD, nelx, nely, nelz, nl, &prm);
    if (!rc) memcpy(CXl, H.L[l].CX, H.L[l].nel * sizeof(double));
    hier_free(&H);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Driver maths (top3D125mod.m)                                        */
/* ------------------------------------------------------------------ */

/* Filter kernel h = max(0, rmin - dist) on offsets -R+1..R-1, R=ceil(rmin)
 * (top3D125mod.m:92-95).  Returns the half width w = R-1; h has
 * (2w+1)^3 entries ordered [dx][dz][dy]. */
int or_filter_kernel(double rmin, double *h)
{
    int w = (int)ceil(rmin) - 1;
    int n = 2 * w + 1, q = 0;
    for (int dx = -w; dx <= w; ++dx)
        for (int dz = -w; dz <= w; ++dz)
            for (int dy = -w; dy <= w; ++dy, ++q) {
                double v = rmin - sqrt((double)(dx * dx + dy * dy + dz * dz));
                h[q] = v > 0 ? v : 0.0;
            }
    (void)n;
    return w;
}

static inline int mirror(int a, int n)
{
    /* 'symmetric' padding: -1 -> 0, -2 -> 1, n -> n-1, n+1 -> n-2 */
    while (a < 0 || a >= n) a = a < 0 ? -a - 1 : 2 * n - a - 1;
    return a;
}

/* imfilter(v, h, 'symmetric') on a (nely, nelz, nelx) column-major array
 * = element order (top3D125mod.m:96,108,140-141): correlation. */
void or_imfilter(const double *in, double *out, int nelx, int nely, int nelz, double rmin)
{
    double h[16 * 16 * 16];
    int w = or_filter_kernel(rmin, h);
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < nelx; ++i)
        for (int k = 0; k < nelz; ++k)
            for (int j = 0; j < nely; ++j) {
                double acc = 0.0;
                int q = 0;
                for (int dx = -w; dx <= w; ++dx)
                    for (int dz = -w; dz <= w; ++dz)
                        for (int dy = -w; dy <= w; ++dy, ++q) {
                            if (h[q] == 0.0) continue;
                            acc += h[q] * in[ELEM(mirror(i + dx, nelx), mirror(j + dy, nely), mirror(k + dz, nelz))];
                        }
                out[ELEM(i, j, k)] = acc;
            }
}

/* Projection helpers (top3D125mod.m:85-88). */
static double prj(double v, double eta, double beta)
{
    return (tanh(beta * eta) + tanh(beta * (v - eta))) / (tanh(beta * eta) + tanh(beta * (1 - eta)));
}
static double deta_f(double v, double eta, double beta)
{
    double sech = 1.0 / cosh(beta * (v - eta));
    return -beta / sinh(beta) * sech * sech * sinh(v * beta) * sinh((1 - v) * beta);
}
static double dprj(double v, double eta, double beta)
{
    double t = tanh(beta * (v - eta));
    return beta * (1 - t * t) / (tanh(beta * eta) + tanh(beta * (1 - eta)));
}

/* eta Newton (top3D125mod.m:110-116): returns eta* with
 * |mean(prj(xPhys)) - volfrac| <= 1e-6; *iters gets the Newton count. */
double or_eta_newton(const double *xPhys, long nel, double eta, double beta, double volfrac, int *iters)
{
    int it = 0;
    double f;
    {
        double s = 0.0;
        for (long e = 0; e < nel; ++e) s += prj(xPhys[e], eta, beta);
        f = s / nel - volfrac;
    }
    while (fabs(f) > 1e-6 && it < 100) {
        double sd = 0.0;
        for (long e = 0; e < nel; ++e) sd += deta_f(xPhys[e], eta, beta);
        eta = eta - f / (sd / nel);
        double s = 0.0;
        for (long e = 0; e < nel; ++e) s += prj(xPhys[e], eta, beta);
        f = s / nel - volfrac;
        ++it;
    }
    if (iters) *iters = it;
    return eta;
}

/* Compliance sensitivity (top3D125mod.m:139) in the MATLAB frame:
 * ce_e = sum((U(cMat)*Ke0).*U(cMat), 2), U 0-based here, cMat 1-based. */
void or_sens_matlab(const double *Ke, const double *U, const int32_t *cmat, long nel, double *ce)
{
#pragma omp parallel for schedule(static)
    for (long e = 0; e < nel; ++e) {
        double ue[24], acc = 0.0;
        for (int q = 0; q < 24; ++q) ue[q] = U[cmat[24 * e + q] - 1];
        for (int c = 0; c < 24; ++c) {
            double t = 0.0;
            for (int r = 0; r < 24; ++r) t += ue[r] * Ke[24 * r + c];
            acc += t * ue[c];
        }
        ce[e] = acc;
    }
}

/* OC update (top3D125mod.m:143-151): x updated in place, returns lmid. */
double or_oc(double *x, const double *dc, const double *dv, long nel, double volfrac, double move, int *iters)
{
    double *ocP = (double *)malloc(nel * sizeof(double));
    double *xT = (double *)malloc(nel * sizeof(double));
    double mean = 0.0;
    for (long e = 0; e < nel; ++e) {
        xT[e] = x[e];
        ocP[e] = x[e] * sqrt(-dc[e] / dv[e]);
        mean += ocP[e];
    }
    mean /= nel;
    double l1 = 0.0, l2 = mean / volfrac, lmid = 0.0;
    int it = 0;
    while ((l2 - l1) / (l2 + l1) > 1e-4) {
        lmid = 0.5 * (l1 + l2);
        double s = 0.0;
        for (long e = 0; e < nel; ++e) {
            double v = ocP[e] / lmid;
            double xu = xT[e] + move, xl = xT[e] - move;
            v = v < xu ? v : xu;
            v = v < 1.0 ? v : 1.0;
            v = v > xl ? v : xl;
            v = v > 0.0 ? v : 0.0;
            x[e] = v;
            s += v;
        }
        if (s / nel > volfrac) l1 = lmid; else l2 = lmid;
        ++it;
    }
    if (iters) *iters = it;
    free(ocP);
    free(xT);
    return lmid;
}

/* One full design iteration of top3D125mod.m:105-165 (without plots and
 * file output) with the MGCG solve in the C frame.  State arrays are
 * updated in place; stats[0..7] = {compliance, mean(xPhys), ch, eta,
 * lmid, relres, nit, seconds}. ft: 1 = density filter, 3 = filter +
 * eta-Newton projection. */
int or_design_iter(int preset, const double *Ke, const double *F, double *U, const double *FDofs, int nFD,
                   int nelx, int nely, int nelz, int nl, double tol, int maxit, double *x, double *xPhys,
                   double *xOld, double volfrac, double penal, double rmin, int ft, double *eta, double beta,
                   double move, double Emin, double E0, double *stats)
{
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    long nel = (long)nelx * nely * nelz;
    long ndof = 3L * (nelx + 1) * (nely + 1) * (nelz + 1);
    double *Hs = (double *)malloc(nel * sizeof(double));
    double *dHs = (double *)malloc(nel * sizeof(double));
    double *ones = (double *)malloc(nel * sizeof(double));
    double *CX = (double *)malloc(nel * sizeof(double));
    double *dsK = (double *)malloc(nel * sizeof(double));
    double *dc = (double *)malloc(nel * sizeof(double));
    double *dv = (double *)malloc(nel * sizeof(double));
    double *tmp = (double *)malloc(nel * sizeof(double));
    int32_t *cmat = (int32_t *)malloc(24 * nel * sizeof(int32_t));
    for (long e = 0; e < nel; ++e) ones[e] = 1.0;
    or_imfilter(ones, Hs, nelx, nely, nelz, rmin);
    memcpy(dHs, Hs, nel * sizeof(double));
    or_imfilter(x, xPhys, nelx, nely, nelz, rmin);
    for (long e = 0; e < nel; ++e) xPhys[e] /= Hs[e];
    if (ft > 1) {
        if (ft == 3) *eta = or_eta_newton(xPhys, nel, *eta, beta, volfrac, NULL);
        for (long e = 0; e < nel; ++e) {
            dHs[e] = Hs[e] / dprj(xPhys[e], *eta, beta);
            xPhys[e] = prj(xPhys[e], *eta, beta);
        }
    }
    double ch = 0.0;
    for (long e = 0; e < nel; ++e) ch += (xPhys[e] - xOld[e]) * (xPhys[e] - xOld[e]);
    ch = sqrt(ch) / nel;
    memcpy(xOld, xPhys, nel * sizeof(double));
    for (long e = 0; e < nel; ++e) {
        dsK[e] = -penal * (E0 - Emin) * pow(xPhys[e], penal - 1);
        CX[e] = Emin + pow(xPhys[e], penal) * (E0 - Emin);
    }
    double res = 0.0;
    int nit = 0;
    int rc = or_mgcg(preset, Ke, F, U, CX, FDofs, nFD, nelx, nely, nelz, nl, tol, maxit, &res, &nit, 1);
    if (rc == 0) {
        or_edofmat_matlab(nelx, nely, nelz, cmat);
        or_sens_matlab(Ke, U, cmat, nel, dc);
        for (long e = 0; e < nel; ++e) {
            dc[e] = dsK[e] * dc[e] / dHs[e];
            tmp[e] = (1.0 / nel / volfrac) / dHs[e];
        }
        or_imfilter(dc, dv, nelx, nely, nelz, rmin); /* dv holds filtered dc */
        memcpy(dc, dv, nel * sizeof(double));
        or_imfilter(tmp, dv, nelx, nely, nelz, rmin);
        double lmid = or_oc(x, dc, dv, nel, volfrac, move, NULL);
        double c = 0.0, vm = 0.0;
        for (long n = 0; n < ndof; ++n) c += F[n] * U[n];
        for (long e = 0; e < nel; ++e) vm += xPhys[e];
        stats[0] = c;
        stats[1] = vm / nel;
        stats[2] = ch;
        stats[3] = *eta;
        stats[4] = lmid;
    }
    stats[5] = res;
    stats[6] = nit;
    clock_gettime(CLOCK_MONOTONIC, &t1);
    stats[7] = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
    free(Hs); free(dHs); free(ones); free(CX); free(dsK); free(dc); free(dv); free(tmp); free(cmat);
    return rc;
}

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_set_num_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}
