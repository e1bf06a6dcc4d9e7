"""Oracle for the periodic unit cell -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and bench.py's CPU baseline may use this
module, and only as the checker; the product path never calls it.

The reference has no periodic solve (Homogenization/QNC/QNC_V1_mod1m.m is a
supported compliance problem on mgcg9), so this restates the textbook
periodic homogenisation of a hex-meshed cell with the reference's element
(Ke0 in the local order ie/je/ke of Cantilever/omp/mgcg1.c:621-624) and a
direct sparse solve:

* periodic node P(i, j, k) = nely*nelz*(i mod nelx) + nely*(k mod nelz)
  + (j mod nely), DOF 3P + d;
* K = sum_e CX_e Ke0 on the wrapped connectivity, node 0 pinned (the null
  space of K is the three rigid translations);
* load case q (unit strain, Voigt xx, yy, zz, xy, yz, zx with engineering
  shears): f_q = sum_e CX_e Ke0 u0_q with u0_q = eps_q . x on the unit
  element's nodes;
* K chi_q = f_q, and CH_ab = 1/|Y| sum_e CX_e (u0_a - chi_a)_e^T Ke0
  (u0_b - chi_b)_e, evaluated element by element (not through the dot
  product identities the device code uses).

Pinned by closed forms in tests/test_periodic_oracle.py: a homogeneous cell
gives CH = E D(nu); a laminate of element layers gives the Backus average
(the FE space holds the exact piecewise-linear solution) in each of the
three layer directions; CH is symmetric and positive definite; translating
the cell leaves CH unchanged and swapping two of its axes permutes CH the
same way.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

IE = np.array([0, 1, 1, 0, 0, 1, 1, 0])
JE = np.array([0, 0, 1, 1, 0, 0, 1, 1])
KE = np.array([0, 0, 0, 0, 1, 1, 1, 1])


def periodic_edofs(nelx, nely, nelz):
    """(nel, 24) periodic DOF indices, element e = nely*nelz*i + nely*k + j."""
    i, k, j = np.meshgrid(np.arange(nelx), np.arange(nelz), np.arange(nely), indexing="ij")
    i, k, j = i.ravel(), k.ravel(), j.ravel()
    nodes = np.stack([nely * nelz * ((i + IE[a]) % nelx) + nely * ((k + KE[a]) % nelz) + (j + JE[a]) % nely
                      for a in range(8)], axis=1)
    return (3 * nodes[:, :, None] + np.arange(3)[None, None, :]).reshape(-1, 24)


def unit_strain_fields():
    """u0[q] (24): nodal displacements eps_q . x of the unit element."""
    u0 = np.zeros((6, 8, 3))
    x, y, z = IE.astype(float), JE.astype(float), KE.astype(float)
    u0[0, :, 0] = x
    u0[1, :, 1] = y
    u0[2, :, 2] = z
    u0[3, :, 0], u0[3, :, 1] = 0.5 * y, 0.5 * x
    u0[4, :, 1], u0[4, :, 2] = 0.5 * z, 0.5 * y
    u0[5, :, 0], u0[5, :, 2] = 0.5 * z, 0.5 * x
    return u0.reshape(6, 24)


def assemble(Ke, CX, nelx, nely, nelz):
    ed = periodic_edofs(nelx, nely, nelz)
    n = 3 * nelx * nely * nelz
    rows = np.repeat(ed, 24, axis=1).ravel()
    cols = np.tile(ed, (1, 24)).ravel()
    vals = (CX[:, None] * np.asarray(Ke, dtype=np.float64).ravel()[None, :]).ravel()
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n)), ed


def homogenize(Ke, CX, nelx, nely, nelz):
    """(CH 6x6, chi 6 x ndof_periodic) with node 0 pinned."""
    Ke = np.asarray(Ke, dtype=np.float64).reshape(24, 24)
    CX = np.asarray(CX, dtype=np.float64)
    K, ed = assemble(Ke, CX, nelx, nely, nelz)
    n = K.shape[0]
    u0 = unit_strain_fields()
    fe = u0 @ Ke.T  # fe[q] = Ke u0_q
    F = np.zeros((6, n))
    for q in range(6):
        np.add.at(F[q], ed.ravel(), (CX[:, None] * fe[q][None, :]).ravel())
    free = np.arange(3, n)
    lu = spla.splu(K[free][:, free].tocsc())
    chi = np.zeros((6, n))
    for q in range(6):
        chi[q, free] = lu.solve(F[q, free])
    CH = np.zeros((6, 6))
    for a in range(6):
        wa = u0[a][None, :] - chi[a][ed]
        for b in range(a, 6):
            wb = u0[b][None, :] - chi[b][ed]
            CH[a, b] = CH[b, a] = np.sum(CX * np.einsum("em,mn,en->e", wa, Ke, wb)) / len(CX)
    return CH, chi


def full_to_periodic(v, nelx, nely, nelz):
    """Master entries of a field on the (nelx+1)(nely+1)(nelz+1)-node layout,
    in the periodic numbering."""
    v = np.asarray(v).reshape(nelx + 1, nelz + 1, nely + 1, 3)
    return v[:nelx, :nelz, :nely, :].ravel()


def isotropic_D(E, nu):
    """6x6 stiffness, Voigt xx, yy, zz, xy, yz, zx, engineering shears."""
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] += 2 * mu
    D[3:, 3:] = mu * np.eye(3)
    return D


def backus_x(E, nu, frac):
    """Exact effective stiffness of isotropic layers normal to x (Backus
    averaging); E, nu, frac are per-layer arrays (volume fractions sum to 1).
    Voigt order xx, yy, zz, xy, yz, zx."""
    E, nu, frac = (np.asarray(a, dtype=np.float64) for a in (E, nu, frac))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    M = lam + 2 * mu

    def avg(v):
        return float(np.sum(frac * v))

    c11 = 1.0 / avg(1.0 / M)
    r = avg(lam / M)
    C = np.zeros((6, 6))
    C[0, 0] = c11
    C[0, 1] = C[1, 0] = C[0, 2] = C[2, 0] = c11 * r
    C[1, 1] = C[2, 2] = avg(4 * mu * (lam + mu) / M) + c11 * r * r
    C[1, 2] = C[2, 1] = avg(2 * mu * lam / M) + c11 * r * r
    C[4, 4] = avg(mu)                 # yz: in the layer plane
    C[3, 3] = C[5, 5] = 1.0 / avg(1.0 / mu)  # xy, zx: across the layers
    return C
