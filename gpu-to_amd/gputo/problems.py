"""Problem set-up of the Cantilever drivers (Cantilever/omp/top3D125mod.m).

All arrays use the reference's numbering: node n = (nely+1)(nelz+1) i +
(nely+1) k + j, DOF 3n + d, fixed DOFs as 1-based doubles.
"""
import numpy as np


def ke_h8_gauss(nu=0.3, h=1.0):
    """24x24 trilinear hex stiffness (E = 1), 2x2x2 Gauss, in the element
    order ie={0,1,1,0,0,1,1,0}, je={0,0,1,1,0,0,1,1}, ke={0,0,0,0,1,1,1,1}
    (Cantilever/omp/mgcg1.c:621-624)."""
    ie = np.array([0, 1, 1, 0, 0, 1, 1, 0])
    je = np.array([0, 0, 1, 1, 0, 0, 1, 1])
    kk = np.array([0, 0, 0, 0, 1, 1, 1, 1])
    sx, sy, sz = 2 * ie - 1, 2 * je - 1, 2 * kk - 1
    c1 = 1.0 / ((1 + nu) * (1 - 2 * nu))
    D = np.zeros((6, 6))
    D[:3, :3] = c1 * nu
    np.fill_diagonal(D[:3, :3], c1 * (1 - nu))
    D[3:, 3:] = np.eye(3) * c1 * (1 - 2 * nu) / 2
    g = 1.0 / np.sqrt(3.0)
    K = np.zeros((24, 24))
    for s in (-g, g):
        for e in (-g, g):
            for t in (-g, g):
                dx = 0.25 * sx * (1 + sy * e) * (1 + sz * t) / h
                dy = 0.25 * sy * (1 + sx * s) * (1 + sz * t) / h
                dz = 0.25 * sz * (1 + sx * s) * (1 + sy * e) / h
                B = np.zeros((6, 24))
                B[0, 0::3], B[1, 1::3], B[2, 2::3] = dx, dy, dz
                B[3, 0::3], B[3, 1::3] = dy, dx
                B[4, 1::3], B[4, 2::3] = dz, dy
                B[5, 0::3], B[5, 2::3] = dz, dx
                K += B.T @ D @ B * (h ** 3 / 8.0)
    return K


def ke_integer_tables():
    """Integer matrices (A, B) with Ke(nu) = (A + nu B) / ((1+nu)(2nu-1) 144)
    for the unit H8 -- the form in which the MATLAB drivers write Ke0
    (top3D125mod.m:50-76) -- derived from the Gauss-point stiffness:
    (1+nu)(1-2nu) Ke(nu) is affine in nu."""
    def khat(nu):
        return (1 + nu) * (1 - 2 * nu) * ke_h8_gauss(nu)
    A = np.rint(-144 * khat(0.0))
    B = np.rint(-144 * (khat(1.0) - khat(0.0)))
    return A, B


def ke0_h8(nu=0.3, h=1.0):
    """Ke0 exactly as the MATLAB drivers build it (top3D125mod.m:50-76):
    1/(1+nu)/(2nu-1)/144 * (A + nu B), evaluated in that order, so entries
    with the same integer pair are bitwise equal.  Times h for an h-cube."""
    A, B = ke_integer_tables()
    K = 1.0 / (1 + nu) / (2 * nu - 1) / 144 * (A + nu * B)
    return K if h == 1.0 else h * K


def node_id(nelx, nely, nelz, i, j, k):
    return (nely + 1) * (nelz + 1) * i + (nely + 1) * k + j


def fixed_x0_face(nelx, nely, nelz):
    """fixed = 1 : 3*(nely+1)*(nelz+1): every DOF of the x = 0 face
    (top3D125mod.m:80), 1-based doubles."""
    return np.arange(1, 3 * (nely + 1) * (nelz + 1) + 1, dtype=np.float64)


def uniform_edge_load(nelx, nely, nelz, total=1.0):
    """Unit total downward load spread over the nodes (i = nelx, k = 0,
    j = 0..nely) on their z-DOFs."""
    F = np.zeros(3 * (nelx + 1) * (nely + 1) * (nelz + 1))
    j = np.arange(nely + 1)
    F[3 * node_id(nelx, nely, nelz, nelx, j, 0) + 2] = -total / (nely + 1)
    return F


def matlab_edge_load(nelx, nely, nelz):
    """F(lcDof) = -sin((0:nely)/nely*pi) on the same z-DOFs
    (top3D125mod.m:78,81)."""
    F = np.zeros(3 * (nelx + 1) * (nely + 1) * (nelz + 1))
    j = np.arange(nely + 1)
    F[3 * node_id(nelx, nely, nelz, nelx, j, 0) + 2] = -np.sin(j / nely * np.pi)
    return F


def cantilever(nelx, nely, nelz, load="uniform"):
    """(Ke0, F, fixed) of the Cantilever drivers."""
    F = uniform_edge_load(nelx, nely, nelz) if load == "uniform" else matlab_edge_load(nelx, nely, nelz)
    return ke0_h8(), F, fixed_x0_face(nelx, nely, nelz)


def perturbed_density(nelx, nely, nelz, volfrac, seed=None, amplitude=0.01):
    """x0 = volfrac (+ U(-amplitude, amplitude) from default_rng(seed))."""
    nel = nelx * nely * nelz
    x = np.full(nel, float(volfrac))
    if seed is not None:
        x += np.random.default_rng(seed).uniform(-amplitude, amplitude, nel)
    return x


def inclusion_cell(n, seed=0, volfrac=0.3, radius=3.0, inclusion=0.25):
    """Initial density of the periodic unit cell (n x n x n elements, element
    e = n*n*i + n*k + j): a seeded U(0, 1) field smoothed with the cone
    filter h = max(0, radius - dist) under periodic wrap, shifted to mean
    volfrac and clipped to [0, 1], with the elements within inclusion*n of
    the cell centre set to 0 (a spherical void inclusion)."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, 1.0, (n, n, n))  # axes (i, k, j)
    r = int(np.ceil(radius)) - 1
    acc = np.zeros_like(x)
    wsum = 0.0
    for di in range(-r, r + 1):
        for dk in range(-r, r + 1):
            for dj in range(-r, r + 1):
                w = max(0.0, radius - np.sqrt(di * di + dj * dj + dk * dk))
                if w > 0.0:
                    acc += w * np.roll(x, (di, dk, dj), axis=(0, 1, 2))
                    wsum += w
    x = acc / wsum
    x = np.clip(x - x.mean() + volfrac, 0.0, 1.0)
    c = np.arange(n) + 0.5 - n / 2.0
    ci, ck, cj = np.meshgrid(c, c, c, indexing="ij")
    x[ci ** 2 + cj ** 2 + ck ** 2 <= (inclusion * n) ** 2] = 0.0
    return x.ravel()


def simp_cx(x, penal=3.0, Emin=1e-9, E0=1.0):
    """E(x) = Emin + x^penal (E0 - Emin)."""
    return Emin + np.asarray(x, dtype=np.float64) ** penal * (E0 - Emin)
