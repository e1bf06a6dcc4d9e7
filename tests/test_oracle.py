"""Pins of the CPU oracle (oracle/gputo_oracle.c) -- CPU only.

What pins each oracle function (chosen so that a dropped term, a wrong sign
or index, or a transposed operand fails one of them):
  or_kem3            reference integer table of Ke0 (tests/golden, from
                     top3D125mod.m:50-76); symmetry, 6 rigid-body modes,
                     Ke(h) = h Ke(1)
  or_prolong_local   reference P1[8][27] table (mgcg1.c:1185-1192)
  or_matvec0/defect0 dense K assembled independently in numpy (brute force),
                     symmetry <Ku,v> = <u,Kv>, K * rigid mode = 0, patch test
  or_prepad          diagonal of the same dense K
  or_restrict/interp restriction is exactly P^T of the prolongation
                     (adjoint identity) and P reproduces linear fields
  or_coarse_l1       uniform density: equals 2 Ke0 (closed form, Ke(2) = 2 Ke(1))
  or_coarse_galerkin uniform density: 2 Ke0 (Q1 nesting is exact); random
                     density: global P^T K P restricted to one coarse element
  or_corrcsa         mirrored table lands on node (i, j-/+1, k)
  or_choldc/cholsl   numpy.linalg.solve
  or_rstrfd          closed form (even-index nodes)
  or_edofmat_matlab  independent numpy construction from top3D125mod.m:37-41
  or_mgcg            dense direct solve of the free DOFs (top3D125.m:125-130)
  or_imfilter        scipy.ndimage.correlate(mode='reflect') (= 'symmetric')
  or_eta_newton      mean(prj(xPhys, eta*)) = volfrac
  or_oc              volume constraint, move limits, bisection bracket
  or_sens_matlab     u^T K u in the C frame (frame invariance)
"""
import json
import os

import numpy as np
import pytest
import scipy.ndimage as ndi

from conftest import ROOT, cantilever_case, dense_K

GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_tables.json")))


def ke_from_table(nu=0.3):
    # evaluated as the MATLAB line does: 1/(1+nu)/(2*nu-1)/144 * (a + nu*b)
    v = 1.0 / (1 + nu) / (2 * nu - 1) / 144 * (np.array(GOLD["ke_a"], float) + nu * np.array(GOLD["ke_nu"], float))
    flat = np.zeros(576)
    lower_colmajor = np.array([(p % 24) >= (p // 24) for p in range(576)])
    flat[lower_colmajor] = v
    K = flat.reshape(24, 24).T  # column-major fill
    return K + K.T - np.diag(np.diag(K))


def rigid_modes(nelx, nely, nelz):
    from gputo.problems import node_id
    ndof = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)
    modes = []
    X = np.zeros((ndof // 3, 3))
    for i in range(nelx + 1):
        for k in range(nelz + 1):
            for j in range(nely + 1):
                X[node_id(nelx, nely, nelz, i, j, k)] = (i, j, k)
    for d in range(3):
        m = np.zeros((ndof // 3, 3))
        m[:, d] = 1
        modes.append(m.ravel())
    for a, b in ((0, 1), (1, 2), (2, 0)):
        m = np.zeros((ndof // 3, 3))
        m[:, a] = -X[:, b]
        m[:, b] = X[:, a]
        modes.append(m.ravel())
    return modes


def test_ke0_matches_reference_table(oracle):
    K = oracle.kem3(0.3, 1.0)
    assert np.abs(K - ke_from_table()).max() < 1e-15
    assert np.allclose(K, K.T, atol=0)
    w = np.linalg.eigvalsh(K)
    assert (np.abs(w) < 1e-12).sum() == 6 and w.min() > -1e-12
    assert np.abs(oracle.kem3(0.3, 2.0) - 2 * K).max() < 1e-14


def test_ke0_python_helper_matches(oracle):
    from gputo.problems import ke0_h8, ke_h8_gauss
    # the integer-table form reproduces the reference's Ke0 bit for bit
    assert np.array_equal(ke0_h8(), ke_from_table())
    assert np.array_equal(ke0_h8(0.2), ke_from_table(0.2))
    assert np.abs(ke_h8_gauss() - ke_from_table()).max() < 1e-15


def test_prolongation_table(oracle):
    P = oracle.prolong_local()
    assert np.array_equal(P[0::3, 0::3], np.array(GOLD["p1"]))
    # block structure: no coupling between components
    assert not P[0::3, 1::3].any() and not P[1::3, 0::3].any()


@pytest.mark.parametrize("mesh", [(2, 2, 2), (4, 2, 2), (3, 2, 5)])
def test_matvec_bruteforce(oracle, mesh):
    nelx, nely, nelz = mesh
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=1)
    K = dense_K(Ke, CX, nelx, nely, nelz)
    u = np.random.default_rng(0).uniform(-1, 1, K.shape[0])
    yd = K @ u
    yd[(fixed - 1).astype(int)] = 0
    y = oracle.matvec0(Ke, CX, u, fixed, nelx, nely, nelz)
    assert np.abs(y - yd).max() < 1e-13 * np.abs(yd).max()
    r = oracle.defect0(Ke, F, u, CX, fixed, nelx, nely, nelz)
    rd = F - K @ u
    rd[(fixed - 1).astype(int)] = 0
    assert np.abs(r - rd).max() < 1e-13 * np.abs(rd).max()
    assert np.abs(oracle.prepad0(Ke, CX, nelx, nely, nelz) - np.diag(K)).max() < 1e-14 * np.diag(K).max()


def test_matvec_symmetry_rigid_patch(oracle):
    nelx, nely, nelz = 4, 4, 2
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=2)
    ndof = F.size
    rng = np.random.default_rng(3)
    u, v = rng.uniform(-1, 1, ndof), rng.uniform(-1, 1, ndof)
    none = np.zeros(0)
    Ku = oracle.matvec0(Ke, CX, u, none, nelx, nely, nelz)
    Kv = oracle.matvec0(Ke, CX, v, none, nelx, nely, nelz)
    assert abs(Ku @ v - u @ Kv) < 1e-12 * abs(Ku @ v)
    for m in rigid_modes(nelx, nely, nelz):
        assert np.abs(oracle.matvec0(Ke, CX, m, none, nelx, nely, nelz)).max() < 1e-12
    # patch test: a linear field with uniform E leaves no interior residual
    from gputo.problems import node_id
    lin = np.zeros(ndof)
    for i in range(nelx + 1):
        for k in range(nelz + 1):
            for j in range(nely + 1):
                n = node_id(nelx, nely, nelz, i, j, k)
                lin[3 * n:3 * n + 3] = (0.3 * i + 0.1 * j, -0.2 * k, 0.5 * j + 0.05 * i)
    y = oracle.matvec0(Ke, np.ones_like(CX), lin, none, nelx, nely, nelz).reshape(-1, 3)
    for i in range(1, nelx):
        for k in range(1, nelz):
            for j in range(1, nely):
                assert np.abs(y[node_id(nelx, nely, nelz, i, j, k)]).max() < 1e-13


def test_transfer_adjoint_and_linear(oracle):
    from gputo.problems import node_id
    fx, fy, fz = 8, 4, 6
    cx, cy, cz = 4, 2, 3
    nf, nc = 3 * (fx + 1) * (fy + 1) * (fz + 1), 3 * (cx + 1) * (cy + 1) * (cz + 1)
    rng = np.random.default_rng(4)
    r, v = rng.uniform(-1, 1, nf), rng.uniform(-1, 1, nc)
    Rr = oracle.restrict(r, np.zeros(0), cx, cy, cz)
    Pv = oracle.interp_add(v, np.zeros(nf), fx, fy, fz)
    assert abs(Rr @ v - r @ Pv) < 1e-12 * abs(Rr @ v)
    lin_c = np.zeros(nc)
    for i in range(cx + 1):
        for k in range(cz + 1):
            for j in range(cy + 1):
                n = node_id(cx, cy, cz, i, j, k)
                lin_c[3 * n:3 * n + 3] = (2 * i + 3 * j, k, 1.0)
    Pl = oracle.interp_add(lin_c, np.zeros(nf), fx, fy, fz).reshape(-1, 3)
    for i in range(fx + 1):
        for k in range(fz + 1):
            for j in range(fy + 1):
                n = node_id(fx, fy, fz, i, j, k)
                assert np.allclose(Pl[n], (i + 1.5 * j, 0.5 * k, 1.0), atol=1e-14)


def test_coarse_operators_uniform(oracle):
    Ke = oracle.kem3()
    cx, cy, cz = 2, 2, 1
    CX = np.full(8 * cx * cy * cz, 0.37)
    A1 = oracle.coarse_l1(CX, cx, cy, cz)
    AG = oracle.coarse_galerkin(cx, cy, cz, Ke=Ke, CX_fine=CX)
    for A in (A1, AG):
        assert np.abs(A - 2 * 0.37 * Ke[None]).max() < 1e-14


def test_galerkin_equals_global_ptkp(oracle):
    from gputo.problems import node_id
    fx, fy, fz = 4, 2, 2
    cx, cy, cz = 2, 1, 1
    Ke = oracle.kem3()
    CX = np.random.default_rng(5).uniform(0.01, 1, fx * fy * fz)
    K = dense_K(Ke, CX, fx, fy, fz)
    nf, nc = K.shape[0], 3 * (cx + 1) * (cy + 1) * (cz + 1)
    P = np.zeros((nf, nc))
    for c in range(nc):
        e = np.zeros(nc)
        e[c] = 1
        P[:, c] = oracle.interp_add(e, np.zeros(nf), fx, fy, fz)
    Kc = P.T @ K @ P
    AG = oracle.coarse_galerkin(cx, cy, cz, Ke=Ke, CX_fine=CX)
    Kc2 = oracle.assemble(AG, cx, cy, cz)
    assert np.abs(Kc - Kc2).max() < 1e-13 * np.abs(Kc).max()
    # a second Galerkin level from stored matrices
    AG2 = oracle.coarse_galerkin(1, 1, 1, A_fine=oracle.coarse_galerkin(2, 2, 2, Ke=Ke, CX_fine=np.ones(64)))
    assert np.abs(AG2[0] - 4 * Ke).max() < 1e-13


def test_corrcsa_mirrored_placement(oracle):
    from gputo.problems import node_id
    nelx, nely, nelz = 2, 2, 2
    A = np.zeros((8, 24, 24))
    i, j, k, d = 1, 1, 1, 2
    dof = 3 * node_id(nelx, nely, nelz, i, j, k) + d
    Am = oracle.corrcsa(A, np.array([dof + 1.0]), nelx, nely, nelz, mirrored=True)
    Ao = oracle.corrcsa(A, np.array([dof + 1.0]), nelx, nely, nelz, mirrored=False)
    K_m = oracle.assemble(Am, nelx, nely, nelz)
    K_o = oracle.assemble(Ao, nelx, nely, nelz)
    hit_m = {int(q) for q in np.nonzero(np.diag(K_m))[0]}
    hit_o = {int(q) for q in np.nonzero(np.diag(K_o))[0]}
    assert hit_o == {dof}
    assert hit_m == {3 * node_id(nelx, nely, nelz, i, j - 1, k) + d, 3 * node_id(nelx, nely, nelz, i, j + 1, k) + d}


def test_cholesky(oracle):
    rng = np.random.default_rng(6)
    M = rng.uniform(-1, 1, (40, 40))
    A = M @ M.T + 40 * np.eye(40)
    b = rng.uniform(-1, 1, 40)
    x, fail = oracle.cholesky_solve(A, b)
    assert fail == 0
    assert np.abs(x - np.linalg.solve(A, b)).max() < 1e-13
    _, fail = oracle.cholesky_solve(-A, b)
    assert fail == 1


def test_rstrfd_closed_form(oracle):
    from gputo.problems import node_id
    nelx, nely, nelz = 8, 4, 4
    fixed = np.array([3 * node_id(nelx, nely, nelz, i, j, k) + d + 1.0
                      for i in (0, 3) for k in range(nelz + 1) for j in range(nely + 1) for d in range(3)])
    out = oracle.rstrfd(fixed, nelx // 2, nely // 2, nelz // 2)
    exp = [3 * node_id(4, 2, 2, i // 2, j // 2, k // 2) + d + 1.0
           for i in (0, 3) for k in range(nelz + 1) for j in range(nely + 1) for d in range(3)
           if i % 2 == 0 and j % 2 == 0 and k % 2 == 0]
    assert out.tolist() == exp


def test_edofmat_closed_form(oracle):
    nelx, nely, nelz = 3, 2, 4
    nodeNrs = np.arange(1, (nelx + 1) * (nely + 1) * (nelz + 1) + 1).reshape(nelx + 1, nelz + 1, nely + 1)
    cvec = 3 * nodeNrs[:nelx, :nelz, :nely].reshape(-1) + 1  # element order j fastest, then k, then i
    s1, s2, s3 = 3 * (nely + 1) * (nelz + 1), 3 * (nely + 1) * (nelz + 2), 3 * (nely + 1)
    off = np.array([0, 1, 2, s1, s1 + 1, s1 + 2, s1 - 3, s1 - 2, s1 - 1, -3, -2, -1, s3, s3 + 1, s3 + 2,
                    s2, s2 + 1, s2 + 2, s2 - 3, s2 - 2, s2 - 1, s3 - 3, s3 - 2, s3 - 1])
    assert np.array_equal(oracle.edofmat_matlab(nelx, nely, nelz), cvec[:, None] + off[None, :])


@pytest.mark.parametrize("preset", [0, 1])
def test_mgcg_direct_solve(oracle, preset):
    nelx, nely, nelz, nl = 8, 4, 4, 3
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=7, kind="uniform")
    K = dense_K(Ke, CX, nelx, nely, nelz)
    free = np.setdiff1d(np.arange(F.size), (fixed - 1).astype(int))
    Ud = np.zeros(F.size)
    Ud[free] = np.linalg.solve(K[np.ix_(free, free)], F[free])
    U, relres, nit, rc = oracle.mgcg(Ke, F, CX, fixed, nelx, nely, nelz, nl, 1e-12, 300, preset=preset,
                                     flip_y=False)
    assert rc == 0 and relres < 1e-12
    assert np.abs(U - Ud).max() < 1e-10 * np.abs(Ud).max()


def test_imfilter_symmetric(oracle):
    nelx, nely, nelz = 6, 5, 4
    x = np.random.default_rng(8).uniform(0, 1, nelx * nely * nelz)
    for rmin in (np.sqrt(3), 2 * np.sqrt(3), 1.2):
        w = int(np.ceil(rmin)) - 1
        d = np.arange(-w, w + 1)
        DX, DZ, DY = np.meshgrid(d, d, d, indexing="ij")
        h = np.maximum(0, rmin - np.sqrt(DX ** 2 + DY ** 2 + DZ ** 2))
        arr = x.reshape(nelx, nelz, nely)  # [i][k][j]
        ref = ndi.correlate(arr, h, mode="reflect").ravel()
        assert np.abs(oracle.imfilter(x, nelx, nely, nelz, rmin) - ref).max() < 1e-13
    ones = oracle.imfilter(np.ones(x.size), nelx, nely, nelz, np.sqrt(3))
    # Hs is constant: centre sqrt3, 6 faces sqrt3-1, 12 edges sqrt3-sqrt2, corners 0
    r3 = np.sqrt(3)
    assert np.allclose(ones, ones[0]) and np.isclose(ones[0], r3 + 6 * (r3 - 1) + 12 * (r3 - np.sqrt(2)))


def test_eta_newton_and_oc(oracle):
    rng = np.random.default_rng(9)
    xp = rng.uniform(0, 1, 5000)
    eta, it = oracle.eta_newton(xp, 0.4, 2.0, 0.3)
    b = 2.0
    prj = (np.tanh(b * eta) + np.tanh(b * (xp - eta))) / (np.tanh(b * eta) + np.tanh(b * (1 - eta)))
    assert abs(prj.mean() - 0.3) <= 1e-6
    x = 0.2 + rng.uniform(-0.1, 0.1, 5000)
    dc = -rng.uniform(0.1, 2, 5000) ** 4
    dv = np.full(5000, 1 / 5000 / 0.2)
    ocp = x * np.sqrt(-dc / dv)

    def vol(lam):
        return np.clip(np.clip(ocp / lam, x - 0.2, x + 0.2), 0, 1).mean()

    lo, hi = 1e-6, 1e6  # brute-force multiplier on a wide bracket
    for _ in range(200):
        mid = np.sqrt(lo * hi)
        lo, hi = (mid, hi) if vol(mid) > 0.2 else (lo, mid)
    assert lo < ocp.mean() / 0.2 * (1 - 1e-2)  # the reference bracket [0, mean(ocP)/volfrac] holds it
    xn, lmid, it = oracle.oc(x, dc, dv, 0.2, 0.2)
    assert abs(lmid - lo) < 2e-4 * lo
    assert abs(xn.mean() - 0.2) < 1e-3
    assert (xn <= np.minimum(x + 0.2, 1) + 1e-15).all() and (xn >= np.maximum(x - 0.2, 0) - 1e-15).all()
    ocp = x * np.sqrt(-dc / dv)
    assert np.allclose(xn, np.clip(np.clip(ocp / lmid, x - 0.2, x + 0.2), 0, 1))


def test_sensitivity_frame_invariance(oracle):
    nelx, nely, nelz = 3, 2, 2
    Ke = oracle.kem3()
    U = np.random.default_rng(10).uniform(-1, 1, 3 * 4 * 3 * 3)
    cmat = oracle.edofmat_matlab(nelx, nely, nelz)
    ce = oracle.sens_matlab(Ke, U, cmat)
    Uc = U.copy()
    Uc[1::3] *= -1
    en = oracle.elem_nodes_c(nelx, nely, nelz)
    for e in range(cmat.shape[0]):
        ue = np.concatenate([Uc[3 * n:3 * n + 3] for n in en[e]])
        assert abs(ue @ Ke @ ue - ce[e]) < 1e-12 * abs(ce[e])


def test_learn_oracle_gradient_finite_differences():
    """The torch restatement of the network: autograd against central
    differences of <gy, y> on a tiny grid."""
    import learn_oracle
    rng = np.random.default_rng(5)
    p = rng.normal(size=learn_oracle.nparams()) * 0.05
    x = rng.normal(size=(4, 5, 3, 3))
    gy = rng.normal(size=(3, 5, 3, 3))
    y, g = learn_oracle.run(p, x, gy)
    assert y.shape == (3, 5, 3, 3)
    eps = 1e-6
    for idx in rng.choice(learn_oracle.nparams(), 12, replace=False):
        pp = p.copy()
        pp[idx] += eps
        y1, _ = learn_oracle.run(pp, x)
        pp[idx] -= 2 * eps
        y2, _ = learn_oracle.run(pp, x)
        fd = ((y1 - y2) * gy).sum() / (2 * eps)
        assert abs(fd - g[idx]) <= 1e-6 * max(1.0, abs(g[idx]))
