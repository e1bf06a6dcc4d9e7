"""Parity of the HIP path (libgputo.so through the C ABI) with the CPU oracle.

Tolerances: index maps are bit-exact; FP64 operators agree to a few ulps
times the operand magnitudes (rtol 1e-13 of max|y|); quantities that pass
through the coarsest explicit inverse agree to the conditioning of the
penalised coarsest matrix (rtol 1e-9); converged solves agree to the CG
tolerance.
"""
import numpy as np
import pytest

from conftest import cantilever_case, dense_K

pytestmark = pytest.mark.gpu

MESHES = [(4, 2, 2, 2), (8, 4, 4, 3), (16, 8, 8, 3), (24, 12, 12, 3)]


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def solvers(gputo):
    made = {}

    def get(nelx, nely, nelz, nl, preset=0):
        key = (nelx, nely, nelz, nl, preset)
        if key not in made:
            made[key] = gputo.Solver(nelx, nely, nelz, nl, preset=preset)
        return made[key]

    yield get
    for s in made.values():
        s.close()


def general_ke(Ke, seed=3):
    """A symmetric Ke0 without the unit-H8 entry classes (exercises the
    level-0 kernel that reads all 576 entries)."""
    E = np.random.default_rng(seed).uniform(-1, 1, (24, 24)) * 1e-3
    return Ke + (E + E.T)


@pytest.mark.parametrize("mesh", MESHES)
@pytest.mark.parametrize("kind", ["perturbed", "binary", "general_ke"])
def test_matvec_defect_diag(solvers, oracle, mesh, kind):
    nelx, nely, nelz, nl = mesh
    print("case", mesh, kind, flush=True)
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=1, kind="perturbed" if kind == "general_ke" else kind)
    if kind == "general_ke":
        Ke = general_ke(Ke)
    s = solvers(nelx, nely, nelz, nl)
    u = np.random.default_rng(2).uniform(-1, 1, s.ndof)
    y = s.matvec(Ke, CX, u, fixed)
    yo = oracle.matvec0(Ke, CX, u, fixed, nelx, nely, nelz)
    assert rel(y, yo) < 1e-13
    assert np.array_equal(y, s.matvec(Ke, CX, u, fixed))  # deterministic
    r = s.defect(Ke, CX, F, u, fixed)
    ro = oracle.defect0(Ke, F, u, CX, fixed, nelx, nely, nelz)
    assert rel(r, ro) < 1e-13
    ad = s.diag(Ke, CX)
    assert rel(ad, oracle.prepad0(Ke, CX, nelx, nely, nelz)) < 1e-14


def test_class_and_general_ke_kernels_agree(solvers):
    """The class-valued level-0 kernel and the general one compute the same
    product for the unit-H8 Ke0 (the general one is forced by a Ke0 whose one
    entry differs in the last bit, which the class check rejects)."""
    nelx, nely, nelz, nl = 16, 8, 8, 3
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=9, kind="perturbed")
    K2 = Ke.copy()
    K2[0, 0] = np.nextafter(K2[0, 0], 1.0)
    s = solvers(nelx, nely, nelz, nl)
    u = np.random.default_rng(10).uniform(-1, 1, s.ndof)
    y1 = s.matvec(Ke, CX, u, fixed)
    y2 = s.matvec(K2, CX, u, fixed)
    assert rel(y1, y2) < 1e-13


def test_matvec_dense_bruteforce(solvers):
    nelx, nely, nelz = 4, 2, 2
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=3)
    s = solvers(nelx, nely, nelz, 2)
    u = np.random.default_rng(4).uniform(-1, 1, s.ndof)
    K = dense_K(Ke, CX, nelx, nely, nelz)
    yd = K @ u
    yd[(fixed - 1).astype(int)] = 0.0
    assert rel(s.matvec(Ke, CX, u, fixed), yd) < 1e-13


@pytest.mark.parametrize("preset", [0, 1])
@pytest.mark.parametrize("mesh", [(8, 4, 4, 3), (16, 8, 8, 3), (16, 8, 8, 4), (32, 16, 16, 5)])
def test_hierarchy_levels(solvers, oracle, mesh, preset):
    nelx, nely, nelz, nl = mesh
    print("case", mesh, preset, flush=True)
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=5)
    s = solvers(nelx, nely, nelz, nl, preset)
    s.setup(Ke, CX, fixed)
    for l in range(nl):
        A, ad, fd = s.level_matrices(l)
        Ao, ado, fdo = oracle.level_data(Ke, CX, fixed, nelx, nely, nelz, nl, l, preset=preset)
        assert np.array_equal(fd, fdo), "fixed list of level %d" % l  # bit-exact, same order
        assert np.array_equal(s.fixed_list(l), fdo)
        if l > 0:
            assert rel(A, Ao) < 1e-12, "element matrices of level %d" % l
        assert rel(ad, ado) < 1e-12


def test_hierarchy_levels_table_galerkin(solvers, oracle):
    """Levels with at least 8192 elements are built from the per-fine-element
    tables (level 2, 32x16x16, here)."""
    nelx, nely, nelz, nl = 128, 64, 64, 4
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=7, kind="perturbed")
    s = solvers(nelx, nely, nelz, nl, 0)
    s.setup(Ke, CX, fixed)
    for l in (2,):
        A, ad, fd = s.level_matrices(l)
        Ao, ado, fdo = oracle.level_data(Ke, CX, fixed, nelx, nely, nelz, nl, l)
        assert rel(A, Ao) < 1e-12, "element matrices of level %d" % l
        assert rel(ad, ado) < 1e-12


def test_galerkin_levels_follow_ke(solvers, oracle):
    """With a Galerkin level 1 the coarse operators are rebuilt when Ke
    changes between calls on one handle (and stay right when it does not)."""
    nelx, nely, nelz, nl = 16, 8, 8, 3
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=8)
    s = solvers(nelx, nely, nelz, nl, 1)
    for K in (Ke, 2.0 * Ke, 2.0 * Ke, Ke):
        s.setup(K, CX, fixed)
        for l in (1, 2):
            A, ad, fd = s.level_matrices(l)
            Ao, ado, fdo = oracle.level_data(K, CX, fixed, nelx, nely, nelz, nl, l, preset=1)
            assert rel(A, Ao) < 1e-12


def test_edofmat_bit_exact(solvers, oracle):
    for nelx, nely, nelz, nl in MESHES[:3]:
        s = solvers(nelx, nely, nelz, nl)
        assert np.array_equal(s.edofmat(), oracle.edofmat_matlab(nelx, nely, nelz))


def test_transfers(solvers, oracle):
    nelx, nely, nelz, nl = 16, 8, 8, 3
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=6)
    s = solvers(nelx, nely, nelz, nl)
    s.setup(Ke, CX, fixed)
    rng = np.random.default_rng(7)
    for l in range(nl - 1):
        fx, fy, fz = nelx >> l, nely >> l, nelz >> l
        cx, cy, cz = fx // 2, fy // 2, fz // 2
        rf = rng.uniform(-1, 1, 3 * (fx + 1) * (fy + 1) * (fz + 1))
        fdc = s.fixed_list(l + 1)
        fc = s.restrict(l, rf)
        assert rel(fc, oracle.restrict(rf, fdc, cx, cy, cz)) < 1e-14
        uc = rng.uniform(-1, 1, 3 * (cx + 1) * (cy + 1) * (cz + 1))
        uf = rng.uniform(-1, 1, rf.size)
        assert rel(s.interp_add(l, uc, uf), oracle.interp_add(uc, uf, fx, fy, fz)) < 1e-14


@pytest.mark.parametrize("mesh", [(16, 8, 8, 3), (48, 24, 24, 3), (32, 16, 16, 2)])
def test_coarse_solve(solvers, oracle, mesh):
    """The coarsest solve through the explicit inverse of the factor against
    numpy's dense solve: a 5x3x3-node grid (5 tiles), the 13x7x7-node grid of
    the 192x96x96 hierarchy (1911 DOFs, band 173, 60 tiles: one-workgroup
    factorisation, six halving levels of the inverse) and a 17x9x9-node grid
    (band 275: tile-parallel factorisation)."""
    nelx, nely, nelz, nl = mesh
    print("case", mesh, flush=True)
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=8)
    s = solvers(nelx, nely, nelz, nl)
    s.setup(Ke, CX, fixed)
    A, _, fd = s.level_matrices(nl - 1)
    cx, cy, cz = nelx >> (nl - 1), nely >> (nl - 1), nelz >> (nl - 1)
    AQ = oracle.assemble(A, cx, cy, cz)
    b = np.random.default_rng(9).uniform(-1, 1, AQ.shape[0])
    b[(fd - 1).astype(int)] = 0.0
    x = s.coarse_solve(b)
    xr = np.linalg.solve(AQ, b)
    xr[(fd - 1).astype(int)] = 0.0
    assert rel(x, xr) < 1e-9


@pytest.mark.parametrize("preset", [0, 1, 2])
@pytest.mark.parametrize("mesh", [(16, 8, 8, 3), (32, 16, 16, 4), (32, 16, 16, 5), (48, 24, 24, 4)])
def test_vcycle(solvers, oracle, preset, mesh):
    """One V-cycle of the FP64 hierarchy against the oracle's, including the
    node-stencil levels 2..nl-2 (nl >= 4) and the mgcg9 engine (nswp = 2)."""
    nelx, nely, nelz, nl = mesh
    print("case", mesh, preset, flush=True)
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=10)
    s = solvers(nelx, nely, nelz, nl, preset)
    r = np.random.default_rng(11).uniform(-1, 1, s.ndof)
    r[(fixed - 1).astype(int)] = 0.0
    z = s.vcycle(Ke, CX, fixed, r)
    oracle.lib().or_set_nswp(2)
    zo = oracle.vcycle_once(Ke, CX, fixed, nelx, nely, nelz, nl, r, preset=preset)
    assert rel(z, zo) < 1e-9


@pytest.mark.parametrize("mesh", [(8, 4, 4, 3), (16, 8, 8, 3), (24, 12, 12, 3), (32, 16, 16, 4)])
@pytest.mark.parametrize("kind", ["perturbed", "smooth"])
def test_mgcg1_vs_oracle(solvers, oracle, mesh, kind):
    nelx, nely, nelz, nl = mesh
    print("case", mesh, kind, flush=True)
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=12, kind=kind)
    s = solvers(nelx, nely, nelz, nl)
    U, relres, nit, rc = s.solve(Ke, F, CX, fixed, 1e-10, 500)
    assert rc == 0
    Uo, relres_o, nit_o, rco = oracle.mgcg(Ke, F, CX, fixed, nelx, nely, nelz, nl, 1e-10, 500)
    assert rco == 0
    assert relres < 1e-10
    assert abs(nit - nit_o) <= 1
    assert rel(U, Uo) < 1e-8
    C, Co = F @ U, F @ Uo
    assert abs(C - Co) <= 1e-9 * abs(Co)
    U2, relres2, nit2, _ = s.solve(Ke, F, CX, fixed, 1e-10, 500)
    assert np.array_equal(U, U2) and nit2 == nit and relres2 == relres  # bit-identical repeat


def test_mgcg1_stagnation_matches_oracle(solvers, oracle):
    """Random stiff islands in near-void (E ratio 1e-9): the reference
    algorithm stagnates (its damped Jacobi factor 0.48 exceeds 2/lambda_max
    of D^-1 K on such fields); the CG trajectories still agree after maxit
    iterations."""
    nelx, nely, nelz, nl = 8, 4, 4, 3
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=12, kind="binary")
    s = solvers(nelx, nely, nelz, nl)
    U, relres, nit, rc = s.solve(Ke, F, CX, fixed, 1e-10, 300)
    Uo, relres_o, nit_o, rco = oracle.mgcg(Ke, F, CX, fixed, nelx, nely, nelz, nl, 1e-10, 300)
    assert rc == 0 and rco == 0 and nit == nit_o == 300
    assert abs(relres - relres_o) <= 1e-6 * relres_o


def test_mgcg1_dense_direct(solvers):
    nelx, nely, nelz, nl = 8, 4, 4, 3
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=13)
    s = solvers(nelx, nely, nelz, nl)
    U, relres, nit, rc = s.solve(Ke, F, CX, fixed, 1e-12, 500)
    assert rc == 0
    K = dense_K(Ke, CX, nelx, nely, nelz)
    free = np.setdiff1d(np.arange(s.ndof), (fixed - 1).astype(int))
    Ud = np.zeros(s.ndof)
    Ud[free] = np.linalg.solve(K[np.ix_(free, free)], F[free])
    Ud[1::3] *= -1  # MATLAB frame
    assert rel(U, Ud) < 1e-9
    assert abs(F @ U - F @ Ud) <= 1e-10 * abs(F @ Ud)


def test_wide_band_coarsest(gputo):
    """Coarsest factorisations on both sides of the single-workgroup band
    limit (256 rows below the diagonal tile): a 2-level hierarchy at
    32x16x16 factors a 17x9x9-node coarsest grid (band 275 rows, tile-parallel
    kernels), a 4-level one a 5x3x3-node grid (band 62 rows, one workgroup).
    Both solves reach the same displacements."""
    nelx, nely, nelz = 32, 16, 16
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=18)
    out = {}
    for nl in (2, 4):
        print("case", nl, flush=True)
        s = gputo.Solver(nelx, nely, nelz, nl)
        U, relres, nit, rc = s.solve(Ke, F, CX, fixed, 1e-11, 300)
        assert rc == 0 and relres < 1e-11
        out[nl] = (U, nit)
        s.close()
    assert out[2][1] <= out[4][1]
    assert rel(out[2][0], out[4][0]) < 1e-8


def test_mgcg1_matlab_interface(gputo, oracle):
    """The reference call forms: CUDA 12/13 args and omp 13 args with WRK."""
    nelx, nely, nelz, nl = 16, 8, 8, 3
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=14)
    ndof = F.size
    U = np.zeros(ndof)
    cgres, cgiters = gputo.mgcg1(Ke, F, U, CX, fixed, len(fixed), nelx, nely, nelz, nl, 1e-9, 200)
    assert cgiters >= 0 and cgres < 1e-9
    U13 = np.zeros(ndof)
    r13 = gputo.mgcg1(Ke, F, U13, CX, fixed, len(fixed), nelx, nely, nelz, nl, 1e-9, 200, 0)
    assert np.array_equal(U, U13) and r13 == (cgres, cgiters)
    Uw = np.zeros(ndof)
    WRK = np.zeros(100 * ndof)
    rw = gputo.mgcg1(Ke, F, Uw, CX, fixed, WRK, len(fixed), nelx, nely, nelz, nl, 1e-9, 200)
    assert np.array_equal(U, Uw) and rw == (cgres, cgiters)
    # too small a workspace: nit = -1 and U untouched (mgcg1.c:89-97)
    Us = np.zeros(ndof)
    assert gputo.mgcg1(Ke, F, Us, CX, fixed, np.zeros(10), len(fixed), nelx, nely, nelz, nl, 1e-9, 200) == (0.0, -1.0)
    Uo, _, _, _ = oracle.mgcg(Ke, F, CX, fixed, nelx, nely, nelz, nl, 1e-9, 200)
    assert rel(U, Uo) < 1e-7


def test_error_codes(gputo):
    nelx, nely, nelz, nl = 8, 4, 4, 3
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=15)
    s = gputo.Solver(nelx, nely, nelz, nl)
    bad = CX.copy()
    bad[3] = np.nan
    _, _, nit, rc = s.solve(Ke, F, bad, fixed, 1e-9, 50)
    assert rc == -3 and nit == -3
    # negative moduli: the coarsest matrix is not positive definite
    _, _, nit, rc = s.solve(Ke, F, -CX, fixed, 1e-9, 50)
    assert rc == -2 and nit == -2
    # zero load: U = 0 exactly
    U, relres, nit, rc = s.solve(Ke, np.zeros_like(F), CX, fixed, 1e-9, 50)
    assert rc == 0 and nit == 0 and not U.any()
    with pytest.raises(gputo.GputoError):
        gputo.Solver(7, 4, 4, 3)  # 7 not divisible by 4
    s.close()


def test_filter_and_sens(solvers, oracle):
    nelx, nely, nelz, nl = 16, 8, 8, 3
    s = solvers(nelx, nely, nelz, nl)
    x = np.random.default_rng(16).uniform(0, 1, s.nel)
    for rmin in (np.sqrt(3), 2 * np.sqrt(3), 1.5):
        assert rel(s.filter(x, rmin), oracle.imfilter(x, nelx, nely, nelz, rmin)) < 1e-14
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=17)
    U, _, _, rc = s.solve(Ke, F, CX, fixed, 1e-10, 300)
    assert rc == 0
    ce_m = oracle.sens_matlab(Ke, U, oracle.edofmat_matlab(nelx, nely, nelz))
    Uc = U.copy()
    Uc[1::3] *= -1  # back to the C frame
    assert rel(s.sens(Ke, Uc), ce_m) < 1e-12
