"""Level operators of the hierarchy through gputo_level_apply: the
matrix-free level-1 operator (Gauss-point form of the rediscretised level-1
element matrices, coarseAe2c, Cantilever/TOP3D125MGCGCUDA/mgcg1.cu:2042-2070)
against the oracle's stored level-1 matrices applied element by element
(multAV, Cantilever/omp/mgcg1.c:667-715), plus size-independent properties:
zero action on rigid translations and agreement with the stored GPU kernel.

Tolerances (relative to max|y|): FP64 1e-13 (summation order only), FP32
2e-6 (FP32 rounding of u, CX and the Gauss-point arithmetic).
"""
import numpy as np
import pytest

from conftest import cantilever_case

pytestmark = pytest.mark.gpu

OMEGA_REF_C = 0.48  # damped-Jacobi weight of the reference smoother (Jacobc, mgcg1.cu:1206)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


MESHES = [(8, 4, 4, 3), (16, 8, 8, 3), (20, 12, 8, 3), (32, 16, 16, 4), (56, 28, 28, 3)]


@pytest.mark.parametrize("mesh", MESHES)
@pytest.mark.parametrize("fp32", [False, True])
def test_level1_matrix_free_vs_oracle(gputo, oracle, mesh, fp32):
    nelx, nely, nelz, nl = mesh
    print("case", mesh, fp32, flush=True)
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=31, kind="smooth")
    s = gputo.Solver(nelx, nely, nelz, nl, precision=1 if fp32 else 0)
    s.setup(Ke, CX, fixed)
    A1, ad1, fd1 = oracle.level_data(Ke, CX, fixed, nelx, nely, nelz, nl, 1)
    nx, ny, nz = nelx // 2, nely // 2, nelz // 2
    rng = np.random.default_rng(32)
    u = rng.uniform(-1, 1, ad1.size)
    f = rng.uniform(-1, 1, ad1.size)
    tol = 2e-6 if fp32 else 1e-13
    y_or = oracle.matvec_el(A1, u, fd1, nx, ny, nz)
    y = s.level_apply(1, 2, 0, u, fp32=fp32)
    assert rel(y, y_or) < tol
    assert np.array_equal(y, s.level_apply(1, 2, 0, u, fp32=fp32))  # deterministic
    # defect f - A u, zero at fixed DOFs
    r_or = f - y_or
    r_or[fd1.astype(np.int64) - 1] = 0.0
    r = s.level_apply(1, 2, 1, u, f, fp32=fp32)
    assert rel(r, r_or) < tol
    # one damped-Jacobi sweep (1-w) u + w (r + AD u)/AD
    j_or = (1 - OMEGA_REF_C) * u + OMEGA_REF_C * (r_or + ad1 * u) / ad1
    j = s.level_apply(1, 2, 2, u, f, fp32=fp32)
    assert rel(j, j_or) < tol
    # the stored-matrix kernel computes the same operator
    assert rel(s.level_apply(1, 1, 0, u, fp32=fp32), y_or) < tol
    s.close()


@pytest.mark.parametrize("mesh", [(16, 8, 8, 3), (32, 16, 16, 4), (48, 24, 24, 4)])
def test_fp32_level_operators_match_fp64(gputo, mesh):
    """Every level's FP32 operator (matvec, defect, Jacobi sweep) against the
    FP64 operator of the same hierarchy: 2e-6 of max|y|."""
    nelx, nely, nelz, nl = mesh
    print("case", mesh, flush=True)
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=37, kind="perturbed")
    s = gputo.Solver(nelx, nely, nelz, nl, precision=1, l0_kernel=1)
    s.setup(Ke, CX, fixed)
    rng = np.random.default_rng(38)
    for l in range(nl - 1):
        n = 3 * ((nelx >> l) + 1) * ((nely >> l) + 1) * ((nelz >> l) + 1)
        u, f = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
        for op in ((1, 2) if l == 1 else (1,)):
            for mode in (0, 1, 2):
                y64 = s.level_apply(l, op, mode, u, f)
                y32 = s.level_apply(l, op, mode, u, f, fp32=True)
                assert rel(y32, y64) < 2e-6, (l, op, mode, rel(y32, y64))
    s.close()


def test_level1_matrix_free_rigid_translation(gputo):
    nelx, nely, nelz, nl = 24, 12, 12, 3
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=33, kind="smooth")
    s = gputo.Solver(nelx, nely, nelz, nl)
    s.setup(Ke, CX, fixed)  # rows of A sum to zero per direction; fixed rows are zeroed
    n1 = 3 * (nelx // 2 + 1) * (nely // 2 + 1) * (nelz // 2 + 1)
    scale = np.abs(s.level_apply(1, 2, 0, np.random.default_rng(34).uniform(-1, 1, n1))).max()
    for d in range(3):
        t = np.zeros(n1)
        t[d::3] = 1.0
        assert np.abs(s.level_apply(1, 2, 0, t)).max() < 1e-13 * scale
    s.close()


def test_level1_kernel_choice_same_solve(gputo):
    """The V-cycle with the matrix-free level 1 and with the stored level-1
    matrices give the same PCG solve (iteration counts within one, U to 1e-9)."""
    nelx, nely, nelz, nl = 32, 16, 16, 4
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=35, kind="smooth")
    out = []
    for l1 in (1, 2):
        s = gputo.Solver(nelx, nely, nelz, nl, l1_kernel=l1)
        U, r, n, rc = s.solve(Ke, F, CX, fixed, 1e-10, 200)
        assert rc == 0
        out.append((U, n))
        s.close()
    assert abs(out[0][1] - out[1][1]) <= 1
    assert rel(out[1][0], out[0][0]) < 1e-9


def test_level_apply_rejects_bad_arguments(gputo):
    nelx, nely, nelz, nl = 8, 4, 4, 3
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=36, kind="smooth")
    s = gputo.Solver(nelx, nely, nelz, nl, preset=gputo.PRESET_MATLAB_CFG0)
    s.setup(Ke, CX, fixed)
    n1 = 3 * 5 * 3 * 3
    with pytest.raises(gputo.GputoError):
        s.level_apply(1, 2, 0, np.zeros(n1))  # Galerkin level 1 has no matrix-free form
    with pytest.raises(gputo.GputoError):
        s.level_apply(2, 1, 0, np.zeros(3 * 3 * 2 * 2))  # the coarsest level is solved directly
    s.close()


PRERELAX_REF_C = 0.7  # preRelc, mgcg1.cu:1190-1197


def level_reference(oracle, A, ad, fd, nx, ny, nz, u, f, mode, omega=OMEGA_REF_C, prerelax=PRERELAX_REF_C):
    """The oracle's product for gputo_level_apply's mode (A None: level 0)."""
    if mode in (3, 4):
        u = prerelax * f / ad
    y = oracle.matvec_el(A, u, fd, nx, ny, nz) if A is not None else None
    return y, u


@pytest.mark.parametrize("mesh", [(32, 16, 16, 4), (32, 16, 16, 5), (48, 32, 32, 5)])
@pytest.mark.parametrize("fp32", [False, True])
def test_stencil_levels_vs_oracle(gputo, oracle, mesh, fp32):
    """The node-stencil operator that every V-cycle runs on levels 2..nl-2
    (k_stencil, assembled by k_stencil_build from the Galerkin element
    matrices) against the oracle's element matrices of the same level
    applied element by element (defectc, mgcg1.cu:1266-1324), in all five
    modes.  Tolerances as for level 1: 1e-13 FP64, 2e-6 FP32 (rel. to max|y|)."""
    nelx, nely, nelz, nl = mesh
    print("case", mesh, fp32, flush=True)
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=39, kind="smooth")
    s = gputo.Solver(nelx, nely, nelz, nl, precision=1 if fp32 else 0)
    s.setup(Ke, CX, fixed)
    rng = np.random.default_rng(40)
    tol = 2e-6 if fp32 else 1e-13
    for l in range(2, nl - 1):
        A, ad, fd = oracle.level_data(Ke, CX, fixed, nelx, nely, nelz, nl, l)
        nx, ny, nz = nelx >> l, nely >> l, nelz >> l
        u, f = rng.uniform(-1, 1, ad.size), rng.uniform(-1, 1, ad.size)
        fx = fd.astype(np.int64) - 1
        for mode in range(5):
            uu = PRERELAX_REF_C * f / ad if mode in (3, 4) else u
            y_or = oracle.matvec_el(A, uu, fd, nx, ny, nz)
            if mode in (1, 4):
                ref = f - y_or
                ref[fx] = 0.0
            elif mode in (2, 3):
                r = f - y_or
                r[fx] = 0.0
                ref = (1 - OMEGA_REF_C) * uu + OMEGA_REF_C * (r + ad * uu) / ad
            else:
                ref = y_or
            y = s.level_apply(l, 1, mode, u, f, fp32=fp32)
            assert rel(y, ref) < tol, (l, mode, rel(y, ref))
            if mode == 0:
                assert np.array_equal(y, s.level_apply(l, 1, mode, u, f, fp32=fp32))  # deterministic
    s.close()


@pytest.mark.parametrize("kernel,tol", [(1, 2e-6), (2, 2e-6), (3, 1e-4), (4, 2e-6)])
@pytest.mark.parametrize("mesh", [(16, 8, 8, 3), (40, 20, 12, 3)])
def test_level0_fp32_kernel_modes_vs_oracle(gputo, oracle, kernel, tol, mesh):
    """Every mode of the FP32 level-0 kernels (node-gather and the MFMA
    flavours) against the FP64 oracle: matvec, defect, damped-Jacobi sweep,
    sweep from the prerelaxed iterate, defect of the prerelaxed iterate
    (multAV0c / defect0c / preRelc / Jacobc, mgcg1.cu:941-997,1190-1264).
    Tolerances relative to max|y|: FP32 rounding (2e-6); the hi+lo split-BF16
    flavour drops lo*lo (1e-4)."""
    nelx, nely, nelz, nl = mesh
    print("case", mesh, kernel, flush=True)
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=41, kind="smooth")
    s = gputo.Solver(nelx, nely, nelz, nl, precision=2, l0_kernel=kernel)
    s.setup(Ke, CX, fixed)
    rng = np.random.default_rng(42)
    ad = oracle.prepad0(Ke, CX, nelx, nely, nelz)
    u, f = rng.uniform(-1, 1, ad.size), rng.uniform(-1, 1, ad.size)
    fx = (fixed - 1).astype(np.int64)
    for mode in range(5):
        uu = PRERELAX_REF_C * f / ad if mode in (3, 4) else u
        Ku = oracle.matvec0(Ke, CX, uu, np.zeros(0), nelx, nely, nelz)
        if mode in (1, 4):
            ref = f - Ku
            ref[fx] = 0.0
        elif mode in (2, 3):
            r = f - Ku
            r[fx] = 0.0
            ref = (1 - OMEGA_REF_C) * uu + OMEGA_REF_C * (r + ad * uu) / ad
        else:
            ref = Ku
            ref[fx] = 0.0
        y = s.level_apply(0, 1, mode, u, f, fp32=True)
        assert rel(y, ref) < tol, (mode, rel(y, ref))
        assert np.array_equal(y, s.level_apply(0, 1, mode, u, f, fp32=True))  # deterministic
    s.close()
