"""Mixed-precision path: FP32 / split-BF16 MFMA level-0 operators and the
mixed PCG (FP32 operators, FP64 iterate, FP64 residual replacement), checked
against the FP64 oracle.

Tolerances: the FP32 node-gather and FP32 MFMA products agree with the FP64
oracle to 1e-6 of max|y| (FP32 rounding of Ke, E, u and the accumulations);
the three-way split-BF16 product (terms below 2^-24 dropped) to 1e-6; the
two-way split (hi*hi + hi*lo + lo*hi, lo*lo dropped) to 1e-4.
The mixed solve stops on an FP64 true residual below the tolerance, so its
displacements and compliance agree with the plain FP64 solve to the
tolerance, in about the FP64 iteration count.
"""
import numpy as np
import pytest

from conftest import cantilever_case

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_mfma_lane_layout(gputo):
    import ctypes
    err = ctypes.c_double(-1.0)
    assert gputo.load().gputo_selftest_mfma(ctypes.byref(err)) == 0
    assert err.value == 0.0


@pytest.mark.parametrize("mesh", [(8, 4, 4, 3), (16, 8, 8, 3), (24, 12, 12, 3), (48, 24, 24, 4)])
@pytest.mark.parametrize("kernel,tol", [(1, 1e-6), (2, 1e-6), (3, 1e-4), (4, 1e-6), (5, 3e-2)])
def test_reduced_precision_operator(gputo, oracle, mesh, kernel, tol):
    nelx, nely, nelz, nl = mesh
    print("case", mesh, kernel, flush=True)
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=21, kind="smooth")
    s = gputo.Solver(nelx, nely, nelz, nl)
    u = np.random.default_rng(22).uniform(-1, 1, s.ndof)
    y = s.matvec_mixed(kernel, Ke, CX, u, fixed)
    yo = oracle.matvec0(Ke, CX, u, fixed, nelx, nely, nelz)
    assert rel(y, yo) < tol
    assert np.array_equal(y, s.matvec_mixed(kernel, Ke, CX, u, fixed))  # deterministic
    s.close()


@pytest.mark.parametrize("prec,kernel", [(1, 1), (1, 2), (2, 3), (2, 4), (2, 5)])
@pytest.mark.parametrize("mesh", [(16, 8, 8, 3), (48, 24, 24, 4)])
def test_mixed_solve_matches_f64(gputo, oracle, prec, kernel, mesh):
    nelx, nely, nelz, nl = mesh
    print("case", mesh, prec, kernel, flush=True)
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=23, kind="perturbed")
    s64 = gputo.Solver(nelx, nely, nelz, nl)
    U64, r64, n64, rc = s64.solve(Ke, F, CX, fixed, 1e-10, 300)
    assert rc == 0
    sm = gputo.Solver(nelx, nely, nelz, nl, precision=prec, l0_kernel=kernel)
    Um, rm, nm, rc = sm.solve(Ke, F, CX, fixed, 1e-10, 300)
    assert rc == 0 and rm < 1e-10
    # FP32 level-0 operators in the preconditioner keep the FP64 iteration
    # count; the BF16 ones perturb the preconditioner (hi+lo: relative error
    # ~1e-5, single product: ~1e-2) and cost iterations, never accuracy
    print("iterations", n64, nm, flush=True)
    assert nm <= n64 + {3: 40, 5: 2 * n64}.get(kernel, 2), (nm, n64)
    # the outer residual is the true FP64 residual: check it with the oracle
    Uc = Um.copy()
    Uc[1::3] *= -1
    r = oracle.defect0(Ke, F, Uc, CX, fixed, nelx, nely, nelz)
    assert np.linalg.norm(r) / np.linalg.norm(F) < 1e-10
    assert rel(Um, U64) < 1e-7
    C64, Cm = F @ U64, F @ Um
    assert abs(Cm - C64) <= 1e-9 * abs(C64)
    s64.close()
    sm.close()


@pytest.mark.parametrize("kernel", [3, 1])
@pytest.mark.parametrize("sweeps", [(1, 3), (2, 2), (1, 1), (1, 1, 6, 6), (1, 1, 2, 2), (1, 3, 2, 2, 8, 8),
                                    (1, 3, -1, 1, 1, 1, 8, 8)])
def test_level0_sweep_schedule(gputo, oracle, sweeps, kernel):
    """Other sweep counts on levels 0 and 1 (and, with four entries, on the
    levels from 2; with six, 2 and from 3 on; with eight, level 0, level 1,
    level 2 and from 3 on, -1 standing for the prerelaxed iterate alone
    before the coarse correction) of the FP32 V-cycle change the
    preconditioner, not the solution: the FP64 true residual meets the
    tolerance and the compliance matches the FP64 solve.  The six- and
    eight-entry cases have 5 levels, so level 3 is smoothed."""
    nelx, nely, nelz, nl = (64, 32, 32, 5) if len(sweeps) > 4 else (48, 24, 24, 4)
    print("case", sweeps, kernel, flush=True)
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=23, kind="perturbed")
    s64 = gputo.Solver(nelx, nely, nelz, nl)
    U64, r64, n64, rc = s64.solve(Ke, F, CX, fixed, 1e-10, 300)
    assert rc == 0
    if len(sweeps) == 8:
        s0, s1, below, below3 = sweeps[0:2], sweeps[2:4], sweeps[4:6], sweeps[6:8]
    else:
        s0 = s1 = sweeps[0:2]
        below = sweeps[2:4] if len(sweeps) > 2 else (0, 0)
        below3 = sweeps[4:6] if len(sweeps) > 4 else (0, 0)
    sm = gputo.Solver(nelx, nely, nelz, nl, precision=2, l0_kernel=kernel, mixed_npre0=s0[0],
                      mixed_npost0=s0[1], mixed_npre1=s1[0], mixed_npost1=s1[1],
                      mixed_npre2=below[0], mixed_npost2=below[1], mixed_npre3=below3[0], mixed_npost3=below3[1])
    Um, rm, nm, rc = sm.solve(Ke, F, CX, fixed, 1e-10, 300)
    print("iterations", n64, nm, flush=True)
    assert rc == 0 and rm < 1e-10 and nm <= n64 + 10
    Uc = Um.copy()
    Uc[1::3] *= -1
    r = oracle.defect0(Ke, F, Uc, CX, fixed, nelx, nely, nelz)
    assert np.linalg.norm(r) / np.linalg.norm(F) < 1e-10
    assert abs(F @ Um - F @ U64) <= 1e-9 * abs(F @ U64)
    s64.close()
    sm.close()


@pytest.mark.parametrize("lag_mode", [1, 2])
def test_lagged_coarse_factor(gputo, oracle, lag_mode):
    """Mixed solves of a sequence of slowly changing designs with the
    coarsest factorisation lagged by one setup (for the whole solve, or for
    its first 2 V-cycles) reach the FP64 solution of each design to the CG
    tolerance, in about the iteration count of the exact factorisation."""
    nelx, nely, nelz, nl = 48, 24, 24, 4
    Ke, F, fixed = gputo.cantilever(nelx, nely, nelz)
    lag = gputo.Solver(nelx, nely, nelz, nl, precision=1, coarse_lag=lag_mode)
    exact = gputo.Solver(nelx, nely, nelz, nl, precision=1)
    s64 = gputo.Solver(nelx, nely, nelz, nl)
    for k in range(4):
        print("design", k, flush=True)
        x = gputo.problems.perturbed_density(nelx, nely, nelz, 0.12 + 0.01 * k, seed=40 + k)
        CX = 1e-9 + x ** 3 * (1 - 1e-9)
        U64, _, n64, rc = s64.solve(Ke, F, CX, fixed, 1e-10, 300)
        assert rc == 0
        Ue, re, ne, rc = exact.solve(Ke, F, CX, fixed, 1e-10, 300)
        assert rc == 0 and re < 1e-10
        Ul, rl, nlag, rc = lag.solve(Ke, F, CX, fixed, 1e-10, 300)
        assert rc == 0 and rl < 1e-10
        print("iterations", n64, ne, nlag, flush=True)
        assert rel(Ul, U64) < 1e-7 and rel(Ue, U64) < 1e-7
        if k == 0 and lag_mode == 1:
            assert np.array_equal(Ul, Ue)  # first setup: no earlier factor, exact
        assert nlag <= ne + 4
    for s in (lag, exact, s64):
        s.close()


def test_lagged_factor_dropped_after_parameter_change(gputo):
    """A lagged coarsest factorisation still in flight when the parameters
    change is not adopted: the first solve after the change factors exactly,
    so it matches a fresh handle with the new parameters (same iteration
    count, U to 1e-12; the pending factor of the old preset's hierarchy
    would change the preconditioner, and U at the 1e-10 tolerance level).
    Not bit for bit: a handle created with the REF_C preset has no node
    stencil for level 1, which the Galerkin preset stores, so the two apply
    level 1 in different summation orders."""
    nelx, nely, nelz, nl = 32, 16, 16, 4
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=61, kind="perturbed")
    x2 = gputo.problems.perturbed_density(nelx, nely, nelz, 0.15, seed=62)
    CX2 = 1e-9 + x2 ** 3 * (1 - 1e-9)
    s = gputo.Solver(nelx, nely, nelz, nl, precision=1, coarse_lag=1)
    for cx in (CX, CX2):  # the second setup leaves a lagged factor pending
        U, r, n, rc = s.solve(Ke, F, cx, fixed, 1e-10, 300)
        assert rc == 0
    # still lagged after the change: without the drop, this setup would use
    # the pending factor of the old preset's hierarchy
    s.set_params(precision=1, coarse_lag=1, preset=gputo.PRESET_MATLAB_CFG0)
    U1, r1, n1, rc = s.solve(Ke, F, CX2, fixed, 1e-10, 300)
    assert rc == 0
    fresh = gputo.Solver(nelx, nely, nelz, nl, precision=1, coarse_lag=1, preset=gputo.PRESET_MATLAB_CFG0)
    U2, r2, n2, rc = fresh.solve(Ke, F, CX2, fixed, 1e-10, 300)
    assert rc == 0
    print("after change: nit %d, fresh handle: nit %d, |dU|/|U| %.2e" % (n1, n2, rel(U1, U2)), flush=True)
    assert n1 == n2 and rel(U1, U2) < 1e-12
    s.close()
    fresh.close()
