"""The benchmark configurations at their full sizes.

cfg0  48x24x24 cantilever, 4 levels, the MATLAB top3D125mgcg smoother
      (omega 0.6, one sweep down and one up: GPUTO_PRESET_MATLAB_CFG0),
      CG tolerance 1e-10, 20 OC design iterations with move 0.2 and
      rmin sqrt(3), uniform x0 = 0.12: the device loop against the oracle's
      loop on identical inputs (compliance 1e-9 at iteration 1, density and
      compliance within 1e-3 over the run).
cfg1  192x96x96 (1,769,472 elements, 5,447,811 DOFs), 5 levels, plain FP64
      MGCG to 1e-8 on x0 = 0.12 + U(-0.01, 0.01) (seed 0): the recursive
      relres below the tolerance and the true FP64 residual b - K u,
      evaluated by the oracle, below 2e-8 (the two differ by rounding).
cfg2  the same with the FP32 V-cycle (split-BF16 MFMA level 0): compliance
      within 1e-6 of the FP64 solve, the error of u stated and bounded.
cfg3  the same with the learned level-2 stage (trained on the level-2
      defects the first solve logs): compliance within 1e-6 of FP64.
N1    50 design iterations of the reference algorithm (ref_c) at 48x24x24:
      the density and compliance trajectories within 1e-3 of the oracle's.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CFG1 = (192, 96, 96, 5)


@pytest.fixture(scope="module")
def cfg1(gputo):
    nelx, nely, nelz, nl = CFG1
    Ke, F, fixed = gputo.cantilever(nelx, nely, nelz)
    x = gputo.problems.perturbed_density(nelx, nely, nelz, 0.12, seed=0)
    CX = gputo.problems.simp_cx(x)
    s = gputo.Solver(nelx, nely, nelz, nl)
    U, r, n, rc = s.solve(Ke, F, CX, fixed, 1e-8, 500)
    s.close()
    assert rc == 0
    return Ke, F, CX, fixed, U, r, n


def true_relres(oracle, Ke, F, U_matlab, CX, fixed, mesh):
    Uc = U_matlab.copy()
    Uc[1::3] *= -1
    r = oracle.defect0(Ke, F, Uc, CX, fixed, *mesh[:3])
    return np.linalg.norm(r) / np.linalg.norm(F)


def test_config1_fp64_solve(oracle, cfg1):
    Ke, F, CX, fixed, U, r, n = cfg1
    print("cfg1 FP64: nit %d relres %.3e" % (n, r), flush=True)
    assert r < 1e-8
    tr = true_relres(oracle, Ke, F, U, CX, fixed, CFG1)
    print("cfg1 FP64 true relres %.3e" % tr, flush=True)
    assert tr < 2e-8


@pytest.mark.parametrize("learned", [False, True])
def test_config2_3_mixed_solve(gputo, oracle, cfg1, learned):
    Ke, F, CX, fixed, U64, r64, n64 = cfg1
    nelx, nely, nelz, nl = CFG1
    s = gputo.Solver(nelx, nely, nelz, nl, precision=2, l0_kernel=3, learned=learned, mixed_npre0=1,
                     mixed_npost0=1, mixed_npre1=1, mixed_npost1=1, mixed_npre2=6, mixed_npost2=6)
    if learned:
        s.learn_reset(5)
        U, r, n, rc = s.solve(Ke, F, CX, fixed, 1e-8, 500)  # logs level-2 defects
        assert rc == 0
        s.learn_train(200, 1e-3)
    U, r, n, rc = s.solve(Ke, F, CX, fixed, 1e-8, 500)
    s.close()
    assert rc == 0 and r < 1e-8
    tr = true_relres(oracle, Ke, F, U, CX, fixed, CFG1)
    C64, C = F @ U64, F @ U
    eu = np.linalg.norm(U - U64) / np.linalg.norm(U64)
    ec = abs(C - C64) / abs(C64)
    print("cfg%d: nit %d (FP64 %d) true relres %.3e |u-u64|/|u64| %.3e |C-C64|/|C64| %.3e"
          % (3 if learned else 2, n, n64, tr, eu, ec), flush=True)
    assert tr < 2e-8
    assert ec < 1e-6
    assert eu < 1e-6


def test_config0_design_loop(gputo, oracle):
    nelx, nely, nelz, nl = 48, 24, 24, 4
    Ke, F, fixed = gputo.cantilever(nelx, nely, nelz)
    x0 = np.full(nelx * nely * nelz, 0.12)
    cfg = dict(volfrac=0.12, penal=3.0, rmin=np.sqrt(3.0), eta=0.4, beta=2.0, move=0.2, ft=1, cg_tol=1e-10,
               cg_maxit=500)
    s = gputo.Solver(nelx, nely, nelz, nl, preset=gputo.PRESET_MATLAB_CFG0)
    run = gputo.TopOpt(s, Ke, x0, F, fixed, **cfg)
    st = oracle.DesignState(nelx, nely, nelz, x0, eta=0.4)
    for it in range(20):
        g = run.iterate()
        o = oracle.design_iter(st, Ke, F, fixed, nl, 1e-10, 500, 0.12, 3.0, cfg["rmin"], 1, 2.0, 0.2,
                               preset=oracle.PRESET_MATLAB_CFG0)
        print("it %d C %.10e (oracle %.10e) cg %d (%d)" % (it + 1, g["compliance"], o["C"], g["cg_iters"],
                                                           o["cgiters"]), flush=True)
        assert o["rc"] == 0 and g["status"] == 0
        assert g["relres"] < 1e-10
        tol = 1e-9 if it == 0 else 1e-3
        assert abs(g["compliance"] - o["C"]) <= tol * abs(o["C"])
        assert abs(g["volume"] - o["V"]) <= 1e-3
    x, xp, U = run.state()
    assert np.abs(xp - st.xPhys).max() <= 1e-3
    run.close()
    s.close()


def test_design_trajectory_50_iterations(gputo, oracle):
    nelx, nely, nelz, nl = 48, 24, 24, 4
    Ke, F, fixed = gputo.cantilever(nelx, nely, nelz)
    x0 = gputo.problems.perturbed_density(nelx, nely, nelz, 0.12, seed=0)
    cfg = dict(volfrac=0.12, penal=3.0, rmin=np.sqrt(3.0), eta=0.4, beta=2.0, move=0.2, ft=1, cg_tol=1e-10,
               cg_maxit=500)
    s = gputo.Solver(nelx, nely, nelz, nl)
    run = gputo.TopOpt(s, Ke, x0, F, fixed, **cfg)
    st = oracle.DesignState(nelx, nely, nelz, x0, eta=0.4)
    worst_c = worst_x = 0.0
    for it in range(50):
        g = run.iterate()
        o = oracle.design_iter(st, Ke, F, fixed, nl, 1e-10, 500, 0.12, 3.0, cfg["rmin"], 1, 2.0, 0.2)
        assert o["rc"] == 0 and g["status"] == 0
        ec = abs(g["compliance"] - o["C"]) / abs(o["C"])
        worst_c = max(worst_c, ec)
        if it == 0:
            assert ec <= 1e-9
        if it % 10 == 9:
            x, xp, U = run.state()
            worst_x = max(worst_x, np.abs(xp - st.xPhys).max())
            print("it %d C %.8e rel.diff %.2e max|dxPhys| %.2e" % (it + 1, g["compliance"], ec, worst_x), flush=True)
    assert worst_c <= 1e-3 and worst_x <= 1e-3
    run.close()
    s.close()
