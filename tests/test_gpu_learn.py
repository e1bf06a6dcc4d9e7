"""Learned level-2 correction: the HIP network (forward and parameter
gradients) against the torch-CPU restatement in oracle/learn_oracle.py,
training on the device, and the solve it accelerates.

Tolerances: forward 1e-5 and gradients 1e-4 relative to the largest entry
(FP32 activations and accumulation against float64).  The learned solve
converges to the FP64 tolerance (the CG decides), so its displacements
agree with the plain solve to that tolerance.
"""
import numpy as np
import pytest

import learn_oracle
from conftest import cantilever_case

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def net_solver(gputo):
    s = gputo.Solver(32, 16, 16, 4, precision=1, learned=True)
    yield s
    s.close()


def test_network_layout(net_solver):
    info = net_solver.learn_info()
    assert info["nparam"] == learn_oracle.nparams()
    assert info["grid"] == (9, 5, 5)  # level-2 nodes of 32x16x16: (nnx, nnz, nny)


@pytest.mark.parametrize("seed", [0, 1])
def test_network_forward_and_gradient_match_cpu_oracle(net_solver, seed):
    rng = np.random.default_rng(40 + seed)
    p = rng.normal(size=learn_oracle.nparams()) * 0.05
    net_solver.learn_params(p)
    n0, n1, n2 = net_solver.learn_info()["grid"]
    x = rng.normal(size=(4, n0, n1, n2))
    gy = rng.normal(size=(3, n0, n1, n2))
    p32 = net_solver.learn_params()  # the FP32 parameters the device holds
    y_ref, g_ref = learn_oracle.run(p32, x.astype(np.float32).astype(np.float64), gy.astype(np.float32))
    y = net_solver.learn_forward(x)
    assert rel(y, y_ref) < 1e-5
    g = net_solver.learn_grad(x, gy)
    assert rel(g, g_ref) < 1e-4
    # per layer as well, so that a small layer's error is not hidden
    off = 0
    for ci, co, _, _ in learn_oracle.LAYERS:
        n = co * ci * 27 + co
        assert rel(g[off:off + n], g_ref[off:off + n]) < 1e-3
        off += n


@pytest.mark.parametrize("shape,nl", [((96, 48, 48), 4), ((192, 96, 96), 5)])
def test_network_on_larger_grids(gputo, shape, nl):
    """The level-2 grids of larger meshes (25x13x13 and 49x25x25 nodes),
    where the convolutions split their positions over many blocks and their
    input channels over fewer waves than on the small grid."""
    s = gputo.Solver(*shape, nl, precision=1, learned=True)
    try:
        rng = np.random.default_rng(7)
        p = rng.normal(size=learn_oracle.nparams()) * 0.05
        s.learn_params(p)
        n0, n1, n2 = s.learn_info()["grid"]
        assert (n0, n1, n2) == (shape[0] // 4 + 1, shape[2] // 4 + 1, shape[1] // 4 + 1)
        x = rng.normal(size=(4, n0, n1, n2))
        gy = rng.normal(size=(3, n0, n1, n2))
        p32 = s.learn_params()
        y_ref, g_ref = learn_oracle.run(p32, x.astype(np.float32).astype(np.float64), gy.astype(np.float32))
        y = s.learn_forward(x)
        print(shape, "forward", rel(y, y_ref))
        assert rel(y, y_ref) < 1e-5
        g = s.learn_grad(x, gy)
        assert rel(g, g_ref) < 1e-4
    finally:
        s.close()


def test_training_lowers_the_level2_energy(gputo):
    nelx, nely, nelz, nl = 48, 24, 24, 4
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=41, kind="smooth")
    s = gputo.Solver(nelx, nely, nelz, nl, precision=1, learned=True)
    U, r, n, rc = s.solve(Ke, F, CX, fixed, 1e-8, 200)
    assert rc == 0 and s.learn_info()["nsamples"] > 0
    l0, l1 = s.learn_train(200, 1e-3)
    print("energy loss", l0, "->", l1)
    assert l1 < l0 and l1 < 0.0  # below zero: the correction lowers the energy
    s.close()


def test_learned_solve_matches_plain(gputo):
    nelx, nely, nelz, nl = 48, 24, 24, 4
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=42, kind="smooth")
    s64 = gputo.Solver(nelx, nely, nelz, nl)
    U64, r64, n64, rc = s64.solve(Ke, F, CX, fixed, 1e-10, 300)
    s = gputo.Solver(nelx, nely, nelz, nl, precision=1, learned=True)
    U0, r0, n0, rc0 = s.solve(Ke, F, CX, fixed, 1e-10, 300)  # untrained: logs samples only
    s.learn_train(300, 1e-3)
    U1, r1, n1, rc1 = s.solve(Ke, F, CX, fixed, 1e-10, 300)
    print("iterations: fp64 %d, mixed %d, mixed+learned %d" % (n64, n0, n1))
    assert rc0 == 0 and rc1 == 0 and r1 < 1e-10
    assert rel(U1, U64) < 1e-7
    C64, C1 = F @ U64, F @ U1
    assert abs(C1 - C64) <= 1e-9 * abs(C64)
    s.close()
    s64.close()


def test_trained_network_gain(gputo):
    """What the trained network gains: the same problem solved by the mixed
    MGCG without and with the stage (learned = 1) after training on the
    defects of the first solve; the stage must not cost iterations here
    (the problem it was trained on)."""
    nelx, nely, nelz, nl = 48, 24, 24, 4
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=43, kind="smooth")
    s = gputo.Solver(nelx, nely, nelz, nl, precision=2, l0_kernel=3, learned=True)
    U0, r0, n0, rc0 = s.solve(Ke, F, CX, fixed, 1e-10, 300)  # untrained: the plain V-cycle, logs samples
    s.learn_train(300, 1e-3)
    U1, r1, n1, rc1 = s.solve(Ke, F, CX, fixed, 1e-10, 300)
    print("CG iterations: mixed %d, mixed + trained stage %d (gain %d)" % (n0, n1, n0 - n1), flush=True)
    assert rc0 == 0 and rc1 == 0 and r1 < 1e-10
    assert n1 <= n0
    s.close()


def design_loop(gputo, s, n, x0, Ke, F, fixed):
    run = gputo.TopOpt(s, Ke, x0, F, fixed, volfrac=0.12, penal=3.0, rmin=np.sqrt(3.0), ft=1, move=0.2,
                       cg_tol=1e-8, cg_maxit=300)
    hist = [run.iterate() for _ in range(n)]
    run.close()
    return hist


def pretrain_small(gputo, s, Ke, fixed, nelx, nely, nelz, problems=4, steps=50):
    """The benchmark's pretraining protocol in small: plain solves of
    cantilevers with random volume fractions and load heights, Adam steps on
    each solve's logged level-2 defects."""
    rng = np.random.default_rng(5)
    nel, nn = nelx * nely * nelz, (nely + 1) * (nelz + 1)
    for _ in range(problems):
        x = np.clip(rng.uniform(0.08, 0.3) + rng.uniform(-0.01, 0.01, nel), 1e-3, 1.0)
        CX = 1e-9 + x ** 3 * (1 - 1e-9)
        F = np.zeros(3 * (nelx + 1) * nn)
        nodes = nelx * nn + (nely + 1) * int(rng.integers(0, nelz + 1)) + np.arange(nely + 1)
        F[3 * nodes + 2] = -1.0 / (nely + 1)
        U, r, n, rc = s.solve(Ke, F, CX, fixed, 1e-8, 300)
        assert rc == 0
        s.learn_train(steps, 1e-3)
    s.learn_logging(False)


@pytest.mark.parametrize("net", ["pretrained", "random"])
def test_fallback_over_a_design_loop(gputo, net):
    """12 successive designs of the 48x24x24 cantilever with the learned
    stage under the solve-level fallback (learned = 2) against the plain
    mixed MGCG.  Every solve reaches the tolerance, so the design
    trajectories agree; the fallback keeps the stage only while it shortens
    the solves, so the CG iterations summed over the loop stay within the
    plain method's (5% and two iterations of slack: the designs differ at
    the solver tolerance, which moves single counts by one).  A random
    network (every V-cycle still takes the energy-optimal step along its
    output) costs time without saving iterations: the fallback must switch
    it off after its trial."""
    nelx, nely, nelz, nl = 48, 24, 24, 4
    Ke, F, fixed = gputo.cantilever(nelx, nely, nelz)
    x0 = gputo.problems.perturbed_density(nelx, nely, nelz, 0.12, seed=0)
    kw = dict(precision=2, l0_kernel=3, mixed_npre0=1, mixed_npost0=1, mixed_npre1=1, mixed_npost1=1,
              mixed_npre2=6, mixed_npost2=6)
    plain = gputo.Solver(nelx, nely, nelz, nl, **kw)
    hp = design_loop(gputo, plain, 12, x0, Ke, F, fixed)
    plain.close()
    s = gputo.Solver(nelx, nely, nelz, nl, learned=2, **kw)
    if net == "pretrained":
        pretrain_small(gputo, s, Ke, fixed, nelx, nely, nelz)
    else:
        s.learn_params(np.random.default_rng(3).normal(size=s.learn_info()["nparam"]) * 0.05)
        s.learn_logging(False)
    hl = design_loop(gputo, s, 12, x0, Ke, F, fixed)
    s.close()
    on = [k + 1 for k, h in enumerate(hl) if h["learned_accepted"] + h["learned_rejected"] > 0]
    np_, nl_ = [h["cg_iters"] for h in hp], [h["cg_iters"] for h in hl]
    print("%s network: stage on in designs %s; CG plain %s (sum %d), learned+fallback %s (sum %d)"
          % (net, on, np_, sum(np_), nl_, sum(nl_)), flush=True)
    for a, b in zip(hp, hl):
        assert b["relres"] < 1e-8
        assert abs(a["compliance"] - b["compliance"]) <= 1e-6 * abs(a["compliance"])
    assert sum(nl_) <= 1.05 * sum(np_) + 2
    assert 1 not in on  # the first solve after training is the plain reference
    if net == "random":
        assert on[:1] == [2] and len(on) <= 2, on  # the trial, and at most one more after the backoff
