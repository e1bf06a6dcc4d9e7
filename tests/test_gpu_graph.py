"""The FP32 V-cycle of the mixed PCG replayed from a recorded HIP graph
(graph_vcycles = 1) against the same V-cycle launched kernel by kernel from
the host (graph_vcycles = 0, the default): the same solves must give the
same displacements, residual and iteration count bit for bit.

Each solve replays its graph once per CG iteration, with the PCG's own
host-launched kernels and a host synchronisation every few iterations
between the replays; successive designs on one handle record the graph
again for the new hierarchy.  The cases cover what changes the recorded
launches: the bench's sweep schedule, a lagged coarsest factor (switched in
mid-solve with coarse_lag = 2), and the learned level-2 stage (V-cycles that
log a sample are launched from the host, later ones replay the correction).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCHEDULE_1314 = dict(mixed_npre0=1, mixed_npost0=3, mixed_npre1=1, mixed_npost1=1, mixed_npre2=4, mixed_npost2=4)


def designs(gputo, nelx, nely, nelz, n):
    for k in range(n):
        x = gputo.problems.perturbed_density(nelx, nely, nelz, 0.12 + 0.01 * k, seed=60 + k)
        yield 1e-9 + x ** 3 * (1 - 1e-9)


def same(a, b):
    Ua, ra, na, rca = a
    Ub, rb, nb, rcb = b
    assert rca == 0 and rcb == 0
    assert na == nb and ra == rb
    assert np.array_equal(Ua, Ub)


@pytest.mark.parametrize("kw", [
    dict(precision=1, l0_kernel=1),
    dict(precision=2, l0_kernel=3, **SCHEDULE_1314),
    dict(precision=2, l0_kernel=3, coarse_lag=1, **SCHEDULE_1314),
    dict(precision=2, l0_kernel=3, coarse_lag=2, **SCHEDULE_1314),
], ids=["f32_gather", "mfma_1314", "mfma_1314_lag1", "mfma_1314_lag2"])
def test_graph_replay_is_bitwise_eager(gputo, kw):
    nelx, nely, nelz, nl = 48, 24, 24, 4
    Ke, F, fixed = gputo.cantilever(nelx, nely, nelz)
    g = gputo.Solver(nelx, nely, nelz, nl, graph_vcycles=1, **kw)
    e = gputo.Solver(nelx, nely, nelz, nl, **kw)
    for k, CX in enumerate(designs(gputo, nelx, nely, nelz, 3)):
        rg = g.solve(Ke, F, CX, fixed, 1e-10, 300)
        re = e.solve(Ke, F, CX, fixed, 1e-10, 300)
        print("design %d: %d CG iterations, relres %.3g" % (k, rg[2], rg[1]), flush=True)
        same(rg, re)
        assert rg[1] < 1e-10
    g.close()
    e.close()


def test_graph_replay_with_learned_stage(gputo):
    """learned = 1 with a fixed random network and sample logging on: the
    first V-cycles of each solve log their defect (host-launched), the later
    ones replay the recorded correction."""
    nelx, nely, nelz, nl = 48, 24, 24, 4
    Ke, F, fixed = gputo.cantilever(nelx, nely, nelz)
    kw = dict(precision=2, l0_kernel=3, learned=1, **SCHEDULE_1314)
    g = gputo.Solver(nelx, nely, nelz, nl, graph_vcycles=1, **kw)
    e = gputo.Solver(nelx, nely, nelz, nl, **kw)
    plain = gputo.Solver(nelx, nely, nelz, nl, **dict(kw, learned=0))
    theta = np.random.default_rng(7).normal(size=g.learn_info()["nparam"]) * 0.05
    for s in (g, e):
        s.learn_params(theta)
        s.learn_logging(True)
    for k, CX in enumerate(designs(gputo, nelx, nely, nelz, 2)):
        rg = g.solve(Ke, F, CX, fixed, 1e-10, 300)
        re = e.solve(Ke, F, CX, fixed, 1e-10, 300)
        rp = plain.solve(Ke, F, CX, fixed, 1e-10, 300)
        print("design %d: %d CG iterations (without the stage %d), relres %.3g" % (k, rg[2], rp[2], rg[1]),
              flush=True)
        same(rg, re)
        assert not np.array_equal(rg[0], rp[0])  # the stage ran
    for s in (g, e, plain):
        s.close()
