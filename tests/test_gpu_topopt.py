"""Design-loop parity: the device-resident top3D125mod.m loop against the
oracle's restatement of the same loop, on identical inputs.

Bars: compliance at design iteration 1 within 1e-9 relative (both solves
converged to 1e-10); density and compliance trajectories within 1e-3 over
the run.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ft", [1, 3])
def test_design_loop_matches_oracle(gputo, oracle, ft):
    nelx, nely, nelz, nl = 24, 12, 12, 3
    iters = 12
    print("case ft=%d" % ft, flush=True)
    Ke, F, fixed = gputo.cantilever(nelx, nely, nelz)
    x0 = gputo.problems.perturbed_density(nelx, nely, nelz, 0.12, seed=0)
    cfg = dict(volfrac=0.12, penal=3.0, rmin=np.sqrt(3.0) if ft == 1 else 2 * np.sqrt(3.0), eta=0.4, beta=2.0,
               move=0.2, ft=ft, cg_tol=1e-10, cg_maxit=300)
    s = gputo.Solver(nelx, nely, nelz, nl)
    run = gputo.TopOpt(s, Ke, x0, F, fixed, **cfg)
    st = oracle.DesignState(nelx, nely, nelz, x0, eta=0.4)
    for it in range(iters):
        g = run.iterate()
        o = oracle.design_iter(st, Ke, F, fixed, nl, 1e-10, 300, 0.12, 3.0, cfg["rmin"], ft, 2.0, 0.2)
        assert o["rc"] == 0 and g["status"] == 0
        tol = 1e-9 if it == 0 else 1e-3
        assert abs(g["compliance"] - o["C"]) <= tol * abs(o["C"]), (it, g["compliance"], o["C"])
        assert abs(g["volume"] - o["V"]) <= 1e-3
        if ft == 3:
            assert abs(g["eta"] - o["eta"]) <= 1e-3
    x, xp, U = run.state()
    assert np.abs(xp - st.xPhys).max() <= 1e-3
    assert np.abs(x - st.x).max() <= 1e-3
    run.close()
    s.close()


def test_design_loop_log(gputo, tmp_path):
    """The loop writes the reference driver's log: header, then one line per
    design iteration carrying that iteration's compliance and CG figures."""
    nelx, nely, nelz, nl = 16, 8, 8, 3
    Ke, F, fixed = gputo.cantilever(nelx, nely, nelz)
    x0 = np.full(nelx * nely * nelz, 0.12)
    s = gputo.Solver(nelx, nely, nelz, nl)
    path = str(tmp_path / gputo.log.filename(nl, nelx, nely, nelz))
    run = gputo.TopOpt(s, Ke, x0, F, fixed, log=path, maxit=3, cg_tol=1e-10, cg_maxit=300)
    hist = [run.iterate() for _ in range(3)]
    run.close()
    s.close()
    text = open(path).read()
    assert text.startswith("Top3D125mod \n\nnl: 3, nelx:  16, nely:   8, nelz:   8, volfrac:0.12, penal: 3,")
    lines = [l for l in text.splitlines() if l.startswith("It.:")]
    assert len(lines) == 3
    for k, (l, h) in enumerate(zip(lines, hist)):
        assert l.startswith("It.:%5i C:%6.5e Time:" % (k + 1, h["compliance"]))
        assert "cgiters: %3d MemUsed:" % h["cg_iters"] in l


@pytest.mark.parametrize("bad", [0.0, -3.0, float("nan"), 1e12])
def test_to_begin_rejects_bad_fixed_list(gputo, bad):
    """A fixed list with an entry outside 1..ndof (or NaN) is rejected before
    any design-loop state exists, and a following iterate call is rejected
    too instead of launching kernels on unset buffers."""
    import ctypes
    from gputo import _lib
    nelx, nely, nelz, nl = 8, 4, 4, 2
    Ke, F, fixed = gputo.cantilever(nelx, nely, nelz)
    s = gputo.Solver(nelx, nely, nelz, nl)
    lib = s.lib
    fx = np.array(fixed, dtype=np.float64)
    fx[len(fx) // 2] = bad
    cfg = _lib.ToConfig(volfrac=0.12, penal=3.0, rmin=np.sqrt(3.0), eta=0.4, beta=2.0, move=0.2, Emin=1e-9, E0=1.0,
                        ft=1, cg_tol=1e-8, cg_maxit=100)
    x0 = np.full(nelx * nely * nelz, 0.12)
    rc = lib.gputo_to_begin(s.h, _lib.ptr(Ke), _lib.ptr(x0), _lib.ptr(F), _lib.ptr(fx), len(fx), ctypes.byref(cfg))
    assert rc == _lib.GPUTO_ERR_ARGUMENT
    st = _lib.ToStats()
    assert lib.gputo_to_iterate(s.h, ctypes.byref(st)) == _lib.GPUTO_ERR_ARGUMENT
    # the handle stays usable: a valid loop starts and iterates
    run = gputo.TopOpt(s, Ke, x0, F, fixed, cg_tol=1e-8, cg_maxit=100)
    assert run.iterate()["status"] == 0
    run.close()
    s.close()
