"""The MEX shims of gpu-to_amd/mex (mgcg1: the CUDA 12- and 13-argument
forms and the OpenMP form with WRK; mgcg9: 18 arguments), compiled against
the MEX API emulation of tests/mex and called as MATLAB would, against the
Python mirror of the same entry points (gputo.mgcg1 / gputo.mgcg9) and the
oracle.  The shims and the mirror call the same library entry point on
equal fresh handles, so the results agree bit for bit."""
import numpy as np
import pytest

from conftest import cantilever_case
from mex_shim import MexFunction

pytestmark = pytest.mark.gpu


def fresh_mirror(gputo):
    gputo.release_handles()
    return gputo


def test_mgcg1_forms(gputo, oracle):
    f = MexFunction("mgcg1")
    nelx, nely, nelz, nl = 16, 8, 8, 3
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=90)
    lw = gputo.load().gputo_est_workspace(len(fixed), nl, nelx, nely, nelz)
    forms = {
        "cuda12": lambda U: (Ke, F, U, CX, fixed, len(fixed), nelx, nely, nelz, nl, 1e-10, 200),
        "cuda13": lambda U: (Ke, F, U, CX, fixed, len(fixed), nelx, nely, nelz, nl, 1e-10, 200, 0),
        "omp13": lambda U: (Ke, F, U, CX, fixed, np.zeros(lw), len(fixed), nelx, nely, nelz, nl, 1e-10, 200),
    }
    Uo, ro, no, rc = oracle.mgcg(Ke, F, CX, fixed, nelx, nely, nelz, nl, 1e-10, 200)
    for name, args in forms.items():
        print("mgcg1 form", name, flush=True)
        U1 = np.zeros(F.size)
        r1, n1 = f(*args(U1))
        U2 = np.zeros(F.size)
        r2, n2 = fresh_mirror(gputo).mgcg1(*args(U2))
        assert n1 >= 0 and (r1, n1) == (r2, n2) and np.array_equal(U1, U2), name
        assert abs(n1 - no) <= 1 and np.abs(U1 - Uo).max() <= 1e-8 * np.abs(Uo).max()
    assert f.unload() == 1  # the shim's mexAtExit function ran (handles destroyed)
    gputo.release_handles()


def test_mgcg9_form(gputo):
    f = MexFunction("mgcg9")
    nelx, nely, nelz, nl = 16, 8, 8, 3
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=91)
    args = lambda U: (Ke, F, U, CX, fixed, len(fixed), nelx, nely, nelz, nl, 1e-8, 300, 2, 0, 0.3, 1.0, 1.0, 1.0)
    U1 = np.zeros(F.size)
    r1, n1 = f(*args(U1))
    U2 = np.zeros(F.size)
    r2, n2 = fresh_mirror(gputo).mgcg9(*args(U2))
    print("mgcg9: nit %d relres %.3e" % (n1, r1), flush=True)
    assert n1 >= 0 and r1 < 1e-8
    assert (r1, n1) == (r2, n2) and np.array_equal(U1, U2)
    assert f.unload() == 1
    gputo.release_handles()
