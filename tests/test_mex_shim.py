"""The MEX shims of gpu-to_amd/mex, compiled against the MEX API emulation
of tests/mex: argument checking and the OpenMP form's workspace test,
which return before any device work (no GPU needed)."""
import numpy as np
import pytest

from conftest import cantilever_case
from mex_shim import MexFunction


@pytest.fixture(scope="module")
def mgcg1():
    return MexFunction("mgcg1")


@pytest.fixture(scope="module")
def mgcg9():
    return MexFunction("mgcg9")


def test_argument_counts(mgcg1, mgcg9):
    with pytest.raises(RuntimeError, match="12 or 13"):
        mgcg1(*([0.0] * 11))
    with pytest.raises(RuntimeError, match="18"):
        mgcg9(*([0.0] * 17))


def test_array_sizes_checked(mgcg1, mgcg9):
    nelx, nely, nelz = 8, 4, 4
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz)
    U = np.zeros(F.size)
    with pytest.raises(RuntimeError, match="sizes"):
        mgcg1(Ke, F[:-3], U, CX, fixed, len(fixed), nelx, nely, nelz, 2, 1e-8, 50)
    with pytest.raises(RuntimeError, match="sizes"):
        mgcg9(Ke, F, U, CX[:-1], fixed, len(fixed), nelx, nely, nelz, 2, 1e-8, 50, 2, 0, 0.3, 1, 1, 1)


def test_openmp_form_short_workspace(mgcg1, gputo):
    """A WRK shorter than estWorkSpace (Cantilever/omp/mgcg1.c:89-97, 181-205)
    gives cgiters = -1 and leaves U untouched; the same as the Python mirror."""
    nelx, nely, nelz, nl = 8, 4, 4, 2
    Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz)
    lw = gputo.load().gputo_est_workspace(len(fixed), nl, nelx, nely, nelz)
    U = np.full(F.size, 7.0)
    WRK = np.zeros(lw - 1)
    res, nit = mgcg1(Ke, F, U, CX, fixed, WRK, len(fixed), nelx, nely, nelz, nl, 1e-8, 50)
    assert (res, nit) == (0.0, -1.0)
    assert np.all(U == 7.0)
    assert gputo.mgcg1(Ke, F, U, CX, fixed, WRK, len(fixed), nelx, nely, nelz, nl, 1e-8, 50) == (0.0, -1.0)
