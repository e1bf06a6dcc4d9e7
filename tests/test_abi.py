"""The C-ABI library loads and exports every symbol include/gputo.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "gputo.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gputo_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported(gputo):
    from gputo import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    from gputo import _lib
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_workspace_estimate_formula(gputo):
    """estWorkSpace of mgcg1.cu:293-322, in doubles (pure host arithmetic)."""
    lib = gputo.load()
    nelx, nely, nelz, nl, nfd = 48, 24, 24, 4, 1875
    ndof0 = 3 * 49 * 25 * 25
    lw = ndof0 * 6 + nelx * nely * nelz + nfd
    for lev in range(nl):
        ex, ey, ez = nelx >> lev, nely >> lev, nelz >> lev
        nd = 3 * (ex + 1) * (ey + 1) * (ez + 1)
        lw += 2 * nd if lev == 0 else 4 * nd + 300 * ex * ey * ez + nfd
        if lev == nl - 1:
            lw += 3 * nd + nd * nd
    assert lib.gputo_est_workspace(nfd, nl, nelx, nely, nelz) == lw


def test_periodic_argument_checks(gputo):
    """A periodic handle needs 2 elements per direction on its coarsest level
    and no learned stage; both are rejected before any device call."""
    import ctypes
    from gputo import _lib
    lib = gputo.load()
    p = _lib.GputoParams()
    lib.gputo_default_params(ctypes.byref(p))
    p.periodic = 1
    h = ctypes.c_void_p()
    # 8 elements, 4 levels: the coarsest level has 1 element per direction
    assert lib.gputo_create(8, 8, 8, 4, ctypes.byref(p), ctypes.byref(h)) == _lib.GPUTO_ERR_ARGUMENT
    p.precision = _lib.PREC_MIXED_F32
    p.learned = 1
    assert lib.gputo_create(32, 32, 32, 4, ctypes.byref(p), ctypes.byref(h)) == _lib.GPUTO_ERR_ARGUMENT
    assert not h.value


def test_odd_sweep_total_rejected(gputo):
    """An odd number of level-0 smoothing steps (the prerelaxation counts as
    the first pre-sweep) makes the V-cycle indefinite: Jacobi's error factor
    is negative on the highest modes.  The handle refuses such schedules."""
    with pytest.raises(gputo.GputoError):
        gputo.Solver(16, 8, 8, 3, precision=2, l0_kernel=3, mixed_npre0=1, mixed_npost0=2)
    with pytest.raises(gputo.GputoError):
        gputo.Solver(16, 8, 8, 3, precision=2, l0_kernel=3, mixed_npre0=2, mixed_npost0=0)
    with pytest.raises(gputo.GputoError):
        gputo.Solver(16, 8, 8, 3, precision=2, l0_kernel=3, mixed_npre1=1, mixed_npost1=2)
    with pytest.raises(gputo.GputoError):
        gputo.Solver(32, 16, 16, 4, precision=2, l0_kernel=3, mixed_npre2=2, mixed_npost2=1)
    with pytest.raises(gputo.GputoError):
        gputo.Solver(64, 32, 32, 5, precision=2, l0_kernel=3, mixed_npre3=2, mixed_npost3=1)
    with pytest.raises(gputo.GputoError):  # -1 (no sweep) is the only negative count
        gputo.Solver(16, 8, 8, 3, precision=2, l0_kernel=3, mixed_npre1=-2, mixed_npost1=1)
    with pytest.raises(gputo.GputoError):  # no pre-sweep: the prerelaxation and two post-sweeps are odd
        gputo.Solver(16, 8, 8, 3, precision=2, l0_kernel=3, mixed_npre1=-1, mixed_npost1=2)
