"""ctypes view of the CPU oracle (liboracle.so, built from gputo_oracle.c).

Test infrastructure only: tests/, the smoke check and bench.py's
cpu_baseline leg use it as the checker and as the CPU baseline.  Every
function restates the reference algorithm cited in gputo_oracle.c.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

PRESET_REF_C = 0
PRESET_MATLAB_CFG0 = 1
PRESET_MGCG9 = 2

_d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_c_int = ctypes.c_int
_c_dbl = ctypes.c_double
_c_long = ctypes.c_long


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        sig = {
            "or_kem3": (None, [_c_dbl, _c_dbl, _c_dbl, _c_dbl, _d]),
            "or_aelm": (None, [_c_dbl, _c_dbl, _d]),
            "or_edofmat_matlab": (None, [_c_int, _c_int, _c_int, _i32]),
            "or_elem_nodes_c": (None, [_c_int, _c_int, _c_int, _i32]),
            "or_rstrfd": (_c_int, [_d, _d, _c_int, _c_int, _c_int, _c_int]),
            "or_matvec0": (None, [_d, _d, _d, _d, _d, _c_int, _c_int, _c_int, _c_int]),
            "or_defect0": (None, [_d, _d, _d, _d, _d, _d, _c_int, _c_int, _c_int, _c_int]),
            "or_matvec_el": (None, [_d, _d, _d, _d, _c_int, _c_int, _c_int, _c_int]),
            "or_defect_el": (None, [_d, _d, _d, _d, _d, _c_int, _c_int, _c_int, _c_int]),
            "or_restrict": (None, [_d, _d, _d, _c_int, _c_int, _c_int, _c_int]),
            "or_interp": (None, [_d, _d, _c_int, _c_int, _c_int]),
            "or_coarse_l1": (None, [_d, _d, _c_dbl, _c_dbl, _c_int, _c_int, _c_int]),
            "or_prolong_local": (None, [_d]),
            "or_corrcsa": (None, [_d, _d, _c_int, _c_int, _c_int, _c_int, _c_int]),
            "or_assemble": (None, [_d, _d, _c_int, _c_int, _c_int]),
            "or_choldc": (_c_int, [_c_int, _d, _d]),
            "or_cholsl": (None, [_c_int, _d, _d, _d, _d]),
            "or_mgcg": (_c_int, [_c_int, _d, _d, _d, _d, _d, _c_int, _c_int, _c_int, _c_int, _c_int,
                                 _c_dbl, _c_int, ctypes.POINTER(_c_dbl), ctypes.POINTER(_c_int), _c_int]),
            "or_vcycle_once": (_c_int, [_c_int, _d, _d, _d, _c_int, _c_int, _c_int, _c_int, _c_int, _d, _d]),
            "or_filter_kernel": (_c_int, [_c_dbl, _d]),
            "or_imfilter": (None, [_d, _d, _c_int, _c_int, _c_int, _c_dbl]),
            "or_eta_newton": (_c_dbl, [_d, _c_long, _c_dbl, _c_dbl, _c_dbl, ctypes.POINTER(_c_int)]),
            "or_sens_matlab": (None, [_d, _d, _i32, _c_long, _d]),
            "or_oc": (_c_dbl, [_d, _d, _d, _c_long, _c_dbl, _c_dbl, ctypes.POINTER(_c_int)]),
            "or_design_iter": (_c_int, [_c_int, _d, _d, _d, _d, _c_int, _c_int, _c_int, _c_int, _c_int,
                                        _c_dbl, _c_int, _d, _d, _d, _c_dbl, _c_dbl, _c_dbl, _c_int,
                                        ctypes.POINTER(_c_dbl), _c_dbl, _c_dbl, _c_dbl, _c_dbl, _d]),
            "or_set_nswp": (None, [_c_int]),
            "or_level_cx": (_c_int, [_d, _d, _d, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _d]),
            "or_coarse_cx": (None, [_d, _d, _c_int, _c_int, _c_int]),
            "or_num_threads": (_c_int, []),
            "or_set_num_threads": (None, [_c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        # or_coarse_galerkin / or_prepad / or_level_data take optional (NULL) arrays
        L.or_coarse_galerkin.restype = None
        L.or_coarse_galerkin.argtypes = [ctypes.c_void_p] * 4 + [_c_int] * 3
        L.or_prepad.restype = None
        L.or_prepad.argtypes = [ctypes.c_void_p] * 4 + [_c_int] * 3
        L.or_level_data.restype = _c_int
        L.or_level_data.argtypes = ([_c_int, _d, _d, _d, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int]
                                    + [ctypes.c_void_p] * 3)
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def ndof_of(nelx, nely, nelz):
    return 3 * (nelx + 1) * (nely + 1) * (nelz + 1)


def kem3(nu=0.3, h=1.0):
    Ke = np.zeros(576)
    lib().or_kem3(nu, h, h, h, Ke)
    return Ke.reshape(24, 24)


def aelm(nu=0.3, h=2.0):
    A = np.zeros(8 * 576)
    lib().or_aelm(nu, h, A)
    return A.reshape(8, 24, 24)


def edofmat_matlab(nelx, nely, nelz):
    c = np.zeros(nelx * nely * nelz * 24, dtype=np.int32)
    lib().or_edofmat_matlab(nelx, nely, nelz, c)
    return c.reshape(-1, 24)


def elem_nodes_c(nelx, nely, nelz):
    c = np.zeros(nelx * nely * nelz * 8, dtype=np.int32)
    lib().or_elem_nodes_c(nelx, nely, nelz, c)
    return c.reshape(-1, 8)


def rstrfd(fd_fine, nelx_c, nely_c, nelz_c):
    fd_fine = np.ascontiguousarray(fd_fine, dtype=np.float64)
    out = np.zeros(max(len(fd_fine), 1))
    n = lib().or_rstrfd(fd_fine, out, nelx_c, nely_c, nelz_c, len(fd_fine))
    return out[:n]


def matvec0(Ke, CX, U, fixed, nelx, nely, nelz):
    R = np.zeros_like(U)
    fixed = np.ascontiguousarray(fixed, dtype=np.float64)
    lib().or_matvec0(np.ascontiguousarray(Ke, dtype=np.float64).ravel(), CX, U, R,
                     fixed if len(fixed) else np.zeros(1), len(fixed), nelx, nely, nelz)
    return R


def defect0(Ke, F, U, CX, fixed, nelx, nely, nelz):
    R = np.zeros_like(U)
    fixed = np.ascontiguousarray(fixed, dtype=np.float64)
    lib().or_defect0(np.ascontiguousarray(Ke, dtype=np.float64).ravel(), F, U, R, CX,
                     fixed if len(fixed) else np.zeros(1), len(fixed), nelx, nely, nelz)
    return R


def matvec_el(A, U, fixed, nelx, nely, nelz):
    R = np.zeros_like(U)
    fixed = np.ascontiguousarray(fixed, dtype=np.float64)
    lib().or_matvec_el(np.ascontiguousarray(A).ravel(), U, R, fixed if len(fixed) else np.zeros(1),
                       len(fixed), nelx, nely, nelz)
    return R


def prepad0(Ke, CX, nelx, nely, nelz):
    AD = np.zeros(ndof_of(nelx, nely, nelz))
    Ke = np.ascontiguousarray(Ke, dtype=np.float64).ravel()
    lib().or_prepad(_p(Ke), None, _p(CX), _p(AD), nelx, nely, nelz)
    return AD


def prepad_el(A, nelx, nely, nelz):
    AD = np.zeros(ndof_of(nelx, nely, nelz))
    A = np.ascontiguousarray(A).ravel()
    lib().or_prepad(None, _p(A), None, _p(AD), nelx, nely, nelz)
    return AD


def restrict(R_fine, fixed_c, nelx_c, nely_c, nelz_c):
    F = np.zeros(ndof_of(nelx_c, nely_c, nelz_c))
    fixed_c = np.ascontiguousarray(fixed_c, dtype=np.float64)
    lib().or_restrict(R_fine, F, fixed_c if len(fixed_c) else np.zeros(1), len(fixed_c), nelx_c, nely_c, nelz_c)
    return F


def interp_add(U_coarse, U_fine, nelx, nely, nelz):
    U = np.array(U_fine, dtype=np.float64, copy=True)
    lib().or_interp(U_coarse, U, nelx, nely, nelz)
    return U


def coarse_l1(CX_fine, nelx_c, nely_c, nelz_c, nu=0.3, h=2.0):
    A = np.zeros(nelx_c * nely_c * nelz_c * 576)
    lib().or_coarse_l1(A, CX_fine, nu, h, nelx_c, nely_c, nelz_c)
    return A.reshape(-1, 24, 24)


def coarse_galerkin(nelx_c, nely_c, nelz_c, A_fine=None, Ke=None, CX_fine=None):
    A = np.zeros(nelx_c * nely_c * nelz_c * 576)
    AP = None if A_fine is None else np.ascontiguousarray(A_fine).ravel()
    Kp = None if Ke is None else np.ascontiguousarray(Ke, dtype=np.float64).ravel()
    lib().or_coarse_galerkin(_p(AP), _p(Kp), _p(CX_fine), _p(A), nelx_c, nely_c, nelz_c)
    return A.reshape(-1, 24, 24)


def prolong_local():
    P = np.zeros(24 * 81)
    lib().or_prolong_local(P)
    return P.reshape(24, 81)


def corrcsa(A, fixed, nelx, nely, nelz, mirrored=True):
    A = np.array(A, dtype=np.float64, copy=True)
    fixed = np.ascontiguousarray(fixed, dtype=np.float64)
    flat = A.reshape(-1)
    lib().or_corrcsa(flat, fixed if len(fixed) else np.zeros(1), len(fixed), nelx, nely, nelz, int(mirrored))
    return flat.reshape(A.shape)


def assemble(A, nelx, nely, nelz):
    n = ndof_of(nelx, nely, nelz)
    AQ = np.zeros(n * n)
    lib().or_assemble(np.ascontiguousarray(A).ravel(), AQ, nelx, nely, nelz)
    return AQ.reshape(n, n)


def cholesky_solve(AQ, b):
    n = AQ.shape[0]
    a = np.array(AQ, dtype=np.float64, copy=True).ravel()
    p = np.zeros(n)
    fail = lib().or_choldc(n, a, p)
    x = np.zeros(n)
    lib().or_cholsl(n, a, p, np.ascontiguousarray(b, dtype=np.float64), x)
    return x, fail


def mgcg(Ke, F, CX, fixed, nelx, nely, nelz, nl, tol, maxit, preset=PRESET_REF_C, flip_y=True, nswp=2):
    """mgcg1 semantics (mgcg9 with preset PRESET_MGCG9 and its nswp):
    returns (U, relres, nit, rc); U is in the MATLAB frame (y negated) when
    flip_y."""
    lib().or_set_nswp(nswp)
    U = np.zeros(ndof_of(nelx, nely, nelz))
    res = ctypes.c_double(0.0)
    nit = ctypes.c_int(0)
    fixed = np.ascontiguousarray(fixed, dtype=np.float64)
    rc = lib().or_mgcg(preset, np.ascontiguousarray(Ke, dtype=np.float64).ravel(),
                       np.ascontiguousarray(F, dtype=np.float64), U, np.ascontiguousarray(CX, dtype=np.float64),
                       fixed if len(fixed) else np.zeros(1), len(fixed), nelx, nely, nelz, nl, tol, maxit,
                       ctypes.byref(res), ctypes.byref(nit), int(flip_y))
    return U, res.value, nit.value, rc


def vcycle_once(Ke, CX, fixed, nelx, nely, nelz, nl, r, preset=PRESET_REF_C):
    z = np.zeros_like(r)
    fixed = np.ascontiguousarray(fixed, dtype=np.float64)
    rc = lib().or_vcycle_once(preset, np.ascontiguousarray(Ke, dtype=np.float64).ravel(), CX, fixed,
                              len(fixed), nelx, nely, nelz, nl, np.ascontiguousarray(r), z)
    if rc:
        raise RuntimeError("oracle hierarchy setup failed: %d" % rc)
    return z


def level_cx(Ke, CX, fixed, nelx, nely, nelz, nl, level):
    """Element scales of level `level` of the mgcg9 hierarchy."""
    out = np.zeros((nelx >> level) * (nely >> level) * (nelz >> level))
    fixed = np.ascontiguousarray(fixed, dtype=np.float64)
    rc = lib().or_level_cx(np.ascontiguousarray(Ke, dtype=np.float64).ravel(), np.ascontiguousarray(CX, float),
                           fixed if len(fixed) else np.zeros(1), len(fixed), nelx, nely, nelz, nl, level, out)
    if rc:
        raise RuntimeError("oracle hierarchy setup failed: %d" % rc)
    return out


def level_data(Ke, CX, fixed, nelx, nely, nelz, nl, level, preset=PRESET_REF_C):
    """(A_l [nel_l,24,24] or None, AD_l, fixed_l) of a fresh hierarchy."""
    nx, ny, nz = nelx >> level, nely >> level, nelz >> level
    nel = nx * ny * nz
    A = np.zeros(nel * 576) if level > 0 else None
    AD = np.zeros(ndof_of(nx, ny, nz))
    FD = np.zeros(max(len(fixed), 1))
    fixed = np.ascontiguousarray(fixed, dtype=np.float64)
    n = lib().or_level_data(preset, np.ascontiguousarray(Ke, dtype=np.float64).ravel(), CX, fixed, len(fixed),
                            nelx, nely, nelz, nl, level, _p(A), _p(AD), _p(FD))
    if n < 0:
        raise RuntimeError("oracle hierarchy setup failed: %d" % n)
    return (None if A is None else A.reshape(nel, 24, 24)), AD, FD[:n]


def imfilter(v, nelx, nely, nelz, rmin):
    out = np.zeros_like(v)
    lib().or_imfilter(np.ascontiguousarray(v, dtype=np.float64), out, nelx, nely, nelz, rmin)
    return out


def eta_newton(xPhys, eta, beta, volfrac):
    it = ctypes.c_int(0)
    e = lib().or_eta_newton(np.ascontiguousarray(xPhys, dtype=np.float64), len(xPhys), eta, beta, volfrac,
                            ctypes.byref(it))
    return e, it.value


def sens_matlab(Ke, U_matlab, cmat):
    ce = np.zeros(cmat.shape[0])
    lib().or_sens_matlab(np.ascontiguousarray(Ke, dtype=np.float64).ravel(),
                         np.ascontiguousarray(U_matlab, dtype=np.float64),
                         np.ascontiguousarray(cmat, dtype=np.int32), cmat.shape[0], ce)
    return ce


def oc(x, dc, dv, volfrac, move):
    x = np.array(x, dtype=np.float64, copy=True)
    it = ctypes.c_int(0)
    lmid = lib().or_oc(x, np.ascontiguousarray(dc, dtype=np.float64), np.ascontiguousarray(dv, dtype=np.float64),
                       len(x), volfrac, move, ctypes.byref(it))
    return x, lmid, it.value


class DesignState:
    """State of the top3D125mod.m design loop, advanced by design_iter."""

    def __init__(self, nelx, nely, nelz, x0, eta=0.4):
        self.nelx, self.nely, self.nelz = nelx, nely, nelz
        nel = nelx * nely * nelz
        self.x = np.array(x0, dtype=np.float64, copy=True)
        self.xPhys = np.array(x0, dtype=np.float64, copy=True)
        self.xOld = np.ones(nel)
        self.U = np.zeros(ndof_of(nelx, nely, nelz))
        self.eta = ctypes.c_double(eta)


def design_iter(state, Ke, F, fixed, nl, tol, maxit, volfrac, penal, rmin, ft, beta, move,
                Emin=1e-9, E0=1.0, preset=PRESET_REF_C):
    """One top3D125mod.m design iteration; returns the stats dict."""
    s = state
    stats = np.zeros(8)
    fixed = np.ascontiguousarray(fixed, dtype=np.float64)
    rc = lib().or_design_iter(preset, np.ascontiguousarray(Ke, dtype=np.float64).ravel(),
                              np.ascontiguousarray(F, dtype=np.float64), s.U, fixed, len(fixed),
                              s.nelx, s.nely, s.nelz, nl, tol, maxit, s.x, s.xPhys, s.xOld, volfrac, penal, rmin,
                              ft, ctypes.byref(s.eta), beta, move, Emin, E0, stats)
    keys = ["C", "V", "ch", "eta", "lmid", "cgres", "cgiters", "seconds"]
    out = dict(zip(keys, stats.tolist()))
    out["rc"] = rc
    return out


def set_num_threads(n):
    lib().or_set_num_threads(int(n))


def num_threads():
    return lib().or_num_threads()
