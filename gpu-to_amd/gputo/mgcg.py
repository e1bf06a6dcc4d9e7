"""MATLAB-facing solve entry points.

`mgcg1` keeps the argument list and semantics of the reference's MEX solve
so driver code ports line for line:

    [cgres, cgiters] = mgcg1(Ke0, F, U, CX, fixed', nfd, nelx, nely, nelz, nl, cgtol, cgmax[, printLev])
        (Cantilever/TOP3D125MGCGCUDA/mgcg1.cu:120-182)
    [cgres, cgiters] = mgcg1(Ke0, F1, U, CX, fixed', WRK, nfd, nelx, nely, nelz, nl, cgtol, cgmax)
        (Cantilever/omp/mgcg1.c:51-105; WRK is accepted and checked against the
         reference's workspace estimate, then ignored)

U is overwritten in place with the solution in the MATLAB frame (y negated),
cgiters is the 0-based exit index, and a failure returns cgiters < 0 instead
of terminating the process.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, f64, ptr

_HANDLES = {}


class Solver:
    """One libgputo handle for an nelx x nely x nelz mesh with nl levels."""

    def __init__(self, nelx, nely, nelz, nl, preset=_lib.PRESET_REF_C, precision=_lib.PREC_F64, learned=False,
                 print_level=0, l0_kernel=0, l1_kernel=0, nswp=0, coarse_lag=0, periodic=False, mixed_npre0=0,
                 mixed_npost0=0, mixed_npre1=0, mixed_npost1=0, mixed_npre2=0, mixed_npost2=0, homog_sequential=0,
                 graph_vcycles=0, mixed_npre3=0, mixed_npost3=0):
        self.lib = _lib.load()
        self.nelx, self.nely, self.nelz, self.nl = int(nelx), int(nely), int(nelz), int(nl)
        self.nel = self.nelx * self.nely * self.nelz
        self.ndof = 3 * (self.nelx + 1) * (self.nely + 1) * (self.nelz + 1)
        self.params = _lib.GputoParams()
        self.lib.gputo_default_params(ctypes.byref(self.params))
        self.params.preset = preset
        self.params.precision = precision
        # False/0: no learned stage; True/1: applied in every V-cycle of a
        # trained network; 2: with the solve-level fallback policy
        self.params.learned = 2 if learned == 2 else int(bool(learned))
        self.params.print_level = print_level
        self.params.l0_kernel = l0_kernel
        self.params.l1_kernel = l1_kernel
        self.params.nswp = nswp
        self.params.coarse_lag = coarse_lag
        self.params.periodic = int(bool(periodic))
        self.params.mixed_npre0 = mixed_npre0
        self.params.mixed_npost0 = mixed_npost0
        self.params.mixed_npre1 = mixed_npre1
        self.params.mixed_npost1 = mixed_npost1
        self.params.mixed_npre2 = mixed_npre2
        self.params.mixed_npost2 = mixed_npost2
        self.params.homog_sequential = int(bool(homog_sequential))
        self.params.graph_vcycles = int(bool(graph_vcycles))
        self.params.mixed_npre3 = mixed_npre3
        self.params.mixed_npost3 = mixed_npost3
        h = ctypes.c_void_p()
        check(self.lib.gputo_create(self.nelx, self.nely, self.nelz, self.nl, ctypes.byref(self.params),
                                    ctypes.byref(h)), "gputo_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.gputo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, **kw):
        for k, v in kw.items():
            setattr(self.params, k, v)
        check(self.lib.gputo_set_params(self.h, ctypes.byref(self.params)), "gputo_set_params")

    # ---- the mgcg1 solve ------------------------------------------------
    def solve(self, Ke0, F, CX, fixed, tol, maxit, U=None):
        """Returns (U_matlab_frame, relres, nit, rc)."""
        Ke0, F, CX = f64(Ke0), f64(F), f64(CX)
        fixed = f64(fixed).ravel()
        if U is None:
            U = np.zeros(self.ndof)
        relres = ctypes.c_double(0.0)
        nit = ctypes.c_int(0)
        rc = self.lib.gputo_mgcg1(self.h, ptr(Ke0), ptr(F), ptr(U), ptr(CX), ptr(fixed), len(fixed), float(tol),
                                  int(maxit), ctypes.byref(relres), ctypes.byref(nit))
        return U, relres.value, nit.value, rc

    # ---- periodic unit cell ----------------------------------------------
    def homogenize(self, Ke0, CX, tol, maxit, fields=False):
        """Homogenised 6x6 elasticity tensor of a periodic cell (handle created
        with periodic=True).  Voigt order xx, yy, zz, xy, yz, zx with
        engineering shears, x along i, y along j, z along k.  Returns
        (CH, relres[6], nit[6], chi) where chi (6 x ndof, full node layout,
        equal on image nodes) is None unless fields=True."""
        Ke0, CX = f64(Ke0), f64(CX)
        CH = np.zeros(36)
        relres = np.zeros(6)
        nit = np.zeros(6, dtype=np.int32)
        chi = np.zeros((6, self.ndof)) if fields else None
        rc = self.lib.gputo_homogenize(self.h, ptr(Ke0), ptr(CX), float(tol), int(maxit), ptr(CH), ptr(relres),
                                       ptr(nit), ptr(chi))
        check(rc, "gputo_homogenize")
        return CH.reshape(6, 6), relres, nit, chi

    # ---- kernel-level entry points --------------------------------------
    def matvec(self, Ke0, CX, u, fixed):
        y = np.zeros(self.ndof)
        fixed = f64(fixed).ravel()
        check(self.lib.gputo_matvec_f64(self.h, ptr(f64(Ke0)), ptr(f64(CX)), ptr(f64(u)), ptr(fixed), len(fixed),
                                        ptr(y)), "gputo_matvec_f64")
        return y

    def matvec_mixed(self, kernel, Ke0, CX, u, fixed):
        y = np.zeros(self.ndof)
        fixed = f64(fixed).ravel()
        check(self.lib.gputo_matvec_mixed(self.h, kernel, ptr(f64(Ke0)), ptr(f64(CX)), ptr(f64(u)), ptr(fixed),
                                          len(fixed), ptr(y)), "gputo_matvec_mixed")
        return y

    def defect(self, Ke0, CX, F, u, fixed):
        r = np.zeros(self.ndof)
        fixed = f64(fixed).ravel()
        check(self.lib.gputo_defect_f64(self.h, ptr(f64(Ke0)), ptr(f64(CX)), ptr(f64(F)), ptr(f64(u)), ptr(fixed),
                                        len(fixed), ptr(r)), "gputo_defect_f64")
        return r

    def diag(self, Ke0, CX):
        ad = np.zeros(self.ndof)
        check(self.lib.gputo_diag_f64(self.h, ptr(f64(Ke0)), ptr(f64(CX)), ptr(ad)), "gputo_diag_f64")
        return ad

    def setup(self, Ke0, CX, fixed):
        fixed = f64(fixed).ravel()
        check(self.lib.gputo_setup(self.h, ptr(f64(Ke0)), ptr(f64(CX)), ptr(fixed), len(fixed)), "gputo_setup")

    def vcycle(self, Ke0, CX, fixed, r):
        z = np.zeros(self.ndof)
        fixed = f64(fixed).ravel()
        check(self.lib.gputo_vcycle_f64(self.h, ptr(f64(Ke0)), ptr(f64(CX)), ptr(fixed), len(fixed), ptr(f64(r)),
                                        ptr(z)), "gputo_vcycle_f64")
        return z

    def level_info(self, level):
        a, b, c, d = (ctypes.c_int() for _ in range(4))
        check(self.lib.gputo_level_info(self.h, level, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                        ctypes.byref(d)), "gputo_level_info")
        return a.value, b.value, c.value, d.value

    def level_matrices(self, level):
        nx, ny, nz, nfd = self.level_info(level)
        nel = nx * ny * nz
        ndof = 3 * (nx + 1) * (ny + 1) * (nz + 1)
        A = np.zeros(nel * 576) if level > 0 else None
        ad = np.zeros(ndof)
        fd = np.zeros(max(nfd, 1))
        check(self.lib.gputo_level_matrices(self.h, level, ptr(A), ptr(ad), ptr(fd)), "gputo_level_matrices")
        return (None if A is None else A.reshape(nel, 24, 24)), ad, fd[:nfd]

    def level_cx(self, level):
        nx, ny, nz, _ = self.level_info(level)
        cx = np.zeros(nx * ny * nz)
        check(self.lib.gputo_level_cx(self.h, level, ptr(cx)), "gputo_level_cx")
        return cx

    def level_apply(self, level, op, mode, u, f=None, fp32=False):
        """One level-`level` operator application (gputo_level_apply): mode 0
        y = A u, 1 y = f - A u, 2 one damped-Jacobi sweep, 3 one sweep from
        prerelax*f/AD, 4 f - A (prerelax*f/AD); op 1 stored (level 0: the
        level-0 kernel), 2 matrix-free (level 1)."""
        y = np.zeros(len(u))
        check(self.lib.gputo_level_apply(self.h, level, op, mode, int(bool(fp32)), ptr(f64(u)),
                                         None if f is None else ptr(f64(f)), ptr(y)), "gputo_level_apply")
        return y

    def bench_level(self, level, op, fp32=False, reps=50):
        avg = ctypes.c_double()
        check(self.lib.gputo_bench_level(self.h, level, op, int(bool(fp32)), reps, ctypes.byref(avg)),
              "gputo_bench_level")
        return avg.value

    # ---- learned level-2 correction -------------------------------------
    def learn_info(self):
        n = ctypes.c_longlong()
        d = [ctypes.c_int() for _ in range(4)]
        check(self.lib.gputo_learn_info(self.h, ctypes.byref(n), *[ctypes.byref(v) for v in d]), "gputo_learn_info")
        return {"nparam": n.value, "grid": (d[0].value, d[1].value, d[2].value), "nsamples": d[3].value}

    def learn_params(self, values=None):
        info = self.learn_info()
        if values is None:
            buf = np.zeros(info["nparam"])
            check(self.lib.gputo_learn_params(self.h, 0, ptr(buf)), "gputo_learn_params")
            return buf
        buf = f64(values).ravel()
        assert buf.size == info["nparam"]
        check(self.lib.gputo_learn_params(self.h, 1, ptr(buf)), "gputo_learn_params")

    def learn_logging(self, on):
        check(self.lib.gputo_learn_logging(self.h, int(bool(on))), "gputo_learn_logging")

    def learn_reset(self, seed=1234):
        check(self.lib.gputo_learn_reset(self.h, seed), "gputo_learn_reset")

    def learn_forward(self, x):
        n0, n1, n2 = self.learn_info()["grid"]
        y = np.zeros(3 * n0 * n1 * n2)
        check(self.lib.gputo_learn_forward(self.h, ptr(f64(x).ravel()), ptr(y)), "gputo_learn_forward")
        return y.reshape(3, n0, n1, n2)

    def learn_grad(self, x, gy):
        g = np.zeros(self.learn_info()["nparam"])
        check(self.lib.gputo_learn_grad(self.h, ptr(f64(x).ravel()), ptr(f64(gy).ravel()), ptr(g)),
              "gputo_learn_grad")
        return g

    def learn_train(self, steps, lr=1e-3):
        l0, l1 = ctypes.c_double(), ctypes.c_double()
        check(self.lib.gputo_learn_train(self.h, steps, lr, ctypes.byref(l0), ctypes.byref(l1)), "gputo_learn_train")
        return l0.value, l1.value

    def fixed_list(self, level):
        n = ctypes.c_int()
        check(self.lib.gputo_fixed_list(self.h, level, None, ctypes.byref(n)), "gputo_fixed_list")
        out = np.zeros(max(n.value, 1))
        check(self.lib.gputo_fixed_list(self.h, level, ptr(out), ctypes.byref(n)), "gputo_fixed_list")
        return out[:n.value]

    def restrict(self, level, r_fine):
        nx, ny, nz, _ = self.level_info(level + 1)
        out = np.zeros(3 * (nx + 1) * (ny + 1) * (nz + 1))
        check(self.lib.gputo_restrict_f64(self.h, level, ptr(f64(r_fine)), ptr(out)), "gputo_restrict_f64")
        return out

    def interp_add(self, level, u_coarse, u_fine):
        out = np.array(u_fine, dtype=np.float64, copy=True)
        check(self.lib.gputo_interp_f64(self.h, level, ptr(f64(u_coarse)), ptr(out)), "gputo_interp_f64")
        return out

    def coarse_solve(self, b):
        x = np.zeros_like(f64(b))
        check(self.lib.gputo_coarse_solve_f64(self.h, ptr(f64(b)), ptr(x)), "gputo_coarse_solve_f64")
        return x

    def edofmat(self):
        c = np.zeros((self.nel, 24), dtype=np.int32)
        check(self.lib.gputo_edofmat(self.h, ptr(c)), "gputo_edofmat")
        return c

    def filter(self, x, rmin):
        out = np.zeros(self.nel)
        check(self.lib.gputo_filter_f64(self.h, ptr(f64(x)), float(rmin), ptr(out)), "gputo_filter_f64")
        return out

    def sens(self, Ke0, U_cframe):
        ce = np.zeros(self.nel)
        check(self.lib.gputo_sens_f64(self.h, ptr(f64(Ke0)), ptr(f64(U_cframe)), ptr(ce)), "gputo_sens_f64")
        return ce

    def device_bytes(self):
        return self.lib.gputo_device_bytes(self.h)


def _handle(nelx, nely, nelz, nl, preset=_lib.PRESET_REF_C, print_level=0, nswp=0):
    # one persistent handle per mesh, as a MEX would keep it between calls
    key = (int(nelx), int(nely), int(nelz), int(nl), preset)
    s = _HANDLES.get(key)
    if s is None:
        s = Solver(nelx, nely, nelz, nl, preset=preset, print_level=print_level, nswp=nswp)
        _HANDLES[key] = s
    elif s.params.nswp != nswp or s.params.print_level != print_level:
        s.set_params(nswp=nswp, print_level=print_level)
    return s


def release_handles():
    for s in _HANDLES.values():
        s.close()
    _HANDLES.clear()


def _scalar(v):
    return float(np.asarray(v).ravel()[0])


def mgcg1(*args):
    """[cgres, cgiters] = mgcg1(...) with the reference's argument lists.

    12/13 arguments: the CUDA MEX form (optional printLev);
    13 arguments with a vector 6th argument: the OpenMP MEX form with WRK.
    U (3rd argument) is overwritten in place.
    """
    if len(args) == 13 and np.asarray(args[5]).size > 1:  # omp form with WRK
        Ke0, F, U, CX, fixed, WRK, nfd, nelx, nely, nelz, nl, tol, maxit = args
        print_level = 0
        lenW = _lib.load().gputo_est_workspace(int(_scalar(nfd)), int(_scalar(nl)), int(_scalar(nelx)),
                                               int(_scalar(nely)), int(_scalar(nelz)))
        if np.asarray(WRK).size < lenW:  # mgcg1.c:89-97
            return 0.0, -1.0
    elif len(args) in (12, 13):
        Ke0, F, U, CX, fixed, nfd, nelx, nely, nelz, nl, tol, maxit = args[:12]
        print_level = int(_scalar(args[12])) if len(args) == 13 else 0
    else:
        raise TypeError("mgcg1 takes 12 or 13 arguments")
    if not (isinstance(U, np.ndarray) and U.dtype == np.float64 and U.flags["C_CONTIGUOUS"]):
        raise TypeError("U must be a contiguous float64 array (it is overwritten in place)")
    nfd = int(_scalar(nfd))
    fixed = f64(fixed).ravel()[:nfd]
    s = _handle(_scalar(nelx), _scalar(nely), _scalar(nelz), _scalar(nl), print_level=print_level)
    _, relres, nit, rc = s.solve(Ke0, F, CX, fixed, _scalar(tol), int(_scalar(maxit)), U=U.reshape(-1))
    if rc != 0:
        return 0.0, float(rc)
    return relres, float(nit)


def mgcg9(Ke0, F, U, CX, fixed, nfd, nelx, nely, nelz, nl, tol, maxit, nswp, printLev, nu, hx, hy, hz):
    """[cgres, cgiters] = mgcg9(...) with the argument list of
    Homogenization/QNC/mgcg9.cu:124-182, solved by the mgcg9 engine
    (GPUTO_PRESET_MGCG9 with its nswp; nu and hx..hz are parsed and unused
    there too).  U (3rd argument) is overwritten in place."""
    if not (isinstance(U, np.ndarray) and U.dtype == np.float64 and U.flags["C_CONTIGUOUS"]):
        raise TypeError("U must be a contiguous float64 array (it is overwritten in place)")
    nfd = int(_scalar(nfd))
    fixed = f64(fixed).ravel()[:nfd]
    s = _handle(_scalar(nelx), _scalar(nely), _scalar(nelz), _scalar(nl), preset=_lib.PRESET_MGCG9,
                print_level=int(_scalar(printLev)), nswp=int(_scalar(nswp)))
    _, relres, nit, rc = s.solve(Ke0, F, CX, fixed, _scalar(tol), int(_scalar(maxit)), U=U.reshape(-1))
    if rc != 0:
        return 0.0, float(rc)
    return relres, float(nit)
