"""ctypes binding of libgputo.so (the C ABI declared in include/gputo.h).

The library is built in-tree (`make -C gpu-to_amd`); loading fails loudly
when it is missing -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libgputo.so")

GPUTO_OK = 0
GPUTO_ERR_WORKSPACE = -1
GPUTO_ERR_NOT_SPD = -2
GPUTO_ERR_NONFINITE = -3
GPUTO_ERR_DEVICE = -4
GPUTO_ERR_ARGUMENT = -5

PRESET_REF_C = 0
PRESET_MATLAB_CFG0 = 1
PRESET_MGCG9 = 2

PREC_F64 = 0
PREC_MIXED_F32 = 1
PREC_MIXED_BF16 = 2

# level-0 kernels (gputo_params.l0_kernel, gputo_matvec_mixed, gputo_bench_operator)
L0_F64_GATHER = 0
L0_GATHER = 1
L0_MFMA_F32 = 2
L0_MFMA_BF16 = 3
L0_MFMA_BF16X3 = 4
L0_MFMA_BF16X1 = 5

# level-1 operator (gputo_params.l1_kernel)
L1_AUTO = 0
L1_STORED = 1
L1_MATRIX_FREE = 2


class GputoParams(ctypes.Structure):
    _fields_ = [
        ("preset", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("learned", ctypes.c_int),
        ("print_level", ctypes.c_int),
        ("l0_kernel", ctypes.c_int),
        ("nswp", ctypes.c_int),
        ("l1_kernel", ctypes.c_int),
        ("coarse_lag", ctypes.c_int),
        ("periodic", ctypes.c_int),
        ("mixed_npre0", ctypes.c_int),
        ("mixed_npost0", ctypes.c_int),
        ("mixed_npre1", ctypes.c_int),
        ("mixed_npost1", ctypes.c_int),
        ("mixed_npre2", ctypes.c_int),
        ("mixed_npost2", ctypes.c_int),
        ("homog_sequential", ctypes.c_int),
        ("graph_vcycles", ctypes.c_int),
        ("mixed_npre3", ctypes.c_int),
        ("mixed_npost3", ctypes.c_int),
    ]


class ToConfig(ctypes.Structure):
    _fields_ = [
        ("volfrac", ctypes.c_double),
        ("penal", ctypes.c_double),
        ("rmin", ctypes.c_double),
        ("eta", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("move", ctypes.c_double),
        ("Emin", ctypes.c_double),
        ("E0", ctypes.c_double),
        ("ft", ctypes.c_int),
        ("cg_tol", ctypes.c_double),
        ("cg_maxit", ctypes.c_int),
    ]


class ToStats(ctypes.Structure):
    _fields_ = [
        ("compliance", ctypes.c_double),
        ("volume", ctypes.c_double),
        ("change", ctypes.c_double),
        ("eta", ctypes.c_double),
        ("lmid", ctypes.c_double),
        ("relres", ctypes.c_double),
        ("cg_iters", ctypes.c_int),
        ("status", ctypes.c_int),
        ("seconds_total", ctypes.c_double),
        ("seconds_solve", ctypes.c_double),
        ("inner_iters", ctypes.c_int),
        ("outer_iters", ctypes.c_int),
        ("learned_accepted", ctypes.c_int),
        ("learned_rejected", ctypes.c_int),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_pI = ctypes.POINTER(ctypes.c_int)
_pD = ctypes.POINTER(ctypes.c_double)

# symbol -> (restype, argtypes); every entry point of include/gputo.h
SIGNATURES = {
    "gputo_default_params": (None, [ctypes.POINTER(GputoParams)]),
    "gputo_create": (_I, [_I, _I, _I, _I, ctypes.POINTER(GputoParams), ctypes.POINTER(_P)]),
    "gputo_destroy": (_I, [_P]),
    "gputo_set_params": (_I, [_P, ctypes.POINTER(GputoParams)]),
    "gputo_mgcg1": (_I, [_P, _P, _P, _P, _P, _P, _I, _D, _I, _pD, _pI]),
    "gputo_solve_d": (_I, [_P, _P, _P, _P, _P, _P, _I, _D, _I, _pD, _pI]),
    "gputo_est_workspace": (ctypes.c_longlong, [_I, _I, _I, _I, _I]),
    "gputo_device_bytes": (ctypes.c_longlong, [_P]),
    "gputo_matvec_f64": (_I, [_P, _P, _P, _P, _P, _I, _P]),
    "gputo_matvec_mixed": (_I, [_P, _I, _P, _P, _P, _P, _I, _P]),
    "gputo_matvec_f64_noalloc": (_I, [_I, _I, _I, _P, _P, _P, _P, _I, _P, _P]),
    "gputo_selftest_mfma": (_I, [_pD]),
    "gputo_defect_f64": (_I, [_P, _P, _P, _P, _P, _P, _I, _P]),
    "gputo_diag_f64": (_I, [_P, _P, _P, _P]),
    "gputo_vcycle_f64": (_I, [_P, _P, _P, _P, _I, _P, _P]),
    "gputo_setup": (_I, [_P, _P, _P, _P, _I]),
    "gputo_level_info": (_I, [_P, _I, _pI, _pI, _pI, _pI]),
    "gputo_level_matrices": (_I, [_P, _I, _P, _P, _P]),
    "gputo_level_cx": (_I, [_P, _I, _P]),
    "gputo_restrict_f64": (_I, [_P, _I, _P, _P]),
    "gputo_interp_f64": (_I, [_P, _I, _P, _P]),
    "gputo_coarse_solve_f64": (_I, [_P, _P, _P]),
    "gputo_bench_operator": (_I, [_P, _I, _I, _pD]),
    "gputo_level_apply": (_I, [_P, _I, _I, _I, _I, _P, _P, _P]),
    "gputo_bench_level": (_I, [_P, _I, _I, _I, _I, _pD]),
    "gputo_learn_info": (_I, [_P, ctypes.POINTER(ctypes.c_longlong), _pI, _pI, _pI, _pI]),
    "gputo_learn_params": (_I, [_P, _I, _P]),
    "gputo_learn_reset": (_I, [_P, ctypes.c_uint]),
    "gputo_learn_logging": (_I, [_P, _I]),
    "gputo_learn_forward": (_I, [_P, _P, _P]),
    "gputo_learn_grad": (_I, [_P, _P, _P, _P]),
    "gputo_learn_train": (_I, [_P, _I, _D, _pD, _pD]),
    "gputo_edofmat": (_I, [_P, _P]),
    "gputo_fixed_list": (_I, [_P, _I, _P, _pI]),
    "gputo_to_begin": (_I, [_P, _P, _P, _P, _P, _I, ctypes.POINTER(ToConfig)]),
    "gputo_to_iterate": (_I, [_P, ctypes.POINTER(ToStats)]),
    "gputo_to_get": (_I, [_P, _P, _P, _P]),
    "gputo_to_end": (_I, [_P]),
    "gputo_filter_f64": (_I, [_P, _P, _D, _P]),
    "gputo_sens_f64": (_I, [_P, _P, _P, _P]),
    "gputo_homogenize": (_I, [_P, _P, _P, _D, _I, _P, _P, _P, _P]),
    "gputo_version": (ctypes.c_char_p, []),
    "gputo_device_synchronize": (_I, []),
}

_LIB = None


def load():
    """Load libgputo.so and set every signature; raises when it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libgputo.so is not built (%s): run `make -C gpu-to_amd`" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def ptr(a):
    """Data pointer of a C-contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(ctypes.c_void_p)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class GputoError(RuntimeError):
    def __init__(self, code, what=""):
        names = {-1: "workspace", -2: "coarsest matrix not SPD", -3: "non-finite value", -4: "device error",
                 -5: "invalid argument"}
        super().__init__("%s failed: %d (%s)" % (what, code, names.get(code, "?")))
        self.code = code


def check(rc, what):
    if rc != 0:
        raise GputoError(rc, what)
    return rc
