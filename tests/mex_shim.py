"""ctypes access to the MEX shims built by tests/mex/Makefile
(gpu-to_amd/mex/mgcg1_gputo.c and mgcg9_gputo.c over the MEX API
emulation of tests/mex)."""
import ctypes
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mex")
_P = ctypes.c_void_p


class MexFunction:
    def __init__(self, name):
        self.lib = ctypes.CDLL(os.path.join(HERE, "lib%s_mex.so" % name))
        L = self.lib
        L.mx_wrap.restype = _P
        L.mx_wrap.argtypes = [_P, ctypes.c_size_t]
        L.mx_free.argtypes = [_P]
        L.mxGetScalar.restype = ctypes.c_double
        L.mxGetScalar.argtypes = [_P]
        L.mex_call.restype = ctypes.c_int
        L.mex_call.argtypes = [ctypes.c_int, ctypes.POINTER(_P), ctypes.c_int, ctypes.POINTER(_P)]
        L.mex_last_error.restype = ctypes.c_char_p
        L.mex_unload.restype = ctypes.c_int

    def __call__(self, *args):
        """Calls the MEX function with MATLAB-like arguments (arrays are
        passed by reference, so an in-place U is updated); returns
        (cgres, cgiters), or raises RuntimeError with mexErrMsgTxt's text."""
        keep, arrs = [], []
        for a in args:
            a = np.asarray(a, dtype=np.float64) if not isinstance(a, np.ndarray) else a
            if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
                a = np.ascontiguousarray(a, dtype=np.float64)
            if a.ndim == 0:
                a = a.reshape(1)
            keep.append(a)
            arrs.append(self.lib.mx_wrap(a.ctypes.data_as(_P), a.size))
        prhs = (_P * len(arrs))(*arrs)
        plhs = (_P * 2)()
        try:
            rc = self.lib.mex_call(2, plhs, len(arrs), prhs)
            if rc != 0:
                raise RuntimeError(self.lib.mex_last_error().decode())
            out = (self.lib.mxGetScalar(plhs[0]), self.lib.mxGetScalar(plhs[1]))
            for p in plhs:
                self.lib.mx_free(p)
            return out
        finally:
            for p in arrs:
                self.lib.mx_free(p)

    def unload(self):
        return self.lib.mex_unload()
