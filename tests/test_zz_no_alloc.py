"""gputo_matvec_f64_noalloc, the level-0 operator entry point that allocates
no device memory (its node mask lives in a static device array), against
gputo_matvec_f64 on the same inputs: bit for bit, on two calls in a row,
the second on a smaller mesh with a different fixed set (the mask of the
first call must not leak into the second)."""
import ctypes

import numpy as np
import pytest

from conftest import cantilever_case

pytestmark = pytest.mark.gpu

_P = ctypes.c_void_p


class Dev:
    """Device copies of host arrays through the HIP runtime (the test
    allocates; the entry point under test does not)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 8))) == 0
        self.bufs.append(p)
        if a.nbytes:
            assert self.hip.hipMemcpy(p, a.ctypes.data_as(_P), ctypes.c_size_t(a.nbytes), 1) == 0
        return p

    def get(self, p, n, dtype=np.float64):
        out = np.zeros(n, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(_P), p, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def sync(self):
        assert self.hip.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def noalloc(lib):
    return lib.gputo_matvec_f64_noalloc  # signature set by gputo.load()


@pytest.mark.parametrize("first,second", [((24, 12, 12), (16, 8, 8))])
def test_noalloc_matvec_matches_handle_matvec(gputo, first, second):
    lib = gputo.load()
    fn = noalloc(lib)
    dev = Dev()
    try:
        for call, (nelx, nely, nelz) in enumerate((first, second)):
            print("no-alloc matvec call %d, mesh %dx%dx%d" % (call, nelx, nely, nelz), flush=True)
            Ke, F, CX, fixed = cantilever_case(nelx, nely, nelz, seed=70 + call)
            Ke = np.ascontiguousarray(Ke, dtype=np.float64)
            u = np.random.default_rng(80 + call).uniform(-1, 1, F.size)
            if call == 1:  # a different fixed set: the clamped face plus one z DOF per load node
                fixed = np.concatenate([fixed, np.nonzero(F)[0] + 1.0])
            fixed0 = (fixed - 1).astype(np.int32)
            dCX, du, dfx, dy = dev.put(CX), dev.put(u), dev.put(fixed0), dev.put(np.full(F.size, np.nan))
            rc = fn(nelx, nely, nelz, Ke.ctypes.data_as(_P), dCX, du, dfx, len(fixed0), dy, None)
            assert rc == 0
            dev.sync()
            y = dev.get(dy, F.size)
            s = gputo.Solver(nelx, nely, nelz, 2)
            y_ref = s.matvec(Ke, CX, u, fixed)
            s.close()
            assert np.array_equal(y, y_ref), np.abs(y - y_ref).max()
            assert np.all(y[fixed0] == 0.0)
            dev.free()
    finally:
        dev.free()


def test_noalloc_matvec_rejects_bad_input(gputo):
    """Host-side checks, before any device work: a mesh beyond the scratch
    size, a Ke0 without the unit-H8 entry structure, missing pointers."""
    lib = gputo.load()
    fn = noalloc(lib)
    Ke, F, CX, fixed = cantilever_case(8, 4, 4)
    Ke = np.ascontiguousarray(Ke, dtype=np.float64)
    dummy = ctypes.c_void_p(8)  # never dereferenced: every case fails before device work
    k = Ke.ctypes.data_as(_P)
    assert fn(256, 256, 256, k, dummy, dummy, dummy, 0, dummy, None) == gputo._lib.GPUTO_ERR_WORKSPACE
    Kb = Ke.copy()
    Kb[0, 1] += 1e-3
    Kb[1, 0] += 1e-3
    assert fn(8, 4, 4, Kb.ctypes.data_as(_P), dummy, dummy, dummy, 0, dummy, None) == gputo._lib.GPUTO_ERR_ARGUMENT
    assert fn(8, 4, 4, k, None, dummy, dummy, 0, dummy, None) == gputo._lib.GPUTO_ERR_ARGUMENT
    assert fn(8, 4, 4, k, dummy, dummy, None, 3, dummy, None) == gputo._lib.GPUTO_ERR_ARGUMENT
