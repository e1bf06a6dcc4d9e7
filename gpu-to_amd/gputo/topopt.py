"""Device-resident topology-optimisation loop (top3D125mod.m:105-165)."""
import ctypes

import numpy as np

from . import _lib
from . import log as _log
from ._lib import check, f64, ptr


class TopOpt:
    """Runs the design loop of the Cantilever drivers on one libgputo handle.

    cfg keys (defaults are the 192x96x96 cantilever benchmark's): volfrac, penal,
    rmin, eta, beta, move, Emin, E0, ft (1: density filter, 3: filter +
    projection), cg_tol, cg_maxit.
    """

    DEFAULTS = dict(volfrac=0.12, penal=3.0, rmin=np.sqrt(3.0), eta=0.4, beta=2.0, move=0.2, Emin=1e-9, E0=1.0, ft=1,
                    cg_tol=1e-8, cg_maxit=200)

    def __init__(self, solver, Ke0, x0, F, fixed, log=None, maxit=100, **cfg):
        """log: None, or a file path (or open text file) that receives the
        reference driver's log: the header now, one line per iterate()."""
        self.s = solver
        self.lib = solver.lib
        c = dict(self.DEFAULTS)
        c.update(cfg)
        self.c = c
        self.cfg = _lib.ToConfig(**{k: c[k] for k, _ in _lib.ToConfig._fields_})
        fixed = f64(fixed).ravel()
        check(self.lib.gputo_to_begin(solver.h, ptr(f64(Ke0)), ptr(f64(x0)), ptr(f64(F)), ptr(fixed), len(fixed),
                                      ctypes.byref(self.cfg)), "gputo_to_begin")
        self.history = []
        self._log = open(log, "w") if isinstance(log, str) else log
        if self._log is not None:
            self._log.write(_log.header(solver.nl, solver.nelx, solver.nely, solver.nelz, c["volfrac"], c["penal"],
                                        c["rmin"], c["ft"], c["eta"], c["move"], maxit, c["beta"], self._mem()))
            self._log.flush()

    def _mem(self):
        return float(self.lib.gputo_device_bytes(self.s.h))

    def iterate(self):
        st = _lib.ToStats()
        rc = self.lib.gputo_to_iterate(self.s.h, ctypes.byref(st))
        d = st.as_dict()
        d["rc"] = rc
        self.history.append(d)
        if rc != 0:
            raise _lib.GputoError(rc, "gputo_to_iterate")
        if self._log is not None:
            c = self.c
            self._log.write(_log.line(len(self.history), d["compliance"], d["seconds_total"], d["volume"], d["change"],
                                      c["penal"], c["beta"], d["eta"], d["lmid"], d["relres"], d["cg_iters"],
                                      self._mem()))
            self._log.flush()
        return d

    def state(self):
        x = np.zeros(self.s.nel)
        xp = np.zeros(self.s.nel)
        U = np.zeros(self.s.ndof)
        check(self.lib.gputo_to_get(self.s.h, ptr(x), ptr(xp), ptr(U)), "gputo_to_get")
        return x, xp, U

    def close(self):
        self.lib.gputo_to_end(self.s.h)
        if self._log is not None and hasattr(self._log, "close"):
            self._log.close()
            self._log = None
