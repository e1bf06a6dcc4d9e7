"""The design-loop log of the reference drivers (Cantilever/omp/
top3D125mod.m:17-26 header, :162-164 per-iteration line), written with the
same printf formats so existing log readers keep working.

The reference's `MemUsed` is MATLAB's `memory().MemUsedMATLAB`; here it is
the device memory the solver handle holds (gputo_device_bytes)."""


def header(nl, nelx, nely, nelz, volfrac, penal, rmin, ft, eta, move, maxit, beta, mem_bytes):
    """The text top3D125mod.m:18-25 writes when it creates the log."""
    s = ("Top3D125mod \n\nnl:%2d, nelx:%4d, nely:%4d, nelz:%4d, volfrac:%.2f, penal:%2d, rmin:%4.4f, ft:%2d, "
         "eta:%.2f, move:%.2f, maxit:%3d, beta:%1.2f \n" % (nl, nelx, nely, nelz, volfrac, int(penal), rmin, ft, eta,
                                                         move, maxit, beta))
    s += "\nInitial Memory USED: %6.4e Bytes = %6.4e GB\n\n" % (mem_bytes, mem_bytes / 1e9)
    return s


def line(loop, compliance, seconds, volume, change, penal, beta, eta, lmid, cgres, cgiters, mem_bytes):
    """One per-iteration line (top3D125mod.m:163-164)."""
    return ("It.:%5i C:%6.5e Time:%4.3f V:%7.3f ch.:%0.2e penal:%7.2f beta:%7.1f eta:%7.2f lm:%0.2e "
            "cgres: %3.4e,cgiters: %3d MemUsed:%6.4e \n" % (loop, compliance, seconds, volume, change, penal, beta, eta,
                                                           lmid, cgres, cgiters, mem_bytes))


def console(loop, compliance, volume, change, penal, beta, eta, lmid):
    """The console line of top3D125mod.m:155-156."""
    return ("It.:%5i C:%6.5e V:%7.3f ch.:%0.2e penal:%7.2f beta:%7.1f eta:%7.2f lm:%0.2e \n" %
            (loop, compliance, volume, change, penal, beta, eta, lmid))


def filename(nl, nelx, nely, nelz):
    """Log file name of top3D125mod.m:17."""
    return "3D125mod_%d_%d_%d_%dcaml.txt" % (nl, nelx, nely, nelz)
