"""The design-loop log keeps the reference drivers' text format
(Cantilever/omp/top3D125mod.m:17-26, 155-156, 162-164): expected strings
written out by hand from those printf formats."""
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpu-to_amd"))

from gputo import log  # noqa: E402


def test_header():
    h = log.header(5, 192, 96, 96, 0.12, 3.0, 3 ** 0.5, 1, 0.4, 0.2, 100, 2.0, 2.5e9)
    assert h == ("Top3D125mod \n\nnl: 5, nelx: 192, nely:  96, nelz:  96, volfrac:0.12, penal: 3, rmin:1.7321, ft: 1, "
                 "eta:0.40, move:0.20, maxit:100, beta:2.00 \n"
                 "\nInitial Memory USED: 2.5000e+09 Bytes = 2.5000e+00 GB\n\n")


def test_iteration_line():
    s = log.line(1, 279.2348, 0.04213, 0.12, 0.0123, 3.0, 2.0, 0.4, 1.5e-3, 4.48e-9, 20, 2.5e9)
    assert s == ("It.:    1 C:2.79235e+02 Time:0.042 V:  0.120 ch.:1.23e-02 penal:   3.00 beta:    2.0 eta:   0.40 "
                 "lm:1.50e-03 cgres: 4.4800e-09,cgiters:  20 MemUsed:2.5000e+09 \n")
    c = log.console(12, 94.39857, 0.1199, 0.00481, 3.0, 2.0, 0.4, 2.25e-4)
    assert c == "It.:   12 C:9.43986e+01 V:  0.120 ch.:4.81e-03 penal:   3.00 beta:    2.0 eta:   0.40 lm:2.25e-04 \n"
    assert log.filename(5, 192, 96, 96) == "3D125mod_5_192_96_96caml.txt"
