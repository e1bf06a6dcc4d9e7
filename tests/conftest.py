import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "gpu-to_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def gputo():
    import gputo
    gputo.load()
    return gputo


def cantilever_case(nelx, nely, nelz, seed=0, kind="perturbed"):
    """(Ke0, F, CX, fixed) of a cantilever with a seeded density field."""
    from gputo.problems import cantilever
    Ke, F, fixed = cantilever(nelx, nely, nelz)
    rng = np.random.default_rng(seed)
    nel = nelx * nely * nelz
    if kind == "uniform":
        x = np.full(nel, 0.12)
    elif kind == "binary":  # disconnected stiff islands in near-void: CG stagnates
        x = np.where(rng.uniform(size=nel) < 0.3, 1.0, 1e-3)
    elif kind == "smooth":  # smoothed random field in [0.05, 1]: E contrast 8e3
        import scipy.ndimage as ndi
        r = ndi.uniform_filter(rng.uniform(size=(nelx, nelz, nely)), size=3, mode="reflect")
        x = (0.05 + 0.95 * (r - r.min()) / (r.max() - r.min())).ravel()
    else:
        x = np.clip(0.12 + rng.uniform(-0.1, 0.1, nel), 0.01, 1.0)
    CX = 1e-9 + x ** 3 * (1 - 1e-9)
    return Ke, F, CX, fixed


def dense_K(Ke, CX, nelx, nely, nelz):
    """Brute-force global stiffness from the C-order element table."""
    from gputo.problems import node_id
    nel = nelx * nely * nelz
    ndof = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)
    ie = [0, 1, 1, 0, 0, 1, 1, 0]
    je = [0, 0, 1, 1, 0, 0, 1, 1]
    ke = [0, 0, 0, 0, 1, 1, 1, 1]
    K = np.zeros((ndof, ndof))
    for i in range(nelx):
        for k in range(nelz):
            for j in range(nely):
                e = nely * nelz * i + nely * k + j
                nodes = [node_id(nelx, nely, nelz, i + ie[a], j + je[a], k + ke[a]) for a in range(8)]
                dofs = np.array([3 * n + d for n in nodes for d in range(3)])
                K[np.ix_(dofs, dofs)] += CX[e] * Ke
    return K
