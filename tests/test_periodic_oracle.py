"""The periodic-cell oracle against closed forms (CPU only).

The reference has no periodic solve, so nothing of the reference pins
these; the closed forms below do: a homogeneous cell (CH = E D), layered
cells in each direction (Backus averages, exact in the FE space),
symmetry and definiteness, translation invariance and axis swaps.
"""
import numpy as np
import pytest

import periodic_oracle as PO
from gputo.problems import ke0_h8

KE = ke0_h8()

# Voigt index permutation of swapping the x and y axes:
# xx <-> yy, zz, xy, yz <-> zx
SWAP_XY = np.array([1, 0, 2, 3, 5, 4])
# swapping x and z: xx <-> zz, yy, xy <-> yz, zx
SWAP_XZ = np.array([2, 1, 0, 4, 3, 5])


def layered(nelx, nely, nelz, axis, E):
    """CX of layers normal to `axis` (0 x/i, 1 y/j, 2 z/k): element layer
    number l gets E[l]."""
    i, k, j = np.meshgrid(np.arange(nelx), np.arange(nelz), np.arange(nely), indexing="ij")
    idx = (i, j, k)[axis].ravel()
    return np.asarray(E, dtype=np.float64)[idx]


def test_unit_strain_fields_give_the_strains():
    # B u0_q at any point of the element is the unit Voigt vector e_q
    u0 = PO.unit_strain_fields()
    g = 0.3
    s, e, t = g, -g, 0.7 * g
    B = np.zeros((6, 24))
    for a in range(8):
        sx, sy, sz = 2 * PO.IE[a] - 1, 2 * PO.JE[a] - 1, 2 * PO.KE[a] - 1
        # physical derivatives of N_a = (1+sx s)(1+sy e)(1+sz t)/8 on the
        # unit cube (x = (s+1)/2) at the parametric point (s, e, t)
        dx = 0.25 * sx * (1 + sy * e) * (1 + sz * t)
        dy = 0.25 * sy * (1 + sx * s) * (1 + sz * t)
        dz = 0.25 * sz * (1 + sx * s) * (1 + sy * e)
        B[0, 3 * a], B[1, 3 * a + 1], B[2, 3 * a + 2] = dx, dy, dz
        B[3, 3 * a], B[3, 3 * a + 1] = dy, dx
        B[4, 3 * a + 1], B[4, 3 * a + 2] = dz, dy
        B[5, 3 * a], B[5, 3 * a + 2] = dz, dx
    np.testing.assert_allclose(B @ u0.T, np.eye(6), atol=1e-14)


def test_periodic_edofs_wrap():
    ed = PO.periodic_edofs(3, 2, 2)
    assert ed.shape == (12, 24)
    assert ed.min() == 0 and ed.max() == 3 * 12 - 1
    # every periodic node belongs to exactly 8 elements
    counts = np.bincount(ed[:, 0::3].ravel() // 3, minlength=12)
    assert (counts == 8).all()


@pytest.mark.parametrize("shape", [(4, 4, 4), (6, 4, 2)])
def test_homogeneous_cell_is_D(shape):
    nelx, nely, nelz = shape
    CX = np.full(nelx * nely * nelz, 2.5)
    CH, chi = PO.homogenize(KE, CX, nelx, nely, nelz)
    np.testing.assert_allclose(CH, 2.5 * PO.isotropic_D(1.0, 0.3), rtol=1e-12, atol=1e-13)
    assert np.abs(chi).max() < 1e-10


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_laminate_is_backus(axis):
    E = [1.0, 1.0, 0.05, 0.3, 0.3, 0.3]  # layers of the 6-element direction
    shape = [4, 4, 4]
    shape[axis] = 6
    CX = layered(*shape, axis, E)
    CH, _ = PO.homogenize(KE, CX, *shape)
    Cx = PO.backus_x(E, [0.3] * 6, [1.0 / 6] * 6)
    perm = {0: np.arange(6), 1: SWAP_XY, 2: SWAP_XZ}[axis]
    np.testing.assert_allclose(CH, Cx[np.ix_(perm, perm)], rtol=1e-10, atol=1e-12)


def random_cx(shape, seed):
    rng = np.random.default_rng(seed)
    x = np.where(rng.uniform(size=int(np.prod(shape))) < 0.5, 1.0, 0.2)
    return x ** 3


def test_symmetric_definite_and_bounded():
    shape = (4, 4, 4)
    CX = random_cx(shape, 1)
    CH, _ = PO.homogenize(KE, CX, *shape)
    np.testing.assert_allclose(CH, CH.T, rtol=0, atol=1e-13 * np.abs(CH).max())
    assert np.linalg.eigvalsh(CH).min() > 0
    # Voigt bound: CH <= mean(CX) D in the Loewner order
    assert np.linalg.eigvalsh(CX.mean() * PO.isotropic_D(1.0, 0.3) - CH).min() > -1e-12


def test_translation_invariance():
    shape = (4, 4, 6)
    CX = random_cx(shape, 2)
    CH, _ = PO.homogenize(KE, CX, *shape)
    nelx, nely, nelz = shape
    rolled = np.roll(CX.reshape(nelx, nelz, nely), (1, 2, 3), axis=(0, 1, 2)).ravel()
    CH2, _ = PO.homogenize(KE, rolled, *shape)
    np.testing.assert_allclose(CH2, CH, rtol=1e-11, atol=1e-13)


def test_axis_swap_permutes():
    shape = (4, 4, 2)
    CX = random_cx(shape, 3)
    CH, _ = PO.homogenize(KE, CX, *shape)
    nelx, nely, nelz = shape
    c = CX.reshape(nelx, nelz, nely)           # axes (i, k, j)
    swapped = np.transpose(c, (2, 1, 0)).ravel()  # new i = old j
    CH2, _ = PO.homogenize(KE, swapped, nely, nelx, nelz)
    np.testing.assert_allclose(CH2, CH[np.ix_(SWAP_XY, SWAP_XY)], rtol=1e-11, atol=1e-13)


def test_a_planted_shear_mistake_is_caught(monkeypatch):
    # tensor shears (eps_xy = 1) instead of engineering shears: the
    # homogeneous closed form must fail
    good = PO.unit_strain_fields

    def wrong():
        u0 = good().copy()
        u0[3:] *= 2.0
        return u0

    monkeypatch.setattr(PO, "unit_strain_fields", wrong)
    CH, _ = PO.homogenize(KE, np.ones(64), 4, 4, 4)
    assert not np.allclose(CH, PO.isotropic_D(1.0, 0.3), rtol=1e-6)
