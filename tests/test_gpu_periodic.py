"""Periodic unit cell on the GPU: homogenised tensors through
gputo_homogenize against the periodic oracle (direct sparse solve) and
closed forms, and the mixed-precision solve against the FP64 one.

Tolerances: the solves run to relres 1e-12 against the oracle's direct
solve, so CH agrees to 1e-8 relative (max-norm) and the fluctuation fields
to 1e-6; the mixed and MFMA solves of the 96^3 cell reach relres 1e-8 and
their CH agrees with the FP64 solve's to 1e-6.
"""
import time

import numpy as np
import pytest

import periodic_oracle as PO

pytestmark = pytest.mark.gpu


def two_phase(shape, seed, p_solid=0.4):
    """SIMP scales of a seeded two-phase field, x in {1, 0.2} (contrast 125)."""
    rng = np.random.default_rng(seed)
    x = np.where(rng.uniform(size=int(np.prod(shape))) < p_solid, 1.0, 0.2)
    return 1e-9 + x ** 3 * (1 - 1e-9)


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def homogenize(gputo, shape, nl, CX, tol=1e-12, maxit=500, fields=True, **kw):
    from gputo.problems import ke0_h8
    s = gputo.Solver(*shape, nl, periodic=True, **kw)
    try:
        return s.homogenize(ke0_h8(), CX, tol, maxit, fields=fields)
    finally:
        s.close()


@pytest.mark.parametrize("shape,nl,seed", [((8, 8, 8), 3, 0), ((12, 8, 16), 3, 1), ((16, 16, 16), 4, 2)])
def test_homogenize_matches_oracle(gputo, shape, nl, seed):
    from gputo.problems import ke0_h8
    print("case", shape, nl, seed, flush=True)
    CX = two_phase(shape, seed)
    CH, relres, nit, chi = homogenize(gputo, shape, nl, CX)
    CHo, chio = PO.homogenize(ke0_h8(), CX, *shape)
    print("nit", nit, "relres", relres.max(), "dCH", rel(CH, CHo))
    assert (nit > 0).all() and (relres < 1e-12).all()
    assert rel(CH, CHo) < 1e-8
    np.testing.assert_allclose(CH, CH.T, rtol=0, atol=1e-14 * np.abs(CH).max())
    for q in range(6):
        c = PO.full_to_periodic(chi[q], *shape)
        assert rel(c, chio[q]) < 1e-6, q
        # images hold their master's value
        full = chi[q].reshape(shape[0] + 1, shape[2] + 1, shape[1] + 1, 3)
        np.testing.assert_array_equal(full[-1], full[0])
        np.testing.assert_array_equal(full[:, -1], full[:, 0])
        np.testing.assert_array_equal(full[:, :, -1], full[:, :, 0])


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_laminate_is_backus(gputo, axis):
    shape = [8, 8, 8]
    shape[axis] = 16
    E = np.repeat([1.0, 0.02, 0.4, 1.0], 4)
    i, k, j = np.meshgrid(np.arange(shape[0]), np.arange(shape[2]), np.arange(shape[1]), indexing="ij")
    CX = E[(i, j, k)[axis].ravel()]
    CH, relres, nit, _ = homogenize(gputo, tuple(shape), 3, CX, fields=False)
    Cx = PO.backus_x(E, [0.3] * 16, [1.0 / 16] * 16)
    perm = [np.arange(6), np.array([1, 0, 2, 3, 5, 4]), np.array([2, 1, 0, 4, 3, 5])][axis]
    np.testing.assert_allclose(CH, Cx[np.ix_(perm, perm)], rtol=1e-8, atol=1e-10)


def test_homogeneous_cell(gputo):
    CH, relres, nit, chi = homogenize(gputo, (8, 8, 8), 3, np.full(512, 0.7), tol=1e-10)
    np.testing.assert_allclose(CH, 0.7 * PO.isotropic_D(1.0, 0.3), rtol=1e-10, atol=1e-12)
    assert np.abs(chi).max() < 1e-8


@pytest.mark.parametrize("kw", [dict(precision=1), dict(precision=2, l0_kernel=3), dict(preset=2),
                                dict(preset=1)], ids=["mixed_f32", "mixed_mfma", "mgcg9", "cfg0"])
def test_periodic_variants_match_f64(gputo, kw):
    shape = (16, 16, 16)
    CX = two_phase(shape, 5)
    ref, _, nit0, _ = homogenize(gputo, shape, 4, CX, tol=1e-12, fields=False)
    CH, relres, nit, _ = homogenize(gputo, shape, 4, CX, tol=1e-12, maxit=2000, fields=False, **kw)
    print(kw, "nit", nit0, nit, "dCH", rel(CH, ref))
    assert (relres < 1e-12).all()
    assert rel(CH, ref) < 1e-9


@pytest.mark.parametrize("kw", [dict(), dict(precision=1), dict(precision=2, l0_kernel=3), dict(preset=2),
                                dict(preset=1)], ids=["f64", "mixed_f32", "mixed_mfma", "mgcg9", "cfg0"])
def test_homogenize_together_matches_one_by_one(gputo, kw):
    """The six load cases solved together (one launch per operator for all
    six vectors) against the same six solved one after the other: the FP64
    solves run the same arithmetic per right-hand side, so they agree to
    rounding (iteration counts equal); the mixed ones apply level 0 with the
    FP32 node-gather kernel when batched, so they agree to the tolerance."""
    shape = (16, 16, 16)
    CX = two_phase(shape, 11)
    CHs, rs, ns, chis = homogenize(gputo, shape, 4, CX, tol=1e-12, maxit=2000, homog_sequential=1, **kw)
    CHb, rb, nb_, chib = homogenize(gputo, shape, 4, CX, tol=1e-12, maxit=2000, **kw)
    print(kw, "nit", list(ns), list(nb_), "dCH", rel(CHb, CHs), "dchi", rel(chib, chis))
    assert (rb < 1e-12).all() and (rs < 1e-12).all()
    if not kw or kw.get("preset") in (1, 2):
        assert list(nb_) == list(ns)
        assert rel(CHb, CHs) < 1e-13
        assert rel(chib, chis) < 1e-12
    else:
        assert rel(CHb, CHs) < 1e-9
        assert rel(chib, chis) < 1e-6


def test_translation_invariance(gputo):
    shape = (8, 16, 8)
    CX = two_phase(shape, 7)
    CH, *_ = homogenize(gputo, shape, 3, CX, fields=False)
    rolled = np.roll(CX.reshape(8, 8, 16), (3, 1, 5), axis=(0, 1, 2)).ravel()
    CH2, *_ = homogenize(gputo, shape, 3, rolled, fields=False)
    assert rel(CH2, CH) < 1e-9


def test_periodic_mgcg1_splits_and_images(gputo):
    """gputo_mgcg1 on a periodic handle: a load split between a node and its
    images acts as one load on the master; the solution is equal on images
    and matches the oracle's pinned solve."""
    from gputo.problems import ke0_h8
    shape = (8, 8, 8)
    nelx, nely, nelz = shape
    CX = two_phase(shape, 9)
    s = gputo.Solver(*shape, 3, periodic=True)
    nn = (nelx + 1, nelz + 1, nely + 1)
    rng = np.random.default_rng(4)
    Fm = rng.normal(size=(nelx, nelz, nely, 3))
    Fm[0, 0, 0] = 0.0  # no load on the pinned node
    F = np.zeros(nn + (3,))
    F[:nelx, :nelz, :nely] = Fm
    # move half of the i = 0 face's load to its images at i = nelx
    F[nelx, :nelz, :nely] = 0.5 * F[0, :nelz, :nely]
    F[0, :nelz, :nely] *= 0.5
    # pin node 0 and its images
    fixed = []
    for a in (0, nelx):
        for c in (0, nelz):
            for b in (0, nely):
                n = (nely + 1) * (nelz + 1) * a + (nely + 1) * c + b
                fixed += [3 * n + 1, 3 * n + 2, 3 * n + 3]
    F = F.ravel()
    U, relres, nit, rc = s.solve(ke0_h8(), F, CX, np.array(fixed, dtype=float), 1e-12, 500)
    s.close()
    assert rc == 0
    U[1::3] *= -1  # back to the C frame
    Uf = U.reshape(nn + (3,))
    np.testing.assert_array_equal(Uf[-1], Uf[0])
    np.testing.assert_array_equal(Uf[:, -1], Uf[:, 0])
    np.testing.assert_array_equal(Uf[:, :, -1], Uf[:, :, 0])
    K, _ = PO.assemble(ke0_h8(), CX, *shape)
    fp = Fm.ravel().copy()
    free = np.arange(3, K.shape[0])
    fp[:3] = 0.0
    import scipy.sparse.linalg as spla
    uo = np.zeros_like(fp)
    uo[free] = spla.spsolve(K[free][:, free].tocsc(), fp[free])
    assert rel(PO.full_to_periodic(U, *shape), uo) < 1e-8


def test_config4_cell_96(gputo):
    """The 96^3 unit cell (two-phase SIMP, smoothed random field with a
    central spherical void): the FP64 MGCG against the mixed-precision solve
    (split-BF16 MFMA level 0 when one by one, FP32 node-gather level 0 when
    batched), each with the six load cases solved together and one after the
    other.  The tensors agree to 1e-6; the times are printed."""
    from gputo.problems import inclusion_cell, simp_cx
    n = 96
    CX = simp_cx(inclusion_cell(n, seed=0))
    out = {}
    for name, kw in (("f64", {}), ("f64_seq", dict(homog_sequential=1)),
                     ("mixed", dict(precision=2, l0_kernel=3)),
                     ("mixed_seq", dict(precision=2, l0_kernel=3, homog_sequential=1))):
        t0 = time.perf_counter()
        CH, relres, nit, _ = homogenize(gputo, (n, n, n), 5, CX, tol=1e-8, maxit=500, fields=False, **kw)
        dt = time.perf_counter() - t0
        print("config4", name, "nit", list(nit), "seconds %.3f" % dt, flush=True)
        assert (relres < 1e-8).all()
        out[name] = CH
    print("CH f64 diag", np.diag(out["f64"]))
    for name in ("f64_seq", "mixed", "mixed_seq"):
        assert rel(out[name], out["f64"]) < 1e-6, name
    assert np.linalg.eigvalsh(out["f64"]).min() > 0
