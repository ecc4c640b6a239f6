"""Pins the oracle's LM pieces: normal equations (numpy brute force), the
pseudo-inverse solve (numpy eigendecomposition / lstsq), the classical
estimates (weak-field closed form, quadrant table, ghost-sample trimming,
NaN guards) and the whole LM (noise-free closed loop; the reference's
invert example recovers run/init.mod from data/synthesis_gpu.per)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, crisp_table, model_row, read_mod

FE6173 = (6173.3356, (2, 1, 1.0, 2, 2, 0.0), 1.0)
GRID = np.arange(-350.0, 665.0 + 1e-9, 35.0)
FREE9 = ("eta0", "B", "vlos", "dopp", "aa", "gm", "az", "S0", "S1")
BENCH_INIT = dict(eta0=8.0, B=400.0, vlos=0.25, dopp=0.034, aa=0.06, gm=30.0, az=30.0,
                  S0=0.15, S1=0.85, mac=0.0, alfa=1.0)


def setup(orc, psf=None, **kw):
    kw.setdefault("init", BENCH_INIT)
    kw.setdefault("free", FREE9)
    return orc.make_setup([FE6173], GRID, psf=psf, **kw)


def _m(**kw):
    d = dict(eta0=10, B=1200, vlos=0.5, dopp=0.03, aa=0.1, gm=60, az=40, S0=0.3, S1=0.6,
             mac=0, alfa=1)
    d.update(kw)
    return model_row(d)


def test_normal_equations_brute_force(orc):
    s = setup(orc, psf=("gauss", 60.0), weights=(1.0, 0.7, 0.8, 1.3), noise=2e-6)
    rng = np.random.default_rng(0)
    m = _m()
    sp, rf = orc.synth_rf(s, m)
    obs = (sp[0] + rng.normal(0, 1e-3, sp[0].shape)).astype(np.float32)
    alpha, beta = orc.covarm(s, obs, sp[0], rf[0])
    w = np.array([1.0, 0.7, 0.8, 1.3]) / 2e-6
    J = rf[0]                                   # [stokes, param, lambda]
    r = sp[0] - obs.astype(np.float64)
    a_ref = np.einsum("s,spl,sql->pq", w, J, J)
    b_ref = np.einsum("s,sl,spl->p", w, r, J)
    np.testing.assert_allclose(alpha, a_ref, rtol=1e-12, atol=1e-6)
    np.testing.assert_allclose(beta, b_ref, rtol=1e-12, atol=1e-6)
    np.testing.assert_array_equal(alpha, alpha.T)
    assert np.all(alpha[9] == 0) and np.all(beta[9] == 0)        # slot 9 fixed
    c2 = orc.chisqr(s, obs, sp[0])
    assert c2 == pytest.approx((w[:, None] * r * r).sum() / (4 * 30 - 10), rel=1e-12)


def _pinv_ref(A, b):
    lam, V = np.linalg.eigh(A)
    w = np.abs(lam)
    inv = np.where(w > 1e-12, 1.0 / np.where(w > 0, w, 1), 0.0)
    return V @ (inv * (V.T @ b))


def test_svd_solve(orc):
    rng = np.random.default_rng(1)
    for trial in range(20):
        J = rng.normal(size=(40, 10)) * rng.uniform(0.1, 100, 10)
        A = J.T @ J
        b = rng.normal(size=10)
        x, w = orc.svd_solve(A, b)
        np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-8)
        np.testing.assert_allclose(np.sort(w), np.sort(np.linalg.eigvalsh(A)), rtol=1e-10)
    # zero rows/columns (fixed parameters): pseudo-inverse, components zeroed
    A[3, :] = A[:, 3] = 0.0
    A[9, :] = A[:, 9] = 0.0
    x, _ = orc.svd_solve(A, b)
    np.testing.assert_allclose(x, np.linalg.pinv(A) @ b, rtol=1e-8, atol=1e-12)
    assert x[3] == 0.0 and x[9] == 0.0
    # eigenvalue below the 1e-12 threshold is dropped, as in milosUtils.cu:311
    Q, _ = np.linalg.qr(rng.normal(size=(10, 10)))
    lam = np.array([5.0, 3.0, 1.0, 0.5, 0.2, 0.1, 1e-3, 1e-6, 1e-9, 1e-14])
    A = (Q * lam) @ Q.T
    x, _ = orc.svd_solve(A, b)
    np.testing.assert_allclose(x, _pinv_ref(A, b), rtol=1e-6)
    # indefinite symmetric matrix: the reference formula uses V |w|^-1 V^T
    A = (Q * np.array([2.0, -1.0, 1, 1, 1, 1, 1, 1, 1, 1])) @ Q.T
    x, _ = orc.svd_solve(A, b)
    np.testing.assert_allclose(x, _pinv_ref(A, b), rtol=1e-9)


def test_classical_estimates_weak_field(orc):
    """Weak field, no PSF: the centre-of-gravity B_los recovers B cos(gamma)
    and the line-centre vlos within the method's known bias."""
    s = setup(orc)
    for B, g in [(50.0, 30.0), (80.0, 140.0)]:
        m = _m(B=B, gm=g, vlos=0.3, eta0=6.0, dopp=0.035, aa=0.05)
        sp = orc.synth(s, m)[0].astype(np.float32)
        est = orc.classical_estimates(s, sp, np.array(s.init_model[:]), for_initial_use=0)
        blos = est[7]                                   # stored in S0 when not initial
        assert blos == pytest.approx(B * np.cos(np.radians(g)), rel=0.1)
        # vlos estimate: centroid / 2 (milosUtils.cu:470)
        assert est[2] == pytest.approx(0.3 / 2.0, abs=0.06)


def test_classical_estimates_azimuth_quadrants(orc):
    s = setup(orc)
    n = 30
    base = np.zeros((4, n), dtype=np.float32)
    base[0] = 0.9
    cases = {(1, 1): (0, 90), (-1, 1): (90, 180), (-1, -1): (45, 135), (1, -1): (45, 135)}
    for (su, sq), (lo, hi) in cases.items():
        sp = base.copy()
        sp[1, 10:14] = sq * 0.01
        sp[2, 10:14] = su * 0.02
        est = orc.classical_estimates(s, sp)
        phi_ref = np.degrees(np.arctan((su * 0.08) / (sq * 0.04))) / 2
        phi_ref += {(1, 1): 0, (-1, 1): 180, (-1, -1): 90, (1, -1): 90}[(su, sq)]
        assert est[6] == pytest.approx(phi_ref, abs=1e-9)
        assert lo <= est[6] <= hi


def test_classical_estimates_ghost_samples(orc):
    """Leading/trailing I < 0 samples are excluded from the centroid sums."""
    s = setup(orc)
    m = _m(B=600.0, gm=50.0, vlos=0.4)
    sp = orc.synth(s, m)[0].astype(np.float32)
    e1 = orc.classical_estimates(s, sp)
    sp2 = sp.copy()
    sp2[0, :2] = -5.0
    sp2[0, -1] = -5.0
    e2 = orc.classical_estimates(s, sp2)
    # same as a grid that simply lacks those samples (azimuth sums cover all)
    sub = orc.make_setup([FE6173], GRID[2:-1], init=BENCH_INIT, free=FREE9)
    e3 = orc.classical_estimates(sub, np.ascontiguousarray(sp[:, 2:-1]))
    np.testing.assert_allclose(e2[[1, 2, 5]], e3[[1, 2, 5]], rtol=1e-9)
    assert e1[1] != e2[1]


def test_lm_nan_guards(orc):
    """An all-zero pixel makes the estimates NaN; the guards of
    milosUtils.cu:621-628 replace them and the LM still returns finite."""
    s = setup(orc, miter=5)
    r = orc.lm_pixel(s, np.zeros((4, 30), dtype=np.float32))
    assert np.all(np.isfinite(r["model"]))


def test_lm_closed_loop_noise_free(orc):
    s = setup(orc, psf=("gauss", 60.0), miter=200)
    m = _m(eta0=9.0, B=1400.0, vlos=0.6, dopp=0.033, aa=0.12, gm=70.0, az=35.0, S0=0.3,
           S1=0.6)
    obs = orc.synth(s, m)[0].astype(np.float32)
    r = orc.lm_pixel(s, obs)
    np.testing.assert_allclose(r["model"][:9], m[:9], rtol=2e-3, atol=1e-3)
    assert r["chisqr"] < 1e-3


def test_lm_recovers_reference_example(orc):
    """README.md:76-89: invert data/synthesis_gpu.per (CRISP PSF, 100 cycles,
    S/N 1000) from run/guess.mod; the answer is run/init.mod."""
    per = np.loadtxt(os.path.join(GOLDEN, "synthesis_gpu.per"))
    obs = np.ascontiguousarray(per[:, 2:6].T.astype(np.float32))
    guess = read_mod(os.path.join(GOLDEN, "guess.mod"))
    truth = read_mod(os.path.join(GOLDEN, "init.mod"))
    dl, val = crisp_table()
    s = orc.make_setup([FE6173], GRID, psf=("table", dl, val), init=guess, free=FREE9,
                       miter=100, noise=1e-6)
    r = orc.lm_pixel(s, obs)
    t = model_row(truth)
    np.testing.assert_allclose(r["model"][:9], t[:9], rtol=5e-3, atol=2e-3)
