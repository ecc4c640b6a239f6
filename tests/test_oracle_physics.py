"""Pins the FP64 oracle's physics routines against external truth:
scipy's Faddeeva function, closed forms, finite differences, numpy brute
force and the reference's own synthesis fixture (data/synthesis_gpu.per)."""
import math
import os

import numpy as np
import pytest
from scipy.special import dawsn, erfc

from conftest import GOLDEN, crisp_table, model_row, read_mod

FE6173 = (6173.3356, (2, 1, 1.0, 2, 2, 0.0), 1.0)
GRID = np.arange(-350.0, 665.0 + 1e-9, 35.0)


def setup(orc, psf=None, **kw):
    return orc.make_setup([FE6173], GRID, psf=psf, **kw)


# ---- Voigt ------------------------------------------------------------------
def test_voigt_vs_faddeeva_golden(orc):
    g = np.load(os.path.join(GOLDEN, "voigt_wofz.npz"))
    err = 0.0
    for i, a in enumerate(g["a"]):
        for j, u in enumerate(g["u"]):
            H, F = orc.voigt(a, u)
            err = max(err, abs(H - g["H"][i, j]), abs(F - g["F"][i, j]))
    # Hui, Armstrong & Wray (1978) quote ~1e-6 accuracy; measured 4.4e-6
    assert err < 6e-6


def test_voigt_closed_forms(orc):
    for u in np.linspace(-6, 6, 25):
        H, F = orc.voigt(1e-9, u)
        assert abs(H - math.exp(-u * u)) < 5e-6                  # H(0,u) = exp(-u^2)
        assert abs(F - dawsn(u) / math.sqrt(math.pi)) < 5e-6     # F(0,u) = D(u)/sqrt(pi)
    for a in [0.01, 0.1, 0.5, 1.0, 3.0]:
        H, F = orc.voigt(a, 0.0)
        assert abs(H - math.exp(a * a) * erfc(a)) < 5e-6         # H(a,0)
        assert abs(F) < 1e-12                                    # F odd in u


def test_voigt_derivative_identities(orc):
    # dH/du = -2uH + 4aF and dF/du = 1/sqrt(pi) - aH - 2uF (me_der.cu:63-64)
    worst = 0.0
    for a in [1e-3, 0.05, 0.3, 1.0]:
        for u in np.linspace(-8, 8, 81):
            H, F = orc.voigt(a, u)
            h = 1e-5
            Hp, Fp = orc.voigt(a, u + h)
            Hm, Fm = orc.voigt(a, u - h)
            worst = max(worst, abs((Hp - Hm) / (2 * h) - (4 * a * F - 2 * u * H)),
                        abs((Fp - Fm) / (2 * h) - (1 / math.sqrt(math.pi) - a * H - 2 * u * F)))
    assert worst < 5e-5


# ---- Zeeman ------------------------------------------------------------------
def test_zeeman_6173_closed_form(orc):
    z = orc.cuanten((2, 1, 1.0, 2, 2, 0.0))
    assert (z.n_pi, z.n_sig) == (1, 1)
    assert z.gl == pytest.approx(2.5) and z.gu == 0.0 and z.geff == pytest.approx(2.5)
    assert z.nup[0] == 0.0 and z.nub[0] == pytest.approx(-2.5) and z.nur[0] == pytest.approx(2.5)
    assert z.wep[0] == 1.0 and z.web[0] == 1.0 and z.wer[0] == 1.0


def test_zeeman_6301_closed_form(orc):
    z = orc.cuanten((2, 1, 2.0, 2, 2, 2.0))
    g1, g2 = 11.0 / 6.0, 1.5
    assert (z.n_pi, z.n_sig) == (5, 4)
    assert z.gl == pytest.approx(g1) and z.gu == pytest.approx(g2)
    assert z.geff == pytest.approx((g1 + g2) / 2)
    # pi strengths ~ M^2 (J -> J), sigma strengths (J-M)(J+M+1)/4 normalised
    np.testing.assert_allclose(z.wep[:5], np.array([4, 1, 0, 1, 4]) / 10.0)
    np.testing.assert_allclose(z.web[:4], np.array([1, 1.5, 1.5, 1]) / 5.0)
    # shifts g1*M_lo - g2*M_up
    np.testing.assert_allclose(z.nub[:4], [g1 * ml - g2 * (ml + 1) for ml in (-2, -1, 0, 1)])
    np.testing.assert_allclose(z.nur[:4], [g1 * ml - g2 * (ml - 1) for ml in (-1, 0, 1, 2)])
    # mirror symmetry used by the synthesis (mil_sinrf.cu:74-76)
    np.testing.assert_allclose([-z.nub[3 - i] for i in range(4)], z.nur[:4])


# ---- PSF ---------------------------------------------------------------------
def test_gaussian_psf_closed_form(orc):
    lam = 6173.3356 + GRID / 1000.0
    p = orc.fgauss_wl(60.0, lam[1] - lam[0], lam[0], lam[15], 30)
    sigma = 60.0 * 0.42466090 / 1000.0
    x = lam[0] + np.arange(30) * (lam[1] - lam[0])        # fgaussC.c:31
    ref = np.exp(-0.5 * ((x - lam[15]) / sigma) ** 2)
    np.testing.assert_allclose(p, ref / ref.sum(), rtol=1e-12, atol=1e-300)
    assert p.argmax() == 15 and p.sum() == pytest.approx(1.0)


def test_table_psf_matches_numpy_interp(orc):
    dl, val = crisp_table()
    p = orc.psf_interp(dl, val, GRID, 175.0)
    ref = np.interp(GRID - 175.0, dl, val)
    np.testing.assert_allclose(p, ref / ref.sum(), rtol=1e-12)
    assert p.argmax() == 15


def test_convolution_brute_force(orc):
    rng = np.random.default_rng(3)
    for n in (7, 30, 31):
        x = rng.normal(size=n)
        h = rng.random(n)
        c = n // 2
        ref = np.array([sum(h[j] * x[k + j - c] for j in range(n) if 0 <= k + j - c < n)
                        for k in range(n)])
        np.testing.assert_allclose(orc.conv(x, h), ref, rtol=1e-13, atol=1e-13)
        # Ic variant: Ic - conv(Ic - x)
        ic = 0.9
        ref_ic = ic - np.array([sum(h[j] * (ic - x[k + j - c]) for j in range(n)
                                    if 0 <= k + j - c < n) for k in range(n)])
        np.testing.assert_allclose(orc.conv(x, h, ic), ref_ic, rtol=1e-13, atol=1e-13)
    # delta kernel is the identity; constants survive the Ic variant
    d = np.zeros(30); d[15] = 1.0
    y = rng.normal(size=30)
    np.testing.assert_array_equal(orc.conv(y, d), y)
    np.testing.assert_allclose(orc.conv(np.full(30, 0.8), rng.random(30), 0.8), 0.8, rtol=1e-14)


# ---- synthesis -----------------------------------------------------------------
def test_synthesis_matches_reference_fixture(orc):
    """data/synthesis_gpu.per: the reference's synthesis of run/init.mod with
    the CRISP PSF on run/malla.grid (E3 path, gmilos.cu:432-628)."""
    per = np.loadtxt(os.path.join(GOLDEN, "synthesis_gpu.per"))
    init = read_mod(os.path.join(GOLDEN, "init.mod"))
    dl, val = crisp_table()
    s = setup(orc, psf=("table", dl, val), init=init)
    sp = orc.synth(s, model_row(init))[0]
    ref = per[:, 2:6].T
    np.testing.assert_allclose(per[:, 1], GRID, atol=1e-6)
    # the fixture is printed with 7 significant digits
    assert np.abs(sp - ref).max() < 2e-7
    assert (np.abs(sp - ref) <= 1e-6 * np.abs(ref) + 2e-8).all()


def _model(**kw):
    d = dict(eta0=10, B=1200, vlos=0.5, dopp=0.03, aa=0.1, gm=60, az=40, S0=0.3, S1=0.6,
             mac=0, alfa=1)
    d.update(kw)
    return model_row(d)


def test_synthesis_invariants(orc):
    s = setup(orc)
    # no absorption: I = S0 + S1 mu, no polarisation
    sp = orc.synth(s, _model(eta0=1e-12))[0]
    np.testing.assert_allclose(sp[0], 0.9, atol=1e-9)
    assert np.abs(sp[1:]).max() < 1e-9
    # no field: no polarisation
    sp = orc.synth(s, _model(B=0.0))[0]
    assert np.abs(sp[1:]).max() < 1e-15
    # azimuth period 180 deg, (Q,U) ~ (cos 2phi, sin 2phi)
    a = orc.synth(s, _model(az=10.0))[0]
    b = orc.synth(s, _model(az=190.0))[0]
    np.testing.assert_allclose(a, b, atol=1e-12)
    q0 = orc.synth(s, _model(az=0.0))[0]
    q45 = orc.synth(s, _model(az=45.0))[0]
    np.testing.assert_allclose(q45[2], q0[1], rtol=1e-9, atol=1e-14)   # U(45) = Q(0)
    # gamma -> 180 - gamma flips V and keeps I (Q, U change through the
    # magneto-optical eta_V rho_U - eta_U rho_V coupling)
    g1 = orc.synth(s, _model(gm=30.0))[0]
    g2 = orc.synth(s, _model(gm=150.0))[0]
    np.testing.assert_allclose(g1[3], -g2[3], atol=1e-12)
    np.testing.assert_allclose(g1[0], g2[0], atol=1e-12)


def test_weak_field_limit(orc):
    """V ~ -C lambda0^2 g_eff B cos(gamma) dI/dlambda for weak fields, with
    dI/dlambda = -(c/lambda0) dI/dvlos taken from the field-free synthesis."""
    s = setup(orc)
    base = dict(gm=0.0, eta0=5.0, dopp=0.04)
    sp = orc.synth(s, _model(B=10.0, **base))[0]
    h = 1e-4
    dIdv = (orc.synth(s, _model(B=0.0, vlos=0.5 + h, **base))[0][0] -
            orc.synth(s, _model(B=0.0, vlos=0.5 - h, **base))[0][0]) / (2 * h)
    dIdl = -(2.99792458e5 / 6173.3356) * dIdv
    pred = -4.6686411e-13 * 6173.3356 ** 2 * 2.5 * 10.0 * dIdl
    assert np.abs(sp[3] - pred).max() < 0.02 * np.abs(pred).max()


# ---- response functions ---------------------------------------------------------
@pytest.mark.parametrize("psf", [None, ("gauss", 60.0)])
def test_response_functions_finite_difference(orc, psf):
    """Exact-derivative mode against central differences of the synthesis.
    The floor (~1e-4) is Hui's approximation not obeying the Faddeeva
    derivative identity exactly; with the PSF the Ic-trick picks the larger
    end sample of the RF (me_der.cu:171-174), not of the spectrum, so vlos
    is excluded there."""
    s = setup(orc, psf=psf, rf_reference_rules=0)
    m = _model(eta0=12, B=1500, vlos=0.7, dopp=0.035, aa=0.2, gm=50, az=30, S0=0.25, S1=0.7)
    sp, rf = orc.synth_rf(s, m)
    for p in range(9):
        if psf and p == 2:
            continue
        h = max(abs(m[p]) * 1e-6, 1e-7)
        mp, mm = m.copy(), m.copy()
        mp[p] += h
        mm[p] -= h
        fd = (orc.synth(s, mp)[0] - orc.synth(s, mm)[0]) / (2 * h)
        err = np.abs(fd - rf[0][:, p, :]).max() / np.abs(fd).max()
        assert err < 1e-3, (p, err)


def test_response_functions_reference_rule(orc):
    """The reference stores w*dF for the magneto-optical Voigt derivatives
    while shi = 2*w*F (me_der.cu:71-74): its rho-derivatives for
    B, vlos, dopp, a are half the exact ones.  The default mode keeps that
    rule; the eta parts and the other parameters are unaffected."""
    m = _model(eta0=12, B=1500, vlos=0.7, dopp=0.035, aa=0.2, gm=50, az=30)
    s_ref = setup(orc, rf_reference_rules=1)
    s_ex = setup(orc, rf_reference_rules=0)
    _, rf_ref = orc.synth_rf(s_ref, m)
    _, rf_ex = orc.synth_rf(s_ex, m)
    for p in (0, 5, 6, 7, 8):
        np.testing.assert_array_equal(rf_ref[0][:, p], rf_ex[0][:, p])
    for p in (1, 2, 3, 4):
        assert np.abs(rf_ref[0][:, p] - rf_ex[0][:, p]).max() > 1e-3 * np.abs(rf_ex[0][:, p]).max()
    # without magneto-optics (no Faraday term: etaq etc. only) they agree:
    # at B = 0 rho_Q,U,V vanish and only the I, V eta-terms remain
    _, a0 = orc.synth_rf(s_ref, _model(B=0.0))
    _, b0 = orc.synth_rf(s_ex, _model(B=0.0))
    np.testing.assert_allclose(a0[0][0, 2], b0[0][0, 2], rtol=1e-12)
