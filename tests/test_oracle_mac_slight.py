"""Pins the oracle's macroturbulence and stray-light routines (the reference's
11-term physics, mil_sinrf.cu:101-140, me_der.cu:119-211, fgauss.cu:16-63,
convolution.cu:90-118) against the circular-convolution theorem (numpy FFT),
closed forms of the Gaussian kernel, conservation laws and central finite
differences of the synthesis."""
import numpy as np
import pytest

from conftest import model_row

FE6173 = (6173.3356, (2, 1, 1.0, 2, 2, 0.0), 1.0)
GRID = np.arange(-350.0, 665.0 + 1e-9, 35.0)
FIELDS = ["eta0", "B", "vlos", "dopp", "aa", "gm", "az", "S0", "S1", "mac", "alfa"]


def setup(orc, psf=None, free=None, **kw):
    return orc.make_setup([FE6173], GRID, psf=psf, free=free, **kw)


def _model(**kw):
    d = dict(eta0=12, B=1500, vlos=0.7, dopp=0.035, aa=0.2, gm=50, az=30, S0=0.25, S1=0.7,
             mac=1.5, alfa=1)
    d.update(kw)
    return model_row(d)


@pytest.mark.parametrize("n", [30, 31, 8])
def test_circular_convolution_theorem(orc, n):
    """y_m = c_{(m + n/2) mod n} with c = ifft(fft(x) fft(h)): the centred
    circular convolution of convolution.cu:90-118 for even and odd n."""
    rng = np.random.default_rng(n)
    x, h = rng.normal(size=n), rng.normal(size=n)
    c = np.real(np.fft.ifft(np.fft.fft(x) * np.fft.fft(h)))
    np.testing.assert_allclose(orc.conv_circ(x, h), np.roll(c, -(n // 2)), atol=1e-12)
    # a delta at the centre is the identity, one sample later a shift by one
    d = np.zeros(n)
    d[n // 2] = 1
    np.testing.assert_allclose(orc.conv_circ(x, d), x, atol=0)
    d = np.roll(d, 1)
    np.testing.assert_allclose(orc.conv_circ(x, d), np.roll(x, 1), atol=0)


def test_mac_kernel_closed_form(orc):
    """fgauss.cu:16-63: exp(-(dl/sigma)^2/2) normalised, sigma = lambda0 mac / c."""
    s = setup(orc)
    lam = np.array(s.lambda_[:30])
    np.testing.assert_allclose(lam, 6173.3356 + GRID / 1000.0, rtol=0, atol=1e-9)
    for mac in (0.3, 1.0, 2.5):
        G, _ = orc.mac_kernels(s, mac)
        sig = 6173.3356 * mac / 2.99792458e5
        g = np.exp(-0.5 * ((lam - lam[15]) / sig) ** 2)
        np.testing.assert_allclose(G, g / g.sum(), rtol=1e-12, atol=1e-300)
        assert abs(G.sum() - 1) < 1e-14 and G.argmax() == 15


def test_mac_kernel_derivative(orc):
    """Exact mode against central differences in mac; the reference's
    (x^2 - 1) G / mac (me_der.cu:124-131) agrees where the grid resolves the
    Gaussian (sum G x^2 = 1) and not where it does not."""
    s_ex = setup(orc, rf_reference_rules=0)
    s_ref = setup(orc, rf_reference_rules=1)
    for mac in (0.8, 2.0, 3.5):
        h = 1e-6 * mac
        fd = (orc.mac_kernels(s_ex, mac + h)[0] - orc.mac_kernels(s_ex, mac - h)[0]) / (2 * h)
        _, gd = orc.mac_kernels(s_ex, mac)
        assert np.abs(gd - fd).max() < 1e-7 * np.abs(fd).max() + 1e-12
    # sigma = 2.4 grid steps: discrete second moment is 1 to ~1e-20
    _, gr = orc.mac_kernels(s_ref, 3.5)
    _, ge = orc.mac_kernels(s_ex, 3.5)
    np.testing.assert_allclose(gr, ge, atol=1e-12 * np.abs(ge).max())
    # sigma = 0.27 steps: the normalisation term matters
    _, gr = orc.mac_kernels(s_ref, 0.4)
    _, ge = orc.mac_kernels(s_ex, 0.4)
    assert np.abs(gr - ge).max() > 1e-3 * np.abs(ge).max()


@pytest.mark.parametrize("psf", [None, ("gauss", 60.0)])
def test_mac_synthesis_conserves_and_switches(orc, psf):
    s = setup(orc, psf=psf)
    sp0 = orc.synth(s, _model(mac=0.0))[0]
    # mac <= 1e-4 leaves the reference's PSF path (mil_sinrf.cu:101)
    np.testing.assert_array_equal(orc.synth(s, _model(mac=1e-4))[0], sp0)
    if psf is None:
        # unit-sum circular kernel: every Stokes parameter keeps its sum
        sp = orc.synth(s, _model(mac=2.0))[0]
        np.testing.assert_allclose(sp.sum(axis=1), sp0.sum(axis=1), rtol=1e-12, atol=1e-14)
        assert np.abs(sp - sp0).max() > 1e-3
    else:
        # PSF and macroturbulence combine into one kernel, the Gaussian
        # correlated with the PSF (mil_sinrf.cu:104-107): on the unconvolved
        # spectrum that is the circular convolution with that kernel
        s_np = setup(orc)
        raw = orc.synth(s_np, _model(mac=0.0))[0]
        G, _ = orc.mac_kernels(s, 2.0)
        sp = orc.synth(s, _model(mac=2.0))[0]
        for st in range(4):
            np.testing.assert_allclose(sp[st], orc.conv_circ(raw[st], G), atol=1e-14)


@pytest.mark.parametrize("psf", [None, ("gauss", 60.0)])
def test_mac_response_functions_finite_difference(orc, psf):
    """Exact-mode RFs with macroturbulence against central differences,
    including d/dmac in slot 9 (kept when the filling-factor flag is set
    and no stray light replaces it, me_der.cu:139-141)."""
    s = setup(orc, psf=psf, rf_reference_rules=0, free=FIELDS[:9] + ["alfa"])
    m = _model(mac=1.7)
    sp, rf = orc.synth_rf(s, m)
    np.testing.assert_allclose(sp[0], orc.synth(s, m)[0], atol=0)
    for p in list(range(9)) + [9]:
        h = max(abs(m[p]) * 1e-6, 1e-7)
        mp, mm = m.copy(), m.copy()
        mp[p] += h
        mm[p] -= h
        fd = (orc.synth(s, mp)[0] - orc.synth(s, mm)[0]) / (2 * h)
        err = np.abs(fd - rf[0][:, p, :]).max() / np.abs(fd).max()
        assert err < 1e-3, (p, err)


def test_stray_light_mixing_and_responses(orc):
    rng = np.random.default_rng(5)
    L = np.abs(rng.normal(0.8, 0.05, size=(4, 30)))
    L[1:] = rng.normal(0, 1e-3, size=(3, 30))
    s = orc.set_slight(setup(orc, psf=("gauss", 60.0), rf_reference_rules=0,
                             free=FIELDS[:9] + ["alfa"]), L)
    s_off = setup(orc, psf=("gauss", 60.0), rf_reference_rules=0, free=FIELDS[:9] + ["alfa"])
    pure = orc.synth(s_off, _model(mac=0.0))[0]
    np.testing.assert_allclose(orc.synth(s, _model(mac=0.0, alfa=1.0))[0], pure, atol=0)
    np.testing.assert_allclose(orc.synth(s, _model(mac=0.0, alfa=0.0))[0], L, atol=1e-15)
    al = 0.65
    m = _model(mac=0.9, alfa=al)
    sp, rf = orc.synth_rf(s, m)
    for p in list(range(9)) + [10]:
        h = max(abs(m[p]) * 1e-6, 1e-7)
        mp, mm = m.copy(), m.copy()
        mp[p] += h
        mm[p] -= h
        fd = (orc.synth(s, mp)[0] - orc.synth(s, mm)[0]) / (2 * h)
        slot = 9 if p == 10 else p
        err = np.abs(fd - rf[0][:, slot, :]).max() / np.abs(fd).max()
        assert err < 1e-3, (p, err)
    # the LM loop's variant takes the mixed spectrum (milosUtils.cu:673):
    # alfa times the exact filling-factor response
    _, rf_mix = orc.synth_rf(s, m, alpha_rf_mixed=1)
    np.testing.assert_allclose(rf_mix[0][:, 9], sp[0] - L, atol=1e-15)
    np.testing.assert_allclose(rf_mix[0][:, 9], al * rf[0][:, 9], atol=1e-14)
    # with the filling factor fixed, slot 9 is zeroed (milosUtils.cu:254-262)
    s_fix = orc.set_slight(setup(orc, psf=("gauss", 60.0)), L)
    _, rf_fix = orc.synth_rf(s_fix, m)
    assert np.abs(rf_fix[0][:, 9]).max() == 0


def test_eleven_term_layout_finite_difference(orc):
    """lm_mils_11 layout (mac and alfa free): slot 9 is d/dmac scaled by alfa,
    slot 10 d/dalfa (me_der.cu:191-207), against central differences of the
    mixed synthesis."""
    L = np.zeros((4, 30))
    L[0] = 0.9 - 0.3 * np.exp(-((GRID - 20.0) / 120.0) ** 2)
    s = orc.set_slight(setup(orc, psf=("gauss", 60.0), rf_reference_rules=0, free=FIELDS), L)
    assert orc.nterms(s) == 11
    m = _model(mac=1.4, alfa=0.7)
    sp, rf = orc.synth_rf(s, m)
    assert rf.shape == (1, 4, 11, 30)
    for p in range(11):
        h = max(abs(m[p]) * 1e-6, 1e-7)
        mp, mm = m.copy(), m.copy()
        mp[p] += h
        mm[p] -= h
        fd = (orc.synth(s, mp)[0] - orc.synth(s, mm)[0]) / (2 * h)
        err = np.abs(fd - rf[0][:, p, :]).max() / np.abs(fd).max()
        assert err < 1e-3, (p, err)
    A, b = orc.covarm(s, sp[0].astype(np.float32), sp[0], rf[0])
    assert A.shape == (11, 11) and b.shape == (11,)


@pytest.mark.parametrize("free", [FIELDS[:9] + ["alfa"], FIELDS])
def test_filling_factor_lm_recovers_truth(orc, free):
    """Noise-free closed loop with stray light: the 10-term run (filling
    factor free, mac fixed at its true value) and the 11-term run (both
    free) recover the synthesising model."""
    L = np.zeros((4, 30))
    L[0] = 0.95 - 0.4 * np.exp(-((GRID - 40.0) / 110.0) ** 2)
    truth = _model(eta0=9, B=900, vlos=0.4, dopp=0.032, aa=0.15, gm=65, az=35, S0=0.3,
                   S1=0.65, mac=1.2, alfa=0.75)
    init = dict(eta0=8, B=400, vlos=0.25, dopp=0.034, aa=0.06, gm=30, az=30, S0=0.15,
                S1=0.85, mac=1.2 if "mac" not in free else 0.8, alfa=0.9)
    s = orc.set_slight(setup(orc, psf=("gauss", 60.0), free=free, init=init, miter=200,
                             use_ce=1), L)
    assert orc.nterms(s) == (11 if "mac" in free else 10)
    obs = orc.synth(s, truth)[0].astype(np.float32)
    r = orc.lm_pixel(s, obs)
    assert r["chisqr"] < 1e-3
    np.testing.assert_allclose(r["model"][[1, 5, 6, 10]], truth[[1, 5, 6, 10]], rtol=2e-3,
                               atol=0.05)
    if "mac" in free:
        assert abs(r["model"][9] - 1.2) < 0.02
