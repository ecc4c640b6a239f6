"""Pins the oracle's step application, range checks and multi-line sums by
what the mathematics fixes, independently of how the oracle writes them:

* ``check`` (milosUtils.cu:160-238) reflects B, gamma and phi into range.
  Each reflection is a symmetry of the Milne-Eddington synthesis --
  (B, gamma) -> (-B, 180 - gamma) swaps the sigma components and flips
  cos(gamma); gamma -> -gamma and gamma -> 360 - gamma keep sin^2 and cos;
  phi -> phi +- 180 keeps (cos 2phi, sin 2phi) -- so a checked model must
  synthesise the same spectrum as the unchecked one and land in range.
* ``apply_delta`` (milosUtils.cu:93-159) subtracts the solved step from the
  free parameters with |dB| <= 300 G and |dgamma|, |dphi| <= 30 deg: a
  strongly damped Levenberg-Marquardt step then has to lower chi^2, and no
  step may move B, gamma, phi by more than those bounds.
* Several lines sharing one atmosphere (create_cuantic.c:17-26 holds the
  multi-line Cuantic layout; the reference's me_der.cu:53-77 keeps only one
  sub-component's Voigt derivative, the oracle and the HIP path sum them as
  the synthesis sums the profiles): central finite differences of every
  response on the 6301.5 + 6302.5 pair (5 pi and 4 + 4 sigma components on
  6301.5), a second line of zero strength reduces to the one-line result,
  and the order of the lines does not matter.
"""
import numpy as np
import pytest

from conftest import model_row

FE6173 = (6173.3356, (2, 1, 1.0, 2, 2, 0.0), 1.0)
FE6301 = (6301.515, (2, 1, 2.0, 2, 2, 2.0), 1.0)
FE6302 = (6302.494, (2, 1, 1.0, 2, 2, 0.0), 0.27)
GRID30 = np.arange(-350.0, 665.0 + 1e-9, 35.0)
GRID112 = -716.0 + 21.5 * np.arange(112)


def _model(**kw):
    d = dict(eta0=10, B=1200, vlos=0.5, dopp=0.03, aa=0.1, gm=60, az=40, S0=0.3, S1=0.6,
             mac=0, alfa=1)
    d.update(kw)
    return model_row(d)


def _rand_models(n, seed):
    rng = np.random.default_rng(seed)
    m = np.zeros((n, 11))
    m[:, 0] = rng.uniform(1.5, 20, n)
    m[:, 1] = rng.uniform(50, 3000, n)
    m[:, 2] = rng.uniform(-3, 3, n)
    m[:, 3] = rng.uniform(0.02, 0.06, n)
    m[:, 4] = rng.uniform(0.01, 0.5, n)
    m[:, 5] = rng.uniform(5, 175, n)
    m[:, 6] = rng.uniform(5, 175, n)
    m[:, 7] = rng.uniform(0.1, 0.4, n)
    m[:, 8] = rng.uniform(0.5, 0.9, n)
    m[:, 10] = 1.0
    return m


# ---- check: reflections are synthesis symmetries ------------------------------
def _setups(orc):
    return [orc.make_setup([FE6173], GRID30, psf=("gauss", 60.0)),
            orc.make_setup([FE6301, FE6302], GRID112)]


def _reflected_inputs(m, kind):
    """An out-of-range model the LM can produce (a step of at most 300 G /
    30 deg from an in-range one) that is physically the same atmosphere."""
    r = m.copy()
    if kind == "B<0":                     # (B, g) ~ (-B, 180 - g)
        r[1] = -min(m[1], 300.0)
        r[5] = 180.0 - m[5]
        m = m.copy()
        m[1] = -r[1]
    elif kind == "gamma<0":               # g ~ -g
        r[5] = -min(m[5], 30.0)
        m = m.copy()
        m[5] = -r[5]
    elif kind == "gamma>180":             # g ~ 360 - g
        r[5] = 180.0 + min(180.0 - m[5], 30.0)
        m = m.copy()
        m[5] = 360.0 - r[5]
    elif kind == "phi<0":                 # phi ~ phi + 180
        r[6] = -min(m[6], 30.0)
        m = m.copy()
        m[6] = r[6] + 180.0
    elif kind == "phi>180":               # phi ~ phi - 180
        r[6] = 180.0 + min(180.0 - m[6], 30.0)
        m = m.copy()
        m[6] = r[6] - 180.0
    elif kind == "B<0,gamma<0":           # (-B, 180 + x) ~ (B, -x) ~ (B, x)
        x = min(m[5], 30.0)
        r[1] = -m[1]
        r[5] = 180.0 + x
        m = m.copy()
        m[5] = x
    return r, m


KINDS = ("B<0", "gamma<0", "gamma>180", "phi<0", "phi>180", "B<0,gamma<0")


@pytest.mark.parametrize("kind", KINDS)
def test_check_reflections_preserve_the_synthesis(orc, kind):
    for s in _setups(orc):
        for m0 in _rand_models(24, seed=KINDS.index(kind)):
            raw, _ = _reflected_inputs(m0, kind)
            chk = orc.check(raw)
            assert 0.0 <= chk[1] <= 5000.0
            assert 0.0 <= chk[5] <= 180.0 and 0.0 <= chk[6] <= 180.0, (kind, raw, chk)
            a = orc.synth(s, raw)[0]
            b = orc.synth(s, chk)[0]
            ic = raw[7] + raw[8]
            assert np.abs(a - b).max() <= 1e-12 * ic, (kind, raw, chk, np.abs(a - b).max())


def test_check_reflection_targets(orc):
    """The checked model is the in-range representative of the same
    atmosphere: B >= 0, gamma and phi in [0, 180]."""
    for kind in KINDS:
        for m0 in _rand_models(16, seed=5):
            raw, expect = _reflected_inputs(m0, kind)
            chk = orc.check(raw)
            np.testing.assert_allclose(chk, expect, rtol=0, atol=1e-12)


def test_check_clamps_ranges(orc):
    """Values beyond the physical ranges are clamped, in-range values pass
    unchanged."""
    m = _model()
    np.testing.assert_array_equal(orc.check(m), m)
    big = _model(B=9000, eta0=5000, vlos=50, dopp=2.0, aa=30, S0=5, S1=5)
    small = _model(eta0=0.1, vlos=-50, dopp=0, aa=0, S0=-1, S1=-1)
    cb, cs = orc.check(big), orc.check(small)
    assert cb[1] == 5000 and cb[0] == 2500 and cb[2] == 20 and cb[3] == 0.6 and cb[4] == 10
    assert cb[7] == 2 and cb[8] == 2
    assert cs[0] == 1 and cs[2] == -20 and cs[3] == 1e-4 and cs[4] == 1e-4
    assert cs[7] == 1e-4 and cs[8] == 1e-4


# ---- apply_delta: bounded steps that descend -----------------------------------
def _obs(orc, s, truth, seed):
    rng = np.random.default_rng(seed)
    sp = orc.synth(s, truth)
    ic = (truth[:, 7] + truth[:, 8])[:, None, None]
    return (sp + rng.normal(size=sp.shape) * 1e-3 * ic).astype(np.float32)


def test_apply_delta_damped_step_descends(orc):
    """delta = (alpha (1 + lambda diag))^+ beta from the oracle's normal
    equations; model - delta with lambda large enough is a short gradient
    step and has to lower chi^2 for every pixel.  A wrong sign, a slot
    applied to the wrong field or a dropped field fails this."""
    s = orc.make_setup([FE6173], GRID30, psf=("gauss", 60.0))
    truth = _rand_models(32, seed=41)
    start = _rand_models(32, seed=42)
    obs = _obs(orc, s, truth, 43)
    nfree = 4 * 30 - 10
    for i in range(32):
        sp, rf = orc.synth_rf(s, start[i])
        A, b = orc.covarm(s, obs[i], sp[0], rf[0])
        c0 = orc.chisqr(s, obs[i], sp[0], nfree)
        for lam in (1e2, 1e3):
            Ad = A[:9, :9].copy()
            Ad[np.diag_indices(9)] *= 1.0 + lam
            x, _ = orc.svd_solve(Ad, b[:9])
            m1 = orc.check(orc.apply_delta(s, start[i], np.append(x, 0.0)))
            c1 = orc.chisqr(s, obs[i], orc.synth(s, m1)[0], nfree)
            assert c1 < c0, (i, lam, c0, c1)


def test_apply_delta_bounds_field_and_angle_steps(orc):
    """|dB| <= 300 G, |dgamma|, |dphi| <= 30 deg (milosUtils.cu:101-142); the
    other free parameters take the whole step; fixed ones none."""
    s = orc.make_setup([FE6173], GRID30)
    m = _model(B=2000, gm=90, az=90)
    rng = np.random.default_rng(3)
    for _ in range(200):
        d = rng.normal(size=10) * np.array([5, 2000, 2, 0.01, 0.1, 150, 150, 0.1, 0.1, 0])
        out = orc.apply_delta(s, m, d)
        step = m - out
        assert abs(step[1]) <= 300.0 and abs(step[5]) <= 30.0 and abs(step[6]) <= 30.0
        for k in (1, 5, 6):
            lim = 300.0 if k == 1 else 30.0
            want = d[k] if abs(d[k]) <= lim else np.sign(d[k]) * lim
            assert step[k] == pytest.approx(want, abs=1e-12)
        for k in (0, 2, 3, 4, 7, 8):
            assert step[k] == pytest.approx(d[k], abs=1e-12)
        assert step[9] == 0 and step[10] == 0
    fixed = orc.make_setup([FE6173], GRID30, free=("B", "gm"))
    out = orc.apply_delta(fixed, m, np.ones(10) * 5.0)
    np.testing.assert_array_equal(np.delete(out, [1, 5]), np.delete(m, [1, 5]))


# ---- multi-line: finite differences, zero-strength reduction, line order --------
@pytest.mark.parametrize("psf", [None, ("gauss", 60.0)])
def test_two_line_response_functions_finite_difference(orc, psf):
    """Exact-derivative mode (rf_reference_rules = 0) on 6301.5 + 6302.5:
    every response against central differences of the FP64 synthesis.  The
    floor (~1e-4) is Hui's rational approximation not obeying the Faddeeva
    derivative identities exactly; summing only one sub-component's
    derivative (the reference's me_der.cu:66-74) is off by O(1)."""
    s = orc.make_setup([FE6301, FE6302], GRID112, psf=psf, rf_reference_rules=0)
    for m in (_model(eta0=12, B=1500, vlos=0.7, dopp=0.035, aa=0.2, gm=50, az=30, S0=0.25,
                     S1=0.7), _model(B=300, gm=120, az=150, vlos=-1.2)):
        _, rf = orc.synth_rf(s, m)
        for p in range(9):
            h = max(abs(m[p]) * 1e-6, 1e-7)
            mp, mm = m.copy(), m.copy()
            mp[p] += h
            mm[p] -= h
            fd = (orc.synth(s, mp)[0] - orc.synth(s, mm)[0]) / (2 * h)
            err = np.abs(fd - rf[0][:, p, :]).max() / np.abs(fd).max()
            assert err < 1e-3, (psf, p, err)


@pytest.mark.parametrize("rules", [1, 0])
def test_zero_strength_second_line_reduces_to_one_line(orc, rules):
    one = orc.make_setup([FE6301], GRID112, psf=("gauss", 60.0), rf_reference_rules=rules)
    two = orc.make_setup([FE6301, (FE6302[0], FE6302[1], 0.0)], GRID112, psf=("gauss", 60.0),
                         rf_reference_rules=rules)
    for m in _rand_models(6, seed=8):
        sp1, rf1 = orc.synth_rf(one, m)
        sp2, rf2 = orc.synth_rf(two, m)
        np.testing.assert_array_equal(sp1, sp2)
        np.testing.assert_allclose(rf2, rf1, rtol=1e-13, atol=1e-13 * np.abs(rf1).max())


@pytest.mark.parametrize("rules", [1, 0])
def test_line_order_does_not_matter(orc, rules):
    a = orc.make_setup([FE6301, FE6302], GRID112, psf=("gauss", 60.0), rf_reference_rules=rules)
    b = orc.make_setup([FE6302, FE6301], GRID112, psf=("gauss", 60.0), rf_reference_rules=rules)
    for i in range(112):                          # the same absolute grid
        b.lambda_[i] = a.lambda_[i]
    for m in _rand_models(6, seed=9):
        spa, rfa = orc.synth_rf(a, m)
        spb, rfb = orc.synth_rf(b, m)
        np.testing.assert_allclose(spb, spa, rtol=0, atol=1e-13)
        np.testing.assert_allclose(rfb, rfa, rtol=1e-11, atol=1e-12 * np.abs(rfa).max())


def test_two_lines_sum_their_opacities(orc):
    """Lines far apart on one grid: each part of the spectrum is the one-line
    synthesis of the nearby line with its strength folded into eta0."""
    far = -400.0 + 20.0 * np.arange(160)          # 6301.1 ... 6304.3 A
    both = orc.make_setup([FE6301, FE6302], far)
    lo = orc.make_setup([FE6301], far)
    hi = orc.make_setup([FE6302[:2] + (1.0,)], far - (FE6302[0] - FE6301[0]) * 1000.0)
    for i in range(160):
        hi.lambda_[i] = both.lambda_[i]
    m = _model(dopp=0.025, aa=0.05, B=800)
    mh = m.copy()
    mh[0] = m[0] * FE6302[2]
    sp = orc.synth(both, m)[0]
    sl = orc.synth(lo, m)[0]
    sh = orc.synth(hi, mh)[0]
    lam = np.array(both.lambda_[:160])
    near_lo = np.abs(lam - FE6301[0]) < 0.25
    near_hi = np.abs(lam - FE6302[0]) < 0.25
    ic = m[7] + m[8]
    # the other line's far wing adds at most a few 1e-4 Ic (damping 0.05)
    assert np.abs(sp[:, near_lo] - sl[:, near_lo]).max() < 2e-3 * ic
    assert np.abs(sp[:, near_hi] - sh[:, near_hi]).max() < 2e-3 * ic
    # and a strong line is visibly present in each window
    assert (ic - sp[0, near_lo]).max() > 0.1 * ic and (ic - sp[0, near_hi]).max() > 0.05 * ic
