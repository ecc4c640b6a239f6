"""GPU parity of the macroturbulence / stray-light kernels (the reference's
11-term physics with mac and alfa fixed: mil_sinrf.cu:101-140,
me_der.cu:119-211) against the FP64 oracle, with the tolerances of
test_gpu_parity.py, and the 9-parameter contexts' contract that per-pixel
mac / alfa are ignored there."""
import numpy as np
import pytest

from test_gpu_parity import DP_TYP, _assert_medians, pair, to_dev
from workload import observations, param_errors

pytestmark = pytest.mark.gpu

GAUSS60 = ("gauss", 60.0)


def stray_profile(n=30, seed=3):
    """A quiet-Sun-like stray-light profile: a broad shallow line in I,
    small Q, U, V."""
    rng = np.random.default_rng(seed)
    x = np.linspace(-1.0, 1.9, n)
    prof = np.zeros((4, n))
    prof[0] = 0.95 - 0.35 * np.exp(-((x - 0.0) / 0.35) ** 2)
    prof[1:] = rng.normal(0, 2e-3, size=(3, n))
    return prof


def ext_pair(gm, orc, psf, mac=1.0, alfa=1.0, slight=None, **kw):
    p, o = pair(gm, orc, "single", psf, init=dict(mac=mac, alfa=alfa), **kw)
    if slight is not None:
        prof = np.asarray(slight, dtype=np.float32).reshape(-1)
        p.use_slight = 1
        for k, v in enumerate(prof):
            p.slight[k] = float(v)
        orc.set_slight(o, slight)
    return p, o


def mac_models(gm, n, seed, slight=False):
    m = gm.random_models(n, seed=seed)
    rng = np.random.default_rng(seed + 1)
    m[:, 9] = rng.uniform(0.0, 3.0, n)
    m[: n // 4, 9] = rng.choice([0.0, 5e-5, 1e-4], n // 4)     # the 1e-4 switch
    if slight:
        m[:, 10] = rng.uniform(0.3, 1.0, n)
    return m


@pytest.mark.parametrize("psf", [None, GAUSS60])
@pytest.mark.parametrize("with_slight", [False, True])
def test_ext_synthesis_matches_oracle(gm, orc, cuda, psf, with_slight):
    L = stray_profile() if with_slight else None
    p, o = ext_pair(gm, orc, psf, slight=L)
    models = mac_models(gm, 256, 31, slight=with_slight)
    ctx = gm.Context(p)
    got = ctx.synth(to_dev(models, cuda)).cpu().numpy()
    ref = orc.synth(o, models)
    ic = (models[:, 7] + models[:, 8])[:, None, None]
    tol = 5e-6 * ic + 2e-5 * np.abs(ref)
    assert (np.abs(got - ref) <= tol).all(), np.abs(got - ref).max()


@pytest.mark.parametrize("psf", [None, GAUSS60])
@pytest.mark.parametrize("with_slight", [False, True])
def test_ext_response_functions_match_oracle(gm, orc, cuda, psf, with_slight):
    L = stray_profile() if with_slight else None
    p, o = ext_pair(gm, orc, psf, slight=L, rf_reference_rules=0)
    models = mac_models(gm, 128, 32, slight=with_slight)
    ctx = gm.Context(p)
    sp, rf = ctx.synth_rf(to_dev(models, cuda))
    rf = rf.cpu().numpy()
    sp_ref, rf_ref = orc.synth_rf(o, models)
    np.testing.assert_allclose(sp.cpu().numpy(), sp_ref, atol=2e-5)
    scale = np.abs(rf_ref).max(axis=-1, keepdims=True)
    ic = (models[:, 7] + models[:, 8])[:, None, None, None]
    step = np.array(DP_TYP + (1.0,))[None, None, :, None]
    tol = np.broadcast_to(5e-4 * scale + 2e-7 * ic / step, rf.shape)
    bad = np.abs(rf - rf_ref) > tol
    if bad.any():
        for i in np.argwhere(bad)[:8]:
            i = tuple(i)
            print("mismatch", i, rf[i], rf_ref[i], tol[i], "model", models[i[0]].round(4))
    assert not bad.any(), np.argwhere(bad)[:5]
    assert np.all(rf[:, :, 9, :] == 0)          # filling factor fixed: slot 9 zeroed


def test_ext_normal_equations_match_oracle(gm, orc, cuda):
    L = stray_profile()
    p, o = ext_pair(gm, orc, GAUSS60, slight=L)
    truth = mac_models(gm, 128, 33, slight=True)
    obs = observations(orc, o, truth, seed=33)
    models = mac_models(gm, 128, 34, slight=True)
    ctx = gm.Context(p)
    a, b, c = ctx.normal_equations(to_dev(models, cuda), to_dev(obs, cuda))
    a, b, c = a.cpu().numpy(), b.cpu().numpy(), c.cpu().numpy()
    sp, rf = orc.synth_rf(o, models)
    for i in range(len(models)):
        A, B = orc.covarm(o, obs[i], sp[i], rf[i])
        A9 = gm.unpack_alpha(a[i])
        assert np.abs(A9 - A[:9, :9]).max() <= 1e-3 * np.abs(A).max()
        assert np.abs(b[i] - B[:9]).max() <= 2e-3 * np.abs(B).max()
        assert c[i] == pytest.approx(orc.chisqr(o, obs[i], sp[i]), rel=2e-5)


@pytest.mark.parametrize("mac,alfa,with_slight,tail", [(1.2, 1.0, False, 1e-9), (0.8, 0.75, True, 1e-9),
                                                       (0.8, 0.75, True, 0.0)])
def test_ext_inversion_matches_oracle(gm, orc, cuda, mac, alfa, with_slight, tail):
    """Fixed macroturbulence (the reference's 10-term run with mac fixed) and
    a fixed filling factor with stray light, inverted from the classical
    estimates: medians of |d| against the oracle's inversion, which applies
    every PSF and mac-kernel tap (tail 0: the product does too; 1e-9: it
    skips the taps of negligible mass)."""
    L = stray_profile() if with_slight else None
    p, o = ext_pair(gm, orc, GAUSS60, mac=mac, alfa=alfa, slight=L, psf_tail_tol=tail)
    truth = gm.random_models(512, seed=35)
    truth[:, 9] = mac
    truth[:, 10] = alfa
    obs = observations(orc, o, truth, seed=35)
    ctx = gm.Context(p)
    got = {k: v.cpu().numpy() for k, v in ctx.invert(to_dev(obs, cuda)).items()}
    ref = orc.invert(o, obs)
    d = param_errors(got["model"], ref["model"])
    _assert_medians(d, "mac %.1f alfa %.2f" % (mac, alfa))
    assert np.all(got["model"][:, 9] == mac) and np.all(got["model"][:, 10] == alfa)
    # accuracy against the truth is that of the plain path
    dt = param_errors(got["model"], truth)
    assert np.median(dt[:, 1]) < 20.0


def test_plain_context_ignores_per_pixel_mac_alfa(gm, orc, cuda):
    """A 9-parameter context (no macroturbulence, no stray light) keeps the
    set-up's mac and alfa for per-pixel start models."""
    p, o = pair(gm, orc, "single", GAUSS60, init_mode=gm.GM_INIT_PER_PIXEL)
    truth = gm.random_models(256, seed=36)
    obs = observations(orc, o, truth, seed=36)
    start = truth.copy()
    start[:, 1] += 100.0
    ctx = gm.Context(p)
    a = {k: v.cpu().numpy() for k, v in ctx.invert(to_dev(obs, cuda), init=to_dev(start, cuda)).items()}
    start[:, 9] = 2.0
    start[:, 10] = 0.5
    b = {k: v.cpu().numpy() for k, v in ctx.invert(to_dev(obs, cuda), init=to_dev(start, cuda)).items()}
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])
    assert np.all(b["model"][:, 9] == 0.0) and np.all(b["model"][:, 10] == 1.0)


# ---- free filling factor: the 10-term (lm_mils) and 11-term (lm_mils_11) runs
FREE10 = ("eta0", "B", "vlos", "dopp", "aa", "gm", "az", "S0", "S1", "alfa")
FREE11 = FREE10 + ("mac",)


def ff_models(gm, n, seed):
    m = gm.random_models(n, seed=seed)
    rng = np.random.default_rng(seed + 2)
    m[:, 9] = rng.uniform(0.3, 2.5, n)
    m[:, 10] = rng.uniform(0.4, 1.0, n)
    return m


@pytest.mark.parametrize("free", [FREE10, FREE11])
def test_free_filling_factor_responses_match_oracle(gm, orc, cuda, free):
    """Reference layout [4][NTERMS][n]: NTERMS 10 keeps d/dalfa in slot 9,
    11 has alfa-scaled d/dmac in 9 and d/dalfa in 10 (me_der.cu:191-207)."""
    L = stray_profile()
    p, o = ext_pair(gm, orc, GAUSS60, slight=L, free=free, rf_reference_rules=0)
    nt = 11 if "mac" in free else 10
    assert orc.nterms(o) == nt
    models = ff_models(gm, 128, 41)
    ctx = gm.Context(p)
    sp, rf = ctx.synth_rf(to_dev(models, cuda))
    rf = rf.cpu().numpy()
    assert rf.shape == (128, 4, nt, 30)
    sp_ref, rf_ref = orc.synth_rf(o, models)
    np.testing.assert_allclose(sp.cpu().numpy(), sp_ref, atol=2e-5)
    scale = np.abs(rf_ref).max(axis=-1, keepdims=True)
    ic = (models[:, 7] + models[:, 8])[:, None, None, None]
    step = np.array(DP_TYP + (0.01,) * (nt - 9))[None, None, :, None]
    tol = np.broadcast_to(5e-4 * scale + 2e-7 * ic / step, rf.shape)
    bad = np.abs(rf - rf_ref) > tol
    if bad.any():
        for i in np.argwhere(bad)[:8]:
            i = tuple(i)
            print("mismatch", i, rf[i], rf_ref[i], tol[i])
    assert not bad.any(), np.argwhere(bad)[:5]


@pytest.mark.parametrize("free", [FREE10, FREE11])
def test_free_filling_factor_normal_equations_match_oracle(gm, orc, cuda, free):
    L = stray_profile()
    p, o = ext_pair(gm, orc, GAUSS60, slight=L, free=free)
    nt = orc.nterms(o)
    truth = ff_models(gm, 96, 42)
    obs = observations(orc, o, truth, seed=42)
    models = ff_models(gm, 96, 43)
    ctx = gm.Context(p)
    a, b, c = ctx.normal_equations(to_dev(models, cuda), to_dev(obs, cuda))
    a, b, c = a.cpu().numpy(), b.cpu().numpy(), c.cpu().numpy()
    assert a.shape == (96, 66) and b.shape == (96, 11)
    sp, rf = orc.synth_rf(o, models)
    # oracle slot -> eleven-slot index (NTERMS 10: alfa in slot 9 -> 10)
    idx = list(range(9)) + ([9, 10] if nt == 11 else [10])
    for i in range(len(models)):
        A, B = orc.covarm(o, obs[i], sp[i], rf[i])
        A11 = gm.unpack_alpha(a[i])
        assert np.abs(A11[np.ix_(idx, idx)] - A).max() <= 1e-3 * np.abs(A).max()
        assert np.abs(b[i][idx] - B).max() <= 2e-3 * np.abs(B).max()
        if nt == 10:
            assert np.all(A11[9] == 0) and np.all(b[i][9] == 0)
        assert c[i] == pytest.approx(orc.chisqr(o, obs[i], sp[i]), rel=2e-5)


def test_ten_term_cordic_step_solve_matches_oracle(gm, orc, cuda):
    """Filling factor free, mac fixed, solver 3: the 11-slot system (slot 9,
    mac, a zero row) solved as the reference's 10-term one -- slots 0-8, then
    alfa in slot 9 -- by the CORDIC solve; mac's step is zero."""
    L = stray_profile()
    p, _ = ext_pair(gm, orc, None, slight=L, free=FREE10, solver=gm.GM_SOLVE_CORDIC)
    ctx = gm.Context(p)
    rng = np.random.default_rng(46)
    n = 128
    packed = np.zeros((n, 66), dtype=np.float32)
    lam = 10.0 ** rng.uniform(-4, 2, n)
    for i in range(n):
        J = rng.normal(size=(120, 11)) * rng.uniform(0.01, 10, 11)
        M = J.T @ J
        M[9, :] = M[:, 9] = 0.0
        for q in range(11):
            for r in range(q, 11):
                packed[i, gm.packed_index(q, r, 11)] = M[q, r]
    b = (rng.normal(size=(n, 11)) * 10).astype(np.float32)
    b[:, 9] = 0.0
    d = ctx.step_solve(to_dev(packed, cuda), to_dev(b, cuda), to_dev(lam, cuda)).cpu().numpy()
    keep = [0, 1, 2, 3, 4, 5, 6, 7, 8, 10]
    for i in range(n):
        A = gm.unpack_alpha(packed[i]).astype(np.float64)
        A[np.diag_indices(11)] *= 1.0 + lam[i]
        x10, _ = orc.cordic_solve10(A[np.ix_(keep, keep)], b[i, keep].astype(np.float64))
        x = np.zeros(11)
        x[keep] = x10
        assert d[i, 9] == 0.0
        assert np.abs(d[i] - x).max() <= 2e-2 * np.abs(x).max() + 1e-12, (i, d[i], x)


def test_eleven_slot_step_solve_matches_oracle(gm, orc, cuda):
    """FP64 LDL^T / Jacobi pseudo-inverse of 11x11 systems: well-conditioned
    damped ones, zero rows, and exactly singular integer ones (A = J^T J / 64
    with dependent columns of small-integer J, exact in float32)."""
    L = stray_profile()
    p, _ = ext_pair(gm, orc, None, slight=L, free=FREE11)
    ctx = gm.Context(p)
    rng = np.random.default_rng(44)
    n = 192
    packed = np.zeros((n, 66), dtype=np.float32)
    lam = np.zeros(n)
    for i in range(n):
        if i % 2 == 0:
            J = rng.normal(size=(120, 11)) * rng.uniform(0.01, 10, 11)
            M = J.T @ J
            if i % 6 == 0:
                k = (i // 6) % 11
                M[k, :] = M[:, k] = 0.0
            lam[i] = 10.0 ** rng.uniform(-4, 2)
        else:
            J = rng.integers(-3, 4, size=(24, 11)).astype(np.float64)
            J[:, 10] = J[:, 0] + J[:, 3]
            if i % 4 == 1:
                J[:, 9] = J[:, 1] - J[:, 2]
            J = J[:, rng.permutation(11)]
            M = J.T @ J / 64.0
        for q in range(11):
            for r in range(q, 11):
                packed[i, gm.packed_index(q, r, 11)] = M[q, r]
    b = (rng.normal(size=(n, 11)) * 10).astype(np.float32)
    d = ctx.step_solve(to_dev(packed, cuda), to_dev(b, cuda), to_dev(lam, cuda)).cpu().numpy()
    for i in range(n):
        A = gm.unpack_alpha(packed[i]).astype(np.float64)
        A[np.diag_indices(11)] *= 1.0 + lam[i]
        x, _ = orc.svd_solve(A, b[i].astype(np.float64))
        if i % 2:
            xp = np.linalg.pinv(A, rcond=1e-10) @ b[i].astype(np.float64)
            assert np.abs(x - xp).max() <= 1e-8 * np.abs(xp).max()
        assert np.abs(d[i] - x).max() <= 1e-6 * np.abs(x).max() + 1e-12, (i, d[i], x)


@pytest.mark.parametrize("free,solver", [(FREE10, 0), (FREE11, 0), (FREE10, 3), (FREE11, 3)])
def test_free_filling_factor_inversion_matches_oracle(gm, orc, cuda, free, solver):
    """Inversions with the filling factor (and mac) free against the oracle's
    10-/11-term LM: medians of |d| for B, gamma, phi, vlos and alfa.  Under
    solver 3 the 10-term run takes the reference's CORDIC solve (the 11-term
    one keeps its FP64 SVD, as lm_mils_11)."""
    L = stray_profile()
    p, o = ext_pair(gm, orc, GAUSS60, mac=1.0, alfa=0.9, slight=L, free=free, solver=solver)
    truth = ff_models(gm, 384, 45)
    if "mac" not in free:
        truth[:, 9] = 1.0
    obs = observations(orc, o, truth, seed=45)
    ctx = gm.Context(p)
    got = {k: v.cpu().numpy() for k, v in ctx.invert(to_dev(obs, cuda)).items()}
    ref = orc.invert(o, obs)
    d = param_errors(got["model"], ref["model"])
    _assert_medians(d, "free " + ",".join(free[9:]))
    print("median |d alfa| %.3g, |d mac| %.3g" % (np.median(d[:, 10]), np.median(d[:, 9])))
    assert np.median(d[:, 10]) <= 1e-3
    if "mac" in free:
        assert np.median(d[:, 9]) <= 1e-2
    else:
        assert np.all(got["model"][:, 9] == 1.0)
    assert np.median(np.abs(got["chisqr"] / ref["chisqr"] - 1)) < 1e-3


def test_stray_light_map_matches_oracle(gm, orc, cuda):
    """Per-pixel stray light (mapStrayLight, gmilos.cu:1212-1226): synthesis
    and an inversion with the filling factor free, each pixel against the
    oracle with that pixel's profile; calls beyond the map are refused."""
    import torch
    L = stray_profile()
    p, o = ext_pair(gm, orc, GAUSS60, mac=0.6, alfa=0.9, slight=L, free=FREE10)
    npix = 48
    rng = np.random.default_rng(46)
    maps = (L[None] * rng.uniform(0.8, 1.2, size=(npix, 1, 1)) +
            rng.normal(0, 1e-3, size=(npix, 4, 30))).astype(np.float32)
    truth = ff_models(gm, npix, 47)
    truth[:, 9] = 0.6
    ctx = gm.Context(p)
    dmap = to_dev(maps, cuda)
    ctx.set_stray_map(dmap)
    got_sp = ctx.synth(to_dev(truth, cuda)).cpu().numpy()
    obs = np.zeros((npix, 4, 30), dtype=np.float32)
    noise = np.random.default_rng(48).normal(size=(npix, 4, 30))
    for i in range(npix):
        orc.set_slight(o, maps[i])
        ref = orc.synth(o, truth[i])[0]
        ic = truth[i, 7] + truth[i, 8]
        assert np.abs(got_sp[i] - ref).max() <= 5e-6 * ic + 2e-5 * np.abs(ref).max()
        obs[i] = (ref + 1e-3 * ic * noise[i]).astype(np.float32)
    got = {k: v.cpu().numpy() for k, v in ctx.invert(to_dev(obs, cuda)).items()}
    d = np.zeros((npix, 11))
    for i in range(npix):
        orc.set_slight(o, maps[i])
        r = orc.lm_pixel(o, obs[i])
        d[i] = param_errors(got["model"][i:i + 1], r["model"][None])[0]
    _assert_medians(d, "stray-light map")
    assert np.median(d[:, 10]) <= 1e-3
    with pytest.raises(RuntimeError):
        ctx.invert(to_dev(np.concatenate([obs, obs[:1]]), cuda))
    ctx.set_stray_map(None)
    back = ctx.synth(to_dev(truth, cuda)).cpu().numpy()
    orc.set_slight(o, L)
    np.testing.assert_allclose(back, orc.synth(o, truth), atol=2e-5)
