"""GPU parity across wavelength counts: the kernels choose their lane layout
by the number of wavelengths (32 lanes per pixel up to 32, 64 lanes up to 64,
then 2 and 4 wavelengths per lane up to 128 and 256), so the synthesis,
response functions, normal equations and whole inversions are held against
the FP64 oracle on each side of every boundary, at the reference's own
ceiling of 50 wavelengths (MAX_LAMBDA, definesCuda.cuh:125) and at the
engine's 256.  Fe I 6173.34 A on -350..665 mA sampled at n points, Gaussian
PSF of FWHM 60 mA on the grid (all n taps with psf_tail_tol 0 at n = 50).
Tolerances are those of test_gpu_parity."""
import numpy as np
import pytest

from test_gpu_parity import DP_TYP, check_inversion, to_dev
from workload import observations

pytestmark = pytest.mark.gpu

NLAMBDA = (31, 32, 33, 50, 64, 65, 128, 129, 256)


def _pair(gm, orc, n, **kw):
    grid = np.linspace(-350.0, 665.0, n)
    tail = kw.pop("psf_tail_tol", 0.0 if n == 50 else 1e-9)
    p = gm.make_setup([gm.FE6173], grid, psf=("gauss", 60.0), psf_tail_tol=tail, **kw)
    o = orc.make_setup([gm.FE6173], grid, psf=("gauss", 60.0), init_defaults=gm.BENCH_INIT,
                       free=gm.FREE9, **{k: v for k, v in kw.items() if k != "precision"})
    return p, o


@pytest.mark.parametrize("n", NLAMBDA)
def test_synthesis_and_responses_across_lambda_counts(gm, orc, cuda, n):
    p, o = _pair(gm, orc, n)
    ctx = gm.Context(p)
    models = gm.random_models(96, seed=40 + n)
    got = ctx.synth(to_dev(models, cuda)).cpu().numpy()
    ref = orc.synth(o, models)
    ic = (models[:, 7] + models[:, 8])[:, None, None]
    assert (np.abs(got - ref) <= 5e-6 * ic + 2e-5 * np.abs(ref)).all(), np.abs(got - ref).max()
    sp, rf = ctx.synth_rf(to_dev(models, cuda))
    rf = rf.cpu().numpy()
    sp_ref, rf_ref = orc.synth_rf(o, models)
    np.testing.assert_allclose(sp.cpu().numpy(), sp_ref, atol=2e-5)
    scale = np.abs(rf_ref).max(axis=-1, keepdims=True)
    step = np.array(DP_TYP + (1.0,))[None, None, :, None]
    tol = 1.25e-4 * scale + 5e-8 * ic[..., None] / step
    ratio = (np.abs(rf - rf_ref) / tol).max()
    print("n %d RF worst |d|/tol %.3g" % (n, ratio), flush=True)
    assert ratio <= 1.0
    assert np.all(rf[:, :, 9, :] == 0)


@pytest.mark.parametrize("n", NLAMBDA)
def test_normal_equations_across_lambda_counts(gm, orc, cuda, n):
    p, o = _pair(gm, orc, n)
    ctx = gm.Context(p)
    truth = gm.random_models(64, seed=60 + n)
    obs = observations(orc, o, truth, seed=60 + n)
    models = gm.random_models(64, seed=61 + n)
    a, b, c = ctx.normal_equations(to_dev(models, cuda), to_dev(obs, cuda))
    a, b, c = a.cpu().numpy(), b.cpu().numpy(), c.cpu().numpy()
    sp, rf = orc.synth_rf(o, models)
    for i in range(len(models)):
        A, B = orc.covarm(o, obs[i], sp[i], rf[i])
        assert np.abs(gm.unpack_alpha(a[i]) - A[:9, :9]).max() <= 1e-3 * np.abs(A).max()
        assert np.abs(b[i] - B[:9]).max() <= 2e-3 * np.abs(B).max()
        assert c[i] == pytest.approx(orc.chisqr(o, obs[i], sp[i]), rel=2e-5)


@pytest.mark.parametrize("n", NLAMBDA)
def test_inversion_across_lambda_counts(gm, orc, cuda, n):
    p, o = _pair(gm, orc, n)
    ctx = gm.Context(p)
    truth = gm.random_models(512, seed=80 + n)
    obs = observations(orc, o, truth, seed=80 + n)
    out = ctx.invert(to_dev(obs, cuda))
    got = {k: v.cpu().numpy() for k, v in out.items()}
    ref = orc.invert(o, obs)
    assert np.all(np.isfinite(got["model"])) and np.all(got["iters"] <= 50)
    check_inversion(orc, o, obs, got, ref, truth, "n %d" % n)
    rel_c2 = np.abs(got["chisqr"] - ref["chisqr"]) / ref["chisqr"]
    assert np.median(rel_c2) < 1e-4
