"""Learned initial guess on the GPU: BF16 MFMA forward/backward against the
numpy restatement (tests/mlp_reference.py, itself pinned by finite
differences), prediction against the reference forward + decode, and
training that lowers the loss on engine-synthesised profiles."""
import numpy as np
import pytest

import mlp_reference as R
from test_gpu_parity import to_dev

pytestmark = pytest.mark.gpu


def _batch(gm, cuda, n, seed, nl=30):
    import torch
    setup = gm.setup_single_line(psf_fwhm=60.0) if nl == 30 else gm.setup_two_lines(psf_fwhm=60.0)
    assert setup.nlambda == nl
    ctx = gm.Context(setup)
    models = gm.random_models(n, seed=seed)
    obs = ctx.synth(torch.from_numpy(models).to(cuda)).contiguous()
    return obs, models


def test_mlp_gradients_match_reference(gm, cuda):
    import torch
    obs, models = _batch(gm, cuda, 128, 31)
    net = gm.MLP(30, seed=5)
    p = net.params()
    rng = np.random.default_rng(2)
    for i, o, wo, bo in net.layout():
        p[bo:bo + o] = rng.normal(size=o) * 0.05        # nonzero biases
    net.set_params(p)
    loss, g = net.gradients(obs, torch.from_numpy(models).to(cuda))
    rloss, rg = R.loss_and_grads(p, 30, obs.cpu().numpy(), models, rounding=True)
    print("loss gpu %.6g ref %.6g" % (loss, rloss), flush=True)
    assert abs(loss - rloss) <= 1e-4 * rloss
    for l, (i, o, wo, bo) in enumerate(net.layout()):
        for name, sl in (("W", slice(wo, wo + i * o)), ("b", slice(bo, bo + o))):
            ref = rg[sl]
            err = np.linalg.norm(g[sl] - ref) / max(np.linalg.norm(ref), 1e-30)
            print("layer %d %s rel err %.3g" % (l, name, err), flush=True)
            assert err < 5e-3, (l, name, err)


@pytest.mark.parametrize("nl,npix", [(30, 320), (30, 1000), (112, 300)])
def test_mlp_predict_matches_reference(gm, cuda, nl, npix):
    """30 wavelengths (input width 128): the fused one-kernel path, on whole
    and ragged 64-pixel tiles; 112 wavelengths (input width 448): the
    layered GEMM path."""
    obs, models = _batch(gm, cuda, npix, 32, nl)
    net = gm.MLP(nl, seed=6)
    p = net.params()
    rng = np.random.default_rng(3)
    for i, o, wo, bo in net.layout():
        p[bo:bo + o] = rng.normal(size=o) * 0.05        # nonzero biases
    net.set_params(p)
    base = np.array([gm.BENCH_INIT[k] for k in ("eta0", "B", "vlos", "dopp", "aa", "gm", "az",
                                                "S0", "S1", "mac", "alfa")])
    got = net.predict(obs, base).cpu().numpy()
    layers = R.unpack(net.params(), nl)
    kin = net.layout()[0][0]
    y, _, _ = R.forward(layers, R.prep_input(obs.cpu().numpy(), kin), rounding=True)
    ref = R.decode(y, base)
    scale = np.array([5.485, 866, 1.732, 0.01155, 0.1414, 5, 5, 0.02, 0.02, 1, 1])
    err = np.abs(got - ref) / scale
    print("max scaled err", err.max(axis=0), flush=True)
    assert np.median(err) < 1e-3 and np.percentile(err, 99) < 2e-2


def test_mlp_fused_predict_equals_layered(gm, cuda):
    """The one-kernel predict (16-byte aligned input) and the layered GEMM
    predict (input moved off 16-byte alignment) round the same products in
    the same order: equal bit for bit."""
    import torch
    obs, _ = _batch(gm, cuda, 777, 34)
    net = gm.MLP(30, seed=8)
    p = net.params()
    rng = np.random.default_rng(4)
    for i, o, wo, bo in net.layout():
        p[bo:bo + o] = rng.normal(size=o) * 0.05
    net.set_params(p)
    base = np.array([gm.BENCH_INIT[k] for k in ("eta0", "B", "vlos", "dopp", "aa", "gm", "az",
                                                "S0", "S1", "mac", "alfa")])
    flat = torch.empty(obs.numel() + 1, dtype=torch.float32, device=obs.device)
    shifted = flat[1:].view(obs.shape)
    shifted.copy_(obs)
    assert shifted.data_ptr() % 16 != 0 and obs.data_ptr() % 16 == 0
    fused = net.predict(obs, base).cpu().numpy()
    layered = net.predict(shifted, base).cpu().numpy()
    assert np.array_equal(fused, layered), np.abs(fused - layered).max(axis=0)


def test_mlp_training_lowers_loss(gm, cuda):
    import torch
    obs, models = _batch(gm, cuda, 8192, 33)
    tm = torch.from_numpy(models).to(cuda)
    net = gm.MLP(30, seed=7)
    losses = []
    for it in range(300):
        j = (it * 1024) % 8192
        losses.append(net.train_step(obs[j:j + 1024], tm[j:j + 1024], lr=1e-3))
    print("loss first %.4g last %.4g" % (np.mean(losses[:8]), np.mean(losses[-8:])), flush=True)
    assert np.all(np.isfinite(losses))
    assert np.mean(losses[-8:]) < 0.3 * np.mean(losses[:8])
