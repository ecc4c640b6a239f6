"""Pins the numpy restatement of the learned-guess network that the GPU
gradient tests compare against: its analytic backward pass against central
finite differences of its own forward pass (FP64, rounding off), its BF16
rounding against torch's, and its target map against the decoder."""
import numpy as np
import torch

import mlp_reference as R
from conftest import load_product


def _random_case(nl=30, b=16, seed=0):
    rng = np.random.default_rng(seed)
    _, n = R.layout(nl)
    params = np.zeros(n)
    for i, o, wo, bo in R.layout(nl)[0]:
        params[wo:wo + i * o] = rng.normal(size=i * o) * np.sqrt(2.0 / i)
        params[bo:bo + o] = rng.normal(size=o) * 0.1
    obs = np.empty((b, 4, nl))
    obs[:, 0] = 1.0 + 0.2 * rng.normal(size=(b, nl))
    obs[:, 1:] = 0.02 * rng.normal(size=(b, 3, nl))
    models = load_product().random_models(b, seed=seed)
    return params, obs, models


def test_backward_matches_finite_differences():
    nl = 30
    params, obs, models = _random_case(nl)
    loss, g = R.loss_and_grads(params, nl, obs, models, rounding=False)
    rng = np.random.default_rng(1)
    for i, o, wo, bo in R.layout(nl)[0]:
        idx = list(rng.integers(wo, wo + min(i, 4 * nl) * o, 6)) + list(rng.integers(bo, bo + o, 3))
        for k in idx:
            h = 1e-5
            p1, p2 = params.copy(), params.copy()
            p1[k] += h
            p2[k] -= h
            fd = (R.loss_and_grads(p1, nl, obs, models, False)[0] -
                  R.loss_and_grads(p2, nl, obs, models, False)[0]) / (2 * h)
            assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (k, fd, g[k])


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(size=4096) * 10.0 ** rng.uniform(-8, 8, 4096),
                        [0.0, -0.0, 1.0, 1.00390625, 1.005859375, 65504.0]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(R.bf16(x), ref)


def test_decode_inverts_targets():
    models = load_product().random_models(512, seed=4)
    models[:, 6] = np.clip(models[:, 6], 0.5, 179.5)
    back = R.decode(R.prep_target(models), base=np.zeros(11))
    for f in range(9):
        np.testing.assert_allclose(back[:, f], models[:, f], rtol=1e-5, atol=1e-3 if f == 1 else 1e-5)
