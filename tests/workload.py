"""Synthetic observations for the parity tests: seeded true models, FP64
oracle synthesis, Gaussian noise of sigma = 1e-3 Ic (Ic = S0 + S1)."""
import numpy as np


def observations(orc, oracle_setup, models, sigma=1e-3, seed=0, high_fraction=0.0):
    clean = orc.synth_batch(oracle_setup, models)
    rng = np.random.default_rng(seed + 101)
    ic = models[:, 7] + models[:, 8]
    sig = np.full(models.shape[0], sigma)
    if high_fraction > 0:
        sig[rng.random(models.shape[0]) < high_fraction] = 3e-3
    noise = rng.normal(size=clean.shape) * (sig * ic)[:, None, None]
    return np.ascontiguousarray((clean + noise).astype(np.float32))


def param_errors(a, b):
    """Per-parameter |a - b| with the azimuth difference taken modulo 180."""
    d = np.abs(a - b)
    d[:, 6] = np.minimum(d[:, 6], 180.0 - d[:, 6])
    return d
