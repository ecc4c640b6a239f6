"""The reference's 10-term step solve (mil_svdf over svdcordicf: 18 steps of
a Brent-Luk CORDIC Jacobi, not run to convergence) as restated by the
oracle, pinned by

* an independent matrix-form statement of one step (A <- P (G^T A G) P^T
  with the diagonal blocks' off-diagonal entry zeroed, V <- P (V G) P^T, P
  the permutation i -> CP[i]) against the oracle's element-wise code, at
  every step count from 1 to 27;
* invariants: V orthogonal, V A_k V^T = P^k A P^-k (so A itself after every
  9 steps, the permutation being a 9-cycle), a diagonal input left alone;
* convergence: run long, the steps reach the eigen-decomposition
  (numpy.linalg.eigh) and the solve reaches numpy.linalg.pinv;
* the LM loop: with the CORDIC solve a noise-free profile is still inverted
  back to its model."""
import numpy as np
import pytest

import oracle as O

CP = [0, 2, 4, 1, 6, 3, 8, 5, 9, 7]


def _perm():
    P = np.zeros((10, 10))
    for i in range(10):
        P[CP[i], i] = 1.0
    return P


def _angle(a11, a12, a22):
    if a11 == a22:
        return np.pi / 4
    return float(np.clip(np.arctan(2 * a12 / (a22 - a11)) / 2, -np.pi / 4, np.pi / 4))


def _matrix_form(A, steps):
    P = _perm()
    a = A.copy()
    V = np.eye(10)
    for _ in range(steps):
        G = np.zeros((10, 10))
        for k in range(5):
            th = _angle(a[2 * k, 2 * k], a[2 * k, 2 * k + 1], a[2 * k + 1, 2 * k + 1])
            c, s = np.cos(th), np.sin(th)
            G[2 * k:2 * k + 2, 2 * k:2 * k + 2] = [[c, s], [-s, c]]
        M = G.T @ a @ G
        for k in range(5):
            M[2 * k, 2 * k + 1] = M[2 * k + 1, 2 * k] = 0.0
        a = P @ M @ P.T
        V = P @ (V @ G) @ P.T
    return a, V


def _systems(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        J = rng.normal(size=(120, 10)) * rng.uniform(0.01, 10, 10)
        out.append(J.T @ J)
    return out


@pytest.fixture(scope="module", autouse=True)
def _lib():
    O.lib()


def test_cordic_steps_match_matrix_form():
    for A in _systems(4, 1):
        for steps in range(1, 28):
            a, v = O.cordic_eig10(A, steps)
            ar, vr = _matrix_form(A, steps)
            sc = np.abs(A).max()
            assert np.abs(a - ar).max() <= 1e-12 * sc, steps
            assert np.abs(v - vr).max() <= 1e-12, steps


def test_cordic_invariants():
    P = _perm()
    for A in _systems(4, 2):
        sc = np.abs(A).max()
        for steps in (1, 4, 9, 13, 18, 27):
            a, v = O.cordic_eig10(A, steps)
            assert np.abs(v @ v.T - np.eye(10)).max() <= 1e-13
            Pk = np.linalg.matrix_power(P, steps)
            assert np.abs(v @ a @ v.T - Pk @ A @ Pk.T).max() <= 1e-12 * sc
            if steps % 9 == 0:
                assert np.abs(v @ a @ v.T - A).max() <= 1e-12 * sc
    # a diagonal matrix with distinct entries is already diagonal: every
    # angle is zero and 18 steps hand it back with V = I
    D = np.diag(np.arange(1.0, 11.0))
    a, v = O.cordic_eig10(D, 18)
    assert np.array_equal(a, D) and np.array_equal(v, np.eye(10))
    x, w = O.cordic_solve10(D, np.ones(10))
    assert np.allclose(x, 1.0 / np.arange(1.0, 11.0), rtol=1e-15)


def test_cordic_converges_to_eigh_and_pinv():
    for A in _systems(6, 3):
        a, v = O.cordic_eig10(A, 9 * 12)
        off = a - np.diag(np.diag(a))
        assert np.abs(off).max() <= 1e-10 * np.abs(A).max()
        assert np.allclose(np.sort(np.diag(a)), np.linalg.eigh(A)[0], rtol=1e-10,
                           atol=1e-12 * np.abs(A).max())


def test_cordic_solve_truncation_and_zero_row():
    """18 steps are two sweeps: on well-conditioned LM-like systems the step
    is close to, but not exactly, the pseudo-inverse step; a zero row (a
    fixed parameter or the unused tenth slot) gets a zero step."""
    rng = np.random.default_rng(4)
    errs = []
    for A in _systems(32, 5):
        A = A.copy()
        A[np.diag_indices(10)] *= 1.1          # Marquardt damping
        b = rng.normal(size=10)
        x, _ = O.cordic_solve10(A, b)
        xr = np.linalg.pinv(A) @ b
        errs.append(np.linalg.norm(x - xr) / np.linalg.norm(xr))
        A[9, :] = A[:, 9] = 0.0
        b[9] = 0.0
        x, w = O.cordic_solve10(A, b)
        assert x[9] == 0.0
        assert np.isfinite(x).all()
    errs = np.array(errs)
    print("18-step CORDIC vs pinv: median rel %.2e, max %.2e" % (np.median(errs), errs.max()))
    assert np.median(errs) > 0.0               # not converged: a different answer
    assert np.median(errs) < 0.5               # but a descent-quality step


def test_lm_with_cordic_closed_loop_and_reference_example():
    """With the CORDIC solve the LM still inverts a noise-free profile back to
    its model, and the reference's own invert example (data/synthesis_gpu.per
    from run/guess.mod, CRISP PSF, README.md:76-89) back to run/init.mod."""
    import os
    from conftest import GOLDEN, crisp_table, model_row, read_mod
    from test_oracle_lm import FE6173, FREE9, GRID, _m, setup
    s = setup(O, psf=("gauss", 60.0), miter=200, solver=O.SOLVE_CORDIC)
    m = _m(eta0=9.0, B=1400.0, vlos=0.6, dopp=0.033, aa=0.12, gm=70.0, az=35.0, S0=0.3,
           S1=0.6)
    obs = O.synth(s, m)[0].astype(np.float32)
    r = O.lm_pixel(s, obs)
    np.testing.assert_allclose(r["model"][:9], m[:9], rtol=2e-3, atol=1e-3)
    assert r["chisqr"] < 1e-3
    r_svd = O.lm_pixel(setup(O, psf=("gauss", 60.0), miter=200), obs)
    print("iterations: CORDIC %d, converged SVD %d" % (r["iters"], r_svd["iters"]))

    per = np.loadtxt(os.path.join(GOLDEN, "synthesis_gpu.per"))
    obs = np.ascontiguousarray(per[:, 2:6].T.astype(np.float32))
    guess = read_mod(os.path.join(GOLDEN, "guess.mod"))
    truth = read_mod(os.path.join(GOLDEN, "init.mod"))
    dl, val = crisp_table()
    s = O.make_setup([FE6173], GRID, psf=("table", dl, val), init=guess, free=FREE9,
                     miter=100, noise=1e-6, solver=O.SOLVE_CORDIC)
    r = O.lm_pixel(s, obs)
    np.testing.assert_allclose(r["model"][:9], model_row(truth)[:9], rtol=5e-3, atol=2e-3)
