"""Serving loop (gm_server): a tile's results are the same alone, in a mix of
tiles of other sizes and setups sharing the launches, and from gm_invert;
admission keeps the in-flight footprint within the HBM budget; oversize tiles
are refused.  Results are compared bit for bit: every pixel's LM trajectory
depends only on its own data."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _obs(gm, ctx, npix, seed, sigma=1e-3):
    import torch
    models = gm.random_models(npix, seed=seed)
    obs = ctx.synth(torch.from_numpy(models).to("cuda"))
    g = torch.Generator(device="cuda").manual_seed(seed)
    ic = torch.from_numpy(models[:, 7] + models[:, 8]).to("cuda").float()[:, None, None]
    obs = obs + sigma * ic * torch.randn(obs.shape, generator=g, device="cuda")
    return obs.contiguous()


def _setups(gm):
    return [gm.setup_single_line(psf_fwhm=60.0, precision=gm.GM_PREC_FP32),
            gm.setup_single_line(precision=gm.GM_PREC_MIXED),
            gm.setup_two_lines(psf_fwhm=60.0, precision=gm.GM_PREC_MIXED),
            gm.setup_two_lines(precision=gm.GM_PREC_FP32)]


def _same(a, b, label):
    for k in ("model", "chisqr", "iters", "naccs", "synth"):
        x, y = a[k].cpu().numpy(), b[k].cpu().numpy()
        assert np.array_equal(x, y, equal_nan=True), (label, k, np.abs(x - y).max())


def _wait_all(srv, tickets, timeout_s=120):
    got = {}
    t0 = time.time()
    while len(got) < len(tickets):
        for r in srv.poll(wait_ms=50):
            got[r["ticket"]] = r
        assert time.time() - t0 < timeout_s, "tiles did not complete"
    return got


def _ext_setup(gm):
    """Macroturbulence fixed at 0.8 km/s, filling factor free against a
    stray-light profile: the eleven-slot kernels."""
    x = np.linspace(-1.0, 1.9, 30)
    stray = np.zeros((4, 30))
    stray[0] = 0.95 - 0.35 * np.exp(-(x / 0.35) ** 2)
    free = ("eta0", "B", "vlos", "dopp", "aa", "gm", "az", "S0", "S1", "alfa")
    return gm.setup_single_line(psf_fwhm=60.0, precision=gm.GM_PREC_MIXED, free=free,
                                init=dict(mac=0.8, alfa=0.9), slight=stray)


@pytest.mark.parametrize("kind", ["plain", "ext", "cordic", "ext_cordic"])
def test_server_tile_matches_invert(gm, cuda, kind):
    """Plain and eleven-slot set-ups, each also with the reference's CORDIC
    step solve (the ext set-up is then a 10-term system, mac fixed)."""
    setup = _setups(gm)[0] if kind in ("plain", "cordic") else _ext_setup(gm)
    if kind.endswith("cordic"):
        setup.solver = gm.GM_SOLVE_CORDIC
    ctx = gm.Context(setup)
    obs = _obs(gm, ctx, 3000, seed=41)
    ref = ctx.invert(obs)
    srv = gm.Server(pool_pixels=1 << 16)
    sid = srv.add_setup(setup)
    t, _ = srv.submit(sid, obs)
    got = _wait_all(srv, [t])
    assert got[t]["status"] == 0
    _same(got[t]["out"], ref, "alone vs gm_invert")
    srv.close()


def test_server_results_same_alone_and_in_mix(gm, cuda):
    setups = _setups(gm)
    ctxs = [gm.Context(s) for s in setups]
    rng = np.random.default_rng(7)
    tiles = []
    for k in range(14):
        sid = int(rng.integers(0, 4))
        npix = int(np.exp(rng.uniform(np.log(256), np.log(20000))))
        tiles.append((sid, _obs(gm, ctxs[sid], npix, seed=100 + k)))
    print("tiles", [(s, o.shape[0]) for s, o in tiles], flush=True)
    # alone: one tile at a time on a fresh server
    alone = []
    srv = gm.Server(pool_pixels=1 << 16)
    ids = [srv.add_setup(s) for s in setups]
    for sid, obs in tiles:
        t, _ = srv.submit(ids[sid], obs)
        alone.append(_wait_all(srv, [t])[t]["out"])
    srv.close()
    # mix: all tiles in flight together, arriving over time
    srv = gm.Server(pool_pixels=1 << 16)
    ids = [srv.add_setup(s) for s in setups]
    tickets = []
    for i, (sid, obs) in enumerate(tiles):
        tickets.append(srv.submit(ids[sid], obs)[0])
        if i % 3 == 2:
            time.sleep(0.002)
    got = _wait_all(srv, tickets)
    st = srv.stats()
    print("server stats", st, flush=True)
    assert st["tiles_done"] == len(tiles)
    for i, t in enumerate(tickets):
        assert got[t]["status"] == 0
        _same(got[t]["out"], alone[i], f"tile {i}")
    srv.close()


def test_server_budget_admission(gm, cuda):
    setup = _setups(gm)[0]
    ctx = gm.Context(setup)
    per_px = 8 * 27 + 4 * (54 + 38) + 4 * 9 + 2 * 4 * 4 * 30  # bytes per pixel slot
    budget = 6000 * per_px
    srv = gm.Server(hbm_budget_bytes=budget, pool_pixels=1 << 15)
    sid = srv.add_setup(setup)
    obs = [_obs(gm, ctx, n, seed=200 + n) for n in (4000, 3000, 2500, 1000, 5000)]
    tickets = [srv.submit(sid, o)[0] for o in obs]
    got = _wait_all(srv, tickets)
    st = srv.stats()
    print("stats", st, flush=True)
    assert all(got[t]["status"] == 0 for t in tickets)
    assert st["peak_inflight_bytes"] <= budget
    # queued tiles waited for budget: later tiles admitted after earlier ones
    assert got[tickets[4]]["queue_ms"] >= 0
    for o, t in zip(obs, tickets):
        ref = ctx.invert(o)
        _same(got[t]["out"], ref, "budgeted")
    with pytest.raises(gm.GmError):
        srv.submit(sid, _obs(gm, ctx, (1 << 15) + 1, seed=9))
    srv.close()
