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


def test_server_back_to_back_single_tiles_on_idle_server(gm, cuda):
    """Each tile submitted to an idle server completes with status 0: a tile
    never fails admission for arriving while the worker is between passes."""
    setup = _setups(gm)[1]
    ctx = gm.Context(setup)
    obs = _obs(gm, ctx, 4096, seed=55)
    ref = ctx.invert(obs[:64].contiguous())
    srv = gm.Server(pool_pixels=1 << 14)
    sid = srv.add_setup(setup)
    for k in range(150):
        n = 1 + (k * 37) % 64
        t, _ = srv.submit(sid, obs[:n].contiguous())
        r = _wait_all(srv, [t])[t]
        assert r["status"] == 0, (k, n, r["status"])
        if n == 64:
            _same(r["out"], ref, "single tile")
    srv.close()


def _bytes_per_slot(nlambda):
    """Pool footprint of one state slot (gm_server.cpp's build_engine)."""
    return 8 * (2 * 12 + 5) + 4 * (80 + 2 * 20) + 4 * (2 + 4 + 3) + 4 * 2 * 4 * nlambda


@pytest.mark.parametrize("aging_ms", [10000, -1])
def test_server_admission_passes_tiles_that_do_not_fit(gm, cuda, aging_ms):
    """Budget for 10 000 slots: a 6 000-pixel tile in flight, a second one
    that does not fit behind it, then a 2 000-pixel tile that does.  With
    aging the small tile enters at once, ahead of the waiting one; in strict
    arrival order it waits behind it."""
    setup = _setups(gm)[1]
    ctx = gm.Context(setup)
    budget = 10000 * _bytes_per_slot(30)
    srv = gm.Server(hbm_budget_bytes=budget, pool_pixels=1 << 15, aging_ms=aging_ms)
    sid = srv.add_setup(setup)
    obs = [_obs(gm, ctx, n, seed=300 + n) for n in (6000, 6000, 2000)]
    tickets = [srv.submit(sid, o)[0] for o in obs]
    got = _wait_all(srv, tickets)
    admit = [got[t]["t_submit_ms"] + got[t]["queue_ms"] for t in tickets]
    print("aging", aging_ms, "admitted at", admit, flush=True)
    assert all(got[t]["status"] == 0 for t in tickets)
    assert srv.stats()["peak_inflight_bytes"] <= budget
    if aging_ms > 0:
        assert admit[2] < admit[1]
    else:
        assert admit[2] >= admit[1]
    for o, t in zip(obs, tickets):
        _same(got[t]["out"], ctx.invert(o), "admission order")
    srv.close()


def test_server_cancel(gm, cuda):
    """A queued tile and a tile in flight are cancelled: both come back with
    GM_CANCELLED and the other tiles are unaffected (equal to gm_invert); a
    completed or unknown ticket cannot be cancelled."""
    setup = _setups(gm)[2]
    ctx = gm.Context(setup)
    budget = 12000 * _bytes_per_slot(112)
    srv = gm.Server(hbm_budget_bytes=budget, pool_pixels=1 << 15, aging_ms=-1)
    sid = srv.add_setup(setup)
    obs = [_obs(gm, ctx, n, seed=400 + n) for n in (8000, 9000, 1500, 700)]
    tickets = [srv.submit(sid, o)[0] for o in obs]
    # tile 1 cannot enter while tile 0 is in flight (strict order holds
    # tiles 2 and 3 behind it): cancel it in the queue and tile 0 in flight
    assert srv.cancel(tickets[1])
    assert srv.cancel(tickets[0])
    assert not srv.cancel(tickets[0])            # already being cancelled
    assert not srv.cancel(987654321)
    got = _wait_all(srv, tickets)
    assert got[tickets[0]]["status"] == gm.GM_CANCELLED
    assert got[tickets[1]]["status"] == gm.GM_CANCELLED
    for i in (2, 3):
        assert got[tickets[i]]["status"] == 0
        _same(got[tickets[i]]["out"], ctx.invert(obs[i]), f"tile {i} beside cancelled ones")
    assert not srv.cancel(tickets[2])            # completed
    st = srv.stats()
    assert st["pending_tiles"] == 0 and st["inflight_tiles"] == 0
    # the pool is whole again: a tile of the full budget runs
    t, _ = srv.submit(sid, obs[1])
    r = _wait_all(srv, [t])[t]
    assert r["status"] == 0
    _same(r["out"], ctx.invert(obs[1]), "after cancellations")
    srv.close()
