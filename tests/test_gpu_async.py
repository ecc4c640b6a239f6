"""gm_invert returns once its work is queued: two contexts on two streams run
side by side and give the results each gives alone; a copy queued behind an
inversion on its stream sees the finished outputs; gm_last_stats waits for
the call it reports."""
import numpy as np
import pytest

from test_gpu_parity import to_dev

pytestmark = pytest.mark.gpu


def _cube(gm, ctx, n, seed, cuda):
    import torch
    truth = torch.from_numpy(gm.random_models(n, seed=seed)).to(cuda)
    obs = ctx.synth(truth)
    gm.add_noise(truth, obs, 1e-3, seed=seed + 1)
    return obs.contiguous()


def test_two_streams_match_serial(gm, cuda):
    import torch
    p = gm.setup_single_line(psf_fwhm=60.0)
    c1, c2 = gm.Context(p), gm.Context(p)
    a = _cube(gm, c1, 20000, 61, cuda)
    b = _cube(gm, c1, 13000, 62, cuda)
    ra = {k: v.clone() for k, v in c1.invert(a).items()}
    rb = {k: v.clone() for k, v in c1.invert(b).items()}
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(s1):
        oa = c1.invert(a)
        host_a = oa["model"].to("cpu", non_blocking=False)
    with torch.cuda.stream(s2):
        ob = c2.invert(b)
    torch.cuda.synchronize()
    for k in ("model", "chisqr", "iters", "synth"):
        assert torch.equal(oa[k], ra[k]), k
        assert torch.equal(ob[k], rb[k]), k
    assert torch.equal(host_a, ra["model"].cpu())
    st1, st2 = c1.stats(), c2.stats()
    assert st1["npix"] == 20000 and st2["npix"] == 13000
    assert st1["lm_rounds"] >= 1 and st2["lm_rounds"] >= 1


def test_stats_wait_for_the_call(gm, cuda):
    import torch
    p = gm.setup_single_line(psf_fwhm=60.0)
    ctx = gm.Context(p)
    obs = _cube(gm, ctx, 50000, 63, cuda)
    torch.cuda.synchronize()
    out = ctx.invert(obs)
    st = ctx.stats()                     # no explicit synchronisation before
    iters = out["iters"].cpu().numpy()
    assert st["npix"] == 50000
    assert st["solves"] == int(iters.sum())
    assert st["lm_rounds"] == int(iters.max())
