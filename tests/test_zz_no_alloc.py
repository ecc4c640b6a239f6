"""gm_invert_noalloc: the inversion with its LM state in the library's fixed
per-device array instead of an allocation, called through the C ABI and
held bit for bit against gm_invert on the same inputs -- two calls in a row,
the second with fewer pixels, on the one-line, two-line and eleven-slot
set-ups -- and refused above GM_NOALLOC_MAX_PIXELS."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _obs(gm, ctx, npix, seed):
    import torch
    models = torch.from_numpy(gm.random_models(npix, seed=seed)).to("cuda")
    obs = ctx.synth(models)
    g = torch.Generator(device="cuda").manual_seed(seed)
    ic = (models[:, 7] + models[:, 8]).float()[:, None, None]
    return (obs + 1e-3 * ic * torch.randn(obs.shape, generator=g, device="cuda")).contiguous()


def _outputs(gm, npix, nlambda):
    import torch
    return {"model": torch.full((npix, 11), np.nan, dtype=torch.float64, device="cuda"),
            "chisqr": torch.full((npix,), np.nan, dtype=torch.float32, device="cuda"),
            "iters": torch.full((npix,), -1, dtype=torch.int32, device="cuda"),
            "naccs": torch.full((npix,), -1, dtype=torch.int32, device="cuda"),
            "synth": torch.full((npix, 4, nlambda), np.nan, dtype=torch.float32, device="cuda")}


def _ext_setup(gm):
    x = np.linspace(-1.0, 1.9, 30)
    stray = np.zeros((4, 30))
    stray[0] = 0.95 - 0.35 * np.exp(-(x / 0.35) ** 2)
    free = ("eta0", "B", "vlos", "dopp", "aa", "gm", "az", "S0", "S1", "alfa")
    return gm.setup_single_line(psf_fwhm=60.0, free=free, init=dict(mac=0.8, alfa=0.9),
                                slight=stray)


@pytest.mark.parametrize("kind", ["one_line", "two_lines", "eleven_slots"])
def test_noalloc_matches_invert(gm, cuda, kind):
    import torch
    if kind == "one_line":
        setup, sizes = gm.setup_single_line(psf_fwhm=60.0), (gm.GM_NOALLOC_MAX_PIXELS, 3001)
    elif kind == "two_lines":
        setup, sizes = gm.setup_two_lines(psf_fwhm=60.0), (20000, 777)
    else:
        setup, sizes = _ext_setup(gm), (12000, 513)
    plain = gm.Context(setup)
    noalloc = gm.Context(setup)
    for k, n in enumerate(sizes):
        print(kind, "call", k, "npix", n, flush=True)
        obs = _obs(gm, plain, n, seed=10 * k + len(kind))
        ref = plain.invert(obs)
        got = noalloc.invert_noalloc(obs, _outputs(gm, n, setup.nlambda))
        torch.cuda.synchronize()
        for f in ("model", "chisqr", "iters", "naccs", "synth"):
            a, b = got[f].cpu().numpy(), ref[f].cpu().numpy()
            assert np.array_equal(a, b), (kind, k, f, np.nanmax(np.abs(a - b)))
        st = noalloc.stats()
        assert st["npix"] == n and st["lm_rounds"] >= 1


def test_noalloc_refuses_more_than_its_maximum(gm, cuda):
    import torch
    setup = gm.setup_single_line()
    ctx = gm.Context(setup)
    n = gm.GM_NOALLOC_MAX_PIXELS + 1
    obs = torch.zeros((n, 4, 30), dtype=torch.float32, device="cuda")
    with pytest.raises(gm.GmError) as e:
        ctx.invert_noalloc(obs, _outputs(gm, n, 30))
    assert e.value.code == 1
    # an empty call is a no-op
    e0 = torch.zeros((0, 4, 30), dtype=torch.float32, device="cuda")
    ctx.invert_noalloc(e0, _outputs(gm, 0, 30))
