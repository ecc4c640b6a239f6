"""`gmilos <file>.mtrol` end to end on the reference's example (README.md:76-89):
synthesis of run/init.mod reproduces data/synthesis_gpu.per, inverting it
from run/guess.mod recovers run/init.mod, the estimates mode writes a model,
and a FITS cube inverts to the same models as the library call."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, model_row, read_mod, write_fits

pytestmark = pytest.mark.gpu
CLI = os.path.join(ROOT, "g-milos_amd", "bin", "gmilos")


@pytest.fixture()
def refdir(tmp_path):
    if not os.path.exists(CLI):
        pytest.skip("gmilos CLI not built")
    (tmp_path / "run").mkdir()
    (tmp_path / "data").mkdir()
    for f in ("LINES", "init.mod", "guess.mod", "malla.grid", "invert.mtrol",
              "synthesis.mtrol", "CRISP_6173_1mA.psf"):
        shutil.copy(os.path.join(GOLDEN, f), tmp_path / "run" / f)
    shutil.copy(os.path.join(GOLDEN, "synthesis_gpu.per"), tmp_path / "data")
    return tmp_path


def run(cwd, *args):
    print("gmilos", *args, flush=True)
    p = subprocess.run([CLI, *args], cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    print(p.stdout[-2000:])
    assert p.returncode == 0, p.stdout
    return p.stdout


def test_cli_synthesis_matches_reference_fixture(refdir):
    run(refdir, "run/synthesis.mtrol")
    got = np.loadtxt(refdir / "data" / "synthesis_gpu.per")      # overwritten, as the reference
    ref = np.loadtxt(os.path.join(GOLDEN, "synthesis_gpu.per"))
    np.testing.assert_allclose(got[:, :2], ref[:, :2], atol=1e-6)
    assert np.abs(got[:, 2:] - ref[:, 2:]).max() < 2e-6


def test_cli_wavelength_fits_file(refdir):
    """The synthesis example with its grid given as a FITS file of absolute
    wavelengths (gmilos.cu:197-209) reproduces the same fixture."""
    lam = 6173.3356 + np.arange(-350.0, 666.0, 35.0) / 1000.0
    write_fits(str(refdir / "run" / "lambda.fits"), np.stack([np.full(30, 1.0), lam]), -64)
    text = open(refdir / "run" / "synthesis.mtrol").read().splitlines()
    assert text[4].startswith("Wavelength grid file")
    text[4] = "Wavelength grid file      (s):run/lambda.fits  ! wavelength file"
    open(refdir / "run" / "synth_fits.mtrol", "w").write("\n".join(text) + "\n")
    run(refdir, "run/synth_fits.mtrol")
    got = np.loadtxt(refdir / "data" / "synthesis_gpu.per")
    ref = np.loadtxt(os.path.join(GOLDEN, "synthesis_gpu.per"))
    np.testing.assert_allclose(got[:, :2], ref[:, :2], atol=1e-6)
    assert np.abs(got[:, 2:] - ref[:, 2:]).max() < 2e-6


def test_cli_synthesis_with_macroturbulence_and_stray_light(refdir, orc):
    """run/synthesis.mtrol with v_mac = 1.5 km/s, filling factor 0.7 and a
    .per stray-light profile (gmilos.cu:434-461): the written .per matches
    the oracle's synthesis (CRISP PSF, circular macroturbulence kernel)."""
    mod = open(refdir / "run" / "init.mod").read()
    lines = [("1.5" if "v_mac" in l.lower() else "0.7" if "filling" in l.lower() else None, l)
             for l in mod.splitlines()]
    out = []
    for v, l in lines:
        if v is None:
            out.append(l)
        else:
            key = l.split(":")[0]
            out.append("%s:%s" % (key, v))
    open(refdir / "run" / "macsl.mod", "w").write("\n".join(out) + "\n")
    off = np.arange(-350.0, 665.0 + 1e-9, 35.0)
    rng = np.random.default_rng(4)
    stray = np.zeros((4, 30))
    stray[0] = 0.9 - 0.3 * np.exp(-((off - 20.0) / 120.0) ** 2)
    stray[1:] = rng.normal(0, 1e-3, size=(3, 30))
    with open(refdir / "run" / "stray.per", "w") as f:
        for k in range(30):
            f.write("1\t%f\t%e\t%e\t%e\t%e\n" % (off[k], *stray[:, k]))
    text = open(refdir / "run" / "synthesis.mtrol").read().splitlines()
    assert text[2].startswith("Stray light file") and text[7].startswith("Initial guess model 1")
    text[2] = "Stray light file             :run/stray.per     ! stray light"
    text[7] = "Initial guess model 1     (*):run/macsl.mod !"
    open(refdir / "run" / "synth_ms.mtrol", "w").write("\n".join(text) + "\n")
    run(refdir, "run/synth_ms.mtrol")
    got = np.loadtxt(refdir / "data" / "synthesis_gpu.per")      # <observed profiles>.per
    m = read_mod(refdir / "run" / "macsl.mod")
    assert m["mac"] == 1.5 and m["alfa"] == 0.7
    from conftest import crisp_table
    dl, val = crisp_table()
    o = orc.make_setup([(6173.3356, (2, 1, 1.0, 2, 2, 0.0), 1.0)], off, psf=("table", dl, val))
    orc.set_slight(o, stray)
    ref = orc.synth(o, model_row(m))[0]
    np.testing.assert_allclose(got[:, 1], off, atol=1e-6)
    assert np.abs(got[:, 2:].T - ref).max() < 2e-6


def test_cli_stray_light_fits_profile(refdir):
    """A 2-D FITS stray-light profile (readFitsStrayLightFile, utilsFits.c:
    1599-1630; wavelength on either axis) gives the same synthesis as the
    .per profile."""
    off = np.arange(-350.0, 665.0 + 1e-9, 35.0)
    stray = np.zeros((4, 30), dtype=np.float32)
    stray[0] = 0.9 - 0.3 * np.exp(-((off - 20.0) / 120.0) ** 2)
    stray[3] = 1e-3 * np.sin(off / 100.0)
    with open(refdir / "run" / "stray.per", "w") as f:
        for k in range(30):
            f.write("1\t%f\t%e\t%e\t%e\t%e\n" % (off[k], *stray[:, k]))
    write_fits(str(refdir / "run" / "stray_a.fits"), stray, -32)          # naxes (30, 4)
    write_fits(str(refdir / "run" / "stray_b.fits"), stray.T.copy(), -32)  # naxes (4, 30)
    mod = open(refdir / "run" / "init.mod").read().splitlines()
    mod = [("filling factor: 0.6" if l.lower().startswith("filling") else l) for l in mod]
    open(refdir / "run" / "ff.mod", "w").write("\n".join(mod) + "\n")
    res = []
    for name in ("stray.per", "stray_a.fits", "stray_b.fits"):
        text = open(refdir / "run" / "synthesis.mtrol").read().splitlines()
        text[2] = "Stray light file             :run/%s     ! stray light" % name
        text[7] = "Initial guess model 1     (*):run/ff.mod !"
        open(refdir / "run" / "s.mtrol", "w").write("\n".join(text) + "\n")
        run(refdir, "run/s.mtrol")
        res.append(np.loadtxt(refdir / "data" / "synthesis_gpu.per"))
    np.testing.assert_array_equal(res[1], res[2])
    np.testing.assert_allclose(res[0], res[1], atol=1e-6)   # .per holds 7 digits
    plain = np.loadtxt(os.path.join(GOLDEN, "synthesis_gpu.per"))
    assert np.abs(res[0][:, 2] - plain[:, 2]).max() > 1e-2       # the stray light is in


def test_cli_inverts_filling_factor_with_stray_light(refdir):
    """Synthesise run/init.mod with filling factor 0.75 and a stray-light
    .per, then invert it from run/guess.mod with "Invert stray light factor"
    on (fix[10], the reference's 10-term run with alfa free): the model and
    the filling factor are recovered."""
    off = np.arange(-350.0, 665.0 + 1e-9, 35.0)
    stray = np.zeros((4, 30))
    stray[0] = 0.95 - 0.35 * np.exp(-((off - 40.0) / 110.0) ** 2)
    with open(refdir / "run" / "stray.per", "w") as f:
        for k in range(30):
            f.write("1\t%f\t%e\t%e\t%e\t%e\n" % (off[k], *stray[:, k]))
    mod = open(refdir / "run" / "init.mod").read().splitlines()
    mod = [("filling factor: 0.75" if l.lower().startswith("filling") else l) for l in mod]
    open(refdir / "run" / "ff.mod", "w").write("\n".join(mod) + "\n")
    text = open(refdir / "run" / "synthesis.mtrol").read().splitlines()
    text[2] = "Stray light file             :run/stray.per     ! stray light"
    text[7] = "Initial guess model 1     (*):run/ff.mod !"
    open(refdir / "run" / "synth_ff.mtrol", "w").write("\n".join(text) + "\n")
    run(refdir, "run/synth_ff.mtrol")
    text = open(refdir / "run" / "invert.mtrol").read().splitlines()
    assert text[35].startswith("Invert stray light factor")
    text[2] = "Stray light file             :run/stray.per     ! stray light"
    text[35] = "Invert stray light factor?   :1                 ! (0 or blank=no, 1=yes)"
    open(refdir / "run" / "inv_ff.mtrol", "w").write("\n".join(text) + "\n")
    run(refdir, "run/inv_ff.mtrol", "--mixed")
    m = read_mod(refdir / "data" / "synthesis_gpu_model.mod")
    truth = read_mod(refdir / "run" / "ff.mod")
    np.testing.assert_allclose(model_row(m)[:9], model_row(truth)[:9], rtol=5e-3, atol=2e-3)
    assert abs(m["alfa"] - 0.75) < 1e-3


def test_cli_inverts_reference_example(refdir):
    run(refdir, "run/invert.mtrol")
    m = read_mod(refdir / "data" / "synthesis_gpu_model.mod")
    truth = read_mod(os.path.join(GOLDEN, "init.mod"))
    np.testing.assert_allclose(model_row(m)[:9], model_row(truth)[:9], rtol=5e-3, atol=2e-3)
    syn = np.loadtxt(refdir / "data" / "synthesis_gpu_stokes.per")
    ref = np.loadtxt(os.path.join(GOLDEN, "synthesis_gpu.per"))
    assert np.abs(syn[:, 2:] - ref[:, 2:]).max() < 1e-3
    run(refdir, "run/invert.mtrol", "--mixed")
    m2 = read_mod(refdir / "data" / "synthesis_gpu_model.mod")
    np.testing.assert_allclose(model_row(m2)[:9], model_row(truth)[:9], rtol=5e-3, atol=2e-3)
    run(refdir, "run/invert.mtrol", "--cordic")          # the reference's own step solve
    m3 = read_mod(refdir / "data" / "synthesis_gpu_model.mod")
    np.testing.assert_allclose(model_row(m3)[:9], model_row(truth)[:9], rtol=5e-3, atol=2e-3)
    run(refdir, "run/invert.mtrol", "--jacobi")          # every step through the SVD
    m4 = read_mod(refdir / "data" / "synthesis_gpu_model.mod")
    np.testing.assert_allclose(model_row(m4)[:9], model_row(truth)[:9], rtol=5e-3, atol=2e-3)


def test_cli_classical_estimates_mode(refdir):
    text = open(refdir / "run" / "invert.mtrol").read().splitlines()
    text[0] = "Number of cycles          (*):-1                         ! estimates"
    open(refdir / "run" / "ce.mtrol", "w").write("\n".join(text) + "\n")
    run(refdir, "run/ce.mtrol")
    m = read_mod(refdir / "data" / "synthesis_gpu_model_ce.mod")
    assert 500 < m["B"] < 3000 and 0 <= m["az"] <= 180
    assert m["eta0"] == 0 and m["S1"] == 0


def test_cli_fits_cube_matches_library(refdir, gm, cuda):
    import torch
    from test_io import _write_cube, IO_LIB
    io = ctypes.CDLL(IO_LIB)
    io.gmio_fits_write_image.argtypes = [ctypes.c_char_p, ctypes.c_int32] + [ctypes.c_void_p] * 3
    io.gmio_fits_read_image.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32),
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
    nr, nc = 6, 5
    setup = gm.setup_single_line(psf_fwhm=60.0, precision=gm.GM_PREC_FP32,
                                 init=read_mod(os.path.join(GOLDEN, "guess.mod")), miter=100)
    ctx = gm.Context(setup)
    truth = torch.from_numpy(gm.random_models(nr * nc, seed=5)).to(cuda)
    obs = ctx.synth(truth)
    cube = obs.cpu().numpy().reshape(nc, nr, 4, 30).transpose(1, 0, 2, 3)   # [row, col, s, l]
    path = str(refdir / "data" / "cube.fits").encode()
    _write_cube(io, path, ("lambda", "stokes", "row", "col"), cube)
    text = open(refdir / "run" / "invert.mtrol").read().splitlines()
    text[1] = "Observed profiles          (*):data/cube.fits    !"
    text[3] = "PSF file                     :60                 ! FWHM"
    open(refdir / "run" / "cube.mtrol", "w").write("\n".join(text) + "\n")
    run(refdir, "run/cube.mtrol")
    naxis = ctypes.c_int32()
    naxes = np.zeros(4, dtype=np.int64)
    buf = np.zeros(nr * nc * 13, dtype=np.float32)
    assert io.gmio_fits_read_image(str(refdir / "data" / "cube_mod.fits").encode(),
                                   ctypes.byref(naxis), naxes.ctypes.data, buf.ctypes.data,
                                   buf.size) == 0
    planes = buf.reshape(13, nr * nc)
    lib = ctx.invert(obs.contiguous())
    np.testing.assert_allclose(planes[:11].T, lib["model"].cpu().numpy().astype(np.float32),
                               rtol=1e-6, atol=1e-5)
    np.testing.assert_array_equal(planes[11], lib["iters"].cpu().numpy().astype(np.float32))
    assert os.path.exists(refdir / "data" / "cube_stokes.fits")
