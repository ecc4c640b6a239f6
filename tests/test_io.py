"""File surface of the gmilos command: the reference's example control file,
models, grid, PSF table and profiles parse to the values their readers
produce (readConfig.c, gmilos.cu), and the FITS codec round-trips cubes in
every axis order with and without CTYPE keywords (utilsFits.c:267-649)."""
import ctypes
import itertools
import os
import shutil

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, write_fits

IO_LIB = os.path.join(ROOT, "g-milos_amd", "lib", "libgmilos_io.so")
P = 4096


class gmio_control(ctypes.Structure):
    _fields_ = ([("number_of_cycles", ctypes.c_int32)] +
                [(n, ctypes.c_char * P) for n in (
                    "observed_profiles", "stray_light_file", "psf_file", "wavelength_file",
                    "atomic_parameters_file", "abundances_file", "initial_guess_model",
                    "initial_guess_model_2")] +
                [("weight", ctypes.c_double * 4), ("automatic_select_of_nodes", ctypes.c_int32),
                 ("fix", ctypes.c_int32 * 11), ("fix2", ctypes.c_int32 * 11),
                 ("invert_filling_factor", ctypes.c_int32), ("mu", ctypes.c_double),
                 ("estimated_sn_for_i", ctypes.c_int32), ("continuum_contrast", ctypes.c_int32),
                 ("initial_diagonal_element", ctypes.c_double), ("use_fft", ctypes.c_int32),
                 ("logclambda", ctypes.c_int32), ("num_streams", ctypes.c_int32),
                 ("device_id", ctypes.c_int32), ("noise", ctypes.c_double),
                 ("toplim", ctypes.c_double), ("convolve_with_psf", ctypes.c_int32),
                 ("fwhm", ctypes.c_double), ("use_malla_grid", ctypes.c_int32),
                 ("save_chisqr", ctypes.c_int32), ("save_synthesis_adjusted", ctypes.c_int32)])


@pytest.fixture(scope="module")
def io():
    if not os.path.exists(IO_LIB):
        pytest.skip("libgmilos_io.so not built")
    L = ctypes.CDLL(IO_LIB)
    vp = ctypes.c_void_p
    L.gmio_read_mtrol.argtypes = [ctypes.c_char_p, ctypes.POINTER(gmio_control)]
    L.gmio_read_mod.argtypes = [ctypes.c_char_p, vp]
    L.gmio_write_mod.argtypes = [ctypes.c_char_p, vp, ctypes.c_int32, ctypes.c_double]
    L.gmio_read_grid.argtypes = [ctypes.c_char_p] + [ctypes.POINTER(ctypes.c_double)] * 3
    L.gmio_read_psf.argtypes = [ctypes.c_char_p, vp, vp, ctypes.c_int32]
    L.gmio_count_lines.argtypes = [ctypes.c_char_p]
    L.gmio_read_per.argtypes = [ctypes.c_char_p, ctypes.c_int32, vp]
    L.gmio_write_per.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, vp, vp]
    L.gmio_fits_read_cube.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(vp)]
    L.gmio_cube_free.argtypes = [vp]
    L.gmio_cube_dims.argtypes = [vp] + [ctypes.POINTER(ctypes.c_int32)] * 3
    L.gmio_cube_data.argtypes = [vp]
    L.gmio_cube_data.restype = ctypes.POINTER(ctypes.c_float)
    L.gmio_fits_write_profiles.argtypes = [ctypes.c_char_p, vp, vp]
    L.gmio_fits_write_models.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, vp,
                                         vp, vp, ctypes.c_int32]
    L.gmio_fits_write_image.argtypes = [ctypes.c_char_p, ctypes.c_int32, vp, vp, vp]
    L.gmio_fits_read_image.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32), vp, vp,
                                       ctypes.c_int64]
    L.gmio_fits_read_lambda.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32), vp,
                                        ctypes.c_int32]
    L.gmio_read_stray_per.argtypes = [ctypes.c_char_p, ctypes.c_int32, vp, vp]
    L.gmio_file_readable.argtypes = [ctypes.c_char_p]
    L.gmio_file_ext.argtypes = [ctypes.c_char_p]
    L.gmio_file_ext.restype = ctypes.c_char_p
    L.gmio_base_name.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32]
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@pytest.fixture()
def refdir(tmp_path):
    """The reference's run/ and data/ layout (paths in the .mtrol files)."""
    (tmp_path / "run").mkdir()
    (tmp_path / "data").mkdir()
    for f in ("LINES", "init.mod", "guess.mod", "malla.grid", "invert.mtrol",
              "synthesis.mtrol", "CRISP_6173_1mA.psf"):
        shutil.copy(os.path.join(GOLDEN, f), tmp_path / "run" / f)
    shutil.copy(os.path.join(GOLDEN, "synthesis_gpu.per"), tmp_path / "data")
    return tmp_path


def test_read_invert_mtrol(io, refdir, monkeypatch):
    monkeypatch.chdir(refdir)
    c = gmio_control()
    assert io.gmio_read_mtrol(b"run/invert.mtrol", ctypes.byref(c)) == 0
    assert c.number_of_cycles == 100
    assert c.observed_profiles == b"data/synthesis_gpu.per"
    assert c.psf_file == b"run/CRISP_6173_1mA.psf" and c.convolve_with_psf == 1
    assert c.wavelength_file == b"run/malla.grid" and c.use_malla_grid == 1
    assert c.atomic_parameters_file == b"run/LINES"
    assert c.initial_guess_model == b"run/guess.mod"
    assert list(c.weight) == [1.0, 1.0, 1.0, 1.0]
    assert list(c.fix) == [1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0]
    assert c.mu == 1.0                                  # blank keeps the default
    assert c.estimated_sn_for_i == 1000
    assert c.noise == pytest.approx(1e-6, rel=1e-7)
    assert c.initial_diagonal_element == pytest.approx(0.1)
    assert c.num_streams == 4 and c.device_id == 0 and c.logclambda == 0
    s = gmio_control()
    assert io.gmio_read_mtrol(b"run/synthesis.mtrol", ctypes.byref(s)) == 0
    assert s.number_of_cycles == 0 and s.initial_guess_model == b"run/init.mod"


def test_mtrol_psf_as_fwhm_and_empty(io, refdir, monkeypatch):
    monkeypatch.chdir(refdir)
    text = open("run/invert.mtrol").read().splitlines()
    text[3] = "PSF file                     :60                 ! FWHM"
    open("a.mtrol", "w").write("\n".join(text) + "\n")
    c = gmio_control()
    assert io.gmio_read_mtrol(b"a.mtrol", ctypes.byref(c)) == 0
    assert c.convolve_with_psf == 1 and c.fwhm == 60.0
    text[3] = "PSF file                     :                  ! none"
    open("b.mtrol", "w").write("\n".join(text) + "\n")
    assert io.gmio_read_mtrol(b"b.mtrol", ctypes.byref(c)) == 0
    assert c.convolve_with_psf == 0
    assert io.gmio_read_mtrol(b"missing.mtrol", ctypes.byref(c)) == -1


def test_mod_grid_psf_per(io, tmp_path):
    m = np.zeros(11)
    assert io.gmio_read_mod(os.path.join(GOLDEN, "init.mod").encode(), _p(m)) == 0
    np.testing.assert_array_equal(m, [20, 1100, 0.2, 0.03, 0.05, 120, 150, 0.35, 0.5, 0, 1])
    f, s, l = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    idx = io.gmio_read_grid(os.path.join(GOLDEN, "malla.grid").encode(), ctypes.byref(f),
                            ctypes.byref(s), ctypes.byref(l))
    assert (idx, f.value, s.value, l.value) == (1, -350.0, 35.0, 665.0)
    path = os.path.join(GOLDEN, "CRISP_6173_1mA.psf").encode()
    n = io.gmio_count_lines(path)
    dl, val = np.zeros(n), np.zeros(n)
    assert io.gmio_read_psf(path, _p(dl), _p(val), n) == 5500
    ref = np.loadtxt(os.path.join(GOLDEN, "CRISP_6173_1mA.psf"))
    np.testing.assert_array_equal(dl, ref[:, 0])
    np.testing.assert_array_equal(val, ref[:, 1])
    sp = np.zeros((4, 30), dtype=np.float32)
    assert io.gmio_read_per(os.path.join(GOLDEN, "synthesis_gpu.per").encode(), 30, _p(sp)) == 0
    per = np.loadtxt(os.path.join(GOLDEN, "synthesis_gpu.per"))
    np.testing.assert_array_equal(sp, per[:, 2:6].T.astype(np.float32))
    out = str(tmp_path / "x.per").encode()
    off = per[:, 1].copy()
    assert io.gmio_write_per(out, 1, 30, _p(off), _p(sp)) == 0
    back = np.loadtxt(out.decode())
    np.testing.assert_allclose(back, per, rtol=1e-6, atol=1e-12)
    mo = str(tmp_path / "x.mod").encode()
    assert io.gmio_write_mod(mo, _p(m), 7, 1.5) == 0
    m2 = np.zeros(11)
    io.gmio_read_mod(mo, _p(m2))
    np.testing.assert_allclose(m2, m)


def test_name_helpers(io):
    assert io.gmio_file_ext(b"data/synthesis_gpu.per") == b".per"
    assert io.gmio_file_ext(b"noext") == b""
    buf = ctypes.create_string_buffer(4096)
    io.gmio_base_name(b"data/cube.fits", buf, 4096)
    assert buf.value == b"data/cube"


CT = {"lambda": b"WAVE-GRI", "row": b"HPLN-TAN", "col": b"HPLT-TAN", "stokes": b"STOKES"}


def _write_cube(io, path, order, cube, header=True):
    """cube[row, col, stokes, lambda] written with axis `order` (NAXIS1 first)."""
    axes = {"row": 0, "col": 1, "stokes": 2, "lambda": 3}
    arr = np.transpose(cube, [axes[a] for a in reversed(order)])   # C order: NAXIS4 first
    arr = np.ascontiguousarray(arr, dtype=np.float32)
    naxes = np.array([cube.shape[axes[a]] for a in order], dtype=np.int64)
    ct = (ctypes.c_char_p * 4)(*[CT[a] for a in order]) if header else None
    assert io.gmio_fits_write_image(path, 4, _p(naxes), ct, _p(arr)) == 0


@pytest.mark.parametrize("order", list(itertools.permutations(["row", "col", "stokes", "lambda"])))
def test_fits_cube_axis_orders(io, tmp_path, order):
    rng = np.random.default_rng(0)
    nr, nc, nl = 5, 3, 30
    cube = rng.normal(size=(nr, nc, 4, nl)).astype(np.float32)
    path = str(tmp_path / "c.fits").encode()
    _write_cube(io, path, order, cube)
    h = ctypes.c_void_p()
    assert io.gmio_fits_read_cube(path, 30, ctypes.byref(h)) == 0
    r, c_, l = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    io.gmio_cube_dims(h, ctypes.byref(r), ctypes.byref(c_), ctypes.byref(l))
    assert (r.value, c_.value, l.value) == (nr, nc, nl)
    data = np.ctypeslib.as_array(io.gmio_cube_data(h), shape=(nr * nc * 4 * nl,)).copy()
    data = data.reshape(nc, nr, 4, nl)               # pixel = col * nrows + row
    np.testing.assert_array_equal(data.transpose(1, 0, 2, 3), cube)
    # the profile writer restores the input's axis order and header
    out = str(tmp_path / "p.fits").encode()
    assert io.gmio_fits_write_profiles(out, h, _p(np.ascontiguousarray(data.reshape(-1)))) == 0
    assert open(out, "rb").read() == open(path, "rb").read()
    io.gmio_cube_free(h)


def test_fits_headerless_inference(io, tmp_path):
    """Without CTYPEs: Stokes = smallest axis, lambda = next smallest."""
    rng = np.random.default_rng(1)
    cube = rng.normal(size=(40, 50, 4, 30)).astype(np.float32)
    path = str(tmp_path / "h.fits").encode()
    _write_cube(io, path, ("lambda", "stokes", "row", "col"), cube, header=False)
    h = ctypes.c_void_p()
    assert io.gmio_fits_read_cube(path, 30, ctypes.byref(h)) == 0
    data = np.ctypeslib.as_array(io.gmio_cube_data(h), shape=(40 * 50 * 4 * 30,)).copy()
    np.testing.assert_array_equal(data.reshape(50, 40, 4, 30).transpose(1, 0, 2, 3), cube)
    io.gmio_cube_free(h)
    # wrong wavelength count and non-4 Stokes axes are rejected
    assert io.gmio_fits_read_cube(path, 31, ctypes.byref(h)) == -3


def test_fits_model_planes(io, tmp_path):
    nr, nc = 4, 6
    rng = np.random.default_rng(2)
    models = rng.normal(size=(nr * nc, 11))
    c2 = rng.random(nr * nc).astype(np.float32)
    it = rng.integers(1, 50, nr * nc).astype(np.int32)
    path = str(tmp_path / "m.fits").encode()
    assert io.gmio_fits_write_models(path, nr, nc, _p(models), _p(c2), _p(it), 1) == 0
    naxis = ctypes.c_int32()
    naxes = np.zeros(4, dtype=np.int64)
    buf = np.zeros(nr * nc * 13, dtype=np.float32)
    assert io.gmio_fits_read_image(path, ctypes.byref(naxis), _p(naxes), _p(buf), buf.size) == 0
    assert naxis.value == 3 and list(naxes[:3]) == [nr, nc, 13]
    planes = buf.reshape(13, nr * nc)                # plane-major, pixel order
    np.testing.assert_array_equal(planes[:11].T, models.astype(np.float32))
    np.testing.assert_array_equal(planes[11], it.astype(np.float32))
    np.testing.assert_array_equal(planes[12], c2)


def test_fits_wavelength_file(io, tmp_path):
    """readFitsLambdaToArray (utilsFits.c:1446-1529): naxes (nlambda, 2), row 1
    the line index, row 2 the wavelengths."""
    lam = 6173.3356 + np.arange(-350.0, 666.0, 35.0) / 1000.0
    img = np.stack([np.full(30, 1.0), lam])
    idx = ctypes.c_int32()
    out = np.zeros(64)
    path = str(tmp_path / "l64.fits")
    write_fits(path, img, -64)
    assert io.gmio_fits_read_lambda(path.encode(), ctypes.byref(idx), _p(out), 64) == 30
    assert idx.value == 1
    np.testing.assert_array_equal(out[:30], lam)
    # float32 storage reads back the rounded values, as a TDOUBLE read does
    write_fits(path, img, -32)
    assert io.gmio_fits_read_lambda(path.encode(), ctypes.byref(idx), _p(out), 64) == 30
    np.testing.assert_array_equal(out[:30], lam.astype(np.float32).astype(np.float64))
    # scaled integers: value = BZERO + BSCALE * stored
    ints = np.stack([np.full(30, 2), np.arange(30) * 7 - 100]).astype(np.int16)
    write_fits(path, ints, 16, bscale=0.5, bzero=6000.0)
    assert io.gmio_fits_read_lambda(path.encode(), ctypes.byref(idx), _p(out), 64) == 30
    np.testing.assert_array_equal(out[:30], 6000.0 + 0.5 * (np.arange(30) * 7 - 100))
    assert idx.value == 6001                  # line-index row is scaled too
    # too many wavelengths for the buffer, a third row, a 1-D image, no file
    write_fits(path, img, -64)
    assert io.gmio_fits_read_lambda(path.encode(), ctypes.byref(idx), _p(out), 20) == -3
    write_fits(path, np.vstack([img, img[1:]]), -64)
    assert io.gmio_fits_read_lambda(path.encode(), ctypes.byref(idx), _p(out), 64) < 0
    write_fits(path, lam, -64)
    assert io.gmio_fits_read_lambda(path.encode(), ctypes.byref(idx), _p(out), 64) < 0
    assert io.gmio_fits_read_lambda(b"missing.fits", ctypes.byref(idx), _p(out), 64) < 0


def test_stray_light_per(io, tmp_path):
    """readPerStrayLightFile (utilsFits.c:1532-1575): four Stokes columns, the
    offsets column checked against the truncated grid offsets."""
    off = np.arange(-350.0, 665.0 + 1e-9, 35.0)
    sp = np.arange(120, dtype=np.float32).reshape(4, 30) / 100.0
    path = str(tmp_path / "s.per")
    with open(path, "w") as f:
        for k in range(30):
            f.write("1\t%f\t%e\t%e\t%e\t%e\n" % (off[k], *sp[:, k]))
    got = np.zeros((4, 30), dtype=np.float32)
    assert io.gmio_read_stray_per(path.encode(), 30, _p(off), _p(got)) == 0
    np.testing.assert_array_equal(got, sp)
    shifted = off + 1.0
    assert io.gmio_read_stray_per(path.encode(), 30, _p(shifted), _p(got)) == -3
    assert io.gmio_read_stray_per(path.encode(), 31, _p(np.zeros(31)), _p(np.zeros((4, 31), np.float32))) < 0
    assert io.gmio_read_stray_per(b"missing.per", 30, _p(off), _p(got)) == -1
    assert io.gmio_file_readable(path.encode()) == 1
    assert io.gmio_file_readable(b"missing.per") == 0 and io.gmio_file_readable(b"") == 0
