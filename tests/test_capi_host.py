"""The C-ABI library loads, exports every entry point declared in
include/gmilos.h, and its host-side set-up helpers (no device work) agree
with the oracle's independent set-up and with the reference's fixtures."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, crisp_table

GRID = np.arange(-350.0, 665.0 + 1e-9, 35.0)


def header_functions():
    src = open(os.path.join(ROOT, "include", "gmilos.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gm_[a-z_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol(gm):
    lib = gm.load_library()
    names = header_functions()
    assert len(names) >= 16
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(gm.EXPORTED)


def test_struct_layouts_match_header(gm):
    # sizes computed by the C compiler's layout rules for the header structs
    assert ctypes.sizeof(gm.gm_zeeman) == 8 + 60 * 8 + 4 * 8
    assert ctypes.sizeof(gm.gm_out) == 5 * 8
    st = gm.gm_setup
    assert st.lambda_.offset == 8 and st.zee.offset % 8 == 0


def test_status_strings(gm):
    lib = gm.load_library()
    assert lib.gm_status_string(0) == b"ok"
    assert lib.gm_status_string(4) == b"unsupported configuration"


@pytest.mark.parametrize("qn", [(2, 1, 1.0, 2, 2, 0.0), (2, 1, 2.0, 2, 2, 2.0),
                                (2, 3, 1.0, 2, 2, 0.0), (3, 2, 1.0, 3, 2, 1.0)])
def test_zeeman_matches_oracle(gm, orc, qn):
    a = gm.zeeman(qn)
    b = orc.cuanten(qn)
    assert (a.n_pi, a.n_sig) == (b.n_pi, b.n_sig)
    for f in ("nub", "nup", "nur", "web", "wep", "wer"):
        np.testing.assert_allclose(getattr(a, f)[:], getattr(b, f)[:], rtol=1e-15, atol=1e-15)
    assert a.geff == pytest.approx(b.geff, rel=1e-15)


def test_lines_file_parse(gm):
    path = os.path.join(GOLDEN, "LINES")
    l0, qn = gm.parse_lines(path, 1)
    assert l0 == 6173.3356 and qn == (2.0, 1.0, 1.0, 2.0, 2.0, 0.0)
    l0, qn = gm.parse_lines(path, 2)
    assert l0 == 6301.515 and qn == (2.0, 1.0, 2.0, 2.0, 2.0, 2.0)
    with pytest.raises(gm.GmError):
        gm.parse_lines(path, 99)


def test_psf_helpers_match_oracle(gm, orc):
    lam = [6173.3356 - 0.35]
    for _ in range(29):
        lam.append(lam[-1] + 0.035)
    a = gm.psf_gaussian(60.0, lam)
    b = orc.fgauss_wl(60.0, lam[1] - lam[0], lam[0], lam[15], 30)
    np.testing.assert_allclose(a, b, rtol=1e-14, atol=1e-300)
    dl, val = crisp_table()
    a = gm.psf_from_table(dl, val, GRID, 35.0)
    b = orc.psf_interp(dl, val, GRID, 175.0)
    np.testing.assert_allclose(a, b, rtol=1e-14)
    # a table narrower than the grid is rejected (gmilos.cu:272-275)
    with pytest.raises(gm.GmError):
        gm.psf_from_table(dl[2000:3000], val[2000:3000], GRID, 35.0)


def test_make_setup_grid_and_limits(gm):
    s = gm.setup_single_line(psf_fwhm=60.0)
    assert s.nlambda == 30 and s.use_psf == 1
    assert s.lambda_[0] == pytest.approx(6173.3356 - 0.35)
    assert s.lambda_[29] == pytest.approx(6173.3356 + 0.665)
    assert np.argmax(np.array(s.psf[:30])) == 15
    s2 = gm.setup_two_lines(psf_fwhm=60.0)
    assert s2.nlambda == 112 and s2.nlines == 2 and s2.zee[0].n_pi == 5
    with pytest.raises(ValueError):
        gm.make_setup([gm.FE6173], np.arange(300.0))
