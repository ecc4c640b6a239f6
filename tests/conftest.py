import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running")


def load_product():
    """Imports the product package (directory g-milos_amd) as gmilos_amd."""
    if "gmilos_amd" in sys.modules:
        return sys.modules["gmilos_amd"]
    path = os.path.join(ROOT, "g-milos_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location(
        "gmilos_amd", path, submodule_search_locations=[os.path.dirname(path)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["gmilos_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def gm():
    return load_product()


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")


def crisp_table():
    t = np.loadtxt(os.path.join(GOLDEN, "CRISP_6173_1mA.psf"))
    return t[:, 0], t[:, 1]


def model_row(d):
    fields = ("eta0", "B", "vlos", "dopp", "aa", "gm", "az", "S0", "S1", "mac", "alfa")
    return np.array([float(d[f]) for f in fields])


def read_mod(path):
    """Minimal .mod reader for the fixtures (key substring : value)."""
    keys = {"eta_0": "eta0", "magnetic field": "B", "LOS velocity": "vlos",
            "Doppler width": "dopp", "damping": "aa", "gamma": "gm", "phi": "az",
            "S_0": "S0", "S_1": "S1", "v_mac": "mac", "filling factor": "alfa"}
    out = {}
    with open(path) as f:
        for line in f:
            if ":" not in line:
                continue
            name, val = line.split(":", 1)
            for k, v in keys.items():
                if k in name:
                    out[v] = float(val.split("!")[0])
    return out


def write_fits(path, data, bitpix, bscale=None, bzero=None):
    """A primary-HDU FITS image written from the standard's definition
    (80-column cards, 2880-byte records, big-endian data); numpy axis order
    is the reverse of NAXISn."""
    cards = ["SIMPLE  = %20s" % "T", "BITPIX  = %20d" % bitpix, "NAXIS   = %20d" % data.ndim]
    cards += ["NAXIS%d  = %20d" % (i + 1, n) for i, n in enumerate(reversed(data.shape))]
    if bscale is not None:
        cards.append("BSCALE  = %20.12G" % bscale)
    if bzero is not None:
        cards.append("BZERO   = %20.12G" % bzero)
    cards.append("END")
    head = "".join(c.ljust(80) for c in cards).encode()
    head += b" " * (-len(head) % 2880)
    dt = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}[bitpix]
    body = np.ascontiguousarray(data, dtype=dt).tobytes()
    body += b"\0" * (-len(body) % 2880)
    with open(path, "wb") as f:
        f.write(head + body)

