"""ctypes binding of the FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the benchmark's
CPU-baseline leg, never by the product path.  The oracle restates the
reference's per-pixel routines in double precision (see gm_oracle.c for the
file:line each routine follows) and builds its own set-up (Zeeman pattern,
PSF) from the same ingredients the product uses, so parity tests compare two
independent computations.

Parity status of each routine (what pins it) is recorded in gm_oracle.h's
companion tests: tests/test_oracle_*.py.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

MAXL, MAXLINES, MAXCOMP, NT, NMODEL = 256, 4, 10, 10, 11
MODEL_FIELDS = ("eta0", "B", "vlos", "dopp", "aa", "gm", "az", "S0", "S1", "mac", "alfa")


class gmo_zeeman(ctypes.Structure):
    _fields_ = [("n_pi", ctypes.c_int), ("n_sig", ctypes.c_int),
                ("nub", ctypes.c_double * MAXCOMP), ("nup", ctypes.c_double * MAXCOMP),
                ("nur", ctypes.c_double * MAXCOMP), ("web", ctypes.c_double * MAXCOMP),
                ("wep", ctypes.c_double * MAXCOMP), ("wer", ctypes.c_double * MAXCOMP),
                ("gl", ctypes.c_double), ("gu", ctypes.c_double), ("geff", ctypes.c_double),
                ("fo", ctypes.c_double)]


class gmo_setup(ctypes.Structure):
    _fields_ = [("nlambda", ctypes.c_int), ("lambda_", ctypes.c_double * MAXL),
                ("nlines", ctypes.c_int), ("lambda0", ctypes.c_double * MAXLINES),
                ("zee", gmo_zeeman * MAXLINES), ("use_psf", ctypes.c_int),
                ("psf", ctypes.c_double * MAXL), ("weight", ctypes.c_double * 4),
                ("noise", ctypes.c_double), ("ilambda", ctypes.c_double),
                ("toplim", ctypes.c_double), ("miter", ctypes.c_int), ("mu", ctypes.c_double),
                ("logclambda", ctypes.c_int), ("fix", ctypes.c_int * 11),
                ("init_model", ctypes.c_double * NMODEL), ("use_ce", ctypes.c_int),
                ("ce_line", ctypes.c_int), ("rf_reference_rules", ctypes.c_int),
                ("use_slight", ctypes.c_int), ("slight", ctypes.c_double * (4 * MAXL)),
                ("nterms", ctypes.c_int), ("solver", ctypes.c_int)]

SOLVE_SVD, SOLVE_CORDIC = 0, 3


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        d, i, vp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        P = ctypes.POINTER
        L.gmo_cuanten.argtypes = [d] * 7 + [P(gmo_zeeman)]
        L.gmo_fgauss_wl.argtypes = [d, d, d, d, i, vp]
        L.gmo_psf_interp.argtypes = [vp, vp, i, vp, i, d, vp]
        L.gmo_voigt.argtypes = [d, d, P(d), P(d)]
        L.gmo_conv.argtypes = [vp, i, vp]
        L.gmo_conv_ic.argtypes = [vp, i, vp, d]
        L.gmo_conv_circ.argtypes = [vp, i, vp]
        L.gmo_mac_kernels.argtypes = [P(gmo_setup), d, vp, vp]
        L.gmo_synth_rf2.argtypes = [P(gmo_setup), vp, vp, vp, i]
        L.gmo_nterms.argtypes = [P(gmo_setup)]
        L.gmo_nterms.restype = i
        L.gmo_synth.argtypes = [P(gmo_setup), vp, vp]
        L.gmo_synth_rf.argtypes = [P(gmo_setup), vp, vp, vp]
        L.gmo_covarm.argtypes = [P(gmo_setup), vp, vp, vp, vp, vp]
        L.gmo_chisqr.argtypes = [P(gmo_setup), vp, vp, i]
        L.gmo_chisqr.restype = d
        L.gmo_svd_solve.argtypes = [vp, vp, i, vp, vp]
        L.gmo_cordic_eig10.argtypes = [vp, i, vp, vp]
        L.gmo_cordic_solve10.argtypes = [vp, vp, vp, vp]
        L.gmo_apply_delta.argtypes = [P(gmo_setup), vp, vp, vp]
        L.gmo_check.argtypes = [vp]
        L.gmo_classical_estimates.argtypes = [P(gmo_setup), vp, vp, i]
        L.gmo_lm_pixel.argtypes = [P(gmo_setup), vp, vp, vp, P(d), P(i), vp, P(i)]
        L.gmo_invert.argtypes = [P(gmo_setup), vp, ctypes.c_long, vp, vp, vp, vp, vp, vp, i]
        L.gmo_synth_batch.argtypes = [P(gmo_setup), vp, ctypes.c_long, vp, i]
        L.gmo_max_threads.restype = i
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def cuanten(qn, fo=1.0):
    z = gmo_zeeman()
    rc = lib().gmo_cuanten(*[float(x) for x in qn], float(fo), ctypes.byref(z))
    if rc != 0:
        raise ValueError("too many Zeeman components")
    return z


def fgauss_wl(fwhm_mA, step, first, central, n):
    out = np.zeros(n)
    lib().gmo_fgauss_wl(fwhm_mA, step, first, central, n, _p(out))
    return out


def psf_interp(dl, val, offsets_mA, offset_mA):
    dl = np.ascontiguousarray(dl, dtype=np.float64)
    val = np.ascontiguousarray(val, dtype=np.float64)
    off = np.ascontiguousarray(offsets_mA, dtype=np.float64)
    out = np.zeros(len(off))
    rc = lib().gmo_psf_interp(_p(dl), _p(val), len(dl), _p(off), len(off), float(offset_mA),
                              _p(out))
    if rc != 0:
        raise ValueError("grid beyond the PSF table")
    return out


def voigt(a, u):
    H, F = ctypes.c_double(), ctypes.c_double()
    lib().gmo_voigt(float(a), float(u), ctypes.byref(H), ctypes.byref(F))
    return H.value, F.value


def nterms(s):
    """NTERMS of the reference kernel: 11 (lm_mils_11) when mac and alfa are
    both free, else 10 (gmilos.cu:1234-1239)."""
    return lib().gmo_nterms(ctypes.byref(s))


def conv_circ(x, h):
    x = np.array(x, dtype=np.float64)
    h = np.ascontiguousarray(h, dtype=np.float64)
    lib().gmo_conv_circ(_p(x), len(x), _p(h))
    return x


def mac_kernels(s, mac):
    G = np.zeros(s.nlambda)
    Gd = np.zeros(s.nlambda)
    lib().gmo_mac_kernels(ctypes.byref(s), float(mac), _p(G), _p(Gd))
    return G, Gd


def set_slight(s, profile):
    """Stray-light profile [4][nlambda] (None switches it off)."""
    if profile is None:
        s.use_slight = 0
        return s
    prof = np.asarray(profile, dtype=np.float64).reshape(4 * s.nlambda)
    s.use_slight = 1
    for k in range(4 * s.nlambda):
        s.slight[k] = float(prof[k])
    return s


def conv(x, h, ic=None):
    x = np.array(x, dtype=np.float64)
    h = np.ascontiguousarray(h, dtype=np.float64)
    if ic is None:
        lib().gmo_conv(_p(x), len(x), _p(h))
    else:
        lib().gmo_conv_ic(_p(x), len(x), _p(h), float(ic))
    return x


def make_setup(lines, offsets_mA, psf=None, init=None, free=None, miter=50, noise=1e-6,
               weights=(1.0, 1.0, 1.0, 1.0), ilambda=0.1, toplim=1e-12, mu=1.0, use_ce=1,
               ce_line=0, rf_reference_rules=1, logclambda=0, init_defaults=None,
               solver=SOLVE_SVD):
    """Oracle set-up from the same ingredients as the product's make_setup:
    grid offsets in mA from lines[0]'s centre (gmilos.cu:171-196), PSF as a
    Gaussian FWHM on the grid (gmilos.cu:244-245) or a (dl, value) table
    (gmilos.cu:247-291)."""
    s = gmo_setup()
    off = np.asarray(offsets_mA, dtype=np.float64)
    n = len(off)
    lam0 = float(lines[0][0])
    steps = np.diff(off)
    if n > 1 and np.allclose(steps, steps[0]):
        lam = [lam0 + off[0] / 1000.0]
        for _ in range(1, n):
            lam.append(lam[-1] + steps[0] / 1000.0)
    else:
        lam = list(lam0 + off / 1000.0)
    s.nlambda = n
    for i, v in enumerate(lam):
        s.lambda_[i] = v
    s.nlines = len(lines)
    for L, (l0, qn, fo) in enumerate(lines):
        s.lambda0[L] = float(l0)
        s.zee[L] = cuanten(qn, fo)
    if psf is None:
        s.use_psf = 0
    else:
        s.use_psf = 1
        if psf[0] == "gauss":
            p = fgauss_wl(psf[1], lam[1] - lam[0], lam[0], lam[n // 2], n)
        elif psf[0] == "table":
            pos = 0
            for i in range(n):
                if pos:
                    break
                if abs(np.trunc(off[i])) == 0:
                    pos = i
            offset = (((n // 2) - pos) * (steps[0] / 1000.0)) * 1000 if pos != n // 2 else 0.0
            p = psf_interp(psf[1], psf[2], off, offset)
        elif psf[0] == "taps":
            p = np.asarray(psf[1], dtype=np.float64)
        else:
            raise ValueError(psf[0])
        for i, v in enumerate(p):
            s.psf[i] = float(v)
    for i in range(4):
        s.weight[i] = float(weights[i])
    s.noise = float(noise)
    s.ilambda = float(ilambda)
    s.toplim = float(toplim)
    s.miter = int(miter)
    s.mu = float(mu)
    s.logclambda = int(logclambda)
    free = free if free is not None else MODEL_FIELDS[:9]
    for name in free:
        s.fix[MODEL_FIELDS.index(name)] = 1
    model = dict(init_defaults or {})
    if init:
        model.update(init)
    for i, f in enumerate(MODEL_FIELDS):
        s.init_model[i] = float(model.get(f, 1.0 if f == "alfa" else 0.0))
    s.use_ce = int(use_ce)
    s.ce_line = int(ce_line)
    s.rf_reference_rules = int(rf_reference_rules)
    s.nterms = 11 if (s.fix[9] and s.fix[10]) else 10
    s.solver = int(solver)
    return s


def psf_taps(s):
    return np.array(s.psf[:s.nlambda])


def synth(s, models):
    models = np.ascontiguousarray(np.atleast_2d(models), dtype=np.float64)
    n = s.nlambda
    out = np.zeros((models.shape[0], 4, n))
    for p in range(models.shape[0]):
        lib().gmo_synth(ctypes.byref(s), _p(models[p]), _p(out[p]))
    return out


def synth_rf(s, models, alpha_rf_mixed=0):
    models = np.ascontiguousarray(np.atleast_2d(models), dtype=np.float64)
    n = s.nlambda
    nt = nterms(s)
    sp = np.zeros((models.shape[0], 4, n))
    rf = np.zeros((models.shape[0], 4, nt, n))
    for p in range(models.shape[0]):
        lib().gmo_synth_rf2(ctypes.byref(s), _p(models[p]), _p(sp[p]), _p(rf[p]),
                            int(alpha_rf_mixed))
    return sp, rf


def synth_batch(s, models, threads=0):
    models = np.ascontiguousarray(models, dtype=np.float64)
    out = np.zeros((models.shape[0], 4, s.nlambda), dtype=np.float32)
    lib().gmo_synth_batch(ctypes.byref(s), _p(models), models.shape[0], _p(out),
                          threads or lib().gmo_max_threads())
    return out


def covarm(s, obs, synth_, rf):
    obs = np.ascontiguousarray(obs, dtype=np.float32)
    synth_ = np.ascontiguousarray(synth_, dtype=np.float64)
    rf = np.ascontiguousarray(rf, dtype=np.float64)
    nt = nterms(s)
    alpha = np.zeros(nt * nt)
    beta = np.zeros(nt)
    lib().gmo_covarm(ctypes.byref(s), _p(obs), _p(synth_), _p(rf), _p(alpha), _p(beta))
    return alpha.reshape(nt, nt), beta


def chisqr(s, obs, synth_, nfree=None):
    obs = np.ascontiguousarray(obs, dtype=np.float32)
    synth_ = np.ascontiguousarray(synth_, dtype=np.float64)
    nfree = nfree if nfree is not None else 4 * s.nlambda - nterms(s)
    return lib().gmo_chisqr(ctypes.byref(s), _p(obs), _p(synth_), nfree)


def cordic_eig10(A, steps):
    """The reference's Brent-Luk CORDIC steps on a symmetric 10 x 10 A:
    (final matrix, V)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    a = np.zeros((10, 10))
    v = np.zeros((10, 10))
    lib().gmo_cordic_eig10(_p(A), int(steps), _p(a), _p(v))
    return a, v


def cordic_solve10(A, b):
    """The reference's 10-term step solve (18 CORDIC steps): (x, w)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(10)
    w = np.zeros(10)
    lib().gmo_cordic_solve10(_p(A), _p(b), _p(x), _p(w))
    return x, w


def svd_solve(A, b):
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    n = len(b)
    x = np.zeros(n)
    w = np.zeros(n)
    lib().gmo_svd_solve(_p(A), _p(b), n, _p(x), _p(w))
    return x, w


def apply_delta(s, model, delta, out=None):
    model = np.ascontiguousarray(model, dtype=np.float64)
    delta = np.ascontiguousarray(delta, dtype=np.float64)
    res = np.array(model if out is None else out, dtype=np.float64)
    lib().gmo_apply_delta(ctypes.byref(s), _p(model), _p(delta), _p(res))
    return res


def check(model):
    m = np.array(model, dtype=np.float64)
    lib().gmo_check(_p(m))
    return m


def classical_estimates(s, spectro, model=None, for_initial_use=1):
    spectro = np.ascontiguousarray(spectro, dtype=np.float32)
    m = np.array(model if model is not None else s.init_model[:], dtype=np.float64)
    lib().gmo_classical_estimates(ctypes.byref(s), _p(spectro), _p(m), int(for_initial_use))
    return m


def lm_pixel(s, obs, init=None):
    obs = np.ascontiguousarray(obs, dtype=np.float32)
    model = np.zeros(NMODEL)
    synth_ = np.zeros(4 * s.nlambda, dtype=np.float32)
    c2 = ctypes.c_double()
    it = ctypes.c_int()
    na = ctypes.c_int()
    ip = None
    if init is not None:
        init = np.ascontiguousarray(init, dtype=np.float64)
        ip = _p(init)
    lib().gmo_lm_pixel(ctypes.byref(s), _p(obs), ip, _p(model), ctypes.byref(c2),
                       ctypes.byref(it), _p(synth_), ctypes.byref(na))
    return dict(model=model, chisqr=c2.value, iters=it.value, naccs=na.value,
                synth=synth_.reshape(4, s.nlambda))


def invert(s, obs, init=None, threads=0, want_synth=True):
    obs = np.ascontiguousarray(obs, dtype=np.float32)
    npix = obs.shape[0]
    models = np.zeros((npix, NMODEL))
    chis = np.zeros(npix, dtype=np.float32)
    iters = np.zeros(npix, dtype=np.int32)
    naccs = np.zeros(npix, dtype=np.int32)
    synth_ = np.zeros_like(obs) if want_synth else None
    ip = None
    if init is not None:
        init = np.ascontiguousarray(init, dtype=np.float64)
        ip = _p(init)
    lib().gmo_invert(ctypes.byref(s), _p(obs), npix, ip, _p(models), _p(chis), _p(iters),
                     _p(synth_) if want_synth else None, _p(naccs),
                     threads or lib().gmo_max_threads())
    return dict(model=models, chisqr=chis, iters=iters, naccs=naccs, synth=synth_)
