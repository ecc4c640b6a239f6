"""MI355X-native G-MILOS inversion engine: Python host mirror of the C ABI.

The reference is driven as ``gmilos <file>.mtrol`` (gmilos.cu:86-1341); its
engine is the per-pixel Levenberg-Marquardt kernel ``lm_mils``
(milosUtils.cu:574-710) fed by constant symbols (gmilos.cu:1184-1198).  This
module mirrors that surface over ``include/gmilos.h``:

* :func:`make_setup` builds the constant set (line, grid, PSF, weights, LM
  constants, free parameters, initial model) from the same ingredients the
  reference reads from LINES / malla.grid / .psf / .mod / .mtrol files;
* :class:`Context` owns a ``gm_ctx`` and runs the inversion or its pieces on
  device buffers (torch tensors are used only as device memory).

The HIP library must be built (``build()``); every entry point raises when it
is missing -- there is no CPU fallback.
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libgmilos.so")

GM_MAX_LAMBDA = 256
GM_MAX_LINES = 4
GM_MAX_COMP = 10
GM_NMODEL = 11
GM_NOALLOC_MAX_PIXELS = 65536

GM_PREC_FP32, GM_PREC_MIXED = 0, 1
GM_SOLVE_SVD, GM_SOLVE_LDL, GM_SOLVE_JACOBI, GM_SOLVE_CORDIC = 0, 1, 2, 3
GM_INIT_CE, GM_INIT_MODEL, GM_INIT_PER_PIXEL = 0, 1, 2

# Init_Model order (defines.h:94-106)
MODEL_FIELDS = ("eta0", "B", "vlos", "dopp", "aa", "gm", "az", "S0", "S1", "mac", "alfa")

STATUS = {0: "ok", 1: "invalid argument", 2: "HIP runtime error", 3: "out of device memory",
          4: "unsupported configuration", 5: "not found", 6: "cancelled"}
GM_OK, GM_ERR_NOT_FOUND, GM_CANCELLED = 0, 5, 6


class GmError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: {STATUS.get(code, code)} ({code})")
        self.code = code


class gm_zeeman(ctypes.Structure):
    _fields_ = [("n_pi", ctypes.c_int32), ("n_sig", ctypes.c_int32),
                ("nub", ctypes.c_double * GM_MAX_COMP), ("nup", ctypes.c_double * GM_MAX_COMP),
                ("nur", ctypes.c_double * GM_MAX_COMP), ("web", ctypes.c_double * GM_MAX_COMP),
                ("wep", ctypes.c_double * GM_MAX_COMP), ("wer", ctypes.c_double * GM_MAX_COMP),
                ("gl", ctypes.c_double), ("gu", ctypes.c_double), ("geff", ctypes.c_double),
                ("fo", ctypes.c_double)]


class gm_setup(ctypes.Structure):
    _fields_ = [("nlambda", ctypes.c_int32), ("lambda_", ctypes.c_double * GM_MAX_LAMBDA),
                ("nlines", ctypes.c_int32), ("lambda0", ctypes.c_double * GM_MAX_LINES),
                ("zee", gm_zeeman * GM_MAX_LINES), ("use_psf", ctypes.c_int32),
                ("psf", ctypes.c_double * GM_MAX_LAMBDA), ("weight", ctypes.c_double * 4),
                ("noise", ctypes.c_double), ("ilambda", ctypes.c_double),
                ("toplim", ctypes.c_double), ("miter", ctypes.c_int32), ("mu", ctypes.c_double),
                ("logclambda", ctypes.c_int32), ("fix", ctypes.c_int32 * 11),
                ("init_model", ctypes.c_double * GM_NMODEL), ("init_mode", ctypes.c_int32),
                ("ce_line", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("solver", ctypes.c_int32), ("rf_reference_rules", ctypes.c_int32),
                ("psf_tail_tol", ctypes.c_double), ("use_slight", ctypes.c_int32),
                ("slight", ctypes.c_float * (4 * GM_MAX_LAMBDA))]


class gm_out(ctypes.Structure):
    _fields_ = [("model", ctypes.c_void_p), ("chisqr", ctypes.c_void_p),
                ("iters", ctypes.c_void_p), ("synth", ctypes.c_void_p),
                ("naccs", ctypes.c_void_p)]


class gm_stats(ctypes.Structure):
    _fields_ = [("npix", ctypes.c_int64), ("synth_evals", ctypes.c_int64),
                ("rf_evals", ctypes.c_int64), ("solves", ctypes.c_int64),
                ("svd_solves", ctypes.c_int64),
                ("lm_rounds", ctypes.c_int32), ("psf_taps", ctypes.c_int32),
                ("ms_total", ctypes.c_float), ("ms_synth", ctypes.c_float),
                ("ms_rf", ctypes.c_float), ("ms_solve", ctypes.c_float),
                ("ms_other", ctypes.c_float)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class gm_server_config(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("pool_pixels", ctypes.c_int64),
                ("max_tiles", ctypes.c_int32), ("device", ctypes.c_int32),
                ("aging_ms", ctypes.c_int32), ("scan", ctypes.c_int32)]


class gm_tile_result(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_uint64), ("setup_id", ctypes.c_int32),
                ("status", ctypes.c_int32), ("npix", ctypes.c_int64),
                ("t_submit_ms", ctypes.c_double), ("t_admit_ms", ctypes.c_double),
                ("t_done_ms", ctypes.c_double)]


class gm_server_stats(ctypes.Structure):
    _fields_ = [("tiles_done", ctypes.c_int64), ("pixels_done", ctypes.c_int64),
                ("lm_rounds", ctypes.c_int64), ("pending_tiles", ctypes.c_int64),
                ("inflight_tiles", ctypes.c_int64), ("inflight_bytes", ctypes.c_int64),
                ("peak_inflight_bytes", ctypes.c_int64)]


class gm_prior(ctypes.Structure):
    _fields_ = [("lo", ctypes.c_double * GM_NMODEL), ("hi", ctypes.c_double * GM_NMODEL),
                ("weak_fraction", ctypes.c_double), ("weak_b_max", ctypes.c_double)]


class gm_guided_stats(ctypes.Structure):
    _fields_ = [("npix", ctypes.c_int64), ("n_fallback", ctypes.c_int64),
                ("n_replaced", ctypes.c_int64), ("iters_learned", ctypes.c_int64),
                ("iters_fallback", ctypes.c_int64), ("ms_predict", ctypes.c_float),
                ("ms_lm", ctypes.c_float), ("ms_fallback", ctypes.c_float),
                ("ms_total", ctypes.c_float)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


EXPORTED = ("gm_ctx_create", "gm_ctx_destroy", "gm_status_string", "gm_ctx_set_timing",
            "gm_last_stats", "gm_invert", "gm_invert_noalloc", "gm_synth", "gm_synth_rf", "gm_normal_equations",
            "gm_step_solve", "gm_classical_estimates", "gm_zeeman_from_quantum",
            "gm_parse_lines_file", "gm_psf_gaussian", "gm_psf_from_table", "gm_setup_defaults",
            "gm_mlp_create", "gm_mlp_destroy", "gm_mlp_num_params", "gm_mlp_get_params",
            "gm_mlp_set_params", "gm_mlp_train_step", "gm_mlp_gradients", "gm_mlp_predict",
            "gm_server_create", "gm_server_add_setup", "gm_server_submit", "gm_server_poll",
            "gm_server_cancel",
            "gm_server_get_stats", "gm_server_clock_ms", "gm_server_destroy",
            "gm_prior_default", "gm_draw_models", "gm_add_noise", "gm_mlp_fit",
            "gm_invert_guided", "gm_ctx_set_stray_map")

_lib = None


def load_library(path=LIB_PATH):
    """Loads libgmilos.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build the HIP library first")
    try:
        # torch ships its own HIP runtime; it has to be the first one loaded
        # so that this library and torch share one runtime in the process
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
    P = ctypes.POINTER
    sig = {
        "gm_ctx_create": ([P(gm_setup), P(vp)], i32),
        "gm_ctx_destroy": ([vp], i32),
        "gm_status_string": ([i32], ctypes.c_char_p),
        "gm_ctx_set_timing": ([vp, i32], i32),
        "gm_ctx_set_stray_map": ([vp, vp, i64], i32),
        "gm_last_stats": ([vp, P(gm_stats)], i32),
        "gm_invert": ([vp, vp, i64, vp, P(gm_out), vp], i32),
        "gm_invert_noalloc": ([vp, vp, i64, vp, P(gm_out), vp], i32),
        "gm_synth": ([vp, vp, i64, vp, vp], i32),
        "gm_synth_rf": ([vp, vp, i64, vp, vp, vp], i32),
        "gm_normal_equations": ([vp, vp, vp, i64, vp, vp, vp, vp], i32),
        "gm_step_solve": ([vp, vp, vp, vp, i64, vp, vp], i32),
        "gm_classical_estimates": ([vp, vp, i64, i32, vp, vp], i32),
        "gm_zeeman_from_quantum": ([ctypes.c_double] * 7 + [P(gm_zeeman)], i32),
        "gm_parse_lines_file": ([ctypes.c_char_p, i32, P(ctypes.c_double),
                                 P(ctypes.c_double)], i32),
        "gm_psf_gaussian": ([ctypes.c_double, P(ctypes.c_double), i32, P(ctypes.c_double)], i32),
        "gm_psf_from_table": ([P(ctypes.c_double), P(ctypes.c_double), i32, P(ctypes.c_double),
                               i32, ctypes.c_double, P(ctypes.c_double)], i32),
        "gm_setup_defaults": ([P(gm_setup)], None),
        "gm_mlp_create": ([i32, ctypes.c_uint64, P(vp)], i32),
        "gm_mlp_destroy": ([vp], i32),
        "gm_mlp_num_params": ([vp], i64),
        "gm_mlp_get_params": ([vp, vp, i64], i32),
        "gm_mlp_set_params": ([vp, vp, i64], i32),
        "gm_mlp_train_step": ([vp, vp, vp, i64, ctypes.c_float, ctypes.c_float,
                               P(ctypes.c_float), vp], i32),
        "gm_mlp_gradients": ([vp, vp, vp, i64, vp, P(ctypes.c_float), vp], i32),
        "gm_mlp_predict": ([vp, vp, i64, vp, vp, vp], i32),
        "gm_server_create": ([P(gm_server_config), P(vp)], i32),
        "gm_server_add_setup": ([vp, P(gm_setup), P(i32)], i32),
        "gm_server_submit": ([vp, i32, vp, i64, vp, P(gm_out), P(ctypes.c_uint64)], i32),
        "gm_server_poll": ([vp, P(gm_tile_result), i32, P(i32), i32], i32),
        "gm_server_cancel": ([vp, ctypes.c_uint64], i32),
        "gm_server_get_stats": ([vp, P(gm_server_stats)], i32),
        "gm_server_clock_ms": ([vp], ctypes.c_double),
        "gm_server_destroy": ([vp], i32),
        "gm_prior_default": ([P(gm_prior)], None),
        "gm_draw_models": ([P(gm_prior), ctypes.c_uint64, i64, i64, vp, vp], i32),
        "gm_add_noise": ([vp, vp, i64, i32, ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                          vp], i32),
        "gm_mlp_fit": ([vp, vp, P(gm_prior), i64, i64, i64, ctypes.c_float, ctypes.c_float,
                        ctypes.c_double, ctypes.c_uint64, vp, vp], i32),
        "gm_invert_guided": ([vp, vp, vp, vp, i64, ctypes.c_double, P(gm_out),
                              P(gm_guided_stats), vp], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _check(code, what):
    if code != 0:
        raise GmError(code, what)


def _dptr(a):
    return (ctypes.c_double * len(a))(*[float(x) for x in a])


# ---------------------------------------------------------------------------
# set-up construction
def zeeman(qn, fo=1.0):
    """Zeeman pattern from (S_lo, L_lo, J_lo, S_up, L_up, J_up)."""
    lib = load_library()
    z = gm_zeeman()
    _check(lib.gm_zeeman_from_quantum(*[float(x) for x in qn], float(fo), ctypes.byref(z)),
           "gm_zeeman_from_quantum")
    return z


def parse_lines(path, index):
    """(lambda0, quantum numbers) of one LINES record (readConfig.c:388-492)."""
    lib = load_library()
    l0 = ctypes.c_double()
    qn = (ctypes.c_double * 6)()
    _check(lib.gm_parse_lines_file(path.encode(), int(index), ctypes.byref(l0), qn),
           "gm_parse_lines_file")
    return l0.value, tuple(qn)


def psf_gaussian(fwhm_mA, lam):
    lib = load_library()
    n = len(lam)
    out = (ctypes.c_double * n)()
    _check(lib.gm_psf_gaussian(float(fwhm_mA), _dptr(lam), n, out), "gm_psf_gaussian")
    return np.array(out[:], dtype=np.float64)


def psf_from_table(dl_mA, val, offsets_mA, step_mA):
    lib = load_library()
    n = len(offsets_mA)
    out = (ctypes.c_double * n)()
    _check(lib.gm_psf_from_table(_dptr(dl_mA), _dptr(val), len(dl_mA), _dptr(offsets_mA), n,
                                 float(step_mA), out), "gm_psf_from_table")
    return np.array(out[:], dtype=np.float64)


# Fixed initial guess of the benchmark workloads (init.mod style).
BENCH_INIT = dict(eta0=8.0, B=400.0, vlos=0.25, dopp=0.034, aa=0.06, gm=30.0, az=30.0,
                  S0=0.15, S1=0.85, mac=0.0, alfa=1.0)
FREE9 = ("eta0", "B", "vlos", "dopp", "aa", "gm", "az", "S0", "S1")


def make_setup(lines, offsets_mA, psf=None, init=None, free=FREE9, miter=50, noise=1e-6,
               weights=(1.0, 1.0, 1.0, 1.0), ilambda=0.1, toplim=1e-12, mu=1.0,
               precision=GM_PREC_MIXED, solver=GM_SOLVE_SVD, init_mode=GM_INIT_CE,
               ce_line=0, rf_reference_rules=1, psf_tail_tol=1e-9, logclambda=0,
               slight=None):
    """Builds a gm_setup.

    lines: list of (lambda0 [A], (S_lo, L_lo, J_lo, S_up, L_up, J_up), fo).
    offsets_mA: grid offsets from lines[0]'s centre in mA (malla.grid style).
    psf: None, ("gauss", fwhm_mA) or ("table", dl_mA, values).
    slight: optional stray-light profile [4][nlambda] mixed with the
    initial model's filling factor alfa (init=dict(alfa=...)).
    """
    lib = load_library()
    s = gm_setup()
    lib.gm_setup_defaults(ctypes.byref(s))
    offsets = np.asarray(offsets_mA, dtype=np.float64)
    n = len(offsets)
    if n > GM_MAX_LAMBDA or len(lines) > GM_MAX_LINES:
        raise ValueError("grid or line count beyond the engine's limits")
    lam0 = float(lines[0][0])
    # absolute grid as gmilos.cu:190-195 (cumulative steps from the first
    # offset) when the grid is uniform, else offsets added directly
    steps = np.diff(offsets)
    if n > 1 and np.allclose(steps, steps[0]):
        lam = [lam0 + offsets[0] / 1000.0]
        for _ in range(1, n):
            lam.append(lam[-1] + steps[0] / 1000.0)
    else:
        lam = list(lam0 + offsets / 1000.0)
    s.nlambda = n
    for i, v in enumerate(lam):
        s.lambda_[i] = v
    s.nlines = len(lines)
    for L, (l0, qn, fo) in enumerate(lines):
        s.lambda0[L] = float(l0)
        s.zee[L] = zeeman(qn, fo)
    if psf is None:
        s.use_psf = 0
    else:
        s.use_psf = 1
        if psf[0] == "gauss":
            p = psf_gaussian(psf[1], lam)
        elif psf[0] == "table":
            p = psf_from_table(psf[1], psf[2], offsets, float(steps[0]) if n > 1 else 1.0)
        elif psf[0] == "taps":
            p = np.asarray(psf[1], dtype=np.float64)
        else:
            raise ValueError(f"unknown psf kind {psf[0]}")
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
    for i in range(11):
        s.fix[i] = 0
    for name in free:
        s.fix[MODEL_FIELDS.index(name)] = 1
    model = dict(BENCH_INIT)
    if init:
        model.update(init)
    for i, f in enumerate(MODEL_FIELDS):
        s.init_model[i] = float(model[f])
    s.init_mode = int(init_mode)
    s.ce_line = int(ce_line)
    s.precision = int(precision)
    s.solver = int(solver)
    s.rf_reference_rules = int(rf_reference_rules)
    s.psf_tail_tol = float(psf_tail_tol)
    if slight is not None:
        prof = np.asarray(slight, dtype=np.float32).reshape(4 * n)
        s.use_slight = 1
        for k in range(4 * n):
            s.slight[k] = float(prof[k])
    return s


# Workload descriptions (synthetic benchmark and test configurations) ------
FE6173 = (6173.3356, (2, 1, 1.0, 2, 2, 0.0), 1.0)     # 5P1-5D0, g_eff 2.5
FE6301 = (6301.515, (2, 1, 2.0, 2, 2, 2.0), 1.0)      # 5P2-5D2, g_eff 1.67
FE6302 = (6302.494, (2, 1, 1.0, 2, 2, 0.0), 1.0)      # 5P1-5D0, g_eff 2.5
# in a 6301.5 + 6302.5 pair the 6302.5 strength relative to 6301.5 follows
# their log(gf) difference in run/LINES (-1.16 vs -0.59): 10**-0.57
FE6302_PAIR = (6302.494, (2, 1, 1.0, 2, 2, 0.0), 0.27)
GRID_6173 = np.arange(-350.0, 665.0 + 1e-9, 35.0)       # 30 samples
GRID_6302 = -716.0 + 21.5 * np.arange(112)              # 112 samples over both lines


def setup_single_line(psf_fwhm=None, **kw):
    """Fe I 6173.3356 A, -350:35:665 mA, optional Gaussian PSF."""
    psf = ("gauss", psf_fwhm) if psf_fwhm else None
    return make_setup([FE6173], GRID_6173, psf=psf, **kw)


def setup_two_lines(psf_fwhm=None, **kw):
    """Fe I 6301.5 + 6302.5 A on 112 samples of 21.5 mA; classical
    estimates use the g_eff = 2.5 line."""
    psf = ("gauss", psf_fwhm) if psf_fwhm else None
    kw.setdefault("ce_line", 1)
    return make_setup([FE6301, FE6302_PAIR], GRID_6302, psf=psf, **kw)


def random_models(n, seed, weak_fraction=0.0):
    """Per-pixel true models drawn uniformly: B 0-3000 G, gamma/phi 0-180 deg,
    vlos -3..3 km/s, dopp 20-60 mA, a 0.01-0.5, eta0 1-20, S0 0.1-0.4,
    S1 0.5-0.9, mac 0, alpha 1; a fraction of pixels with B < 100 G."""
    rng = np.random.default_rng(seed)
    m = np.zeros((n, GM_NMODEL), dtype=np.float64)
    m[:, 0] = rng.uniform(1.0, 20.0, n)
    m[:, 1] = rng.uniform(0.0, 3000.0, n)
    m[:, 2] = rng.uniform(-3.0, 3.0, n)
    m[:, 3] = rng.uniform(0.020, 0.060, n)
    m[:, 4] = rng.uniform(0.01, 0.5, n)
    m[:, 5] = rng.uniform(0.0, 180.0, n)
    m[:, 6] = rng.uniform(0.0, 180.0, n)
    m[:, 7] = rng.uniform(0.1, 0.4, n)
    m[:, 8] = rng.uniform(0.5, 0.9, n)
    m[:, 9] = 0.0
    m[:, 10] = 1.0
    if weak_fraction > 0:
        weak = rng.random(n) < weak_fraction
        m[weak, 1] = rng.uniform(0.0, 100.0, int(weak.sum()))
    return m


def noise_sigma(models, sigma=1e-3, seed=0, high_fraction=0.0, high_sigma=3e-3):
    """Per-pixel noise sigma in units of Ic = S0 + S1 (mu = 1)."""
    rng = np.random.default_rng(seed + 7919)
    sig = np.full(models.shape[0], sigma)
    if high_fraction > 0:
        sig[rng.random(models.shape[0]) < high_fraction] = high_sigma
    return sig * (models[:, 7] + models[:, 8])


# ---------------------------------------------------------------------------
class Context:
    """Owns one gm_ctx.  Buffers are torch tensors on the current device."""

    def __init__(self, setup):
        self.lib = load_library()
        self.setup = setup
        h = ctypes.c_void_p()
        _check(self.lib.gm_ctx_create(ctypes.byref(setup), ctypes.byref(h)), "gm_ctx_create")
        self.h = h
        self.nlambda = setup.nlambda
        # a free filling factor solves eleven slots (mac, alfa in 9, 10); the
        # reference layout of the responses has NTERMS = 11 only when mac is
        # free too (lm_mils_11), else 10
        self.np = 11 if setup.fix[10] else 9
        self.nterms = 11 if (setup.fix[9] and setup.fix[10]) else 10

    def close(self):
        if self.h:
            self.lib.gm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _stream():
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def set_stray_map(self, stray_map):
        """Per-pixel stray light: float32 [npix, 4, nlambda] CUDA tensor
        (kept referenced here), or None for the set-up's profile."""
        if stray_map is None:
            _check(self.lib.gm_ctx_set_stray_map(self.h, None, 0), "gm_ctx_set_stray_map")
            self._stray_map = None
            return
        import torch
        assert stray_map.dtype == torch.float32 and stray_map.is_contiguous()
        assert stray_map.shape[1:] == (4, self.nlambda)
        _check(self.lib.gm_ctx_set_stray_map(self.h, ctypes.c_void_p(stray_map.data_ptr()),
                                             stray_map.shape[0]), "gm_ctx_set_stray_map")
        self._stray_map = stray_map

    def set_timing(self, on):
        _check(self.lib.gm_ctx_set_timing(self.h, 1 if on else 0), "gm_ctx_set_timing")

    def stats(self):
        st = gm_stats()
        _check(self.lib.gm_last_stats(self.h, ctypes.byref(st)), "gm_last_stats")
        return st.as_dict()

    def invert_noalloc(self, obs, out, init=None):
        """gm_invert_noalloc into caller-allocated outputs (a dict as
        returned by invert); npix at most GM_NOALLOC_MAX_PIXELS."""
        return self.invert(obs, init=init, out=out, noalloc=True)

    def invert(self, obs, init=None, want_synth=True, out=None, noalloc=False):
        """obs: float32 [npix, 4, nlambda] CUDA tensor.  Returns a dict of
        tensors: model [npix, 11] f64, chisqr, iters, naccs, synth."""
        import torch
        npix = obs.shape[0]
        dev = obs.device
        if out is None:
            out = {
                "model": torch.empty((npix, GM_NMODEL), dtype=torch.float64, device=dev),
                "chisqr": torch.empty(npix, dtype=torch.float32, device=dev),
                "iters": torch.empty(npix, dtype=torch.int32, device=dev),
                "naccs": torch.empty(npix, dtype=torch.int32, device=dev),
            }
            if want_synth:
                out["synth"] = torch.empty_like(obs)
        o = gm_out(out["model"].data_ptr(), out["chisqr"].data_ptr(), out["iters"].data_ptr(),
                   out["synth"].data_ptr() if out.get("synth") is not None else None,
                   out["naccs"].data_ptr())
        assert obs.dtype == torch.float32 and obs.is_contiguous()
        assert obs.shape[1] == 4 and obs.shape[2] == self.nlambda
        ip = None
        if init is not None:
            assert init.dtype == torch.float64 and init.shape == (npix, GM_NMODEL)
            ip = ctypes.c_void_p(init.data_ptr())
        fn = self.lib.gm_invert_noalloc if noalloc else self.lib.gm_invert
        _check(fn(self.h, ctypes.c_void_p(obs.data_ptr()), npix, ip, ctypes.byref(o),
                  self._stream()), "gm_invert_noalloc" if noalloc else "gm_invert")
        return out

    def synth(self, models):
        import torch
        npix = models.shape[0]
        sp = torch.empty((npix, 4, self.nlambda), dtype=torch.float32, device=models.device)
        _check(self.lib.gm_synth(self.h, ctypes.c_void_p(models.data_ptr()), npix,
                                 ctypes.c_void_p(sp.data_ptr()), self._stream()), "gm_synth")
        return sp

    def synth_rf(self, models):
        import torch
        npix = models.shape[0]
        sp = torch.empty((npix, 4, self.nlambda), dtype=torch.float32, device=models.device)
        rf = torch.empty((npix, 4, self.nterms, self.nlambda), dtype=torch.float32,
                         device=models.device)
        _check(self.lib.gm_synth_rf(self.h, ctypes.c_void_p(models.data_ptr()), npix,
                                    ctypes.c_void_p(sp.data_ptr()), ctypes.c_void_p(rf.data_ptr()),
                                    self._stream()), "gm_synth_rf")
        return sp, rf

    def normal_equations(self, models, obs):
        import torch
        npix = models.shape[0]
        dev = models.device
        na = self.np * (self.np + 1) // 2
        a = torch.empty((npix, na), dtype=torch.float32, device=dev)
        b = torch.empty((npix, self.np), dtype=torch.float32, device=dev)
        c = torch.empty(npix, dtype=torch.float64, device=dev)
        _check(self.lib.gm_normal_equations(self.h, ctypes.c_void_p(models.data_ptr()),
                                            ctypes.c_void_p(obs.data_ptr()), npix,
                                            ctypes.c_void_p(a.data_ptr()),
                                            ctypes.c_void_p(b.data_ptr()),
                                            ctypes.c_void_p(c.data_ptr()), self._stream()),
               "gm_normal_equations")
        return a, b, c

    def step_solve(self, alpha, beta, flambda):
        import torch
        npix = alpha.shape[0]
        d = torch.empty((npix, self.np), dtype=torch.float64, device=alpha.device)
        _check(self.lib.gm_step_solve(self.h, ctypes.c_void_p(alpha.data_ptr()),
                                      ctypes.c_void_p(beta.data_ptr()),
                                      ctypes.c_void_p(flambda.data_ptr()), npix,
                                      ctypes.c_void_p(d.data_ptr()), self._stream()),
               "gm_step_solve")
        return d

    def classical_estimates(self, obs, for_initial_use=True):
        import torch
        npix = obs.shape[0]
        m = torch.empty((npix, GM_NMODEL), dtype=torch.float64, device=obs.device)
        _check(self.lib.gm_classical_estimates(self.h, ctypes.c_void_p(obs.data_ptr()), npix,
                                               1 if for_initial_use else 0,
                                               ctypes.c_void_p(m.data_ptr()), self._stream()),
               "gm_classical_estimates")
        return m


def default_prior(weak_fraction=0.0):
    pr = gm_prior()
    load_library().gm_prior_default(ctypes.byref(pr))
    pr.weak_fraction = float(weak_fraction)
    return pr


def draw_models(npix, seed, prior=None, offset=0, device="cuda"):
    """Device-drawn true models [npix, 11] (float64 tensor)."""
    import torch
    pr = prior if prior is not None else default_prior()
    m = torch.empty((npix, GM_NMODEL), dtype=torch.float64, device=device)
    _check(load_library().gm_draw_models(ctypes.byref(pr), int(seed), int(offset), int(npix),
                                         ctypes.c_void_p(m.data_ptr()), Context._stream()),
           "gm_draw_models")
    return m


def add_noise(models, spectra, sigma, seed, mu=1.0):
    """In place: spectra += N(0, (sigma Ic)^2), Ic = S0 + S1 mu."""
    _check(load_library().gm_add_noise(ctypes.c_void_p(models.data_ptr()),
                                       ctypes.c_void_p(spectra.data_ptr()), spectra.shape[0],
                                       spectra.shape[2], float(sigma), float(mu), int(seed),
                                       Context._stream()), "gm_add_noise")
    return spectra


def invert_guided(learned_ctx, plain_ctx, net, obs, fallback_chisqr, want_synth=True):
    """Learned start + LM with the fixed-guess fallback; returns (out, stats)."""
    import torch
    npix, dev = obs.shape[0], obs.device
    out = {"model": torch.empty((npix, GM_NMODEL), dtype=torch.float64, device=dev),
           "chisqr": torch.empty(npix, dtype=torch.float32, device=dev),
           "iters": torch.empty(npix, dtype=torch.int32, device=dev),
           "naccs": torch.empty(npix, dtype=torch.int32, device=dev)}
    if want_synth:
        out["synth"] = torch.empty_like(obs)
    o = gm_out(out["model"].data_ptr(), out["chisqr"].data_ptr(), out["iters"].data_ptr(),
               out["synth"].data_ptr() if want_synth else None, out["naccs"].data_ptr())
    st = gm_guided_stats()
    _check(load_library().gm_invert_guided(learned_ctx.h, plain_ctx.h, net.h,
                                           ctypes.c_void_p(obs.data_ptr()), npix,
                                           float(fallback_chisqr), ctypes.byref(o),
                                           ctypes.byref(st), Context._stream()),
           "gm_invert_guided")
    return out, st.as_dict()


class MLP:
    """Learned initial guess (BF16 MFMA network); device buffers are torch
    tensors on the current device."""

    def __init__(self, nlambda, seed=0):
        self.lib = load_library()
        h = ctypes.c_void_p()
        _check(self.lib.gm_mlp_create(int(nlambda), int(seed), ctypes.byref(h)), "gm_mlp_create")
        self.h = h
        self.nlambda = nlambda
        self.nparam = self.lib.gm_mlp_num_params(h)

    def close(self):
        if self.h:
            self.lib.gm_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.nparam, dtype=np.float32)
        _check(self.lib.gm_mlp_get_params(self.h, out.ctypes.data_as(ctypes.c_void_p),
                                          self.nparam), "gm_mlp_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(self.lib.gm_mlp_set_params(self.h, p.ctypes.data_as(ctypes.c_void_p), p.size),
               "gm_mlp_set_params")

    def train_step(self, obs, models, lr, weight_decay=1e-4):
        loss = ctypes.c_float()
        _check(self.lib.gm_mlp_train_step(self.h, ctypes.c_void_p(obs.data_ptr()),
                                          ctypes.c_void_p(models.data_ptr()), obs.shape[0],
                                          float(lr), float(weight_decay), ctypes.byref(loss),
                                          Context._stream()), "gm_mlp_train_step")
        return loss.value

    def gradients(self, obs, models):
        g = np.zeros(self.nparam, dtype=np.float32)
        loss = ctypes.c_float()
        _check(self.lib.gm_mlp_gradients(self.h, ctypes.c_void_p(obs.data_ptr()),
                                         ctypes.c_void_p(models.data_ptr()), obs.shape[0],
                                         g.ctypes.data_as(ctypes.c_void_p), ctypes.byref(loss),
                                         Context._stream()), "gm_mlp_gradients")
        return loss.value, g

    def fit(self, ctx, n_train, batch=4096, steps=2000, lr=2e-3, weight_decay=1e-4,
            sigma=1e-3, seed=1, prior=None, curve=True):
        """Synthesise n_train profiles with ctx and train; returns the loss
        curve (numpy) when curve is set."""
        pr = prior if prior is not None else default_prior()
        lc = np.zeros(steps, dtype=np.float32) if curve else None
        _check(self.lib.gm_mlp_fit(self.h, ctx.h, ctypes.byref(pr), int(n_train), int(batch),
                                   int(steps), float(lr), float(weight_decay), float(sigma),
                                   int(seed),
                                   lc.ctypes.data_as(ctypes.c_void_p) if curve else None,
                                   Context._stream()), "gm_mlp_fit")
        return lc

    def predict(self, obs, base):
        import torch
        npix = obs.shape[0]
        m = torch.empty((npix, GM_NMODEL), dtype=torch.float64, device=obs.device)
        b = np.ascontiguousarray(base, dtype=np.float64)
        _check(self.lib.gm_mlp_predict(self.h, ctypes.c_void_p(obs.data_ptr()), npix,
                                       b.ctypes.data_as(ctypes.c_void_p),
                                       ctypes.c_void_p(m.data_ptr()), Context._stream()),
               "gm_mlp_predict")
        return m

    def layout(self):
        """[(in, out, w_offset, b_offset)] of the flat parameter vector."""
        kin = ((4 * self.nlambda + 63) // 64) * 64
        dims = [kin, 256, 256, 256, 64]
        out, off = [], 0
        for l in range(4):
            wo = off
            off += dims[l] * dims[l + 1]
            out.append((dims[l], dims[l + 1], wo, off))
            off += dims[l + 1]
        return out


class Server:
    """Serving loop (gm_server): tiles of any size for registered setups are
    submitted as they arrive and complete independently.  The caller keeps a
    tile's device buffers alive until its result is polled; this wrapper holds
    references to the tensors for that purpose."""

    def __init__(self, hbm_budget_bytes=0, pool_pixels=0, max_tiles=0, device=0, aging_ms=0,
                 scan=0):
        self.lib = load_library()
        cfg = gm_server_config(int(hbm_budget_bytes), int(pool_pixels), int(max_tiles), int(device),
                               int(aging_ms), int(scan))
        h = ctypes.c_void_p()
        _check(self.lib.gm_server_create(ctypes.byref(cfg), ctypes.byref(h)), "gm_server_create")
        self.h = h
        self.nlambda = {}
        self.keep = {}

    def add_setup(self, setup):
        sid = ctypes.c_int32()
        _check(self.lib.gm_server_add_setup(self.h, ctypes.byref(setup), ctypes.byref(sid)),
               "gm_server_add_setup")
        self.nlambda[sid.value] = setup.nlambda
        return sid.value

    def submit(self, setup_id, obs, init=None, want_synth=True):
        """obs: float32 [npix, 4, nlambda] device tensor, already written (the
        producer stream synchronised).  Returns (ticket, outputs dict)."""
        import torch
        assert obs.dtype == torch.float32 and obs.is_contiguous()
        assert obs.shape[1] == 4 and obs.shape[2] == self.nlambda[setup_id]
        npix, dev = obs.shape[0], obs.device
        out = {"model": torch.empty((npix, GM_NMODEL), dtype=torch.float64, device=dev),
               "chisqr": torch.empty(npix, dtype=torch.float32, device=dev),
               "iters": torch.empty(npix, dtype=torch.int32, device=dev),
               "naccs": torch.empty(npix, dtype=torch.int32, device=dev)}
        if want_synth:
            out["synth"] = torch.empty_like(obs)
        torch.cuda.current_stream(dev).synchronize()      # obs written, outputs allocated
        o = gm_out(out["model"].data_ptr(), out["chisqr"].data_ptr(), out["iters"].data_ptr(),
                   out["synth"].data_ptr() if want_synth else None, out["naccs"].data_ptr())
        t = ctypes.c_uint64()
        ip = ctypes.c_void_p(init.data_ptr()) if init is not None else None
        _check(self.lib.gm_server_submit(self.h, int(setup_id), ctypes.c_void_p(obs.data_ptr()),
                                         npix, ip, ctypes.byref(o), ctypes.byref(t)),
               "gm_server_submit")
        self.keep[t.value] = (obs, init, out)
        return t.value, out

    def poll(self, max_results=256, wait_ms=0):
        """Completed tiles as dicts (ticket, setup_id, status, npix, latency_ms,
        queue_ms, outputs)."""
        buf = (gm_tile_result * max_results)()
        n = ctypes.c_int32()
        _check(self.lib.gm_server_poll(self.h, buf, max_results, ctypes.byref(n), int(wait_ms)),
               "gm_server_poll")
        res = []
        for k in range(n.value):
            r = buf[k]
            kept = self.keep.pop(r.ticket, (None, None, None))
            res.append({"ticket": r.ticket, "setup_id": r.setup_id, "status": r.status,
                        "npix": r.npix, "t_submit_ms": r.t_submit_ms,
                        "latency_ms": r.t_done_ms - r.t_submit_ms,
                        "queue_ms": r.t_admit_ms - r.t_submit_ms, "out": kept[2]})
        return res

    def cancel(self, ticket):
        """True when the tile was queued or in flight (its result then comes
        back through poll with status GM_CANCELLED), False when it is unknown
        or already complete."""
        code = self.lib.gm_server_cancel(self.h, ctypes.c_uint64(ticket))
        if code == GM_ERR_NOT_FOUND:
            return False
        _check(code, "gm_server_cancel")
        return True

    def clock_ms(self):
        return self.lib.gm_server_clock_ms(self.h)

    def stats(self):
        st = gm_server_stats()
        _check(self.lib.gm_server_get_stats(self.h, ctypes.byref(st)), "gm_server_get_stats")
        return {f: getattr(st, f) for f, _ in gm_server_stats._fields_}

    def close(self):
        if self.h:
            self.lib.gm_server_destroy(self.h)
            self.h = None
            self.keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def packed_index(p, q, n=9):
    """Index of (p, q) in the packed n x n upper triangle (n = 9 or 11)."""
    if p > q:
        p, q = q, p
    return p * n - (p * (p - 1)) // 2 + (q - p)


def unpack_alpha(a):
    """[..., 45] or [..., 66] packed upper triangle -> [..., n, n] symmetric."""
    a = np.asarray(a)
    n = 11 if a.shape[-1] == 66 else 9
    out = np.zeros(a.shape[:-1] + (n, n), dtype=a.dtype)
    for p in range(n):
        for q in range(p, n):
            out[..., p, q] = a[..., packed_index(p, q, n)]
            out[..., q, p] = a[..., packed_index(p, q, n)]
    return out


def build(verbose=False):
    """Compiles libgmilos.so (and the CLI when its sources exist)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("gmilos_amd_build",
                                                  os.path.join(HERE, "build.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.build(verbose)
