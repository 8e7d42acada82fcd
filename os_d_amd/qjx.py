"""ctypes mirror of include/qjx.h.

The reference drives one ensemble through strategy objects (NewDelBehavior,
ExpNewDelBehavior, ExperimentBehavior, PropagateBehavior, CoreBehavior; see
source/cpp/QJX/QJX/processor.cpp:50-75).  `Engine` exposes the same steps at
ensemble granularity: build the model and propagator tree, initialise the
ensemble, transient, observation periods, observables, jump logs.

The library must be built (os_d_amd/build.py); there is no CPU fallback: a
missing or unloadable libqjx.so raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libqjx.so")

FP64, FP32, BF16X3 = 0, 1, 2

QJX_OK = 0
ERRORS = {-1: "QJX_ERR_ARG", -2: "QJX_ERR_HIP", -3: "QJX_ERR_NORM", -4: "QJX_ERR_LOG",
          -5: "QJX_ERR_IO", -6: "QJX_ERR_NOMEM"}


class QJXError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


class ModelDesc(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("prm_E", ctypes.c_double), ("prm_U", ctypes.c_double),
                ("prm_J", ctypes.c_double), ("drv_ampl", ctypes.c_double),
                ("drv_freq", ctypes.c_double), ("diss_gamma", ctypes.c_double),
                ("ex_deep", ctypes.c_int), ("steps_per_period", ctypes.c_int),
                ("precision", ctypes.c_int)]


class RkDesc(ctypes.Structure):
    _fields_ = [("rk_ns", ctypes.c_int), ("drv_type", ctypes.c_int), ("drv_phase", ctypes.c_double)]


class LpnDesc(ctypes.Structure):
    _fields_ = [("lpn_type", ctypes.c_int), ("lpn_eps_deep", ctypes.c_int), ("lpn_eps_error", ctypes.c_double),
                ("lpn_eps_high", ctypes.c_double), ("lpn_eps_low", ctypes.c_double),
                ("lpn_delta_s", ctypes.c_double), ("lpn_delta_f_high", ctypes.c_double),
                ("lpn_delta_f_low", ctypes.c_double), ("lambda_per_periods", ctypes.c_int),
                ("save_lambdas", ctypes.c_int), ("num_random_obs", ctypes.c_int),
                ("random_obs_seed", ctypes.c_int), ("random_obs_mns", ctypes.c_int),
                ("random_obs_type", ctypes.c_int)]


class EnsembleDesc(ctypes.Structure):
    _fields_ = [("num_trajectories", ctypes.c_int), ("seed", ctypes.c_int), ("mns", ctypes.c_int),
                ("start_type", ctypes.c_int), ("start_state", ctypes.c_int), ("jump", ctypes.c_int),
                ("log_capacity", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("matvecs", ctypes.c_int64), ("iterations", ctypes.c_int64),
                ("jumps", ctypes.c_int64), ("gemm_flops", ctypes.c_double),
                ("gemm_ms", ctypes.c_double), ("decide_ms", ctypes.c_double),
                ("gemm_launches", ctypes.c_int64), ("gemm_bn", ctypes.c_int32),
                ("gemm_grid", ctypes.c_int32), ("gemm_band_flops", ctypes.c_double),
                ("gemm_waves", ctypes.c_int32), ("pad", ctypes.c_int32)]


def _stats_dict(s):
    return {k: getattr(s, k) for k, _ in Stats._fields_}


class JobResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("num_trajectories", ctypes.c_int), ("n_periods", ctypes.c_int),
                ("dim", ctypes.c_int), ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_done", ctypes.c_double), ("matvecs", ctypes.c_int64), ("jumps", ctypes.c_int64),
                ("mean_avg", ctypes.c_double), ("mean_std", ctypes.c_double),
                ("energy_avg", ctypes.c_double), ("energy_std", ctypes.c_double),
                ("norm_avg", ctypes.c_double)]


class LindbladDrive(ctypes.Structure):
    _fields_ = [("drv_type", ctypes.c_int), ("drv_phase", ctypes.c_double), ("per_stage", ctypes.c_int)]


class PoolUsage(ctypes.Structure):
    _fields_ = [("used_bytes", ctypes.c_int64), ("budget_bytes", ctypes.c_int64), ("trajectories", ctypes.c_int64),
                ("capacity", ctypes.c_int64), ("queued", ctypes.c_int32), ("building", ctypes.c_int32),
                ("running", ctypes.c_int32), ("done", ctypes.c_int32), ("failed", ctypes.c_int32),
                ("pad", ctypes.c_int32)]


_lib = None

# (name, restype, argtypes) of every symbol declared in include/qjx.h
_P = ctypes.POINTER
_D = _P(ctypes.c_double)
SIGNATURES = [
    ("qjx_last_error", ctypes.c_char_p, []),
    ("qjx_ctx_create", ctypes.c_int, [_P(ModelDesc), _P(ctypes.c_void_p)]),
    ("qjx_ctx_create_rk", ctypes.c_int, [_P(ModelDesc), _P(RkDesc), _P(ctypes.c_void_p)]),
    ("qjx_ctx_destroy", None, [ctypes.c_void_p]),
    ("qjx_ensemble_init", ctypes.c_int, [ctypes.c_void_p, _P(EnsembleDesc)]),
    ("qjx_trans", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("qjx_run_periods", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    ("qjx_run_steps", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    ("qjx_chars", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D, _D, _D, _D, _D, _D, _D]),
    ("qjx_density", ctypes.c_int, [ctypes.c_void_p, _D, _D]),
    ("qjx_density_noalloc", ctypes.c_int, [ctypes.c_void_p, _D, _D]),
    ("qjx_random_obs", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 4 + [_D]),
    ("qjx_get_state", ctypes.c_int, [ctypes.c_void_p, _D, _D, _D, _P(ctypes.c_int),
                                     _P(ctypes.c_uint64)]),
    ("qjx_set_state", ctypes.c_int, [ctypes.c_void_p, _D, _D, _D, _P(ctypes.c_uint64)]),
    ("qjx_get_jumps", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D, _P(ctypes.c_int), _D, _D,
                                     ctypes.c_int]),
    ("qjx_jump_interval_pdf", ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int, _D,
                                             _P(ctypes.c_int64), _P(ctypes.c_int64)]),
    ("qjx_get_propagator", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _D]),
    ("qjx_set_propagator", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _D]),
    ("qjx_get_stats", ctypes.c_int, [ctypes.c_void_p, _P(Stats)]),
    ("qjx_dim", ctypes.c_int, [ctypes.c_void_p]),
    ("qjx_peak_probe", ctypes.c_int, [_D, _D, _D, _D, _D]),
    ("qjx_set_profiling", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("qjx_mfma_probe_f32x16", ctypes.c_int, [_P(ctypes.c_float), _P(ctypes.c_float), _P(ctypes.c_float)]),
    ("qjx_mfma_probe", ctypes.c_int, [_P(ctypes.c_float), _P(ctypes.c_float), _P(ctypes.c_float), _D, _D,
                                      _D, _P(ctypes.c_float), _P(ctypes.c_float), _P(ctypes.c_float)]),
    ("qjx_run_config", ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]),
    ("qjx_batch_create", ctypes.c_int, [_P(ctypes.c_void_p), ctypes.c_int, _P(ctypes.c_void_p)]),
    ("qjx_batch_destroy", None, [ctypes.c_void_p]),
    ("qjx_batch_run_steps", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    ("qjx_batch_run_periods", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    ("qjx_batch_get_stats", ctypes.c_int, [ctypes.c_void_p, _P(Stats)]),
    ("qjx_batch_set_profiling", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("qjx_lindblad_create", ctypes.c_int, [_P(ModelDesc), ctypes.c_int, _P(ctypes.c_void_p)]),
    ("qjx_lindblad_destroy", None, [ctypes.c_void_p]),
    ("qjx_lindblad_set_rho", ctypes.c_int, [ctypes.c_void_p, _D]),
    ("qjx_lindblad_get_rho", ctypes.c_int, [ctypes.c_void_p, _D]),
    ("qjx_lindblad_run_periods", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("qjx_lindblad_observables", ctypes.c_int, [ctypes.c_void_p, _D, _D, _D, _D, _D]),
    ("qjx_lindblad_dim", ctypes.c_int, [ctypes.c_void_p]),
    ("qjx_lindblad_create_drive", ctypes.c_int, [_P(ModelDesc), ctypes.c_int, _P(LindbladDrive),
                                                 _P(ctypes.c_void_p)]),
    ("qjx_lindblad_save_rho", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    ("qjx_lpn_create", ctypes.c_int, [ctypes.c_void_p, _P(LpnDesc), _P(ctypes.c_void_p)]),
    ("qjx_lpn_destroy", None, [ctypes.c_void_p]),
    ("qjx_lpn_trans", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("qjx_lpn_run_periods", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("qjx_lpn_get", ctypes.c_int, [ctypes.c_void_p, _D, _D, _P(ctypes.c_int), _D, _D, _D]),
    ("qjx_lpn_random_obs", ctypes.c_int, [ctypes.c_void_p, _D, _D]),
    ("qjx_lpn_lambdas", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D, _D, _D, ctypes.c_int]),
    ("qjx_lpn_random_obs_matrix", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D]),
    ("qjx_lpn_chars", ctypes.c_int, [ctypes.c_void_p, _D, _D, _D, _D, _D, _D, ctypes.c_int, ctypes.c_int, _D]),
    ("qjx_pool_create", ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_int, _P(ctypes.c_void_p)]),
    ("qjx_pool_destroy", None, [ctypes.c_void_p]),
    ("qjx_pool_submit", ctypes.c_int, [ctypes.c_void_p, _P(ModelDesc), _P(EnsembleDesc), ctypes.c_int,
                                       _P(ctypes.c_int)]),
    ("qjx_pool_step", ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_int), _P(ctypes.c_int)]),
    ("qjx_pool_result", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _P(JobResult), _D, _D]),
    ("qjx_pool_clock", ctypes.c_double, [ctypes.c_void_p]),
    ("qjx_pool_get_usage", ctypes.c_int, [ctypes.c_void_p, _P(PoolUsage)]),
]


def load_library(path=LIB_PATH):
    """Load libqjx.so and bind every C-ABI symbol (raises if absent)."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise QJXError(-2, "libqjx.so not built at %s (run os_d_amd/build.py)" % path)
    lib = ctypes.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib


def _check(lib, rc):
    if rc < 0:
        raise QJXError(rc, lib.qjx_last_error().decode())
    return rc


def _ptr(a, ctype=ctypes.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctype))


class Engine:
    """One dimer model + one trajectory ensemble on the current HIP device.

    prop="qj" (prop_id 0) propagates with the bisection tree of exp(-i H_eff dt);
    prop="rk" (prop_id 1) takes rk_ns RK4 steps per period with drive type drv_type
    (1: E + A sin(omega t + drv_phase), else the square drive), FP64 only."""

    def __init__(self, N, prm_E=1.0, prm_U=0.5, prm_J=1.0, drv_ampl=1.5, drv_freq=1.0,
                 diss_gamma=0.1, ex_deep=16, steps_per_period=4, precision=FP64, prop="qj",
                 rk_ns=1000, drv_type=0, drv_phase=0.0):
        self.lib = load_library()
        self.model = ModelDesc(N, prm_E, prm_U, prm_J, drv_ampl, drv_freq, diss_gamma, ex_deep,
                               steps_per_period, precision)
        h = ctypes.c_void_p()
        if prop == "rk":
            self.rk = RkDesc(rk_ns, drv_type, drv_phase)
            _check(self.lib, self.lib.qjx_ctx_create_rk(ctypes.byref(self.model), ctypes.byref(self.rk),
                                                        ctypes.byref(h)))
        elif prop == "qj":
            _check(self.lib, self.lib.qjx_ctx_create(ctypes.byref(self.model), ctypes.byref(h)))
        else:
            raise ValueError("prop must be 'qj' or 'rk'")
        self.prop = prop
        self.h = h
        self.d = N + 1
        self.M = 0

    def close(self):
        if self.h:
            self.lib.qjx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def ensemble_init(self, num_trajectories, seed=0, mns=1000000, start_type=0, start_state=0,
                      jump=0, log_capacity=0):
        self.ens = EnsembleDesc(num_trajectories, seed, mns, start_type, start_state, jump,
                                log_capacity)
        _check(self.lib, self.lib.qjx_ensemble_init(self.h, ctypes.byref(self.ens)))
        self.M = num_trajectories

    def trans(self, num_tp_periods):
        _check(self.lib, self.lib.qjx_trans(self.h, num_tp_periods))

    def run_periods(self, n, is_obs=1):
        _check(self.lib, self.lib.qjx_run_periods(self.h, n, is_obs))

    def run_steps(self, n, is_obs=1):
        _check(self.lib, self.lib.qjx_run_steps(self.h, n, is_obs))

    def chars(self, is_start=0, adr_avg=False, adr=False):
        M, d = self.M, self.d
        out = {k: np.zeros(M) for k in ("norm", "mean", "dispersion", "m2", "energy")}
        aa = np.zeros(d) if adr_avg else None
        ad = np.zeros((M, d)) if adr else None
        _check(self.lib, self.lib.qjx_chars(self.h, is_start, _ptr(out["norm"]), _ptr(out["mean"]),
                                            _ptr(out["dispersion"]), _ptr(out["m2"]),
                                            _ptr(out["energy"]), _ptr(aa), _ptr(ad)))
        if adr_avg:
            out["adr_avg"] = aa
        if adr:
            out["adr"] = ad
        return out

    def density(self, rho=True):
        """Ensemble density matrix (d x d complex, or None) and purity Tr rho^2."""
        d = self.d
        r = np.zeros((d, d), dtype=np.complex128) if rho else None
        p = np.zeros(1)
        _check(self.lib, self.lib.qjx_density(self.h, None if r is None else _ptr(r.view(np.float64)), _ptr(p)))
        return r, float(p[0])

    def random_obs(self, num, seed=1, mns=1000000, typ=2):
        """Random observables of every trajectory's current state (M x num complex)."""
        out = np.zeros((self.M, num), dtype=np.complex128)
        _check(self.lib, self.lib.qjx_random_obs(self.h, num, seed, mns, typ, _ptr(out.view(np.float64))))
        return out

    def density_noalloc(self, rho=True):
        """density() through the allocation-free entry point (fixed device scratch)."""
        d = self.d
        r = np.zeros((d, d), dtype=np.complex128) if rho else None
        p = np.zeros(1)
        _check(self.lib, self.lib.qjx_density_noalloc(self.h, None if r is None else _ptr(r.view(np.float64)),
                                                      _ptr(p)))
        return r, float(p[0])

    def get_state(self, psi=True):
        M, d = self.M, self.d
        ps = np.zeros((M, d), dtype=np.complex128) if psi else None
        eta, t = np.zeros(M), np.zeros(M)
        nj = np.zeros(M, dtype=np.int32)
        rng = np.zeros(M, dtype=np.uint64)
        _check(self.lib, self.lib.qjx_get_state(self.h, None if ps is None else _ptr(ps.view(np.float64)),
                                                _ptr(eta), _ptr(t), _ptr(nj, ctypes.c_int),
                                                _ptr(rng, ctypes.c_uint64)))
        out = {"eta": eta, "t": t, "njumps": nj, "rng": rng}
        if psi:
            out["psi"] = ps
        return out

    def set_state(self, psi=None, eta=None, t=None, rng=None):
        psi = None if psi is None else np.ascontiguousarray(psi, dtype=np.complex128)
        _check(self.lib, self.lib.qjx_set_state(
            self.h, None if psi is None else _ptr(psi.view(np.float64)),
            _ptr(None if eta is None else np.ascontiguousarray(eta, dtype=np.float64)),
            _ptr(None if t is None else np.ascontiguousarray(t, dtype=np.float64)),
            _ptr(None if rng is None else np.ascontiguousarray(rng, dtype=np.uint64), ctypes.c_uint64)))

    def get_jumps(self, tr):
        n = _check(self.lib, self.lib.qjx_get_jumps(self.h, tr, None, None, None, None, 0))
        times, norms, etas = np.zeros(n), np.zeros(n), np.zeros(n)
        types = np.zeros(n, dtype=np.int32)
        _check(self.lib, self.lib.qjx_get_jumps(self.h, tr, _ptr(times), _ptr(types, ctypes.c_int),
                                                _ptr(norms), _ptr(etas), n))
        return {"times": times, "types": types, "norms": norms, "etas": etas}

    def get_propagator(self, b, lvl):
        U = np.zeros((self.d, self.d), dtype=np.complex128)
        _check(self.lib, self.lib.qjx_get_propagator(self.h, b, lvl, _ptr(U.view(np.float64))))
        return U

    def set_propagator(self, b, lvl, U):
        U = np.ascontiguousarray(U, dtype=np.complex128)
        _check(self.lib, self.lib.qjx_set_propagator(self.h, b, lvl, _ptr(U.view(np.float64))))

    def jump_interval_pdf(self, bin_begin=1e-10, num_decades=15, bins_per_decade=10):
        """Waiting-time pdf of the logged jumps (times_between_jumps.m binning):
        returns (bin borders, pdf, counts, outside)."""
        nb = num_decades * bins_per_decade
        pdf = np.zeros(nb)
        cnt = np.zeros(nb, dtype=np.int64)
        out = np.zeros(1, dtype=np.int64)
        _check(self.lib, self.lib.qjx_jump_interval_pdf(self.h, bin_begin, num_decades, bins_per_decade, _ptr(pdf),
                                                        _ptr(cnt, ctypes.c_int64), _ptr(out, ctypes.c_int64)))
        borders = bin_begin * 10.0 ** (np.arange(nb + 1) / bins_per_decade)
        return borders, pdf, cnt, int(out[0])

    def stats(self):
        s = Stats()
        _check(self.lib, self.lib.qjx_get_stats(self.h, ctypes.byref(s)))
        return _stats_dict(s)

    def set_profiling(self, on=True):
        _check(self.lib, self.lib.qjx_set_profiling(self.h, 1 if on else 0))


class Batch:
    """Several engines (one precision, any dimensions) advanced together: one
    grouped GEMM per iteration over all of their trajectories.  The engines stay
    usable on their own and must outlive the batch."""

    def __init__(self, engines):
        self.lib = load_library()
        self.engines = list(engines)
        arr = (ctypes.c_void_p * len(self.engines))(*[e.h.value for e in self.engines])
        h = ctypes.c_void_p()
        _check(self.lib, self.lib.qjx_batch_create(arr, len(self.engines), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.qjx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run_steps(self, n, is_obs=0):
        _check(self.lib, self.lib.qjx_batch_run_steps(self.h, n, is_obs))

    def run_periods(self, n, is_obs=0):
        _check(self.lib, self.lib.qjx_batch_run_periods(self.h, n, is_obs))

    def stats(self):
        s = Stats()
        _check(self.lib, self.lib.qjx_batch_get_stats(self.h, ctypes.byref(s)))
        return _stats_dict(s)

    def set_profiling(self, on=True):
        _check(self.lib, self.lib.qjx_batch_set_profiling(self.h, 1 if on else 0))


class Lindblad:
    """The dimer's Lindblad equation integrated in rho space on the GPU (FP64 RK4)."""

    def __init__(self, N, steps_per_period=2000, prm_E=1.0, prm_U=0.5, prm_J=1.0, drv_ampl=1.5, drv_freq=1.0,
                 diss_gamma=0.1, drv_type=0, drv_phase=0.0, per_stage=0):
        """per_stage=1: the drive at every RK4 stage time as CQdiss evaluates it
        (drv_type 1 sine, else the sign of the sine); per_stage=0: the square drive
        held constant over each step."""
        self.lib = load_library()
        m = ModelDesc(N, prm_E, prm_U, prm_J, drv_ampl, drv_freq, diss_gamma, 1, 4, 0)
        h = ctypes.c_void_p()
        dr = LindbladDrive(drv_type, drv_phase, per_stage)
        _check(self.lib, self.lib.qjx_lindblad_create_drive(ctypes.byref(m), steps_per_period, ctypes.byref(dr),
                                                            ctypes.byref(h)))
        self.h = h
        self.d = N + 1

    def close(self):
        if self.h:
            self.lib.qjx_lindblad_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_rho(self, rho):
        r = np.ascontiguousarray(rho, dtype=np.complex128)
        _check(self.lib, self.lib.qjx_lindblad_set_rho(self.h, _ptr(r.view(np.float64))))

    def get_rho(self):
        r = np.zeros((self.d, self.d), dtype=np.complex128)
        _check(self.lib, self.lib.qjx_lindblad_get_rho(self.h, _ptr(r.view(np.float64))))
        return r

    def run_periods(self, n):
        _check(self.lib, self.lib.qjx_lindblad_run_periods(self.h, n))

    def save_rho(self, path):
        _check(self.lib, self.lib.qjx_lindblad_save_rho(self.h, path.encode()))

    def observables(self):
        v = np.zeros(4)
        adr = np.zeros(self.d)
        _check(self.lib, self.lib.qjx_lindblad_observables(self.h, _ptr(v[0:]), _ptr(v[1:]), _ptr(v[2:]),
                                                           _ptr(v[3:]), _ptr(adr)))
        return {"mean": v[0], "m2": v[1], "energy": v[2], "purity": v[3], "adr": adr}


class Pool:
    """Continuous batching of ensemble jobs on the current HIP device."""

    def __init__(self, precision=FP32, hbm_budget_gib=0.0, max_trajectories=1 << 20):
        self.lib = load_library()
        h = ctypes.c_void_p()
        _check(self.lib, self.lib.qjx_pool_create(precision, hbm_budget_gib, max_trajectories, ctypes.byref(h)))
        self.h = h
        self.sizes = {}

    def close(self):
        if self.h:
            self.lib.qjx_pool_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, N, num_trajectories, n_periods, seed=0, mns=1000000, start_type=0, start_state=0,
               prm_E=1.0, prm_U=0.5, prm_J=1.0, drv_ampl=1.5, drv_freq=1.0, diss_gamma=0.1, ex_deep=16,
               steps_per_period=4):
        m = ModelDesc(N, prm_E, prm_U, prm_J, drv_ampl, drv_freq, diss_gamma, ex_deep, steps_per_period, 0)
        e = EnsembleDesc(num_trajectories, seed, mns, start_type, start_state, 0, 0)
        jid = ctypes.c_int()
        _check(self.lib, self.lib.qjx_pool_submit(self.h, ctypes.byref(m), ctypes.byref(e), n_periods,
                                                  ctypes.byref(jid)))
        self.sizes[jid.value] = num_trajectories
        return jid.value

    def step(self):
        r, q = ctypes.c_int(), ctypes.c_int()
        _check(self.lib, self.lib.qjx_pool_step(self.h, ctypes.byref(r), ctypes.byref(q)))
        return r.value, q.value

    def clock(self):
        return self.lib.qjx_pool_clock(self.h)

    def result(self, jid, arrays=False):
        res = JobResult()
        M = self.sizes[jid]
        mean = np.zeros(M) if arrays else None
        energy = np.zeros(M) if arrays else None
        _check(self.lib, self.lib.qjx_pool_result(self.h, jid, ctypes.byref(res), _ptr(mean), _ptr(energy)))
        out = {k: getattr(res, k) for k, _ in JobResult._fields_}
        if arrays:
            out["mean"] = mean
            out["energy"] = energy
        return out

    def usage(self):
        u = PoolUsage()
        _check(self.lib, self.lib.qjx_pool_get_usage(self.h, ctypes.byref(u)))
        return {k: getattr(u, k) for k, _ in PoolUsage._fields_ if k != "pad"}

    def drain(self, max_cycles=1000000):
        for _ in range(max_cycles):
            r, q = self.step()
            if r == 0 and q == 0:
                return
        raise QJXError(-1, "pool did not drain")


def mfma_probe(A32, B32, A64, B64, Abf, Bbf):
    """Run the matrix-core layout probe; returns (D32, D64, Dbf)."""
    lib = load_library()
    A32, B32, Abf, Bbf = (np.ascontiguousarray(x, dtype=np.float32) for x in (A32, B32, Abf, Bbf))
    A64, B64 = (np.ascontiguousarray(x, dtype=np.float64) for x in (A64, B64))
    D32 = np.zeros((32, 32), dtype=np.float32)
    D64 = np.zeros((16, 16), dtype=np.float64)
    Dbf = np.zeros((32, 32), dtype=np.float32)
    _check(lib, lib.qjx_mfma_probe(_ptr(A32, ctypes.c_float), _ptr(B32, ctypes.c_float),
                                   _ptr(D32, ctypes.c_float), _ptr(A64), _ptr(B64), _ptr(D64),
                                   _ptr(Abf, ctypes.c_float), _ptr(Bbf, ctypes.c_float),
                                   _ptr(Dbf, ctypes.c_float)))
    return D32, D64, Dbf


def mfma_probe_f32x16(A, B):
    """The 16x16x4 FP32 matrix-core product of one wavefront: D = A (16x4) B (4x16)."""
    lib = load_library()
    A, B = (np.ascontiguousarray(x, dtype=np.float32) for x in (A, B))
    D = np.zeros((16, 16), dtype=np.float32)
    _check(lib, lib.qjx_mfma_probe_f32x16(_ptr(A, ctypes.c_float), _ptr(B, ctypes.c_float), _ptr(D, ctypes.c_float)))
    return D


def run_config(config_path, params_path, precision=FP64, steps_per_period=0):
    """Drop-in runner: the reference's config.txt / params.txt in, its output files out."""
    lib = load_library()
    return _check(lib, lib.qjx_run_config(config_path.encode(), params_path.encode(), precision,
                                          steps_per_period))


LPN_DEFAULTS = dict(lpn_type=-1, lpn_eps_deep=100, lpn_eps_error=1e-10, lpn_eps_high=100.0, lpn_eps_low=-100.0,
                    lpn_delta_s=1e-2, lpn_delta_f_high=1e-2, lpn_delta_f_low=1e-2, lambda_per_periods=1,
                    save_lambdas=0, num_random_obs=0, random_obs_seed=1, random_obs_mns=1000000,
                    random_obs_type=2)


class Lpn:
    """The LPN task (task_id 7) over an Engine whose ensemble is initialised with an
    even number of trajectories; keyword arguments are the params.txt keys."""

    def __init__(self, engine, **kw):
        self.lib = engine.lib
        self.engine = engine
        p = dict(LPN_DEFAULTS, **kw)
        self.desc = LpnDesc(**p)
        self.nobs = p["num_random_obs"]
        h = ctypes.c_void_p()
        _check(self.lib, self.lib.qjx_lpn_create(engine.h, ctypes.byref(self.desc), ctypes.byref(h)))
        self.h = h
        self.M = engine.M

    def close(self):
        if self.h:
            self.lib.qjx_lpn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def trans(self, n):
        _check(self.lib, self.lib.qjx_lpn_trans(self.h, n))

    def run_periods(self, n):
        _check(self.lib, self.lib.qjx_lpn_run_periods(self.h, n))

    def get(self):
        M = self.M
        out = {k: np.zeros(M) for k in ("lambda_now", "lambda", "mean_lpn", "energy_lpn", "delta_s")}
        nr = np.zeros(M, dtype=np.int32)
        _check(self.lib, self.lib.qjx_lpn_get(self.h, _ptr(out["lambda_now"]), _ptr(out["lambda"]),
                                              _ptr(nr, ctypes.c_int), _ptr(out["mean_lpn"]),
                                              _ptr(out["energy_lpn"]), _ptr(out["delta_s"])))
        out["num_renorms"] = nr
        return out

    def random_obs(self):
        o = np.zeros((self.M, max(self.nobs, 1)), dtype=np.complex128)
        ol = np.zeros_like(o)
        _check(self.lib, self.lib.qjx_lpn_random_obs(self.h, _ptr(o.view(np.float64)), _ptr(ol.view(np.float64))))
        return o[:, :self.nobs], ol[:, :self.nobs]

    def lambdas(self, tr):
        n = _check(self.lib, self.lib.qjx_lpn_lambdas(self.h, tr, None, None, None, 0))
        r, ds, df = np.zeros(n), np.zeros(n), np.zeros(n)
        self.lib.qjx_lpn_lambdas(self.h, tr, _ptr(r), _ptr(ds), _ptr(df), n)
        return r, ds, df

    def chars(self, begin=0, end=None):
        """calc_chars_std of the last period (before its renormalisations), abs_diag_rho
        and its average over trajectories [begin, end)."""
        M, d = self.M, self.engine.d
        end = M // 2 if end is None else end
        out = {k: np.zeros(M) for k in ("norm", "mean", "dispersion", "m2", "energy")}
        adr = np.zeros((M, d))
        avg = np.zeros(d)
        _check(self.lib, self.lib.qjx_lpn_chars(self.h, _ptr(out["norm"]), _ptr(out["mean"]), _ptr(out["dispersion"]),
                                                _ptr(out["m2"]), _ptr(out["energy"]), _ptr(adr), begin, end,
                                                _ptr(avg)))
        out["adr"] = adr
        out["adr_avg"] = avg
        return out

    def random_obs_matrix(self, obs):
        d = self.engine.d
        R = np.zeros((d, d), dtype=np.complex128)
        _check(self.lib, self.lib.qjx_lpn_random_obs_matrix(self.h, obs, _ptr(R.view(np.float64))))
        return R


def peak_probe():
    """Measured device ceilings: FP32 / FP64 / BF16 MFMA and FP64 vector TFLOP/s,
    HBM copy GB/s."""
    lib = load_library()
    v = [ctypes.c_double() for _ in range(5)]
    _check(lib, lib.qjx_peak_probe(*[ctypes.byref(x) for x in v]))
    return dict(zip(("f32_mfma_tflops", "f64_mfma_tflops", "bf16_mfma_tflops", "f64_valu_tflops", "hbm_gbps"),
                    (x.value for x in v)))
