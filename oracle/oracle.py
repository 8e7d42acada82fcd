"""ctypes loader for the FP64 CPU oracle (oracle/qjx_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the os_d_amd package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "libqjx_oracle.so")
SOURCES = ["qjx_oracle.cpp", "lindblad_oracle.cpp"]


def build():
    """Compile the oracle with g++ (OpenMP)."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = [os.path.join(HERE, s) for s in SOURCES if os.path.exists(os.path.join(HERE, s))]
    tmp = LIB + ".tmp"
    # fma() is the hardware instruction (-mfma); nothing else is contracted.
    cmd = ["g++", "-O2", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-mfma", "-ffp-contract=off"] + \
        srcs + ["-o", tmp]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    os.replace(tmp, LIB)
    return LIB


class LpnDesc(ctypes.Structure):
    _fields_ = [("lpn_type", ctypes.c_int), ("lpn_eps_deep", ctypes.c_int), ("lpn_eps_error", ctypes.c_double),
                ("lpn_eps_high", ctypes.c_double), ("lpn_eps_low", ctypes.c_double),
                ("lpn_delta_s", ctypes.c_double), ("lpn_delta_f_high", ctypes.c_double),
                ("lpn_delta_f_low", ctypes.c_double), ("lambda_per_periods", ctypes.c_int),
                ("save_lambdas", ctypes.c_int), ("num_random_obs", ctypes.c_int),
                ("random_obs_seed", ctypes.c_int), ("random_obs_mns", ctypes.c_int),
                ("random_obs_type", ctypes.c_int)]


LPN_DEFAULTS = dict(lpn_type=-1, lpn_eps_deep=100, lpn_eps_error=1e-10, lpn_eps_high=100.0, lpn_eps_low=-100.0,
                    lpn_delta_s=1e-2, lpn_delta_f_high=1e-2, lpn_delta_f_low=1e-2, lambda_per_periods=1,
                    save_lambdas=0, num_random_obs=0, random_obs_seed=1, random_obs_mns=1000000,
                    random_obs_type=2)


class Model(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("prm_E", ctypes.c_double), ("prm_U", ctypes.c_double),
                ("prm_J", ctypes.c_double), ("drv_ampl", ctypes.c_double),
                ("drv_freq", ctypes.c_double), ("gamma", ctypes.c_double), ("S", ctypes.c_int),
                ("D", ctypes.c_int)]


_lib = None
_P = ctypes.POINTER
_D = _P(ctypes.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.qjo_model_matrices.argtypes = [_P(Model), _D, _D, _D, _D, _D]
        L.qjo_expm.argtypes = [ctypes.c_int, _D, _D]
        L.qjo_mcg59_uniform.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _D]
        L.qjo_create.restype = ctypes.c_void_p
        L.qjo_create.argtypes = [_P(Model)] + [ctypes.c_int] * 8
        L.qjo_destroy.argtypes = [ctypes.c_void_p]
        L.qjo_error.restype = ctypes.c_char_p
        L.qjo_error.argtypes = [ctypes.c_void_p]
        L.qjo_set_propagator.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _D]
        L.qjo_get_propagator.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _D]
        L.qjo_trans.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.qjo_periods.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.qjo_steps.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.qjo_periods_one.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.qjo_chars.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, _D, _D, _D, _D, _D]
        L.qjo_get_state.argtypes = [ctypes.c_void_p, _D, _D, _D, _P(ctypes.c_int), _P(ctypes.c_uint64)]
        L.qjo_set_state.argtypes = [ctypes.c_void_p, _D, _D, _D, _P(ctypes.c_uint64)]
        L.qjo_get_jumps.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, _P(ctypes.c_int), _D, _D,
                                    ctypes.c_int]
        L.qjo_num_matvecs.restype = ctypes.c_int64
        L.qjo_num_matvecs.argtypes = [ctypes.c_void_p]
        L.qjo_num_jumps.restype = ctypes.c_int64
        L.qjo_num_jumps.argtypes = [ctypes.c_void_p]
        L.qjo_rk_setup.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.qjo_rk_sin.restype = ctypes.c_double
        L.qjo_rk_sin.argtypes = [ctypes.c_double]
        L.qjo_rk_drive.restype = ctypes.c_double
        L.qjo_rk_drive.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.qjo_rk_rhs.argtypes = [ctypes.c_void_p, ctypes.c_double, _D, _D]
        L.qjo_rk_norm.restype = ctypes.c_double
        L.qjo_rk_norm.argtypes = [ctypes.c_int, _D]
        L.qjo_rk_int.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        L.qjo_mcg31_uniform.argtypes = [ctypes.c_uint, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, _D]
        L.qjo_mcg31_gaussian.argtypes = [ctypes.c_uint, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D]
        L.qjo_random_obs_matrix.argtypes = [ctypes.c_int] * 5 + [_D]
        L.qjo_lpn_setup.argtypes = [ctypes.c_void_p, _P(LpnDesc)]
        L.qjo_lpn_trans.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.qjo_lpn_period.argtypes = [ctypes.c_void_p]
        L.qjo_lpn_get.argtypes = [ctypes.c_void_p, _D, _D, _P(ctypes.c_int), _D, _D, _D, _D, _D]
        L.qjo_lpn_random_obs.argtypes = [ctypes.c_void_p, _D, _D]
        L.qjo_lpn_lambdas.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, _D, _D, ctypes.c_int]
        L.qjo_random_obs.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 4 + [_D]
        L.qjo_lpn_exp10.restype = ctypes.c_double
        L.qjo_lpn_exp10.argtypes = [ctypes.c_double]
        L.qjo_lpn_log.restype = ctypes.c_double
        L.qjo_lpn_log.argtypes = [ctypes.c_double]
        L.qjl_evolve.argtypes = [_P(Model), ctypes.c_int, ctypes.c_int, _D]
        L.qjl_evolve_drive.argtypes = [_P(Model), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_int, _D]
        L.qjl_rhs.argtypes = [_P(Model), ctypes.c_double, _D, _D]
        _lib = L
    return _lib


def _p(a, t=ctypes.c_double):
    return None if a is None else a.ctypes.data_as(_P(t))


def default_threads():
    """OpenMP threads of the oracle: OMP_NUM_THREADS when set (the host share a
    job may use), else every CPU."""
    n = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    return min(n, os.cpu_count() or n)


def model(N, prm_E=1.0, prm_U=0.5, prm_J=1.0, drv_ampl=1.5, drv_freq=1.0, gamma=0.1, S=4, D=16):
    return Model(N, prm_E, prm_U, prm_J, drv_ampl, drv_freq, gamma, S, D)


def model_matrices(m):
    d = m.N + 1
    out = [np.zeros((d, d), dtype=np.complex128) for _ in range(5)]
    lib().qjo_model_matrices(ctypes.byref(m), *[_p(x.view(np.float64)) for x in out])
    return dict(zip(("H", "Hdrv", "L", "Hqj0", "Hqj1"), out))


def expm(A):
    A = np.ascontiguousarray(A, dtype=np.complex128)
    E = np.zeros_like(A)
    rc = lib().qjo_expm(A.shape[0], _p(A.view(np.float64)), _p(E.view(np.float64)))
    assert rc == 0
    return E


def lindblad_evolve(m, rho, n_periods, steps_per_period=2000):
    """rho-space RK4 of the dimer Lindblad equation over n_periods periods."""
    r = np.ascontiguousarray(rho, dtype=np.complex128).copy()
    rc = lib().qjl_evolve(ctypes.byref(m), steps_per_period, n_periods, _p(r.view(np.float64)))
    assert rc == 0
    return r


def lindblad_evolve_drive(m, rho, n_periods, steps_per_period=2000, drv_type=0, drv_phase=0.0, per_stage=1,
                          period0=0):
    """rho-space RK4 with CQdiss's per-stage drive (per_stage=1) or the per-step
    square drive (per_stage=0), from period period0 on."""
    r = np.ascontiguousarray(rho, dtype=np.complex128).copy()
    rc = lib().qjl_evolve_drive(ctypes.byref(m), steps_per_period, period0, n_periods, drv_type, drv_phase,
                                per_stage, _p(r.view(np.float64)))
    assert rc == 0
    return r


def lindblad_rhs(m, E, rho):
    r = np.ascontiguousarray(rho, dtype=np.complex128)
    out = np.zeros_like(r)
    lib().qjl_rhs(ctypes.byref(m), E, _p(r.view(np.float64)), _p(out.view(np.float64)))
    return out


def lindblad_observables(m, rho):
    """<k>, <(k - <k>)^2>, energy <H + E H_drv>, purity, diag(rho) of a density matrix."""
    d = m.N + 1
    k = np.arange(d)
    p = np.real(np.diag(rho))
    mean = float(p @ k)
    mm = model_matrices(m)
    Hs = mm["H"] + m.prm_E * mm["Hdrv"]
    return {"mean": mean, "m2": float(p @ (k - mean) ** 2), "energy": float(np.real(np.trace(rho @ Hs))),
            "purity": float(np.real(np.trace(rho @ rho))), "adr": p}


def mcg31_uniform(seed, k, n, count, lo=0.0, hi=1.0):
    out = np.zeros(count)
    lib().qjo_mcg31_uniform(seed, k, n, count, lo, hi, _p(out))
    return out


def mcg31_gaussian(seed, k, n, count):
    out = np.zeros(count)
    lib().qjo_mcg31_gaussian(seed, k, n, count, _p(out))
    return out


def random_obs_matrix(d, seed, mns, typ, obs):
    R = np.zeros((d, d), dtype=np.complex128)
    lib().qjo_random_obs_matrix(d, seed, mns, typ, obs, _p(R.view(np.float64)))
    return R


def rk_sin(x):
    return lib().qjo_rk_sin(float(x))


def rk_norm(x):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    return lib().qjo_rk_norm(x.shape[0], _p(x.view(np.float64)))


def mcg59_uniform(k, mns, n):
    out = np.zeros(n)
    lib().qjo_mcg59_uniform(k, mns, n, _p(out))
    return out


class Ensemble:
    """The oracle ensemble: one trajectory per OpenMP iteration, as the reference."""

    def __init__(self, m, num_traj, seed=0, mns=1000000, start_type=0, start_state=0, jump=1,
                 threads=0, build_propagators=True):
        self.L = lib()
        self.m = m
        self.d = m.N + 1
        self.M = num_traj
        threads = threads or default_threads()
        self.h = self.L.qjo_create(ctypes.byref(m), num_traj, seed, mns, start_type, start_state,
                                   jump, threads, 1 if build_propagators else 0)
        if not self.h:
            raise ValueError("qjo_create rejected the arguments")

    def close(self):
        if self.h:
            self.L.qjo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ok(self, rc):
        if rc != 0:
            raise RuntimeError("oracle: " + self.L.qjo_error(self.h).decode())

    def set_propagator(self, b, lvl, U):
        U = np.ascontiguousarray(U, dtype=np.complex128)
        self._ok(self.L.qjo_set_propagator(self.h, b, lvl, _p(U.view(np.float64))))

    def get_propagator(self, b, lvl):
        U = np.zeros((self.d, self.d), dtype=np.complex128)
        self._ok(self.L.qjo_get_propagator(self.h, b, lvl, _p(U.view(np.float64))))
        return U

    def trans(self, n):
        self._ok(self.L.qjo_trans(self.h, n))

    def periods(self, n, is_obs=1):
        self._ok(self.L.qjo_periods(self.h, n, is_obs))

    def steps(self, n, is_obs=1):
        self._ok(self.L.qjo_steps(self.h, n, is_obs))

    def periods_one(self, tr, n, is_obs=1):
        self._ok(self.L.qjo_periods_one(self.h, tr, n, is_obs))

    def chars(self, is_start=0, adr=False):
        M, d = self.M, self.d
        out = {k: np.zeros(M) for k in ("norm", "mean", "dispersion", "m2", "energy")}
        a = np.zeros((M, d)) if adr else None
        self._ok(self.L.qjo_chars(self.h, is_start, _p(out["norm"]), _p(out["mean"]),
                                  _p(out["dispersion"]), _p(out["m2"]), _p(out["energy"]), _p(a)))
        if adr:
            out["adr"] = a
        return out

    def get_state(self):
        M, d = self.M, self.d
        psi = np.zeros((M, d), dtype=np.complex128)
        eta, t = np.zeros(M), np.zeros(M)
        nj = np.zeros(M, dtype=np.int32)
        rng = np.zeros(M, dtype=np.uint64)
        self._ok(self.L.qjo_get_state(self.h, _p(psi.view(np.float64)), _p(eta), _p(t),
                                      _p(nj, ctypes.c_int), _p(rng, ctypes.c_uint64)))
        return {"psi": psi, "eta": eta, "t": t, "njumps": nj, "rng": rng}

    def set_state(self, psi=None, eta=None, t=None, rng=None):
        psi = None if psi is None else np.ascontiguousarray(psi, dtype=np.complex128)
        self._ok(self.L.qjo_set_state(
            self.h, None if psi is None else _p(psi.view(np.float64)),
            _p(None if eta is None else np.ascontiguousarray(eta, dtype=np.float64)),
            _p(None if t is None else np.ascontiguousarray(t, dtype=np.float64)),
            _p(None if rng is None else np.ascontiguousarray(rng, dtype=np.uint64), ctypes.c_uint64)))

    def rk_setup(self, rk_ns, drv_type=0, drv_phase=0.0):
        """Switch to Runge-Kutta propagation (prop_id 1)."""
        self._ok(self.L.qjo_rk_setup(self.h, rk_ns, drv_type, drv_phase))

    def rk_drive(self, t):
        return self.L.qjo_rk_drive(self.h, t)

    def rk_rhs(self, t, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        k = np.zeros_like(x)
        self._ok(self.L.qjo_rk_rhs(self.h, t, _p(x.view(np.float64)), _p(k.view(np.float64))))
        return k

    def rk_int(self, tr, step):
        self._ok(self.L.qjo_rk_int(self.h, tr, step))

    def random_obs(self, num, seed=1, mns=1000000, typ=2):
        """calc_random_obs of the STD path for every trajectory (M x num complex)."""
        out = np.zeros((self.M, num), dtype=np.complex128)
        self._ok(self.L.qjo_random_obs(self.h, num, seed, mns, typ, _p(out.view(np.float64))))
        return out

    def lpn_setup(self, **kw):
        """LPN bookkeeping (task_id 7); keywords are the params.txt keys."""
        p = dict(LPN_DEFAULTS, **kw)
        self.lpn_nobs = p["num_random_obs"]
        self._ok(self.L.qjo_lpn_setup(self.h, ctypes.byref(LpnDesc(**p))))

    def lpn_trans(self, n):
        self._ok(self.L.qjo_lpn_trans(self.h, n))

    def lpn_period(self):
        self._ok(self.L.qjo_lpn_period(self.h))

    def lpn_get(self):
        M = self.M
        out = {k: np.zeros(M) for k in ("lambda_now", "lambda", "mean_lpn", "energy_lpn", "delta_s", "mean",
                                        "energy")}
        nr = np.zeros(M, dtype=np.int32)
        self._ok(self.L.qjo_lpn_get(self.h, _p(out["lambda_now"]), _p(out["lambda"]), _p(nr, ctypes.c_int),
                                    _p(out["mean_lpn"]), _p(out["energy_lpn"]), _p(out["delta_s"]),
                                    _p(out["mean"]), _p(out["energy"])))
        out["num_renorms"] = nr
        return out

    def lpn_random_obs(self):
        o = np.zeros((self.M, max(self.lpn_nobs, 1)), dtype=np.complex128)
        ol = np.zeros_like(o)
        self._ok(self.L.qjo_lpn_random_obs(self.h, _p(o.view(np.float64)), _p(ol.view(np.float64))))
        return o[:, :self.lpn_nobs], ol[:, :self.lpn_nobs]

    def lpn_lambdas(self, tr):
        n = self.L.qjo_lpn_lambdas(self.h, tr, None, None, None, 0)
        r, ds, df = np.zeros(n), np.zeros(n), np.zeros(n)
        self.L.qjo_lpn_lambdas(self.h, tr, _p(r), _p(ds), _p(df), n)
        return r, ds, df

    def get_jumps(self, tr):
        n = self.L.qjo_get_jumps(self.h, tr, None, None, None, None, 0)
        times, norms, etas = np.zeros(n), np.zeros(n), np.zeros(n)
        types = np.zeros(n, dtype=np.int32)
        self.L.qjo_get_jumps(self.h, tr, _p(times), _p(types, ctypes.c_int), _p(norms), _p(etas), n)
        return {"times": times, "types": types, "norms": norms, "etas": etas}

    def num_matvecs(self):
        return self.L.qjo_num_matvecs(self.h)

    def num_jumps(self):
        return self.L.qjo_num_jumps(self.h)
