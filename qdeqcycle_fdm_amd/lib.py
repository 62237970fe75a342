"""ctypes bindings of libqdeq.so (include/qdeq.h) and a small host-side
mirror of the reference's call surface.

The reference is one C++ executable whose hot path is
``RK43 -> setBoundary -> ode_solve -> eval_stress_on_fault -> NRsearch ->
cal_dpsi`` (odeSolver.cpp:1277-1674).  :class:`Engine` exposes the same
functions (names and argument meaning kept) over the HIP library, plus the
device-resident RK43 driver.  There is no CPU fallback: if libqdeq.so is
missing or no GPU is visible, every call raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

from . import config as cfgmod

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqdeq.so")

QDEQ_OK = 0
QDEQ_ERR_ARG = 1
QDEQ_ERR_HIP = 2
QDEQ_ERR_NAN_FRICTION = 3
QDEQ_ERR_MG_NOT_CONVERGED = 4
QDEQ_ERR_STATE = 5
QDEQ_ERR_NOMEM = 6
QDEQ_HOST, QDEQ_DEVICE = 0, 1

# Every entry point include/qdeq.h declares.
EXPORTS = (
    "qdeq_default_config", "qdeq_create", "qdeq_destroy", "qdeq_last_error",
    "qdeq_global_error", "qdeq_hip_runtime_version", "qdeq_get_mu_equi",
    "qdeq_build_mu_equi", "qdeq_build_mu_equi_noalloc",
    "qdeq_ode_solve", "qdeq_set_boundary", "qdeq_eval_stress",
    "qdeq_nr_search", "qdeq_cal_dpsi", "qdeq_compute_error",
    "qdeq_compute_step_size", "qdeq_rk43_setup", "qdeq_rk43_step",
    "qdeq_rk43_get_state", "qdeq_set_tiles", "qdeq_uses_tiled_cycle",
    "qdeq_set_kernel_timing", "qdeq_get_kernel_timing", "qdeq_copy_probe",
    "qdeq_device_synchronize", "qdeq_autotune_tiles", "qdeq_get_tile_shapes",
    "qdeq_kernel_breakdown", "qdeq_set_tile_shapes", "qdeq_ensemble_step",
    "qdeq_output_open", "qdeq_output_close", "qdeq_events_enable",
    "qdeq_events_get", "qdeq_projection_enable", "qdeq_projection_stats",
    "qdeq_projection_fit", "qdeq_projection_inject", "qdeq_launch_ahead_stats",
    "qdeq_folded_stop_tests",
    # include/qdeq_guess.h
    "qdeq_guess_default_config", "qdeq_train_default_config",
    "qdeq_guess_create", "qdeq_guess_destroy", "qdeq_guess_last_error",
    "qdeq_guess_num_params", "qdeq_guess_dims", "qdeq_guess_get_params",
    "qdeq_guess_set_params", "qdeq_guess_pairs", "qdeq_guess_clear_pairs",
    "qdeq_guess_read_pairs", "qdeq_guess_train", "qdeq_guess_forward_backward",
    "qdeq_guess_forward", "qdeq_attach_guess", "qdeq_guess_stats",
    "qdeq_guess_inject", "qdeq_guess_infer",
)
TIME_OFF, TIME_UP0, TIME_TAIL, TIME_DOWN1, TIME_FAULT = 0, 1, 2, 3, 4
GUESS_OFF, GUESS_HARVEST, GUESS_PREDICT = 0, 1, 2


class QdeqError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"qdeq status {status}: {msg}")
        self.status = status


class QdeqConfig(ctypes.Structure):
    _fields_ = [
        ("ni", ctypes.c_int), ("nk", ctypes.c_int), ("dh", ctypes.c_double),
        ("n_levels", ctypes.c_int), ("n_pre", ctypes.c_int),
        ("n_post", ctypes.c_int), ("n_coarse", ctypes.c_int),
        ("relax", ctypes.c_double), ("cycle", ctypes.c_int),
        ("precision", ctypes.c_int), ("tol_mode", ctypes.c_int),
        ("tol", ctypes.c_double), ("max_vcycles", ctypes.c_int),
        ("ini_dt", ctypes.c_double), ("mindt", ctypes.c_double),
        ("maxdt", ctypes.c_double), ("totTol", ctypes.c_double),
        ("tmax", ctypes.c_double), ("f0", ctypes.c_double),
        ("V0", ctypes.c_double), ("L", ctypes.c_double),
        ("Vini", ctypes.c_double), ("device", ctypes.c_int),
    ]


class StepInfo(ctypes.Structure):
    _fields_ = [
        ("t", ctypes.c_double), ("dt", ctypes.c_double),
        ("totErr", ctypes.c_double), ("errA0", ctypes.c_double),
        ("errA1", ctypes.c_double), ("V_max", ctypes.c_double),
        ("Ts_max", ctypes.c_double), ("it", ctypes.c_int),
        ("attempts", ctypes.c_int), ("vcycles", ctypes.c_longlong),
        ("solves", ctypes.c_longlong), ("ir_sweeps", ctypes.c_longlong),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Event(ctypes.Structure):
    _fields_ = [
        ("t_onset", ctypes.c_double), ("t_end", ctypes.c_double),
        ("t_peak", ctypes.c_double), ("v_peak", ctypes.c_double),
        ("slip_max", ctypes.c_double), ("it_onset", ctypes.c_int),
        ("it_end", ctypes.c_int), ("node_onset", ctypes.c_int),
        ("node_peak", ctypes.c_int),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class GuessConfig(ctypes.Structure):
    _fields_ = [
        ("nk", ctypes.c_int), ("factor", ctypes.c_int), ("channels", ctypes.c_int),
        ("width", ctypes.c_int), ("hidden", ctypes.c_int), ("capacity", ctypes.c_int),
        ("stride", ctypes.c_int), ("seed", ctypes.c_ulonglong), ("device", ctypes.c_int),
    ]


class TrainConfig(ctypes.Structure):
    _fields_ = [
        ("epochs", ctypes.c_int), ("batch", ctypes.c_int), ("lr", ctypes.c_double),
        ("lr_final", ctypes.c_double), ("beta1", ctypes.c_double),
        ("beta2", ctypes.c_double), ("eps", ctypes.c_double),
        ("seed", ctypes.c_ulonglong),
    ]


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_lib = None


def check_runtime(built: int, running: int) -> None:
    """Refuses a HIP runtime other than the one the library was built
    against (major.minor, HIP_VERSION encoding).  A library loaded earlier
    can bring its own libamdhip64 under the same soname (torch's wheels
    bundle one), and libqdeq then binds to that runtime."""
    def mm(v):
        return v // 10000000, (v // 100000) % 100
    if mm(built) != mm(running):
        raise RuntimeError(
            "libqdeq.so was built against HIP %d.%d but this process runs HIP "
            "%d.%d, loaded by another library under the same soname (for "
            "example torch's bundled runtime): load qdeqcycle_fdm_amd.lib before "
            "that library, or run without it" % (*mm(built), *mm(running)))


def load(path: str = LIB_PATH):
    """Loads libqdeq.so (raises if it has not been built, or if the process
    runs another HIP runtime than the one it was built against)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with qdeqcycle_fdm_amd.build.build()")
    L = ctypes.CDLL(path)
    L.qdeq_hip_runtime_version.argtypes = [_ip, _ip]
    built, running = ctypes.c_int(0), ctypes.c_int(0)
    if L.qdeq_hip_runtime_version(ctypes.byref(built), ctypes.byref(running)) == QDEQ_OK:
        check_runtime(built.value, running.value)
    L.qdeq_default_config.argtypes = [ctypes.POINTER(QdeqConfig)]
    L.qdeq_default_config.restype = None
    L.qdeq_create.argtypes = [ctypes.POINTER(QdeqConfig), _dp, _dp,
                              ctypes.POINTER(_vp)]
    L.qdeq_destroy.argtypes = [_vp]
    L.qdeq_last_error.argtypes = [_vp]
    L.qdeq_last_error.restype = ctypes.c_char_p
    L.qdeq_global_error.argtypes = []
    L.qdeq_global_error.restype = ctypes.c_char_p
    L.qdeq_get_mu_equi.argtypes = [_vp, _vp, ctypes.c_int]
    L.qdeq_build_mu_equi.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp,
                                     ctypes.c_int]
    L.qdeq_build_mu_equi_noalloc.argtypes = [ctypes.c_int, ctypes.c_int, _vp, _vp,
                                             ctypes.c_int]
    L.qdeq_ode_solve.argtypes = [_vp, _vp, ctypes.c_int, _ip, _dp, _dp]
    L.qdeq_set_boundary.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_int]
    L.qdeq_eval_stress.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int]
    L.qdeq_nr_search.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int]
    L.qdeq_cal_dpsi.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.c_int]
    L.qdeq_compute_error.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
    L.qdeq_compute_error.restype = ctypes.c_double
    L.qdeq_compute_step_size.argtypes = [ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, _dp]
    L.qdeq_compute_step_size.restype = ctypes.c_double
    L.qdeq_set_tiles.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.qdeq_uses_tiled_cycle.argtypes = [_vp]
    L.qdeq_autotune_tiles.argtypes = [_vp, ctypes.c_int]
    L.qdeq_get_tile_shapes.argtypes = [_vp, _ip, _ip]
    L.qdeq_kernel_breakdown.argtypes = [_vp, ctypes.c_int, _ip, _dp]
    L.qdeq_set_tile_shapes.argtypes = [_vp, ctypes.c_int, _ip]
    L.qdeq_ensemble_step.argtypes = [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, _ip]
    L.qdeq_output_open.argtypes = [_vp, ctypes.c_char_p, ctypes.c_int]
    L.qdeq_output_close.argtypes = [_vp]
    L.qdeq_events_enable.argtypes = [_vp, ctypes.c_double, ctypes.c_int]
    L.qdeq_events_get.argtypes = [_vp, ctypes.POINTER(Event), ctypes.c_int, _ip]
    L.qdeq_set_kernel_timing.argtypes = [_vp, ctypes.c_int]
    L.qdeq_get_kernel_timing.argtypes = [_vp, _dp, ctypes.POINTER(ctypes.c_longlong)]
    L.qdeq_copy_probe.argtypes = [_vp, ctypes.c_longlong, ctypes.c_int, _dp]
    L.qdeq_device_synchronize.argtypes = [ctypes.c_int]
    L.qdeq_rk43_setup.argtypes = [_vp, _dp, _dp, _dp, _dp, _dp]
    L.qdeq_rk43_step.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(StepInfo)]
    L.qdeq_rk43_get_state.argtypes = [_vp, ctypes.POINTER(StepInfo), _dp, _dp,
                                      _dp, _dp, _dp, _dp]
    _fp = ctypes.POINTER(ctypes.c_float)
    _llp = ctypes.POINTER(ctypes.c_longlong)
    L.qdeq_guess_default_config.argtypes = [ctypes.POINTER(GuessConfig), ctypes.c_int]
    L.qdeq_guess_default_config.restype = None
    L.qdeq_train_default_config.argtypes = [ctypes.POINTER(TrainConfig)]
    L.qdeq_train_default_config.restype = None
    L.qdeq_guess_create.argtypes = [ctypes.POINTER(GuessConfig), ctypes.POINTER(_vp)]
    L.qdeq_guess_destroy.argtypes = [_vp]
    L.qdeq_guess_last_error.argtypes = [_vp]
    L.qdeq_guess_last_error.restype = ctypes.c_char_p
    L.qdeq_guess_num_params.argtypes = [_vp]
    L.qdeq_guess_num_params.restype = ctypes.c_longlong
    L.qdeq_guess_dims.argtypes = [_vp, _ip, _ip]
    L.qdeq_guess_get_params.argtypes = [_vp, _fp]
    L.qdeq_guess_set_params.argtypes = [_vp, _fp]
    L.qdeq_guess_pairs.argtypes = [_vp, _ip]
    L.qdeq_guess_clear_pairs.argtypes = [_vp]
    L.qdeq_guess_read_pairs.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _fp, _fp, _dp]
    L.qdeq_guess_train.argtypes = [_vp, ctypes.POINTER(TrainConfig), _dp]
    L.qdeq_guess_forward_backward.argtypes = [_vp, ctypes.c_int, _fp, _fp, _dp, _fp]
    L.qdeq_guess_forward.argtypes = [_vp, ctypes.c_int, _fp, _fp]
    L.qdeq_attach_guess.argtypes = [_vp, _vp, ctypes.c_int]
    L.qdeq_guess_stats.argtypes = [_vp, _llp, _llp, _llp]
    L.qdeq_guess_inject.argtypes = [_vp, ctypes.c_double]
    L.qdeq_guess_infer.argtypes = [_vp, _fp, _fp]
    L.qdeq_projection_enable.argtypes = [_vp, ctypes.c_int]
    L.qdeq_projection_stats.argtypes = [_vp, _llp, _llp, _dp, _ip]
    L.qdeq_launch_ahead_stats.argtypes = [_vp, _llp, _llp, _llp]
    L.qdeq_folded_stop_tests.argtypes = [_vp, _llp]
    L.qdeq_projection_fit.argtypes = [_vp, _dp, _dp, ctypes.c_int, _dp]
    L.qdeq_projection_inject.argtypes = [_vp, ctypes.c_double]
    for name in EXPORTS:
        getattr(L, name)  # every declared symbol must resolve
    _lib = L
    return L


def _c(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _pf(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _hip():
    """The HIP runtime libqdeq.so runs on (already loaded, same soname)."""
    load()
    return ctypes.CDLL("libamdhip64.so.7")


class DeviceArray:
    """A contiguous FP64 array in device memory, for the QDEQ_DEVICE path of
    the drop-in entry points.  Allocated with hipMalloc of the runtime
    libqdeq.so runs on; copies are synchronous."""

    def __init__(self, shape, device: int = 0):
        self.shape = tuple(int(n) for n in np.atleast_1d(shape))
        self.size = int(np.prod(self.shape))
        self.device = device
        self.ptr = ctypes.c_void_p()
        hip = _hip()
        if hip.hipSetDevice(device) != 0 or \
                hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(8 * self.size)) != 0:
            self.ptr = ctypes.c_void_p()
            raise QdeqError(QDEQ_ERR_HIP, f"hipMalloc of {self.size} doubles failed")

    @classmethod
    def from_host(cls, a, device: int = 0) -> "DeviceArray":
        a = _c(a)
        d = cls(a.shape, device)
        d.copy_from(a)
        return d

    def copy_from(self, a) -> None:
        a = _c(a)
        if a.size != self.size:
            raise ValueError(f"copy of {a.size} values into {self.size}")
        if _hip().hipMemcpy(self.ptr, ctypes.c_void_p(a.ctypes.data),
                            ctypes.c_size_t(8 * self.size), 1) != 0:
            raise QdeqError(QDEQ_ERR_HIP, "hipMemcpy to the device failed")

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape)
        if _hip().hipMemcpy(ctypes.c_void_p(out.ctypes.data), self.ptr,
                            ctypes.c_size_t(8 * self.size), 2) != 0:
            raise QdeqError(QDEQ_ERR_HIP, "hipMemcpy to the host failed")
        return out

    def free(self) -> None:
        if self.ptr:
            _hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _vptrs(arrays, sizes, device: int):
    """Pointers and the common pointer space of the arrays of one call: all
    numpy FP64 arrays (QDEQ_HOST) or all DeviceArrays on `device`
    (QDEQ_DEVICE), each with the number of values the entry point reads."""
    host = [isinstance(a, np.ndarray) for a in arrays]
    dev = [isinstance(a, DeviceArray) for a in arrays]
    if not (all(host) or all(dev)):
        raise TypeError("the arrays of one call must all be numpy float64 arrays "
                        "or all DeviceArrays")
    out = []
    for a, n in zip(arrays, sizes):
        if a.size != n:
            raise ValueError(f"array of {a.size} values where {n} are needed")
        if all(host):
            if a.dtype != np.float64 or not a.flags.c_contiguous:
                raise TypeError("expected contiguous float64 arrays")
            out.append(ctypes.c_void_p(a.ctypes.data))
        else:
            if a.device != device:
                raise ValueError(f"array on device {a.device}, context on {device}")
            out.append(a.ptr)
    return out, (QDEQ_HOST if all(host) else QDEQ_DEVICE)


def make_config(params: cfgmod.SolverParams, device: int = 0) -> QdeqConfig:
    c = QdeqConfig()
    load().qdeq_default_config(ctypes.byref(c))
    c.ni, c.nk, c.dh = params.ni, params.nk, params.dh
    c.n_levels = params.n_levels
    c.n_pre, c.n_post, c.n_coarse = params.n_pre, params.n_post, params.n_coarse
    c.relax = params.relax
    c.cycle = params.cycle
    c.precision = params.precision
    c.tol_mode, c.tol = params.tol_mode, params.tol
    c.max_vcycles = params.max_vcycles
    c.ini_dt, c.mindt, c.maxdt = params.ini_dt, params.mindt, params.maxdt
    c.totTol, c.tmax = params.totTol, params.tmax
    c.f0, c.V0, c.L, c.Vini = params.f0, params.V0, params.L, params.Vini
    c.device = device
    return c


def device_synchronize(device: int = 0) -> None:
    """Waits for all work on the device (host-side timing barrier)."""
    st = load().qdeq_device_synchronize(device)
    if st != QDEQ_OK:
        raise QdeqError(st, load().qdeq_global_error().decode())


def build_mu_equi(mu: np.ndarray, device: int = 0) -> np.ndarray:
    """SetMedium's equivalent-medium array, built on the GPU."""
    mu = _c(mu)
    nz, nx = mu.shape
    out = np.zeros((2 * nz - 1, 2 * nx - 1))
    st = load().qdeq_build_mu_equi(nx - 2, nz - 2, _p(mu), _p(out), device)
    if st != QDEQ_OK:
        raise QdeqError(st, load().qdeq_global_error().decode())
    return out


class Engine:
    """One simulation context on one GPU.

    ``problem`` supplies the grid, medium and friction arrays
    (:mod:`qdeqcycle_fdm_amd.config`); ``mu_equi`` is optional (built on the
    device when omitted).
    """

    def __init__(self, problem: cfgmod.Problem, mu_equi: Optional[np.ndarray] = None,
                 *, device: int = 0, **overrides):
        import dataclasses
        params = dataclasses.replace(problem.params, **overrides)
        self.problem = problem
        self.params = params
        self.cfg = make_config(params, device)
        self.nx, self.nz = params.nx, params.nz
        self._mu = _c(problem.mu)
        me = None if mu_equi is None else _c(mu_equi)
        h = _vp()
        L = load()
        st = L.qdeq_create(ctypes.byref(self.cfg), _p(self._mu),
                           None if me is None else _p(me), ctypes.byref(h))
        if st != QDEQ_OK:
            raise QdeqError(st, L.qdeq_global_error().decode())
        self.h = h

    # -- errors --
    def _check(self, st: int):
        if st != QDEQ_OK:
            raise QdeqError(st, load().qdeq_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            load().qdeq_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_tiles(self, tiled: bool = True, tile_x: int = 64, tile_y: int = 32):
        """Fast-cycle execution shape (identical results either way)."""
        self._check(load().qdeq_set_tiles(self.h, 1 if tiled else 0, tile_x, tile_y))

    def autotune_tiles(self, reps: int = 5):
        """Per-level tile shapes of the fast cycle, fastest of the candidates
        (identical results for every shape)."""
        self._check(load().qdeq_autotune_tiles(self.h, reps))
        return self.tile_shapes()

    def kernel_breakdown(self, reps: int = 20) -> list:
        """Average microseconds of each tiled-cycle kernel launched back to
        back: down kernels (finest first), tail, up kernels (to level 0)."""
        n = ctypes.c_int(0)
        buf = (ctypes.c_double * 25)()
        self._check(load().qdeq_kernel_breakdown(self.h, reps, ctypes.byref(n), buf))
        return [buf[i] for i in range(n.value)]

    def set_tile_shapes(self, shapes):
        """Pins per-level shapes [((tx, ty, rows), (tx, ty, rows)), ...]."""
        flat = [v for down, up in shapes for v in (*down, *up)]
        buf = (ctypes.c_int * max(len(flat), 1))(*flat)
        self._check(load().qdeq_set_tile_shapes(self.h, len(shapes), buf))

    def output_open(self, directory: str, every: int = 2):
        """Reference fault output records (output_fault_<k>.bin) in directory."""
        self._check(load().qdeq_output_open(self.h, directory.encode(), every))

    def output_close(self):
        self._check(load().qdeq_output_close(self.h))

    def events_enable(self, v_threshold: float = 1e-3, capacity: int = 4096):
        self._cap = capacity
        self._check(load().qdeq_events_enable(self.h, v_threshold, capacity))

    def events(self) -> list:
        cap = getattr(self, "_cap", 4096)
        buf = (Event * cap)()
        n = ctypes.c_int(0)
        self._check(load().qdeq_events_get(self.h, buf, cap, ctypes.byref(n)))
        return [buf[i].as_dict() for i in range(min(n.value, cap))]

    def tile_shapes(self):
        """[(down (tx, ty, rows), up (tx, ty, rows)) per tile level]."""
        n = ctypes.c_int(0)
        buf = (ctypes.c_int * 72)()
        self._check(load().qdeq_get_tile_shapes(self.h, ctypes.byref(n), buf))
        return [((buf[6 * l], buf[6 * l + 1], buf[6 * l + 2]),
                 (buf[6 * l + 3], buf[6 * l + 4], buf[6 * l + 5]))
                for l in range(n.value)]

    @property
    def uses_tiled_cycle(self) -> bool:
        return bool(load().qdeq_uses_tiled_cycle(self.h))

    def set_kernel_timing(self, which: int):
        """HIP-event timing of one fast-cycle kernel (TIME_* selector)."""
        self._check(load().qdeq_set_kernel_timing(self.h, which))

    def kernel_timing(self):
        """(total_ms, calls) accumulated since set_kernel_timing."""
        ms = ctypes.c_double(0)
        n = ctypes.c_longlong(0)
        self._check(load().qdeq_get_kernel_timing(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def copy_probe(self, n_doubles: int, reps: int = 20) -> float:
        """Average ms of one plain device copy of n_doubles doubles."""
        ms = ctypes.c_double(0)
        self._check(load().qdeq_copy_probe(self.h, n_doubles, reps, ctypes.byref(ms)))
        return ms.value

    def mu_equi(self) -> np.ndarray:
        out = np.zeros((2 * self.nz - 1, 2 * self.nx - 1))
        self._check(load().qdeq_get_mu_equi(self.h, ctypes.c_void_p(out.ctypes.data),
                                            QDEQ_HOST))
        return out

    # -- drop-in functions (reference names) --
    def ode_solve(self, U):
        """ode_solve (odeSolver.cpp:605): in place on U (nz, nx), a numpy
        array or a DeviceArray; returns (vcycles, res0, res)."""
        (ptr,), where = _vptrs([U], [self.nx * self.nz], self.cfg.device)
        vc = ctypes.c_int(0)
        r0 = ctypes.c_double(0)
        r = ctypes.c_double(0)
        self._check(load().qdeq_ode_solve(self.h, ptr, where, ctypes.byref(vc),
                                          ctypes.byref(r0), ctypes.byref(r)))
        return vc.value, r0.value, r.value

    def setBoundary(self, U, slip, slip_shift, t, ratio, dt):
        """setBoundary (odeSolver.cpp:1150): in place on U; U, slip and
        slip_shift all numpy arrays or all DeviceArrays."""
        (pu, ps, pss), where = _vptrs([U, slip, slip_shift],
                                      [self.nx * self.nz, self.nz, self.nz],
                                      self.cfg.device)
        self._check(load().qdeq_set_boundary(self.h, pu, ps, pss, t, ratio, dt, where))

    def eval_stress_on_fault(self, U, mu, tau_n0):
        U, mu, tau_n0 = _c(U), _c(mu), _c(tau_n0)
        Ts = np.zeros(self.nz)
        Tn = np.zeros(self.nz)
        args = [ctypes.c_void_p(a.ctypes.data) for a in (U, mu, tau_n0, Ts, Tn)]
        self._check(load().qdeq_eval_stress(self.h, *args, QDEQ_HOST))
        return Ts, Tn

    def NRsearch(self, Ts, Tn, psi, a, eta, V=None):
        V = np.zeros(self.nz) if V is None else _c(V).copy()
        arrs = [_c(x) for x in (Ts, Tn, psi, a, eta)]
        args = [ctypes.c_void_p(x.ctypes.data) for x in arrs]
        self._check(load().qdeq_nr_search(self.h, *args, ctypes.c_void_p(V.ctypes.data),
                                          QDEQ_HOST))
        return V

    def cal_dpsi(self, psi, V, b):
        arrs = [_c(x) for x in (psi, V, b)]
        d = np.zeros(self.nz)
        args = [ctypes.c_void_p(x.ctypes.data) for x in arrs]
        self._check(load().qdeq_cal_dpsi(self.h, *args, ctypes.c_void_p(d.ctypes.data),
                                         QDEQ_HOST))
        return d

    @staticmethod
    def computeError(s3, p3, s4, p4) -> float:
        s3, p3, s4, p4 = (_c(x) for x in (s3, p3, s4, p4))
        return load().qdeq_compute_error(len(s3) - 2, _p(s3), _p(p3), _p(s4), _p(p4))

    @staticmethod
    def computeStepSize(it, dt, mindt, maxdt, totErr, totTol, errA) -> float:
        e = _c(errA)
        return load().qdeq_compute_step_size(it, dt, mindt, maxdt, totErr, totTol, _p(e))

    # -- RK43 driver --
    def rk43_setup(self):
        pb = self.problem
        self._fault = [_c(pb.a), _c(pb.b), _c(pb.eta), _c(pb.tau_n0), _c(pb.U0)]
        self._check(load().qdeq_rk43_setup(self.h, *[_p(x) for x in self._fault]))

    def rk43_step(self, nsteps: int = 1) -> dict:
        info = StepInfo()
        self._check(load().qdeq_rk43_step(self.h, nsteps, ctypes.byref(info)))
        return info.as_dict()

    def state(self) -> dict:
        info = StepInfo()
        U = np.zeros((self.nz, self.nx))
        out = {k: np.zeros(self.nz) for k in ("slip", "psi", "V", "dpsi", "Ts")}
        self._check(load().qdeq_rk43_get_state(
            self.h, ctypes.byref(info), _p(U), _p(out["slip"]), _p(out["psi"]),
            _p(out["V"]), _p(out["dpsi"]), _p(out["Ts"])))
        out["U"] = U
        out["info"] = info.as_dict()
        return out


def ensemble_step(engines, nsteps: int):
    """Advances nsteps accepted steps of every engine together
    (qdeq_ensemble_step); returns the per-engine status list and raises if
    any failed."""
    n = len(engines)
    hs = (_vp * n)(*[e.h for e in engines])
    st = (ctypes.c_int * n)()
    rc = load().qdeq_ensemble_step(hs, n, nsteps, st)
    status = list(st)
    if rc != QDEQ_OK:
        bad = [i for i, s in enumerate(status) if s != QDEQ_OK]
        msgs = "; ".join(f"member {i}: {load().qdeq_last_error(engines[i].h).decode()}"
                         for i in bad)
        raise QdeqError(rc, msgs or load().qdeq_global_error().decode())
    return status


class Guess:
    """Learned initial guess network (include/qdeq_guess.h): owns the
    weights, the harvested training pairs and the training loop on the GPU."""

    def __init__(self, nk: int, *, device: int = 0, **overrides):
        L = load()
        c = GuessConfig()
        L.qdeq_guess_default_config(ctypes.byref(c), nk)
        c.device = device
        for k, v in overrides.items():
            setattr(c, k, v)
        self.cfg = c
        h = _vp()
        st = L.qdeq_guess_create(ctypes.byref(c), ctypes.byref(h))
        if st != QDEQ_OK:
            raise QdeqError(st, L.qdeq_global_error().decode())
        self.h = h
        nout, f0 = ctypes.c_int(0), ctypes.c_int(0)
        L.qdeq_guess_dims(h, ctypes.byref(nout), ctypes.byref(f0))
        self.nk, self.nout, self.F0 = nk, nout.value, f0.value
        self.nc = (nk - 1) // c.factor + 1

    def _check(self, st: int):
        if st != QDEQ_OK:
            raise QdeqError(st, load().qdeq_guess_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            load().qdeq_guess_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_params(self) -> int:
        return int(load().qdeq_guess_num_params(self.h))

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.num_params, dtype=np.float32)
        self._check(load().qdeq_guess_get_params(self.h, _pf(out)))
        return out

    def set_params(self, flat):
        flat = _f32(flat)
        assert flat.size == self.num_params
        self._check(load().qdeq_guess_set_params(self.h, _pf(flat)))

    @property
    def pairs(self) -> int:
        n = ctypes.c_int(0)
        self._check(load().qdeq_guess_pairs(self.h, ctypes.byref(n)))
        return n.value

    def clear_pairs(self):
        self._check(load().qdeq_guess_clear_pairs(self.h))

    def read_pairs(self, first: int, n: int):
        x = np.zeros((n, self.nk), dtype=np.float32)
        t = np.zeros((n, self.nout), dtype=np.float32)
        s = np.zeros(n)
        self._check(load().qdeq_guess_read_pairs(self.h, first, n, _pf(x), _pf(t), _p(s)))
        return x, t, s

    def train(self, **overrides) -> float:
        tc = TrainConfig()
        load().qdeq_train_default_config(ctypes.byref(tc))
        for k, v in overrides.items():
            setattr(tc, k, v)
        loss = ctypes.c_double(0)
        self._check(load().qdeq_guess_train(self.h, ctypes.byref(tc), ctypes.byref(loss)))
        return loss.value

    def forward_backward(self, x, t):
        x, t = _f32(x), _f32(t)
        B = x.shape[0]
        loss = ctypes.c_double(0)
        g = np.zeros(self.num_params, dtype=np.float32)
        self._check(load().qdeq_guess_forward_backward(
            self.h, B, _pf(x), _pf(t), ctypes.byref(loss), _pf(g)))
        return loss.value, g

    def forward(self, x) -> np.ndarray:
        x = _f32(x)
        y = np.zeros((x.shape[0], self.nout), dtype=np.float32)
        self._check(load().qdeq_guess_forward(self.h, x.shape[0], _pf(x), _pf(y)))
        return y


def _engine_attach_guess(self, guess: Optional[Guess], mode: int):
    """Attaches a Guess for the RK43 driver's solves (GUESS_* mode)."""
    self._guess = guess
    self._check(load().qdeq_attach_guess(self.h, guess.h if guess else None, mode))


def _engine_guess_stats(self) -> dict:
    p, a, h = ctypes.c_longlong(0), ctypes.c_longlong(0), ctypes.c_longlong(0)
    self._check(load().qdeq_guess_stats(self.h, ctypes.byref(p), ctypes.byref(a),
                                        ctypes.byref(h)))
    return dict(predicted=p.value, accepted=a.value, harvested=h.value)


def _engine_projection_enable(self, slots: int):
    """Projection start of the RK43 stage solves from the last `slots`
    solutions (0 disables); runs one solve for the unit remote response."""
    self._check(load().qdeq_projection_enable(self.h, slots))


def _engine_projection_stats(self) -> dict:
    p, a = ctypes.c_longlong(0), ctypes.c_longlong(0)
    r, n = ctypes.c_double(0.0), ctypes.c_int(0)
    self._check(load().qdeq_projection_stats(self.h, ctypes.byref(p), ctypes.byref(a),
                                             ctypes.byref(r), ctypes.byref(n)))
    return dict(offered=p.value, used=a.value, fit_ratio=r.value, columns=n.value)


def _engine_projection_fit(self, g_new, g_old):
    """Coefficients of the projection start's least-squares fit (the RK43
    driver's kernel) for g_new (N) against g_old (cnt x N, newest first):
    returns (c (cnt,), misfit ratio, columns kept)."""
    g_new = np.ascontiguousarray(g_new, dtype=np.float64)
    g_old = np.ascontiguousarray(g_old, dtype=np.float64)
    cnt = g_old.shape[0]
    out = np.zeros(cnt + 2)
    self._check(load().qdeq_projection_fit(self.h, _p(g_new), _p(g_old), cnt, _p(out)))
    return out[:cnt], float(out[cnt]), int(out[cnt + 1])


def _engine_guess_inject(self, factor: float):
    self._check(load().qdeq_guess_inject(self.h, factor))


def _engine_guess_infer(self, x) -> np.ndarray:
    x = _f32(x)
    y = np.zeros(self._guess.nout, dtype=np.float32)
    self._check(load().qdeq_guess_infer(self.h, _pf(x), _pf(y)))
    return y


Engine.attach_guess = _engine_attach_guess
Engine.guess_stats = _engine_guess_stats
Engine.guess_inject = _engine_guess_inject
Engine.projection_enable = _engine_projection_enable
Engine.projection_stats = _engine_projection_stats


def _engine_launch_ahead_stats(self) -> dict:
    """Solves launched ahead behind the previous stage's solve and taken
    over, or abandoned (the previous solve needed more V-cycles); of those
    taken, next_step counts the next step's stage 2 queued behind an
    accepted step's final stage 1."""
    t, a, n = ctypes.c_longlong(0), ctypes.c_longlong(0), ctypes.c_longlong(0)
    self._check(load().qdeq_launch_ahead_stats(self.h, ctypes.byref(t), ctypes.byref(a),
                                               ctypes.byref(n)))
    return dict(taken=t.value, abandoned=a.value, next_step=n.value)


def _engine_folded_stop_tests(self) -> int:
    """Stop tests run inside the next V-cycle of their batch so far."""
    n = ctypes.c_longlong(0)
    self._check(load().qdeq_folded_stop_tests(self.h, ctypes.byref(n)))
    return n.value


Engine.launch_ahead_stats = _engine_launch_ahead_stats
Engine.folded_stop_tests = _engine_folded_stop_tests
Engine.projection_fit = _engine_projection_fit
Engine.projection_inject = lambda self, factor: self._check(
    load().qdeq_projection_inject(self.h, factor))
Engine.guess_infer = _engine_guess_infer
