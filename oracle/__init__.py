"""ctypes bindings to the CPU FP64 oracle (liboracle_qdeq.so).

TEST INFRASTRUCTURE ONLY: the oracle is the checker the HIP product library
is compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module; the product package never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# QDEQ_ORACLE_LIB selects another build of the same source (the RK43 checks
# load builds with planted mistakes this way, in a subprocess)
LIB_PATH = os.environ.get("QDEQ_ORACLE_LIB") or os.path.join(_HERE, "liboracle_qdeq.so")

ORC_TOL_ABS, ORC_TOL_REL = 0, 1
ORC_OK, ORC_ERR_ARG, ORC_ERR_NAN_FRICTION, ORC_ERR_MG_NOT_CONVERGED = 0, 1, 2, 3

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("ni", ctypes.c_int), ("nk", ctypes.c_int), ("dh", ctypes.c_double),
        ("n_levels", ctypes.c_int), ("n_cycles", ctypes.c_int),
        ("relax", ctypes.c_double), ("tol_mode", ctypes.c_int),
        ("tol", ctypes.c_double), ("max_vcycles", ctypes.c_int),
        ("ini_dt", ctypes.c_double), ("mindt", ctypes.c_double),
        ("maxdt", ctypes.c_double), ("totTol", ctypes.c_double),
        ("tmax", ctypes.c_double), ("f0", ctypes.c_double),
        ("V0", ctypes.c_double), ("L", ctypes.c_double),
        ("Vini", ctypes.c_double), ("cycle", ctypes.c_int),
        ("n_pre", ctypes.c_int), ("n_post", ctypes.c_int),
        ("n_coarse", ctypes.c_int),
    ]


def build(force: bool = False) -> str:
    """Compiles the oracle with its Makefile (gcc)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all", "-C", _HERE],
                       check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_mu_equi.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_level_dims.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_int, _ip, _ip, _ip, _dp]
        L.orc_rbgs.argtypes = [_dp, _dp, _dp, ctypes.c_int, ctypes.c_int,
                               ctypes.c_double, ctypes.c_int, ctypes.c_int,
                               ctypes.c_double]
        L.orc_residual.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.orc_restrict.argtypes = [_dp, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int]
        L.orc_prolong.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int]
        L.orc_nrm2.argtypes = [_dp, ctypes.c_long]
        L.orc_nrm2.restype = ctypes.c_double
        L.orc_mg_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    _dp]
        L.orc_mg_create.restype = ctypes.c_void_p
        L.orc_mg_destroy.argtypes = [ctypes.c_void_p]
        L.orc_mg_set_cycle.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int]
        L.orc_mg_set_cycle.restype = None
        L.orc_prolong_full.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int]
        L.orc_mg_vcycle.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_ode_solve.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_int, _ip, _dp, _dp]
        L.orc_ode_solve.restype = ctypes.c_int
        L.orc_extend_crew.argtypes = [_dp, ctypes.c_int, ctypes.c_int]
        L.orc_set_boundary.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                       ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double]
        L.orc_eval_stress.argtypes = [_dp, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_double, _dp, _dp]
        L.orc_nr_search.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp,
                                    ctypes.c_double, _dp, _ip]
        L.orc_nr_search.restype = ctypes.c_int
        L.orc_cal_dpsi.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, _dp]
        L.orc_compute_error.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_compute_error.restype = ctypes.c_double
        L.orc_compute_step_size.argtypes = [ctypes.c_int, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_double, _dp]
        L.orc_compute_step_size.restype = ctypes.c_double
        L.orc_set_libm.argtypes = [ctypes.c_int]
        L.orc_set_libm.restype = None
        L.orc_elem.argtypes = [ctypes.c_int, ctypes.c_double]
        L.orc_elem.restype = ctypes.c_double
        L.orc_tableau.argtypes = [_dp]
        L.orc_tableau.restype = None
        L.orc_sim_create.argtypes = [ctypes.POINTER(OrcParams), _dp, _dp, _dp,
                                     _dp, _dp, _dp, _dp]
        L.orc_sim_create.restype = ctypes.c_void_p
        L.orc_sim_destroy.argtypes = [ctypes.c_void_p]
        L.orc_sim_init.argtypes = [ctypes.c_void_p]
        L.orc_sim_init.restype = ctypes.c_int
        L.orc_sim_set_state.argtypes = [ctypes.c_void_p, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
        L.orc_sim_set_state.restype = ctypes.c_int
        L.orc_sim_step.argtypes = [ctypes.c_void_p]
        L.orc_sim_step.restype = ctypes.c_int
        L.orc_sim_scalars.argtypes = [ctypes.c_void_p, _dp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_threads.restype = ctypes.c_int
        L.orc_sim_set_dt.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_sim_set_dt.restype = None
        L.orc_sim_last_attempt.argtypes = [ctypes.c_void_p] + [_dp] * 8
        L.orc_sim_last_attempt.restype = None
        L.orc_sim_stage_log.argtypes = [ctypes.c_void_p, _dp]
        L.orc_sim_stage_log.restype = ctypes.c_int
        L.orc_sim_arrays.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _c(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


ELEM = {"exp": 0, "sinh": 1, "cosh": 2, "log": 3, "log1p": 4, "asinh": 5}


def elem(name: str, x: float) -> float:
    """The engine's elementary function ``name`` (oracle copy)."""
    return lib().orc_elem(ELEM[name], float(x))


def set_libm(use_glibc: bool) -> None:
    """Selects glibc (True) or the engine's algorithms (False) for the
    exp/sinh/cosh/asinh calls of NRsearch and cal_dpsi."""
    lib().orc_set_libm(1 if use_glibc else 0)


def set_threads(n: int) -> int:
    """OpenMP threads of the oracle's multigrid loops (results do not depend
    on it); returns the count in use."""
    return lib().orc_set_threads(int(n))


def tableau() -> dict:
    """The oracle's RK43 coefficients: c, b, hb (6,), a (6, 6)."""
    out = np.zeros(54)
    lib().orc_tableau(_p(out))
    return dict(c=out[0:6], b=out[6:12], hb=out[12:18], a=out[18:].reshape(6, 6))


def level_dims(ni, nk, dh, n_levels):
    N = (ctypes.c_int * n_levels)()
    K = (ctypes.c_int * n_levels)()
    FC = (ctypes.c_int * n_levels)()
    DH = (ctypes.c_double * n_levels)()
    lib().orc_level_dims(ni, nk, dh, n_levels, N, K, FC, DH)
    return list(N), list(K), list(FC), list(DH)


def mu_equi(mu: np.ndarray) -> np.ndarray:
    """SetMedium's equivalent-medium array from the nodal (nz, nx) mu."""
    mu = _c(mu)
    nz, nx = mu.shape
    out = np.zeros((2 * nz - 1, 2 * nx - 1))
    lib().orc_mu_equi(nx - 2, nz - 2, _p(mu), _p(out))
    return out


def rbgs(u, f, me, ni, nk, dh, level, phase, relax=1.0):
    lib().orc_rbgs(_p(u), _p(_c(f)), _p(_c(me)), ni, nk, dh, level, phase, relax)


def residual(r, u, f, me, ni, nk, dh, level):
    lib().orc_residual(_p(r), _p(_c(u)), _p(_c(f)), _p(_c(me)), ni, nk, dh, level)


def nrm2(x) -> float:
    """Euclidean norm as the oracle's solves take it (orc_nrm2)."""
    x = _c(x)
    return lib().orc_nrm2(_p(x), x.size)


def restrict(r2, r, e2, ni, nk, level):
    lib().orc_restrict(_p(r2), _p(_c(r)), _p(e2), ni, nk, level)


def prolong(u, e2, ni, nk, level):
    lib().orc_prolong(_p(u), _p(_c(e2)), ni, nk, level)


def set_boundary(U, ni, nk, slip, slip_shift, Vini, t, ratio, dt):
    lib().orc_set_boundary(_p(U), ni, nk, _p(_c(slip)), _p(_c(slip_shift)),
                           Vini, t, ratio, dt)


def eval_stress(U, mu, tau_n0, ni, nk, dh):
    nz = nk + 2
    Ts = np.zeros(nz)
    Tn = np.zeros(nz)
    lib().orc_eval_stress(_p(_c(U)), _p(_c(mu)), _p(_c(tau_n0)), ni, nk, dh,
                          _p(Ts), _p(Tn))
    return Ts, Tn


def nr_search(Ts, Tn, psi, a, eta, V0, V=None):
    """Returns (status, V, iterations)."""
    nz = len(Ts)
    V = np.zeros(nz) if V is None else _c(V).copy()
    iters = np.zeros(nz, dtype=np.int32)
    st = lib().orc_nr_search(nz - 2, _p(_c(Ts)), _p(_c(Tn)), _p(_c(psi)),
                             _p(_c(a)), _p(_c(eta)), V0, _p(V),
                             iters.ctypes.data_as(_ip))
    return st, V, iters


def cal_dpsi(psi, V, b, f0, V0, L):
    nz = len(psi)
    d = np.zeros(nz)
    lib().orc_cal_dpsi(nz - 2, _p(_c(psi)), _p(_c(V)), _p(_c(b)), f0, V0, L, _p(d))
    return d


def compute_error(s3, p3, s4, p4):
    nz = len(s3)
    return lib().orc_compute_error(nz - 2, _p(_c(s3)), _p(_c(p3)), _p(_c(s4)),
                                   _p(_c(p4)))


def compute_step_size(it, dt, mindt, maxdt, totErr, totTol, errA):
    e = _c(errA)
    return lib().orc_compute_step_size(it, dt, mindt, maxdt, totErr, totTol, _p(e))


class MG:
    """Oracle multigrid solver (ode_solve)."""

    def __init__(self, ni, nk, dh, n_levels, mu_equi_arr, n_cycles=32, relax=1.0):
        self._me = _c(mu_equi_arr)
        self.ni, self.nk = ni, nk
        self.h = lib().orc_mg_create(ni, nk, dh, n_levels, n_cycles, relax,
                                     _p(self._me))
        if not self.h:
            raise ValueError("orc_mg_create failed")

    def set_cycle(self, n_pre, n_post, n_coarse, full_top=True):
        """The engine's fast cycle shape (no reference counterpart)."""
        lib().orc_mg_set_cycle(self.h, n_pre, n_post, n_coarse, 1 if full_top else 0)

    def vcycle(self, u, f):
        lib().orc_mg_vcycle(self.h, _p(u), _p(_c(f)))

    def solve(self, U, tol_mode=ORC_TOL_ABS, tol=1e-9, max_vcycles=0):
        """Solves in place on the (nz, nx) array U; returns
        (status, vcycles, res0, res)."""
        assert U.dtype == np.float64 and U.flags.c_contiguous
        vc = ctypes.c_int(0)
        r0 = ctypes.c_double(0)
        r = ctypes.c_double(0)
        st = lib().orc_ode_solve(self.h, _p(U), tol_mode, tol, max_vcycles,
                                 ctypes.byref(vc), ctypes.byref(r0), ctypes.byref(r))
        return st, vc.value, r0.value, r.value

    def __del__(self):
        try:
            if self.h:
                lib().orc_mg_destroy(self.h)
                self.h = None
        except Exception:
            pass


SCALAR_NAMES = ("t", "dt", "it", "errA0", "errA1", "totErr", "attempts",
                "vcycles", "solves", "V_max", "Ts_max")


class Sim:
    """Oracle RK43 simulation (odeSolver.cpp:1277-1674)."""

    def __init__(self, problem, mu_equi_arr=None, *, tol_mode=None, tol=None,
                 max_vcycles=None, n_cycles=None, cycle=None, n_pre=None,
                 n_post=None, n_coarse=None, relax=None):
        p = problem.params
        self.problem = problem
        me = mu_equi(problem.mu) if mu_equi_arr is None else mu_equi_arr
        self._arrs = [_c(problem.mu), _c(me), _c(problem.a), _c(problem.b),
                      _c(problem.eta), _c(problem.tau_n0), _c(problem.U0)]
        self.params = OrcParams(
            ni=p.ni, nk=p.nk, dh=p.dh, n_levels=p.n_levels,
            n_cycles=p.n_cycles if n_cycles is None else n_cycles,
            relax=p.relax if relax is None else relax,
            tol_mode=p.tol_mode if tol_mode is None else tol_mode,
            tol=p.tol if tol is None else tol,
            max_vcycles=p.max_vcycles if max_vcycles is None else max_vcycles,
            ini_dt=p.ini_dt, mindt=p.mindt, maxdt=p.maxdt, totTol=p.totTol,
            tmax=p.tmax, f0=p.f0, V0=p.V0, L=p.L, Vini=p.Vini,
            cycle=p.cycle if cycle is None else cycle,
            n_pre=p.n_pre if n_pre is None else n_pre,
            n_post=p.n_post if n_post is None else n_post,
            n_coarse=p.n_coarse if n_coarse is None else n_coarse)
        self.h = lib().orc_sim_create(ctypes.byref(self.params),
                                      *[_p(a) for a in self._arrs])
        self.nx, self.nz = p.nx, p.nz

    def init(self):
        st = lib().orc_sim_init(self.h)
        if st != ORC_OK:
            raise RuntimeError(f"oracle init failed: status {st}")

    def set_state(self, t, dt, it, errA0, errA1, U, slip, psi):
        """Restart from a saved state (recomputes stage 1)."""
        self._keep = [_c(U), _c(slip), _c(psi)]
        st = lib().orc_sim_set_state(self.h, t, dt, int(it), errA0, errA1,
                                     *[_p(a) for a in self._keep])
        if st != ORC_OK:
            raise RuntimeError(f"oracle restart failed: status {st}")

    def step(self):
        st = lib().orc_sim_step(self.h)
        if st != ORC_OK:
            raise RuntimeError(f"oracle step failed: status {st}")

    def set_dt(self, dt: float):
        """Step size of the next attempt (test hook)."""
        lib().orc_sim_set_dt(self.h, float(dt))

    def last_attempt(self) -> dict:
        """Stage rates V, dpsi and stage states slip, psi of stages 2..6 of
        the last attempt (rows 2..6 of (7, nz) arrays; rows 0-1 zero) and its
        y3 / y4 solutions."""
        nz = self.nz
        bufs = {k: np.zeros((6, nz)) for k in ("V", "dpsi", "slip", "psi")}
        ys = {k: np.zeros(nz) for k in ("y3s", "y3p", "y4s", "y4p")}
        lib().orc_sim_last_attempt(self.h, _p(bufs["V"]), _p(bufs["dpsi"]),
                                   _p(bufs["slip"]), _p(bufs["psi"]),
                                   *[_p(ys[k]) for k in ("y3s", "y3p", "y4s", "y4p")])
        out = {k: np.vstack([np.zeros((1, nz)), v]) for k, v in bufs.items()}
        out.update(ys)
        return out

    def stage_log(self) -> np.ndarray:
        """(n, 5): t, dt, ratio, U at row 1 of the remote and of the fault
        column right after setBoundary, for the last n <= 8 stage calls."""
        o = np.zeros(40)
        n = lib().orc_sim_stage_log(self.h, _p(o))
        return o[:5 * n].reshape(n, 5)

    def scalars(self) -> dict:
        o = (ctypes.c_double * 11)()
        lib().orc_sim_scalars(self.h, o)
        return dict(zip(SCALAR_NAMES, list(o)))

    def arrays(self) -> dict:
        U = np.zeros((self.nz, self.nx))
        out = {k: np.zeros(self.nz) for k in ("slip", "psi", "V", "dpsi", "Ts")}
        lib().orc_sim_arrays(self.h, _p(U), _p(out["slip"]), _p(out["psi"]),
                             _p(out["V"]), _p(out["dpsi"]), _p(out["Ts"]))
        out["U"] = U
        return out

    def __del__(self):
        try:
            if self.h:
                lib().orc_sim_destroy(self.h)
                self.h = None
        except Exception:
            pass
