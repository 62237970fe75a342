"""Problem set-up for the antiplane quasi-dynamic earthquake-cycle engine.

The reference hard-codes one problem (SetParams.cpp:3-30, fault.cpp:73-135,
odeSolver.cpp:272-319) and reads its initial displacement from a file that is
not part of the repository (odeSolver.cpp:103).  This module builds the same
set of inputs as plain arrays, in the reference's layouts:

* grid arrays: ``(nz, nx)`` row-major, x fastest, one ghost layer
  (``nx = ni + 2``, ``nz = nk + 2``; odeSolver.cpp:24-37);
* fault arrays: length ``nz``, valid on rows ``1..nk`` (Alloc.cpp:3-40);
* units: displacement m, spacing m, coordinates km, mu GPa, stress MPa,
  eta MPa*s/m, slip rate m/s (coord.cpp:18,27; fault.cpp:153).

Named workloads:

* ``bp1_129``   -- 48 km square, 129^2 nodes, uniform mu = 32 GPa, 4 levels.
* ``basin_1025`` -- same physics, 1025^2 nodes, low-rigidity basin
  (mu = 20 GPa in the top 2 km within 10 km of the fault, 500 m tanh edges),
  7 levels, 1e-10 relative elliptic residual.
* ``basin_2049`` -- 2049^2 nodes, 8 levels, same medium.
* ``ref_basin_801`` -- the reference's own hard-coded problem (801^2,
  dh = 30 m, elliptical basin, b tapering 0.020 -> 0 over 12-17 km).
"""
from __future__ import annotations

import dataclasses
import json
import math
import os
from typing import Optional

import numpy as np

TOL_ABS = 0  # reference stop rule ||r||_2 <= tol (odeSolver.cpp:689-696)
TOL_REL = 1  # ||r||_2 <= tol * ||r(u_interior = 0)||_2

PREC_FP64 = 0     # plain FP64 V-cycles
PREC_FP32_IR = 1  # FP32 V-cycles inside FP64 iterative refinement

CYCLE_REFERENCE = 0  # the reference V-cycle exactly (odeSolver.cpp:1101-1147)
CYCLE_FAST = 1       # same smoother/transfers, tunable sweeps, full prolongation


def tmax_reference() -> float:
    """tmax of SetParams.cpp:5-12: 3000 years of 360 days plus one minute."""
    t_year, t_month, t_day, t_hour, t_min, t_sec = 3000, 0, 0, 0, 1, 0
    return t_sec + 60.0 * (t_min + 60.0 * (t_hour + 24.0 * (
        t_day + 30.0 * (t_month + 12.0 * t_year))))


@dataclasses.dataclass
class SolverParams:
    """Scalars of one simulation (SetParams.cpp:5-27, fault.cpp:85-89,
    odeSolver.cpp:2-4)."""

    ni: int
    nk: int
    dh: float
    n_levels: int
    n_cycles: int = 32          # sweeps per smoothing, odeSolver.cpp:4
    relax: float = 1.0          # odeSolver.cpp:2
    tol_mode: int = TOL_ABS
    tol: float = 1.0e-9         # odeSolver.cpp:689
    max_vcycles: int = 200
    ini_dt: float = 1.0e-3      # SetParams.cpp:14
    mindt: float = 1.0e-5
    maxdt: float = 1.0e8
    totTol: float = 1.0e-7
    tmax: float = dataclasses.field(default_factory=tmax_reference)
    f0: float = 0.6             # fault.cpp:85-89
    V0: float = 1.0e-6
    L: float = 0.008
    Vini: float = 1.0e-9
    max_steps: int = 100000000  # maxNumSteps, SetParams.cpp:18
    it_save_jump: int = 2       # output record every it_save_jump steps, :27
    # engine-only knobs (no reference counterpart)
    precision: int = PREC_FP64
    cycle: int = CYCLE_REFERENCE
    n_pre: int = 32
    n_post: int = 32
    n_coarse: int = 32

    @property
    def nx(self) -> int:
        return self.ni + 2

    @property
    def nz(self) -> int:
        return self.nk + 2


@dataclasses.dataclass
class Problem:
    """All inputs of one simulation."""

    name: str
    params: SolverParams
    mu: np.ndarray        # (nz, nx) nodal shear modulus, GPa; ghosts 0
    a: np.ndarray         # (nz,)
    b: np.ndarray         # (nz,)
    eta: np.ndarray       # (nz,) radiation damping, MPa*s/m
    tau_n0: np.ndarray    # (nz,) normal stress, MPa (negative = compressive)
    U0: np.ndarray        # (nz, nx) initial displacement with ghosts, m
    mu_equi: Optional[np.ndarray] = None  # (2nz-1, 2nx-1), filled by caller


def coords(ni: int, nk: int, dh: float):
    """Node coordinates in km (coord.cpp:13-29): x = (i-1) dh, z = (k-nk) dh."""
    nx, nz = ni + 2, nk + 2
    nk2 = nk + 1
    i = np.arange(nx, dtype=np.float64)
    k = np.arange(nz, dtype=np.float64)
    x = (i - 1) * dh * 1.0e-3
    z = (k + 1 - nk2) * dh * 1.0e-3
    X = np.broadcast_to(x[None, :], (nz, nx)).copy()
    Z = np.broadcast_to(z[:, None], (nz, nx)).copy()
    return X, Z


def fill_ghosts(U: np.ndarray) -> np.ndarray:
    """Ghost layer of an (nz, nx) field: linear extrapolation of the ghost
    columns, then mirrored ghost rows (the rule of odeSolver.cpp:612-650)."""
    U = U.copy()
    U[:, 0] = 2.0 * U[:, 1] - U[:, 2]
    U[:, -1] = 2.0 * U[:, -2] - U[:, -3]
    U[0, :] = U[2, :]
    U[-1, :] = U[-3, :]
    return U


def steady_prestress(sigma_n: float, a_max: float, b: float, f0: float,
                     V0: float, Vini: float, eta: float) -> float:
    """Uniform prestress that keeps V = Vini where a = a_max at steady state:
    tau0 = sigma_n a_max asinh[Vini/(2V0) exp((f0 + b ln(V0/Vini))/a_max)]
    + eta Vini."""
    psi0 = f0 + b * math.log(V0 / Vini)
    return sigma_n * a_max * math.asinh(
        Vini / (2.0 * V0) * math.exp(psi0 / a_max)) + eta * Vini


def _bp1_friction(depth_km: np.ndarray):
    a = np.where(depth_km <= 15.0, 0.010,
                 np.where(depth_km < 18.0, 0.010 + 0.015 * (depth_km - 15.0) / 3.0,
                          0.025))
    b = np.full_like(depth_km, 0.015)
    return a, b


def make_bp1_problem(n: int = 129, *, levels: Optional[int] = None,
                     basin: bool = False, name: Optional[str] = None,
                     Dc: float = 0.008, sigma_n: float = 50.0,
                     b_seis: Optional[float] = None,
                     tol_mode: int = TOL_ABS, tol: float = 1.0e-9,
                     cs: float = 3.464) -> Problem:
    """48 km x 48 km antiplane problem on an n x n grid (n = 2^p + 1).

    Friction: a = 0.010 above 15 km depth, rising linearly to 0.025 at 18 km;
    b = 0.015 (or ``b_seis`` above 15 km); sigma_n, Dc as given; f0 = 0.6,
    V0 = 1e-6, Vini = 1e-9.  Radiation damping eta = mu/(2 cs) with the nodal
    mu on the fault.  ``basin`` adds the low-rigidity layer
    mu = 32 - 12 S, S = 1/4 [1 - tanh((x-10)/0.5)] [1 - tanh((d-2)/0.5)].
    Initial displacement: the uniform-prestress field U0 = tau0 x/(32e3)."""
    ni = nk = n
    dh = 48000.0 / (n - 1)
    if levels is None:
        levels = int(round(math.log2((n - 1) / 16))) + 1  # down to 17^2
    p = SolverParams(ni=ni, nk=nk, dh=dh, n_levels=levels, tol_mode=tol_mode,
                     tol=tol, L=Dc)
    nx, nz = p.nx, p.nz
    X, Z = coords(ni, nk, dh)
    depth = -Z
    mu = np.zeros((nz, nx))
    interior = (slice(1, nk + 1), slice(1, ni + 1))
    if basin:
        S = 0.25 * (1.0 - np.tanh((X - 10.0) / 0.5)) * (1.0 - np.tanh((depth - 2.0) / 0.5))
        mu[interior] = (32.0 - 12.0 * S)[interior]
    else:
        mu[interior] = 32.0
    dcol = depth[:, 1]
    a, b = _bp1_friction(dcol)
    if b_seis is not None:
        b = np.where(dcol <= 15.0, b_seis, b)
    a[0] = a[-1] = 0.0
    b[0] = b[-1] = 0.0
    eta = np.zeros(nz)
    eta[1:nk + 1] = mu[1:nk + 1, 1] / (2.0 * cs)
    tau_n0 = np.zeros(nz)
    tau_n0[1:nk + 1] = -sigma_n
    eta_far = 32.0 / (2.0 * cs)
    tau0 = steady_prestress(sigma_n, 0.025, 0.015, p.f0, p.V0, p.Vini, eta_far)
    xm = (np.arange(nx) - 1) * dh
    U = np.zeros((nz, nx))
    U[1:nk + 1, 1:ni + 1] = (tau0 * xm[1:ni + 1] / (32.0 * 1.0e3))[None, :]
    U0 = fill_ghosts(U)
    if name is None:
        name = f"{'basin' if basin else 'bp1'}_{n}"
    return Problem(name=name, params=p, mu=mu, a=a, b=b, eta=eta,
                   tau_n0=tau_n0, U0=U0)


def make_reference_problem(ni: int = 801, nk: int = 801, dh: float = 30.0,
                           U0: Optional[np.ndarray] = None) -> Problem:
    """The reference's own hard-coded problem: elliptical tanh basin
    (odeSolver.cpp:294-319), b = 0.020 tapering to 0 between 12 and 17 km,
    a = 0.015, sigma_n = 50 MPa, eta = sqrt(rho mu)/2 (fault.cpp:85-120),
    6 levels (odeSolver.cpp:3).  The reference's U0 file is not available;
    if ``U0`` (nk, ni) is not given, a uniform-prestress linear field is
    used."""
    p = SolverParams(ni=ni, nk=nk, dh=dh, n_levels=6)
    nx, nz = p.nx, p.nz
    X, Z = coords(ni, nk, dh)
    mu_out, mu_in, rho_out, rho_in = 36.0, 8.0, 2.8, 2.0
    W, D = 24.0, 8.0
    c = 0.5 * W / D
    rr = 0.25 * W * W
    rw = 1.0 + 0.5 * W / D
    r = X * X + c * c * Z * Z
    temp = (r - rr) / rw
    mu = np.zeros((nz, nx))
    rho = np.zeros((nz, nx))
    sl = (slice(1, nk + 1), slice(1, ni + 1))
    mu[sl] = (0.5 * (mu_out - mu_in) * (np.tanh(temp) + 1.0) + mu_in)[sl]
    rho[sl] = (0.5 * (rho_out - rho_in) * (np.tanh(temp) + 1.0) + rho_in)[sl]
    depth = -Z[:, 1]
    b0, b_max = 0.020, 0.0
    b = np.where(depth <= 12.0, b0,
                 np.where(depth < 17.0, b0 + (b_max - b0) * (depth - 12.0) / 5.0, b_max))
    a = np.full(nz, 0.015)
    eta = 0.5 * np.sqrt(rho[:, 1] * mu[:, 1])
    for arr in (a, b, eta):
        arr[0] = arr[-1] = 0.0
    tau_n0 = np.zeros(nz)
    tau_n0[1:nk + 1] = -50.0
    U = np.zeros((nz, nx))
    if U0 is None:
        tau0 = steady_prestress(50.0, 0.015, 0.0, p.f0, p.V0, p.Vini, 0.0)
        xm = (np.arange(nx) - 1) * dh
        U[1:nk + 1, 1:ni + 1] = (tau0 * xm[1:ni + 1] / (mu_out * 1.0e3))[None, :]
    else:
        U[1:nk + 1, 1:ni + 1] = np.asarray(U0, dtype=np.float64).reshape(nk, ni)
    return Problem(name="ref_basin_801", params=p, mu=mu, a=a, b=b, eta=eta,
                   tau_n0=tau_n0, U0=fill_ghosts(U))


def read_reference_u0(path: str, ni: int, nk: int) -> np.ndarray:
    """Reads the reference's initial-displacement file: ni*nk FP64 values,
    x fastest, rows from the bottom (odeSolver.cpp:96-127).  Returns
    (nk, ni)."""
    data = np.fromfile(path, dtype=np.float64)
    if data.size != ni * nk:
        raise ValueError(f"{path}: expected {ni * nk} doubles, found {data.size}")
    return data.reshape(nk, ni)


def ensemble_members(count: int = 64, seed: int = 0,
                     sizes=(257, 513, 1025)):
    """Seeded ensemble draws: grid from ``sizes``, Dc log-uniform in
    [4, 16] mm, sigma_n uniform in [25, 75] MPa, b uniform in
    [0.013, 0.018] above 15 km depth."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(count):
        n = int(rng.choice(np.asarray(sizes)))
        Dc = float(math.exp(rng.uniform(math.log(0.004), math.log(0.016))))
        sig = float(rng.uniform(25.0, 75.0))
        bs = float(rng.uniform(0.013, 0.018))
        out.append(dict(n=n, Dc=Dc, sigma_n=sig, b_seis=bs, member=m))
    return out


# ---------------------------------------------------------------------------
# Parameter files
# ---------------------------------------------------------------------------
#
# The reference sets its parameters in code (setparams, SetParams.cpp:3-30;
# the friction and medium in fault.cpp:73-135 and odeSolver.cpp:272-319) and
# its README announces a parameter-file reader.  A parameter file here is a
# JSON object whose keys keep the reference's names where it has them:
#
#   "problem"      "bp1" | "basin" | "reference"  (required: which set-up)
#   "ni", "nk", "dh"                      grid (GRID, SetParams.cpp:21-23);
#                                         bp1/basin: ni = nk, dh = 48 km/(n-1)
#   "t_year" .. "t_sec" or "tmax"         run length (SetParams.cpp:5-12)
#   "ini_dt", "mindt", "maxdt", "totTol", "maxNumSteps", "it_save_jump"
#   "f0", "V0", "Dc", "Vini"              friction constants (fault.cpp:85-89)
#   "sigma_n", "b_seis"                   bp1/basin: normal stress, b above 15 km
#   "n_levels", "n_cycles", "relax"       multigrid (odeSolver.cpp:2-4)
#   "tol_mode" ("abs" | "rel"), "tol", "max_vcycles"
#   "cycle" ("reference" | "fast"), "n_pre", "n_post", "n_coarse"
#   "precision" ("fp64" | "fp32_ir")
#   "u0_file"                             initial displacement in the
#                                         reference's format (odeSolver.cpp:
#                                         96-128), relative to the file
#
# Unknown keys and values of the wrong type are errors.

_TIME_KEYS = ("t_year", "t_month", "t_day", "t_hour", "t_min", "t_sec")
_INT_KEYS = {"ni", "nk", "maxNumSteps", "it_save_jump", "n_levels", "n_cycles",
             "max_vcycles", "n_pre", "n_post", "n_coarse", *_TIME_KEYS}
_FLOAT_KEYS = {"dh", "tmax", "ini_dt", "mindt", "maxdt", "totTol", "f0", "V0",
               "Dc", "Vini", "sigma_n", "b_seis", "relax", "tol"}
_STR_KEYS = {"problem": ("bp1", "basin", "reference"), "tol_mode": ("abs", "rel"),
             "cycle": ("reference", "fast"), "precision": ("fp64", "fp32_ir"),
             "u0_file": None, "name": None}


class ParamsError(ValueError):
    pass


def _check_types(d: dict, where: str) -> None:
    for k, v in d.items():
        if k in _INT_KEYS:
            if isinstance(v, bool) or not isinstance(v, int):
                raise ParamsError(f"{where}: {k} must be an integer, got {v!r}")
        elif k in _FLOAT_KEYS:
            if isinstance(v, bool) or not isinstance(v, (int, float)):
                raise ParamsError(f"{where}: {k} must be a number, got {v!r}")
        elif k in _STR_KEYS:
            allowed = _STR_KEYS[k]
            if not isinstance(v, str) or (allowed and v not in allowed):
                raise ParamsError(f"{where}: {k} must be one of {allowed}, got {v!r}")
        else:
            raise ParamsError(f"{where}: unknown parameter {k!r}")


def problem_from_params(d: dict, base_dir: str = ".", where: str = "parameters") -> Problem:
    """Builds the Problem a parameter dictionary describes (keys above)."""
    _check_types(d, where)
    if "problem" not in d:
        raise ParamsError(f"{where}: 'problem' is required")
    kind = d["problem"]
    if any(k in d for k in _TIME_KEYS) and "tmax" in d:
        raise ParamsError(f"{where}: give either tmax or t_year..t_sec")
    if kind == "reference":
        for k in ("sigma_n", "b_seis"):
            if k in d:
                raise ParamsError(f"{where}: {k} is fixed by the reference problem")
        ni, nk, dh = d.get("ni", 801), d.get("nk", 801), float(d.get("dh", 30.0))
        u0 = None
        if "u0_file" in d:
            u0 = read_reference_u0(os.path.join(base_dir, d["u0_file"]), ni, nk)
        pb = make_reference_problem(ni, nk, dh, U0=u0)
    else:
        n = d.get("ni", d.get("nk", 129))
        if d.get("nk", n) != n:
            raise ParamsError(f"{where}: the {kind} problem is square (ni = nk)")
        if "dh" in d and abs(float(d["dh"]) - 48000.0 / (n - 1)) > 1e-9:
            raise ParamsError(f"{where}: dh is fixed at 48 km / (n - 1) = "
                              f"{48000.0 / (n - 1)} m for the {kind} problem")
        if "u0_file" in d:
            raise ParamsError(f"{where}: u0_file applies to the reference problem")
        pb = make_bp1_problem(n, levels=d.get("n_levels"), basin=kind == "basin",
                              Dc=float(d.get("Dc", 0.008)),
                              sigma_n=float(d.get("sigma_n", 50.0)),
                              b_seis=None if "b_seis" not in d else float(d["b_seis"]))
    p = pb.params
    upd = {}
    if any(k in d for k in _TIME_KEYS):
        t = {k: d.get(k, 0) for k in _TIME_KEYS}
        upd["tmax"] = t["t_sec"] + 60.0 * (t["t_min"] + 60.0 * (t["t_hour"] + 24.0 * (
            t["t_day"] + 30.0 * (t["t_month"] + 12.0 * t["t_year"]))))
    names = {"tmax": "tmax", "ini_dt": "ini_dt", "mindt": "mindt", "maxdt": "maxdt",
             "totTol": "totTol", "maxNumSteps": "max_steps",
             "it_save_jump": "it_save_jump", "f0": "f0", "V0": "V0", "Dc": "L",
             "Vini": "Vini", "n_levels": "n_levels", "n_cycles": "n_cycles",
             "relax": "relax", "tol": "tol", "max_vcycles": "max_vcycles",
             "n_pre": "n_pre", "n_post": "n_post", "n_coarse": "n_coarse"}
    for k, field in names.items():
        if k in d:
            upd[field] = type(getattr(p, field))(d[k])
    if "tol_mode" in d:
        upd["tol_mode"] = TOL_ABS if d["tol_mode"] == "abs" else TOL_REL
    if "cycle" in d:
        upd["cycle"] = CYCLE_REFERENCE if d["cycle"] == "reference" else CYCLE_FAST
    if "precision" in d:
        upd["precision"] = PREC_FP64 if d["precision"] == "fp64" else PREC_FP32_IR
    params = dataclasses.replace(p, **upd)
    if params.ini_dt <= 0 or params.mindt <= 0 or params.maxdt < params.mindt:
        raise ParamsError(f"{where}: need 0 < mindt <= maxdt and ini_dt > 0")
    if params.tol <= 0 or params.totTol <= 0 or params.max_vcycles < 1:
        raise ParamsError(f"{where}: tol, totTol and max_vcycles must be positive")
    n = min(params.ni, params.nk)
    for _ in range(params.n_levels - 1):
        if n < 3:
            raise ParamsError(f"{where}: too many levels ({params.n_levels}) for the grid")
        n = (n - 1) // 2 + 1
    if n < 3:
        raise ParamsError(f"{where}: too many levels ({params.n_levels}) for the grid")
    name = d.get("name") or pb.name
    return dataclasses.replace(pb, params=params, name=name)


def load_params_file(path: str) -> Problem:
    """Reads a JSON parameter file (keys above) into a Problem."""
    with open(path) as fh:
        try:
            d = json.load(fh)
        except json.JSONDecodeError as e:
            raise ParamsError(f"{path}: not valid JSON ({e})") from None
    if not isinstance(d, dict):
        raise ParamsError(f"{path}: expected a JSON object")
    return problem_from_params(d, os.path.dirname(os.path.abspath(path)), path)


PARAMS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "params")


def workload(name: str) -> Problem:
    """A named workload from the package's parameter files (params/*.json):
    bp1_129, basin_1025, basin_2049, ref_basin_801."""
    path = os.path.join(PARAMS_DIR, name + ".json")
    if not os.path.exists(path):
        have = sorted(f[:-5] for f in os.listdir(PARAMS_DIR) if f.endswith(".json"))
        raise KeyError(f"unknown workload {name!r}; have {have}")
    return load_params_file(path)


WORKLOADS = {n: (lambda n=n: workload(n)) for n in
             ("bp1_129", "basin_1025", "basin_2049", "ref_basin_801")}
