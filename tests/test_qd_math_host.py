"""The engine's elementary functions (qdeqcycle_fdm_amd/csrc/qd_math.h, the
code the fault kernels run) built for the host with g++ -ffp-contract=off
(tests/c/qd_math_check.cpp):

* qd_sinh_cosh, which the Newton iteration of NRsearch (fault.cpp:180-228)
  calls once per iteration, gives exactly the bits of qd_sinh and qd_cosh
  called separately, at every branch edge of the two functions, at the
  infinities and at NaN;
* every function of the header gives the bits of the oracle's copy
  (oracle/qdeq_oracle.c), which the GPU-vs-oracle parity tests rely on;
* a header with one planted mistake (built in a temporary directory) fails
  each of these comparisons.
"""
import math
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c", "qd_math_check.cpp")
NAMES = ("exp", "sinh", "cosh", "log", "log1p", "asinh")


def _bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def _val(b: int) -> float:
    return struct.unpack("<d", struct.pack("<Q", b))[0]


CSRC = os.path.join(ROOT, "qdeqcycle_fdm_amd", "csrc")


def _build(out: str, include_dir: str) -> str:
    gxx = shutil.which("g++")
    if gxx is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("g++ or the ROCm headers are missing")
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include", "-I", include_dir, SRC, "-o", out],
                   check=True, capture_output=True)
    return out


@pytest.fixture(scope="module")
def check_bin(tmp_path_factory):
    return _build(str(tmp_path_factory.mktemp("qdm") / "qd_math_check"), CSRC)


def _inputs():
    edges = [0.5, 22.0, 709.0, 709.782712893383973096, -745.133219101941108420, 1.0]
    xs = [0.0, -0.0, math.inf, -math.inf, math.nan, 1e-300, 5e-324, 1e300]
    for e in edges:
        for v in (e, math.nextafter(e, 0.0), math.nextafter(e, math.inf), 0.5 * e, 2.0 * e):
            xs += [v, -v]
    rng = np.random.default_rng(7)
    xs += list(rng.uniform(-50.0, 50.0, 2000))          # Newton variable of NRsearch
    xs += list(rng.uniform(-1.0, 1.0, 500))
    mag = 10.0 ** rng.uniform(-300, 300, 1000)
    xs += list(mag * rng.choice([-1.0, 1.0], 1000))
    return [float(x) for x in xs]


def _run(check_bin, xs):
    text = "\n".join(format(_bits(x), "016x") for x in xs) + "\n"
    res = subprocess.run([check_bin], input=text, capture_output=True, text=True, check=True)
    rows = [[int(w, 16) for w in line.split()] for line in res.stdout.splitlines()]
    assert len(rows) == len(xs)
    return rows


def _same(a: int, b: int) -> bool:
    fa, fb = _val(a), _val(b)
    return a == b or (math.isnan(fa) and math.isnan(fb))


def test_sinh_cosh_together_equal_separate_calls(check_bin):
    xs = _inputs()
    for x, row in zip(xs, _run(check_bin, xs)):
        assert _same(row[6], row[1]), ("sinh", x)
        assert _same(row[7], row[2]), ("cosh", x)


def test_engine_header_equals_oracle_copies(check_bin):
    xs = _inputs()
    for x, row in zip(xs, _run(check_bin, xs)):
        for i, name in enumerate(NAMES):
            assert _same(row[i], _bits(oracle.elem(name, x))), (name, x)


@pytest.mark.parametrize("old,new,col,ref", [
    # sinh of the shared branch given cosh's sum: caught against qd_sinh
    ("      r = 0.5 * (e - ie);", "      r = 0.5 * (e + ie);", 6, 1),
    # one Taylor coefficient of exp changed: caught against the oracle's exp
    ("  p = p * r + 1.0 / 720.0;", "  p = p * r + 1.0 / 721.0;", 0, None),
])
def test_planted_mistakes_are_caught(tmp_path, check_bin, old, new, col, ref):
    """Each comparison above fails for a header with one planted mistake."""
    text = open(os.path.join(CSRC, "qd_math.h")).read()
    assert text.count(old) == 1
    d = tmp_path / "inc"
    d.mkdir()
    (d / "qd_math.h").write_text(text.replace(old, new))
    mutant = _build(str(tmp_path / "mutant"), str(d))
    xs = _inputs()
    rows = _run(mutant, xs)
    if ref is not None:
        bad = [x for x, r in zip(xs, rows) if not _same(r[col], r[ref])]
    else:
        bad = [x for x, r in zip(xs, rows) if not _same(r[col], _bits(oracle.elem("exp", x)))]
    assert bad, "the planted mistake went unnoticed"
