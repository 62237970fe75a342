"""The C ABI library builds, loads, and exports every entry point that
include/qdeq.h declares (no compute calls: no GPU needed)."""
import os
import re
import subprocess

import pytest

from qdeqcycle_fdm_amd import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


HEADERS = ("qdeq.h", "qdeq_guess.h")


def _declared():
    names = set()
    for h in HEADERS:
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(qdeq_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declarations_match_binding_list():
    assert _declared() == sorted(lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("libqdeq.so not built")
    L = lib.load()
    for name in _declared():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", lib.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(qdeq_\w+)", out))
    assert set(_declared()) <= exported


def test_host_only_entry_points():
    """computeError / computeStepSize are host functions of the library."""
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("libqdeq.so not built")
    import numpy as np
    s3 = np.array([0, 1.0, 2.0, 0]); s4 = np.array([0, 1.5, 2.0, 0])
    p3 = np.array([0, 0.1, 0.2, 0]); p4 = np.array([0, 0.1, 0.0, 0])
    e = lib.Engine.computeError(s3, p3, s4, p4)
    assert abs(e - (0.5 + 0.2) / 2 ** 0.5) < 1e-15
    dt = lib.Engine.computeStepSize(1, 10.0, 1e-5, 1e8, 1e-7, 1e-7, [0.0, 0.0])
    assert abs(dt - 9.0) < 1e-12


def test_header_compiles_as_c():
    """include/*.h are plain C (no HIP or C++ types in the signatures)."""
    src = ('#include "qdeq.h"\n#include "qdeq_guess.h"\n'
           'int main(void){qdeq_config c; qdeq_default_config(&c);'
           ' qdeq_guess_config g; qdeq_guess_default_config(&g, 1025); return 0;}\n')
    p = subprocess.run(["gcc", "-std=c99", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        "-x", "c", "-"], input=src, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr


def test_hip_runtime_version_check():
    """The library reports the HIP version it was built against and the one
    the process runs; lib.load() refuses a different major.minor."""
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("libqdeq.so not built")
    import ctypes
    L = lib.load()
    built, running = ctypes.c_int(0), ctypes.c_int(0)
    assert L.qdeq_hip_runtime_version(ctypes.byref(built), ctypes.byref(running)) == 0
    hdr = open("/opt/rocm/include/hip/hip_version.h").read()
    v = {k: int(re.search(rf"#define HIP_VERSION_{k} (\d+)", hdr).group(1))
         for k in ("MAJOR", "MINOR", "PATCH")}
    assert built.value == v["MAJOR"] * 10000000 + v["MINOR"] * 100000 + v["PATCH"]
    assert running.value // 100000 == built.value // 100000
    lib.check_runtime(70226015, 70226999)           # same 7.2: accepted
    with pytest.raises(RuntimeError, match=r"built against HIP 7\.2 .* runs HIP 7\.0"):
        lib.check_runtime(70226015, 70051831)


def test_load_after_another_hip_runtime_is_refused():
    """torch's wheel bundles its own libamdhip64 under the same soname; a
    process that imported torch first runs that runtime, and loading the
    library there raises instead of running on it."""
    if not os.path.exists(lib.LIB_PATH):
        pytest.skip("libqdeq.so not built")
    code = ("import torch\n"
            "print('TORCH_HIP', torch.version.hip)\n"
            "from qdeqcycle_fdm_amd import lib\n"
            "try:\n"
            "    lib.load()\n"
            "    print('LOADED')\n"
            "except RuntimeError as e:\n"
            "    print('REFUSED', e)\n")
    out = subprocess.run([os.environ.get("PYTHON", "python3"), "-c", code], cwd=ROOT,
                         capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr
    tv = re.search(r"TORCH_HIP (\S+)", out.stdout).group(1)
    if tv.startswith("7.2"):
        assert "LOADED" in out.stdout
    else:
        assert "REFUSED" in out.stdout and "load qdeqcycle_fdm_amd.lib before" in out.stdout


def test_device_path_program_builds():
    """The C caller of the device path (run by tests/test_gpu_device_path.py)
    is built with the library."""
    from qdeqcycle_fdm_amd import build
    assert os.path.exists(build.DEVICE_PATH_BIN) or not os.path.exists(lib.LIB_PATH)
