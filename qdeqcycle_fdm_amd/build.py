"""Builds libqdeq.so (HIP, gfx950) in-tree with hipcc.

Every HIP source is compiled with -ffp-contract=off (so FP64 expressions are
evaluated in the order written, as the reference writes them) and with
packed FP32 instructions disabled.  The library lands next to this file.

Also builds _build/device_path, a C program of the test suite that calls the
drop-in entry points with host and with device arrays (tests/c).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD_DIR = os.path.join(ROOT, "_build")
LIB_PATH = os.path.join(HERE, "libqdeq.so")
SOURCES = ["mg.hip", "mg_tiled.hip", "coarse.hip", "fault.hip", "medium.hip", "api.hip",
           "nn.hip", "guess.hip", "output.hip", "proj.hip"]
ARCH = os.environ.get("QDEQ_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
DEVICE_PATH_SRC = os.path.join(ROOT, "tests", "c", "device_path.c")
DEVICE_PATH_BIN = os.path.join(BUILD_DIR, "device_path")
COMMON_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def build(verbose: bool = False) -> str:
    """Compiles every source (in parallel) and links libqdeq.so."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(BUILD_DIR, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC, *COMMON_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # the host pass reports the GPU-only target feature; that is expected
        lines = [ln for ln in text.splitlines()
                 if "packed-fp32-ops" not in ln and ln.strip()]
        if p.returncode != 0:
            failed.append((src, text))
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        for src, text in failed:
            sys.stderr.write(f"--- {src} ---\n{text}\n")
        raise RuntimeError("hipcc failed: " + ", ".join(s for s, _ in failed))
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    build_device_path()
    return LIB_PATH


def build_device_path() -> str:
    """The C caller of tests/c/device_path.c, linked against libqdeq.so (found
    through an $ORIGIN-relative rpath) and the HIP runtime."""
    cmd = ["gcc", "-std=c99", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROCM, "include"),
           DEVICE_PATH_SRC, "-o", DEVICE_PATH_BIN, "-L", HERE, "-lqdeq",
           "-L", os.path.join(ROCM, "lib"), "-lamdhip64", "-lm",
           "-Wl,-rpath,$ORIGIN/../qdeqcycle_fdm_amd:" + os.path.join(ROCM, "lib")]
    subprocess.run(cmd, check=True)
    return DEVICE_PATH_BIN


if __name__ == "__main__":
    print(build(verbose=True))
