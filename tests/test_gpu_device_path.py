"""GPU checks of the drop-in entry points with device arrays (QDEQ_DEVICE).

* A C program calls setBoundary / ode_solve / eval_stress / NRsearch /
  cal_dpsi on device buffers the caller filled with asynchronous copies on
  the null stream right before each call (tests/c/device_path.c), and gets
  the host path's results bit for bit.
* The Python binding's DeviceArray path equals its numpy path; arrays of
  mixed pointer spaces and arrays of the wrong size are refused.
* A level-0 tiling with more tiles than the norm-partial buffer holds is
  refused (qdeq_set_tiles).
"""
import subprocess

import numpy as np
import pytest

from qdeqcycle_fdm_amd import build, config, lib

pytestmark = pytest.mark.gpu


def test_c_program_device_path_equals_host_path():
    print("running", build.DEVICE_PATH_BIN, flush=True)
    out = subprocess.run([build.DEVICE_PATH_BIN, "0"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "device path OK" in out.stdout


def test_python_device_arrays_equal_host_arrays():
    pb = config.make_bp1_problem(129, basin=True, tol_mode=config.TOL_REL, tol=1e-10)
    p = pb.params
    rng = np.random.default_rng(2)
    U = pb.U0.copy()
    U[1:-1, 2:-2] += rng.normal(scale=0.1, size=(p.nk, p.ni - 2))
    slip = np.r_[0.0, rng.uniform(0.0, 2.0, p.nk), 0.0]
    ss = pb.U0[:, p.ni].copy()
    with lib.Engine(pb, cycle=config.CYCLE_FAST, n_pre=1, n_post=3, n_coarse=0,
                    relax=1.15) as eng:
        Uh = U.copy()
        eng.setBoundary(Uh, slip, ss, 1e8, 0.5, 1e6)
        out_h = eng.ode_solve(Uh)
        dU = lib.DeviceArray.from_host(U)
        eng.setBoundary(dU, lib.DeviceArray.from_host(slip),
                        lib.DeviceArray.from_host(ss), 1e8, 0.5, 1e6)
        out_d = eng.ode_solve(dU)
        assert out_h == out_d
        np.testing.assert_array_equal(dU.to_host(), Uh)
        with pytest.raises(TypeError):
            eng.setBoundary(dU, slip, ss, 1e8, 0.5, 1e6)       # mixed spaces
        with pytest.raises(ValueError):
            eng.ode_solve(lib.DeviceArray((p.nz, p.nx - 1)))   # wrong size
        with pytest.raises(ValueError):
            eng.ode_solve(U[:, :-1].copy())
        dU.free()


def test_small_level0_tiles_are_refused_at_1025():
    """8 x 2 tiles would make 66177 level-0 tiles, more norm partials than
    the context holds (16384): refused before any launch uses them."""
    pb = config.make_bp1_problem(1025, basin=True, tol_mode=config.TOL_REL, tol=1e-10)
    with lib.Engine(pb, cycle=config.CYCLE_FAST, n_pre=1, n_post=3, n_coarse=0,
                    relax=1.15) as eng:
        for tile in ((8, 2), (16, 4)):
            with pytest.raises(lib.QdeqError) as ei:
                eng.set_tiles(True, *tile)
            assert ei.value.status == lib.QDEQ_ERR_ARG
            assert "too small" in str(ei.value)
        eng.set_tiles(True, 64, 16)   # 17 x 65 tiles: accepted
