"""The library's host-only code (SCF driver, dense linear algebra, host Becke grid) built with
g++ under AddressSanitizer and UndefinedBehaviorSanitizer, run on the CPU: no error report, and
the results are right (eigen/LU residuals, SCF convergence and idempotency, grid point counts
equal to the oracle's build_becke_grid)."""
import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "iqcp_amd" / "csrc"


@pytest.fixture(scope="module")
def sanitized(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("asan") / "sanitize_driver"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           str(ROOT / "tests" / "host" / "sanitize_driver.cpp"), str(SRC / "scf.cpp"), str(SRC / "grid.cpp"),
           "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("sanitizer build failed:\n" + r.stderr[-4000:])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([str(out), "6"], capture_output=True, text=True, env=env, timeout=600)
    assert run.returncode == 0, run.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in run.stderr and "runtime error" not in run.stderr, run.stderr[-4000:]
    vals = {}
    for line in run.stdout.splitlines():
        k, v = line.split()
        vals[k] = float(v)
    return vals


def test_linear_algebra_under_sanitizers(sanitized):
    for n in (1, 2, 7, 40, 97):
        assert sanitized[f"eigen_residual_{n}"] < 1e-12 * max(1, n)
        assert sanitized[f"lu_ok_{n}"] == 1
        assert sanitized[f"lu_residual_{n}"] < 1e-11


def test_scf_driver_under_sanitizers(sanitized):
    for v in (0, 1):
        assert sanitized[f"scf_rc_{v}"] == 0
        assert sanitized[f"scf_converged_{v}"] == 1
        assert sanitized[f"scf_idempotency_{v}"] < 1e-8
        assert abs(sanitized[f"scf_electrons_{v}"] - 10.0) < 1e-9
    # DIIS and damping + level shift reach the same stationary state
    assert abs(sanitized["scf_energy_0"] - sanitized["scf_energy_1"]) < 1e-8


def test_host_grid_under_sanitizers(sanitized):
    pos = np.array([[0.0, 0.0, 0.0], [1.43, 1.1, 0.0], [-1.43, 1.1, 0.0],
                    [5.5, 0.3, 0.2], [6.9, 1.4, 0.1], [4.1, 1.4, -0.3]])
    Z = np.array([8, 1, 1, 8, 1, 1], dtype=np.int32)
    for pruning in (0, 1):
        assert sanitized[f"grid_rc_{pruning}"] == 0
        _, w, _ = oracle.becke_grid(Z, pos, 75, 0, pruning)
        assert int(sanitized[f"grid_points_{pruning}"]) == len(w)
        assert abs(sanitized[f"grid_wsum_{pruning}"] - w.sum()) < 1e-6 * w.sum()
