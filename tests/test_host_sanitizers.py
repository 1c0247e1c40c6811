"""The oracle's C++ under AddressSanitizer and UndefinedBehaviorSanitizer
(CPU).  `make -C oracle sanitize` builds the same source with both sanitizers
(every finding fatal); the oracle's CPU tests then run in a child process
with the ASan runtime preloaded and NDT_ORACLE_LIB pointing at that build.
A planted overflow (an export buffer one voxel short, written by the
oracle) shows that the sanitizer is live in that setup.

The library's own host code runs under the same sanitizers in the GPU tests
(tests/test_gpu_0_c_abi.py), since it needs a device."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_LIB = os.path.join(ROOT, "oracle", "_san", "libndt_oracle.so")


def _asan_runtime():
    gxx = shutil.which("g++")
    if not gxx:
        return None
    path = subprocess.run([gxx, "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture(scope="module")
def san_env():
    rt = _asan_runtime()
    if rt is None:
        pytest.skip("g++ with the AddressSanitizer runtime is not available")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    env = dict(os.environ)
    env["LD_PRELOAD"] = rt
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:exitcode=23"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1:exitcode=24"
    env["NDT_ORACLE_LIB"] = SAN_LIB
    env["PYTHONPATH"] = ROOT + os.pathsep + env.get("PYTHONPATH", "")
    return env


def test_oracle_tests_under_sanitizers(san_env):
    tests = [os.path.join(ROOT, "tests", f) for f in ("test_oracle.py", "test_parity_helper.py",
                                                      "test_ingest_host.py")]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                        *tests], capture_output=True, text=True, env=san_env, cwd=ROOT, timeout=600)
    tail = (r.stdout[-3000:] + r.stderr[-3000:])
    assert r.returncode == 0, tail
    assert "passed" in r.stdout and "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, tail


PLANTED = r"""
import numpy as np
from oracle import oracle as O
lib = O.lib()
assert lib._name.endswith("_san/libndt_oracle.so"), lib._name
rng = np.random.default_rng(0)
pts = rng.uniform(0, 20, (20000, 3)).astype(np.float32)
g = lib.oracle_grid_build(pts, pts.shape[0], 2.0, 6, 0.01)
v = lib.oracle_grid_size(g)
short = np.zeros(3 * (v - 1), np.int32)  # one voxel short: the export overruns it
lib.oracle_grid_export(g, short, np.zeros(3 * v, np.float32), np.zeros(9 * v, np.float32),
                       np.zeros(9 * v, np.float32), np.zeros(3 * v, np.float32), np.zeros(v, np.uint32),
                       np.zeros(3 * v), np.zeros(9 * v))
print("not caught")
"""


def test_planted_overflow_is_caught(san_env):
    r = subprocess.run([sys.executable, "-c", PLANTED], capture_output=True, text=True, env=san_env, cwd=ROOT,
                       timeout=300)
    assert r.returncode == 23, r.stdout + r.stderr[-3000:]
    assert "heap-buffer-overflow" in r.stderr and "oracle_grid_export" in r.stderr
    assert "not caught" not in r.stdout
