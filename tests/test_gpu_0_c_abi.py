"""The library from plain C (GPU).

* c_abi_example: a gcc-built program that includes only include/ndt_hip.h and
  links -lndt_hip aligns a scan of a known pose, checks batch against single
  alignments, NVTL evaluation and an error path.
* abi_host_check: drives every host-array entry point and its error paths
  (map rebuilds, lookups, derivatives, single/batch/uneven alignments,
  scoring, filtering, regularisation, the server, graph replay, autotuner,
  phase profile, the MLP's parameter copies); once against the ordinary
  library, once against san/libndt_hip.so, whose host code is built under
  AddressSanitizer and UndefinedBehaviorSanitizer, with the program linked to
  the sanitizer runtimes (every finding fatal; leak detection off, since the
  HIP runtime keeps its own allocations until exit).

The programs are further processes on the GPU.  Their alignments use
cooperative launches, which the runtime serves to one process at a time: once
this test process has made one (any other GPU test does), a child's are
refused until this process exits.  The file name makes these tests the first
ones collected."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cuda_ndt_matcher_amd", "c_abi_example")
HOST_CHECK = os.path.join(ROOT, "cuda_ndt_matcher_amd", "abi_host_check")
SAN_CHECK = os.path.join(ROOT, "cuda_ndt_matcher_amd", "san", "abi_host_check")


def _run(exe, env=None):
    assert os.path.exists(exe), f"{exe} is not built (make -C cuda_ndt_matcher_amd/csrc)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=180, env=env)
    print(r.stdout, file=sys.stderr)
    print(r.stderr[-4000:], file=sys.stderr)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "PASSED" in r.stdout
    return r


def test_c_program_through_the_abi():
    print("[case] c_abi_example", file=sys.stderr, flush=True)
    _run(EXE)


def test_every_host_entry_point():
    print("[case] abi_host_check", file=sys.stderr, flush=True)
    _run(HOST_CHECK)


def test_host_code_under_sanitizers():
    print("[case] san/abi_host_check (AddressSanitizer + UBSan host code)", file=sys.stderr, flush=True)
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:abort_on_error=0:exitcode=23"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1:exitcode=24"
    r = _run(SAN_CHECK, env)
    assert "built with AddressSanitizer" in r.stdout
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
