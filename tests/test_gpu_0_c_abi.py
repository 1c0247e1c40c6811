"""The library from plain C (GPU): a gcc-built program that includes only
include/ndt_hip.h and links -lndt_hip aligns a scan of a known pose, checks
batch against single alignments, NVTL evaluation and an error path.

The program is a second process on the GPU.  Its alignments use cooperative
launches, which the runtime serves to one process at a time: once this test
process has made one (any other GPU test does), the child's are refused until
this process exits.  The file name makes this test the first one collected."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cuda_ndt_matcher_amd", "c_abi_example")


def test_c_program_through_the_abi():
    print("[case] c_abi_example", file=sys.stderr, flush=True)
    assert os.path.exists(EXE), f"{EXE} is not built (make -C cuda_ndt_matcher_amd/csrc)"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout, file=sys.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASSED" in r.stdout
