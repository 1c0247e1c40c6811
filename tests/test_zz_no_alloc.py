"""The allocation-free alignment entry point against ndt_hip_align_device,
through the C ABI: two calls in a row (the second with a smaller scan), each
bit-identical to the allocating entry point's result with the same
workgroup count; sizes beyond the documented limit are refused."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _print_case(msg):
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def test_align_device_noalloc_matches_align_device():
    import torch

    from cuda_ndt_matcher_amd import _lib, synth

    lib = _lib.load()
    w = synth.make_workload("small", 0)
    cfg = _lib.NdtConfig()
    lib.ndt_hip_config_default(C.byref(cfg))
    h = C.c_void_p()
    assert lib.ndt_hip_create(C.byref(cfg), C.byref(h)) == 0
    try:
        mp = np.ascontiguousarray(w.map_points, np.float32)
        assert lib.ndt_hip_set_target(h, mp, mp.shape[0]) == 0
        rng = np.random.default_rng(5)
        sub = np.ascontiguousarray(w.scan[np.sort(rng.choice(w.scan.shape[0], w.scan.shape[0] // 2, replace=False))])
        d_out = torch.zeros(C.sizeof(_lib.NdtResult), dtype=torch.uint8, device="cuda:0")
        for k, (scan, pose) in enumerate(((w.scan, w.guess), (sub, w.guess + np.array([0.1, -0.1, 0, 0, 0, 0])))):
            _print_case(f"noalloc call {k}: {scan.shape[0]} points")
            d_scan = torch.from_numpy(np.ascontiguousarray(scan, np.float32)).to("cuda:0")
            torch.cuda.synchronize()
            pv = np.ascontiguousarray(pose, np.float64)
            ref = _lib.NdtResult()
            assert lib.ndt_hip_align_device(h, C.c_void_p(d_scan.data_ptr()), scan.shape[0], pv, C.byref(ref)) == 0
            d_out.zero_()
            torch.cuda.synchronize()
            rc = lib.ndt_hip_align_device_noalloc(h, C.c_void_p(d_scan.data_ptr()), scan.shape[0], pv,
                                                  C.c_void_p(d_out.data_ptr()))
            assert rc == 0, lib.ndt_hip_last_error(h)
            got = _lib.NdtResult.from_buffer_copy(d_out.cpu().numpy().tobytes())
            assert bytes(got) == bytes(ref), f"call {k}: results differ"
            assert got.iterations > 0 and got.status in (0, 1)
        # n at or beyond 2^31 points is refused before any launch
        rc = lib.ndt_hip_align_device_noalloc(h, C.c_void_p(d_scan.data_ptr()), 1 << 31, pv,
                                              C.c_void_p(d_out.data_ptr()))
        assert rc == -7
        # a null output is an argument error
        assert lib.ndt_hip_align_device_noalloc(h, C.c_void_p(d_scan.data_ptr()), scan.shape[0], pv, None) == -1
    finally:
        lib.ndt_hip_destroy(h)
