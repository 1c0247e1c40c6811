"""The C-ABI library loads and exports every symbol include/ndt_hip.h declares.

CPU-only: no compute call reaches a GPU here.  The Laplace helpers are pure
host code and are checked against closed forms.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "ndt_hip.h")).read()
    return sorted(set(re.findall(r"\b(ndt_hip_[a-z0-9_]+)\s*\(", text)))


def test_header_lists_match_binding():
    from cuda_ndt_matcher_amd import _lib

    assert sorted(_lib.EXPORTS) == declared_symbols()


def test_library_exports_every_declared_symbol():
    from cuda_ndt_matcher_amd import _lib

    lib = _lib.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_struct_layouts(tmp_path):
    """The ctypes mirrors have the header's sizes and field offsets (the
    header compiled by gcc prints them)."""
    import subprocess

    from cuda_ndt_matcher_amd import _lib

    assert C.sizeof(_lib.NdtResult) == 6 * 8 + 8 + 3 * 8 + 16 + 36 * 8
    mirrors = {"ndt_config": _lib.NdtConfig, "ndt_result": _lib.NdtResult, "ndt_timing": _lib.NdtTiming,
               "ndt_phase_profile": _lib.NdtPhaseProfile, "ndt_tuning_entry": _lib.NdtTuningEntry,
               "ndt_filter_params": _lib.NdtFilterParams, "ndt_filter_counts": _lib.NdtFilterCounts}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ndt_hip.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            a, b, v = ln.split()
            got[(a, b)] = int(v)
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_config_defaults_match_reference():
    # NdtScanMatcherConfig::default (src/ndt_cuda/src/ndt.rs:90-107)
    from cuda_ndt_matcher_amd import _lib

    c = _lib.NdtConfig()
    _lib.load().ndt_hip_config_default(C.byref(c))
    assert c.resolution == 2.0 and c.max_iterations == 30 and c.trans_epsilon == 0.01
    assert c.step_size == 0.1 and c.outlier_ratio == 0.55 and c.regularization == 0.001
    assert c.use_line_search == 0 and c.min_points_per_voxel == 6 and c.regularization_enabled == 0
    assert c.regularization_scale_factor == 0.01 and c.search_mode == 0 and c.precision == 0


def test_laplace_helpers_closed_form(oracle_mod):
    from cuda_ndt_matcher_amd import laplace_covariance_6x6, laplace_covariance_xy

    rng = np.random.default_rng(0)
    A = rng.normal(size=(6, 6))
    H = -(A @ A.T + np.eye(6))
    np.testing.assert_allclose(laplace_covariance_xy(H), -np.linalg.inv(H[:2, :2]), rtol=1e-12)
    np.testing.assert_allclose(laplace_covariance_xy(H), oracle_mod.laplace_xy(H), rtol=1e-14)
    np.testing.assert_allclose(laplace_covariance_6x6(H), -np.linalg.inv(H), rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(laplace_covariance_xy(np.zeros((6, 6))), np.eye(2))
    np.testing.assert_allclose(laplace_covariance_6x6(np.zeros((6, 6))), np.eye(6))


def test_isometry_conversions_match_oracle(oracle_mod):
    from cuda_ndt_matcher_amd import Isometry3

    rng = np.random.default_rng(1)
    for _ in range(50):
        pose = np.concatenate([rng.normal(0, 10, 3), rng.uniform(-1.4, 1.4, 3)])
        iso = Isometry3.from_pose_vector(pose)
        np.testing.assert_allclose(np.c_[iso.R, iso.t], oracle_mod.pose_to_matrix(pose), atol=1e-15)
        np.testing.assert_allclose(iso.to_pose_vector(), oracle_mod.matrix_to_pose(np.c_[iso.R, iso.t]),
                                   atol=1e-15)


def test_invalid_config_rejected_without_device():
    from cuda_ndt_matcher_amd import _lib

    lib = _lib.load()
    c = _lib.NdtConfig()
    lib.ndt_hip_config_default(C.byref(c))
    c.search_mode = 9
    h = C.c_void_p()
    assert lib.ndt_hip_create(C.byref(c), C.byref(h)) == -1
    assert not h.value


def test_product_does_not_import_oracle():
    """The product path never loads the oracle or a CPU fallback."""
    pkg = os.path.join(ROOT, "cuda_ndt_matcher_amd")
    pattern = re.compile(r"(import\s+oracle|from\s+oracle|libndt_oracle|ndt_oracle\.h)")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                assert not pattern.search(open(os.path.join(dirpath, f)).read()), f


def test_result_conversions_read_every_field():
    """The one-result path (field by field) and the batch path (one
    structured numpy view) turn ndt_result structs into identical
    AlignResults, field for field, with every field of the struct set to a
    distinct value."""
    from cuda_ndt_matcher_amd import _lib, matcher

    rng = np.random.default_rng(3)
    arr = (_lib.NdtResult * 9)()
    for i in range(9):
        r = arr[i]
        for k in range(6):
            r.pose[k] = rng.normal()
        for k in range(36):
            r.hessian[k] = rng.normal()
        r.status, r.converged = i % 4, i % 2
        r.score, r.transform_probability, r.nvtl = rng.normal(size=3)
        r.iterations, r.num_correspondences = 3 + i, 1000 + i
        r.oscillation_count, r.pair_evaluations = i, (1 << 31) + i
    batch = matcher._results(arr, 7)
    assert len(batch) == 7
    for i, b in enumerate(batch):
        a = matcher._result(arr[i])
        r = arr[i]
        assert np.array_equal(a.pose_vector, list(r.pose)) and np.array_equal(b.pose_vector, a.pose_vector)
        assert np.array_equal(a.hessian, np.reshape(list(r.hessian), (6, 6)))
        assert np.array_equal(b.hessian, a.hessian)
        for f in ("converged", "status", "score", "transform_probability", "nvtl", "iterations",
                  "num_correspondences", "oscillation_count", "pair_evaluations"):
            assert getattr(a, f) == getattr(b, f) and type(getattr(a, f)) is type(getattr(b, f)), f
        assert a.status == matcher.STATUS_NAMES[i % 4] and a.converged == bool(i % 2)
        assert a.score == r.score and a.iterations == 3 + i and a.pair_evaluations == (1 << 31) + i
