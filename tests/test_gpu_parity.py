"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

Bar: voxel keys, sort order, point-to-voxel assignment and counts bit-exact;
per-voxel means and covariances within 1e-5 relative; neighbour sets equal;
alignment iterations and status exact, pose within 1 mm / 0.01 deg, NVTL and
TP within 1e-4.
"""
import math

import numpy as np
import pytest

from tests import npref
from oracle import parity

pytestmark = pytest.mark.gpu

POSE_MM = 1e-3
POSE_RAD = np.deg2rad(0.01)
SCORE_TOL = 1e-4


def _print_case(name):
    print(f"[gpu-case] {name}", flush=True)


@pytest.fixture(scope="module")
def small():
    from cuda_ndt_matcher_amd import synth

    return synth.make_workload("small", 0)


@pytest.fixture(scope="module")
def oracle_grid(small, oracle_mod):
    return oracle_mod.Grid(small.map_points, 2.0)


def _matcher(points=None, **cfg):
    from cuda_ndt_matcher_amd import NdtScanMatcher, NdtScanMatcherConfig

    m = NdtScanMatcher(NdtScanMatcherConfig(**cfg))
    if points is not None:
        m.set_target(points)
    return m


@pytest.fixture(scope="module")
def gpu_small(small):
    _print_case("set_target small")
    return _matcher(small.map_points)


def morton_np(keys):
    """Compact Morton code of integer keys relative to their bounding box."""
    kmin = keys.min(0)
    r = (keys - kmin).astype(np.uint64)
    span = (keys.max(0) - kmin)
    bits = [int(s).bit_length() for s in span]
    out = np.zeros(keys.shape[0], np.uint64)
    pos = 0
    for i in range(max(bits)):
        for a in range(3):
            if i < bits[a]:
                out |= ((r[:, a] >> np.uint64(i)) & np.uint64(1)) << np.uint64(pos)
                pos += 1
    return out


def _check_build(matcher, points, oracle_grid, oracle_mod):
    n = points.shape[0]
    dbg = matcher.build_debug(n)
    keys = oracle_mod.point_keys(points, 2.0)
    # voxel key of every point, bit-exact
    assert np.array_equal(dbg["point_keys"], keys)
    # stable sort order of the packed key, bit-exact
    mk = morton_np(keys.astype(np.int64))
    order = np.argsort(mk, kind="stable")
    assert np.array_equal(dbg["sorted_index"], order.astype(np.uint32))
    assert np.array_equal(dbg["sorted_keys"], mk[order])
    # voxels: same set of coordinates, same counts, statistics within 1e-5
    g = matcher.target_grid()
    ex = oracle_grid.export()
    gi = np.lexsort(g["coords"].T[::-1])
    oi = np.lexsort(ex["coords"].T[::-1])
    assert np.array_equal(g["coords"][gi], ex["coords"][oi])
    assert np.array_equal(g["counts"][gi], ex["counts"][oi])
    for name in ("means", "covs", "invs"):
        a = g[name][gi].astype(np.float64).reshape(len(gi), -1)
        b = ex[name][oi].astype(np.float64).reshape(len(oi), -1)
        scale = np.maximum(np.abs(b).max(1, keepdims=True), 1e-30)
        err = np.abs(a - b) / scale
        assert err.max() <= 1e-5, f"{name}: max rel err {err.max():.3e}"
    # principal axis up to sign
    dot = np.abs(np.sum(g["axes"][gi].astype(np.float64) * ex["axes"][oi], 1))
    assert np.percentile(dot, 99) > 1 - 1e-4
    return gi, oi


def test_build_parity_small(gpu_small, small, oracle_grid, oracle_mod):
    _print_case("build parity small")
    _check_build(gpu_small, small.map_points, oracle_grid, oracle_mod)


@pytest.mark.parametrize("n", [6, 100, 4095, 4096, 4097, 12289])
def test_build_parity_tile_edges(oracle_mod, n):
    """Maps whose point counts sit on and around the 4096-point sort tiles (a
    single partial tile, exact tiles, one point past a tile): keys, order,
    counts and statistics as the oracle's."""
    _print_case(f"build parity n={n}")
    rng = np.random.default_rng(n)
    # six points inside one voxel; about 25 per voxel at n = 100; then a 6 x 5 x 2 voxel box
    extent = [1.5, 1.5, 1.5] if n < 50 else ([4.0, 4.0, 2.0] if n <= 1000 else [12.0, 10.0, 4.0])
    pts = (rng.random((n, 3)) * np.array(extent) + [100.2, -51.8, 2.1]).astype(np.float32)
    m = _matcher(pts)
    _check_build(m, pts, oracle_mod.Grid(pts, 2.0), oracle_mod)
    m.close()


def test_build_parity_wide_keys(small, oracle_mod):
    """Two copies of the small map 4.1 km apart on every axis: 12 key bits per
    axis, 36 in all, so the build sorts 64-bit keys in five passes; the box is
    far past the cell tables, so the lookups take the hash.  Keys, order,
    counts and statistics as the oracle's; lookups around both copies as the
    oracle's search."""
    _print_case("build parity 64-bit keys")
    off = np.float32(4100.0)
    pts = np.concatenate([small.map_points, small.map_points + off]).astype(np.float32)
    span = np.ptp(np.floor(pts / np.float32(2.0)).astype(np.int64), axis=0)
    assert sum(int(s).bit_length() for s in span) > 32
    grid = oracle_mod.Grid(pts, 2.0)
    m = _matcher(pts)
    _check_build(m, pts, grid, oracle_mod)
    rng = np.random.default_rng(21)
    q = small.scan.astype(np.float32)[:400]
    q = np.concatenate([q, q + off, rng.uniform(-100, 100, (200, 3)).astype(np.float32) + off])
    g = m.target_grid()
    ex = grid.export()
    for mode in (0, 1):
        c, i = m.lookup(q, mode, max_k=40)
        co, io = grid.search(q, mode, max_k=40)
        assert np.array_equal(c, co), mode
        assert _coord_sets(g["coords"], c, i) == _coord_sets(ex["coords"], co, io), mode
    m.close()


def _coord_sets(coords, cnt, idx):
    out = []
    for i in range(len(cnt)):
        out.append(sorted(tuple(coords[j]) for j in idx[i, :cnt[i]]))
    return out


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_lookup_parity(gpu_small, small, oracle_grid, mode):
    _print_case(f"lookup mode {mode}")
    rng = np.random.default_rng(mode)
    R = npref.rot_xyz(*small.guess[3:])
    q = (small.scan.astype(np.float64) @ R.T + small.guess[:3]).astype(np.float32)
    q = np.concatenate([q, (q[:500] + rng.normal(0, 1.0, (500, 3))).astype(np.float32)])
    cnt_g, idx_g = gpu_small.lookup(q, mode, max_k=40)
    cnt_o, idx_o = oracle_grid.search(q, mode, max_k=40)
    assert np.array_equal(cnt_g, cnt_o)
    g = gpu_small.target_grid()
    ex = oracle_grid.export()
    assert _coord_sets(g["coords"], cnt_g, idx_g) == _coord_sets(ex["coords"], cnt_o, idx_o)


@pytest.mark.gpu
def test_cell_tables_match_hash_lookups(gpu_small, small):
    """The per-cell tables and the hash + bitmap path visit the same voxels in
    the same order: lookups agree exactly; derivatives and alignments agree to
    rounding (the two inlined paths may contract multiply-adds differently)."""
    _print_case("cell tables vs hash")
    hashed = _matcher(small.map_points, cell_tables=False)
    rng = np.random.default_rng(11)
    R = npref.rot_xyz(*small.guess[3:])
    q = (small.scan.astype(np.float64) @ R.T + small.guess[:3]).astype(np.float32)
    q = np.concatenate([q, (q[:500] + rng.normal(0, 1.0, (500, 3))).astype(np.float32),
                        rng.uniform(-500, 500, (200, 3)).astype(np.float32)])
    for mode in (0, 1, 2, 3):
        c1, i1 = gpu_small.lookup(q, mode, max_k=40)
        c2, i2 = hashed.lookup(q, mode, max_k=40)
        assert np.array_equal(c1, c2) and np.array_equal(i1, i2), mode
    s1, g1, H1, n1 = gpu_small.derivatives(small.scan, small.guess, True)
    s2, g2, H2, n2 = hashed.derivatives(small.scan, small.guess, True)
    assert n1 == n2 and s1 == pytest.approx(s2, rel=1e-12)
    np.testing.assert_allclose(g1, g2, rtol=0, atol=1e-12 * np.abs(g2).max())
    np.testing.assert_allclose(H1, H2, rtol=0, atol=1e-12 * np.abs(H2).max())
    r1 = gpu_small.align(small.scan, small.guess)
    r2 = hashed.align(small.scan, small.guess)
    assert r1.status == r2.status and r1.iterations == r2.iterations
    np.testing.assert_allclose(r1.pose_vector, r2.pose_vector, rtol=0, atol=1e-9)
    assert r1.nvtl == pytest.approx(r2.nvtl, rel=1e-12)


@pytest.mark.parametrize("far", [False, True])
def test_sparse_outliers_leave_the_grid_unchanged(gpu_small, small, far):
    """Isolated map points, each alone in its voxel (too few points to be valid),
    widen the points' voxel box.  Near ones stay inside a box small enough for the
    cell tables, which the build then uses without waiting for the valid voxels'
    box; far ones push it past the cell-table limit, so the build waits for the
    valid box, which still fits.  Either way the valid voxels, the lookups, the
    derivatives and the alignment equal those of the map without them."""
    _print_case(f"sparse outliers far={far}")
    off = np.array([5000.0, 5000.0, 200.0]) if far else np.array([30.0, 30.0, 5.0])
    extra = small.map_points.max(0) + off + np.array([[0.0, 0.0, 0.0], [7.0, 0.0, 0.0], [0.0, 9.0, 0.0]])
    m = _matcher(np.concatenate([small.map_points, extra.astype(np.float32)]))
    g0, g1 = gpu_small.target_grid(), m.target_grid()
    assert g1["coords"].shape == g0["coords"].shape
    o0, o1 = np.lexsort(g0["coords"].T[::-1]), np.lexsort(g1["coords"].T[::-1])
    for k in ("coords", "means", "covs", "invs", "axes", "counts"):
        assert np.array_equal(g0[k][o0], g1[k][o1]), k
    rng = np.random.default_rng(5)
    R = npref.rot_xyz(*small.guess[3:])
    q = (small.scan.astype(np.float64) @ R.T + small.guess[:3]).astype(np.float32)
    q = np.concatenate([q, rng.uniform(-150, 150, (300, 3)).astype(np.float32), extra.astype(np.float32)])
    for mode in (0, 1, 2, 3):
        c0, i0 = gpu_small.lookup(q, mode, max_k=40)
        c1, i1 = m.lookup(q, mode, max_k=40)
        assert np.array_equal(c0, c1), mode
        assert _coord_sets(g0["coords"], c0, i0) == _coord_sets(g1["coords"], c1, i1), mode
    s0, gr0, H0, n0 = gpu_small.derivatives(small.scan, small.guess, True)
    s1, gr1, H1, n1 = m.derivatives(small.scan, small.guess, True)
    assert n0 == n1 and s1 == pytest.approx(s0, rel=1e-12)
    np.testing.assert_allclose(gr1, gr0, rtol=0, atol=1e-12 * np.abs(gr0).max())
    np.testing.assert_allclose(H1, H0, rtol=0, atol=1e-12 * np.abs(H0).max())
    r0, r1 = gpu_small.align(small.scan, small.guess), m.align(small.scan, small.guess)
    assert r0.status == r1.status and r0.iterations == r1.iterations
    np.testing.assert_allclose(r1.pose_vector, r0.pose_vector, rtol=0, atol=1e-9)
    m.close()


def test_voxel_arrays_grow_when_the_segment_count_exceeds_the_first_estimate(gpu_small, small):
    """A fresh handle sizes its per-voxel arrays for n/4 + 1024 voxels without
    waiting for the segment count.  A lattice of single points above the map (one
    per cell: too few to be valid) pushes the count past that, so the build grows
    the arrays and runs the per-voxel stage again; a second build on the same
    handle fits at once.  Both give the map's own voxels, lookups and alignment."""
    _print_case("voxel arrays regrow")
    i, j, l = np.meshgrid(np.arange(100), np.arange(100), np.arange(4), indexing="ij")
    lattice = np.stack([-99.0 + 2 * i, -99.0 + 2 * j, 31.0 + 4 * l], -1).reshape(-1, 3).astype(np.float32)
    pts = np.concatenate([small.map_points, lattice])
    keys = np.floor(pts / np.float32(2.0)).astype(np.int64)
    segments = np.unique(keys, axis=0).shape[0]
    assert segments > pts.shape[0] // 4 + 1024  # the premise: the first estimate is too small
    g0 = gpu_small.target_grid()
    o0 = np.lexsort(g0["coords"].T[::-1])
    r0 = gpu_small.align(small.scan, small.guess)
    m = _matcher(pts)
    for build in range(2):
        if build:
            m.set_target(pts)
        g1 = m.target_grid()
        o1 = np.lexsort(g1["coords"].T[::-1])
        for k in ("coords", "means", "covs", "invs", "axes", "counts"):
            assert np.array_equal(g0[k][o0], g1[k][o1]), (build, k)
        c0, i0 = gpu_small.lookup(small.scan, 0, max_k=40)
        c1, i1 = m.lookup(small.scan, 0, max_k=40)
        assert np.array_equal(c0, c1)
        assert _coord_sets(g0["coords"], c0, i0) == _coord_sets(g1["coords"], c1, i1)
        r1 = m.align(small.scan, small.guess)
        assert r0.status == r1.status and r0.iterations == r1.iterations
        np.testing.assert_allclose(r1.pose_vector, r0.pose_vector, rtol=0, atol=1e-9)
    m.close()


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_derivatives_parity(gpu_small, small, oracle_grid, oracle_mod, mode):
    _print_case(f"derivatives mode {mode}")
    from cuda_ndt_matcher_amd import NdtScanMatcherConfig  # noqa: F401

    gpu = _matcher(small.map_points, search_mode=mode) if mode else gpu_small
    rng = np.random.default_rng(3)
    p = oracle_mod.params(search_mode=mode)
    for k in range(4):
        pose = small.guess + np.concatenate([rng.normal(0, 0.2, 3), rng.normal(0, 0.02, 3)])
        for hess in (True, False):
            s_g, g_g, H_g, c_g = gpu.derivatives(small.scan, pose, hess)
            s_o, g_o, H_o, c_o = oracle_grid.derivatives(small.scan, pose, p, hess)
            assert c_g == c_o
            assert s_g == pytest.approx(s_o, rel=1e-10)
            np.testing.assert_allclose(g_g, g_o, rtol=0, atol=1e-9 * np.abs(g_o).max())
            if hess:
                np.testing.assert_allclose(H_g, H_o, rtol=0, atol=1e-9 * np.abs(H_o).max())


def _compare_align(r, o, n_points):
    assert r.status == ["Converged", "MaxIterations", "NoCorrespondences", "SingularHessian"][o.status]
    assert r.iterations == o.iterations
    assert np.abs(r.pose_vector[:3] - o.pose[:3]).max() <= POSE_MM
    assert np.abs(r.pose_vector[3:] - o.pose[3:]).max() <= POSE_RAD
    assert r.nvtl == pytest.approx(o.nvtl, abs=SCORE_TOL)
    assert r.transform_probability == pytest.approx(o.transform_probability, abs=SCORE_TOL)
    if o.status == 0:
        # Converged: the reported Hessian and pair count are those of the final
        # pose, which both sides share.  At max iterations they come from the
        # last iterate, which is not part of the parity bar.
        assert r.num_correspondences == o.num_correspondences
        assert r.oscillation_count == o.oscillation_count
        np.testing.assert_allclose(r.hessian, o.hessian, rtol=0, atol=1e-6 * max(1.0, np.abs(o.hessian).max()))


def test_fp32_derivatives_close_to_fp64_oracle(small, oracle_grid, oracle_mod):
    """Precision variant: per-point terms in FP32 with compensated sums.  The
    neighbour sets are the FP64 ones (the lookups do not change); score,
    gradient and Hessian agree with the FP64 oracle to FP32 rounding of the
    per-pair terms (tolerance 2e-5 of the largest entry)."""
    _print_case("fp32 derivatives")
    gpu = _matcher(small.map_points, precision=1)
    rng = np.random.default_rng(31)
    p = oracle_mod.params()
    for k in range(3):
        pose = small.guess + np.concatenate([rng.normal(0, 0.2, 3), rng.normal(0, 0.02, 3)])
        s_g, g_g, H_g, c_g = gpu.derivatives(small.scan, pose, True)
        s_o, g_o, H_o, c_o = oracle_grid.derivatives(small.scan, pose, p, True)
        assert c_g == c_o
        assert s_g == pytest.approx(s_o, rel=2e-5)
        np.testing.assert_allclose(g_g, g_o, rtol=0, atol=2e-5 * np.abs(g_o).max())
        np.testing.assert_allclose(H_g, H_o, rtol=0, atol=2e-5 * np.abs(H_o).max())


def test_fp32_alignment_matches_fp64_oracle(small, oracle_grid, oracle_mod):
    """Precision variant on the small workload: same status and iteration
    count as the FP64 oracle, pose within 1 mm / 0.01 deg (or, where the
    oracle moves under FP32-size perturbations, the checks of
    oracle/parity.py for such starts)."""
    _print_case("fp32 align")
    p = oracle_mod.params()
    r = _matcher(small.map_points, precision=1).align(small.scan, small.guess)
    parity.check_alignment(r, oracle_grid, small.scan, small.guess, p, noise=1e-6, hessian_rel=1e-4, score_rel=2e-5)


@pytest.mark.parametrize("mode,ls", [(0, False), (1, False), (2, False), (3, False), (0, True)])
def test_align_parity_small(small, oracle_grid, oracle_mod, mode, ls):
    _print_case(f"align small mode {mode} ls {ls}")
    gpu = _matcher(small.map_points, search_mode=mode, use_line_search=ls)
    o = oracle_grid.align(small.scan, small.guess, oracle_mod.params(search_mode=mode, use_line_search=ls))
    r = gpu.align(small.scan, small.guess)
    _compare_align(r, o, small.scan.shape[0])


def test_align_workgroup_counts_agree(small, oracle_grid, oracle_mod):
    """One workgroup per alignment (block barrier) and many (slot barrier) give the oracle's answer."""
    o = oracle_grid.align(small.scan, small.guess, oracle_mod.params())
    for w in (1, 3, 16):
        _print_case(f"align small W={w}")
        r = _matcher(small.map_points, workgroups_per_alignment=w).align(small.scan, small.guess)
        _compare_align(r, o, small.scan.shape[0])


def test_scoring_parity(gpu_small, small, oracle_grid, oracle_mod):
    _print_case("scoring")
    p = oracle_mod.params()
    pose = small.true_pose + np.array([0.05, -0.04, 0.02, 0.001, -0.002, 0.003])
    assert gpu_small.evaluate_nvtl(small.scan, pose) == pytest.approx(oracle_grid.nvtl(small.scan, pose, p),
                                                                       abs=1e-10)
    assert gpu_small.evaluate_transform_probability(small.scan, pose) == pytest.approx(
        oracle_grid.transform_probability(small.scan, pose, p), abs=1e-10)
    np.testing.assert_allclose(gpu_small.compute_per_point_scores(small.scan, pose),
                               oracle_grid.per_point_scores(small.scan, pose, p), rtol=1e-12, atol=1e-12)
    poses = [pose + np.array([d, 0, 0, 0, 0, 0]) for d in np.linspace(-1, 1, 9)]
    np.testing.assert_allclose(gpu_small.evaluate_nvtl_batch(small.scan, poses),
                               [oracle_grid.nvtl_cell27(small.scan, q, p) for q in poses], rtol=1e-12)


def test_align_batch_matches_oracle(gpu_small, small, oracle_grid, oracle_mod):
    _print_case("align batch")
    rng = np.random.default_rng(21)
    # z exact: starts with a z error let the reference's eigenvalue clamp drive
    # large roll/pitch steps and are ill-conditioned (tests/parity.py probe)
    poses = [small.true_pose + np.array([rng.normal(0, 0.3), rng.normal(0, 0.3), 0, 0, 0, rng.normal(0, 0.03)])
             for _ in range(16)]
    res = gpu_small.align_batch(small.scan, poses)
    out = parity.check_many(res, oracle_grid, [(small.scan, q) for q in poses], oracle_mod.params())
    assert [o[0] for o in out].count("match") >= 4, out


def test_align_scans_uneven(gpu_small, small, oracle_grid, oracle_mod):
    _print_case("align scans")
    rng = np.random.default_rng(22)
    scans = []
    for k, n in enumerate((50, 400, 1300, small.scan.shape[0])):
        sel = np.sort(rng.choice(small.scan.shape[0], n, replace=False))
        scans.append((small.scan[sel], small.guess + np.array([0.1 * k, 0, 0, 0, 0, 0])))
    res = gpu_small.align_parallel_scans(scans)
    out = parity.check_many(res, oracle_grid, scans, oracle_mod.params())
    assert [o[0] for o in out].count("match") >= 2, out


def test_error_classes(small):
    _print_case("errors")
    from cuda_ndt_matcher_amd import NdtError

    m = _matcher()
    with pytest.raises(NdtError) as e:
        m.align(small.scan, small.guess)
    assert e.value.code == -3
    with pytest.raises(NdtError) as e:
        m.set_target(np.zeros((0, 3), np.float32))
    assert e.value.code == -2
    with pytest.raises(NdtError) as e:
        m.set_target(np.array([[0, 0, 0], [5, 5, 5]], np.float32))
    assert e.value.code == -4
    m.set_target(small.map_points)
    with pytest.raises(NdtError) as e:
        m.align(np.zeros((0, 3), np.float32), small.guess)
    assert e.value.code == -2


def _voxelize(pts, leaf):
    k = np.floor(pts * (np.float32(1.0) / np.float32(leaf))).astype(np.int64)
    keys, inv = np.unique(k, axis=0, return_inverse=True)
    s = np.zeros((len(keys), 3))
    c = np.zeros(len(keys))
    np.add.at(s, inv.ravel(), pts.astype(np.float64))
    np.add.at(c, inv.ravel(), 1)
    return (s / c[:, None]).astype(np.float32)


def test_half_cubic_reference_sequence(oracle_mod):
    """ndt.rs:1490-1600 through the GPU API, against the oracle."""
    _print_case("half cubic")
    from cuda_ndt_matcher_amd import Isometry3

    mp = npref.half_cubic() + np.array([100.0, 100.0, 0.0], np.float32)
    scan = _voxelize(npref.half_cubic(), 1.0)
    m = _matcher(mp)
    og = oracle_mod.Grid(mp, 2.0)
    good = m.align(scan, Isometry3.translation(100.0, 100.0, 0.0))
    assert abs(good.pose.t[0] - 100) < 2 and abs(good.pose.t[1] - 100) < 2 and abs(good.pose.t[2]) < 2
    _compare_align(good, og.align(scan, [100, 100, 0, 0, 0, 0], oracle_mod.params()), scan.shape[0])
    bad = m.align(scan, Isometry3.translation(-100.0, -100.0, 0.0))
    assert bad.status == "NoCorrespondences" and bad.iterations == 0 and bad.nvtl == 0.0
    assert good.score > bad.score


@pytest.mark.slow
def test_full_size_build_parity(full_wl, full_oracle, oracle_mod):
    """The 5M-point map (about 400k voxels): keys, sort order, counts and
    statistics.  Alignments at this size: tests/test_gpu_configs.py."""
    _print_case("full build")
    m = _matcher(full_wl.map_points)
    _check_build(m, full_wl.map_points, full_oracle, oracle_mod)
    m.close()


def test_pair_evaluations_count_the_derivative_passes(small):
    """pair_evaluations sums the point-voxel pairs of the alignment's
    derivative passes: with one iteration it is the pass at the initial pose;
    over a whole alignment it is at least the final pass's count and, for a
    start near the optimum, close to iterations times it."""
    from cuda_ndt_matcher_amd import NdtScanMatcher

    _print_case("pair evaluations")
    one = NdtScanMatcher.builder().resolution(2.0).max_iterations(1).search_mode(2).build()
    one.set_target(small.map_points)
    _, _, _, c0 = one.derivatives(small.scan, small.guess)
    r1 = one.align(small.scan, small.guess)
    assert r1.iterations == 1 and r1.pair_evaluations == c0 == r1.num_correspondences
    full = NdtScanMatcher.builder().resolution(2.0).max_iterations(30).search_mode(2).build()
    full.set_target(small.map_points)
    r = full.align(small.scan, small.true_pose)
    assert r.pair_evaluations >= r.num_correspondences
    assert abs(r.pair_evaluations - r.iterations * r.num_correspondences) <= 0.05 * r.pair_evaluations
