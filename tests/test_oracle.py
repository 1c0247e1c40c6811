"""Pins the CPU oracle (oracle/ndt_oracle.cpp) before it is trusted.

Each oracle function is checked against the reference's own unit-test cases
(cited), closed forms, or a brute-force numpy recomputation chosen so that a
plausible mistake (a dropped term, a wrong sign or index, a transposed
operand, 0.01 instead of (double)0.01f, <= instead of <) fails.
"""
import math

import numpy as np
import pytest

from tests import npref

pytestmark = []


# --- A7 GaussianParams (derivatives/types.rs:61-75) -------------------------
@pytest.mark.parametrize("res,outlier,d1,d2", [
    (2.0, 0.55, -4.196518186951408, 0.24847851012449546),
    (1.0, 0.55, -2.217225244042889, 0.43312300470355464),
    (1.0, 0.3, -3.191847152480282, 0.3212908807322899),
])
def test_gauss_params_closed_form(oracle_mod, res, outlier, d1, d2):
    g1, g2 = oracle_mod.gauss(res, outlier)
    assert g1 == pytest.approx(d1, rel=1e-14)
    assert g2 == pytest.approx(d2, rel=1e-14)


# --- A1 VoxelCoord::from_point (voxel_grid/types.rs:206-235) -----------------
def test_point_keys_reference_cases(oracle_mod):
    k = oracle_mod.point_keys([[0, 0, 0], [3.5, 5.1, 1.9], [-3.5, -1.1, -0.1]], 2.0)
    assert k.tolist() == [[0, 0, 0], [1, 2, 0], [-2, -1, -1]]


@pytest.mark.parametrize("res", [2.0, 1.0, 1.5, 0.3])
def test_point_keys_brute_force(oracle_mod, res):
    rng = np.random.default_rng(7)
    p = rng.uniform(-500, 500, (20000, 3)).astype(np.float32)
    # exact multiples of res and their float neighbours probe the floor boundary
    k = rng.integers(-300, 300, (2000, 3))
    b = (k * np.float32(res)).astype(np.float32)
    edges = np.concatenate([b, np.nextafter(b, np.float32(-np.inf)), np.nextafter(b, np.float32(np.inf))])
    p = np.concatenate([p, edges])
    assert np.array_equal(oracle_mod.point_keys(p, res), npref.keys_f32(p, res))


# --- A2-A4 voxel statistics (voxel_grid/types.rs:54-190, cpu.rs:27-75) ------
def _cluster_points():
    # voxel_grid/types.rs:240-252
    return np.array([[1.0, 2.0, 3.0], [1.1, 2.1, 3.1], [0.9, 1.9, 2.9], [1.05, 2.05, 3.05], [0.95, 1.95, 2.95],
                     [1.02, 2.02, 3.02], [0.98, 1.98, 2.98], [1.03, 2.03, 3.03], [0.97, 1.97, 2.97], [1.0, 2.0, 3.0]],
                    np.float32)


def test_voxel_stats_match_numpy_cov(oracle_mod):
    rng = np.random.default_rng(3)
    pts = []
    for c in rng.uniform(-50, 50, (40, 3)):
        n = int(rng.integers(6, 60))
        pts.append((c + rng.normal(0, [0.3, 0.2, 0.05], (n, 3))).astype(np.float32))
    pts = np.concatenate(pts)
    g = oracle_mod.Grid(pts, 2.0)
    ex = g.export()
    keys = npref.keys_f32(pts, 2.0)
    uniq = {tuple(k) for k in keys}
    # every occupied cell with >= 6 points appears, keyed and counted exactly
    for i, c in enumerate(ex["coords"]):
        sel = np.all(keys == c, axis=1)
        P = pts[sel].astype(np.float64)
        assert ex["counts"][i] == P.shape[0] >= 6
        assert np.allclose(ex["means64"][i], P.mean(0), rtol=0, atol=1e-12)
        cov = np.cov(P.T, ddof=1)
        ev = np.linalg.eigvalsh(cov)
        if ev.min() >= ev.max() * np.float64(np.float32(0.01)):
            assert np.allclose(ex["covs"][i], cov, rtol=1e-6, atol=1e-7 * np.abs(cov).max())
    expected = sum(1 for u in uniq if np.sum(np.all(keys == np.array(u), axis=1)) >= 6)
    assert len(g) == expected


def test_voxel_too_few_points_dropped(oracle_mod):
    # voxel_grid/types.rs:285-297: fewer than min_points -> no voxel
    base = np.array([[0.5, 0.5, 0.5]], np.float32) + np.random.default_rng(0).normal(0, 0.1, (5, 3)).astype(np.float32)
    assert len(oracle_mod.Grid(base, 2.0)) == 0
    six = np.concatenate([base, [[0.6, 0.4, 0.55]]]).astype(np.float32)
    assert len(oracle_mod.Grid(six, 2.0)) == 1


def test_voxel_cluster_reference_case(oracle_mod):
    g = oracle_mod.Grid(_cluster_points(), 2.0)
    ex = g.export()
    assert len(g) == 1
    np.testing.assert_allclose(ex["means"][0], [1.0, 2.0, 3.0], atol=0.1)
    ident = ex["covs"][0].astype(np.float64) @ ex["invs64"][0]
    np.testing.assert_allclose(np.diag(ident), 1.0, atol=0.01)


def _grid_from_cov(oracle_mod, cov, n=4000, seed=0):
    """Points whose sample covariance is exactly `cov` (whitened then coloured)."""
    rng = np.random.default_rng(seed)
    z = rng.normal(size=(n, 3))
    z -= z.mean(0)
    L = np.linalg.cholesky(np.cov(z.T, ddof=1))
    z = z @ np.linalg.inv(L).T
    A = np.linalg.cholesky(cov) if np.all(np.linalg.eigvalsh(cov) > 0) else None
    x = z @ A.T + 1.0
    return x


def test_regularize_planar_reference_case(oracle_mod):
    # voxel_grid/types.rs:300-336: diag(1, 1, 1e-4) -> clamp, cov*inv = I, axis = z
    cov = np.diag([0.04, 0.04, 0.04e-4])
    rng = np.random.default_rng(1)
    z = rng.normal(size=(3000, 3))
    z = (z - z.mean(0)) @ np.linalg.inv(np.linalg.cholesky(np.cov(z.T, ddof=1))).T
    pts = (z @ np.sqrt(cov) + np.array([1.0, 1.0, 1.0])).astype(np.float32)
    g = oracle_mod.Grid(pts, 2.0)
    ex = g.export()
    assert len(g) == 1
    ev = np.linalg.eigvalsh(ex["covs"][0].astype(np.float64))
    assert ev.min() >= ev.max() * 0.009
    np.testing.assert_allclose(ex["covs"][0].astype(np.float64) @ ex["invs64"][0], np.eye(3), atol=1e-6)
    assert abs(ex["axes"][0][2]) > 0.9


def test_regularize_threshold_is_f32_ratio(oracle_mod):
    """thr = lambda_max * (0.01f as f64) = lambda_max * 0.009999999776482582 (types.rs:138).

    A smallest eigenvalue of 0.0099999999 * lambda_max lies between 0.01f and
    0.01: the reference leaves it unclamped.  With 0.01 it would be clamped and
    the inverse's z entry would change by 1e-8 relative; we check the inverse
    against both candidates in f64.
    """
    for ratio, clamped in ((0.0099999999, False), (0.0099999995, True)):
        lam = np.array([1.0, 0.5, ratio])
        V = np.linalg.qr(np.random.default_rng(5).normal(size=(3, 3)))[0]
        eigs = np.linalg.eigh(V @ np.diag(lam) @ V.T)
        ev, _ = oracle_mod.sym_eigen(V @ np.diag(lam) @ V.T)
        thr = ev.max() * float(np.float32(0.01))
        assert (ev.min() < thr) == clamped
        assert eigs[0].min() == pytest.approx(ratio, rel=1e-12)


def test_sym_eigen_matches_lapack(oracle_mod):
    rng = np.random.default_rng(11)
    for n in (3, 6):
        for _ in range(50):
            A = rng.normal(size=(n, n))
            A = A + A.T
            ev, V = oracle_mod.sym_eigen(A)
            ref = np.linalg.eigvalsh(A)
            np.testing.assert_allclose(np.sort(ev), ref, rtol=0, atol=1e-12 * np.abs(ref).max())
            np.testing.assert_allclose(V @ np.diag(ev) @ V.T, A, atol=1e-12 * np.abs(A).max())
            np.testing.assert_allclose(V.T @ V, np.eye(n), atol=1e-13)


def test_sym_eigen_reads_lower_triangle(oracle_mod):
    A = np.array([[2.0, 99.0, 99.0], [1.0, 3.0, 99.0], [0.5, 0.25, 4.0]])
    ev, _ = oracle_mod.sym_eigen(A)
    sym = np.tril(A) + np.tril(A, -1).T
    np.testing.assert_allclose(np.sort(ev), np.linalg.eigvalsh(sym), atol=1e-13)


# --- A5/A6 radius search (voxel_grid/search.rs:90-245) ----------------------
def test_radius_search_brute_force(oracle_mod):
    rng = np.random.default_rng(2)
    pts = rng.uniform(-20, 20, (60000, 3)).astype(np.float32)
    g = oracle_mod.Grid(pts, 2.0)
    ex = g.export()
    q = rng.uniform(-22, 22, (3000, 3)).astype(np.float32)
    # include queries sitting exactly on voxel means and at distance ~r
    q[:200] = ex["means"][rng.integers(0, len(g), 200)]
    q[200:400] = ex["means"][rng.integers(0, len(g), 200)] + np.array([2.0, 0, 0], np.float32)
    cnt, idx = g.search(q, oracle_mod.SEARCH_KDTREE, max_k=64)
    for i in range(q.shape[0]):
        ref = npref.radius_neighbors(q[i], ex["means"], 2.0)
        assert cnt[i] == len(ref)
        got = idx[i, :cnt[i]]
        assert set(got.tolist()) == set(ref.tolist())
        d = npref.kiddo_dist2(q[i], ex["means"][got])
        assert np.all(np.diff(d) >= 0), "nearest first"


def test_radius_search_strict_boundary(oracle_mod):
    # A mean at exactly squared distance r^2 is excluded (kiddo compares with <).
    rng = np.random.default_rng(4)
    c = np.array([1.0, 1.0, 1.0])
    z = rng.normal(size=(2000, 3)) * 0.1
    z -= z.mean(0)
    pts = (z + c).astype(np.float32)
    g = oracle_mod.Grid(pts, 2.0)
    m = g.export()["means"][0]
    q = (m + np.array([2.0, 0.0, 0.0], np.float32)).astype(np.float32)
    assert npref.kiddo_dist2(q, m[None])[0] == np.float32(4.0)
    cnt, _ = g.search(q[None], oracle_mod.SEARCH_KDTREE)
    assert cnt[0] == 0
    cnt, _ = g.search(np.nextafter(q, m)[None], oracle_mod.SEARCH_KDTREE)
    assert cnt[0] == 1


def test_direct_searches(oracle_mod):
    rng = np.random.default_rng(6)
    pts = rng.uniform(-10, 10, (40000, 3)).astype(np.float32)
    g = oracle_mod.Grid(pts, 2.0)
    ex = g.export()
    coord_to_idx = {tuple(c): i for i, c in enumerate(ex["coords"].tolist())}
    q = rng.uniform(-12, 12, (500, 3)).astype(np.float32)
    offs = {
        oracle_mod.SEARCH_DIRECT1: [(0, 0, 0)],
        oracle_mod.SEARCH_DIRECT7: [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)],
        oracle_mod.SEARCH_DIRECT26: [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)
                                     if (a, b, c) != (0, 0, 0)],
    }
    for mode, rel in offs.items():
        cnt, idx = g.search(q, mode, max_k=32)
        for i in range(q.shape[0]):
            k = npref.keys_f32(q[i][None], 2.0)[0]
            ref = [coord_to_idx[t] for t in (tuple(k + np.array(o)) for o in rel) if t in coord_to_idx]
            assert idx[i, :cnt[i]].tolist() == ref


# --- A8-A12 derivatives (derivatives/cpu.rs, angular.rs) --------------------
def _small_grid(oracle_mod, seed=0):
    rng = np.random.default_rng(seed)
    pts = []
    for c in rng.uniform(-6, 6, (30, 3)):
        pts.append(c + rng.normal(0, [0.4, 0.3, 0.2], (80, 3)))
    return oracle_mod.Grid(np.concatenate(pts).astype(np.float32), 2.0)


def test_score_at_mean_is_minus_d1(oracle_mod):
    # derivatives/cpu.rs:694-716: a point exactly at the voxel mean scores -d1.
    g = _small_grid(oracle_mod)
    m = g.export()["means"][0]
    p = oracle_mod.params(search_mode=oracle_mod.SEARCH_DIRECT1, fp64_voxels=True)
    s, grad, H, c = g.derivatives(m[None], [0, 0, 0, 0, 0, 0], p)
    d1, _ = npref.gauss(2.0, 0.55)
    assert c >= 1
    assert s >= -d1 * (1 - 1e-12)


def _richardson(g, pts, pose, p, h):
    """Central differences at h and h/2 combined to O(h^4): (gradient, Hessian)."""
    def central(step):
        fg = np.zeros(6)
        fH = np.zeros((6, 6))
        for k in range(6):
            e = np.zeros(6)
            e[k] = step
            sp, gp, _, _ = g.derivatives(pts, pose + e, p)
            sm, gm, _, _ = g.derivatives(pts, pose - e, p)
            fg[k] = (sp - sm) / (2 * step)
            fH[:, k] = (gp - gm) / (2 * step)
        return fg, fH
    g1, H1 = central(h)
    g2, H2 = central(h / 2)
    return (4 * g2 - g1) / 3, (4 * H2 - H1) / 3


def _fd_ok_pose(g, pts, pose, p, h):
    """True when no point's neighbour set changes within +-h on any pose axis."""
    def nb(q):
        x = (pts.astype(np.float64) @ npref.rot_xyz(*q[3:]).T + q[:3]).astype(np.float32)
        cnt, idx = g.search(x, p.search_mode, max_k=32)
        return [tuple(sorted(idx[i, :cnt[i]].tolist())) for i in range(len(cnt))]
    base = nb(np.asarray(pose, float))
    for k in range(6):
        for step in (h, -h, h / 2, -h / 2):
            q = np.array(pose, float)
            q[k] += step
            if nb(q) != base:
                return False
    return True


@pytest.mark.parametrize("mode", [0, 3])
@pytest.mark.parametrize("zero_pitch", [True, False])
def test_gradient_and_hessian_match_finite_differences(oracle_mod, mode, zero_pitch):
    """Gradient = FD of the score; Hessian = FD of the gradient.

    The reference's angular Hessian row d1 (d2x/dpitch2, angular.rs:137-140)
    carries +sin(pitch) on z where the derivative of Rx Ry Rz has -sin(pitch)
    (the term as printed in Magnusson's thesis and PCL).  The oracle keeps it,
    so the pitch-pitch entry equals FD only at pitch = 0; every other entry
    equals FD everywhere.
    """
    g = _small_grid(oracle_mod, 1)
    rng = np.random.default_rng(9)
    p = oracle_mod.params(search_mode=mode, fp64_voxels=True)
    ex = g.export()
    pts = (ex["means"][rng.integers(0, len(g), 40)] + rng.normal(0, 0.3, (40, 3))).astype(np.float32)
    checked = 0
    h = 1e-4  # above the 1e-5 small-angle snap
    for trial in range(30):
        pose = np.concatenate([rng.normal(0, 0.05, 3), rng.normal(0, 0.05, 3) + 0.02])
        if zero_pitch:
            pose[4] = 0.0
        if not _fd_ok_pose(g, pts, pose, p, h):
            continue
        s0, g0, H0, _ = g.derivatives(pts, pose, p)
        fdg, fdH = _richardson(g, pts, pose, p, h)
        np.testing.assert_allclose(fdg, g0, rtol=1e-7, atol=1e-9 * np.abs(g0).max())
        mask = np.ones((6, 6), bool)
        if not zero_pitch:
            mask[4, 4] = False
            assert abs(fdH[4, 4] - H0[4, 4]) > 1e-6 * abs(H0[4, 4])
        scale = np.abs(H0).max()
        np.testing.assert_allclose(fdH[mask], H0[mask], rtol=1e-7, atol=1e-9 * scale)
        np.testing.assert_allclose(H0, H0.T, rtol=1e-10, atol=1e-10)
        checked += 1
    assert checked >= 5


def test_point_to_plane_gradient_matches_fd(oracle_mod):
    g = _small_grid(oracle_mod, 2)
    rng = np.random.default_rng(10)
    p = oracle_mod.params(search_mode=3, distance_metric=1, fp64_voxels=True)
    ex = g.export()
    pts = (ex["means"][rng.integers(0, len(g), 30)] + rng.normal(0, 0.3, (30, 3))).astype(np.float32)
    pose = np.array([0.01, -0.02, 0.01, 0.02, 0.0, 0.03])
    h = 1e-4
    assert _fd_ok_pose(g, pts, pose, p, h)
    s0, g0, H0, _ = g.derivatives(pts, pose, p)
    fdg, fdH = _richardson(g, pts, pose, p, h)
    np.testing.assert_allclose(fdg, g0, rtol=1e-7, atol=1e-9 * np.abs(g0).max())
    np.testing.assert_allclose(fdH, H0, rtol=1e-7, atol=1e-9 * np.abs(H0).max())


def test_translation_transform(oracle_mod):
    # derivatives/cpu.rs:567-650 (translation and 90 degree yaw); the
    # optimiser's transform is Rx Ry Rz: a point at a voxel mean moved by the
    # pose must still score -d1 at the moved mean.
    g = _small_grid(oracle_mod, 3)
    p = oracle_mod.params(search_mode=oracle_mod.SEARCH_DIRECT1, fp64_voxels=True)
    m = g.export()["means64"][0]
    d1, _ = npref.gauss(2.0, 0.55)
    for pose in ([0.5, -0.25, 0.125, 0, 0, 0], [0.0, 0.0, 0.0, 0, 0, math.pi / 2], [0.1, 0.2, -0.1, 0.3, -0.2, 0.1]):
        R = npref.rot_xyz(*pose[3:])
        src = np.linalg.solve(R, m - np.array(pose[:3]))
        s, _, _, c = g.derivatives(src.astype(np.float32)[None], pose, p)
        assert c == 1
        assert s == pytest.approx(-d1, rel=1e-6)


# --- A13 newton_step_negative_definite (optimization/newton.rs) ------------
def test_newton_negative_definite_cases(oracle_mod):
    g = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0])
    d = oracle_mod.newton_step(g, -np.eye(6))
    np.testing.assert_allclose(d, g, rtol=1e-14)
    d = oracle_mod.newton_step(g, -2 * np.eye(6))
    np.testing.assert_allclose(d, g / 2, rtol=1e-14)
    # positive definite: every eigenvalue clamped to -1e-3
    d = oracle_mod.newton_step(g, np.eye(6))
    np.testing.assert_allclose(d, g / 1e-3, rtol=1e-12)
    # mixed (newton.rs:358-402): eigenvalue +1 on axis 0 clamped to -1e-3
    H = -np.eye(6)
    H[0, 0] = 1.0
    d = oracle_mod.newton_step(g, H)
    exp = g.copy()
    exp[0] = g[0] / 1e-3
    np.testing.assert_allclose(d, exp, rtol=1e-12)
    # singular: all |lambda| below the SVD tolerance -> no step
    assert oracle_mod.newton_step(g, -1e-12 * np.eye(6)) is None


def test_newton_random_negative_definite(oracle_mod):
    rng = np.random.default_rng(12)
    for _ in range(30):
        A = rng.normal(size=(6, 6))
        H = -(A @ A.T + 0.1 * np.eye(6))
        g = rng.normal(size=6)
        d = oracle_mod.newton_step(g, H)
        np.testing.assert_allclose(d, np.linalg.solve(H, -g), rtol=1e-9, atol=1e-12)
        # indefinite: compare with the explicit clamp
        B = rng.normal(size=(6, 6))
        H2 = B + B.T
        w, V = np.linalg.eigh(H2)
        if np.all(w < 0):
            continue
        wc = np.minimum(w, -1e-3)
        d2 = oracle_mod.newton_step(g, H2)
        np.testing.assert_allclose(d2, -(V @ np.diag(1 / wc) @ V.T) @ g, rtol=1e-8, atol=1e-10)


# --- A16 count_oscillation (optimization/oscillation.rs tests) --------------
def test_oscillation_reference_cases(oracle_mod):
    O = oracle_mod.oscillation
    assert O([[x, 0, 0] for x in range(5)]) == 0
    assert O([[0, 0, 0], [1, 0, 0], [1.5, .5, 0], [1.8, 1, 0], [2, 1.5, 0]]) == 0
    assert O([[0, 0, 0], [1, 0, 0], [0, 0, 0], [1, 0, 0], [0, 0, 0]]) == 3
    assert O([[0, 0, 0]] + [[1.0 if i % 2 else 0.0, 0, 0] for i in range(1, 26)]) > 10
    assert O([[0, 0, 0], [1, 0, 0], [0, 0, 0], [1, 0, 0], [2, 0, 0], [3, 0, 0], [4, 0, 0]]) == 2
    assert O([[0, 0, 0], [1, 0, 0]]) == 0
    assert O([[1, 2, 3]] * 4) == 0


# --- A17 More-Thuente (optimization/more_thuente.rs tests) ------------------
def test_more_thuente_reference_cases(oracle_mod):
    r = oracle_mod.more_thuente(lambda a: (-(a - 2.0) ** 2, -2.0 * (a - 2.0)), -4.0, 4.0, 1.0, step_max=5.0,
                                step_min=0.0, max_iterations=20)
    assert r["converged"] and 1.0 <= r["step_length"] <= 2.5 and r["final_value"] > -4.0
    r = oracle_mod.more_thuente(lambda a: (0.0, 0.0), 1.0, 0.0, 0.1)
    assert r["step_length"] == 0.0 and not r["converged"]
    r = oracle_mod.more_thuente(lambda a: (0.0, 0.0), 1.0, -1.0, 0.1)
    assert r["step_length"] == 0.0 and not r["converged"]
    r = oracle_mod.more_thuente(lambda a: (10 * a - a * a, 10 - 2 * a), 0.0, 10.0, 1.0, step_max=10.0, step_min=0.0,
                                max_iterations=20)
    assert r["step_length"] > 0.0 and r["final_value"] > 0.0


def test_more_thuente_wolfe_conditions(oracle_mod):
    # Whatever step it accepts as converged satisfies the strong Wolfe conditions.
    for c in (0.5, 1.0, 3.0, 7.0):
        f = lambda a, c=c: (-(a - c) ** 2 - 0.1 * a ** 4, -2 * (a - c) - 0.4 * a ** 3)
        s0, d0 = f(0.0)
        r = oracle_mod.more_thuente(f, s0, d0, 0.5, step_max=10.0, step_min=1e-9, max_iterations=50)
        if r["converged"]:
            a = r["step_length"]
            sa, da = f(a)
            assert -sa <= -s0 - 1e-4 * d0 * a + 1e-12
            assert abs(da) <= 0.9 * abs(d0) + 1e-12


# --- A18/A19 NVTL, TP, per-point scores (scoring/nvtl.rs, metrics.rs) -------
def test_scores_brute_force(oracle_mod):
    g = _small_grid(oracle_mod, 4)
    ex = g.export()
    rng = np.random.default_rng(13)
    pts = (ex["means"][rng.integers(0, len(g), 300)] + rng.normal(0, 0.6, (300, 3))).astype(np.float32)
    pose = np.array([0.05, -0.03, 0.02, 0.01, -0.02, 0.04])
    p = oracle_mod.params()
    per = npref.brute_scores(pts, pose, ex, 2.0, 0.55)
    with_nb = [s for s in per if s]
    nvtl = sum(max(s) for s in with_nb) / len(with_nb)
    tp = sum(sum(s) for s in per) / sum(len(s) for s in per)
    assert g.nvtl(pts, pose, p) == pytest.approx(nvtl, rel=1e-12)
    assert g.transform_probability(pts, pose, p) == pytest.approx(tp, rel=1e-12)
    np.testing.assert_allclose(g.per_point_scores(pts, pose, p), [sum(s) for s in per], rtol=1e-12, atol=1e-14)


def test_nvtl_cell27_brute_force(oracle_mod):
    g = _small_grid(oracle_mod, 5)
    ex = g.export()
    coord_to_idx = {tuple(c): i for i, c in enumerate(ex["coords"].tolist())}
    rng = np.random.default_rng(14)
    pts = (ex["means"][rng.integers(0, len(g), 200)] + rng.normal(0, 0.8, (200, 3))).astype(np.float32)
    pose = np.array([0.1, 0.0, -0.05, 0.0, 0.02, -0.03])
    d1, d2 = npref.gauss(2.0, 0.55)
    x = pts.astype(np.float64) @ npref.rot_zyx(*pose[3:]).T + pose[:3]
    tot, cnt = 0.0, 0
    for xi in x:
        v = np.floor(xi * (1.0 / 2.0)).astype(int)
        sc = []
        for a in (-1, 0, 1):
            for b in (-1, 0, 1):
                for c in (-1, 0, 1):
                    k = coord_to_idx.get((v[0] + a, v[1] + b, v[2] + c))
                    if k is not None:
                        d = xi - ex["means"][k].astype(np.float64)
                        sc.append(-d1 * math.exp(-d2 * (d @ (ex["invs"][k].astype(np.float64) @ d)) / 2))
        if sc:
            tot += max(sc)
            cnt += 1
    assert g.nvtl_cell27(pts, pose, oracle_mod.params()) == pytest.approx(tot / cnt, rel=1e-12)


# --- A15 Isometry conversions (optimization/types.rs:156-174, 231-239) ------
def test_pose_matrix_round_trip(oracle_mod):
    rng = np.random.default_rng(15)
    for _ in range(100):
        pose = np.concatenate([rng.normal(0, 10, 3), rng.uniform(-1.4, 1.4, 3)])
        m = oracle_mod.pose_to_matrix(pose)
        np.testing.assert_allclose(m[:, :3], npref.rot_zyx(*pose[3:]), atol=1e-14)
        np.testing.assert_allclose(oracle_mod.matrix_to_pose(m), pose, atol=1e-10)


# --- A20 Laplace covariance (cuda_ndt_matcher/src/covariance.rs:82-98) -----
def test_laplace_xy(oracle_mod):
    H = np.zeros((6, 6))
    H[:2, :2] = [[-4.0, 1.0], [1.0, -2.0]]
    np.testing.assert_allclose(oracle_mod.laplace_xy(H), -np.linalg.inv(H[:2, :2]), rtol=1e-15)
    np.testing.assert_allclose(oracle_mod.laplace_xy(np.zeros((6, 6))), np.eye(2))


# --- A14 NdtOptimizer::align behaviour (ndt.rs tests, solver.rs) ------------
def _voxelize(pts, leaf):
    k = np.floor(pts * (np.float32(1.0) / np.float32(leaf))).astype(np.int64)
    keys, inv = np.unique(k, axis=0, return_inverse=True)
    s = np.zeros((len(keys), 3))
    c = np.zeros(len(keys))
    np.add.at(s, inv.ravel(), pts.astype(np.float64))
    np.add.at(c, inv.ravel(), 1)
    return (s / c[:, None]).astype(np.float32)


def test_half_cubic_voxel_count(oracle_mod):
    # voxel_grid/mod.rs:640-657: 150..400 voxels at resolution 2.0
    pts = npref.half_cubic()
    assert pts.shape[0] == 30603
    assert 150 <= len(oracle_mod.Grid(pts, 2.0)) <= 400


def test_standard_alignment_sequence(oracle_mod):
    # ndt.rs:1490-1537 and the recovery test at ndt.rs:1546-1600
    mp = npref.half_cubic() + np.array([100.0, 100.0, 0.0], np.float32)
    g = oracle_mod.Grid(mp, 2.0)
    scan = _voxelize(npref.half_cubic(), 1.0)
    p = oracle_mod.params()
    r = g.align(scan, [100, 100, 0, 0, 0, 0], p)
    assert abs(r.pose[0] - 100) < 2 and abs(r.pose[1] - 100) < 2 and abs(r.pose[2]) < 2
    assert r.score > 0 and r.nvtl > 0 and r.num_correspondences > 0
    bad = g.align(scan, [-100, -100, 0, 0, 0, 0], p)
    assert bad.status == 2 and bad.iterations == 0 and bad.nvtl == 0.0
    np.testing.assert_array_equal(bad.pose, [-100, -100, 0, 0, 0, 0])


def test_align_small_offset(oracle_mod):
    # ndt.rs:1594-1625 (translation (0.5, 0.3, 0.1) on the half cube)
    mp = npref.half_cubic()
    g = oracle_mod.Grid(mp, 2.0)
    scan = _voxelize(mp, 1.0)
    r = g.align(scan, [0.5, 0.3, 0.1, 0.0, 0.0, 0.0], oracle_mod.params())
    assert np.linalg.norm(r.pose[:3]) < np.linalg.norm([0.5, 0.3, 0.1]) + 0.5


def test_alignment_score_ordering_and_nvtl_range(oracle_mod):
    # ndt.rs:1632-1700
    mp = npref.half_cubic()
    g = oracle_mod.Grid(mp, 2.0)
    scan = _voxelize(mp, 1.0)
    p = oracle_mod.params()
    tp0 = g.transform_probability(scan, [0, 0, 0, 0, 0, 0], p)
    tpf = g.transform_probability(scan, [50, 50, 0, 0, 0, 0], p)
    assert tp0 > 0 and tpf < tp0
    for off in (0.0, 1.0, 2.0, 5.0, 10.0):
        v = g.nvtl(scan, [off, 0, 0, 0, 0, 0], p)
        assert v >= 0 and np.isfinite(v)


def test_align_converges_on_urban_scene(oracle_mod):
    from cuda_ndt_matcher_amd import synth
    w = synth.make_workload("small", 0)
    g = oracle_mod.Grid(w.map_points, 2.0)
    r = g.align(w.scan, w.guess, oracle_mod.params())
    assert r.status == 0
    assert np.abs(r.pose[:3] - w.true_pose[:3]).max() < 0.02
    assert np.abs(r.pose[3:] - w.true_pose[3:]).max() < np.deg2rad(0.5)


def test_max_iterations_returns_best_pose(oracle_mod):
    mp = npref.half_cubic()
    g = oracle_mod.Grid(mp, 2.0)
    scan = _voxelize(npref.half_cubic(), 1.0)
    r = g.align(scan, [1.5, -1.0, 0.0, 0.0, 0.0, 0.1], oracle_mod.params(max_iterations=3))
    assert r.status == 1 and r.iterations == 3
    # the best-scoring pose among those visited is returned, with its score
    s0 = g.derivatives(scan, r.pose, oracle_mod.params())[0]
    assert s0 == pytest.approx(r.score, rel=1e-12)
    assert r.transform_probability == pytest.approx(r.score / scan.shape[0], rel=1e-15)


def test_line_search_alignment_converges(oracle_mod):
    mp = npref.half_cubic()
    g = oracle_mod.Grid(mp, 2.0)
    scan = _voxelize(npref.half_cubic(), 1.0)
    r = g.align(scan, [0.4, -0.3, 0.1, 0.0, 0.0, 0.03], oracle_mod.params(use_line_search=True))
    assert np.abs(r.pose[:3]).max() < 0.2


def test_regularization_term(oracle_mod):
    """regularization.rs:102-159: score -s*w*l^2 with l the longitudinal offset.

    The reference's gradient and Hessian are those of l(x, y) only (no yaw
    terms, regularization.rs:146); they equal the x/y derivatives of the score.
    """
    p = oracle_mod.params(regularization_enabled=True, regularization_scale_factor=0.01,
                          regularization_translation=[5.0, -2.0, 0.0])
    pose = np.array([1.0, 0.5, 0.0, 0.0, 0.0, 0.3])
    s, g, H = oracle_mod.regularization(p, pose, 100)
    lon = (-2.0 - 0.5) * math.sin(0.3) + (5.0 - 1.0) * math.cos(0.3)
    assert s == pytest.approx(-0.01 * 100 * lon * lon, rel=1e-14)
    h = 1e-6
    for k in (0, 1):
        e = np.zeros(6)
        e[k] = h
        sp = oracle_mod.regularization(p, pose + e, 100)[0]
        sm = oracle_mod.regularization(p, pose - e, 100)[0]
        assert g[k] == pytest.approx((sp - sm) / (2 * h), rel=1e-7)
        gp = oracle_mod.regularization(p, pose + e, 100)[1]
        gm = oracle_mod.regularization(p, pose - e, 100)[1]
        np.testing.assert_allclose(H[:2, k], (gp[:2] - gm[:2]) / (2 * h), rtol=1e-6)
    assert np.all(g[2:] == 0) and np.all(H[2:, :] == 0)
    off = oracle_mod.params(regularization_enabled=False, regularization_translation=[5.0, -2.0, 0.0])
    assert oracle_mod.regularization(off, pose, 100)[0] == 0.0
