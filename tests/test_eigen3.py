"""The voxel covariances' 3x3 eigendecomposition (csrc/eigen3.hpp: cyclic
Jacobi whose rotations use the device's estimate-and-refine reciprocals),
built for the host from the same source the device compiles, against LAPACK
(numpy.linalg.eigvalsh), the defining identities, closed forms and the
oracle's Jacobi in IEEE arithmetic, whose eigenpairs the reference's
regularisation (voxel_grid/types.rs:115-190) consumes."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
P = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("eigen3") / "libeigen3.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-D__host__=",
                    "-D__device__=", os.path.join(HERE, "native", "eigen6_capi.cpp"), "-o", str(out)], check=True)
    return ctypes.CDLL(str(out))


def eig3(lib, A, mode=1):
    A = np.ascontiguousarray(np.tril(A), np.float64)  # only the lower triangle is read
    ev, V = np.zeros(3), np.zeros(9)
    lib.sym3_full(A.ctypes.data_as(P), ev.ctypes.data_as(P), V.ctypes.data_as(P), mode)
    return ev, V.reshape(3, 3)


def covariances(rng, count):
    """Sample covariances of voxel-like point sets: isotropic, planar (walls,
    ground), linear (poles), rotated, badly scaled and exactly planar."""
    for t in range(count):
        n = int(rng.integers(6, 60))
        X = rng.normal(size=(n, 3))
        kind = t % 6
        if kind == 1:
            X[:, 2] *= 1e-4
        elif kind == 2:
            X[:, 1:] *= 1e-5
        elif kind == 3:
            X *= 10 ** rng.uniform(-4, 2)
        elif kind == 4:
            X = (X * [1.0, 1.0, 1e-3]) @ np.linalg.qr(rng.normal(size=(3, 3)))[0]
        elif kind == 5:
            X[:, 2] = 0.0
        Xc = X - X.mean(0)
        yield Xc.T @ Xc / (n - 1)


@pytest.mark.parametrize("mode", [0, 1])
def test_decomposition_identities(lib, mode):
    rng = np.random.default_rng(7)
    for A in covariances(rng, 6000):
        ev, V = eig3(lib, A, mode)
        nrm = np.abs(A).max()
        assert np.abs(A @ V - V * ev).max() <= 1e-14 * nrm
        assert np.abs(V.T @ V - np.eye(3)).max() <= 1e-14
        assert np.abs(np.sort(ev) - np.linalg.eigvalsh(A)).max() <= 1e-14 * nrm


def test_closed_forms(lib):
    # diagonal: no rotation, eigenvalues exact, V = I
    D = np.diag([3.0, 1e-6, 250.0])
    ev, V = eig3(lib, D)
    assert np.array_equal(ev, np.diag(D)) and np.array_equal(V, np.eye(3))
    # isotropic: c I stays as it is
    ev, V = eig3(lib, 0.37 * np.eye(3))
    assert np.array_equal(ev, [0.37] * 3) and np.array_equal(V, np.eye(3))
    # [[a, c], [c, a]] block: eigenvalues a +- c, vectors (1, +-1) / sqrt 2
    A = np.diag([4.0, 4.0, 9.0])
    A[0, 1] = A[1, 0] = 1.5
    ev, V = eig3(lib, A)
    assert np.allclose(np.sort(ev), [2.5, 5.5, 9.0], rtol=0, atol=1e-15)
    k = int(np.argmin(ev))
    assert np.allclose(np.abs(V[:2, k]), [2 ** -0.5, 2 ** -0.5], rtol=0, atol=1e-16)
    # rank one u u^T: |u|^2 once, zero twice
    u = np.array([1.0, -2.0, 0.5])
    ev, _ = eig3(lib, np.outer(u, u))
    assert abs(ev.max() - u @ u) <= 2e-15 * (u @ u)
    assert np.abs(np.sort(ev)[:2]).max() <= 2e-15 * (u @ u)
    # an off-diagonal far below the diagonal gap takes the IEEE branch
    A = np.diag([1.0, 2.0, 3.0])
    A[1, 0] = A[0, 1] = 1e-160
    ev, V = eig3(lib, A)
    assert np.array_equal(np.sort(ev), [1.0, 2.0, 3.0]) and np.abs(V - np.eye(3)).max() <= 1e-159


def regularise(ev, V, ratio=np.float64(np.float32(0.01))):
    """voxel_grid/types.rs:115-190 as voxel_stats applies it: eigenvalues below
    ratio * max are raised to it; the rebuilt covariance and its inverse."""
    mx = max(0.0, ev.max())
    cl = np.where(ev < mx * ratio, mx * ratio, ev)
    return V @ np.diag(cl) @ V.T, V @ np.diag(1.0 / cl) @ V.T


def test_regularised_covariance_matches_the_oracle(lib, oracle_mod):
    """The quantities the voxel statistics carry (regularised covariance and its
    inverse) from the device arithmetic agree with the oracle's IEEE Jacobi far
    inside the 1e-5 relative bar the build's parity tests hold them to."""
    rng = np.random.default_rng(13)
    worst = 0.0
    for A in covariances(rng, 3000):
        ev, V = eig3(lib, A, 1)
        ev_o, V_o = oracle_mod.sym_eigen(A)
        nrm = np.abs(A).max()
        assert np.abs(np.sort(ev) - np.sort(ev_o)).max() <= 1e-14 * nrm
        if ev_o.max() <= 0.0:
            continue
        C, Ci = regularise(ev, V)
        Co, Cio = regularise(ev_o, V_o)
        rel = max(np.abs(C - Co).max() / np.abs(Co).max(), np.abs(Ci - Cio).max() / np.abs(Cio).max())
        worst = max(worst, rel)
    assert worst <= 1e-10
