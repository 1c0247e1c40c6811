"""The Newton solve's 6x6 eigendecomposition (csrc/eigen6.hpp: Householder
tridiagonalisation + implicit QL), built for the host from the same source
the device compiles, against LAPACK (numpy.linalg.eigvalsh), the defining
identities, closed forms and the oracle's clamped Newton step
(optimization/newton.rs:119-176)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
P = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("eigen6") / "libeigen6.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-D__host__=",
                    "-D__device__=", os.path.join(HERE, "native", "eigen6_capi.cpp"), "-o", str(out)], check=True)
    L = ctypes.CDLL(str(out))
    L.refine_rsqrt_from.restype = L.refine_rcp_from.restype = ctypes.c_double
    L.refine_rsqrt_from.argtypes = L.refine_rcp_from.argtypes = [ctypes.c_double, ctypes.c_double]
    return L


def eig(lib, H, b, mode=1):
    H = np.ascontiguousarray(np.tril(H), np.float64)
    b = np.ascontiguousarray(b, np.float64)
    lam, Q, y = np.zeros(6), np.zeros(36), np.zeros(6)
    it = lib.eigen6_full(H.ctypes.data_as(P), b.ctypes.data_as(P), lam.ctypes.data_as(P), Q.ctypes.data_as(P),
                         y.ctypes.data_as(P), mode)
    return it, lam, Q.reshape(6, 6), y


def clamped_step(lam, Q, y):
    """newton.rs:119-176 in the eigenbasis, as newton_step_wg forms it."""
    lc = lam.copy()
    if not np.all(lam < 0):
        lc = np.where(lc > -1e-3, -1e-3, lc)
    if np.abs(lc).max() < 1e-10:
        return None
    return Q @ np.where(np.abs(lc) > 1e-10, y / lc, 0.0)


def spectra(rng, count):
    for t in range(count):
        Qr, _ = np.linalg.qr(rng.normal(size=(6, 6)))
        kind = t % 4
        if kind == 0:
            A = rng.normal(size=(6, 6))
            yield A + A.T
            continue
        if kind == 1:  # graded, mostly negative: NDT Hessians
            ev = -10 ** rng.uniform(-3, 5, 6) * rng.choice([1, 1, 1, -1], 6)
        elif kind == 2:  # clusters and a zero eigenvalue
            ev = np.array([-1e4, -1e4 * (1 + 1e-12), 3.0, 3.0, -1e-3, 0.0])
        else:  # translation and rotation blocks of different scale
            A = rng.normal(size=(6, 6))
            S = np.diag([1e2, 1e2, 1e2, 1e3, 1e3, 1e3])
            yield S @ (A + A.T) @ S
            continue
        H = Qr @ np.diag(ev) @ Qr.T
        yield 0.5 * (H + H.T)


def test_refinement_steps(lib):
    """One third-order step from an estimate with relative error up to 2^-22
    (the hardware's estimates are within 2^-24) lands within 2 ulps of the
    exact value."""
    rng = np.random.default_rng(5)
    for x in np.concatenate([10 ** rng.uniform(-30, 30, 2000), [1.0, 2.0, 0.25, 1e-300, 1e300]]):
        for rel in (2.0 ** -22, -(2.0 ** -22), 0.0):
            exact = 1.0 / np.sqrt(x)
            y = lib.refine_rsqrt_from(x, exact * (1 + rel))
            assert abs(y - exact) <= 2 * np.spacing(exact)
            exact = 1.0 / x
            y = lib.refine_rcp_from(x, exact * (1 + rel))
            assert abs(y - exact) <= 2 * np.spacing(exact)


@pytest.mark.parametrize("mode", [0, 1])
def test_decomposition_identities(lib, mode):
    rng = np.random.default_rng(11)
    for H in spectra(rng, 4000):
        b = rng.normal(size=6)
        it, lam, Q, y = eig(lib, H, b, mode)
        assert it >= 0
        nrm = np.abs(H).max()
        assert np.abs(H @ Q - Q * lam).max() <= 1e-13 * nrm
        assert np.abs(Q.T @ Q - np.eye(6)).max() <= 1e-13
        assert np.abs(Q.T @ b - y).max() <= 1e-13 * np.abs(b).max()
        assert np.abs(np.sort(lam) - np.linalg.eigvalsh(H)).max() <= 1e-13 * nrm


def test_closed_forms(lib):
    # diagonal: no reflection, no QL step, eigenvalues exact, Q = I
    D = np.diag([-3.0, 5.0, -1e-4, 2e6, 0.0, -7.5])
    it, lam, Q, y = eig(lib, D, np.arange(6.0))
    assert it == 0 and np.array_equal(lam, np.diag(D)) and np.array_equal(Q, np.eye(6))
    assert np.array_equal(y, np.arange(6.0))
    # 2x2 block [[a, c], [c, a]]: eigenvalues a +- c, vectors (1, +-1)/sqrt 2
    H = np.diag([4.0, 4.0, -1.0, -2.0, -3.0, -4.0])
    H[0, 1] = H[1, 0] = 1.5
    it, lam, Q, _ = eig(lib, H, np.zeros(6))
    assert np.allclose(np.sort(lam), [-4, -3, -2, -1, 2.5, 5.5], rtol=0, atol=1e-14)
    k = int(np.argmax(lam))
    assert np.allclose(np.abs(Q[:2, k]), [2 ** -0.5, 2 ** -0.5], atol=1e-15)
    # rank one: u u^T has eigenvalue |u|^2 once and zero five times
    u = np.array([1.0, -2.0, 0.5, 3.0, 0.0, 1.0])
    _, lam, _, _ = eig(lib, np.outer(u, u), np.zeros(6))
    assert abs(lam.max() - u @ u) <= 1e-14 * (u @ u)
    assert np.abs(np.sort(lam)[:5]).max() <= 1e-14 * (u @ u)


def test_step_matches_oracle(lib, oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(3)
    n = 0
    for t in range(3000):
        Qr, _ = np.linalg.qr(rng.normal(size=(6, 6)))
        # eigenvalues at least 0.1 from zero and from the clamp boundary
        # -1e-3; the step's relative error is bounded by a few ulps times
        # the condition kappa = max |lambda| / min |lambda'| for either method
        ev = -10 ** rng.uniform(-1, 5, 6) * (rng.choice([1, 1, -1], 6) if t % 2 else 1)
        H = Qr @ np.diag(ev) @ Qr.T
        H = 0.5 * (H + H.T)
        g = rng.normal(size=6) * 10 ** rng.uniform(-2, 3)
        _, lam, Q, y = eig(lib, H, -g)
        d = clamped_step(lam, Q, y)
        ref = O.newton_step(g, H)
        assert (d is None) == (ref is None)
        if d is not None:
            lc = np.where(lam > -1e-3, -1e-3, lam) if not np.all(lam < 0) else lam
            kappa = np.abs(lc).max() / np.abs(lc).min()
            assert np.abs(d - ref).max() <= 100 * 2.2e-16 * kappa * np.abs(ref).max()
            n += 1
    assert n == 3000
