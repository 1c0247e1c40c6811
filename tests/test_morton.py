"""The voxel build's compact Morton key (csrc/morton.hpp: magic-number bit
spreading in up to three phases), built for the host from the same source,
equals the key's definition for every per-axis width combination up to 21
bits, and so does its 32-bit form for every combination of at most 32 bits;
the per-axis keys it interleaves may be formed by multiplication at a
power-of-two resolution."""
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def test_spread_morton_equals_definition(tmp_path):
    exe = tmp_path / "morton_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-D__host__=", "-D__device__=",
                    os.path.join(HERE, "native", "morton_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 mismatches" in r.stdout


def test_power_of_two_resolution_keys_multiply_exactly():
    """For a power-of-two resolution the build forms voxel keys as
    floor(p * (1 / res)) instead of floor(p / res): the two are the same
    IEEE f32 value for every p (scaling by 2^-k rounds the one real value
    either way), checked over random bit patterns (subnormals, infinities and
    NaNs included) and edge values."""
    rng = np.random.default_rng(11)
    bits = rng.integers(0, 1 << 32, size=4_000_000, dtype=np.uint64).astype(np.uint32)
    p = bits.view(np.float32)
    edge = np.array([0.0, -0.0, 1e-45, -1e-45, 1.1754942e-38, 3.4028235e38, -3.4028235e38, np.inf, -np.inf,
                     np.nan, 0.5, 1.0, 2.0, 3.9999998, 4.0, -2.0000002, 1e7, -1e7], dtype=np.float32)
    p = np.concatenate([p, edge])
    with np.errstate(all="ignore"):
        for k in range(-8, 9):
            res = np.float32(2.0 ** k)
            inv = np.float32(1.0) / res
            assert inv * res == np.float32(1.0)
            q_div = np.floor(p / res)
            q_mul = np.floor(p * inv)
            assert np.array_equal(q_div, q_mul, equal_nan=True), k
    # and a resolution that is not a power of two does differ somewhere
    res = np.float32(0.3)
    with np.errstate(all="ignore"):
        assert not np.array_equal(np.floor(p / res), np.floor(p * (np.float32(1.0) / res)), equal_nan=True)
