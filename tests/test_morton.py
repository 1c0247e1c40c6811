"""The voxel build's compact Morton key (csrc/morton.hpp: magic-number bit
spreading in up to three phases), built for the host from the same source,
equals the key's definition for every per-axis width combination up to 21
bits, and so does its 32-bit form for every combination of at most 32 bits."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_spread_morton_equals_definition(tmp_path):
    exe = tmp_path / "morton_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-D__host__=", "-D__device__=",
                    os.path.join(HERE, "native", "morton_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 mismatches" in r.stdout
