"""The 63-bit Morton key is assembled from 32-bit words (the spread of a
21-bit coordinate as the spreads of its low 11 and high 10 bits;
bvhtest_amd/csrc/types.hpp expandBits21Halves / morton63) instead of 64-bit
shifts and masks.  The product's own helpers are compiled for the host and
compared with the 64-bit spread (expandBits21, the reference's expandBits of
lib/builder/LBVH_CUDA.cu:132-138 widened to 21 bits per axis) for every
21-bit coordinate.  The GPU parity tests pin the device build (63-bit keys
bit-exact against the oracle)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_morton63_halves_equal_the_64bit_spread(tmp_path):
    exe = tmp_path / "morton63_check"
    subprocess.run(["g++", "-std=c++17", "-O2", os.path.join(HERE, "morton63_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
