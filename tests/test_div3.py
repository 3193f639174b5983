"""The build's centroid divides by three with a 3-operation correctly
rounded method (bvhtest_amd/csrc/types.hpp div3) instead of a general f32
division.  Its result must equal IEEE x / 3.0f bit for bit (the reference's
centroid, lib/builder/LBVH_CUDA.cu:92).  The product's own helper is compiled
for the host with g++ (no FMA contraction; the fused operations are explicit)
and compared on a strided sample of all 2^32 bit patterns plus every pattern
near a power of two; BVHT_EXHAUSTIVE=1 checks all of them (about a minute).
The GPU parity tests pin the device build of the same code (Morton keys
bit-exact against the oracle)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_div3_equals_ieee_division(tmp_path):
    exe = tmp_path / "div3_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", os.path.join(HERE, "div3_check.cpp"), "-o",
                    str(exe)], check=True)
    stride = "1" if os.environ.get("BVHT_EXHAUSTIVE") == "1" else "101"
    r = subprocess.run([str(exe), stride], stdout=subprocess.PIPE, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
