"""Builds the scalar CPU oracle (TEST INFRASTRUCTURE ONLY) into
oracle/liboracle_sgm.so with g++ -O2 (the SGM path in sgm_oracle.cpp, the
depth filters and the disparity-to-depth conversion in depth_oracle.cpp, the
remap and rectify maps in rectify_oracle.cpp, the sparse stereo matcher in
stereo_match_oracle.cpp).

The reference (MisterEkole/vx-rs) is Metal + Rust and cannot be compiled in
this environment (no Metal toolchain, no Rust toolchain), so there is no
oracle/_ref build; the oracle is the C++ restatement in sgm_oracle.cpp.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "liboracle_sgm.so")


def build():
    tmp = OUT + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-fPIC", "-shared", "-pthread",
           "-o", tmp, os.path.join(HERE, "sgm_oracle.cpp"), os.path.join(HERE, "depth_oracle.cpp"),
           os.path.join(HERE, "rectify_oracle.cpp"), os.path.join(HERE, "stereo_match_oracle.cpp")]
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build()
