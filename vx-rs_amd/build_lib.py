"""Builds the gfx950 SGM library (vxsgm/libvxsgm.so) in-tree with hipcc.

Run as a script (``python vx-rs_amd/build_lib.py``) or call ``build()``.
Every HIP source is compiled for gfx950 only, with packed-fp32 instruction
selection disabled and IEEE fp32 division (no fast-math), then linked into one
shared library that exports the C ABI declared in include/vxh/*.h.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "vxsgm", "libvxsgm.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["sgm_paths_c0.hip", "sgm_paths_c1.hip", "sgm_paths_c2.hip", "sgm_sweep_c0.hip",
               "sgm_sweep_c1.hip", "sgm_sweep_c2.hip", "sgm_kernels.hip",
               "sgm_runtime.hip", "sgm_batch.hip", "sgm_server.cpp", "depth.hip", "pipeline.hip",
               "rectify.hip", "datasets.cpp", "stereo_match.hip"]
CXX_SOURCES = ["synth.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-Wall"]
HIP_FLAGS = ["--offload-arch=gfx950", "-ffp-contract=off",
             "-fhip-fp32-correctly-rounded-divide-sqrt",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True, jobs=None, out=OUT, defines=(), objdir=OBJ, csrc=CSRC):
    """Compiles every source (from `csrc`) and links `out`; `defines`
    (NAME=VALUE strings) select compile-time variants of the kernels, built in
    their own `objdir`."""
    os.makedirs(objdir, exist_ok=True)
    cmds, objs = [], []
    dflags = ["-D" + d for d in defines]
    for src in HIP_SOURCES:
        o = os.path.join(objdir, src + ".o")
        cmds.append([HIPCC, "-x", "hip", *HIP_FLAGS, *COMMON, "-I", csrc, *dflags, "-c",
                     os.path.join(csrc, src), "-o", o])
        objs.append(o)
    for src in CXX_SOURCES:
        o = os.path.join(objdir, src + ".o")
        cmds.append(["g++", *COMMON, "-I", csrc, "-c", os.path.join(csrc, src), "-o", o])
        objs.append(o)
    # Independent translation units compile side by side (at most 8 at once:
    # each hipcc instance holds a few GB while optimising the kernels).
    jobs = jobs or min(8, os.cpu_count() or 1, len(cmds))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, c) for c in cmds]:
            f.result()
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lz"])
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
