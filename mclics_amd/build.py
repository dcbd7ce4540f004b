"""Builds the native parts in-tree: libmclgpu (HIP, gfx950), the mclics host program and the CPU oracle.

    python -m mclics_amd.build            # everything
    python -m mclics_amd.build lib        # only libmclgpu

Outputs: mclics_amd/lib/libmclgpu.so, mclics_amd/bin/mclics, oracle/build/libmcloracle.so.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MCL_OFFLOAD_ARCH", "gfx950")
# -packed-fp32-ops: no packed FP32 instructions in the generated code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

LIB = os.path.join(HERE, "lib", "libmclgpu.so")
BIN = os.path.join(HERE, "bin", "mclics")
ORACLE = os.path.join(ROOT, "oracle", "build", "libmcloracle.so")


def _run(cmd, cwd=ROOT):
    print("+ " + " ".join(cmd), flush=True)
    subprocess.run(cmd, cwd=cwd, check=True)


def build_lib():
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    srcs = [os.path.join(HERE, "csrc", f) for f in ("mcl_engine.hip", "mcl_mlp.hip")]
    tmp = LIB + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-shared", "-o", tmp] + srcs)
    os.replace(tmp, LIB)


def build_host():
    srcs = sorted(glob.glob(os.path.join(HERE, "host", "*.cpp")))
    if not srcs:
        return
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    tmp = BIN + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-fopenmp", "-I", os.path.join(ROOT, "include"), "-I",
          os.path.join(HERE, "csrc"), "-o", tmp] + srcs +
         ["-L", os.path.dirname(LIB), "-lmclgpu", "-Wl,-rpath,$ORIGIN/../lib", "-pthread"])
    os.replace(tmp, BIN)


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-B"])


def build_all():
    build_lib()
    build_host()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "lib": build_lib, "host": build_host, "oracle": build_oracle}[w]()
