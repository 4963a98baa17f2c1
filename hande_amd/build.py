"""In-tree build of the engine library ``hande_amd/libfqmc.so`` for gfx950.

    python -m hande_amd.build

Device code is compiled with -ffp-contract=off (the replay mode reproduces the
reference's floating-point expression order, which has no fused multiply-add)
and with packed FP32 instructions disabled.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libfqmc.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FQMC_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include"),
          "-I" + CSRC, "-I" + os.path.join(CSRC, "host")]
HIP_FLAGS = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in ("host/systems.cpp", "host/driver.cpp"):
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        run(["g++", *COMMON, "-Wall", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in ("engine.hip", "mlp.hip"):
        obj = os.path.join(BUILD, src + ".o")
        run([HIPCC, *COMMON, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = OUT + ".tmp"
    run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs])
    shutil.move(tmp, OUT)
    return OUT


if __name__ == "__main__":
    sys.exit(0 if build() else 1)
