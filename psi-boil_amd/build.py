"""Build libpsb.so (hipcc, gfx950) in place.

    python psi-boil_amd/build.py

Flags: -ffp-contract=off keeps the FP64 per-cell arithmetic rounding like the
reference's g++ build (bit-exact SpMV / residual); packed FP32 instructions
are disabled for the device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpsb.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
SOURCES = ["psb_api.hip", "psb_batch.hip", "psb_gen.cpp", "psb_io.cpp"]


def build(verbose=False):
    cmd = [HIPCC, *FLAGS, "-o", OUT + ".tmp"] + [os.path.join(CSRC, s) for s in SOURCES] + \
          ["-lpthread"]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
