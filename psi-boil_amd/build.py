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

FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
SOURCES = ["psb_api.hip", "psb_batch.hip", "psb_gen.cpp", "psb_io.cpp"]
OBJDIR = os.path.join(HERE, "_build")


def build(verbose=False):
    """Compiles the sources to objects in parallel, then links the library."""
    os.makedirs(OBJDIR, exist_ok=True)
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *FLAGS, "-c", "-o", obj, os.path.join(CSRC, src)])
    procs = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd))
        procs.append(subprocess.Popen(cmd))
    bad = [c for c, p in zip(cmds, procs) if p.wait() != 0]
    if bad:
        raise subprocess.CalledProcessError(1, bad[0])
    link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUT + ".tmp", *objs,
            "-lpthread"]
    if verbose:
        print(" ".join(link))
    subprocess.check_call(link)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
