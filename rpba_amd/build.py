"""In-tree build of the HIP library (librpba_hip.so, gfx950) and the `rpba`
host driver.  Compiles every time it is called."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librpba_hip.so")
BIN = os.path.join(HERE, "bin", "rpba")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no FMA contraction, so the FP64 path follows the
# operation order of the reference formulas; packed FP32 instructions off.
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-result"]
SOURCES = ["kernels.hip", "noalloc.hip", "dense.hip", "lp.hip", "structure_dev.hip", "partition_kernels.hip", "solver.cpp", "partition.cpp", "batch.cpp", "structure.cpp", "bal_io.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return r


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(BIN), exist_ok=True)

    def comp(src):
        out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        _run([HIPCC, *HIPFLAGS, "-c", "-o", out, os.path.join(CSRC, src)])
        return out

    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(comp, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    _run([HIPCC, "-O2", "-std=c++17", "-o", BIN, os.path.join(CSRC, "rpba_main.cpp"),
          "-L" + HERE, "-lrpba_hip", "-Wl,-rpath,$ORIGIN/.."])
    if verbose:
        print("built", LIB, BIN)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
