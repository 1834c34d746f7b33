"""Build the CPU oracle (oracle/libpsfun_oracle.so) with gcc + OpenMP.

Test infrastructure only: the oracle is the checker for the HIP library and
the CPU baseline of bench.py; the product never loads it.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "psfun_oracle.c")
OUT = os.path.join(HERE, "libpsfun_oracle.so")


def build(verbose=True):
    tmp = OUT + ".tmp"
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-Wall", "-Wno-unused-function",
           SRC, "-o", tmp, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
