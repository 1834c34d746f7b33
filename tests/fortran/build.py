"""Build the Fortran callers with flang (test infrastructure):
  tests/fortran/psfun_driver.f90 -- flat calls through include/psfun_hip_mod.f90
  tests/fortran/psfun_typed.f90  -- the reference's example drivers against
                                    the typed layer psfun_amd/fortran (built
                                    first by psfun_amd/fortran/build.py)
Both link psfun_amd/libpsfun_hip.so.  Output: tests/fortran/_build/.
Usage: python tests/fortran/build.py"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build")
EXE = os.path.join(OUT, "psfun_driver")
TYPED = os.path.join(OUT, "psfun_typed")


def flang():
    for c in (os.environ.get("FLANG"), "/opt/rocm/llvm/bin/flang", shutil.which("flang"), shutil.which("amdflang")):
        if c and os.path.exists(c):
            return c
    return None


def build(verbose=True):
    fc = flang()
    if fc is None:
        if verbose:
            print("flang not found: Fortran driver not built")
        return None
    os.makedirs(OUT, exist_ok=True)
    mod_src = os.path.join(REPO, "include", "psfun_hip_mod.f90")
    mod_obj = os.path.join(OUT, "psfun_hip_mod.o")
    steps = [
        [fc, "-O2", "-module-dir", OUT, "-c", mod_src, "-o", mod_obj],
        [fc, "-O2", "-I", OUT, "-module-dir", OUT, os.path.join(HERE, "psfun_driver.f90"), mod_obj, "-o", EXE,
         "-L", os.path.join(REPO, "psfun_amd"), "-lpsfun_hip", "-Wl,-rpath,$ORIGIN/../../../psfun_amd"],
    ]
    sys.path.insert(0, REPO)
    from psfun_amd.fortran import build as fb
    flib = fb.build(verbose=verbose)
    steps.append([fc, "-O2", "-I", fb.OUT, "-module-dir", OUT, os.path.join(HERE, "psfun_typed.f90"), flib, "-o", TYPED,
                  "-L", os.path.join(REPO, "psfun_amd"), "-lpsfun_hip", "-Wl,-rpath,$ORIGIN/../../../psfun_amd"])
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"flang failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {EXE} and {TYPED}")
    return EXE


if __name__ == "__main__":
    sys.exit(0 if build() else 1)
