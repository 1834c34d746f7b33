"""Build the Fortran drop-in layer: the PSBLAS subset (psb_base_mod,
psb_prec_mod, psb_util_mod) and PSFUN's typed modules (psfun_d_serial_mod,
psfun_d_krylov_mod, psfun_krylov_mod, psfun_z_quadrature_mod,
psfun_utils_mod) over include/psfun_hip_mod.f90, with flang, into
psfun_amd/fortran/_build/libpsfun_f.a and the .mod files a Fortran caller
compiles against (-I psfun_amd/fortran/_build).  The archive calls
libpsfun_hip.so, which the caller links too.
Usage: python -m psfun_amd.fortran.build"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libpsfun_f.a")

# compilation order follows the module dependencies
SOURCES = [
    os.path.join(REPO, "include", "psfun_hip_mod.f90"),
    os.path.join(HERE, "psb_base_mod.f90"),
    os.path.join(HERE, "psb_prec_mod.f90"),
    os.path.join(HERE, "psb_util_mod.f90"),
    os.path.join(HERE, "psfun_utils_mod.f90"),
    os.path.join(HERE, "psfun_d_serial_mod.f90"),
    os.path.join(HERE, "psfun_d_krylov_mod.f90"),
    os.path.join(HERE, "psfun_krylov_mod.f90"),
    os.path.join(HERE, "psfun_z_quadrature_mod.f90"),
]


def flang():
    for c in (os.environ.get("FLANG"), "/opt/rocm/llvm/bin/flang", shutil.which("flang"), shutil.which("amdflang")):
        if c and os.path.exists(c):
            return c
    return None


def build(verbose=True):
    fc = flang()
    if fc is None:
        if verbose:
            print("flang not found: Fortran layer not built")
        return None
    os.makedirs(OUT, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(OUT, os.path.basename(src)[:-4] + ".o")
        cmd = [fc, "-O2", "-fPIC", "-module-dir", OUT, "-I", OUT, "-c", src, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"flang failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        objs.append(obj)
    tmp = LIB + ".tmp"
    if os.path.exists(tmp):
        os.remove(tmp)
    r = subprocess.run(["ar", "rcs", tmp] + objs, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"ar failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    try:
        sys.exit(0 if build() else 1)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
