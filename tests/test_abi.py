"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

from conftest import REPO


def declared_symbols():
    names = []
    inc = os.path.join(REPO, "include")
    for f in sorted(os.listdir(inc)):
        if f.endswith(".h"):
            txt = open(os.path.join(inc, f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names += re.findall(r"\b(psfun_\w+)\s*\(", txt)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("psfun_ctx_create", "psfun_spmv", "psfun_dot", "psfun_nrm2", "psfun_axpby",
                 "psfun_lanczos_apply", "psfun_strerror"):
        assert must in names


def test_library_exports_all_declared_symbols():
    from psfun_amd import _lib
    L = _lib.lib()
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from psfun_amd import _lib
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_strerror_without_device():
    from psfun_amd import _lib
    L = _lib.lib()
    assert L.psfun_strerror(0) == b"success"
    assert b"istop" in L.psfun_strerror(_lib.PSFUN_ERR_INVALID_ISTOP)


def test_ctx_create_fails_cleanly_without_device():
    import torch
    if torch.cuda.is_available():
        return
    from psfun_amd import _lib
    h = ctypes.c_void_p()
    st = _lib.lib().psfun_ctx_create(0, ctypes.byref(h))
    assert st == _lib.PSFUN_ERR_NO_DEVICE


def test_fortran_module_binds_declared_symbols():
    """Every bind(C, name=...) of the Fortran interface module names a symbol
    of the C header, and the constants it defines match the header's."""
    src = open(os.path.join(REPO, "include", "psfun_hip_mod.f90")).read()
    binds = re.findall(r"bind\(C,\s*name='(\w+)'\)", src)
    assert len(binds) >= 20
    assert set(binds) <= set(declared_symbols())
    from psfun_amd import _lib
    consts = dict((k, int(v)) for k, v in re.findall(r"\b(PSFUN_[A-Z_0-9]+)\s*=\s*(\d+)", src))
    for k, v in consts.items():
        if hasattr(_lib, k):
            assert getattr(_lib, k) == v, k
    hdr = re.sub(r"/\*.*?\*/", "", open(os.path.join(REPO, "include", "psfun_hip.h")).read(), flags=re.S)
    hconst = dict((k, int(v)) for k, v in re.findall(r"\b(PSFUN_[A-Z_0-9]+)\s*=\s*(\d+)", hdr))
    for k, v in consts.items():
        assert hconst.get(k) == v, k
