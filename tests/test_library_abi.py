"""The C-ABI library loads and exports every entry point include/tdyhip.h
declares (no GPU needed: nothing is called)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(tdyhip_[a-z0-9_]+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    from tdycore_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libtdyhip.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(_lib.LIB_PATH)
    declared = _declared()
    assert len(declared) >= 17
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing
    assert set(_lib.EXPORTS) <= declared


def test_loader_raises_without_library(tmp_path, monkeypatch):
    from tdycore_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.TdyHipError):
        _lib.load()


def test_create_rejects_bad_mesh_without_touching_the_gpu():
    """Argument checks run on the host before any device work."""
    from tdycore_amd import _lib
    L = _lib.load()
    h = ctypes.c_void_p()
    ms = _lib.Mesh()
    mt = _lib.Materials()
    rc = L.tdyhip_create(ctypes.byref(ms), ctypes.byref(mt), ctypes.byref(h))
    assert rc == 1
    assert b"empty mesh" in L.tdyhip_last_error()


def test_petsc_binding_calls_only_declared_entry_points():
    """The reference-side PETSc binding (tdycore_amd/petsc/tdympfao_hip.c, not
    compiled here: PETSc is absent) calls only entry points include/tdyhip.h
    declares, and defines the five callbacks with the reference's names."""
    src = open(os.path.join(ROOT, "tdycore_amd", "petsc", "tdympfao_hip.c")).read()
    code = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    called = set(re.findall(r"\b(tdyhip_[a-z0-9_]+)\s*\(", code))
    assert called and called <= _declared(), called - _declared()
    for cb in ("TDyMPFAOIFunction_HIP(TS ts, PetscReal t, Vec U, Vec U_t, Vec R, void *ctx)",
               "TDyMPFAOIJacobian_HIP(TS ts, PetscReal t, Vec U, Vec U_t, PetscReal shift, Mat A, Mat B, void *ctx)",
               "TDyMPFAOSNESFunction_HIP(SNES snes, Vec U, Vec R, void *ctx)",
               "TDyMPFAOSNESJacobian_HIP(SNES snes, Vec U, Mat A, Mat B, void *ctx)",
               "TDyMPFAOSNESPreSolve_HIP(TDy tdy)"):
        assert "PetscErrorCode " + cb + " {" in code, cb
