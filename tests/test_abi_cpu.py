"""The C ABI of libabie_hip.so: it loads and exports every function that
include/abie_hip.h declares (no compute calls: no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "abie_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\([^;{]*\)\s*;", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while")))


def test_header_parses():
    names = declared_functions()
    for n in ("initialize_code", "integrator_gr", "calculate_accelerations", "abie_hip_accel_soa",
              "abie_ensemble_run", "abie_mlp_loss_grad"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from abie_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libabie_hip.so not built (run python -m abie_amd.build)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # every signature the Python binding declares is a real symbol of the header
    assert set(_lib.SIGNATURES) <= set(declared_functions())


def test_binding_fails_loudly_without_library(monkeypatch, tmp_path):
    from abie_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.AbieHipError):
        _lib.load()
