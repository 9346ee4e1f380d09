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


def test_step_root7_within_two_ulp():
    """The seventh root that the massive-only step controllers take of
    estim / epsb (abie_hip_step_root7, the device code evaluated on the
    host): within 2 ulp of the exact root (mpmath, 40 digits) over the whole
    double range, exact on 2^(7k) and 1, and 0, +inf, NaN map to themselves."""
    import math
    import mpmath as mp
    import numpy as np
    from abie_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libabie_hip.so not built (run python -m abie_amd.build)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    f = lib.abie_hip_step_root7
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_double]
    mp.mp.dps = 40
    rng = np.random.default_rng(7)
    xs = list(10.0 ** rng.uniform(-307, 307, 400)) + list(10.0 ** rng.uniform(-12, 4, 400))
    xs += [5e-324, 2.2250738585072014e-308, 1.7976931348623157e308, 1e-4, 3.0, 128.0 - 2 ** -45]
    worst = 0.0
    for x in xs:
        y = f(x)
        exact = mp.root(mp.mpf(x), 7)
        worst = max(worst, float(abs(mp.mpf(y) - exact) / mp.mpf(math.ulp(float(exact)))))
    assert worst <= 2.0, worst
    for k in range(-40, 41):
        assert f(2.0 ** (7 * k)) == 2.0 ** k
    assert f(1.0) == 1.0 and f(0.0) == 0.0 and f(math.inf) == math.inf and math.isnan(f(math.nan))
