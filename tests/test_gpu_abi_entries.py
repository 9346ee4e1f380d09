"""The remaining reference entry points of the path, called directly through
the C ABI: integrator_gauss_radau15 on caller arrays
(src/integrator_gauss_radau15.c:221), ode_n_body_second_order_gpu with the
gpu_init / gpu_finalize lifetime (src/gpuforce.cu:88-170), and the error
codes that replace the reference's exit() calls."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = 4 * np.pi ** 2


def same_bits(a, b):
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    return bool(np.all(a.view(np.int64) == b.view(np.int64)))


@pytest.fixture()
def lib():
    from abie_amd.clibabie import CLibABIE
    c = CLibABIE()
    c.lib.finalize_code()
    yield c
    c.lib.abie_hip_set_force_mode(0)
    c.lib.finalize_code()


def test_integrator_gauss_radau15_on_caller_arrays(lib, golden):
    """One 1-yr chunk of the reference's golden run, by the direct call:
    positions, velocities and model time bit for bit (exact forces)."""
    d = golden("gr15")
    m = d["small_m"].copy()
    pos = d["small_pos0"].reshape(-1).copy()
    vel = d["small_vel0"].reshape(-1).copy()
    lib.initialize_code(G, 0.0, len(m), 1, 1, 0.0, deviceID=0)
    lib.lib.abie_hip_set_force_mode(1)
    ret = lib.integrator_gauss_radau15(pos, vel, m, len(m), G, 0.0, 1.0, 1.0)
    assert ret == d["small_ret"][0]
    assert lib.get_model_time() == d["small_t"][0]
    assert same_bits(pos, d["small_pos"][0]) and same_bits(vel, d["small_vel"][0])


def test_ode_n_body_second_order_gpu_lifetime(lib, golden):
    """gpu_init brings the device up without initialize_code; the force
    entry point then matches the reference bit for bit in exact mode;
    after gpu_finalize it reports an error instead of exiting."""
    from abie_amd import _lib
    d = golden("forces")
    pos, m = d["c0_pos"].reshape(-1), np.ascontiguousarray(d["c0_m"])
    Gc = float(d["G"])
    lib.lib.gpu_init(len(m), 0)
    lib.lib.abie_hip_set_force_mode(1)
    acc = np.zeros(pos.size)
    call = lambda: ctypes_call(lib, pos, m, Gc, acc)
    rc = call()
    assert rc == 0
    assert same_bits(acc.reshape(-1, 3), d["c0_acc"])
    lib.lib.gpu_finalize()
    rc = call()
    assert rc != 0 and "initialize_code" in _lib.load().abie_hip_last_error().decode()


def ctypes_call(lib, pos, m, Gc, acc):
    import ctypes
    from abie_amd import _lib
    return ctypes.c_long(lib.lib.ode_n_body_second_order_gpu(
        np.ascontiguousarray(pos).ctypes.data_as(_lib.c_dp), len(m), Gc, m.ctypes.data_as(_lib.c_dp),
        np.zeros(len(m)).ctypes.data_as(_lib.c_dp), acc.ctypes.data_as(_lib.c_dp))).value


def test_invalid_device_and_unsupported_calls(lib):
    """Where the reference falls back to the CPU or exits, the library
    returns an error code and names it."""
    from abie_amd import _lib
    L = _lib.load()
    assert L.initialize_code(G, 0.0, 8, 1, 1, -1) == -2  # ABIE_HIP_ENODEV: no CPU fallback
    assert L.initialize_code(G, 0.0, 8, 1, 1, 99) == -2
    L.finalize_code()
    assert L.initialize_code(G, 0.0, 8, 1, 1, 0) == 0
    assert L.integrator_rk(0.0, 1.0, 0.1) < 0 and L.integrator_wh(0.0, 1.0, 0.1) < 0
    assert b"outside" in L.abie_hip_last_error()


def test_integrator_gr_refuses_mixed_forces(lib):
    """Mixed forces are built for the massive-only controllers only: the
    reference's controller (every body in its error norms) refuses them with
    ABIE_HIP_EUNSUP before touching the state, and runs again once the mode
    is back to fp64; calculate_accelerations keeps honouring the mode."""
    from abie_amd import _lib, ics
    pos, vel, m = ics.config0(n_test=40)
    lib.initialize_code(G, 0.0, len(m), 1, 1, 0.0, deviceID=0)
    p, v = pos.reshape(-1).copy(), vel.reshape(-1).copy()
    lib.set_state(p, v, m, np.zeros_like(m), len(m), G, 0.0)
    assert lib.lib.abie_hip_set_force_mode(2) == 0
    assert lib.lib.integrator_gr(0.0, 0.01, 1.0) == -5  # ABIE_HIP_EUNSUP
    assert b"mixed" in _lib.load().abie_hip_last_error()
    assert lib.get_model_time() == 0.0
    acc = lib.calculate_accelerations(p, v, len(m), G, m)
    assert np.all(np.isfinite(acc))
    assert lib.lib.abie_hip_set_force_mode(0) == 0
    assert lib.lib.integrator_gr(0.0, 0.01, 1.0) == 0
    assert lib.get_model_time() > 0.01


def test_issue_rate_measurement():
    """abie_hip_issue_rate at 8 waves per SIMD: a wave64 v_fma_f64 holds a
    SIMD 4 cycles and v_rsq_f64 (quarter rate) 16, so FMA-only chains run at
    4-4.5 cycles per instruction, rsq-only at 16-17, and one rsq per 16 FMAs
    at (16 x 4 + 16) / 17 = 4.7 (+10 %); clocks lie in (0.5, 2.6) GHz; bad
    arguments are refused."""
    import ctypes
    from abie_amd import _lib, engine
    r = {k: engine.issue_rate(k, 5.0) for k in ("fma", "rsq", "mix")}
    print(r)
    for ghz, _ in r.values():
        assert 0.5 < ghz < 2.6
    assert 4.0 * 0.98 <= r["fma"][1] <= 4.5
    assert 16.0 * 0.98 <= r["rsq"][1] <= 17.0
    assert 80.0 / 17 * 0.98 <= r["mix"][1] <= 80.0 / 17 * 1.1
    L = _lib.load()
    a, b = ctypes.c_double(), ctypes.c_double()
    assert L.abie_hip_issue_rate(0, 0.0, ctypes.byref(a), ctypes.byref(b)) == -1  # ABIE_HIP_EINVAL
    assert L.abie_hip_issue_rate(3, 5.0, ctypes.byref(a), ctypes.byref(b)) == -1
    assert L.abie_hip_issue_rate(0, 5.0, None, ctypes.byref(b)) == -1


def test_time_accel_soa_reads_the_kernel_clock():
    """time_accel_soa(clock=True): the clocked instance of the scalar-source
    kernel writes the same bits as the product kernel and reports a shader
    clock within (0.5, 2.6) GHz; plans it does not cover report 0."""
    import torch
    from abie_amd import engine, ics
    pos, vel, m = ics.config1a(n_test=1 << 16)
    src, sinks, n_self, _ = engine.soa_from_bodies(pos, m, G)
    ref = engine.accel_soa(src, sinks, n_self, "fp64")
    out = tuple(torch.full_like(ref[0], float("nan")) for _ in range(3))
    ms, ghz = engine.time_accel_soa(src, sinks, n_self, out, mode="fp64", reps=20, clock=True)
    torch.cuda.synchronize()
    print("64 + 65,536: %.2f us per launch at %.3f GHz" % (ms * 1e3, ghz))
    assert 0.5 < ghz < 2.6
    for a, b in zip(out, ref):
        assert torch.equal(a, b)
    pp, _, pm = ics.plummer(4096)  # 4,096 sources: the LDS-tiled kernel
    src, sinks, n_self, _ = engine.soa_from_bodies(pp, pm, 1.0)
    out = engine.accel_soa(src, sinks, n_self, "fp64")
    _, ghz = engine.time_accel_soa(src, sinks, n_self, out, mode="fp64", reps=5, clock=True)
    assert ghz == 0.0
