"""abie_hip_accel_soa_noalloc (the allocation-free force entry point) against
abie_hip_accel_soa through the C ABI: the same kernels, so the same bits,
on two calls in a row whose second is smaller (the library's scratch keeps
the first call's partial sums), on shapes that split the source range and
so use the scratch, and on one that does not."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = 4 * np.pi ** 2


def _call(name, src, sinks, n_self, mode):
    import ctypes
    import torch
    from abie_amd import _lib, engine
    sx, sy, sz, gm = src
    px, py, pz = sinks
    out = tuple(torch.full((px.numel(),), float("nan"), dtype=torch.float64, device="cuda") for _ in range(3))
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = getattr(_lib.load(), name)(p(sx), p(sy), p(sz), p(gm), sx.numel(), p(px), p(py), p(pz), px.numel(),
                                    int(n_self), p(out[0]), p(out[1]), p(out[2]), engine.FORCE_MODES[mode], stream)
    _lib.check(rc, name)
    torch.cuda.synchronize()
    return torch.stack(out, 1)


def _case(kind, n, seed):
    from abie_amd import engine, ics
    if kind == "plummer":
        pos, vel, m = ics.plummer(n, seed=seed)
        return engine.soa_from_bodies(pos, m, 1.0)[:3]
    nm, nt = n
    pos, vel, m = ics.planetary_system(seed, nm, nt)
    return engine.soa_from_bodies(pos, m, G)[:3]


@pytest.mark.parametrize("mode,first,second", [
    ("fp64", ("plummer", 4096, 1), ("plummer", 2500, 2)),     # 4 then 2 source parts
    ("mixed", ("planet", (2048, 6000), 3), ("planet", (64, 50000), 4)),  # split, then no scratch
    ("exact", ("planet", (300, 20000), 5), ("plummer", 2200, 6)),
])
def test_noalloc_matches_allocating_entry_point(mode, first, second):
    import torch
    for kind, n, seed in (first, second):
        src, sinks, n_self = _case(kind, n, seed)
        a = _call("abie_hip_accel_soa", src, sinks, n_self, mode)
        b = _call("abie_hip_accel_soa_noalloc", src, sinks, n_self, mode)
        assert torch.isfinite(a).all()
        assert torch.equal(a, b), (mode, kind, n)
