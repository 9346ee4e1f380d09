"""Python wrappers of the device entry points of libabie_hip.so (part 2 of
include/abie_hip.h): the SoA force kernel, the batched GR15 ensemble engine
and the step-proposer MLP kernels.

Device memory for the force / MLP entry points comes from PyTorch tensors
(torch is used for allocation and streams only); the ensemble engine owns its
device memory and takes host numpy arrays.
"""
import ctypes

import numpy as np

from . import _lib

FORCE_MODES = {"fp64": _lib.FORCE_FP64, "exact": _lib.FORCE_FP64_EXACT, "mixed": _lib.FORCE_MIXED,
               "fp32": _lib.FORCE_FP32}  # fp32: force entry points only
CONTROLLERS = {"plain": _lib.CTRL_PLAIN, "massive": _lib.CTRL_MASSIVE, "learned": _lib.CTRL_LEARNED}


def _mode(m):
    return FORCE_MODES[m] if isinstance(m, str) else int(m)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


def _dp(a):
    return a.ctypes.data_as(_lib.c_dp)


# ---------------------------------------------------------------- forces
def accel_soa(src, sinks, n_self, mode="fp64", out=None, stream=None):
    """Accelerations of sinks from sources on the GPU.

    src = (sx, sy, sz, gm) float64 cuda tensors (gm = G*m, massive bodies),
    sinks = (px, py, pz); sinks [0, n_self) are the first n_self sources.
    Returns (ax, ay, az) (allocated unless ``out`` is given)."""
    import torch
    sx, sy, sz, gm = src
    px, py, pz = sinks
    for t in (sx, sy, sz, gm, px, py, pz):
        if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
            raise ValueError("accel_soa needs contiguous float64 cuda tensors")
    n_src, n_sink = sx.numel(), px.numel()
    if out is None:
        out = tuple(torch.empty(n_sink, dtype=torch.float64, device=px.device) for _ in range(3))
    ax, ay, az = out
    lib = _lib.load()
    _lib.check(lib.abie_hip_accel_soa(_ptr(sx), _ptr(sy), _ptr(sz), _ptr(gm), n_src, _ptr(px), _ptr(py), _ptr(pz),
                                      n_sink, int(n_self), _ptr(ax), _ptr(ay), _ptr(az), _mode(mode),
                                      _stream(stream)), "abie_hip_accel_soa")
    return out


def time_accel_soa(src, sinks, n_self, out, mode="fp64", reps=100, stream=None, each=False, clock=False):
    """Mean milliseconds per launch of ``reps`` back-to-back launches, timed
    with HIP events recorded on the launch stream.  each=True also returns
    the mean and the median duration of a launch bracketed by its own event
    pair: (back_to_back, each_mean, each_median).  clock=True appends the
    shader clock (GHz) the kernel ran at, read inside an instance of the same
    kernel (0.0 for plans other than the unsplit scalar-source kernel)."""
    sx, sy, sz, gm = src
    px, py, pz = sinks
    ax, ay, az = out
    ms = ctypes.c_double(0.0)
    ms1 = ctypes.c_double(0.0)
    med = ctypes.c_double(0.0)
    ghz = ctypes.c_double(0.0)
    lib = _lib.load()
    _lib.check(lib.abie_hip_time_accel_soa(_ptr(sx), _ptr(sy), _ptr(sz), _ptr(gm), sx.numel(), _ptr(px), _ptr(py),
                                           _ptr(pz), px.numel(), int(n_self), _ptr(ax), _ptr(ay), _ptr(az),
                                           _mode(mode), _stream(stream), int(reps), ctypes.byref(ms),
                                           ctypes.byref(ms1) if each else None,
                                           ctypes.byref(med) if each else None,
                                           ctypes.byref(ghz) if clock else None),
               "abie_hip_time_accel_soa")
    r = (ms.value, ms1.value, med.value) if each else (ms.value,)
    if clock:
        r = r + (ghz.value,)
    return r if len(r) > 1 else r[0]


ISSUE_KINDS = {"fma": 0, "rsq": 1, "mix": 2}


def issue_rate(kind="fma", milliseconds=20.0):
    """(shader GHz, SIMD cycles per wave instruction) of FP64 VALU work at 8
    waves per SIMD: kind "fma" (v_fma_f64), "rsq" (v_rsq_f64) or "mix" (one
    v_rsq_f64 per 16 v_fma_f64), measured by a kernel of about
    ``milliseconds``."""
    ghz = ctypes.c_double(0.0)
    cyc = ctypes.c_double(0.0)
    lib = _lib.load()
    _lib.check(lib.abie_hip_issue_rate(ISSUE_KINDS[kind], float(milliseconds), ctypes.byref(ghz),
                                       ctypes.byref(cyc)), "abie_hip_issue_rate")
    return ghz.value, cyc.value


def soa_from_bodies(pos, m, G, device="cuda"):
    """Split bodies into (sources, sinks, n_self) tensors: massive bodies
    (m > 0, ascending) first, then the others; returns the permutation too."""
    import torch
    m = np.asarray(m, dtype=np.float64)
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    massive = np.nonzero(m > 0)[0]
    rest = np.nonzero(~(m > 0))[0]
    perm = np.concatenate([massive, rest])
    p = pos[perm]
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=device)
    src = (t(p[: len(massive), 0]), t(p[: len(massive), 1]), t(p[: len(massive), 2]), t(G * m[massive]))
    sinks = (t(p[:, 0]), t(p[:, 1]), t(p[:, 2]))
    return src, sinks, len(massive), perm


# ---------------------------------------------------------------- ensemble
class Ensemble:
    """Independent systems integrated by the device GR15 engine.

    ``systems`` is a list of (pos (n,3), vel (n,3), m (n,)[, r (n,)])."""

    def __init__(self, systems, G=4 * np.pi ** 2, force_mode="fp64", controller="plain"):
        self.lib = _lib.load()
        sizes = [len(s[2]) for s in systems]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
        self.n_sys = len(systems)
        self.total = int(self.offsets[-1])
        pos = np.ascontiguousarray(np.concatenate([np.asarray(s[0], np.float64).reshape(-1, 3) for s in systems]))
        vel = np.ascontiguousarray(np.concatenate([np.asarray(s[1], np.float64).reshape(-1, 3) for s in systems]))
        m = np.ascontiguousarray(np.concatenate([np.asarray(s[2], np.float64) for s in systems]))
        r = np.ascontiguousarray(np.concatenate([np.asarray(s[3], np.float64) if len(s) > 3 else np.zeros(len(s[2]))
                                                 for s in systems]))
        self.masses = m
        handle = ctypes.c_void_p()
        _lib.check(self.lib.abie_ensemble_create(ctypes.byref(handle), self.n_sys,
                                                 self.offsets.ctypes.data_as(_lib.c_ip), _dp(pos), _dp(vel), _dp(m),
                                                 _dp(r), float(G), _mode(force_mode),
                                                 CONTROLLERS[controller] if isinstance(controller, str)
                                                 else int(controller)), "abie_ensemble_create")
        self.h = handle

    def close(self):
        if self.h:
            self.lib.abie_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, t_end, max_seconds=0.0):
        done = ctypes.c_int(0)
        _lib.check(self.lib.abie_ensemble_run(self.h, float(t_end), float(max_seconds), ctypes.byref(done)),
                   "abie_ensemble_run")
        return bool(done.value)

    def state(self):
        pos = np.zeros(3 * self.total)
        vel = np.zeros(3 * self.total)
        t = np.zeros(self.n_sys)
        st = np.zeros(self.n_sys, dtype=np.int32)
        _lib.check(self.lib.abie_ensemble_get_state(self.h, _dp(pos), _dp(vel), _dp(t),
                                                    st.ctypes.data_as(_lib.c_ip)), "abie_ensemble_get_state")
        return pos.reshape(-1, 3), vel.reshape(-1, 3), t, st

    def system(self, k, pos, vel):
        a, b = self.offsets[k], self.offsets[k + 1]
        return pos[a:b], vel[a:b]

    def counters(self):
        c = np.zeros((self.n_sys, 6), dtype=np.int64)
        _lib.check(self.lib.abie_ensemble_get_counters(self.h, c.ctypes.data_as(_lib.c_llp)),
                   "abie_ensemble_get_counters")
        return c

    def finish_times(self):
        """Seconds from the first unit's start to each system's end in the
        last run (device clock); -1 for a system still running."""
        f = np.zeros(self.n_sys)
        _lib.check(self.lib.abie_ensemble_get_finish_times(self.h, _dp(f)), "abie_ensemble_get_finish_times")
        return f

    def energy(self):
        e = np.zeros(self.n_sys)
        _lib.check(self.lib.abie_ensemble_energy(self.h, _dp(e)), "abie_ensemble_energy")
        return e

    def set_steplog(self, capacity):
        _lib.check(self.lib.abie_ensemble_set_steplog(self.h, int(capacity)), "abie_ensemble_set_steplog")
        self.log_cap = int(capacity)

    def steplog(self):
        n = ctypes.c_longlong(0)
        _lib.check(self.lib.abie_ensemble_get_steplog(self.h, None, ctypes.byref(n)), "abie_ensemble_get_steplog")
        rec = np.zeros((max(n.value, 1), _lib.STEPLOG_FIELDS))
        _lib.check(self.lib.abie_ensemble_get_steplog(self.h, _dp(rec), ctypes.byref(n)), "abie_ensemble_get_steplog")
        return rec[: n.value]

    def set_proposer(self, weights):
        if weights is None:
            _lib.check(self.lib.abie_ensemble_set_proposer(self.h, None, 0), "abie_ensemble_set_proposer")
            return
        w = np.ascontiguousarray(weights, dtype=np.float32)
        _lib.check(self.lib.abie_ensemble_set_proposer(self.h, w.ctypes.data_as(_lib.c_fp), w.size),
                   "abie_ensemble_set_proposer")


# ---------------------------------------------------------------- MLP
def mlp_forward(w, x, stream=None):
    import torch
    y = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
    _lib.check(_lib.load().abie_mlp_forward(_ptr(w), _ptr(x), _ptr(y), x.shape[0], _stream(stream)),
               "abie_mlp_forward")
    return y


def mlp_loss_grad(w, x, target, sw=None, stream=None):
    import torch
    g = torch.empty(_lib.MLP_NPARAM, dtype=torch.float32, device=x.device)
    loss = torch.empty(1, dtype=torch.float32, device=x.device)
    _lib.check(_lib.load().abie_mlp_loss_grad(_ptr(w), _ptr(x), _ptr(target), None if sw is None else _ptr(sw),
                                              x.shape[0], _ptr(g), _ptr(loss), _stream(stream)), "abie_mlp_loss_grad")
    return g, loss


def mlp_adam(w, g, m, v, step, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
    _lib.check(_lib.load().abie_mlp_adam(_ptr(w), _ptr(g), _ptr(m), _ptr(v), w.numel(), int(step), lr, beta1, beta2,
                                         eps, _stream(stream)), "abie_mlp_adam")
