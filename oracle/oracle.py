"""ctypes front end of the CPU oracle (``oracle/liborc.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s CPU-baseline leg, as the checker (or, for the baseline, the
thing timed on the host).  The product never imports it.

The oracle restates ``src/common.c:39-158,192-208,340-357`` and
``src/integrator_gauss_radau15.c`` of the reference in C (see
``abie_oracle.c``); this module only marshals numpy arrays.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libabie_ref.so")

_dp = ctypes.POINTER(ctypes.c_double)


class Attempt(ctypes.Structure):
    _fields_ = [("t", ctypes.c_double), ("h", ctypes.c_double),
                ("estim_b6", ctypes.c_double), ("err", ctypes.c_double),
                ("dtreq", ctypes.c_double), ("accepted", ctypes.c_int),
                ("pc_iters", ctypes.c_int)]


class OrcSim(ctypes.Structure):
    _fields_ = [("n", ctypes.c_size_t), ("G", ctypes.c_double), ("C", ctypes.c_double),
                ("pos", _dp), ("vel", _dp), ("m", _dp), ("r", _dp),
                ("ext_acc", _dp), ("ext_acc_on", ctypes.c_int),
                ("t", ctypes.c_double), ("ce_dist", ctypes.c_double),
                ("max_ce", ctypes.c_size_t), ("max_coll", ctypes.c_size_t),
                ("n_ce", ctypes.c_size_t), ("n_coll", ctypes.c_size_t),
                ("buf_ce", _dp), ("buf_coll", _dp),
                ("n_evals", ctypes.c_longlong), ("n_attempts", ctypes.c_longlong),
                ("n_accepted", ctypes.c_longlong), ("n_pc_iters", ctypes.c_longlong),
                ("n_hmin", ctypes.c_longlong),
                ("log", ctypes.POINTER(Attempt)), ("log_cap", ctypes.c_size_t),
                ("log_len", ctypes.c_size_t), ("skip_first_node", ctypes.c_int),
                ("pow_fn", ctypes.c_void_p)]


POWFN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_double, ctypes.c_double)


def _cr_pow(x, y):
    """x**y rounded to nearest from a 40-digit mpmath evaluation."""
    import math
    import mpmath as mp
    if x != x or x < 0:
        return math.nan
    if x == 0.0:
        return 0.0
    if math.isinf(x):
        return x
    with mp.workdps(40):
        return float(mp.power(mp.mpf(x), mp.mpf(y)))


CR_POW = POWFN(_cr_pow)


def build():
    """Compile liborc.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE, "liborc.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_accel.argtypes = [_dp, ctypes.c_size_t, ctypes.c_double, _dp, _dp]
        L.orc_accel.restype = None
        L.orc_accel_range.argtypes = [_dp, ctypes.c_size_t, ctypes.c_double, _dp, _dp, ctypes.c_size_t,
                                      ctypes.c_size_t]
        L.orc_accel_range.restype = None
        L.orc_gr15.argtypes = [ctypes.POINTER(OrcSim), ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_gr15.restype = ctypes.c_int
        L.orc_energy.argtypes = [ctypes.POINTER(OrcSim)]
        L.orc_energy.restype = ctypes.c_double
        L.orc_check_events.argtypes = [ctypes.POINTER(OrcSim), _dp, ctypes.c_double]
        L.orc_check_events.restype = ctypes.c_int
        L.orc_calculate_accelerations.argtypes = [ctypes.POINTER(OrcSim), _dp, _dp, _dp]
        L.orc_calculate_accelerations.restype = ctypes.c_int
        L.orc_div_const.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_div_const.restype = ctypes.c_double
        L.orc_vector_max_abs.argtypes = [_dp, ctypes.c_size_t]
        L.orc_vector_max_abs.restype = ctypes.c_double
        for name in ("orc_hs", "orc_rs", "orc_cs"):
            getattr(L, name).restype = _dp
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_dp)


def accel(pos, G, m, acc_init=None):
    """``ode_n_body_second_order`` (common.c:39-77) on an (n, 3) array.

    Deleted sinks (m < 0) keep ``acc_init`` (default zeros)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    m = np.ascontiguousarray(m, dtype=np.float64)
    n = m.size
    acc = np.zeros(3 * n) if acc_init is None else np.array(acc_init, dtype=np.float64).reshape(-1)
    lib().orc_accel(_ptr(pos), n, G, _ptr(m), _ptr(acc))
    return acc.reshape(n, 3)


def accel_threads(pos, G, m, threads):
    """``accel`` with the sinks split into ``threads`` contiguous ranges, each
    on its own host thread (ctypes releases the GIL during the C call); the
    result is bit-identical to ``accel``."""
    from concurrent.futures import ThreadPoolExecutor
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    m = np.ascontiguousarray(m, dtype=np.float64)
    n = m.size
    acc = np.zeros(3 * n)
    L = lib()
    bounds = np.linspace(0, n, threads + 1).astype(np.int64)
    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(lambda k: L.orc_accel_range(_ptr(pos), n, G, _ptr(m), _ptr(acc), int(bounds[k]),
                                                int(bounds[k + 1])), range(threads)))
    return acc.reshape(n, 3)


def tables():
    L = lib()
    hs = np.ctypeslib.as_array(L.orc_hs(), shape=(8,)).copy()
    rs = np.ctypeslib.as_array(L.orc_rs(), shape=(64,)).reshape(8, 8).copy()
    cs = np.ctypeslib.as_array(L.orc_cs(), shape=(64,)).reshape(8, 8).copy()
    return hs, rs, cs


class Sim:
    """One simulation held by the oracle (the reference's global ``sim``,
    common.h:26-65, made explicit)."""

    def __init__(self, pos, vel, m, r=None, G=4 * np.pi ** 2, C=0.0,
                 max_ce=1, max_coll=1, ce_dist=0.0, ext_acc=None,
                 log_cap=0, skip_first_node=False, cr_pow=False):
        n = len(m)
        self.pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1).copy()
        self.vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1).copy()
        self.m = np.ascontiguousarray(m, dtype=np.float64).copy()
        self.r = np.zeros(n) if r is None else np.ascontiguousarray(r, dtype=np.float64).copy()
        self.buf_ce = np.zeros(4 * max_ce) if max_ce > 0 else None
        self.buf_coll = np.zeros(4 * max_coll) if max_coll > 0 else None
        self.ext = None if ext_acc is None else np.ascontiguousarray(ext_acc, dtype=np.float64).reshape(-1).copy()
        self.log = (Attempt * log_cap)() if log_cap > 0 else None
        s = OrcSim()
        s.n = n; s.G = G; s.C = C
        s.pos = _ptr(self.pos); s.vel = _ptr(self.vel); s.m = _ptr(self.m); s.r = _ptr(self.r)
        s.ext_acc = _ptr(self.ext) if self.ext is not None else _dp()
        s.ext_acc_on = 1 if self.ext is not None else 0
        s.t = 0.0; s.ce_dist = ce_dist
        s.max_ce = max_ce; s.max_coll = max_coll
        s.buf_ce = _ptr(self.buf_ce) if self.buf_ce is not None else _dp()
        s.buf_coll = _ptr(self.buf_coll) if self.buf_coll is not None else _dp()
        if self.log is not None:
            s.log = ctypes.cast(self.log, ctypes.POINTER(Attempt))
        s.log_cap = log_cap
        s.skip_first_node = 1 if skip_first_node else 0
        # cr_pow: correctly rounded step-size exponents (glibc's pow is not
        # correctly rounded for ~0.1% of x**(1/7)); the device engine rounds
        # them correctly, so this is the oracle the device must match bitwise
        s.pow_fn = ctypes.cast(CR_POW, ctypes.c_void_p).value if cr_pow else None
        self.s = s

    @property
    def t(self):
        return self.s.t

    def integrate(self, t_end, t=None):
        """``integrator_gr(t, t_end, dt)`` (common.c:400-403)."""
        t0 = self.s.t if t is None else t
        return lib().orc_gr15(ctypes.byref(self.s), t0, t_end, 1.0)

    def energy(self):
        return lib().orc_energy(ctypes.byref(self.s))

    def check_events(self, pos=None, t=0.0):
        p = self.pos if pos is None else np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        return lib().orc_check_events(ctypes.byref(self.s), _ptr(p), t)

    def accelerations(self, pos=None, acc=None):
        """calculate_accelerations into a fresh array, or into a copy of the
        caller's ``acc`` (whose deleted-body entries the force loop leaves)."""
        p = self.pos if pos is None else np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        acc = np.zeros_like(p) if acc is None else np.ascontiguousarray(acc, dtype=np.float64).reshape(-1).copy()
        rc = lib().orc_calculate_accelerations(ctypes.byref(self.s), _ptr(p), _ptr(self.vel), _ptr(acc))
        if rc != 0:
            raise ValueError("oracle: unsupported configuration (C > 0)")
        return acc.reshape(-1, 3)

    def attempts(self):
        if self.log is None:
            return []
        return [self.log[i] for i in range(self.s.log_len)]

    def counters(self):
        s = self.s
        return dict(evals=s.n_evals, attempts=s.n_attempts, accepted=s.n_accepted,
                    pc_iters=s.n_pc_iters, hmin=s.n_hmin, n_ce=s.n_ce, n_coll=s.n_coll)
