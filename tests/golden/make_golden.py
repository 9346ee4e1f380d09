"""Write the golden fixtures under tests/golden/ from the reference's own C.

Run in a container that holds the reference sources:

    make -C oracle ref            # compiles /root/reference/src/*.c -> oracle/_ref/
    python tests/golden/make_golden.py

The reference library keeps one global simulation (``src/common.h:70``) that
``initialize_code`` sets up once per process (``src/common.c:225-227``), so
every case runs in a child process of its own.  Each fixture holds the inputs
and the reference's outputs only; nothing of the reference's source travels.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from abie_amd import ics  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "libabie_ref.so")
G = 4 * np.pi ** 2
_dp = ctypes.POINTER(ctypes.c_double)


def ref_lib():
    L = ctypes.CDLL(REF)
    L.initialize_code.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_size_t,
                                  ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int]
    L.initialize_code.restype = ctypes.c_int
    L.set_state.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_size_t, ctypes.c_double, ctypes.c_double]
    L.set_state.restype = None
    L.get_state.argtypes = [_dp, _dp, _dp, _dp]
    L.get_state.restype = ctypes.c_int
    L.integrator_gr.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
    L.integrator_gr.restype = ctypes.c_int
    L.get_model_time.restype = ctypes.c_double
    L.calculate_energy.restype = ctypes.c_double
    L.set_close_encounter_distance.argtypes = [ctypes.c_double]
    L.get_close_encounter_buffer.argtypes = [_dp]
    L.get_collision_buffer.argtypes = [_dp]
    L.ode_n_body_second_order.argtypes = [_dp, ctypes.c_size_t, ctypes.c_double, _dp, _dp, _dp]
    L.ode_n_body_second_order.restype = ctypes.c_size_t
    L.check_collisions_close_encounters.argtypes = [_dp, _dp, ctypes.c_size_t, ctypes.c_double]
    L.check_collisions_close_encounters.restype = ctypes.c_size_t
    L.set_additional_forces.argtypes = [ctypes.c_size_t, _dp]
    L.set_additional_forces.restype = ctypes.c_size_t
    return L


def p(a):
    return a.ctypes.data_as(_dp)


def ref_accel(L, pos, Gc, m, r=None, acc_init=None):
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    m = np.ascontiguousarray(m, dtype=np.float64)
    r = np.zeros_like(m) if r is None else np.ascontiguousarray(r, dtype=np.float64)
    acc = np.zeros_like(pos) if acc_init is None else np.array(acc_init, dtype=np.float64).reshape(-1)
    L.ode_n_body_second_order(p(pos), m.size, Gc, p(m), p(r), p(acc))
    return acc.reshape(-1, 3)


def save(name, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    print("wrote", os.path.relpath(path, ROOT), "%.1f KB" % (os.path.getsize(path) / 1024))


# ---------------------------------------------------------------- cases
def case_forces():
    L = ref_lib()
    out = {}
    pos, vel, m = ics.config0(n_test=1000)
    out.update(c0_pos=pos, c0_m=m, c0_acc=ref_accel(L, pos, G, m))
    pos, vel, m = ics.config1a(n_test=4096)
    out.update(c1a_pos=pos, c1a_m=m, c1a_acc=ref_accel(L, pos, G, m))
    pos, vel, m = ics.plummer(512)
    out.update(pl_pos=pos, pl_m=m, pl_acc=ref_accel(L, pos, 1.0, m))
    save("forces", G=np.float64(G), **out)


def case_forces_edge():
    L = ref_lib()
    out = {}
    rng = np.random.default_rng(7)
    # one body: no sources
    out["one_pos"] = np.array([[0.3, -0.2, 0.1]]); out["one_m"] = np.array([1.0])
    out["one_acc"] = ref_accel(L, out["one_pos"], G, out["one_m"])
    # zero-mass and negative-mass sources are skipped; deleted sink untouched
    pos = rng.normal(size=(6, 3)); m = np.array([1.0, 0.0, 3e-3, -1.0, 0.0, 2e-5])
    sentinel = np.full((6, 3), 7.25)
    out["mix_pos"] = pos; out["mix_m"] = m; out["mix_init"] = sentinel
    out["mix_acc"] = ref_accel(L, pos, G, m, acc_init=sentinel)
    # coincident massive pair -> 0/0
    pos = np.array([[1.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 2.0, 0.0]]); m = np.array([1.0, 0.5, 0.25])
    out["coinc_pos"] = pos; out["coinc_m"] = m; out["coinc_acc"] = ref_accel(L, pos, G, m)
    # test particle on top of a massive body
    pos = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [3.0, 0.0, 0.0]]); m = np.array([1.0, 0.0, 1e-3])
    out["tpon_pos"] = pos; out["tpon_m"] = m; out["tpon_acc"] = ref_accel(L, pos, G, m)
    # widely spread magnitudes
    pos = np.concatenate([rng.normal(size=(8, 3)) * 1e-3, rng.normal(size=(8, 3)) * 1e3])
    m = np.exp(rng.uniform(np.log(1e-12), 0, 16)); m[::5] = 0.0
    out["wide_pos"] = pos; out["wide_m"] = m; out["wide_acc"] = ref_accel(L, pos, G, m)
    save("forces_edge", G=np.float64(G), **out)


def run_chunks(L, pos, vel, m, r, t_end, store_dt, Gc=G, ext=None):
    """The Python driver of the reference (abie/integrator.py:158-203 with
    abie/integrator_gauss_radau15.py:20-45): chunks of store_dt."""
    n = m.size
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1).copy()
    vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1).copy()
    m = m.astype(np.float64).copy(); r = r.astype(np.float64).copy()
    L.set_state(p(pos), p(vel), p(m), p(r), n, Gc, 0.0)
    if ext is not None:
        ext = np.ascontiguousarray(ext, dtype=np.float64).reshape(-1).copy()
        L.set_additional_forces(n, p(ext))
    t = 0.0
    dt = min(store_dt, t_end - t)
    rows = []
    e0 = L.calculate_energy()
    while t < t_end:
        next_t = t + dt - ((t + dt) % dt)
        ret = L.integrator_gr(t, next_t, 1.0)
        L.get_state(p(pos), p(vel), p(m), p(r))
        t = L.get_model_time()
        rows.append((next_t, t, ret, pos.copy(), vel.copy(), L.calculate_energy()))
        if ret != 0:
            break
    return e0, rows


def pack_rows(prefix, e0, rows):
    return {prefix + "e0": np.float64(e0),
            prefix + "next_t": np.array([r[0] for r in rows]),
            prefix + "t": np.array([r[1] for r in rows]),
            prefix + "ret": np.array([r[2] for r in rows]),
            prefix + "pos": np.array([r[3] for r in rows]),
            prefix + "vel": np.array([r[4] for r in rows]),
            prefix + "energy": np.array([r[5] for r in rows])}


def case_gr15():
    L = ref_lib()
    L.initialize_code(G, 0.0, 4096, 1, 1, -1)
    out = {}
    pos, vel, m = ics.planetary_system(11, 5, 20)
    r = np.zeros_like(m)
    out.update(small_pos0=pos, small_vel0=vel, small_m=m)
    e0, rows = run_chunks(L, pos, vel, m, r, 1.0, 1.0)
    out.update(pack_rows("small_", e0, rows))
    # several store_dt chunks; every chunk restarts GR15 (initial step, b = 0)
    e0, rows = run_chunks(L, pos, vel, m, r, 1.0, 0.25)
    out.update(pack_rows("chunks_", e0, rows))
    # config-0 shape with fewer test particles, two chunks of 2.5 yr
    pos, vel, m = ics.config0(n_test=64)
    r = np.zeros_like(m)
    out.update(c0_pos0=pos, c0_vel0=vel, c0_m=m)
    e0, rows = run_chunks(L, pos, vel, m, r, 5.0, 2.5)
    out.update(pack_rows("c0_", e0, rows))
    # constant external acceleration
    pos, vel, m = ics.planetary_system(12, 3, 6)
    ext = np.random.default_rng(3).normal(size=pos.shape) * 1e-3
    r = np.zeros_like(m)
    out.update(ext_pos0=pos, ext_vel0=vel, ext_m=m, ext_acc=ext)
    e0, rows = run_chunks(L, pos, vel, m, r, 0.5, 0.5, ext=ext)
    out.update(pack_rows("ext_", e0, rows))
    save("gr15", G=np.float64(G), **out)


def case_events():
    """Close encounters and collisions: check_collisions_close_encounters
    directly, then inside GR15 (integrator_gauss_radau15.c:342-344)."""
    L = ref_lib()
    L.initialize_code(G, 0.0, 64, 3, 2, -1)
    out = {}
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.05, 0.0, 0.0], [5.0, 5.0, 0.0], [5.0, 5.0, 0.02]])
    radii = np.array([0.0, 0.03, 0.03, 0.0, 0.0])
    L.set_close_encounter_distance(0.06)
    codes, ce_bufs, coll_bufs = [], [], []
    for t in (0.5, 1.5, 2.5):
        codes.append(L.check_collisions_close_encounters(p(pos.reshape(-1).copy()), p(radii.copy()), 5, t))
        b1 = np.zeros(4 * 3); b2 = np.zeros(4 * 2)
        L.get_close_encounter_buffer(p(b1)); L.get_collision_buffer(p(b2))
        ce_bufs.append(b1); coll_bufs.append(b2)
    out.update(chk_pos=pos, chk_r=radii, chk_ce_dist=np.float64(0.06), chk_times=np.array([0.5, 1.5, 2.5]),
               chk_codes=np.array(codes), chk_ce=np.array(ce_bufs), chk_coll=np.array(coll_bufs))
    save("events_check", **out)


def case_events_gr15():
    L = ref_lib()
    L.initialize_code(G, 0.0, 64, 5, 1, -1)
    L.set_close_encounter_distance(0.0)
    # a planet and a massive body on a head-on course; radii make it collide
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 0.3, 0.0], [2.0, 0.0, 0.0]])
    vel = np.array([[0.0, 0.0, 0.0], [0.0, 2 * np.pi, 0.0], [0.0, -2 * np.pi, 0.0], [0.0, 4.44, 0.0]])
    m = np.array([1.0, 1e-3, 1e-3, 0.0])
    r = np.array([0.0, 0.01, 0.01, 0.0])
    e0, rows = run_chunks(L, pos, vel, m, r, 1.0, 1.0)
    b = np.zeros(4)
    L.get_collision_buffer(p(b))
    out = pack_rows("coll_", e0, rows)
    out.update(coll_pos0=pos, coll_vel0=vel, coll_m=m, coll_r=r, coll_buf=b)
    save("events_gr15", **out)


def case_energy():
    L = ref_lib()
    L.initialize_code(G, 0.0, 2048, 1, 1, -1)
    pos, vel, m = ics.config0(n_test=1000)
    m = m.copy(); m[7] = -1e-9  # a deleted body still counts (common.c:344)
    r = np.zeros_like(m)
    L.set_state(p(pos.reshape(-1).copy()), p(vel.reshape(-1).copy()), p(m), p(r), m.size, G, 0.0)
    save("energy", pos=pos, vel=vel, m=m, G=np.float64(G), energy=np.float64(L.calculate_energy()))


def case_tables():
    """The reference's compiled Gauss-Radau tables (integrator_gauss_radau15.h:24-54)."""
    L = ref_lib()
    hs = np.array((ctypes.c_double * 8).in_dll(L, "hs"))
    rs = np.array((ctypes.c_double * 64).in_dll(L, "rs")).reshape(8, 8)
    cs = np.array((ctypes.c_double * 64).in_dll(L, "cs")).reshape(8, 8)
    scal = {name: np.float64(ctypes.c_double.in_dll(L, name).value)
            for name in ("epsb", "fac", "exponent", "tolpc", "h_min")}
    save("radau_tables", hs=hs, rs=rs, cs=cs, **scal)


CASES = {k[5:]: v for k, v in globals().items() if k.startswith("case_")}


if __name__ == "__main__":
    if len(sys.argv) > 1:
        CASES[sys.argv[1]]()
    else:
        if not os.path.exists(REF):
            sys.exit("build the reference first: make -C oracle ref")
        for name in CASES:
            subprocess.run([sys.executable, __file__, name], check=True)
