"""Parity of the device GR15 engine with the reference (golden fixtures) and
the CPU oracle, through the C ABI.

With ABIE_FORCE_FP64_EXACT forces every trajectory is bit-identical to the
reference's; with the default FP64 forces positions agree to 1e-7 relative
after the short runs used here (rounding differences of ~1e-16 per force are
amplified by the integration: a 1-ulp change of one initial coordinate moves
the oracle's own 5-yr run by 5e-9)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = 4 * np.pi ** 2


def same_bits(a, b):
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    nan = np.isnan(a) & np.isnan(b)
    return bool(np.all((a.view(np.int64) == b.view(np.int64)) | nan))


@pytest.fixture()
def clib():
    from abie_amd.clibabie import CLibABIE
    c = CLibABIE()
    c.lib.finalize_code()
    yield c
    c.lib.abie_hip_set_force_mode(0)
    c.lib.finalize_code()


def drive(clib, pos, vel, m, r, t_end, store_dt, max_ce=1, max_coll=1, ext=None, mode=1):
    """abie/integrator.py:158-203 chunk loop over integrator_gr."""
    from abie_amd.events import ParticleException
    clib.initialize_code(G, 0.0, len(m), max_ce, max_coll, 0.0, deviceID=0)
    clib.lib.abie_hip_set_force_mode(mode)
    n = len(m)
    p = np.ascontiguousarray(pos, np.float64).reshape(-1).copy()
    v = np.ascontiguousarray(vel, np.float64).reshape(-1).copy()
    mm = np.ascontiguousarray(m, np.float64).copy(); rr = np.ascontiguousarray(r, np.float64).copy()
    clib.set_state(p, v, mm, rr, n, G, 0.0)
    if ext is not None:
        clib.set_additional_forces(ext.reshape(-1))
    t = 0.0
    dt = min(store_dt, t_end - t)
    rows = []
    e0 = clib.get_total_energy()
    while t < t_end:
        next_t = t + dt - ((t + dt) % dt)
        ret = 0
        try:
            clib.integrator_gr(t, next_t, 1.0)
        except ParticleException as ex:
            ret = 1 if "Close" in type(ex).__name__ else 2
        clib.get_state(p, v, mm, rr)
        t = clib.get_model_time()
        rows.append((next_t, t, ret, p.copy(), v.copy(), clib.get_total_energy()))
        if ret:
            break
    return e0, rows


@pytest.mark.parametrize("case,t_end,store_dt", [("small", 1.0, 1.0), ("chunks", 1.0, 0.25),
                                                  ("c0", 5.0, 2.5), ("ext", 0.5, 0.5)])
def test_exact_gr15_bitwise_reference(clib, golden, case, t_end, store_dt):
    d = golden("gr15")
    pre = "small" if case == "chunks" else case
    m = d[pre + "_m"]
    ext = d["ext_acc"] if case == "ext" else None
    e0, rows = drive(clib, d[pre + "_pos0"], d[pre + "_vel0"], m, np.zeros_like(m), t_end, store_dt, ext=ext)
    assert len(rows) == len(d[case + "_t"])
    for i, (nt, t, ret, pos, vel, en) in enumerate(rows):
        assert t == d[case + "_t"][i]
        assert ret == d[case + "_ret"][i]
        assert same_bits(pos, d[case + "_pos"][i]), i
        assert same_bits(vel, d[case + "_vel"][i]), i
        assert en == pytest.approx(d[case + "_energy"][i], rel=1e-14, abs=0)
    assert e0 == pytest.approx(d[case + "_e0"], rel=1e-14)


def test_fp64_gr15_close_to_reference(clib, golden):
    d = golden("gr15")
    m = d["c0_m"]
    e0, rows = drive(clib, d["c0_pos0"], d["c0_vel0"], m, np.zeros_like(m), 5.0, 2.5, mode=0)
    last = rows[-1]
    ref = d["c0_pos"][-1].reshape(-1, 3)
    got = last[3].reshape(-1, 3)
    # a 1-ulp perturbation of one coordinate moves this run by ~5e-9 (oracle)
    assert np.max(np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)) < 1e-7
    assert abs((last[5] - e0) / e0) < 1e-13


def test_collision_event_reference(clib, golden):
    """Collision inside GR15: return code, time, state and ring as the reference."""
    d = golden("events_gr15")
    e0, rows = drive(clib, d["coll_pos0"], d["coll_vel0"], d["coll_m"], d["coll_r"], 1.0, 1.0, max_ce=5, max_coll=1)
    assert rows[-1][2] == 2
    assert rows[-1][1] == d["coll_t"][0]
    assert same_bits(rows[-1][3], d["coll_pos"][0]) and same_bits(rows[-1][4], d["coll_vel"][0])
    assert same_bits(clib.get_collision_data().reshape(-1), d["coll_buf"])


def test_check_events_reference(clib, golden):
    d = golden("events_check")
    clib.initialize_code(G, 0.0, 8, 3, 2, float(d["chk_ce_dist"]), deviceID=0)
    for i, t in enumerate(d["chk_times"]):
        ret = clib.check_collisions_close_encounters(d["chk_pos"].reshape(-1), d["chk_r"], 5, float(t))
        assert ret == d["chk_codes"][i]
        assert same_bits(clib.get_close_encounter_data().reshape(-1), d["chk_ce"][i])
        assert same_bits(clib.get_collision_data().reshape(-1), d["chk_coll"][i])


def test_energy_reference(clib, golden):
    d = golden("energy")
    clib.initialize_code(G, 0.0, len(d["m"]), 1, 1, 0.0, deviceID=0)
    clib.set_state(d["pos"].reshape(-1), d["vel"].reshape(-1), d["m"], np.zeros(len(d["m"])), len(d["m"]), G, 0.0)
    assert clib.get_total_energy() == pytest.approx(float(d["energy"]), rel=1e-14)


def test_ensemble_exact_matches_oracle_per_system(orc):
    """Systems of every engine configuration in one ensemble, exact forces:
    each system bit-identical to the oracle's GR15 on its own (the oracle with
    correctly rounded step-size exponents, as the device computes them)."""
    from abie_amd import engine, ics
    sizes = [(3, 5), (5, 100), (12, 200), (30, 400), (64, 900)]
    systems = [ics.planetary_system(100 + k, nm, nt) for k, (nm, nt) in enumerate(sizes)]
    ens = engine.Ensemble(systems, G=G, force_mode="exact", controller="plain")
    assert ens.run(0.3)
    pos, vel, t, st = ens.state()
    cnt = ens.counters()
    for k, (p0, v0, m) in enumerate(systems):
        s = orc.Sim(p0, v0, m, cr_pow=True)
        s.integrate(0.3, 0.0)
        p, v = ens.system(k, pos, vel)
        assert t[k] == s.t
        assert same_bits(p, s.pos.reshape(-1, 3)), k
        assert same_bits(v, s.vel.reshape(-1, 3)), k
        assert cnt[k, 2] == s.counters()["accepted"] and cnt[k, 1] == s.counters()["attempts"]
    ens.close()


def test_ensemble_resume_is_seamless():
    """Stopping at an intermediate time and continuing gives the same
    trajectory as one uninterrupted run (state persists on the device)."""
    from abie_amd import engine, ics
    systems = [ics.planetary_system(7, 6, 50), ics.planetary_system(8, 3, 10)]
    a = engine.Ensemble(systems, force_mode="fp64")
    a.run(0.7)
    b = engine.Ensemble(systems, force_mode="fp64")
    b.run(0.2)
    b.run(0.45)
    b.run(0.7)
    pa, va, ta, _ = a.state()
    pb, vb, tb, _ = b.state()
    assert np.array_equal(ta, tb) and same_bits(pa, pb) and same_bits(va, vb)


def test_ensemble_energy_conservation_mixed():
    from abie_amd import engine, ics
    systems = [ics.planetary_system(40 + k, 5, 200) for k in range(4)]
    ens = engine.Ensemble(systems, force_mode="mixed", controller="massive")
    e0 = ens.energy()
    ens.run(2.0)
    e1 = ens.energy()
    assert np.all(np.abs((e1 - e0) / e0) < 1e-12)


def kepler_position(r0, v0, mu, t):
    """Universal-variable two-body solution (mpmath, 40 digits)."""
    import mpmath as mp
    mp.mp.dps = 40
    r0 = [mp.mpf(float(x)) for x in r0]; v0 = [mp.mpf(float(x)) for x in v0]
    mu = mp.mpf(mu)
    rn = mp.sqrt(sum(x * x for x in r0)); vn2 = sum(x * x for x in v0)
    a = 1 / (2 / rn - vn2 / mu)
    n = mp.sqrt(mu / a ** 3)
    rv = sum(x * y for x, y in zip(r0, v0))
    e_cos = 1 - rn / a; e_sin = rv / mp.sqrt(mu * a)
    e = mp.sqrt(e_cos ** 2 + e_sin ** 2)
    E0 = mp.atan2(e_sin, e_cos)
    M = E0 - e_sin + n * mp.mpf(t)
    E = M
    for _ in range(60):
        E = E - (E - e * mp.sin(E) - M) / (1 - e * mp.cos(E))
    dE = E - E0
    F = 1 - a / rn * (1 - mp.cos(dE))
    Gt = mp.mpf(t) - mp.sqrt(a ** 3 / mu) * (dE - mp.sin(dE))
    return np.array([float(F * r0[i] + Gt * v0[i]) for i in range(3)])


@pytest.mark.parametrize("controller", ["plain", "massive"])
def test_kepler_orbit_closed_form(controller):
    """A test particle around a star (one massive body: the massive-only
    controller then controls on every body) for 10 orbits of e ~ 0.21,
    against the closed-form solution."""
    from abie_amd import engine
    r0 = np.array([1.0, 0.0, 0.0]); v0 = np.array([0.0, 2 * np.pi * 1.1, 0.0])
    pos = np.array([[0.0, 0.0, 0.0], r0]); vel = np.array([[0.0, 0.0, 0.0], v0])
    m = np.array([1.0, 0.0])
    ens = engine.Ensemble([(pos, vel, m)], G=G, controller=controller)
    assert ens.run(10.0)
    p, v, t, st = ens.state()
    want = kepler_position(r0, v0, G, t[0])
    assert np.all(np.isfinite(p)) and st[0] == 0
    assert np.linalg.norm(p[1] - want) < 1e-9
    assert np.array_equal(p[0], np.zeros(3))


def test_force_free_body_moves_freely():
    """One body alone has no acceleration (the reference's error estimate
    would be 0/0): it moves at constant velocity and the run ends."""
    from abie_amd import engine
    pos = np.array([[1.0, 2.0, 3.0]]); vel = np.array([[0.5, -0.25, 0.125]])
    for ctrl in ("plain", "massive"):
        ens = engine.Ensemble([(pos, vel, np.array([1.0]))], G=G, controller=ctrl)
        assert ens.run(2.0)
        p, v, t, st = ens.state()
        assert t[0] >= 2.0 and np.allclose(p[0], pos[0] + vel[0] * t[0], rtol=1e-14, atol=1e-14)
        assert np.array_equal(v[0], vel[0])


def test_exact_gr15_with_deleted_and_massless_bodies(orc):
    """A deleted body (m < 0: neither source nor sink) and test particles in
    one system: the device GR15 with exact forces equals the oracle bit for
    bit (both give the deleted body zero acceleration)."""
    from abie_amd import engine, ics
    pos, vel, m = ics.planetary_system(31, 6, 20)
    m = m.copy()
    m[3] = -1.0
    ens = engine.Ensemble([(pos, vel, m)], G=G, force_mode="exact", controller="plain")
    assert ens.run(0.4)
    p, v, t, _ = ens.state()
    s = orc.Sim(pos, vel, m, cr_pow=True)
    s.integrate(0.4, 0.0)
    assert t[0] == s.t
    assert same_bits(p, s.pos.reshape(-1, 3)) and same_bits(v, s.vel.reshape(-1, 3))


@pytest.mark.parametrize("pair", [(3, 900), (600, 601), (700, 1100), (1190, 1195)])
def test_close_encounter_scan_across_tiles_matches_oracle(clib, orc, pair):
    """integrator_gr on 2 massive bodies + 1,200 test particles, one pair of
    test particles kept within the close-encounter distance (the scan stages
    the 1,202 bodies in three LDS tiles of 512; the pairs sit across tiles
    and on a tile's diagonal): return code, time, state and ring bit-identical
    to the oracle's GR15 with exact forces."""
    from abie_amd import ics
    from abie_amd.events import ParticleException
    pos, vel, m = ics.planetary_system(909, 2, 1200)
    i, j = pair
    pos = pos.copy()
    vel = vel.copy()
    pos[j] = pos[i] + np.array([3e-4, -2e-4, 1e-4])
    vel[j] = vel[i]
    n = len(m)
    r = np.zeros(n)
    clib.initialize_code(G, 0.0, n, 2, 1, 1e-3, deviceID=0)
    clib.lib.abie_hip_set_force_mode(1)
    p, v, mm, rr = pos.reshape(-1).copy(), vel.reshape(-1).copy(), m.copy(), r.copy()
    clib.set_state(p, v, mm, rr, n, G, 0.0)
    ret = 0
    try:
        clib.integrator_gr(0.0, 0.5, 1.0)
    except ParticleException as ex:
        ret = 1 if "Close" in type(ex).__name__ else 2
    clib.get_state(p, v, mm, rr)
    s = orc.Sim(pos, vel, m, r, max_ce=2, max_coll=1, ce_dist=1e-3, cr_pow=True)
    ref = s.integrate(0.5, 0.0)
    assert ret == ref == 1
    assert clib.get_model_time() == s.t
    assert same_bits(p, s.pos) and same_bits(v, s.vel)
    ce = clib.get_close_encounter_data().reshape(-1)
    assert same_bits(ce, s.buf_ce)
    assert sorted(ce[1:3]) == [i, j]


@pytest.mark.parametrize("pair", [(5, 1000), (1100, 1101)])
def test_collision_scan_with_radii_across_tiles_matches_oracle(clib, orc, pair):
    """The same 1,202-body system with radii on every body (the scan's
    radius path): the close pair collides on the first accepted step; return
    code, time, state and both rings as the oracle's GR15."""
    from abie_amd import ics
    from abie_amd.events import ParticleException
    pos, vel, m = ics.planetary_system(910, 2, 1200)
    i, j = pair
    pos = pos.copy()
    vel = vel.copy()
    pos[j] = pos[i] + np.array([3e-4, -2e-4, 1e-4])
    vel[j] = vel[i]
    n = len(m)
    r = np.random.default_rng(4).uniform(1e-7, 1e-6, n)
    r[i] = r[j] = 2e-4
    clib.initialize_code(G, 0.0, n, 2, 1, 1e-3, deviceID=0)
    clib.lib.abie_hip_set_force_mode(1)
    p, v, mm, rr = pos.reshape(-1).copy(), vel.reshape(-1).copy(), m.copy(), r.copy()
    clib.set_state(p, v, mm, rr, n, G, 0.0)
    ret = 0
    try:
        clib.integrator_gr(0.0, 0.5, 1.0)
    except ParticleException as ex:
        ret = 1 if "Close" in type(ex).__name__ else 2
    clib.get_state(p, v, mm, rr)
    s = orc.Sim(pos, vel, m, r, max_ce=2, max_coll=1, ce_dist=1e-3, cr_pow=True)
    ref = s.integrate(0.5, 0.0)
    assert ret == ref == 2
    assert clib.get_model_time() == s.t
    assert same_bits(p, s.pos) and same_bits(v, s.vel)
    assert same_bits(clib.get_close_encounter_data().reshape(-1), s.buf_ce)
    coll = clib.get_collision_data().reshape(-1)
    assert same_bits(coll, s.buf_coll) and sorted(coll[1:3]) == [i, j]


@pytest.mark.parametrize("dup", [True, False])
@pytest.mark.parametrize("pair", [(4, 900), (600, 601)])
def test_zero_encounter_distance_scan_matches_oracle(clib, orc, pair, dup):
    """Close-encounter distance 0 and no radii (the default set-up): a pair
    of test particles at one position is still a close encounter (d = 0 <=
    0, common.c:133); a pair that shares its x coordinate at every step
    (mirror images through the plane of the massive bodies, so x and y stay
    bitwise equal and z exactly opposite) passes the scan's x-slab
    prefilter after every step but is no event.  Return code, time, state
    and ring bit-identical to the oracle's GR15 with exact forces."""
    from abie_amd import ics
    from abie_amd.events import ParticleException
    pos, vel, m = ics.planetary_system(912, 2, 1200)
    i, j = pair
    pos = pos.copy()
    vel = vel.copy()
    pos[:2, 2] = 0.0
    vel[:2, 2] = 0.0
    if dup:
        pos[j], vel[j] = pos[i], vel[i]
    else:
        pos[i, 2] = 0.01
        mirror = np.array([1.0, 1.0, -1.0])
        pos[j], vel[j] = pos[i] * mirror, vel[i] * mirror
    n = len(m)
    r = np.zeros(n)
    clib.initialize_code(G, 0.0, n, 1, 1, 0.0, deviceID=0)
    clib.lib.abie_hip_set_force_mode(1)
    p, v, mm, rr = pos.reshape(-1).copy(), vel.reshape(-1).copy(), m.copy(), r.copy()
    clib.set_state(p, v, mm, rr, n, G, 0.0)
    ret = 0
    try:
        clib.integrator_gr(0.0, 0.05, 1.0)
    except ParticleException as ex:
        ret = 1 if "Close" in type(ex).__name__ else 2
    clib.get_state(p, v, mm, rr)
    s = orc.Sim(pos, vel, m, r, max_ce=1, max_coll=1, ce_dist=0.0, cr_pow=True)
    ref = s.integrate(0.05, 0.0)
    assert ret == ref == (1 if dup else 0)
    assert clib.get_model_time() == s.t
    assert same_bits(p, s.pos) and same_bits(v, s.vel)
    ce = clib.get_close_encounter_data().reshape(-1)
    assert same_bits(ce, s.buf_ce)
    if dup:
        assert sorted(ce[1:3]) == [i, j] and ce[3] == 0.0
    else:
        q = s.pos.reshape(-1, 3)
        assert q[i, 0] == q[j, 0] and q[i, 1] == q[j, 1] and q[i, 2] == -q[j, 2] != 0.0


def test_many_events_in_one_step_wrap_rings_as_oracle(clib, orc):
    """60 close test-particle pairs, 25 of them colliding, in one step: the
    rings (7 and 3 slots) wrap many times inside one chunk of the parallel
    exact pass; rings, counters and state as the oracle's GR15."""
    from abie_amd import ics
    from abie_amd.events import ParticleException
    pos, vel, m = ics.planetary_system(911, 2, 400)
    pos = pos.copy()
    vel = vel.copy()
    n = len(m)
    r = np.zeros(n)
    for q in range(60):
        a, b = 2 + 5 * q, 4 + 5 * q
        pos[b] = pos[a] + np.array([2e-4, 1e-4 * (q % 3), -1e-4])
        vel[b] = vel[a]
        if q % 12 < 5:
            r[a] = r[b] = 2e-4
    clib.initialize_code(G, 0.0, n, 7, 3, 1e-3, deviceID=0)
    clib.lib.abie_hip_set_force_mode(1)
    p, v, mm, rr = pos.reshape(-1).copy(), vel.reshape(-1).copy(), m.copy(), r.copy()
    clib.set_state(p, v, mm, rr, n, G, 0.0)
    ret = 0
    try:
        clib.integrator_gr(0.0, 0.5, 1.0)
    except ParticleException as ex:
        ret = 1 if "Close" in type(ex).__name__ else 2
    clib.get_state(p, v, mm, rr)
    s = orc.Sim(pos, vel, m, r, max_ce=7, max_coll=3, ce_dist=1e-3, cr_pow=True)
    ref = s.integrate(0.5, 0.0)
    assert ret == ref == 1
    assert clib.get_model_time() == s.t
    assert same_bits(p, s.pos) and same_bits(v, s.vel)
    assert same_bits(clib.get_close_encounter_data().reshape(-1), s.buf_ce)
    assert same_bits(clib.get_collision_data().reshape(-1), s.buf_coll)


def test_check_events_many_bodies_matches_oracle(clib, orc):
    """check_collisions_close_encounters on 1,500 bodies (six filter blocks),
    50 close pairs, 20 of them touching, rings of 7 and 3 slots: return
    codes, rings and counters over two calls as the oracle."""
    rng = np.random.default_rng(12)
    n = 1500
    pos = rng.uniform(-10, 10, size=(n, 3))
    r = np.zeros(n)
    for q in range(50):
        a, b = 7 + 29 * q, 20 + 29 * q
        pos[b] = pos[a] + rng.normal(size=3) * 2e-3
        if q % 5 < 2:
            r[a] = r[b] = 5e-3
    clib.initialize_code(G, 0.0, n, 7, 3, 1e-2, deviceID=0)
    s = orc.Sim(pos, np.zeros_like(pos), np.ones(n), r, max_ce=7, max_coll=3, ce_dist=1e-2)
    for t in (1.0, 2.0):
        ret = clib.check_collisions_close_encounters(pos.reshape(-1), r, n, t)
        assert ret == s.check_events(t=t)
        assert same_bits(clib.get_close_encounter_data().reshape(-1), s.buf_ce)
        assert same_bits(clib.get_collision_data().reshape(-1), s.buf_coll)


def test_integrator_gr_reuses_engine_until_state_changes(clib, orc):
    """integrator_gr keeps its engine between calls; set_state (new bodies)
    and set_additional_forces drop it: each run matches the oracle started
    from the state the caller last set (exact forces, bit for bit)."""
    from abie_amd import ics
    pos, vel, m = ics.planetary_system(77, 4, 30)
    n = len(m)
    r = np.zeros(n)
    clib.initialize_code(G, 0.0, n, 1, 1, 0.0, deviceID=0)
    clib.lib.abie_hip_set_force_mode(1)
    p, v = pos.reshape(-1).copy(), vel.reshape(-1).copy()
    clib.set_state(p, v, m.copy(), r.copy(), n, G, 0.0)
    s = orc.Sim(pos, vel, m, cr_pow=True)
    for t0, t1 in ((0.0, 0.2), (0.2, 0.4)):  # second call on the kept engine
        clib.integrator_gr(clib.get_model_time() if t0 else 0.0, t1, 1.0)
        s.integrate(t1, s.t if t0 else 0.0)
        clib.get_state(p, v, m.copy(), r.copy())
        assert clib.get_model_time() == s.t
        assert same_bits(p, s.pos) and same_bits(v, s.vel)
    # new bodies: the engine must start from them
    pos2, vel2 = pos * 1.01, vel * 0.99
    clib.set_state(pos2.reshape(-1).copy(), vel2.reshape(-1).copy(), m.copy(), r.copy(), n, G, 0.0)
    s2 = orc.Sim(pos2, vel2, m, cr_pow=True)
    clib.integrator_gr(0.0, 0.3, 1.0)
    s2.integrate(0.3, 0.0)
    clib.get_state(p, v, m.copy(), r.copy())
    assert same_bits(p, s2.pos) and same_bits(v, s2.vel)
    # extra forces switched on: the engine must see them
    ext = np.random.default_rng(3).normal(size=3 * n) * 1e-3
    clib.set_additional_forces(ext)
    s2 = orc.Sim(p.reshape(-1, 3), v.reshape(-1, 3), m, ext_acc=ext, cr_pow=True)
    t = clib.get_model_time()
    clib.integrator_gr(t, t + 0.2, 1.0)
    s2.integrate(t + 0.2, t)
    clib.get_state(p, v, m.copy(), r.copy())
    assert same_bits(p, s2.pos) and same_bits(v, s2.vel)
    clib.set_additional_forces(np.zeros(3 * n))
