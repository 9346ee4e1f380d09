"""The quad form of the GR15 engine (four lanes per body, gr15_quad.hpp) and
the hybrid schedule of the massive-only controllers.

Both engine forms sum each body's sources with the same fixed summation tree
(the partial sums over j = c mod 4 in order, then (P0 + P1) + (P2 + P3)) and
share every other operation, so a unit gives the same bits in either form,
in any group or slice, and however its run is cut into launches."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = 4 * np.pi ** 2


def same_bits(a, b):
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    return bool(np.all(a.view(np.int64) == b.view(np.int64)))


def _systems():
    from abie_amd import ics
    # norm_all (one massive body), small and large massive counts, systems
    # that fit one group and systems sliced over several
    shapes = [(1, 40), (2, 0), (3, 10), (5, 300), (12, 60), (24, 700), (41, 0), (64, 190), (33, 1500)]
    return [ics.planetary_system(4100 + k, nm, nt) for k, (nm, nt) in enumerate(shapes)]


def _run(systems, years, force_mode, controller, monkeypatch=None, quad=True, proposer=None):
    from abie_amd import engine
    if monkeypatch is not None:
        monkeypatch.setenv("ABIE_HIP_QUAD", "1" if quad else "0")
    e = engine.Ensemble(systems, G=G, force_mode=force_mode, controller=controller)
    if proposer is not None:
        e.set_proposer(proposer)
    e0 = e.energy()
    done = e.run(years)
    p, v, t, st = e.state()
    e1 = e.energy()
    c = e.counters()
    e.close()
    # a lone star with test particles has zero energy: its |dE| is absolute
    de = np.abs(e1 - e0) / np.where(e0 != 0, np.abs(e0), 1.0)
    return dict(done=done, p=p, v=v, t=t, st=st, de=de, cnt=c)


def _run_mode(monkeypatch, mode, systems, years, force_mode, controller, proposer=None):
    from abie_amd import engine
    monkeypatch.setenv("ABIE_HIP_QUAD", mode)
    e = engine.Ensemble(systems, G=G, force_mode=force_mode, controller=controller)
    if proposer is not None:
        e.set_proposer(proposer)
    done = e.run(years)
    p, v, t, st = e.state()
    c = e.counters()
    e.close()
    return done, p, v, t, st, c


@pytest.mark.parametrize("force_mode,controller", [("fp64", "massive"), ("mixed", "massive"), ("mixed", "learned")])
def test_one_lane_and_quad_forms_same_bits(monkeypatch, force_mode, controller):
    """The one-lane form (ABIE_HIP_QUAD=0), the quad form from the start (2)
    and the hybrid schedule (1: one-lane while more units run than there
    are CUs, then the quad form; 320 systems make about 600 units, so the
    switch happens mid-run) share the force sums' summation tree and every
    other operation: identical bits, step counts and proposer decisions."""
    import os
    from abie_amd import ics
    w = None
    if controller == "learned":
        w = np.load(os.path.join(os.path.dirname(__file__), "..", "abie_amd", "data", "proposer.npz"))["weights"]
    systems = ics.ensemble(320, seed=20261203 + 31000)
    runs = {m: _run_mode(monkeypatch, m, systems, 0.1, force_mode, controller, w) for m in ("0", "1", "2")}
    d0, p0, v0, t0, s0, c0 = runs["0"]
    assert d0 and np.all(s0 == 0)
    for m in ("1", "2"):
        d, p, v, t, st, c = runs[m]
        assert d and np.array_equal(t, t0) and np.array_equal(st, s0)
        assert same_bits(p, p0) and same_bits(v, v0), m
        assert np.array_equal(c, c0), m


def test_deadline_stops_and_resumes_with_the_same_bits():
    """max_seconds ends a call at a device-clock deadline (state kept); calls
    repeated until done give the same bits as one uninterrupted call."""
    from abie_amd import engine, ics
    systems = [ics.planetary_system(4200 + k, 6 + 3 * k, 50 * k) for k in range(5)]
    for ctrl in ("plain", "massive"):
        a = engine.Ensemble(systems, G=G, controller=ctrl)
        assert a.run(0.4)
        b = engine.Ensemble(systems, G=G, controller=ctrl)
        calls = 1
        while not b.run(0.4, max_seconds=1e-4):
            calls += 1
            assert calls < 100000
        assert calls > 1
        pa, va, ta, _ = a.state()
        pb, vb, tb, _ = b.state()
        assert np.array_equal(ta, tb) and same_bits(pa, pb) and same_bits(va, vb)
        assert np.array_equal(a.counters(), b.counters())


@pytest.mark.parametrize("force_mode,controller", [("fp64", "massive"), ("mixed", "learned")])
def test_quad_bits_independent_of_grouping_and_launches(monkeypatch, force_mode, controller):
    """A system alone (a lone system is cut into slices of 128 bodies) and
    the same system inside an ensemble (slices of 256), and the ensemble run
    in launches of a few attempts each: identical bits for every body."""
    import os
    w = None
    if controller == "learned":
        w = np.load(os.path.join(os.path.dirname(__file__), "..", "abie_amd", "data", "proposer.npz"))["weights"]
    systems = _systems()
    k = 8  # 33 massive + 1,500 test particles
    whole = _run(systems, 0.2, force_mode, controller, monkeypatch, proposer=w)
    alone = _run([systems[k]], 0.2, force_mode, controller, monkeypatch, proposer=w)
    off = np.concatenate([[0], np.cumsum([len(s[2]) for s in systems])])
    a, b = off[k], off[k + 1]
    assert whole["t"][k] == alone["t"][0]
    assert same_bits(whole["p"][a:b], alone["p"]) and same_bits(whole["v"][a:b], alone["v"])
    monkeypatch.setenv("ABIE_HIP_MAX_ATTEMPTS", "7")
    chunked = _run(systems, 0.2, force_mode, controller, monkeypatch, proposer=w)
    assert np.array_equal(chunked["t"], whole["t"])
    assert same_bits(chunked["p"], whole["p"]) and same_bits(chunked["v"], whole["v"])
    assert np.array_equal(chunked["cnt"], whole["cnt"])


def test_quad_deleted_body_moves_freely(monkeypatch):
    """A deleted body (m < 0) is neither source nor sink: it keeps its
    velocity exactly and moves on a straight line."""
    from abie_amd import ics
    pos, vel, m = ics.planetary_system(77, 6, 40)
    m = m.copy()
    m[4] = -1.0
    r = _run([(pos, vel, m)], 0.5, "fp64", "massive", monkeypatch)
    assert r["done"]
    assert np.array_equal(r["v"][4], vel[4])
    assert np.allclose(r["p"][4], pos[4] + vel[4] * r["t"][0], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("force_mode,controller", [("fp64", "massive"), ("mixed", "learned")])
def test_tail_split_same_bits(monkeypatch, force_mode, controller):
    """The split tail (ABIE_HIP_TAIL_SPLIT=1, the default): once no more than
    an eighth of the CUs' count of units run -- here from the first launch,
    these nine systems make about 18 units -- each unit with enough test
    particles runs as children that carry all its massive bodies and a share
    of its test particles, on CUs of their own, and is merged back at the
    end.  Identical bits, times, statuses and counters with the split off,
    also when the children stop every 7 attempts and catch up with each
    other before the merge."""
    import os
    w = None
    if controller == "learned":
        w = np.load(os.path.join(os.path.dirname(__file__), "..", "abie_amd", "data", "proposer.npz"))["weights"]
    systems = _systems()
    monkeypatch.setenv("ABIE_HIP_TAIL_SPLIT", "0")
    ref = _run(systems, 0.2, force_mode, controller, proposer=w)
    monkeypatch.setenv("ABIE_HIP_TAIL_SPLIT", "1")
    for chunk in (None, "7"):
        if chunk:
            monkeypatch.setenv("ABIE_HIP_MAX_ATTEMPTS", chunk)
        r = _run(systems, 0.2, force_mode, controller, proposer=w)
        assert r["done"] and np.array_equal(r["t"], ref["t"]) and np.array_equal(r["st"], ref["st"])
        assert same_bits(r["p"], ref["p"]) and same_bits(r["v"], ref["v"]), chunk
        assert np.array_equal(r["cnt"], ref["cnt"]), chunk


@pytest.mark.parametrize("force_mode,controller", [("fp64", "massive"), ("mixed", "learned")])
def test_promoted_stragglers_same_bits(monkeypatch, force_mode, controller):
    """Straggler promotion during the one-lane phase (more units than CUs):
    the run proceeds in rounds, and units projected to finish late move to
    the quad form in a tail set of their own (split by the rows rule) for the
    rest of the run.  Forced here with 5-ms rounds and every unit above the
    median promoted (up to a quarter of the CUs): identical bits, times,
    statuses and counters with promotion off."""
    import os
    from abie_amd import ics
    w = None
    if controller == "learned":
        w = np.load(os.path.join(os.path.dirname(__file__), "..", "abie_amd", "data", "proposer.npz"))["weights"]
    rng = np.random.default_rng(5)
    systems = [ics.planetary_system(5200 + k, int(rng.integers(3, 40)), int(rng.integers(0, 300))) for k in range(300)]
    monkeypatch.setenv("ABIE_HIP_PROMOTE", "0")
    ref = _run(systems, 0.3, force_mode, controller, proposer=w)
    monkeypatch.setenv("ABIE_HIP_PROMOTE", "1")
    monkeypatch.setenv("ABIE_HIP_PROMOTE_FACTOR", "1.0")
    monkeypatch.setenv("ABIE_HIP_ROUND_SECONDS", "0.005")
    r = _run(systems, 0.3, force_mode, controller, proposer=w)
    assert ref["done"] and r["done"]
    assert np.array_equal(r["t"], ref["t"]) and np.array_equal(r["st"], ref["st"])
    assert same_bits(r["p"], ref["p"]) and same_bits(r["v"], ref["v"])
    assert np.array_equal(r["cnt"], ref["cnt"])


@pytest.mark.parametrize("force_mode,controller", [("fp64", "massive"), ("mixed", "learned")])
def test_start_order_and_probe_round_same_bits(monkeypatch, force_mode, controller):
    """The start order of the queued units (by size, or by predicted work
    from the initial orbital time scales) and the optional probe round (a
    first round of a few attempts per unit, then the queue re-sorted by
    measured rates) change when each unit runs, never what it computes:
    1,200 small systems (more one-lane units than the GPU holds at once, so
    the probe round runs), identical bits, times, statuses and counters."""
    import os
    from abie_amd import ics
    w = None
    if controller == "learned":
        w = np.load(os.path.join(os.path.dirname(__file__), "..", "abie_amd", "data", "proposer.npz"))["weights"]
    rng = np.random.default_rng(11)
    systems = [ics.planetary_system(7100 + k, int(rng.integers(3, 12)), int(rng.integers(0, 120))) for k in range(1200)]
    monkeypatch.setenv("ABIE_HIP_ORDER", "0")
    ref = _run(systems, 0.1, force_mode, controller, proposer=w)
    monkeypatch.setenv("ABIE_HIP_ORDER", "1")
    r1 = _run(systems, 0.1, force_mode, controller, proposer=w)
    monkeypatch.setenv("ABIE_HIP_PROBE_ATTEMPTS", "4")
    r2 = _run(systems, 0.1, force_mode, controller, proposer=w)
    assert ref["done"]
    for r in (r1, r2):
        assert r["done"]
        assert np.array_equal(r["t"], ref["t"]) and np.array_equal(r["st"], ref["st"])
        assert same_bits(r["p"], ref["p"]) and same_bits(r["v"], ref["v"])
        assert np.array_equal(r["cnt"], ref["cnt"])
