"""Ensemble engine beyond one workgroup: test-particle slicing under the
massive-only controller, and the plumbing of the learned proposer."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = 4 * np.pi ** 2


def same_bits(a, b):
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    return bool(np.all(a.view(np.int64) == b.view(np.int64)))


def test_sliced_system_matches_unsliced_subsets():
    """A 5 + 3,000 body system is split into slices that each carry the
    massive bodies.  Under the massive-only controller test particles never
    influence the step, so (i) the massive bodies evolve exactly as in a run
    without test particles and (ii) any subset of test particles evolves
    exactly as in a run with only that subset."""
    from abie_amd import engine, ics
    pos, vel, m = ics.planetary_system(77, 5, 3000)
    big = engine.Ensemble([(pos, vel, m)], force_mode="fp64", controller="massive")
    big.run(0.5)
    pb, vb, tb, _ = big.state()
    sub = np.concatenate([np.arange(5), 5 + np.array([0, 1, 999, 1500, 2999])])
    small = engine.Ensemble([(pos[:5], vel[:5], m[:5]), (pos[sub], vel[sub], m[sub])], force_mode="fp64",
                            controller="massive")
    small.run(0.5)
    ps, vs, ts, _ = small.state()
    assert tb[0] == ts[0] == ts[1]
    assert same_bits(pb[:5], ps[:5]) and same_bits(vb[:5], vs[:5])
    assert same_bits(pb[sub], ps[5:]) and same_bits(vb[sub], vs[5:])


def test_plain_controller_rejects_unlinkable_system():
    """Under the plain controller a system larger than one workgroup runs as
    linked slices of test particles; 3,000 massive bodies (no test particles
    to slice) are refused with a clear error."""
    from abie_amd import engine, ics, _lib
    pos, vel, m = ics.planetary_system(3, 3000, 0)
    with pytest.raises(_lib.AbieHipError):
        engine.Ensemble([(pos, vel, m)], controller="plain")


def test_learned_with_neutral_proposer_equals_massive():
    """With a proposer that always predicts ratio 1 the learned controller
    proposes dtreq, which is the plain rule away from its clamps."""
    from abie_amd import engine, ics, _lib
    systems = [ics.planetary_system(500 + k, 4, 30) for k in range(3)]
    a = engine.Ensemble(systems, force_mode="fp64", controller="massive")
    b = engine.Ensemble(systems, force_mode="fp64", controller="learned")
    b.set_proposer(np.zeros(_lib.MLP_NPARAM, dtype=np.float32))
    a.run(0.3)
    b.run(0.3)
    pa, va, ta, _ = a.state()
    pb, vb, tb, _ = b.state()
    ca, cb = a.counters(), b.counters()
    assert np.array_equal(ta, tb) and same_bits(pa, pb)
    assert np.all(cb[:, 4] > 0)  # the proposer was consulted
    assert np.array_equal(ca[:, :3], cb[:, :3])


def test_steplog_records_every_attempt():
    from abie_amd import engine, ics
    systems = [ics.planetary_system(900 + k, 3, 10) for k in range(2)]
    e = engine.Ensemble(systems, controller="plain")
    e.set_steplog(100000)
    e.run(0.2)
    rec = e.steplog()
    cnt = e.counters()
    for s in range(2):
        rs = rec[rec[:, 0] == s]
        assert len(rs) == cnt[s, 1]
        assert int(rs[:, 6].sum()) == cnt[s, 2]
        assert np.all(np.diff(rs[:, 1]) >= 0)


def test_skipped_event_scan_changes_nothing(monkeypatch):
    """Without event rings the engine skips the close-encounter scan; forcing
    the scan on gives the same trajectories bit for bit."""
    from abie_amd import engine, ics
    systems = [ics.planetary_system(620 + k, 4 + k, 40 * k) for k in range(3)]
    a = engine.Ensemble(systems, controller="plain")
    a.run(0.3)
    monkeypatch.setenv("ABIE_HIP_EVENTS", "1")
    b = engine.Ensemble(systems, controller="plain")
    b.run(0.3)
    pa, va, ta, _ = a.state()
    pb, vb, tb, _ = b.state()
    assert np.array_equal(ta, tb) and same_bits(pa, pb) and same_bits(va, vb)


@pytest.mark.parametrize("force_mode,controller", [("fp64", "massive"), ("mixed", "massive"), ("fp64", "learned")])
def test_prefetching_source_loop_same_bits(monkeypatch, force_mode, controller):
    """Units of more than one body per thread (systems of more than 512
    bodies that cannot be sliced: 211 and 220 massive bodies) sum their sources with the
    prefetching loop under the massive-only controllers; with it on for
    every unit and off for all, the same bits (odd and even source
    counts)."""
    from abie_amd import engine, ics
    systems = [ics.planetary_system(800 + k, nm, nt) for k, (nm, nt) in
               enumerate([(220, 400), (211, 500), (2, 40), (31, 200)])]
    w = None
    if controller == "learned":
        import os
        w = np.load(os.path.join(os.path.dirname(__file__), "..", "abie_amd", "data", "proposer.npz"))["weights"]
    out = []
    for flag in ("2", "0"):  # from 2 massive bodies on, and never
        monkeypatch.setenv("ABIE_HIP_PREFETCH", flag)
        e = engine.Ensemble(systems, force_mode=force_mode, controller=controller)
        if w is not None:
            e.set_proposer(w)
        e.run(0.2)
        out.append(e.state())
        e.close()
    (pa, va, ta, _), (pb, vb, tb, _) = out
    assert np.array_equal(ta, tb) and same_bits(pa, pb) and same_bits(va, vb)
