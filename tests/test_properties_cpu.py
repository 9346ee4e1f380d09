"""Property tests (hypothesis) of the oracle's force restatement
(src/common.c:39-77): invariants that a dropped term, a wrong sign or index,
or a transposed operand would break, over random small systems with massive
bodies, test particles and deleted bodies."""
import numpy as np
import pytest
from hypothesis import given, settings, strategies as st, HealthCheck

G = 4 * np.pi ** 2


@st.composite
def systems(draw, max_n=12):
    n = draw(st.integers(2, max_n))
    seed = draw(st.integers(0, 2 ** 31 - 1))
    rng = np.random.default_rng(seed)
    pos = rng.normal(size=(n, 3)) * draw(st.sampled_from([0.01, 1.0, 30.0]))
    kind = rng.choice([0, 1, 2], size=n, p=[0.6, 0.3, 0.1])  # massive, test particle, deleted
    m = np.where(kind == 0, np.exp(rng.uniform(np.log(1e-9), 0.0, n)), np.where(kind == 1, 0.0, -1.0))
    return pos, m


common = settings(max_examples=60, deadline=None, suppress_health_check=[HealthCheck.too_slow])


@common
@given(systems())
def test_permutation_equivariance(orc, sysm):
    pos, m = sysm
    perm = np.random.default_rng(len(m)).permutation(len(m))
    a = orc.accel(pos, G, m)
    b = orc.accel(pos[perm], G, m[perm])
    ok = m[perm] >= 0
    # same pairs, different summation order: equal to rounding
    s = np.abs(a[perm][ok]).max() if ok.any() else 1.0
    assert np.allclose(b[ok], a[perm][ok], rtol=1e-12, atol=1e-12 * s)


@common
@given(systems(), st.sampled_from([0.5, 2.0, 8.0]))
def test_inverse_square_scaling_and_linearity_in_G(orc, sysm, lam):
    pos, m = sysm
    a = orc.accel(pos, G, m)
    # powers of two scale exactly: a(lam x) = a(x) / lam^2 bit for bit
    b = orc.accel(pos * lam, G, m)
    assert np.array_equal(b * lam * lam, a)
    c = orc.accel(pos, 2.0 * G, m)
    assert np.array_equal(c, 2.0 * a)


@common
@given(systems())
def test_test_particles_do_not_act_and_momentum_is_conserved(orc, sysm):
    pos, m = sysm
    a = orc.accel(pos, G, m)
    massive = m > 0
    # dropping every test particle leaves the massive accelerations unchanged
    keep = m != 0
    a2 = orc.accel(pos[keep], G, m[keep])
    assert np.array_equal(a2[m[keep] > 0], a[massive])
    # Newton's third law among the massive bodies
    if massive.sum() >= 2:
        f = (m[massive][:, None] * a[massive]).sum(0)
        scale = (np.abs(m[massive][:, None] * a[massive])).sum()
        assert np.abs(f).max() <= 1e-12 * scale
    # deleted sinks are left untouched (zero here)
    assert np.all(a[m < 0] == 0.0)


def test_oracle_gr15_deleted_body_feels_only_the_external_acceleration(orc):
    """Inside GR15 a deleted body (m < 0) is neither source nor sink; with an
    external acceleration set it moves under that constant acceleration
    alone: x(t) = x0 + v0 t + a t^2 / 2 (GR15 integrates a constant
    acceleration exactly up to rounding).  (Its components still enter the
    controller's norms, as every body's do in the reference.)"""
    from abie_amd import ics
    pos, vel, m = ics.planetary_system(41, 4, 6)
    m = m.copy()
    m[2] = -1.0
    ext = np.zeros((len(m), 3))
    ext[2] = [3e-3, -2e-3, 1e-3]
    s = orc.Sim(pos, vel, m, ext_acc=ext)
    s.integrate(1.0, 0.0)
    t = s.t
    want = pos[2] + vel[2] * t + 0.5 * ext[2] * t * t
    got = s.pos.reshape(-1, 3)[2]
    assert np.allclose(got, want, rtol=1e-12, atol=1e-13)
