"""Drop-in facade pieces that need no GPU: particles, element conversions,
plugin discovery, snapshot buffer."""
import numpy as np
import pytest

G = 4 * np.pi ** 2


def test_elements_round_trip_and_invariants():
    from abie_amd.tools import cartesian_to_elements, elements_to_cartesian
    rng = np.random.default_rng(0)
    n = 200
    a = rng.uniform(0.5, 50, n); e = rng.uniform(0.01, 0.8, n); i = rng.uniform(0.01, 3.0, n)
    Om = rng.uniform(-3, 3, n); om = rng.uniform(-3, 3, n); f = rng.uniform(-3, 3, n)
    mp = rng.uniform(0, 1e-3, n)
    pos, vel = elements_to_cartesian(mp, 1.0, a, e, i, Om, om, f, G=G)
    mu = G * (1.0 + mp)
    r = np.linalg.norm(pos, axis=1); v2 = (vel ** 2).sum(1)
    # vis-viva and |h|^2 = mu a (1 - e^2), periapsis distance
    np.testing.assert_allclose(v2, mu * (2 / r - 1 / a), rtol=1e-12)
    h2 = (np.cross(pos, vel) ** 2).sum(1)
    np.testing.assert_allclose(h2, mu * a * (1 - e ** 2), rtol=1e-12)
    a2, e2, i2, Om2, om2, f2 = cartesian_to_elements(mp, 1.0, pos, vel, G)
    np.testing.assert_allclose(a2, a, rtol=1e-10)
    np.testing.assert_allclose(e2, e, rtol=1e-9)
    np.testing.assert_allclose(i2, i, rtol=1e-10)
    wrap = lambda x: (x + np.pi) % (2 * np.pi) - np.pi
    assert np.abs(wrap(Om2 - Om)).max() < 1e-9 and np.abs(wrap(om2 - om)).max() < 1e-8
    assert np.abs(wrap(f2 - f)).max() < 1e-8


def test_particles_add_primary_com_and_merge():
    from abie_amd.particles import Particles
    p = Particles(G)
    p.add(mass=1.0, name="Sun")
    p.add(mass=1e-3, a=1.0, e=0.0, f=0.0, name="P")
    # default primary: COM of the bodies added before (here the Sun alone)
    assert np.allclose(p["P"].pos, [1.0, 0, 0])
    p.add(mass=0.0, a=2.0, f=0.0)
    com = (1.0 * np.zeros(3) + 1e-3 * np.array([1.0, 0, 0])) / 1.001
    assert np.allclose(p[2].pos, com + [2.0, 0, 0])
    assert p.N == 3 and p.positions.shape == (9,)
    p.add(mass=2e-3, pos=[1.0, 0.0, 0.0], vel=[0, 1.0, 0], radius=0.2, name="Q")
    v_before = (1e-3 * p[1].vel + 2e-3 * p[3].vel) / 3e-3
    p.merge_particles_inelastically(1, 3)
    assert p.N == 3
    assert np.allclose(p[2].vel, v_before) and p.masses[2] == pytest.approx(3e-3)


def test_plugins_keep_reference_names():
    from abie_amd.integrator import Integrator
    names = set(Integrator.load_integrators())
    assert names >= {"GaussRadau15", "WisdomHolman", "RungeKutta", "AdamsBashforth", "LeapFrog", "Euler"}


def test_tools_energy_and_aei_closed_forms():
    """Tools.compute_energy on a two-body circular orbit (E = -G m1 m2 / 2a)
    and Tools.from_cartesian_to_aei on known orbits, one body and (3, n)."""
    from abie_amd.tools import Tools
    G = 4 * np.pi ** 2
    m1, m2, a = 1.0, 1e-3, 2.0
    v = np.sqrt(G * (m1 + m2) / a)
    pos = np.array([[-a * m2 / (m1 + m2), 0, 0], [a * m1 / (m1 + m2), 0, 0]])
    vel = np.array([[0, -v * m2 / (m1 + m2), 0], [0, v * m1 / (m1 + m2), 0]])
    row = np.concatenate([pos.reshape(-1), vel.reshape(-1)])
    e = Tools.compute_energy(np.stack([row, row]), [m1, m2], G)
    assert np.allclose(e, -G * m1 * m2 / (2 * a), rtol=1e-13, atol=0)
    assert Tools.compute_energy(np.zeros((2, 5)), [1.0], G) is None
    # circular orbit of radius 3 inclined by 0.4 rad, and an eccentric one
    r0, inc = 3.0, 0.4
    vc = np.sqrt(G / r0)
    p1 = np.array([r0, 0.0, 0.0])
    v1 = np.array([0.0, vc * np.cos(inc), vc * np.sin(inc)])
    aa, ee, ii = Tools.from_cartesian_to_aei(0.0, 1.0, p1, v1, G=G)
    assert abs(aa - r0) < 1e-12 and ee < 1e-12 and abs(ii - inc) < 1e-12
    ecc, ap = 0.3, 2.0
    p2 = np.array([0.0, ap * (1 - ecc), 0.0])
    v2 = np.array([-np.sqrt(G / ap * (1 + ecc) / (1 - ecc)), 0.0, 0.0])
    A, E, I = Tools.from_cartesian_to_aei(0.0, 1.0, np.stack([p1, p2], 1), np.stack([v1, v2], 1), G=G)
    assert np.allclose(A, [r0, ap], rtol=1e-12) and np.allclose(E, [0, ecc], atol=1e-12)
    assert np.allclose(I, [inc, 0.0], atol=1e-12)
    assert all(np.isnan(Tools.from_cartesian_to_aei(0.0, 1.0, np.zeros(3), v1, G=G)))


def test_particle_module_and_add_particle():
    from abie_amd.particle import Particle
    from abie_amd.particles import Particles
    ps = Particles(4 * np.pi ** 2)
    ps.add(mass=1.0, name="Sun")
    ps.add_particle(Particle(mass=1e-3, pos=np.array([5.2, 0.0, 0.0]), vel=np.array([0.0, 2.75, 0.0]),
                             name="Jupiter", radius=4.7e-4))
    assert ps.N == 2 and ps.masses[1] == 1e-3 and ps.radii[1] == 4.7e-4
    assert np.array_equal(ps.positions.reshape(-1, 3)[1], [5.2, 0.0, 0.0])
    with pytest.raises(TypeError):
        ps.add_particle("not a particle")


def test_out_of_path_integrators_raise_clearly():
    from abie_amd.clibabie import CLibABIE
    c = CLibABIE()
    x = np.zeros(6)
    with pytest.raises(NotImplementedError, match="RungeKutta"):
        c.integrator_runge_kutta(x.copy(), x.copy(), np.ones(2), 2, 1.0, 0.0, 1.0, 0.1)
    with pytest.raises(NotImplementedError, match="WisdomHolman"):
        c.integrator_wisdom_holman(x.copy(), x.copy(), np.ones(2), 2, 1.0, 0.0, 1.0, 0.1)
