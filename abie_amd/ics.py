"""Synthetic initial conditions for the benchmark and test workloads.

Units are the reference's: Msun, au, yr, G = 4 pi^2 (``abie/__init__.py:24``),
except the Plummer sphere, which uses Henon units (G = 1, M = 1).

Orbital elements are converted with the textbook two-body relations (the same
relations as ``abie/tools.py:36-111``): perifocal unit vectors P, Q rotated by
(Omega, i, omega), r = a(1-e^2)/(1+e cos f), v = sqrt(G(m_p+m_s)/(a(1-e^2))).
"""
import numpy as np

G_SOLAR = 4.0 * np.pi ** 2
DEFAULT_SEED = 20261203


def kepler_to_cartesian(mp, ms, a, e, inc, Omega, omega, f, G=G_SOLAR):
    """Vectorised elements -> (pos, vel) relative to the primary, shape (n, 3)."""
    a, e, inc, Omega, omega, f = (np.atleast_1d(np.asarray(v, dtype=np.float64))
                                  for v in (a, e, inc, Omega, omega, f))
    mp = np.broadcast_to(np.asarray(mp, dtype=np.float64), a.shape)
    cO, sO = np.cos(Omega), np.sin(Omega)
    co, so = np.cos(omega), np.sin(omega)
    ci, si = np.cos(inc), np.sin(inc)
    P = np.stack([cO * co - sO * so * ci, sO * co + cO * so * ci, so * si], axis=1)
    Q = np.stack([-cO * so - sO * co * ci, -sO * so + cO * co * ci, co * si], axis=1)
    p = a * (1.0 - e ** 2)
    r = p / (1.0 + e * np.cos(f))
    pos = (r * np.cos(f))[:, None] * P + (r * np.sin(f))[:, None] * Q
    vt = np.sqrt(G * (mp + ms) / p)
    vel = (-vt * np.sin(f))[:, None] * P + (vt * (e + np.cos(f)))[:, None] * Q
    return pos, vel


def _elements(rng, n, a_range, a_log, e_max, i_max):
    lo, hi = a_range
    if a_log:
        a = np.exp(rng.uniform(np.log(lo), np.log(hi), n))
    else:
        a = rng.uniform(lo, hi, n)
    e = rng.uniform(0.0, e_max, n)
    inc = rng.uniform(0.0, i_max, n)
    ang = rng.uniform(0.0, 2 * np.pi, (3, n))
    return a, e, inc, ang[0], ang[1], ang[2]


def planetary_system(seed, n_massive, n_test, planet_mass=(1e-7, 1e-3),
                     planet_a=(0.5, 30.0), planet_e=0.1, planet_i=0.05,
                     tp_a=(1.0, 50.0), tp_e=0.2, tp_i=0.1, G=G_SOLAR):
    """Star (1 Msun, at rest at the origin) + (n_massive-1) planets + test
    particles, elements relative to the star.  Returns pos, vel (n, 3), m (n,)."""
    rng = np.random.default_rng(seed)
    npl = n_massive - 1
    lo, hi = planet_mass
    mpl = np.exp(rng.uniform(np.log(lo), np.log(hi), npl))
    el = _elements(rng, npl, planet_a, True, planet_e, planet_i)
    ppos, pvel = kepler_to_cartesian(mpl, 1.0, *el, G=G)
    elt = _elements(rng, n_test, tp_a, False, tp_e, tp_i)
    tpos, tvel = kepler_to_cartesian(0.0, 1.0, *elt, G=G)
    pos = np.concatenate([np.zeros((1, 3)), ppos, tpos])
    vel = np.concatenate([np.zeros((1, 3)), pvel, tvel])
    m = np.concatenate([[1.0], mpl, np.zeros(n_test)])
    return pos, vel, m


def config0(seed=DEFAULT_SEED, n_test=1000):
    """1 star + 4 planets + n_test test particles (100-yr GR15 workload)."""
    return planetary_system(seed, 5, n_test)


def config1a(seed=DEFAULT_SEED, n_test=1 << 20):
    """64 massive (star + 63 bodies, m log-U[1e-9, 1e-3]) + n_test test
    particles with a ~ U[1, 100] au: the force-kernel workload."""
    return planetary_system(seed, 64, n_test, planet_mass=(1e-9, 1e-3), tp_a=(1.0, 100.0))


def plummer(n, seed=DEFAULT_SEED):
    """Plummer sphere, Aarseth, Henon & Wielen (1974) sampling, Henon units
    (G = 1, total mass 1, E = -1/4), equal masses 1/n, centred on the COM."""
    rng = np.random.default_rng(seed)
    m = np.full(n, 1.0 / n)
    x1 = rng.uniform(1e-10, 1.0, n)
    r = 1.0 / np.sqrt(x1 ** (-2.0 / 3.0) - 1.0)
    cth = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2 * np.pi, n)
    sth = np.sqrt(1.0 - cth ** 2)
    pos = r[:, None] * np.stack([sth * np.cos(phi), sth * np.sin(phi), cth], axis=1)
    # speed by von Neumann rejection on g(q) = q^2 (1 - q^2)^(7/2)
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        x4 = rng.uniform(0.0, 1.0, todo.size)
        x5 = rng.uniform(0.0, 0.1, todo.size)
        ok = x5 < x4 ** 2 * (1.0 - x4 ** 2) ** 3.5
        q[todo[ok]] = x4[ok]
        todo = todo[~ok]
    ve = np.sqrt(2.0) * (1.0 + r ** 2) ** (-0.25)
    v = q * ve
    cth = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2 * np.pi, n)
    sth = np.sqrt(1.0 - cth ** 2)
    vel = v[:, None] * np.stack([sth * np.cos(phi), sth * np.sin(phi), cth], axis=1)
    scale = 3.0 * np.pi / 16.0
    pos *= scale
    vel /= np.sqrt(scale)
    pos -= (m[:, None] * pos).sum(0) / m.sum()
    vel -= (m[:, None] * vel).sum(0) / m.sum()
    return pos, vel, m


def ensemble(n_systems, seed=DEFAULT_SEED, massive=(3, 64), test=(0, 2048)):
    """Independent planetary systems of uneven size: n_massive ~ U{3..64}
    (star included), n_test ~ U{0..2048}, elements as in config0, per-system
    seed = seed + k.  Returns a list of (pos, vel, m)."""
    out = []
    for k in range(n_systems):
        rng = np.random.default_rng(seed + k)
        nm = int(rng.integers(massive[0], massive[1] + 1))
        nt = int(rng.integers(test[0], test[1] + 1))
        out.append(planetary_system(seed + k + 1_000_000_007, nm, nt))
    return out
