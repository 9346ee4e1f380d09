"""Two-body element conversions (textbook relations; the reference's are
abie/tools.py:36-230), vectorised over bodies."""
import numpy as np

from .ics import kepler_to_cartesian

G_SOLAR = 4.0 * np.pi ** 2


def elements_to_cartesian(mp, ms, a, e, i, Omega, omega, f, G=G_SOLAR):
    """(pos, vel) relative to the primary, shape (n, 3)."""
    return kepler_to_cartesian(mp, ms, a, e, i, Omega, omega, f, G=G)


def cartesian_to_elements(mp, ms, pos, vel, G=G_SOLAR):
    """Arrays a, e, i, Omega, omega, f of relative states (n, 3).

    Circular orbits (e == 0) return omega = 0 and f measured from the node;
    equatorial orbits take the node along +x."""
    pos = np.atleast_2d(np.asarray(pos, dtype=np.float64))
    vel = np.atleast_2d(np.asarray(vel, dtype=np.float64))
    mu = G * (np.asarray(mp, dtype=np.float64) + np.asarray(ms, dtype=np.float64))
    mu = np.broadcast_to(mu, (len(pos),))
    r = np.linalg.norm(pos, axis=1)
    v2 = np.sum(vel * vel, axis=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        a = 1.0 / (2.0 / r - v2 / mu)
        hvec = np.cross(pos, vel)
        h = np.linalg.norm(hvec, axis=1)
        inc = np.arccos(np.clip(hvec[:, 2] / h, -1.0, 1.0))
        nvec = np.stack([-hvec[:, 1], hvec[:, 0], np.zeros(len(pos))], axis=1)
        nn = np.linalg.norm(nvec, axis=1)
        evec = np.cross(vel, hvec) / mu[:, None] - pos / r[:, None]
        e = np.linalg.norm(evec, axis=1)
        nunit = np.where(nn[:, None] > 0, nvec / np.where(nn > 0, nn, 1)[:, None], np.array([1.0, 0.0, 0.0]))
        Omega = np.arctan2(nunit[:, 1], nunit[:, 0])
        hunit = hvec / h[:, None]
        eunit = np.where(e[:, None] > 1e-15, evec / np.where(e > 0, e, 1)[:, None], nunit)
        omega = np.arctan2(np.sum(hunit * np.cross(nunit, eunit), axis=1), np.sum(nunit * eunit, axis=1))
        runit = pos / r[:, None]
        f = np.arctan2(np.sum(hunit * np.cross(eunit, runit), axis=1), np.sum(eunit * runit, axis=1))
    return a, e, inc, Omega, omega, f


class Tools(object):
    """Reference-style names (abie/tools.py)."""

    @staticmethod
    def from_orbital_elements_to_cartesian(mp, ms, semimajor_axis, eccentricity, true_anomaly, inclination,
                                           argument_of_periapsis, longitude_of_ascending_node, G=G_SOLAR):
        p, v = elements_to_cartesian(mp, ms, semimajor_axis, eccentricity, inclination,
                                     longitude_of_ascending_node, argument_of_periapsis, true_anomaly, G=G)
        return p[0], v[0]

    @staticmethod
    def from_cartesian_to_orbital_elements(mp, ms, position, velocity, G=G_SOLAR):
        a, e, i, Om, om, f = cartesian_to_elements(mp, ms, position, velocity, G)
        return a[0], e[0], f[0], i[0], om[0], Om[0]
