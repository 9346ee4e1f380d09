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
    def compute_energy(sol_state, masses, const_g):
        """Total energy of every row of ``sol_state`` = [positions (3N) |
        velocities (3N)] (abie/tools.py:5-33), vectorised over rows; None for
        a state whose width is not a multiple of 6 (as the reference)."""
        sol_state = np.atleast_2d(np.asarray(sol_state, dtype=np.float64))
        n_rows, n_cols = sol_state.shape
        if n_cols % 6 != 0 or n_rows == 0:
            return None
        n = n_cols // 6
        m = np.asarray(masses, dtype=np.float64)[:n]
        pos = sol_state[:, :3 * n].reshape(n_rows, n, 3)
        vel = sol_state[:, 3 * n:].reshape(n_rows, n, 3)
        ke = 0.5 * np.sum(m * np.sum(vel * vel, axis=2), axis=1)
        d = np.linalg.norm(pos[:, :, None, :] - pos[:, None, :, :], axis=3)
        with np.errstate(divide="ignore"):
            inv = np.where(np.eye(n, dtype=bool), 0.0, 1.0 / d)
        return ke - 0.5 * const_g * np.einsum("i,j,tij->t", m, m, inv)

    @staticmethod
    def from_cartesian_to_aei(mp, ms, position, velocity, G=0.0172020989 ** 2, z_vector=np.array([0.0, 0.0, 1.0])):
        """Semi-major axis, eccentricity and inclination (abie/tools.py:235-318),
        with the reference's default G (the Gaussian constant squared: au, day)
        and its layouts: one body as 3-vectors, or many as (3, n) arrays (one
        column per body); a body at r = 0 gives NaNs."""
        position = np.asarray(position, dtype=np.float64)
        velocity = np.asarray(velocity, dtype=np.float64)
        mu = G * (mp + ms)
        if position.ndim == 1 and velocity.ndim == 1:
            r = np.sqrt(np.dot(position, position))
            if r == 0:
                return np.nan, np.nan, np.nan
            pos, vel = position[:, None], velocity[:, None]
        else:
            pos, vel = position, velocity
        v2 = np.sum(vel * vel, axis=0)
        r = np.sqrt(np.sum(pos * pos, axis=0))
        with np.errstate(divide="ignore", invalid="ignore"):
            a = -mu / (2.0 * (0.5 * v2 - mu / r))
            h = np.cross(pos.T, vel.T).T
            hunit = h / np.sqrt(np.sum(h * h, axis=0))
            evec = ((v2 - mu / r) * pos - np.sum(pos * vel, axis=0) * vel) / mu
            e = np.linalg.norm(evec, axis=0)
            inc = np.arccos(np.matmul(z_vector, hunit))
        if position.ndim == 1 and velocity.ndim == 1:
            return float(a[0]), float(e[0]), float(inc[0])
        return a, e, inc

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
