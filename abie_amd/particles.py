"""Particle container of the drop-in facade (abie/particles.py, abie/particle.py
of the reference), stored as contiguous arrays so that adding N bodies is
O(N) and the orbital-element queries are vectorised.

Semantics kept from the reference:
  * ``add`` with ``a`` given converts elements relative to ``primary``; the
    default primary is the centre of mass of the bodies added so far
    (abie/particles.py:143-161, 378-397);
  * masses 0 are test particles; m < 0 marks a deleted body;
  * ``merge_particles_inelastically`` conserves momentum, combines radii as
    (r1^3 + r2^3)^(1/3) and removes the lighter body (abie/particles.py:215-297).
"""
import numpy as np

from .tools import cartesian_to_elements, elements_to_cartesian


class Particle(object):
    """A view of one body (abie/particle.py:4-79)."""

    def __init__(self, ptype=0, mass=0.0, pos=None, vel=None, radius=0.0, name=None, primary=None):
        self.ptype = ptype
        self.mass = float(mass)
        self.radius = float(radius)
        self.name = name
        self.hash = int(np.random.randint(100000000, 999999999))
        self.primary = primary
        self.pos = np.zeros(3) if pos is None else np.asarray(pos, dtype=np.float64).reshape(3)
        self.vel = np.zeros(3) if vel is None else np.asarray(vel, dtype=np.float64).reshape(3)

    x = property(lambda s: s.pos[0])
    y = property(lambda s: s.pos[1])
    z = property(lambda s: s.pos[2])
    vx = property(lambda s: s.vel[0])
    vy = property(lambda s: s.vel[1])
    vz = property(lambda s: s.vel[2])

    def __repr__(self):
        return ("Particle(m=%g, x=%g, y=%g, z=%g, vx=%g, vy=%g, vz=%g, r=%g, name='%s', hash=%d)"
                % (self.mass, self.x, self.y, self.z, self.vx, self.vy, self.vz, self.radius, self.name, self.hash))


class Particles(object):
    def __init__(self, const_g):
        self.CONST_G = const_g
        self.primary = "#COM#"  # '#COM#', '#M_MAX#', '#M_MIN#', a name/ID, or a list of IDs
        self._pos = np.zeros((0, 3))
        self._vel = np.zeros((0, 3))
        self._m = np.zeros(0)
        self._r = np.zeros(0)
        self._ptype = np.zeros(0, dtype=np.int64)
        self._hash = np.zeros(0, dtype=np.int64)
        self._primary = []
        self._names = {}
        self._n = 0

    def _grow(self, k):
        need = self._n + k
        if need <= len(self._m):
            return
        cap = max(need, 2 * len(self._m), 16)
        for name, shape in (("_pos", (cap, 3)), ("_vel", (cap, 3)), ("_m", (cap,)), ("_r", (cap,)),
                            ("_ptype", (cap,)), ("_hash", (cap,))):
            old = getattr(self, name)
            new = np.zeros(shape, dtype=old.dtype)
            new[: self._n] = old[: self._n]
            setattr(self, name, new)

    # ------------------------------------------------------------ views
    @property
    def N(self):
        return self._n

    @property
    def particles(self):
        return self

    @property
    def positions(self):
        return self._pos[: self._n].reshape(-1)

    @positions.setter
    def positions(self, p):
        self._pos[: self._n] = np.asarray(p, dtype=np.float64).reshape(-1, 3)

    @property
    def velocities(self):
        return self._vel[: self._n].reshape(-1)

    @velocities.setter
    def velocities(self, v):
        self._vel[: self._n] = np.asarray(v, dtype=np.float64).reshape(-1, 3)

    @property
    def masses(self):
        return self._m[: self._n]

    @property
    def radii(self):
        return self._r[: self._n]

    @property
    def ptypes(self):
        return self._ptype[: self._n]

    @property
    def hashes(self):
        return self._hash[: self._n]

    @property
    def names(self):
        inv = {v: k for k, v in self._names.items()}
        return [inv.get(i) for i in range(self._n)]

    def __len__(self):
        return self._n

    def __getitem__(self, item):
        if isinstance(item, (int, np.integer)):
            if not 0 <= item < self._n:
                raise ValueError("Particle #%d does not exist!" % item)
            i = int(item)
            p = Particle(ptype=int(self._ptype[i]), mass=self._m[i], pos=self._pos[i], vel=self._vel[i],
                         radius=self._r[i], name=self.names[i], primary=self._primary[i])
            p.hash = int(self._hash[i])
            return p
        if isinstance(item, str):
            if item not in self._names:
                raise ValueError("Particle %s not exist!" % item)
            return self[self._names[item]]
        return None

    def __repr__(self):
        return "Total number of particles: %d\n" % self._n + "\n".join(repr(self[i]) for i in range(self._n))

    # ------------------------------------------------------------ adding
    def add(self, pos=None, vel=None, mass=0.0, name=None, radius=0.0, ptype=0, a=None, e=0.0, i=0.0,
            Omega=0.0, omega=0.0, f=0.0, primary=None):
        """Add one body by Cartesian state or by orbital elements (a given)."""
        pos = np.zeros(3) if pos is None else np.asarray(pos, dtype=np.float64).reshape(3)
        vel = np.zeros(3) if vel is None else np.asarray(vel, dtype=np.float64).reshape(3)
        if a is not None:
            prim = self.determine_primary_body(primary)
            dp, dv = elements_to_cartesian(mass, prim.mass, a, e, i, Omega, omega, f, G=self.CONST_G)
            pos = dp[0] + prim.pos
            vel = dv[0] + prim.vel
        self._grow(1)
        k = self._n
        self._pos[k] = pos
        self._vel[k] = vel
        self._m[k] = mass
        self._r[k] = radius
        self._ptype[k] = ptype
        self._hash[k] = np.random.randint(100000000, 999999999)
        self._primary.append(primary)
        if name is not None:
            self._names[name] = k
        self._n += 1
        return k

    def add_particle(self, particle):
        """Add a Particle object (abie/particles.py:181-198); anything else is
        a TypeError, as in the reference."""
        if not isinstance(particle, Particle):
            raise TypeError("Incompatible particle type.")
        return self.add(pos=particle.pos, vel=particle.vel, mass=particle.mass, name=particle.name,
                        radius=particle.radius, ptype=particle.ptype, primary=particle.primary)

    def add_many(self, pos, vel, masses, radii=None):
        """Bulk add of Cartesian states (n, 3) -- O(n)."""
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        vel = np.asarray(vel, dtype=np.float64).reshape(-1, 3)
        n = len(pos)
        self._grow(n)
        s = slice(self._n, self._n + n)
        self._pos[s] = pos
        self._vel[s] = vel
        self._m[s] = masses
        self._r[s] = 0.0 if radii is None else radii
        self._hash[s] = np.random.randint(100000000, 999999999, n)
        self._primary.extend([None] * n)
        self._n += n

    def remove_particle(self, pid):
        keep = np.ones(self._n, dtype=bool)
        keep[pid] = False
        for name in ("_pos", "_vel", "_m", "_r", "_ptype", "_hash"):
            arr = getattr(self, name)[: self._n][keep]
            setattr(self, name, arr.copy())
        del self._primary[pid]
        self._names = {k: (v if v < pid else v - 1) for k, v in self._names.items() if v != pid}
        self._n -= 1

    def merge_particles_inelastically(self, pid1, pid2):
        if not (0 <= pid1 < self._n and 0 <= pid2 < self._n):
            return -1
        keep, gone = (pid1, pid2) if self._m[pid1] >= self._m[pid2] else (pid2, pid1)
        m1, m2 = self._m[keep], self._m[gone]
        self._vel[keep] = (m1 * self._vel[keep] + m2 * self._vel[gone]) / (m1 + m2)
        self._m[keep] = m1 + m2
        self._r[keep] = np.power(self._r[keep] ** 3 + self._r[gone] ** 3, 1.0 / 3)
        self._primary = [keep if p == gone else p for p in self._primary]
        print("Merging particles inelastically: #%d + #%d ==> #%d" % (keep, gone, keep))
        self.remove_particle(gone)
        return 0

    # ------------------------------------------------------------ primaries
    def get_center_of_mass(self, subset=None):
        idx = np.arange(self._n) if subset is None else np.asarray(subset)
        m = self._m[idx]
        mt = m.sum()
        return Particle(mass=mt, pos=(m[:, None] * self._pos[idx]).sum(0) / mt,
                        vel=(m[:, None] * self._vel[idx]).sum(0) / mt)

    def determine_primary_body(self, primary):
        if primary is None:
            primary = self.primary
            if self._n == 0:
                return Particle(mass=0.0)
            if primary == "#COM#":
                return self.get_center_of_mass()
            if primary == "#M_MAX#":
                return self[int(np.argmax(self.masses))]
            if primary == "#M_MIN#":
                return self[int(np.argmin(self.masses))]
        if isinstance(primary, list):
            return self.get_center_of_mass(subset=primary)
        if isinstance(primary, Particle):
            return primary
        return self[primary]

    # ------------------------------------------------------------ elements
    def _primary_arrays(self, primary):
        """Primary position, velocity and mass for every body (vectorised)."""
        n = self._n
        pp = np.zeros((n, 3)); pv = np.zeros((n, 3)); pm = np.zeros(n)
        cache = {}
        for k in range(n):
            key = self._primary[k] if self._primary[k] is not None else ("global", repr(primary))
            if key not in cache:
                prim = self.determine_primary_body(self._primary[k] if self._primary[k] is not None else primary)
                cache[key] = (prim.pos, prim.vel, prim.mass)
            pp[k], pv[k], pm[k] = cache[key]
        return pp, pv, pm

    def calculate_orbital_elements(self, primary=None):
        if self._n < 2:
            return np.full(6, np.nan)
        pp, pv, pm = self._primary_arrays(primary)
        a, e, i, Om, om, f = cartesian_to_elements(self.masses, pm, self._pos[: self._n] - pp,
                                                   self._vel[: self._n] - pv, self.CONST_G)
        return np.stack([a, e, i, Om, om, f], axis=1)

    def calculate_aei(self, primary=None):
        if self._n < 2:
            return np.array([np.nan, np.nan, np.nan])
        return self.calculate_orbital_elements(primary)[:, :3]

    @property
    def energy(self):
        """Total energy over all pairs (abie/particles.py:399-416)."""
        m = self.masses
        ke = 0.5 * np.sum(m * np.sum(self._vel[: self._n] ** 2, axis=1))
        d = self._pos[: self._n, None, :] - self._pos[None, : self._n, :]
        r = np.sqrt((d ** 2).sum(-1))
        np.fill_diagonal(r, np.inf)
        return ke - 0.5 * self.CONST_G * np.sum(np.outer(m, m) / r)
