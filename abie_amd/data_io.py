"""Snapshot buffer and HDF5 output of the facade (abie/data_io.py).

File layout, as the reference writes it (abie/data_io.py:73-117): a root
attribute ``G`` (float64) and one group ``Step#k`` per buffer flush, holding
float64 datasets ``time``, ``energy`` of shape (rows,) and ``mass``,
``ptype``, ``hash``, ``radius``, ``x``, ``y``, ``z``, ``vx``, ``vy``, ``vz``,
``a``, ``ecc``, ``inc`` of shape (rows, N).  ``ptype`` and ``hash`` are
float64 there too (the reference's integer buffers are multiplied by NaN at
allocation, abie/data_io.py:47-62), and a field never supplied stays NaN.
The file is written through the HDF5 C library (``abie_amd.h5file``).

Behaviour kept from the reference:
* ``store_state`` skips a time equal to the previous snapshot's, starting
  from 0.0, so a snapshot at t = 0 is not stored (abie/data_io.py:135-137);
* a full buffer is flushed before the next row (abie/data_io.py:139-142);
* close-encounter and collision files are rewritten from the whole event
  buffer on each call, ``%g, %d, %d, %g`` rows under the header
  ``Time, Particle 1, Particle 2, Distance`` (abie/data_io.py:239-255).

Differences: the event files are removed only when the first buffer is
allocated, not again when a merge re-allocates it (the reference's
re-allocation after a merge deletes the collision file it has just
written, abie/data_io.py:64-68 with abie/integrator.py:264-276); ``store``
of whole state blocks, whose buffer the reference never allocates
(abie/data_io.py:159-219), is not provided.
"""
import os
import warnings

import numpy as np

from . import h5file
from .recorder import SimulationDataRecorder

# the order in which the reference fills its per-flush dictionary
STEP_FIELDS = ("time", "mass", "ptype", "hash", "radius", "x", "y", "z", "vx", "vy", "vz", "a", "ecc", "inc",
               "energy")
_PER_BODY = ("mass", "ptype", "hash", "radius", "x", "y", "z", "vx", "vy", "vz", "a", "ecc", "inc")


class DataIO(object):
    def __init__(self, buf_len=1024, output_file_name="data.hdf5", collision_output_file_name="collisions.txt",
                 close_encounter_output_file_name="close_encounters.txt", CONST_G=1):
        self.recorder = SimulationDataRecorder()
        self.buf_len = int(buf_len)
        self.buf_initialized = False
        self.store_t = 0.0
        self.buf = None
        self.buf_cursor = 0
        self.output_file_name = output_file_name
        self.collision_output_file_name = collision_output_file_name
        self.close_encounter_output_file_name = close_encounter_output_file_name
        self.h5_file = None
        self.h5_step_id = 0
        self.CONST_G = CONST_G
        self._event_files_cleared = False

    # ------------------------------------------------------------ buffer
    def initialize_buffer(self, n_particles):
        if self.buf_initialized:
            return
        L = self.buf_len
        self.buf = {"time": np.full(L, np.nan), "energy": np.full(L, np.nan)}
        for key in _PER_BODY:
            self.buf[key] = np.full((L, n_particles), np.nan)
        self.buf_cursor = 0
        if not self._event_files_cleared:
            for path in (self.close_encounter_output_file_name, self.collision_output_file_name):
                if path is not None and os.path.isfile(path):
                    os.remove(path)
            self._event_files_cleared = True
        self.buf_initialized = True

    def reset_buffer(self):
        self.buf_initialized = False

    def store_state(self, t, pos, vel, masses, radii=None, names=None, ptypes=None, a=None, e=None, i=None,
                    energy=None):
        if self.store_t == t:
            return
        if self.buf is None:
            self.initialize_buffer(len(np.asarray(masses)))
        if self.buf_cursor == self.buf_len:
            self.flush()
        row = self.buf_cursor
        b = self.buf
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        vel = np.asarray(vel, dtype=np.float64).reshape(-1, 3)
        b["time"][row] = t
        b["x"][row], b["y"][row], b["z"][row] = pos[:, 0], pos[:, 1], pos[:, 2]
        b["vx"][row], b["vy"][row], b["vz"][row] = vel[:, 0], vel[:, 1], vel[:, 2]
        b["mass"][row] = masses
        for key, val in (("radius", radii), ("hash", names), ("ptype", ptypes), ("a", a), ("ecc", e),
                         ("inc", i), ("energy", energy)):
            if val is not None:
                b[key][row] = val
        self.store_t = t
        self.buf_cursor += 1

    # ------------------------------------------------------------ output
    def _open(self):
        if self.h5_file is not None:
            return True
        if self.output_file_name is None:
            return False
        if not h5file.available():
            warnings.warn("libhdf5 not found: snapshots of %s are kept in memory only (set ABIE_HDF5_LIB)"
                          % self.output_file_name)
            self.output_file_name = None
            return False
        self.h5_file = h5file.File(self.output_file_name, "w")
        self.h5_file.set_attr("G", float(self.CONST_G))
        return True

    def flush(self):
        """Write the buffered rows as group ``Step#k`` and empty the buffer."""
        if self.buf_cursor == 0:
            return
        n = self.buf_cursor
        state = {key: self.buf[key][:n] for key in STEP_FIELDS}
        if self._open():
            group = "Step#%d" % self.h5_step_id
            self.h5_file.create_group(group)
            for key in STEP_FIELDS:
                self.h5_file.create_dataset("%s/%s" % (group, key), state[key])
            self.h5_file.flush()
        self.h5_step_id += 1
        self.buf_cursor = 0
        if self.recorder is not None:
            self.recorder.record(state_dict=state)

    def close(self):
        if self.buf_cursor > 0:
            self.flush()
        if self.h5_file is not None:
            self.h5_file.close()
            self.h5_file = None

    def store(self, buf_t, buf_state, buf_len, n_particles):
        raise NotImplementedError("block stores are not provided; use store_state")

    # ------------------------------------------------------------ events
    @staticmethod
    def _write_events(path, buf):
        if path is None:
            return
        rows = np.asarray(buf, dtype=np.float64).reshape(-1, 4)
        np.savetxt(path, rows, fmt="%g, %d, %d, %g", header="Time, Particle 1, Particle 2, Distance")

    def store_collisions(self, collision_buffer):
        self._write_events(self.collision_output_file_name, collision_buffer)

    def store_close_encounters(self, ce_buffer):
        self._write_events(self.close_encounter_output_file_name, ce_buffer)

    # ------------------------------------------------- initial conditions
    @staticmethod
    def parse_config_file(config_file):
        """A TOML configuration as a dict (abie/data_io.py:257-262)."""
        try:
            import tomllib as _toml
        except ImportError:
            import tomli as _toml
        with open(config_file, "rb") as f:
            return _toml.load(f)

    @staticmethod
    def ic_populate_from_rebound(reb_ic_filename, sim_abie):
        """Initial conditions from a REBOUND snapshot (abie/data_io.py:264-283)."""
        try:
            import rebound
        except ImportError:
            print("Rebound is not installed. Cannot populate IC from rebound.")
            return None
        sim_reb = rebound.Simulation(reb_ic_filename)
        sim_abie.CONST_G = sim_reb.G
        sim_abie.h = sim_reb.dt
        for p in sim_reb.particles:
            sim_abie.add(mass=p.m, x=p.x, y=p.y, z=p.z, vx=p.vx, vy=p.vy, vz=p.vz, name=p.hash.value, radius=p.r)
        return sim_abie

    @staticmethod
    def ic_populate(ic_dict, sim, names=None):
        """Add bodies from ``{name: [x, y, z, vx, vy, vz, m]}``, in ``names``
        order when given (abie/data_io.py:314-325)."""
        for p_name in (names if names is not None else list(ic_dict)):
            row = np.asarray(ic_dict[p_name], dtype=np.float64)
            sim.add(pos=row[0:3], vel=row[3:6], mass=row[6], name=p_name)

    @staticmethod
    def toml_generator(state_vec, G, t0, tf, h, masses, filename):
        """Write a solar-system configuration file from a state vector laid out
        as 9 positions then 9 velocities (abie/data_io.py:327-350)."""
        bodies = ["Sun", "Mercury", "Venus", "Earth", "Mars", "Jupiter", "Saturn", "Uranus", "Neptune"]
        state_vec = np.asarray(state_vec, dtype=np.float64)
        nb = len(bodies)
        lines = ["[physical_params]", "G = %r" % float(G), "", "[integration]", "t0 = %r" % float(t0),
                 "tf = %r" % float(tf), "h = %r" % float(h), "", "[initial_conds]"]
        for k, name in enumerate(bodies):
            row = list(state_vec[3 * k:3 * k + 3]) + list(state_vec[3 * (nb + k):3 * (nb + k) + 3]) + [masses[k]]
            lines.append("%s = [ %s,]" % (name, ", ".join(repr(float(v)) for v in row)))
        with open(filename, "w") as f:
            f.write("\n".join(lines) + "\n")
