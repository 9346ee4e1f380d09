"""Snapshot buffer of the facade (abie/data_io.py of the reference).

The reference writes HDF5 groups ``Step#n`` through h5py; h5py is not
available here, so snapshots are kept in memory and ``flush`` writes them to
``<output>.npz`` with the same field names (t, x, y, z, vx, vy, vz, mass,
radius, ptype, hash, a, ecc, inc, energy).  Close-encounter and collision
rows are appended to text files as in the reference.
"""
import os

import numpy as np


class DataIO(object):
    def __init__(self, buf_len=1024, output_file_name="data.hdf5", close_encounter_output_file_name=None,
                 collision_output_file_name=None, CONST_G=4 * np.pi ** 2, append=False):
        self.buf_len = buf_len
        self.output_file_name = output_file_name
        self.ce_file = close_encounter_output_file_name
        self.coll_file = collision_output_file_name
        self.CONST_G = CONST_G
        self.snapshots = []
        self.recorder = None
        self.n = 0

    def initialize_buffer(self, n_particles):
        self.n = n_particles

    def reset_buffer(self):
        self.snapshots = []

    def store_state(self, t, pos, vel, masses, radii=None, names=None, ptypes=None, a=None, e=None, i=None,
                    energy=None):
        pos = np.asarray(pos).reshape(-1, 3)
        vel = np.asarray(vel).reshape(-1, 3)
        snap = dict(t=t, x=pos[:, 0].copy(), y=pos[:, 1].copy(), z=pos[:, 2].copy(), vx=vel[:, 0].copy(),
                    vy=vel[:, 1].copy(), vz=vel[:, 2].copy(), mass=np.array(masses, copy=True),
                    radius=None if radii is None else np.array(radii, copy=True),
                    ptype=None if ptypes is None else np.array(ptypes, copy=True),
                    hash=None if names is None else np.array(names, copy=True),
                    a=None if a is None else np.array(a, copy=True), ecc=None if e is None else np.array(e, copy=True),
                    inc=None if i is None else np.array(i, copy=True), energy=energy)
        self.snapshots.append(snap)
        if len(self.snapshots) >= self.buf_len:
            self.flush()

    def _append_rows(self, path, buf):
        if path is None:
            return
        with open(path, "a") as f:
            for row in np.asarray(buf).reshape(-1, 4):
                f.write("%g %d %d %g\n" % (row[0], int(row[1]), int(row[2]), row[3]))

    def store_collisions(self, collision_buffer):
        self._append_rows(self.coll_file, collision_buffer)

    def store_close_encounters(self, ce_buffer):
        self._append_rows(self.ce_file, ce_buffer)

    def flush(self):
        if not self.snapshots or self.output_file_name is None:
            return
        base = os.path.splitext(self.output_file_name)[0] + ".npz"
        old = {}
        if os.path.exists(base):
            with np.load(base) as d:
                old = {k: d[k] for k in d.files}
        k0 = len([k for k in old if k.endswith("/t")])
        out = dict(old)
        for j, s in enumerate(self.snapshots):
            for key, val in s.items():
                if val is not None:
                    out["Step#%d/%s" % (k0 + j, key)] = np.asarray(val)
        np.savez(base, **out)
        self.snapshots = []

    def close(self):
        self.flush()
