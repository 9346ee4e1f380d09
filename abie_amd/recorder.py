"""In-memory recorder of flushed snapshots (abie/recorder.py).

``DataIO.flush`` hands every flushed block to ``record``; while recording
(after ``start()``), each monitored quantity is appended to ``data[qty]``:
rows holding a NaN are dropped (abie/recorder.py:63-69), and per-body
columns are matched by particle hash to the bodies of the first recorded
block, so that after a merge the surviving bodies keep their columns and
the removed ones read NaN (abie/recorder.py:71-101).

``particles`` (indices into the first recorded block's bodies) restricts the
per-body quantities to those columns.  The reference applies that index to
the time axis instead (abie/recorder.py:80-84, 102-106); here it selects
bodies, which is what the name promises.
"""
import numpy as np

DEFAULT_QUANTITIES = ["a", "ecc", "inc", "x", "y", "z", "vx", "vy", "vz", "mass", "time", "energy"]


class SimulationDataRecorder(object):
    def __init__(self, particles=None, quantities=["a", "ecc", "inc"], buffer_len=0):
        self._monitored_particles = particles
        self._monitored_quantities = quantities
        self._recording = False
        self._particles_hash_vec = None
        self.data = dict()

    def reset(self):
        self.data = dict()
        self._particles_hash_vec = None

    def start(self):
        self._recording = True

    def stop(self):
        self._recording = False

    def set_monitored_quantities(self, quantities):
        self._monitored_quantities = quantities

    def get_monitored_quantities(self):
        return self._monitored_quantities

    def set_monitored_particles(self, particles):
        self._monitored_particles = particles

    def get_monitored_particles(self):
        return self._monitored_particles

    def record(self, state_dict):
        if not self._recording:
            return
        quantities = DEFAULT_QUANTITIES if self._monitored_quantities is None else self._monitored_quantities
        hashes = np.asarray(state_dict["hash"])
        cur_hash = hashes[0] if hashes.ndim == 2 and len(hashes) else None
        if self._particles_hash_vec is None and cur_hash is not None:
            self._particles_hash_vec = cur_hash.copy()
        ref = self._particles_hash_vec
        cols = None
        if self._monitored_particles is not None:
            cols = np.atleast_1d(np.asarray(self._monitored_particles, dtype=np.int64))
        for qty in quantities:
            if qty not in state_dict:
                continue
            d = np.asarray(state_dict[qty], dtype=np.float64)
            if d.ndim == 2:
                d = d[~np.isnan(d).any(axis=1)]
                if ref is not None and cur_hash is not None and (len(cur_hash) != len(ref)
                                                                 or not np.array_equal(cur_hash, ref)):
                    # columns of the current bodies in the first block's order
                    full = np.full((d.shape[0], len(ref)), np.nan)
                    where = np.nonzero(np.isin(ref, cur_hash))[0]
                    src = np.array([int(np.nonzero(cur_hash == ref[w])[0][0]) for w in where], dtype=np.int64)
                    full[:, where] = d[:, src]
                    d = full
                if cols is not None:
                    d = d[:, cols]
            else:
                d = d[~np.isnan(d)]
            if qty in self.data and self.data[qty] is not None:
                self.data[qty] = np.append(self.data[qty], d, axis=0)
            else:
                self.data[qty] = d.copy()
