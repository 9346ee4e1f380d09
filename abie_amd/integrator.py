"""Integrator plugin framework of the drop-in facade (abie/integrator.py:25-289
of the reference).

Plugins are modules named ``integrator_*.py`` (in this package or in the
current directory) that define ``__integrator__ = "<Name>"`` and a class of
that name deriving from ``Integrator``.  ``integrate(to_time)`` advances in
chunks of ``store_dt``: each chunk calls the subclass's ``integrate_ctypes``
(which restarts GR15, as the reference does), then stores a snapshot and
reports the energy drift dE/E0.
"""
import glob
import importlib
import importlib.util
import os
import sys

import numpy as np

from .clibabie import CLibABIE
from .data_io import DataIO
from .particles import Particles

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))


class Integrator(object):
    def __init__(self, particles=None, buffer=None, CONST_G=4 * np.pi ** 2, CONST_C=0.0, buf_len=10240,
                 deviceID=0):
        self.CONST_G = CONST_G
        self.CONST_C = CONST_C
        self._t = 0.0
        self.t_start = 0.0
        self.t_end = 0.0
        self.h = 0.01
        self.store_dt = 100
        self._particles = particles
        self.acceleration_method = "ctypes"
        self.output_file = "data.hdf5"
        self.collision_output_file = "collisions.txt"
        self.close_encounter_output_file = "close_encounters.txt"
        self.max_close_encounter_events = 1
        self.max_collision_events = 1
        self.close_encounter_distance = 0.0
        self._energy_init = 0.0
        self._energy = 0.0
        self._buf = buffer
        self.buffer_len = buf_len
        self._initialized = False
        self._device_id = deviceID
        self.verbose = True
        self.libabie = CLibABIE()

    @property
    def t(self):
        return self._t

    @property
    def particles(self):
        if self._particles is None:
            self._particles = Particles(self.CONST_G)
        return self._particles

    @property
    def buf(self):
        if self._buf is None:
            self._buf = DataIO(buf_len=self.buffer_len, output_file_name=self.output_file, CONST_G=self.CONST_G)
        return self._buf

    @staticmethod
    def load_integrators():
        """Map integrator name -> module, from integrator_*.py here and in the CWD."""
        mods = {}
        for path in sorted(glob.glob(os.path.join(_PKG_DIR, "integrator_*.py"))):
            name = os.path.basename(path)[:-3]
            mod = importlib.import_module("abie_amd." + name)
            if hasattr(mod, "__integrator__"):
                mods[mod.__integrator__] = mod
        cwd = os.getcwd()
        if os.path.abspath(cwd) != _PKG_DIR:
            for path in sorted(glob.glob(os.path.join(cwd, "integrator_*.py"))):
                name = os.path.basename(path)[:-3]
                spec = importlib.util.spec_from_file_location(name, path)
                mod = importlib.util.module_from_spec(spec)
                sys.modules[name] = mod
                spec.loader.exec_module(mod)
                if hasattr(mod, "__integrator__"):
                    mods[mod.__integrator__] = mod
        return mods

    def initialize(self):
        if self._buf is None:
            self._buf = DataIO(buf_len=self.buffer_len, output_file_name=self.output_file,
                               close_encounter_output_file_name=self.close_encounter_output_file,
                               collision_output_file_name=self.collision_output_file, CONST_G=self.CONST_G)
        if self.particles.N > 0:
            self.libabie.initialize_code(self.CONST_G, self.CONST_C, self.particles.N,
                                         MAX_CE_EVENTS=self.max_close_encounter_events,
                                         MAX_COLLISION_EVENTS=self.max_collision_events,
                                         close_encounter_distance=self.close_encounter_distance,
                                         deviceID=self._device_id)
            self.buf.initialize_buffer(self.particles.N)

    def stop(self):
        if self._buf is not None:
            self._buf.close()

    def calculate_orbital_elements(self, primary=None):
        return self.particles.calculate_orbital_elements(primary)

    def calculate_energy(self):
        if self.acceleration_method == "ctypes":
            return self.libabie.get_total_energy()
        return self.particles.energy

    def set_additional_forces(self, ext_acc):
        self.libabie.set_additional_forces(ext_acc)

    def integrator_warmup(self):
        self.libabie.set_state(self.particles.positions.copy(), self.particles.velocities.copy(),
                               self.particles.masses.copy(), self.particles.radii.copy(), self.particles.N,
                               self.CONST_G, self.CONST_C)

    def integrate(self, to_time=None):
        """Advance to ``to_time`` (or t_end) in store_dt chunks
        (abie/integrator.py:158-213)."""
        if to_time is not None:
            self.t_end = to_time
        if not self._initialized:
            self.initialize()
            self.integrator_warmup()
            self._initialized = True
        if self.t >= self.t_end:
            return 0
        dt = min(self.store_dt, self.t_end - self.t)
        ret = 0
        while self.t < self.t_end:
            if self._energy_init == 0:
                self._energy_init = self.calculate_energy()
            next_t = self.t + dt - ((self.t + dt) % dt)
            if self.acceleration_method == "ctypes":
                ret = self.integrate_ctypes(next_t)
            else:
                ret = self.integrate_numpy(next_t)
            self._energy = self.calculate_energy()
            if self.verbose:
                print("t = %f, N = %d, dE/E0 = %g" % (self.t, self.particles.N,
                                                      np.abs(self._energy - self._energy_init) / self._energy_init))
            if os.path.isfile("STOP"):
                break
        if to_time is None:
            self.stop()
        return ret

    def integrate_numpy(self, to_time):
        raise NotImplementedError("integrate_numpy() is not provided by the device backend")

    def integrate_ctypes(self, to_time):
        raise NotImplementedError("integrate_ctypes() method not implemented!")

    def store_state(self):
        self.buf.initialize_buffer(self.particles.N)
        elem = self.particles.calculate_aei()
        if self._energy == 0:
            self._energy = self.calculate_energy()
        elem = np.atleast_2d(elem)
        self.buf.store_state(self.t, self.particles.positions, self.particles.velocities, self.particles.masses,
                             radii=self.particles.radii, names=self.particles.hashes, ptypes=self.particles.ptypes,
                             a=elem[:, 0], e=elem[:, 1], i=elem[:, 2], energy=self._energy)

    def store_collisions(self, collision_buffer):
        self.buf.store_collisions(collision_buffer)

    def store_close_encounters(self, ce_buffer):
        self.buf.store_close_encounters(ce_buffer)

    def handle_collisions(self, collision_buffer, actions=None):
        actions = ["merge", "store"] if actions is None else actions
        if "store" in actions:
            self.store_state()
            self.store_collisions(collision_buffer)
        if "merge" in actions:
            collision_buffer = np.asarray(collision_buffer).reshape(-1, 4)
            for row in collision_buffer:
                self.particles.merge_particles_inelastically(int(row[1]), int(row[2]))
                self.libabie.reset_collision_buffer()
                self.integrator_warmup()
                self.buf.flush()
                self.buf.reset_buffer()
                self.buf.initialize_buffer(self.particles.N)
        if "halt" in actions:
            print("Simulation terminated due to a collision event.")
            sys.exit(0)

    def handle_close_encounters(self, ce_buffer, actions=None):
        actions = ["merge", "store"] if actions is None else actions
        if "store" in actions:
            self.store_state()
            self.store_close_encounters(ce_buffer)
        if "halt" in actions:
            print("Simulation terminated due to a close encounter event.")
            sys.exit(0)
