"""abie_amd -- MI355X (gfx950) backend of ABIE's all-pairs gravity path.

The ``ABIE`` class keeps the reference's user API (abie/__init__.py:21-403):
constructor keywords, ``add``, ``integrator = "GaussRadau15"``,
``integrate``, ``calculate_energy`` and the event properties.  ``deviceID``
(or its README spelling ``device``) selects the GPU; the accelerations and
the GR15 integrator run in libabie_hip.so.

    import abie_amd as abie
    sim = abie.ABIE(CONST_G=4 * np.pi ** 2, device=0)
    sim.add(mass=1.0, name="Sun")
    sim.add(mass=1e-3, a=5.2, e=0.05, f=0.0, name="Jupiter")
    sim.integrator = "GaussRadau15"
    sim.store_dt = 10
    sim.integrate(100)
"""
import argparse
import sys

import numpy as np

from .data_io import DataIO
from .events import CloseEncounterException, CollisionException  # noqa: F401
from .integrator import Integrator
from .particles import Particle, Particles  # noqa: F401
from .recorder import SimulationDataRecorder

__all__ = ["ABIE", "Particles", "Particle", "Integrator", "CollisionException", "CloseEncounterException"]


class ABIE(object):
    def __init__(self, CONST_G=4 * np.pi ** 2, CONST_C=0.0, integrator="GaussRadau15", h=0.1, store_dt=100,
                 name="simulation", buffer_len=10240, deviceID=None, device=None):
        if deviceID is None:
            deviceID = 0 if device is None else device
        self.__CONST_G = CONST_G
        self.__CONST_C = CONST_C
        self.__t = 0.0
        self.__t_start = 0.0
        self.__t_end = 0.0
        self.__h = h
        self._particles = None
        self.__buf = None
        self.__store_dt = store_dt
        self.__buffer_len = buffer_len
        self.__max_close_encounter_events = 1
        self.__max_collision_events = 1
        self.__close_encounter_distance = 0.0
        self.__device_id = deviceID
        self.__integrator_modules = None
        self.__integrator_instances = {}
        self.__integrator = None
        self.__integrator_name = integrator
        self.output_file = "%s.hdf5" % name
        self.__close_encounter_output_file = "%s_ce.txt" % name
        self.__collision_output_file = "%s_collisions.txt" % name

    # ------------------------------------------------------------ properties
    @property
    def particles(self):
        if self._particles is None:
            self._particles = Particles(self.CONST_G)
        return self._particles

    @property
    def buffer(self):
        if self.__buf is None:
            self.__buf = DataIO(buf_len=self.__buffer_len, output_file_name=self.output_file, CONST_G=self.CONST_G)
        return self.__buf

    def _forward(name):  # noqa: N805 -- property factory forwarding to the integrator
        priv = "_ABIE__" + name

        def get(self):
            if self.__integrator is not None:
                setattr(self, priv, getattr(self.__integrator, name))
            return getattr(self, priv)

        def set_(self, value):
            setattr(self, priv, value)
            if self.__integrator is not None:
                setattr(self.__integrator, name, value)
        return property(get, set_)

    max_close_encounter_events = _forward("max_close_encounter_events")
    max_collision_events = _forward("max_collision_events")
    close_encounter_distance = _forward("close_encounter_distance")
    close_encounter_output_file = _forward("close_encounter_output_file")
    collision_output_file = _forward("collision_output_file")
    h = _forward("h")
    store_dt = _forward("store_dt")
    del _forward

    @property
    def t(self):
        if self.__integrator is not None:
            self.__t = self.__integrator.t
        return self.__t

    @property
    def CONST_G(self):
        return self.__CONST_G

    @CONST_G.setter
    def CONST_G(self, value):
        self.__CONST_G = value
        if self.__integrator is not None:
            self.__integrator.CONST_G = value

    @property
    def CONST_C(self):
        return self.__CONST_C

    @CONST_C.setter
    def CONST_C(self, value):
        self.__CONST_C = value
        if self.__integrator is not None:
            self.__integrator.CONST_C = value

    @property
    def t_end(self):
        return self.__t_end

    @t_end.setter
    def t_end(self, tf):
        self.__t_end = tf
        if self.__integrator is not None:
            self.__integrator.t_end = tf

    @property
    def acceleration_method(self):
        return self.integrator.acceleration_method

    @acceleration_method.setter
    def acceleration_method(self, value):
        self.integrator.acceleration_method = value

    @property
    def integrator(self):
        if self.__integrator is None:
            raise RuntimeError("Integrator not set!")
        return self.__integrator

    @integrator.setter
    def integrator(self, name):
        if self.__integrator_modules is None:
            self.__integrator_modules = Integrator.load_integrators()
        print("Selecting %s as the active integrator." % name)
        if name not in self.__integrator_modules:
            raise ValueError("Unknown integrator %s. Supported integrators are %s"
                             % (name, sorted(self.__integrator_modules)))
        if name not in self.__integrator_instances:
            cls = getattr(self.__integrator_modules[name], name)
            self.__integrator_instances[name] = cls(self.particles, self.buffer, self.CONST_G, self.CONST_C,
                                                    deviceID=self.__device_id)
        it = self.__integrator_instances[name]
        self.__integrator = it
        it.CONST_G = self.CONST_G
        it.t_end = self.__t_end
        it.h = self.__h
        it._t = self.__t
        it.t_start = self.__t_start
        it.store_dt = self.__store_dt
        it.output_file = self.output_file
        it.collision_output_file = self.__collision_output_file
        it.close_encounter_output_file = self.__close_encounter_output_file
        it.max_close_encounter_events = self.__max_close_encounter_events
        it.max_collision_events = self.__max_collision_events
        it.close_encounter_distance = self.__close_encounter_distance

    @property
    def data(self):
        """Flush the snapshot buffer and return what the recorder holds
        (abie/__init__.py:266-269)."""
        self.buffer.flush()
        return self.buffer.recorder.data

    def record_simulation(self, particles=None, quantities=None):
        """Choose what the recorder keeps and return it; recording begins at
        its ``start()`` (abie/__init__.py:271-280)."""
        if self.buffer.recorder is None:
            self.buffer.recorder = SimulationDataRecorder(particles, quantities)
        else:
            self.buffer.recorder.set_monitored_particles(particles)
            self.buffer.recorder.set_monitored_quantities(quantities)
        return self.buffer.recorder

    # ------------------------------------------------------------ actions
    def initialize(self, config=None):
        if self.__integrator is None:
            self.integrator = self.__integrator_name
        if config is None:
            self.integrator.initialize()
            return
        # a parsed configuration file (abie/__init__.py:297-325)
        self.CONST_G = float(config["physical_params"]["G"])
        self.integrator = config["integration"]["integrator"]
        self.integrator.h = float(config["integration"]["h"])
        self.integrator.acceleration_method = config["integration"].get("acc_method", "ctypes")
        if self.integrator.t_start == 0:
            self.integrator.t_start = float(config["integration"]["t0"])
        if self.t_end == 0:
            self.t_end = float(config["integration"]["tf"])
        DataIO.ic_populate(config["initial_conds"], self, names=config.get("names"))
        self.integrator.initialize()

    def stop(self):
        self.integrator.stop()

    def set_additional_forces(self, ext_acc):
        if ext_acc.ndim == 1 and ext_acc.shape[0] == 3 * self.integrator.particles.N:
            self.integrator.set_additional_forces(ext_acc)
        else:
            print("WARNING: Additional forces array needs to be 3 * N vector, where N is the number of particles.")

    def integrate(self, to_time=None):
        if self.__integrator is None:
            self.integrator = self.__integrator_name
        try:
            ret = self.integrator.integrate(to_time)
            self.__t = self.integrator.t
            return ret
        except KeyboardInterrupt:
            print("Keyboard Interruption detected (Ctrl+C). Simulation stopped. Stopping the code...")
            self.stop()
            sys.exit(0)

    def calculate_orbital_elements(self, primary=None):
        return self.integrator.calculate_orbital_elements(primary)

    def calculate_energy(self):
        return self.integrator.calculate_energy()

    def add(self, pos=None, vel=None, x=0.0, y=0.0, z=0.0, vx=0.0, vy=0.0, vz=0.0, mass=0.0, name=None, radius=0.0,
            ptype=0, a=None, e=0.0, i=0.0, Omega=0.0, omega=0.0, f=None, primary=None):
        """Add a body by position/velocity or by orbital elements.  Unlike the
        reference (whose default ``f`` is drawn once at import and shared by
        every call, abie/__init__.py:376), a missing ``f`` is drawn per call."""
        if pos is None:
            pos = np.array([x, y, z], dtype=np.double)
        if vel is None:
            vel = np.array([vx, vy, vz], dtype=np.double)
        if f is None:
            f = 2 * np.pi * np.random.rand()
        return self.particles.add(pos=pos, vel=vel, mass=mass, name=name, radius=radius, ptype=ptype, a=a, e=e,
                                  i=i, Omega=Omega, omega=omega, f=f, primary=primary)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("-c", "--config", help="config file (TOML)", default=None)
    ap.add_argument("-o", "--output_file", dest="output_file", default="data.hdf5")
    ap.add_argument("-r", "--rebound_file", help="REBOUND simulation file", default=None)
    ap.add_argument("-t", "--t_end", type=float, dest="t_end")
    ap.add_argument("-d", "--dt", type=float, dest="dt", default=None)
    ap.add_argument("-s", "--store_dt", type=float, dest="store_dt", default=100)
    ap.add_argument("-i", "--integrator", dest="integrator", default="GaussRadau15")
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args()
    sim = ABIE(deviceID=args.device)
    sim.integrator = args.integrator
    sim.output_file = args.output_file
    if args.t_end is not None:
        sim.t_end = args.t_end
    if args.config is not None:
        sim.initialize(DataIO.parse_config_file(args.config))
    elif args.rebound_file is not None:
        DataIO.ic_populate_from_rebound(args.rebound_file, sim)
    if args.dt is not None:
        sim.h = args.dt
    sim.store_dt = args.store_dt
    sim.integrate()


if __name__ == "__main__":
    main()
