"""GaussRadau15 plugin: each store_dt chunk is one integrator_gr call of
libabie_hip.so, i.e. GR15 on the GPU (abie/integrator_gauss_radau15.py:20-45
of the reference)."""
import numpy as np

from .events import CloseEncounterException, CollisionException
from .integrator import Integrator

__integrator__ = "GaussRadau15"


class GaussRadau15(Integrator):
    def __init__(self, particles=None, buffer=None, CONST_G=4 * np.pi ** 2, CONST_C=0.0, deviceID=0):
        super().__init__(particles=particles, buffer=buffer, CONST_G=CONST_G, CONST_C=CONST_C, deviceID=deviceID)
        self.tol = 1.0e-9

    def integrate_ctypes(self, to_time=None):
        ret = 0
        try:
            self.libabie.integrator_gr(self.t, to_time, 1.0)
        except CollisionException as e:
            print(e)
            self.handle_collisions(self.libabie.get_collision_data())
            ret = 1
        except CloseEncounterException as e:
            print(e)
            self.store_close_encounters(self.libabie.get_close_encounter_data())
            ret = 2
        finally:
            pos = self.particles.positions.copy()
            vel = self.particles.velocities.copy()
            self.libabie.get_state(pos, vel, self.particles.masses.copy(), self.particles.radii.copy())
            self.particles.positions = pos
            self.particles.velocities = vel
            self._t = self.libabie.get_model_time()
            self.store_state()
        return ret
