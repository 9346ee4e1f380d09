"""CLibABIE: the ctypes front end of the device backend, with the method
names and exception behaviour of the reference's ``abie/clibabie.py:8-194``.

Differences from the reference, all deliberate:
  * it binds ``libabie_hip.so`` with explicit ``argtypes`` (the reference
    passes sizes as ``c_int`` into ``size_t`` parameters, clibabie.py:43-45);
  * ``deviceID`` must name a GPU: this library has no CPU path;
  * negative return codes raise ``AbieHipError`` instead of being ignored.
"""
import ctypes

import numpy as np

from . import _lib
from .events import CloseEncounterException, CollisionException


def _dp(a):
    return a.ctypes.data_as(_lib.c_dp)


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


class CLibABIE(object):
    def __init__(self):
        self.lib = _lib.load()
        self.max_close_encounter_events = 1
        self.max_collision_events = 1

    def initialize_code(self, G, C, N_MAX, MAX_CE_EVENTS=1, MAX_COLLISION_EVENTS=1,
                        close_encounter_distance=0, deviceID=0):
        _lib.check(self.lib.initialize_code(G, C, int(N_MAX), int(MAX_CE_EVENTS), int(MAX_COLLISION_EVENTS),
                                            int(deviceID)), "initialize_code")
        self.lib.set_close_encounter_distance(float(close_encounter_distance))
        self.max_close_encounter_events = MAX_CE_EVENTS
        self.max_collision_events = MAX_COLLISION_EVENTS

    def finalize_code(self):
        self.lib.finalize_code()

    def set_state(self, pos, vel, masses, radii, N, G, C):
        pos, vel, masses, radii = _f64(pos), _f64(vel), _f64(masses), _f64(radii)
        self.lib.set_state(_dp(pos), _dp(vel), _dp(masses), _dp(radii), int(N), G, C)

    def get_state(self, pos, vel, masses, radii):
        for a in (pos, vel, masses, radii):
            if a.dtype != np.float64 or not a.flags.c_contiguous:
                raise ValueError("get_state needs C-contiguous float64 arrays")
        return self.lib.get_state(_dp(pos), _dp(vel), _dp(masses), _dp(radii))

    def get_model_time(self):
        return self.lib.get_model_time()

    def get_close_encounter_data(self):
        buf = np.zeros(4 * self.max_close_encounter_events)
        self.lib.get_close_encounter_buffer(_dp(buf))
        return buf.reshape(self.max_close_encounter_events, 4)

    def get_collision_data(self):
        buf = np.zeros(4 * self.max_collision_events)
        self.lib.get_collision_buffer(_dp(buf))
        return buf.reshape(self.max_collision_events, 4)

    def reset_close_encounter_buffer(self):
        self.lib.reset_close_encounter_buffer()

    def reset_collision_buffer(self):
        self.lib.reset_collision_buffer()

    def set_close_encounter_distance(self, value):
        self.lib.set_close_encounter_distance(float(value))

    def get_total_energy(self):
        return self.lib.calculate_energy()

    def set_additional_forces(self, ext_acc):
        ext_acc = _f64(ext_acc)
        self.lib.set_additional_forces(ext_acc.size // 3, _dp(ext_acc))

    def integrator_gr(self, t, t_end, dt):
        ret = _lib.check(self.lib.integrator_gr(t, t_end, dt), "integrator_gr")
        if ret == 1:
            buf = self.get_close_encounter_data()
            raise CloseEncounterException(buf[-1, 0], int(buf[-1, 1]), int(buf[-1, 2]), buf[-1, 3])
        elif ret == 2:
            buf = self.get_collision_data()
            raise CollisionException(buf[-1, 0], int(buf[-1, 1]), int(buf[-1, 2]), buf[-1, 3])
        return ret

    def integrator_gauss_radau15(self, pos, vel, masses, N, G, t, t_end, dt, radii=None):
        radii = np.zeros(int(N)) if radii is None else _f64(radii)
        rc = ctypes.c_long(self.lib.integrator_gauss_radau15(_dp(pos), _dp(vel), _dp(_f64(masses)), _dp(radii),
                                                             int(N), G, t, t_end, dt)).value
        return _lib.check(rc, "integrator_gauss_radau15")

    def integrator_rk(self, t, t_end, dt):
        return _lib.check(self.lib.integrator_rk(t, t_end, dt), "integrator_rk")

    def integrator_wh(self, t, t_end, dt):
        return _lib.check(self.lib.integrator_wh(t, t_end, dt), "integrator_wh")

    def integrator_runge_kutta(self, pos, vel, masses, N, G, t, t_end, dt):
        """abie/clibabie.py:112-122: not part of the device path; the library
        only records the error, raised here instead of returning silently."""
        r = np.zeros(int(N))
        self.lib.integrator_runge_kutta(_dp(pos), _dp(vel), _dp(_f64(masses)), _dp(r), int(N), G, t, t_end, dt)
        raise NotImplementedError(self.lib.abie_hip_last_error().decode())

    def integrator_wisdom_holman(self, pos, vel, masses, N, G, t, t_end, dt):
        """abie/clibabie.py:136-146: as integrator_runge_kutta."""
        r = np.zeros(int(N))
        self.lib.integrator_wisdom_holman(_dp(pos), _dp(vel), _dp(_f64(masses)), _dp(r), int(N), G, t, t_end, dt)
        raise NotImplementedError(self.lib.abie_hip_last_error().decode())

    def ode_n_body_first_order(self, x, N, G, masses):
        """abie/clibabie.py:174-183: not part of the device path."""
        dxdt = np.zeros(np.asarray(x).size)
        rc = ctypes.c_long(self.lib.ode_n_body_first_order(_dp(_f64(x)), int(N), G, _dp(_f64(masses)),
                                                           _dp(dxdt))).value
        _lib.check(rc, "ode_n_body_first_order")
        return dxdt

    def ode_n_body_second_order(self, x, N, G, masses):
        x = _f64(x)
        acc = np.zeros(x.size, dtype=np.double)
        rc = ctypes.c_long(self.lib.ode_n_body_second_order(_dp(x), int(N), G, _dp(_f64(masses)),
                                                            _dp(np.zeros(int(N))), _dp(acc))).value
        _lib.check(rc, "ode_n_body_second_order")
        return acc

    def calculate_accelerations(self, x, v, N, G, masses, radii=None, acc=None):
        x = _f64(x)
        acc = np.zeros(x.size, dtype=np.double) if acc is None else acc
        radii = np.zeros(int(N)) if radii is None else _f64(radii)
        rc = ctypes.c_long(self.lib.calculate_accelerations(_dp(x), _dp(_f64(v)), int(N), G, _dp(_f64(masses)),
                                                            _dp(radii), _dp(acc))).value
        _lib.check(rc, "calculate_accelerations")
        return acc

    def check_collisions_close_encounters(self, x, radii, N, t):
        rc = ctypes.c_long(self.lib.check_collisions_close_encounters(_dp(_f64(x)), _dp(_f64(radii)), int(N),
                                                                      float(t))).value
        return _lib.check(rc, "check_collisions_close_encounters")
