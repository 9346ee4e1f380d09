"""WisdomHolman: the integrator name is kept for the drop-in, but only GR15 runs
on the device path; this plugin reports that it is not provided here (the
reference's CPU library provides it)."""
from .integrator import Integrator

__integrator__ = "WisdomHolman"


class WisdomHolman(Integrator):
    def integrate_ctypes(self, to_time=None):
        raise NotImplementedError("WisdomHolman is not part of the MI355X device path; use GaussRadau15 "
                                  "(or the reference's CPU libabie for WisdomHolman)")
