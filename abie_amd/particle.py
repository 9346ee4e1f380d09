"""``abie.particle`` of the reference (abie/particle.py): the Particle class,
defined with the particle set in ``abie_amd.particles``."""
from .particles import Particle  # noqa: F401

__all__ = ["Particle"]
