"""Event exceptions raised when GR15 stops on a close encounter or a
collision (same classes and messages as the reference's abie/events.py:1-34)."""


class ParticleException(Exception):
    """An event between two particles: time, the two indices, their distance."""

    def __init__(self, t, obj1, obj2, distance):
        super().__init__(t, obj1, obj2, distance)
        self.t = t
        self.obj1 = obj1
        self.obj2 = obj2
        self.distance = distance


class CloseEncounterException(ParticleException):
    def __str__(self):
        return ("Close encounter detected between particle #%d and particle #%d with a distance of %g at t = %g"
                % (self.obj1, self.obj2, self.distance, self.t))


class CollisionException(ParticleException):
    def __str__(self):
        return ("Collision detected between particle #%d and particle #%d with a distance of %g at t = %g"
                % (self.obj1, self.obj2, self.distance, self.t))
