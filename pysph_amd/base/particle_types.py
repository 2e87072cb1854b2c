"""Particle types and tags (base/particle_types.py:1-38, base/particle_array.pxd:20-26)."""


class ParticleType:
    """Enum of particle array types, the values of base/particle_types.py:24-28."""
    Fluid = 0
    Solid = 1
    DummyFluid = 2
    Probe = 3
    Boundary = 4

    def __init__(self):
        raise SystemError('Do not instantiate the ParticleType class')


# ParticleTag enum, base/particle_array.pxd:20-26
LocalReal = 0
Dummy = 1
LocalDummy = 2
RemoteReal = 3
RemoteDummy = 4
GhostParticle = 5


def is_local_real(tag):
    return tag == LocalReal


def is_local_dummy(tag):
    return tag == LocalDummy


def is_remote_real(tag):
    return tag == RemoteReal
