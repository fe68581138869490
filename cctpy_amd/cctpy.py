"""Facade mirroring final_code/cctpy.py: ``from cctpy_amd.cctpy import *``."""
from . import *  # noqa: F401,F403
from . import (P2, P3, ValueWithDistance, BaseUtils, LocalCoordinateSystem, Line2, StraightLine2, ArcLine2,  # noqa
               Trajectory, ApertureObject, Magnet, UniformMagnet, CombinedMagnet, QS, CCT, Protons,
               RunningParticle, ParticleRunner, PhaseSpaceParticle, ParticleFactory, Beamline, HUST_SC_GANTRY,
               GPU_ACCELERATOR)
from .constants import M, MM, LIGHT_SPEED, RAD, MRAD, J, eV, MeV, MeV_PER_C  # noqa: F401
