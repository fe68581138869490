"""cctpy_amd — MI355X-native Biot-Savart field and RK4 proton tracking for CCT gantry beamlines.

The public classes keep the names and call signatures of the reference
toolkit's ``final_code/packages`` (``P2``, ``P3``, ``LocalCoordinateSystem``,
``Trajectory``, ``CCT``, ``QS``, ``Beamline``, ``ParticleRunner``,
``ParticleFactory``, ``PhaseSpaceParticle``, ``GPU_ACCELERATOR``,
``HUST_SC_GANTRY``); fields and tracking run on hand-written HIP kernels in
``libcctamd.so`` (see ``include/cctamd.h``).
"""
from .constants import *  # noqa: F401,F403
from .point import P2, P3, ValueWithDistance
from .base_utils import BaseUtils
from .local_coordinate_system import LocalCoordinateSystem
from .line2s import Line2, StraightLine2, ArcLine2
from .trajectory import Trajectory
from .magnets import ApertureObject, Magnet, UniformMagnet, CombinedMagnet, QS
from .cct import CCT
from .particles import Protons, RunningParticle, ParticleRunner, PhaseSpaceParticle, ParticleFactory
from .beamline import Beamline
from .hust_sc_gantry import HUST_SC_GANTRY
from .gpu_accelerator import GPU_ACCELERATOR

__version__ = "0.1.0"
