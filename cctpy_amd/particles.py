"""Protons, particle state, phase-space mapping and GPU RK4 tracking
(final_code/packages/particles.py:29-1616, tracking path subset).

``ParticleRunner`` keeps the reference's signatures but every call is one
batched launch sequence of the HIP RK4 kernels over all given particles (FP64
Biot-Savart and FP64 state), instead of a per-particle Python loop or a
process pool.
"""
from __future__ import annotations

import math
import warnings
from typing import List, Optional, Sequence, Union

import numpy

from .base_utils import BaseUtils
from .constants import LIGHT_SPEED, MeV, MM, MRAD
from .line2s import Line2
from .local_coordinate_system import LocalCoordinateSystem
from .magnets import Magnet, device_parts
from .point import P2, P3


class Protons:
    """Proton constants and kinematics (particles.py:29-283)."""

    STATIC_MASS_KG = 1.672621898e-27
    STATIC_ENERGY_J = STATIC_MASS_KG * LIGHT_SPEED * LIGHT_SPEED
    STATIC_ENERGY_eV = STATIC_ENERGY_J / (1.6021766208e-19)
    STATIC_ENERGY_MeV = 938.2720813
    CHARGE_QUANTITY = 1.6021766208e-19

    @classmethod
    def get_total_energy_MeV(cls, kinetic_energy_MeV: float) -> float:
        return cls.STATIC_ENERGY_MeV + kinetic_energy_MeV

    @classmethod
    def get_total_energy_J(cls, kinetic_energy_MeV: float) -> float:
        return cls.get_total_energy_MeV(kinetic_energy_MeV) * MeV

    @classmethod
    def get_relativistic_mass(cls, kinetic_energy_MeV: float) -> float:
        return cls.get_total_energy_J(kinetic_energy_MeV) / LIGHT_SPEED / LIGHT_SPEED

    @classmethod
    def get_speed_m_per_s(cls, kinetic_energy_MeV: float) -> float:
        return LIGHT_SPEED * math.sqrt(1 - (cls.STATIC_MASS_KG / cls.get_relativistic_mass(kinetic_energy_MeV)) ** 2)

    @classmethod
    def get_momentum_kg_m_pre_s(cls, kinetic_energy_MeV: float) -> float:
        return cls.get_relativistic_mass(kinetic_energy_MeV) * cls.get_speed_m_per_s(kinetic_energy_MeV)

    @classmethod
    def getMomentum_MeV_pre_c(cls, kinetic_energy_MeV: float) -> float:
        from .constants import MeV_PER_C
        return cls.get_momentum_kg_m_pre_s(kinetic_energy_MeV) / MeV_PER_C

    @classmethod
    def get_magnetic_stiffness(cls, kinetic_energy_MeV: float) -> float:
        return cls.get_momentum_kg_m_pre_s(kinetic_energy_MeV) / cls.CHARGE_QUANTITY

    @classmethod
    def get_kinetic_energy_MeV(cls, momentum_KG_M_PER_S: float) -> float:
        speed = momentum_KG_M_PER_S / math.sqrt(cls.STATIC_MASS_KG ** 2 + (momentum_KG_M_PER_S / LIGHT_SPEED) ** 2)
        rm = cls.STATIC_MASS_KG / math.sqrt(1 - (speed / LIGHT_SPEED) ** 2)
        return (rm * LIGHT_SPEED * LIGHT_SPEED - cls.STATIC_ENERGY_J) / MeV

    @classmethod
    def get_kinetic_energy_MeV_after_momentum_dispersion(cls, old_kinetic_energy_MeV: float,
                                                         momentum_dispersion: float) -> float:
        p = cls.get_momentum_kg_m_pre_s(old_kinetic_energy_MeV) * (1 + momentum_dispersion)
        return cls.get_kinetic_energy_MeV(p)

    @classmethod
    def convert_momentum_dispersion_to_energy_dispersion(cls, momentum_dispersion: float,
                                                         kinetic_energy_MeV: float) -> float:
        k = (kinetic_energy_MeV + cls.STATIC_ENERGY_MeV) / (kinetic_energy_MeV + 2 * cls.STATIC_ENERGY_MeV)
        return momentum_dispersion / k

    @classmethod
    def convert_energy_dispersion_to_momentum_dispersion(cls, energyDispersion: float,
                                                         kineticEnergy_MeV: float) -> float:
        k = (kineticEnergy_MeV + cls.STATIC_ENERGY_MeV) / (kineticEnergy_MeV + 2 * cls.STATIC_ENERGY_MeV)
        return energyDispersion * k


class RunningParticle:
    """Particle in the global frame: position, velocity, relativistic mass, charge, speed, distance."""

    def __init__(self, position: P3, velocity: P3, relativistic_mass: float, e: float, speed: float,
                 distance: float = 0.0):
        self.position = position
        self.velocity = velocity
        self.relativistic_mass = relativistic_mass
        self.e = e
        self.speed = speed
        self.distance = distance

    def copy(self) -> "RunningParticle":
        return RunningParticle(self.position.copy(), self.velocity.copy(), self.relativistic_mass, self.e,
                               self.speed, self.distance)

    def compute_scalar_momentum(self) -> float:
        return self.speed * self.relativistic_mass

    def change_scalar_momentum(self, scalar_momentum: float) -> None:
        """New speed and mass at fixed rest mass and direction (particles.py:380-422)."""
        m0 = self.relativistic_mass * math.sqrt(1 - (self.speed ** 2) / (LIGHT_SPEED ** 2))
        new_speed = scalar_momentum / math.sqrt(m0 ** 2 + (scalar_momentum / LIGHT_SPEED) ** 2)
        self.relativistic_mass = m0 / math.sqrt(1 - (new_speed / LIGHT_SPEED) ** 2)
        self.speed = new_speed
        self.velocity = self.velocity.change_length(new_speed)
        BaseUtils.equal(scalar_momentum, self.compute_scalar_momentum(), err=1e-6,
                        msg="RunningParticle.change_scalar_momentum: momentum mismatch")
        BaseUtils.equal(self.speed, self.velocity.length(),
                        msg="RunningParticle.change_scalar_momentum: speed mismatch")

    def get_natural_coordinate_system(self, y_direction: P3 = None) -> LocalCoordinateSystem:
        return LocalCoordinateSystem.create_by_y_and_z_direction(
            self.position, P3.z_direct() if y_direction is None else y_direction, self.velocity)

    def __str__(self) -> str:
        return f"p={self.position},v={self.velocity},v0={self.speed}"

    __repr__ = __str__

    def to_numpy_array_data(self, numpy_dtype=numpy.float64) -> numpy.ndarray:
        """[px, py, pz, vx, vy, vz, rm, e, speed, distance] (particles.py:443-457)."""
        return numpy.array(self.position.to_list() + self.velocity.to_list()
                           + [self.relativistic_mass, self.e, self.speed, self.distance], dtype=numpy_dtype)

    def detailed_info(self) -> str:
        return (f"Particle[p={self.position}, v={self.velocity}], rm={self.relativistic_mass}, e={self.e}, "
                f"speed={self.speed}, distance={self.distance}]")

    @staticmethod
    def from_numpy_array_data(a) -> "RunningParticle":
        return RunningParticle(P3(a[0], a[1], a[2]), P3(a[3], a[4], a[5]), float(a[6]), float(a[7]), float(a[8]),
                               float(a[9]))

    def populate(self, other: "RunningParticle") -> None:
        self.position.populate(other.position)
        self.velocity.populate(other.velocity)
        self.relativistic_mass = other.relativistic_mass
        self.e = other.e
        self.speed = other.speed
        self.distance = other.distance

    def __sub__(self, o: "RunningParticle") -> "RunningParticle":
        return RunningParticle(self.position - o.position, self.velocity - o.velocity,
                               self.relativistic_mass - o.relativistic_mass, self.e - o.e, self.speed - o.speed,
                               self.distance - o.distance)


def track_particles(ps: Sequence[RunningParticle], magnets: Sequence[Magnet], design_of: Sequence[int],
                    length: float, footstep: float, precision: int = 64, record: bool = False,
                    ppl: Optional[int] = None):
    """Track particles (each through the design index design_of[i]) on the GPU.

    Returns (state (n, 6), record or None, per-particle step counts and dts),
    without mutating the inputs.  Designs holding magnets the device does not
    know (user subclasses of Magnet) are tracked with those magnets' fields
    added on the host at every RK4 stage (``_track_with_host_magnets``).
    ``ppl`` (points-per-lane code) defaults to the block shape that fits the
    largest design's particle count.
    """
    from . import engine
    specs, foreign = [], []
    for m in magnets:
        parts = device_parts(m)
        foreign.append(parts.foreign)
        specs.append(parts.spec())
    order = sorted(range(len(ps)), key=lambda i: design_of[i])
    npt = numpy.bincount(numpy.asarray(design_of, numpy.int64), minlength=len(magnets))
    state = numpy.zeros((len(ps), 6))
    kq = numpy.zeros(len(ps))
    dt = numpy.zeros(len(ps))
    ns = numpy.zeros(len(ps), numpy.int64)
    for row, i in enumerate(order):
        p = ps[i]
        state[row, :3] = p.position.to_list()
        state[row, 3:] = p.velocity.to_list()
        kq[row] = p.e / p.relativistic_mass
        ns[row], dt[row] = engine.rk4_steps(length, footstep, p.speed)
    model = engine.model_for(specs, fp32=(precision != 64))
    if ppl is None:
        ppl = engine.block_shape_for(npt, precision)
    if any(foreign):
        design_rows = numpy.repeat(numpy.arange(len(magnets)), npt)
        out, rec = _track_with_host_magnets(model, [foreign[d] for d in design_rows], state, kq, dt, ns, npt,
                                            precision, ppl, record)
    else:
        out, rec = model.track(state, kq, dt, ns, npt, precision=precision, ppl=ppl, record=record)
    inv = numpy.empty(len(ps), numpy.int64)
    inv[numpy.asarray(order, numpy.int64)] = numpy.arange(len(ps))
    out = out[inv]
    if rec is not None:
        rec = rec[:, inv, :]
    return out, rec, ns[inv], dt[inv]


def _track_with_host_magnets(model, foreign_of, state, kq, dt, ns, npt, precision, ppl, record):
    """RK4 tracking through designs that hold magnets the device does not know (user subclasses of
    Magnet, whose field only their own ``magnetic_field_at`` gives): at every RK4 stage the device
    sums the CCT, QS and uniform fields at all particles' stage points in one launch, each
    particle's foreign magnets add theirs through their Python method (the reference's callback
    contract, particles.py:514-537), and the stage combination follows the reference's RK4
    (base_utils.py:303-309) in the device stage kernel's operation order.  foreign_of[i] lists
    the foreign magnets of particle i's design."""
    P = state.shape[0]
    p0, v0 = state[:, :3].copy(), state[:, 3:].copy()
    max_steps = int(ns.max()) if P else 0
    rec = numpy.zeros((max_steps + 1, P, 6)) if record else None
    if record:
        rec[0] = state
    k, h = kq[:, None], dt[:, None]
    h2, h6 = h / 2, h / 6

    def accel(p, v):
        b = model.field(p, npt, precision=precision, ppl=ppl)
        for i, mags in enumerate(foreign_of):
            for m in mags:
                f = m.magnetic_field_at(P3(*p[i]))
                b[i] = b[i] + numpy.array([f.x, f.y, f.z])
        cr = numpy.stack([v[:, 1] * b[:, 2] - v[:, 2] * b[:, 1], -v[:, 0] * b[:, 2] + v[:, 2] * b[:, 0],
                          v[:, 0] * b[:, 1] - v[:, 1] * b[:, 0]], axis=1)
        return cr * k

    for step in range(max_steps):
        live = (step < ns)[:, None]
        a = accel(p0, v0)
        kp, kv = v0, a
        p1, v1 = p0 + v0 * h2, v0 + a * h2
        a = accel(p1, v1)
        kp, kv = kp + v1 * 2, kv + a * 2
        p2, v2 = p0 + v1 * h2, v0 + a * h2
        a = accel(p2, v2)
        kp, kv = kp + v2 * 2, kv + a * 2
        p3, v3 = p0 + v2 * h, v0 + a * h
        a = accel(p3, v3)
        p0 = numpy.where(live, p0 + (kp + v3) * h6, p0)
        v0 = numpy.where(live, v0 + (kv + a) * h6, v0)
        if record:
            rec[step + 1] = numpy.where(live, numpy.concatenate([p0, v0], axis=1), rec[step + 1])
    return numpy.concatenate([p0, v0], axis=1), rec


class ParticleRunner:
    """RK4 tracking through a magnet (particles.py:509-870), run on the GPU."""

    @staticmethod
    def run_only(p: Union[RunningParticle, List[RunningParticle]], m: Magnet, length: float,
                 footstep: float = 20 * MM, concurrency_level: int = 1,
                 report: bool = True) -> Union[RunningParticle, List[RunningParticle]]:
        """Move p (one particle or a list) by ``length`` along its path in m's field, step ``footstep``.

        Per particle: dt = footstep/speed, t_end = length/speed, ceil(t_end/dt)
        RK4 steps of dt = t_end/steps; position and velocity are updated in
        place and ``distance`` grows by ``length`` (particles.py:585-596).
        ``concurrency_level`` and ``report`` are accepted for call compatibility.
        """
        single = isinstance(p, RunningParticle)
        ps = [p] if single else list(p)
        if not ps:
            return p
        out, _, _, _ = track_particles(ps, [m], [0] * len(ps), length, footstep)
        for q, row in zip(ps, out):
            q.position = P3(*row[:3])
            q.velocity = P3(*row[3:])
            q.distance += length
        return p

    @staticmethod
    def run_only_ode(*args, **kwargs):
        raise NotImplementedError("run_only_ode (scipy adaptive steps) is not part of the GPU tracking path")

    @staticmethod
    def run_get_trajectory(p: RunningParticle, m: Magnet, length: float, footstep: float = 20 * MM) -> List[P3]:
        """Positions after every RK4 step, starting point included (particles.py:680-711)."""
        _, rec, ns, _ = track_particles([p], [m], [0], length, footstep, record=True)
        p.distance += length
        return [P3(*rec[k, 0, :3]) for k in range(int(ns[0]) + 1)]

    @staticmethod
    def run_get_all_info(p: Union[RunningParticle, List[RunningParticle]], m: Magnet, length: float,
                         footstep: float = 1 * MM, concurrency_level: Optional[int] = None,
                         report: bool = False):
        """Particle copies after every RK4 step with distance = d0 + t*speed (particles.py:713-778)."""
        single = isinstance(p, RunningParticle)
        ps = [p] if single else list(p)
        _, rec, ns, dts = track_particles(ps, [m], [0] * len(ps), length, footstep, record=True)
        result = []
        for i, q in enumerate(ps):
            d0 = q.distance
            q.distance += length
            t = 0.0
            infos = []
            for k in range(int(ns[i]) + 1):
                c = q.copy()
                c.position = P3(*rec[k, i, :3])
                c.velocity = P3(*rec[k, i, 3:])
                c.distance = d0 + t * c.speed
                infos.append(c)
                t = t + dts[i]
            result.append(infos)
        return result[0] if single else result


class PhaseSpaceParticle:
    """(x, xp, y, yp, z, delta) relative to an ideal particle (particles.py:873-1315)."""

    XXP_PLANE = 1
    YYP_PLANE = 2

    def __init__(self, x: float = 0.0, xp: float = 0.0, y: float = 0.0, yp: float = 0.0, z: float = 0.0,
                 delta: float = 0.0):
        self.x, self.xp, self.y, self.yp, self.z, self.delta = x, xp, y, yp, z, delta

    def project_to_xxp_plane(self, convert_to_mm: bool = False) -> P2:
        k = 1000 if convert_to_mm else 1
        return P2(self.x * k, self.xp * k)

    def project_to_yyp_plane(self, convert_to_mm: bool = False) -> P2:
        k = 1000 if convert_to_mm else 1
        return P2(self.y * k, self.yp * k)

    def project_to_plane(self, plane_id: int, convert_to_mm: bool = False) -> P2:
        if plane_id == PhaseSpaceParticle.XXP_PLANE:
            return self.project_to_xxp_plane(convert_to_mm)
        if plane_id == PhaseSpaceParticle.YYP_PLANE:
            return self.project_to_yyp_plane(convert_to_mm)
        raise ValueError(f"unknown plane_id({plane_id})")

    @staticmethod
    def phase_space_particles_along_positive_ellipse_in_xxp_plane(xMax: float, xpMax: float, delta: float,
                                                                   number: int) -> List["PhaseSpaceParticle"]:
        e = BaseUtils.Ellipse(1 / (xMax ** 2), 0, 1 / (xpMax ** 2), 1)
        return [PhaseSpaceParticle(q.x, q.y, 0, 0, 0, delta) for q in e.uniform_distribution_points_along_edge(number)]

    @staticmethod
    def phase_space_particles_along_positive_ellipse_in_yyp_plane(yMax: float, ypMax: float, delta: float,
                                                                   number: int) -> List["PhaseSpaceParticle"]:
        e = BaseUtils.Ellipse(1 / (yMax ** 2), 0, 1 / (ypMax ** 2), 1)
        return [PhaseSpaceParticle(0, 0, q.x, q.y, 0, delta) for q in e.uniform_distribution_points_along_edge(number)]

    @staticmethod
    def phase_space_particles_along_positive_ellipse_in_plane(plane_id: int, xMax: float, xpMax: float,
                                                              delta: float, number: int):
        if plane_id == PhaseSpaceParticle.XXP_PLANE:
            return PhaseSpaceParticle.phase_space_particles_along_positive_ellipse_in_xxp_plane(xMax, xpMax, delta,
                                                                                                number)
        if plane_id == PhaseSpaceParticle.YYP_PLANE:
            return PhaseSpaceParticle.phase_space_particles_along_positive_ellipse_in_yyp_plane(xMax, xpMax, delta,
                                                                                                number)
        raise ValueError(f"unknown plane_id({plane_id})")

    @staticmethod
    def phase_space_particles_project_to_xxp_plane(pps: List, convert_to_mm: bool = False) -> List[P2]:
        return [q.project_to_xxp_plane(convert_to_mm) for q in pps]

    @staticmethod
    def phase_space_particles_project_to_yyp_plane(pps: List, convert_to_mm: bool = False) -> List[P2]:
        return [q.project_to_yyp_plane(convert_to_mm) for q in pps]

    @staticmethod
    def phase_space_particles_project_to_plane(pps: List, plane_id: int, convert_to_mm: bool = False) -> List[P2]:
        if plane_id == PhaseSpaceParticle.XXP_PLANE:
            return PhaseSpaceParticle.phase_space_particles_project_to_xxp_plane(pps, convert_to_mm)
        if plane_id == PhaseSpaceParticle.YYP_PLANE:
            return PhaseSpaceParticle.phase_space_particles_project_to_yyp_plane(pps, convert_to_mm)
        raise ValueError(f"unknown plane_id({plane_id})")

    @staticmethod
    def create_from_running_particle(ideal_particle: RunningParticle, coordinate_system: LocalCoordinateSystem,
                                     running_particle: RunningParticle) -> "PhaseSpaceParticle":
        """Projection into the ideal particle's frame; xp = v_x/sqrt(speed^2 - v_x^2) (particles.py:1077-1121)."""
        rel = running_particle.position - ideal_particle.position
        x, y, z = coordinate_system.XI * rel, coordinate_system.YI * rel, coordinate_system.ZI * rel
        dv = running_particle.velocity - ideal_particle.velocity
        vx, vy = coordinate_system.XI * dv, coordinate_system.YI * dv
        xp = vx / math.sqrt(running_particle.speed ** 2 - vx ** 2)
        yp = vy / math.sqrt(running_particle.speed ** 2 - vy ** 2)
        rm, im = running_particle.compute_scalar_momentum(), ideal_particle.compute_scalar_momentum()
        return PhaseSpaceParticle(x, xp, y, yp, z, (rm - im) / im)

    @staticmethod
    def create_from_running_particles(ideal_particle: RunningParticle, coordinate_system: LocalCoordinateSystem,
                                      running_particles: List[RunningParticle]) -> List["PhaseSpaceParticle"]:
        return [PhaseSpaceParticle.create_from_running_particle(ideal_particle, coordinate_system, r)
                for r in running_particles]

    @staticmethod
    def convert_delta_from_momentum_dispersion_to_energy_dispersion(pp: "PhaseSpaceParticle",
                                                                    centerKineticEnergy_MeV) -> "PhaseSpaceParticle":
        c = pp.copy()
        c.delta = Protons.convert_momentum_dispersion_to_energy_dispersion(c.delta, centerKineticEnergy_MeV)
        return c

    @staticmethod
    def convert_delta_from_momentum_dispersion_to_energy_dispersion_for_list(pps, centerKineticEnergy_MeV):
        return [PhaseSpaceParticle.convert_delta_from_momentum_dispersion_to_energy_dispersion(
            q, centerKineticEnergy_MeV) for q in pps]

    @staticmethod
    def convert_delta_from_energy_dispersion_to_momentum_dispersion(pp: "PhaseSpaceParticle",
                                                                    centerKineticEnergy_MeV: float):
        c = pp.copy()
        c.delta = Protons.convert_energy_dispersion_to_momentum_dispersion(c.delta, centerKineticEnergy_MeV)
        return c

    @staticmethod
    def convert_delta_from_energy_dispersion_to_momentum_dispersion_for_list(pps, centerKineticEnergy_MeV: float):
        return [PhaseSpaceParticle.convert_delta_from_energy_dispersion_to_momentum_dispersion(
            q, centerKineticEnergy_MeV) for q in pps]

    def __str__(self) -> str:
        return f"x={self.x},xp={self.xp},y={self.y},yp={self.yp},z={self.z},d={self.delta}"

    __repr__ = __str__

    def copy(self) -> "PhaseSpaceParticle":
        return PhaseSpaceParticle(self.x, self.xp, self.y, self.yp, self.z, self.delta)

    def getDelta(self) -> float:
        return self.delta

    def get_length(self) -> float:
        return math.sqrt(self.x ** 2 + self.xp ** 2 + self.y ** 2 + self.yp ** 2 + self.z ** 2 + self.delta ** 2)


class ParticleFactory:
    """Proton construction helpers (particles.py:1318-1616)."""

    @staticmethod
    def create_proton(position: P3, direct: P3, kinetic_MeV: float = 250) -> RunningParticle:
        speed = LIGHT_SPEED * math.sqrt(
            1.0 - (Protons.STATIC_ENERGY_MeV / (Protons.STATIC_ENERGY_MeV + kinetic_MeV)) ** 2)
        rm = Protons.STATIC_MASS_KG / math.sqrt(1.0 - (speed ** 2) / (LIGHT_SPEED ** 2))
        return RunningParticle(position, direct.copy().change_length(speed), rm, Protons.CHARGE_QUANTITY, speed)

    @staticmethod
    def create_proton_by_position_and_velocity(position: P3, velocity: P3) -> RunningParticle:
        speed = velocity.length()
        rm = Protons.STATIC_MASS_KG / math.sqrt(1.0 - (speed ** 2) / (LIGHT_SPEED ** 2))
        return RunningParticle(position, velocity, rm, Protons.CHARGE_QUANTITY, speed)

    @staticmethod
    def create_proton_along(trajectory: Line2, s: float = 0.0, kinetic_MeV: float = 250) -> RunningParticle:
        return ParticleFactory.create_proton(trajectory.point_at(s).to_p3(), trajectory.direct_at(s).to_p3(),
                                             kinetic_MeV=kinetic_MeV)

    @staticmethod
    def create_from_phase_space_particle(ideal_particle: RunningParticle, coordinate_system: LocalCoordinateSystem,
                                         phase_space_particle: PhaseSpaceParticle) -> RunningParticle:
        """Phase space -> particle: offsets along the frame axes, momentum scaled by (1+delta),
        angles added to the velocity, speed restored (particles.py:1397-1437)."""
        pp = phase_space_particle
        p = ideal_particle.copy()
        p.position += coordinate_system.XI * pp.x
        p.position += coordinate_system.YI * pp.y
        p.position += coordinate_system.ZI * pp.z
        if pp.delta != 0.0:
            p.change_scalar_momentum(p.compute_scalar_momentum() * (1.0 + pp.delta))
        p.velocity += coordinate_system.XI * (pp.xp * p.speed)
        p.velocity += coordinate_system.YI * (pp.yp * p.speed)
        p.velocity = p.velocity.change_length(p.speed)
        return p

    @staticmethod
    def create_from_phase_space_particles(ideal_particle: RunningParticle, coordinate_system: LocalCoordinateSystem,
                                          phase_space_particles: List[PhaseSpaceParticle]) -> List[RunningParticle]:
        return [ParticleFactory.create_from_phase_space_particle(ideal_particle, coordinate_system, q)
                for q in phase_space_particles]
