"""Population evaluation for gantry design searches (the geatpy loop).

``GantryPopulationProblem.run(Vars)`` is the drop-in for ``run(params)`` of
final_code/work/optim/A05run.py:227-340 (called from the geatpy problem's
``aimFunc``, A05geatpy_problem.py:28-30): every row of ``Vars`` (27 values) is
one HUST gantry design (``create_gantry_beamline``, A05run.py:26-180); all
designs track the same phase-space particles (4 momentum dispersions x 12
particles x 2 planes) through the whole beamline as one batched device
population; the end-plane phase space is projected on the device; and the 11
objectives per design are the reference's (absolute maxima and centres of x,
xp, y, yp in mm / mrad, |max|x| - max|y||, and the spread of the per-group
half widths).  geatpy itself is optional: ``run`` takes any numpy population,
and ``geatpy_problem()`` builds the geatpy ``Problem`` only if geatpy imports.
"""
from __future__ import annotations

from typing import List, Sequence

import numpy
import torch

from . import engine
from .beam import BeamGroups
from .constants import MM, MRAD
from .hust_sc_gantry import HUST_SC_GANTRY
from .magnets import device_parts
from .particles import ParticleFactory, PhaseSpaceParticle

# variable bounds and types of the geatpy problem (A05geatpy_problem.py:9-25)
DIM = 27
VAR_TYPES = [0] * 12 + [1] * 2 + [0] * 10 + [1] * 3
LB = [-10, -10, -100, -100, 60, 60, 80, 60, 60, 80, -15000, 4000, 15, 15] + [-10, -100] + \
     [60, 60, 80, 60, 60, 80] + [5000, -12000] + [20, 35, 29]
UB = [10, 10, 100, 100, 120, 120, 100, 120, 120, 100, -6000, 16000, 25, 25] + [10, 100] + \
     [120, 120, 100, 120, 120, 100] + [15000, -5000] + [30, 45, 39]
OBJECTIVES = ("width_max_x", "width_max_xp", "width_max_y", "width_max_yp", "center_x", "center_xp", "center_y",
              "center_yp", "diff_xy", "fluct_x", "fluct_y")


def create_gantry_beamline(param: Sequence[float] = (), part_per_winding: int = 60):
    """The gantry of one 27-value design vector; no vector gives the default gantry (A05run.py:26-180)."""
    p = list(param)
    have = len(p) > 0

    def v(i, default):
        return p[i] if have else default

    ag1, ag2 = (int(p[12]), int(p[13])) if have else (20, 20)
    ag3, ag4, ag5 = (int(p[24]), int(p[25]), int(p[26])) if have else (20, 20, 20)
    ag345 = ag3 + ag4 + ag5
    return HUST_SC_GANTRY(
        DL1=0.8541973, GAP1=0.45, GAP2=0.45,
        qs1_length=0.27, qs1_aperture_radius=60 * MM, qs1_gradient=v(0, 0), qs1_second_gradient=v(2, 0),
        qs2_length=0.27, qs2_aperture_radius=60 * MM, qs2_gradient=v(1, 0), qs2_second_gradient=v(3, 0),
        cct12_big_r=0.95,
        agcct12_inner_small_r=92.5 * MM - 20 * MM + 17 * MM, agcct12_outer_small_r=108.5 * MM - 20 * MM + 17 * MM,
        dicct12_inner_small_r=124.5 * MM - 20 * MM + 17 * MM, dicct12_outer_small_r=140.5 * MM - 20 * MM + 17 * MM,
        agcct1_winding_number=ag1, agcct2_winding_number=ag2, dicct12_winding_number=42,
        dicct12_bending_angle=22.5, agcct1_bending_angle=22.5 * (ag1 / (ag1 + ag2)),
        agcct2_bending_angle=22.5 * (ag2 / (ag1 + ag2)),
        dicct12_tilt_angles=[30, v(4, 90), v(5, 90), v(6, 90)],
        agcct12_tilt_angles=[v(7, 90), 30, v(8, 90), v(9, 90)],
        dicct12_current=v(10, 10000), agcct12_current=v(11, 10000),
        DL2=2.6, GAP3=0.45,
        qs3_length=0.27, qs3_aperture_radius=60 * MM, qs3_gradient=v(14, 0), qs3_second_gradient=v(15, 0),
        cct345_big_r=0.95,
        agcct345_inner_small_r=92.5 * MM + 0.1 * MM + 17 * MM, agcct345_outer_small_r=108.5 * MM + 0.1 * MM + 17 * MM,
        dicct345_inner_small_r=124.5 * MM + 0.1 * MM + 17 * MM,
        dicct345_outer_small_r=140.5 * MM + 0.1 * MM + 17 * MM,
        agcct3_winding_number=ag3, agcct4_winding_number=ag4, agcct5_winding_number=ag5,
        dicct345_winding_number=128,
        dicct345_bending_angle=-67.5, agcct3_bending_angle=-67.5 * (ag3 / ag345),
        agcct4_bending_angle=-67.5 * (ag4 / ag345), agcct5_bending_angle=-67.5 * (ag5 / ag345),
        dicct345_tilt_angles=[30, v(16, 90), v(17, 90), v(18, 90)],
        agcct345_tilt_angles=[v(19, 90), 30, v(20, 90), v(21, 90)],
        dicct345_current=v(22, 10000), agcct345_current=v(23, 10000),
        part_per_winding=part_per_winding,
    ).create_total_beamline()


class GantryPopulationProblem:
    """Batched evaluation of gantry design populations on the GPU."""

    def __init__(self, momentum_dispersions: Sequence[float] = (-0.05, -0.0167, 0.0167, 0.05),
                 particle_number_per_plane_per_dp: int = 12, footstep: float = 50 * MM, kinetic_MeV: float = 215,
                 precision: int = 64, part_per_winding: int = 60):
        self.part_per_winding = part_per_winding
        self.footstep = footstep
        self.precision = precision
        self.group = particle_number_per_plane_per_dp
        default = create_gantry_beamline(part_per_winding=part_per_winding)
        self.length = default.get_length()
        self.ip_start = ParticleFactory.create_proton_along(default, 0.0, kinetic_MeV)
        self.ip_end = ParticleFactory.create_proton_along(default, self.length, kinetic_MeV)
        pps = []
        for dp in momentum_dispersions:
            pps += PhaseSpaceParticle.phase_space_particles_along_positive_ellipse_in_xxp_plane(
                3.5 * MM, 7.5 * MM, dp, particle_number_per_plane_per_dp)
            pps += PhaseSpaceParticle.phase_space_particles_along_positive_ellipse_in_yyp_plane(
                3.5 * MM, 7.5 * MM, dp, particle_number_per_plane_per_dp)
        self.particles = ParticleFactory.create_from_phase_space_particles(
            self.ip_start, self.ip_start.get_natural_coordinate_system(), pps)

    def beamlines(self, params: numpy.ndarray) -> list:
        return [create_gantry_beamline(row, self.part_per_winding) for row in numpy.asarray(params)]

    def track(self, params: numpy.ndarray):
        """Final states (6, D*n) on the device and the groups/projection helper."""
        params = numpy.atleast_2d(numpy.asarray(params, numpy.float64))
        if params.shape[1] != DIM:
            raise ValueError(f"each design needs {DIM} variables")
        bls = self.beamlines(params)
        D, n = len(bls), len(self.particles)
        model = engine.DeviceModel([device_parts(bl).spec() for bl in bls], fp32=(self.precision != 64))
        dev = engine.device()
        one = numpy.array([p.position.to_list() + p.velocity.to_list() for p in self.particles])
        state = torch.from_numpy(numpy.ascontiguousarray(numpy.tile(one, (D, 1)).T)).to(dev)
        sd = [engine.rk4_steps(self.length, self.footstep, p.speed) for p in self.particles] * D
        kq = torch.tensor([p.e / p.relativistic_mass for p in self.particles] * D, dtype=torch.float64, device=dev)
        dt = torch.tensor([d for _, d in sd], dtype=torch.float64, device=dev)
        ns = torch.tensor([s for s, _ in sd], dtype=torch.int64, device=dev)
        plan = model.plan([n] * D, engine.block_shape_for([n], self.precision))
        scratch, partial = model.workspace(plan)
        model.track_tensors(plan, self.precision, state, kq, dt, ns, int(ns.max()), scratch, partial)
        frame = self.ip_end.get_natural_coordinate_system()
        bg = BeamGroups([self.ip_end] * D, [frame] * D, [n] * D, [p.speed for p in self.particles] * D,
                        [p.compute_scalar_momentum() for p in self.particles] * D, group_size=self.group)
        # the reference resets each final speed to |v| before projecting (A05run.py:268-270)
        bg.speed = torch.linalg.vector_norm(state[3:6], dim=0).contiguous()
        return state, bg

    def run(self, params: numpy.ndarray) -> numpy.ndarray:
        """ObjV (D, 11) for a population Vars (D, 27)."""
        state, bg = self.track(params)
        pps, stats = bg.project(state)
        ph = pps.cpu().numpy()
        st = stats.cpu().numpy()
        D = numpy.atleast_2d(params).shape[0]
        n = len(self.particles)
        x = ph[0].reshape(D, n) / MM
        xp = ph[1].reshape(D, n) / MRAD
        y = ph[2].reshape(D, n) / MM
        yp = ph[3].reshape(D, n) / MRAD
        hw = ((st[:, 2] - st[:, 1]) / 2 / MM).reshape(D, -1)   # half width of x per group
        hwy = ((st[:, 4] - st[:, 3]) / 2 / MM).reshape(D, -1)
        wx, wxp = numpy.abs(x).max(1), numpy.abs(xp).max(1)
        wy, wyp = numpy.abs(y).max(1), numpy.abs(yp).max(1)
        obj = numpy.stack([wx, wxp, wy, wyp,
                           numpy.abs(x.sum(1) / n), numpy.abs(xp.sum(1) / n), numpy.abs(y.sum(1) / n),
                           numpy.abs(yp.sum(1) / n),
                           numpy.abs(wx - wy),
                           hw.max(1) - hw.min(1), hwy.max(1) - hwy.min(1)], axis=1)
        return obj


def geatpy_problem(problem: GantryPopulationProblem = None):
    """A geatpy ``Problem`` whose ``aimFunc`` calls ``run`` (A05geatpy_problem.py:6-30).
    Raises ImportError when geatpy is not installed; ``run`` works without it."""
    import geatpy as ea  # optional dependency
    prob = problem or GantryPopulationProblem()

    class Myproblem(ea.Problem):
        def __init__(self):
            M = len(OBJECTIVES)
            ea.Problem.__init__(self, "FirstSection", M, [1] * M, DIM, VAR_TYPES, LB, UB, [1] * DIM, [1] * DIM)

        def aimFunc(self, pop):
            pop.ObjV = prob.run(pop.Phen)

    return Myproblem()
