"""Synthetic workloads of the benchmark configurations.

Each workload is plain data: a list of beamline-builder operations (applied
in order to a builder with the Beamline API) and a seeded beam in phase-space
coordinates.  ``build_beamline`` applies the operations to this package's
``Beamline``; a test can replay the same operations on any other builder with
the same method names.

Configurations (250 MeV protons, RK4, phase space in the ideal particle's
natural frame at the start):

* ``single_cct``   one single-layer bending CCT: R=0.95 m, r=83 mm, 30 deg,
  64 windings x 360 segments (23,040), -9000 A (the sign that bends protons
  along the arc), tilts [30, 80]; 64 protons,
  x,y ~ N(0, 1.75 mm), x',y' ~ N(0, 2 mrad), delta ~ U(+-2%), 1 mm steps
  along the 30 deg arc (0.4974 m, 498 steps).
* ``agcct_section`` a gantry section: drift 0.25 m, QS (0.27 m, +20 T/m),
  drift 0.2 m, AG-CCT (dipole pair 119/107 mm and AG pair 95/83 mm, R=0.95 m,
  -67.5 deg over three AG parts [22, 53, 53] windings, dipole 128 windings,
  720 segments per winding: 368,640 segments, currents 9200/-9200 A, tilts
  seeded in [20, 160] deg), drift 0.2 m, QS (0.27 m, -20 T/m), drift 0.2 m
  (2.509 m); 512 protons, same transverse beam, delta ~ U(+-8%), 1 mm steps.
* ``population``   64 designs of the dipole-pair/AG form: layers in {2, 4},
  windings in [100, 160], {360, 720} segments per winding, currents in
  [8000, 10000] A (AG layers -2000 to -4000 A), tilts [30, t1, t2, t3] with
  harmonic tilts seeded in [70, 110] deg (wider ranges put multipoles as strong
  as the dipole into the dipole layers and the beam into the coils); 72,000 -
  460,800 segments each; each
  design bends by the angle its dipole field sustains for 250 MeV protons on
  R = 0.95 m (41-103 deg) and its 32 protons (single_cct beam) are tracked
  along that arc in 1 mm steps.
* ``hust_total``   the full two-section HUST gantry at 720 segments per winding
  (959,040 segments); 4096 protons, delta ~ U(+-8%), 0.5 mm steps over 6 m.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, List, Sequence, Tuple

import numpy

MM = 0.001
SEEDS = {"single_cct": 20261201, "agcct_section": 20261202, "population": 20261203, "adjoint_section": 20261204,
         "hust_total": 20261205}

Op = Tuple[str, Dict]


@dataclass
class Beam:
    """Phase-space coordinates (n, 5): x, xp, y, yp, delta; ideal particle at s0 with kinetic energy;
    ``length`` overrides the workload's tracking length for this beam."""
    coords: numpy.ndarray
    s0: float
    kinetic_mev: float
    length: float = None


@dataclass
class Workload:
    name: str
    designs: List[List[Op]]
    beams: List[Beam]
    length: float
    footstep: float

    @property
    def n_particles(self) -> int:
        return sum(b.coords.shape[0] for b in self.beams)

    def beam_length(self, i: int) -> float:
        b = self.beams[i]
        return self.length if b.length is None else b.length


def gaussian_beam(rng, n: int, delta: float) -> numpy.ndarray:
    c = numpy.empty((n, 5))
    c[:, 0] = rng.normal(0.0, 1.75 * MM, n)
    c[:, 1] = rng.normal(0.0, 2.0 * MM, n)
    c[:, 2] = rng.normal(0.0, 1.75 * MM, n)
    c[:, 3] = rng.normal(0.0, 2.0 * MM, n)
    c[:, 4] = rng.uniform(-delta, delta, n)
    return c


def _tilt(rng, lo: float = 20.0, hi: float = 160.0) -> float:
    t = float(rng.uniform(lo, hi))
    return t if abs(t - 90.0) > 1e-3 else 90.5


def single_cct(n_particles: int = 64, seed: int = SEEDS["single_cct"]) -> Workload:
    rng = numpy.random.default_rng(seed)
    arc = 0.95 * math.radians(30.0)
    ops: List[Op] = [
        ("first_drift", dict(length=0.1)),
        ("append_single_cct", dict(big_r=0.95, small_r=83 * MM, bending_angle=30.0, tilt_angles=[30.0, 80.0],
                                   winding_number=64, current=-9000.0, disperse_number_per_winding=360)),
        ("append_drift", dict(length=0.1)),
    ]
    return Workload("single_cct", [ops], [Beam(gaussian_beam(rng, n_particles, 0.02), 0.1, 250.0)], arc, 1 * MM)


def agcct_section(n_particles: int = 512, seed: int = SEEDS["agcct_section"],
                  disperse_number_per_winding: int = 720) -> Workload:
    rng = numpy.random.default_rng(seed)
    dip_tilts = [30.0, _tilt(rng), _tilt(rng), _tilt(rng)]
    ag_tilts = [_tilt(rng), 30.0, _tilt(rng), _tilt(rng)]
    ops: List[Op] = [
        ("first_drift", dict(length=0.25)),
        ("append_qs", dict(length=0.27, gradient=20.0, second_gradient=0.0, aperture_radius=60 * MM)),
        ("append_drift", dict(length=0.2)),
        ("append_agcct", dict(big_r=0.95, small_rs=[119 * MM, 107 * MM, 95 * MM, 83 * MM],
                              bending_angles=[-(8 + 3.716404), -(8 + 19.93897), -(8 + 19.844626)],
                              tilt_angles=[dip_tilts, ag_tilts], winding_numbers=[[128], [22, 53, 53]],
                              currents=[9200.0, -9200.0], disperse_number_per_winding=disperse_number_per_winding)),
        ("append_drift", dict(length=0.2)),
        ("append_qs", dict(length=0.27, gradient=-20.0, second_gradient=0.0, aperture_radius=60 * MM)),
        ("append_drift", dict(length=0.2)),
    ]
    length = 0.25 + 0.27 + 0.2 + 0.95 * math.radians(67.5) + 0.2 + 0.27 + 0.2
    return Workload("agcct_section", [ops], [Beam(gaussian_beam(rng, n_particles, 0.08), 0.0, 250.0)], length,
                    1 * MM)


def population(n_designs: int = 64, n_particles: int = 32, seed: int = SEEDS["population"]) -> Workload:
    """Bend angle of each design from its own dipole strength: a CCT dipole layer pair gives
    B = mu0 I cot(t) / pitch with pitch = R theta / W, so theta = mu0 I cot(t) W / (R B_ref)
    with B_ref = p/(qR) of 250 MeV protons; each design is tracked along its own arc."""
    rng = numpy.random.default_rng(seed)
    designs, beams = [], []
    big_r, mu0 = 0.95, 4e-7 * math.pi
    b_ref = 2.5601351895878315  # 250 MeV proton rigidity / 0.95 m
    for _ in range(n_designs):
        layers = int(rng.choice([2, 4]))
        w = int(rng.integers(100, 161))
        n = int(rng.choice([360, 720]))
        i_dip = float(rng.uniform(8000.0, 10000.0))
        tilts = [30.0, _tilt(rng, 70.0, 110.0), _tilt(rng, 70.0, 110.0), _tilt(rng, 70.0, 110.0)]
        theta = round(math.degrees(mu0 * i_dip / math.tan(math.radians(30.0)) * w / (big_r * b_ref)), 1)
        ops: List[Op] = [("first_drift", dict(length=0.1))]
        if layers == 2:
            ops.append(("append_dipole_cct", dict(big_r=big_r, small_r_inner=95 * MM, small_r_outer=107 * MM,
                                                  bending_angle=-theta, tilt_angles=tilts, winding_number=w,
                                                  current=i_dip, disperse_number_per_winding=n)))
        else:
            w1 = w // 3
            ops.append(("append_agcct", dict(big_r=big_r, small_rs=[119 * MM, 107 * MM, 95 * MM, 83 * MM],
                                             bending_angles=[-theta * w1 / w, -theta * (w - w1) / w],
                                             tilt_angles=[tilts, [90.0, 30.0]],
                                             winding_numbers=[[w], [w1, w - w1]],
                                             currents=[i_dip, -float(rng.uniform(2000.0, 4000.0))],
                                             disperse_number_per_winding=n)))
        ops.append(("append_drift", dict(length=0.1)))
        designs.append(ops)
        beams.append(Beam(gaussian_beam(rng, n_particles, 0.02), 0.1, 250.0, big_r * math.radians(theta)))
    return Workload("population", designs, beams, max(b.length for b in beams), 1 * MM)


def hust_total(n_particles: int = 4096, seed: int = SEEDS["hust_total"], length: float = 6.0,
               footstep: float = 0.5 * MM) -> Workload:
    rng = numpy.random.default_rng(seed)
    ops: List[Op] = [("hust_total", dict(part_per_winding=720))]
    return Workload("hust_total", [ops], [Beam(gaussian_beam(rng, n_particles, 0.08), 0.0, 250.0)], length,
                    footstep)


def adjoint_section(n_particles: int = 256, seed: int = SEEDS["adjoint_section"]) -> Workload:
    """The AG-CCT section at its design vector plus a 1.1 m drift to the objective plane
    (rms x + rms y + |dispersion| there, differentiated w.r.t. the 10-parameter vector)."""
    rng = numpy.random.default_rng(seed)
    length = agcct_section(n_particles=1).length + 1.1
    return Workload("adjoint_section", [section_ops(section_theta0())],
                    [Beam(gaussian_beam(rng, n_particles, 0.08), 0.0, 250.0)], length, 1 * MM)


def by_name(name: str, **kw) -> Workload:
    return {"single_cct": single_cct, "agcct_section": agcct_section, "population": population,
            "adjoint_section": adjoint_section, "hust_total": hust_total}[name](**kw)


# ------------------------------------------------------------- builders
def build_beamline(ops: Sequence[Op]):
    """Apply builder operations to this package's Beamline."""
    from .beamline import Beamline
    from .cct import CCT
    from .hust_sc_gantry import HUST_SC_GANTRY
    from .point import P2
    bl = None
    for name, kw in ops:
        if name == "hust_total":
            return HUST_SC_GANTRY(**kw).create_total_beamline()
        if name == "first_drift":
            bl = Beamline.set_start_point(P2.origin()).first_drift(P2.x_direct(), kw["length"])
        elif name == "append_single_cct":
            s0 = bl.get_length()
            bl.trajectory.add_arc_line(radius=kw["big_r"], clockwise=False, angle_deg=kw["bending_angle"])
            cct = CCT.create_cct_along(bl.trajectory, s0, kw["big_r"], kw["small_r"], kw["bending_angle"],
                                       kw["tilt_angles"], kw["winding_number"], kw["current"], P2(0, 0),
                                       P2(2 * math.pi * kw["winding_number"], math.radians(kw["bending_angle"])),
                                       kw["disperse_number_per_winding"])
            bl.magnets.append(cct)
            bl.elements.append((s0, cct, kw["big_r"] * math.radians(kw["bending_angle"])))
        else:
            getattr(bl, name)(**kw)
    return bl


def beam_particles(beam: Beam, bl) -> list:
    """RunningParticles of a beam placed at s0 of the beamline's orbit."""
    from .particles import ParticleFactory, PhaseSpaceParticle
    ip = ParticleFactory.create_proton_along(bl, beam.s0, beam.kinetic_mev)
    cs = ip.get_natural_coordinate_system()
    return [ParticleFactory.create_from_phase_space_particle(ip, cs, PhaseSpaceParticle(c[0], c[1], c[2], c[3], 0.0,
                                                                                        c[4]))
            for c in beam.coords]


# ---------------------------------------------------- design gradients
THETA_NAMES = ["qs1_gradient", "qs2_gradient", "dipole_tilt_2", "dipole_tilt_3", "dipole_tilt_4",
               "ag_tilt_1", "ag_tilt_3", "ag_tilt_4", "dipole_current", "ag_current"]


def section_theta0(seed: int = SEEDS["agcct_section"]) -> numpy.ndarray:
    """The 10-parameter design vector of the agcct_section workload (2 QS gradients, 3 free
    tilts of the dipole pair, 3 of the AG pair, 2 currents)."""
    ops = agcct_section(n_particles=1, seed=seed).designs[0]
    ag = ops[3][1]
    qs1, qs2 = ops[1][1], ops[5][1]
    dt, at = ag["tilt_angles"]
    return numpy.array([qs1["gradient"], qs2["gradient"], dt[1], dt[2], dt[3], at[0], at[2], at[3],
                        ag["currents"][0], ag["currents"][1]], numpy.float64)


def section_ops(theta: Sequence[float], disperse_number_per_winding: int = 720, drift_after: float = 1.1,
                seed: int = SEEDS["agcct_section"]) -> List[Op]:
    """agcct_section builder operations at design vector theta, plus a drift of ``drift_after``."""
    ops = [(n, dict(kw)) for n, kw in agcct_section(n_particles=1, seed=seed,
                                                    disperse_number_per_winding=disperse_number_per_winding).designs[0]]
    t = [float(v) for v in theta]
    ops[1][1]["gradient"] = t[0]
    ops[5][1]["gradient"] = t[1]
    ag = ops[3][1]
    ag["tilt_angles"] = [[30.0, t[2], t[3], t[4]], [t[5], 30.0, t[6], t[7]]]
    ag["currents"] = [t[8], t[9]]
    if drift_after:
        ops.append(("append_drift", dict(length=drift_after)))
    return ops


def beam_spot_objective(final: "torch.Tensor", ideal_position, frame_x, frame_y, delta) -> "torch.Tensor":
    """rms x + rms y + |D| at the end plane, D = least-squares slope of x against delta; x, y are
    offsets from the ideal particle along the ideal particle's frame axes."""
    import torch
    pos = final[:, :3] - torch.as_tensor(ideal_position, dtype=final.dtype)
    x = pos @ torch.as_tensor(frame_x, dtype=final.dtype)
    y = pos @ torch.as_tensor(frame_y, dtype=final.dtype)
    d = torch.as_tensor(delta, dtype=final.dtype)
    dc = d - d.mean()
    xc = x - x.mean()
    slope = (xc * dc).sum() / (dc * dc).sum()
    return torch.sqrt((xc * xc).mean()) + torch.sqrt(((y - y.mean()) ** 2).mean()) + slope.abs()
