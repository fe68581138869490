"""Beamline: design orbit + magnets in append order (final_code/packages/beamline.py:31-804)."""
from __future__ import annotations

import math
from typing import Callable, List, Optional, Tuple

from .base_utils import BaseUtils
from .cct import CCT
from .constants import MM, MRAD
from .line2s import Line2
from .magnets import ApertureObject, DeviceParts, Magnet, QS
from .particles import ParticleFactory, ParticleRunner, PhaseSpaceParticle
from .point import P2, P3, ValueWithDistance
from .trajectory import Trajectory


class Beamline(Line2, Magnet, ApertureObject):
    def __init__(self, trajectory: Trajectory) -> None:
        self.magnets: List[Magnet] = []
        self.trajectory = trajectory
        self.elements: List[Tuple[float, Optional[Magnet], float]] = []

    def magnetic_field_at(self, point: P3) -> P3:
        return Magnet.magnetic_field_at(self, point)

    def _device_parts(self, out: DeviceParts) -> None:
        for m in self.magnets:
            m._device_parts(out)

    def magnetic_field_along(self, line2: Optional[Line2] = None,
                             p2_t0_p3: Callable[[P2], P3] = lambda p2: P3(p2.x, p2.y, 0.0),
                             step: float = 1 * MM) -> List[ValueWithDistance[P3]]:
        return Magnet.magnetic_field_along(self, self.trajectory if line2 is None else line2, p2_t0_p3, step)

    def magnetic_field_bz_along(self, line2: Optional[Line2] = None,
                                p2_t0_p3: Callable[[P2], P3] = lambda p2: P3(p2.x, p2.y, 0.0),
                                step: float = 1 * MM) -> List[P2]:
        return Magnet.magnetic_field_bz_along(self, self.trajectory if line2 is None else line2, p2_t0_p3, step)

    def graident_field_along(self, line2: Optional[Line2] = None, good_field_area_width: float = 10 * MM,
                             step: float = 1 * MM, point_number: int = 4) -> List[P2]:
        """Magnet.graident_field_along on the beamline's own orbit by default (beamline.py:96-115)."""
        return Magnet.graident_field_along(self, self.trajectory if line2 is None else line2, good_field_area_width,
                                           step, point_number)

    def second_graident_field_along(self, line2: Optional[Line2] = None, good_field_area_width: float = 10 * MM,
                                    step: float = 1 * MM, point_number: int = 4) -> List[P2]:
        """Magnet.second_graident_field_along on the orbit by default (beamline.py:117-136)."""
        return Magnet.second_graident_field_along(self, self.trajectory if line2 is None else line2,
                                                  good_field_area_width, step, point_number)

    # --------------------------------------------------------- tracking
    def track_ideal_particle(self, kinetic_MeV: float, s: float = 0.0, length: Optional[float] = None,
                             footstep: float = 5 * MM) -> List[P3]:
        if length is None:
            length = self.trajectory.get_length() - s
        ip = ParticleFactory.create_proton_along(self.trajectory, s, kinetic_MeV)
        return ParticleRunner.run_get_trajectory(ip, self, length, footstep)

    def track_phase_space_particle(self, x_mm: float, xp_mrad: float, y_mm: float, yp_mrad, delta: float,
                                   kinetic_MeV: float, s: float = 0.0, length: Optional[float] = None,
                                   footstep: float = 10 * MM) -> List[ValueWithDistance[PhaseSpaceParticle]]:
        """Phase-space coordinates after every step (beamline.py:158-227)."""
        if length is None:
            length = self.trajectory.get_length() - s
        pp = PhaseSpaceParticle(x=x_mm * MM, xp=xp_mrad * MM, y=y_mm * MM, yp=yp_mrad * MM, z=0.0, delta=delta)
        ip = ParticleFactory.create_proton_along(self.trajectory, s, kinetic_MeV)
        rp = ParticleFactory.create_from_phase_space_particle(ip, ip.get_natural_coordinate_system(), pp)
        out = []
        for cp in ParticleRunner.run_get_all_info(p=rp, m=self, length=length, footstep=footstep):
            cip = ParticleFactory.create_proton_along(self.trajectory, cp.distance + s, kinetic_MeV)
            out.append(ValueWithDistance(PhaseSpaceParticle.create_from_running_particle(
                cip, cip.get_natural_coordinate_system(), cp), cp.distance))
        return out

    def track_phase_ellipse(self, x_sigma_mm: float, xp_sigma_mrad: float, y_sigma_mm: float, yp_sigma_mrad,
                            delta: float, particle_number: int, kinetic_MeV: float, s: float = 0.0,
                            length: Optional[float] = None, footstep: float = 10 * MM, concurrency_level: int = 1,
                            report: bool = True) -> Tuple[List[P2], List[P2]]:
        """Track the edges of an x-xp and a y-yp phase ellipse; returns their end projections in mm/mrad
        (beamline.py:229-326)."""
        if length is None:
            length = self.trajectory.get_length() - s
        ip_start = ParticleFactory.create_proton_along(self.trajectory, s, kinetic_MeV)
        ip_end = ParticleFactory.create_proton_along(self.trajectory, s + length, kinetic_MeV)
        pp_x = PhaseSpaceParticle.phase_space_particles_along_positive_ellipse_in_xxp_plane(
            xMax=x_sigma_mm * MM, xpMax=xp_sigma_mrad * MRAD, delta=delta, number=particle_number)
        pp_y = PhaseSpaceParticle.phase_space_particles_along_positive_ellipse_in_yyp_plane(
            yMax=y_sigma_mm * MM, ypMax=yp_sigma_mrad * MRAD, delta=delta, number=particle_number)
        cs = ip_start.get_natural_coordinate_system()
        rp_x = ParticleFactory.create_from_phase_space_particles(ip_start, cs, pp_x)
        rp_y = ParticleFactory.create_from_phase_space_particles(ip_start, cs, pp_y)
        ParticleRunner.run_only(p=rp_x + rp_y, m=self, length=length, footstep=footstep,
                                concurrency_level=concurrency_level, report=report)
        ce = ip_end.get_natural_coordinate_system()
        xs = [q.project_to_xxp_plane() / MM for q in PhaseSpaceParticle.create_from_running_particles(ip_end, ce, rp_x)]
        ys = [q.project_to_yyp_plane() / MM for q in PhaseSpaceParticle.create_from_running_particles(ip_end, ce, rp_y)]
        st = BaseUtils.Statistic()
        print(f"delta={delta},avg_size_x={st.clear().add_all(P2.extract(xs)[0]).half_width()}mm,"
              f"avg_size_y={st.clear().add_all(P2.extract(ys)[0]).half_width()}mm")
        return xs, ys

    # ------------------------------------------------------ orbit access
    def is_out_of_aperture(self, point: P3) -> bool:
        for m in self.magnets:
            if isinstance(m, ApertureObject) and m.is_out_of_aperture(point):
                return True
        return False

    def trace_is_out_of_aperture(self, trace_with_distance: List[ValueWithDistance[P3]]) -> bool:
        return any(self.is_out_of_aperture(pd.value) for pd in trace_with_distance)

    def get_length(self) -> float:
        return self.trajectory.get_length()

    def point_at(self, s: float) -> P2:
        return self.trajectory.point_at(s)

    def direct_at(self, s: float) -> P2:
        return self.trajectory.direct_at(s)

    class _Builder:
        def __init__(self, start_point: P2) -> None:
            self.start_point = start_point

        def first_drift(self, direct: P2 = None, length: float = 1.0) -> "Beamline":
            direct = P2.x_direct() if direct is None else direct
            bl = Beamline(Trajectory.set_start_point(self.start_point).first_line(direct=direct, length=length))
            bl.elements.append((0, None, length))
            return bl

    @staticmethod
    def set_start_point(start_point: P2 = None) -> "Beamline._Builder":
        return Beamline._Builder(P2.origin() if start_point is None else start_point)

    # --------------------------------------------------------- builders
    def append_drift(self, length: float) -> "Beamline":
        old = self.trajectory.get_length()
        self.trajectory.add_strait_line(length=length)
        self.elements.append((old, None, length))
        return self

    def append_qs(self, length: float, gradient: float, second_gradient: float, aperture_radius: float) -> "Beamline":
        old = self.trajectory.get_length()
        self.trajectory.add_strait_line(length=length)
        qs = QS.create_qs_along(self.trajectory, old, length, gradient, second_gradient, aperture_radius)
        self.magnets.append(qs)
        self.elements.append((old, qs, length))
        return self

    def append_dipole_cct(self, big_r: float, small_r_inner: float, small_r_outer: float, bending_angle: float,
                          tilt_angles: List[float], winding_number: int, current: float,
                          disperse_number_per_winding: int = 120) -> "Beamline":
        """Two-layer dipole CCT: inner layer, then outer layer with negated tilts and ksi end
        (beamline.py:507-574)."""
        old = self.trajectory.get_length()
        cct_length = big_r * abs(BaseUtils.angle_to_radian(bending_angle))
        self.trajectory.add_arc_line(radius=big_r, clockwise=bending_angle < 0, angle_deg=abs(bending_angle))
        for small_r, sign in ((small_r_inner, 1), (small_r_outer, -1)):
            cct = CCT.create_cct_along(
                trajectory=self.trajectory, s=old, big_r=big_r, small_r=small_r, bending_angle=abs(bending_angle),
                tilt_angles=tilt_angles if sign == 1 else BaseUtils.list_multiply(tilt_angles, -1),
                winding_number=winding_number, current=current,
                starting_point_in_ksi_phi_coordinate=P2.origin(),
                end_point_in_ksi_phi_coordinate=P2(sign * 2 * math.pi * winding_number,
                                                   BaseUtils.angle_to_radian(bending_angle)),
                disperse_number_per_winding=disperse_number_per_winding)
            self.magnets.append(cct)
            self.elements.append((old, cct, cct_length))
        return self

    def append_agcct(self, big_r: float, small_rs: List[float], bending_angles: List[float],
                     tilt_angles: List[List[float]], winding_numbers: List[List[int]], currents: List[float],
                     disperse_number_per_winding: int = 120) -> "Beamline":
        """Dipole CCT pair + alternating-gradient quadrupole CCT parts (beamline.py:576-793).

        Append order: dipole outer, dipole inner, then for every AG part its
        inner and outer layer; part k starts where part k-1 ended shifted by
        one winding pitch in phi, with the winding sense alternating.
        """
        if len(small_rs) != 4:
            raise ValueError(f"small_rs({small_rs}) must hold 4 radii: dipole outer, inner, quadrupole outer, inner")
        if not BaseUtils.is_sorted(small_rs[::-1]):
            raise ValueError(f"small_rs({small_rs}) must be sorted from large to small")
        total = sum(bending_angles)
        old = self.trajectory.get_length()
        length = big_r * abs(BaseUtils.angle_to_radian(total))
        self.trajectory.add_arc_line(radius=big_r, clockwise=total < 0, angle_deg=abs(total))
        n = disperse_number_per_winding
        wd = winding_numbers[0][0]
        total_rad = BaseUtils.angle_to_radian(total)
        for small_r, tilts, sign in ((small_rs[0], BaseUtils.list_multiply(tilt_angles[0], -1), -1),
                                     (small_rs[1], tilt_angles[0], 1)):
            cct = CCT.create_cct_along(self.trajectory, old, big_r, small_r, abs(total), tilts, wd, currents[0],
                                       P2.origin(), P2(sign * 2 * math.pi * wd, total_rad), n)
            self.magnets.append(cct)
            self.elements.append((old, cct, length))
        ws = winding_numbers[1]
        rads = BaseUtils.angle_to_radian(list(bending_angles))
        ag_tilts = tilt_angles[1]
        s_part = old
        end_in = end_out = P2.origin()
        for k in range(len(bending_angles)):
            if k == 0:
                start_in, start_out = P2.origin(), P2.origin()
                end_in = P2(((-1.0) ** k) * 2 * math.pi * ws[k], rads[k])
                end_out = P2(((-1.0) ** (k + 1)) * 2 * math.pi * ws[k], rads[k])
            else:
                shift = P2(0, rads[k - 1] / ws[k - 1])
                start_in, start_out = end_in + shift, end_out + shift
                end_in = start_in + P2(((-1) ** k) * 2 * math.pi * ws[k], rads[k])
                end_out = start_out + P2(((-1) ** (k + 1)) * 2 * math.pi * ws[k], rads[k])
            part_length = big_r * BaseUtils.angle_to_radian(abs(bending_angles[k]))
            for small_r, tilts, st, en in ((small_rs[3], BaseUtils.list_multiply(ag_tilts, -1), start_in, end_in),
                                           (small_rs[2], ag_tilts, start_out, end_out)):
                cct = CCT.create_cct_along(self.trajectory, old, big_r, small_r, abs(bending_angles[k]), tilts,
                                           ws[k], currents[1], st, en, n)
                self.magnets.append(cct)
                self.elements.append((s_part, cct, part_length))
            s_part += part_length
        return self

    def get_magnets(self) -> List[Magnet]:
        return self.magnets

    def get_trajectory(self) -> Trajectory:
        return self.trajectory

    def __str__(self) -> str:
        return f"beamline(magnet_size={len(self.magnets)}, traj_len={self.trajectory.get_length()})"

    __repr__ = __str__
