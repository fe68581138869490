"""Bent CCT layer (final_code/packages/cct.py:31-575).

The constructor keeps the reference's arguments, checks and scalar geometry
(a, eta, phi0); the winding itself — W*n straight segments on the
bipolar-toroidal surface — is generated on the GPU when the field or the
path arrays are first needed.
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy

from . import _native
from .base_utils import BaseUtils
from .line2s import Line2, StraightLine2
from .local_coordinate_system import LocalCoordinateSystem
from .magnets import ApertureObject, DeviceParts, Magnet
from .point import P2, P3


class CCT(Magnet, ApertureObject):
    def __init__(self, local_coordinate_system: LocalCoordinateSystem, big_r: float, small_r: float,
                 bending_angle: float, tilt_angles: List[float], winding_number: int, current: float,
                 starting_point_in_ksi_phi_coordinate: P2, end_point_in_ksi_phi_coordinate: P2,
                 disperse_number_per_winding: int = 120):
        if bending_angle < 0:
            raise ValueError(f"CCT bending angle must be positive, not {bending_angle}; reverse bends are set "
                             f"through starting_point_in_ksi_phi_coordinate and end_point_in_ksi_phi_coordinate")
        if big_r < 0:
            raise ValueError(f"big_r = {big_r} is negative")
        if small_r < 0:
            raise ValueError(f"small_r = {small_r} is negative")
        if small_r >= big_r:
            raise ValueError(f"small_r {small_r} >= big_r {big_r}")
        if len(tilt_angles) > _native.MAX_TILTS:
            raise ValueError(f"at most {_native.MAX_TILTS} tilt angles are supported")
        self.local_coordinate_system = local_coordinate_system
        self.big_r = float(big_r)
        self.small_r = float(small_r)
        self.bending_angle = float(bending_angle)
        self.tilt_angles = [float(e) for e in tilt_angles]
        self.winding_number = int(winding_number)
        self.current = float(current)
        self.starting_point_in_ksi_phi_coordinate = starting_point_in_ksi_phi_coordinate
        self.end_point_in_ksi_phi_coordinate = end_point_in_ksi_phi_coordinate
        self.disperse_number_per_winding = int(disperse_number_per_winding)
        self.bending_radian = BaseUtils.angle_to_radian(self.bending_angle)
        self.tilt_radians = BaseUtils.angle_to_radian(self.tilt_angles)
        self.phi0 = self.bending_radian / self.winding_number
        self.a = math.sqrt(self.big_r ** 2 - self.small_r ** 2)
        self.eta = 0.5 * math.log((self.big_r + self.a) / (self.big_r - self.a))
        self.bipolar_toroidal_coordinate_system = CCT.BipolarToroidalCoordinateSystem(
            self.a, self.eta, self.big_r, self.small_r)
        self.total_disperse_number = self.winding_number * self.disperse_number_per_winding
        self._path_cache = None

    # ---------------------------------------------------------- geometry
    def phi_ksi_function(self, ksi: float) -> float:
        """phi(ksi) = k ksi + b + sum_j cot(t_j)/((j+1) sinh eta) sin((j+1) ksi) (cct.py:166-188)."""
        k, b = self._phi_line()
        phi = k * ksi + b
        for j, c in self._tilt_terms():
            phi += c * math.sin((j + 1) * ksi)
        return phi

    def p2_function(self, ksi: float) -> P2:
        return P2(ksi, self.phi_ksi_function(ksi))

    def p3_function(self, ksi: float) -> P3:
        return self.bipolar_toroidal_coordinate_system.convert(self.p2_function(ksi))

    def _phi_line(self) -> Tuple[float, float]:
        x1, y1 = self.starting_point_in_ksi_phi_coordinate.x, self.starting_point_in_ksi_phi_coordinate.y
        x2, y2 = self.end_point_in_ksi_phi_coordinate.x, self.end_point_in_ksi_phi_coordinate.y
        k = (y2 - y1) / (x2 - x1)
        return k, -k * x1 + y1

    def _tilt_terms(self) -> List[Tuple[int, float]]:
        terms = []
        for j, t in enumerate(self.tilt_angles):
            if BaseUtils.equal(t, 90.0):
                continue
            terms.append((j, (1 / math.tan(self.tilt_radians[j])) / ((j + 1) * math.sinh(self.eta))))
        return terms

    def device_record(self) -> numpy.ndarray:
        """The 32-double record consumed by cctamd_winding_generate (layout in include/cctamd.h)."""
        r = numpy.zeros(_native.CCT_PARAMS)
        r[0:12] = self.local_coordinate_system.to_array()
        r[12] = self.a
        r[13] = math.sinh(self.eta)
        r[14] = math.cosh(self.eta)
        x0 = self.starting_point_in_ksi_phi_coordinate.x
        x1 = self.end_point_in_ksi_phi_coordinate.x
        r[15] = x0
        r[16] = (x1 - x0) * (1 / ((self.total_disperse_number + 1) - 1))
        r[17], r[18] = self._phi_line()
        r[19] = 1e-7 * self.current
        r[20] = len(self.tilt_angles)
        skip = 0
        terms = dict(self._tilt_terms())
        for j in range(len(self.tilt_angles)):
            if j in terms:
                r[21 + j] = terms[j]
            else:
                skip |= 1 << j
        r[29] = skip
        return r

    def _device_parts(self, out: DeviceParts) -> None:
        out.cct_params.append(self.device_record())
        out.cct_objects.append(self)
        # |I| <= 1e-6 gives no field (cct.py:254-255): no segments are generated
        out.cct_counts.append(0 if BaseUtils.equal(self.current, 0, err=1e-6) else self.total_disperse_number)

    def _local_path(self) -> numpy.ndarray:
        if self._path_cache is None or self._path_cache[0] != self._key():
            from .engine import winding_points
            self._path_cache = (self._key(), winding_points(self.device_record(), self.total_disperse_number + 1))
        return self._path_cache[1]

    def _key(self):
        return self.device_record().tobytes()

    @property
    def dispersed_path3(self) -> numpy.ndarray:
        return self._local_path()

    @property
    def dispersed_path3_points(self) -> List[P3]:
        return [P3(*p) for p in self._local_path()]

    @property
    def dispersed_path2(self) -> numpy.ndarray:
        ksi = numpy.array(BaseUtils.linspace(self.starting_point_in_ksi_phi_coordinate.x,
                                             self.end_point_in_ksi_phi_coordinate.x,
                                             self.total_disperse_number + 1))
        return numpy.stack([ksi, numpy.array([self.phi_ksi_function(float(k)) for k in ksi])], axis=1)

    @property
    def elementary_current(self) -> numpy.ndarray:
        p = self._local_path()
        return 1e-7 * self.current * (p[1:] - p[:-1])

    @property
    def elementary_current_position(self) -> numpy.ndarray:
        p = self._local_path()
        return 0.5 * (p[1:] + p[:-1])

    def global_path3(self) -> List[P3]:
        return [self.local_coordinate_system.point_to_global_coordinate(p) for p in self.dispersed_path3_points]

    def global_current_elements_and_elementary_current_positions(self, numpy_dtype=numpy.float64):
        """Global-frame segments (kl, midpoints) flattened (cct.py:434-457)."""
        g = numpy.array([p.to_list() for p in self.global_path3()], dtype=numpy_dtype)
        return (1e-7 * self.current * (g[1:] - g[:-1])).flatten(), (0.5 * (g[1:] + g[:-1])).flatten()

    def conductor_length(self, line_number: int = 2 * 7, disperse_number_per_winding: int = 360) -> float:
        """Polyline length of the conductor path times line_number (cct.py:475-491)."""
        from .engine import winding_points
        rec = self.device_record()
        n = disperse_number_per_winding * self.winding_number
        rec[16] = (self.end_point_in_ksi_phi_coordinate.x - self.starting_point_in_ksi_phi_coordinate.x) * (
            1 / ((n + 1) - 1))
        p = winding_points(rec, n + 1)
        return float(numpy.sum(numpy.linalg.norm(numpy.diff(p, axis=0), axis=1))) * line_number

    # ------------------------------------------------------------- field
    def magnetic_field_at(self, point: P3) -> P3:
        if BaseUtils.equal(self.current, 0, err=1e-6):
            return P3.zeros()
        return Magnet.magnetic_field_at(self, point)

    def is_out_of_aperture(self, point: P3) -> bool:
        """cct.py:287-330."""
        lp = self.local_coordinate_system.point_to_local_coordinate(point)
        lp2 = lp.to_p2()
        clockwise = self.end_point_in_ksi_phi_coordinate.y < 0
        phi = lp2.angle_to_x_axis()
        inside = phi > (2 * math.pi - self.bending_radian) if clockwise else phi < self.bending_radian
        if not inside:
            return False
        return (abs(lp.z) > self.small_r or lp2.length() > (self.big_r + self.small_r)
                or lp2.length() < (self.big_r - self.small_r))

    def __str__(self) -> str:
        return (f"CCT: local_coordinate_system({self.local_coordinate_system})big_r({self.big_r})"
                f"small_r({self.small_r})bending_angle({self.bending_angle})tilt_angles({self.tilt_angles})"
                f"winding_number({self.winding_number})current({self.current})"
                f"starting_point_in_ksi_phi_coordinate({self.starting_point_in_ksi_phi_coordinate})"
                f"end_point_in_ksi_phi_coordinate({self.end_point_in_ksi_phi_coordinate})"
                f"disperse_number_per_winding({self.disperse_number_per_winding})")

    __repr__ = __str__

    @staticmethod
    def as_cct(anything) -> "CCT":
        return anything

    class BipolarToroidalCoordinateSystem:
        """(ksi, phi) -> (x, y, z) on the torus of radii (R, r) (cct.py:190-247)."""

        def __init__(self, a: float, eta: float, big_r: float, small_r: float):
            self.a, self.eta, self.big_r, self.small_r = a, eta, big_r, small_r
            BaseUtils.equal(big_r, math.sqrt(a * a / (1 - 1 / math.pow(math.cosh(eta), 2))),
                            msg=f"BipolarToroidalCoordinateSystem: inconsistent a({a}) eta({eta}) R({big_r})")
            BaseUtils.equal(small_r, big_r / math.cosh(eta),
                            msg=f"BipolarToroidalCoordinateSystem: inconsistent R({big_r}) r({small_r}) eta({eta})")

        def convert(self, p: P2) -> P3:
            t = self.a / (math.cosh(self.eta) - math.cos(p.x))
            return P3(t * math.sinh(self.eta) * math.cos(p.y), t * math.sinh(self.eta) * math.sin(p.y),
                      t * math.sin(p.x))

        def main_normal_direction_at(self, p: P2) -> P3:
            center = P3(self.big_r * math.cos(p.y), self.big_r * math.sin(p.y), 0)
            return (self.convert(p) - center).normalize()

        def __str__(self) -> str:
            return f"BipolarToroidalCoordinateSystem a({self.a})eta({self.eta})R({self.big_r})r({self.small_r})"

        __repr__ = __str__

    @staticmethod
    def create_cct_along(trajectory: Line2, s: float, big_r: float, small_r: float, bending_angle: float,
                         tilt_angles: List[float], winding_number: int, current: float,
                         starting_point_in_ksi_phi_coordinate: P2, end_point_in_ksi_phi_coordinate: P2,
                         disperse_number_per_winding: int = 120) -> "CCT":
        """CCT on the orbit from s: centre from two three-point circle fits, local y along
        +/- the start direction, z = global z (cct.py:341-421)."""
        start = trajectory.point_at(s)
        arc = big_r * BaseUtils.angle_to_radian(bending_angle)
        end = trajectory.point_at(s + arc)
        m0 = trajectory.point_at(s + arc / 3 * 1)
        m1 = trajectory.point_at(s + arc / 3 * 2)
        c1, r1 = BaseUtils.circle_center_and_radius(start, m0, m1)
        c2, r2 = BaseUtils.circle_center_and_radius(m0, m1, end)
        BaseUtils.equal(c1, c2, msg=f"create_cct_along: inconsistent centres c1{c1}, c2{c2}")
        BaseUtils.equal(r1, r2, msg=f"create_cct_along: inconsistent radii r1{r1}, r2{r2}")
        center = (c1 + c2) * 0.5
        sd = trajectory.direct_at(s)
        side = StraightLine2(1.0, sd, start).position_of(center)
        if side == 0:
            raise ValueError(f"centre {center} lies on the design orbit {trajectory}")
        ydir = -sd.to_p3() if side == 1 else sd.to_p3()
        lcs = LocalCoordinateSystem.create_by_y_and_z_direction(location=center.to_p3(), y_direction=ydir,
                                                                z_direction=P3.z_direct())
        return CCT(lcs, big_r, small_r, bending_angle, tilt_angles, winding_number, current,
                   starting_point_in_ksi_phi_coordinate, end_point_in_ksi_phi_coordinate,
                   disperse_number_per_winding)
