"""Magnet interface and the analytic magnets of the tracking path
(final_code/packages/magnets.py:49-67, 417-437, 539-759).

Field values are computed on the GPU: a magnet (or a combination of magnets)
is flattened into a device design (CCT windings, QS records, uniform field)
and evaluated by libcctamd.  Magnet kinds the device does not know are still
accepted by the field API, which adds their own ``magnetic_field_at`` on top;
tracking through them is rejected (as the reference GPU path does,
final_code/packages/gpu_accelerator.py:1970).
"""
from __future__ import annotations

import math
from typing import Callable, Iterable, List, Sequence, Tuple

import numpy

from .base_utils import BaseUtils
from .constants import MM
from .line2s import Line2
from .local_coordinate_system import LocalCoordinateSystem
from .point import P2, P3, ValueWithDistance


class ApertureObject:
    def is_out_of_aperture(self, point: P3) -> bool:
        raise NotImplementedError


class Magnet:
    """Anything with a field: ``magnetic_field_at(point: P3) -> P3``."""

    def magnetic_field_at(self, point: P3) -> P3:
        return self.magnetic_field_at_points([point])[0]

    def magnetic_field_at_points(self, points: Sequence[P3], precision: int = 64) -> List[P3]:
        """Batched field evaluation (one GPU launch for all points)."""
        from . import engine
        arr = numpy.array([p.to_list() for p in points], dtype=numpy.float64).reshape(-1, 3)
        b = field_of(self, arr, precision=precision)
        return [P3(*row) for row in b]

    def magnetic_field_along(self, line2: Line2, p2_t0_p3: Callable[[P2], P3] = lambda p2: P3(p2.x, p2.y, 0.0),
                             step: float = 1 * MM) -> List[ValueWithDistance[P3]]:
        """Field sampled at int(length/step)+1 points of line2 (magnets.py:69-91), one GPU launch."""
        length = line2.get_length()
        distances = BaseUtils.linspace(0, length, int(length / step) + 1)
        pts = [line2.point_at(d).to_p3(transformation=p2_t0_p3) for d in distances]
        return [ValueWithDistance(b, d) for b, d in zip(self.magnetic_field_at_points(pts), distances)]

    def magnetic_field_bz_along(self, line2: Line2, p2_t0_p3: Callable[[P2], P3] = lambda p2: P3(p2.x, p2.y, 0.0),
                                step: float = 1 * MM) -> List[P2]:
        return [P2(v.distance, v.value.z) for v in self.magnetic_field_along(line2, p2_t0_p3=p2_t0_p3, step=step)]

    # -- field analysis along a line (magnets.py:112-390) --------------------
    def _transverse_bz(self, side_points: Callable[[float], Tuple[P3, P3]], length: float,
                       good_field_area_width: float, step: float, point_number: int):
        """Bz on point_number points of the transverse segment (right -> left) at every
        s = linspace(0, length, int(length/step)+1): one GPU launch for all points."""
        xs = BaseUtils.linspace(-good_field_area_width / 2, good_field_area_width / 2, point_number)
        distances = BaseUtils.linspace(0, length, int(length / step) + 1)
        pts = []
        for s in distances:
            right, left = side_points(s)
            pts += BaseUtils.linspace(right, left, point_number)
        bz = numpy.array([b.z for b in self.magnetic_field_at_points(pts)]).reshape(len(distances), point_number)
        return xs, distances, bz

    def _line2_sides(self, line2: Line2, good_field_area_width: float):
        d = good_field_area_width / 2
        return lambda s: (line2.right_hand_side_point(s=s, d=d).to_p3(), line2.left_hand_side_point(s=s, d=d).to_p3())

    @staticmethod
    def _fit(xs, bz: numpy.ndarray, order: int) -> numpy.ndarray:
        """numpy.polyfit of every row (BaseUtils.polynomial_fitting), constant term first."""
        if bz.shape[0] == 0:
            return numpy.zeros((0, order + 1))
        return numpy.polyfit(numpy.asarray(xs), bz.T, order)[::-1].T

    def graident_field_along(self, line2: Line2, good_field_area_width: float = 10 * MM, step: float = 1 * MM,
                             point_number: int = 4) -> List[P2]:
        """Quadrupole gradient along line2 from a linear fit of Bz across it (magnets.py:112-164)."""
        xs, ds, bz = self._transverse_bz(self._line2_sides(line2, good_field_area_width), line2.get_length(),
                                         good_field_area_width, step, point_number)
        c = self._fit(xs, bz, 1)
        return [P2(s, float(c[i, 1])) for i, s in enumerate(ds)]

    def second_graident_field_along(self, line2: Line2, good_field_area_width: float = 10 * MM,
                                    step: float = 1 * MM, point_number: int = 4) -> List[P2]:
        """Sextupole gradient: 2 x the quadratic coefficient of the fit (magnets.py:166-219)."""
        xs, ds, bz = self._transverse_bz(self._line2_sides(line2, good_field_area_width), line2.get_length(),
                                         good_field_area_width, step, point_number)
        c = self._fit(xs, bz, 2)
        return [P2(s, float(c[i, 2]) * 2.0) for i, s in enumerate(ds)]

    @staticmethod
    def _multipoles(c: numpy.ndarray, ds) -> List[ValueWithDistance[List[float]]]:
        out = []
        for i, s in enumerate(ds):
            g = c[i].tolist()
            for k in range(2, len(g)):
                g[k] = g[k] * math.factorial(k)
            out.append(ValueWithDistance(value=g, distance=s))
        return out

    def multipole_field_along(self, line2: Line2, order: int, good_field_area_width: float = 10 * MM,
                              step: float = 1 * MM, point_number: int = 10) -> List[ValueWithDistance[List[float]]]:
        """[dipole, quadrupole, sextupole, ...] up to ``order`` along line2; the i-th fit
        coefficient times i! (magnets.py:221-287)."""
        xs, ds, bz = self._transverse_bz(self._line2_sides(line2, good_field_area_width), line2.get_length(),
                                         good_field_area_width, step, point_number)
        return self._multipoles(self._fit(xs, bz, order), ds)

    def multipole_field_along_line3(self, line3, order: int, good_field_area_width: float = 10 * MM,
                                    step: float = 1 * MM, point_number: int = 10,
                                    plane_direct: P3 = None) -> List[ValueWithDistance[List[float]]]:
        """As ``multipole_field_along`` on a 3D curve, transverse in the plane normal to
        ``plane_direct`` (default global z) (magnets.py:289-350)."""
        pd = P3.z_direct() if plane_direct is None else plane_direct
        d = good_field_area_width / 2
        sides = lambda s: (line3.right_hand_side_point(s=s, d=d, plane_direct=pd),  # noqa: E731
                           line3.left_hand_side_point(s=s, d=d, plane_direct=pd))
        xs, ds, bz = self._transverse_bz(sides, line3.get_length(), good_field_area_width, step, point_number)
        return self._multipoles(self._fit(xs, bz, order), ds)

    def integration_field(self, line2: Line2, p2_t0_p3: Callable[[P2], P3] = lambda p2: P3(p2.x, p2.y, 0.0),
                          step: float = 1 * MM) -> float:
        """Trapezoid integral of Bz along line2 with the sampling step (magnets.py:352-374)."""
        fields = self.magnetic_field_bz_along(line2, p2_t0_p3, step)
        ret = 0.0
        for i in range(len(fields) - 1):
            ret += (fields[i].y + fields[i + 1].y) / 2 * step
        return ret

    @staticmethod
    def no_magnet() -> "Magnet":
        return UniformMagnet(P3.zeros())

    @staticmethod
    def uniform_magnet(magnetic_field: P3 = None) -> "Magnet":
        return UniformMagnet(P3.zeros() if magnetic_field is None else magnetic_field)

    @staticmethod
    def combine(*magnets) -> "CombinedMagnet":
        for m in magnets:
            if not isinstance(m, Magnet):
                raise ValueError(f"{m} is not a magnet")
        return CombinedMagnet(*magnets)

    # -- device description ------------------------------------------------
    def _device_parts(self, out: "DeviceParts") -> None:
        out.foreign.append(self)


class DeviceParts:
    """Flattened magnet list of one design, in append order."""

    def __init__(self):
        self.cct_params: List[numpy.ndarray] = []
        self.cct_counts: List[int] = []
        self.qs: List[numpy.ndarray] = []
        self.bconst = numpy.zeros(3)
        self.foreign: List[Magnet] = []
        # the magnet objects behind cct_params and qs, same order (parameter gradients map onto them)
        self.cct_objects: List[Magnet] = []
        self.qs_objects: List[Magnet] = []

    def spec(self):
        from .engine import DesignSpec
        from ._native import CCT_PARAMS
        return DesignSpec(
            cct_params=numpy.array(self.cct_params, numpy.float64).reshape(len(self.cct_params), CCT_PARAMS),
            cct_counts=numpy.array(self.cct_counts, numpy.int64),
            qs=numpy.array(self.qs, numpy.float64).reshape(len(self.qs), 16),
            bconst=self.bconst.copy())


def device_parts(magnet: Magnet) -> DeviceParts:
    parts = DeviceParts()
    magnet._device_parts(parts)
    return parts


def field_of(magnet: Magnet, points: numpy.ndarray, precision: int = 64) -> numpy.ndarray:
    """B (n, 3) of one magnet at global points (n, 3)."""
    from . import engine
    parts = device_parts(magnet)
    pts = numpy.asarray(points, numpy.float64).reshape(-1, 3)
    model = engine.model_for([parts.spec()], fp32=(precision == 32))
    b = model.field(pts, [pts.shape[0]], precision=precision)
    for m in parts.foreign:
        for i in range(pts.shape[0]):
            f = m.magnetic_field_at(P3(*pts[i]))
            b[i] += (f.x, f.y, f.z)
    return b


class UniformMagnet(Magnet):
    """Same field everywhere (magnets.py:417-437)."""

    def __init__(self, magnetic_field: P3 = None) -> None:
        self.magnetic_field = P3.zeros() if magnetic_field is None else magnetic_field

    def magnetic_field_at(self, point: P3) -> P3:
        return self.magnetic_field.copy()

    def _device_parts(self, out: DeviceParts) -> None:
        out.bconst = out.bconst + numpy.array(self.magnetic_field.to_list())

    def __str__(self) -> str:
        return f"UniformMagnet, field {self.magnetic_field}"


class CombinedMagnet(Magnet):
    """Sum of several magnets (magnets.py:539-586)."""

    def __init__(self, *magnets) -> None:
        for m in magnets:
            if not isinstance(m, Magnet):
                raise ValueError(f"{m} is not a magnet")
        self._magnets: List[Magnet] = list(magnets)

    def add(self, magnet: Magnet) -> "CombinedMagnet":
        self._magnets.append(magnet)
        return self

    def add_all(self, magnets: Iterable[Magnet]) -> "CombinedMagnet":
        self._magnets.extend(magnets)
        return self

    def get_magnets(self) -> List[Magnet]:
        return self._magnets

    def remove(self, m: Magnet) -> "CombinedMagnet":
        self._magnets.remove(m)
        return self

    def magnetic_field_at(self, point: P3) -> P3:
        return Magnet.magnetic_field_at(self, point)

    def _device_parts(self, out: DeviceParts) -> None:
        for m in self._magnets:
            m._device_parts(out)


class QS(Magnet, ApertureObject):
    """Hard-edge quadrupole + sextupole (magnets.py:589-759).

    Local frame: origin at the entrance centre, z along the orbit; inside
    0 <= z <= length and the aperture, Bx = g y + s x y, By = g x + s (x^2 - y^2)/2.
    """

    def __init__(self, local_coordinate_system: LocalCoordinateSystem, length: float, gradient: float,
                 second_gradient: float, aperture_radius: float):
        self.local_coordinate_system = local_coordinate_system
        self.length = float(length)
        self.gradient = float(gradient)
        self.second_gradient = float(second_gradient)
        self.aperture_radius = float(aperture_radius)

    def __str__(self) -> str:
        return (f"QS:local_coordinate_system={self.local_coordinate_system}, length={self.length}, "
                f"gradient={self.gradient}, second_gradient={self.second_gradient}, "
                f"aperture_radius={self.aperture_radius}")

    __repr__ = __str__

    def magnetic_field_at(self, point: P3) -> P3:
        return Magnet.magnetic_field_at(self, point)

    def is_out_of_aperture(self, point: P3) -> bool:
        lp = self.local_coordinate_system.point_to_local_coordinate(point)
        if 0 <= lp.z <= self.length:
            return (lp.x ** 2 + lp.y ** 2) > self.aperture_radius ** 2
        return False

    @staticmethod
    def create_qs_along(trajectory: Line2, s: float, length: float, gradient: float, second_gradient: float,
                        aperture_radius: float) -> "QS":
        """Entrance at trajectory(s), z along the orbit, x = z rotated +90 deg (magnets.py:687-724)."""
        origin = trajectory.point_at(s)
        z = trajectory.direct_at(s)
        x = z.rotate(BaseUtils.angle_to_radian(90))
        lcs = LocalCoordinateSystem(location=origin.to_p3(), x_direction=x.to_p3(), z_direction=z.to_p3())
        return QS(lcs, length, gradient, second_gradient, aperture_radius)

    def to_numpy_array(self, numpy_dtype=numpy.float64) -> numpy.ndarray:
        """[O, XI, YI, ZI, length, gradient, second_gradient, aperture_radius] (magnets.py:726-748)."""
        return numpy.array(self.local_coordinate_system.to_array()
                           + [self.length, self.gradient, self.second_gradient, self.aperture_radius],
                           dtype=numpy_dtype)

    def _device_parts(self, out: DeviceParts) -> None:
        out.qs.append(self.to_numpy_array())
        out.qs_objects.append(self)

    @staticmethod
    def as_qs(anything) -> "QS":
        return anything
