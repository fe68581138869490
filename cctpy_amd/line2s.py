"""Directed 2-D curves of the design orbit (final_code/packages/line2s.py:20-560)."""
from __future__ import annotations

import math
from typing import List, Tuple

from .base_utils import BaseUtils
from .constants import MM
from .point import P2, P3, ValueWithDistance


class Line2:
    """A directed curve in the xy plane: get_length, point_at(s), direct_at(s)."""

    def get_length(self) -> float:
        raise NotImplementedError

    def point_at(self, s: float) -> P2:
        raise NotImplementedError

    def direct_at(self, s: float) -> P2:
        delta = 1e-7
        return (self.point_at(s + delta) - self.point_at(s)).normalize()

    def right_hand_side_point(self, s: float, d: float) -> P2:
        return self.point_at(s) + self.direct_at(s).copy().rotate(-math.pi / 2).change_length(d)

    def left_hand_side_point(self, s: float, d: float) -> P2:
        return self.right_hand_side_point(s, -d)

    def point_at_start(self) -> P2:
        return self.point_at(0.0)

    def point_at_end(self) -> P2:
        return self.point_at(self.get_length())

    def point_at_middle(self) -> P2:
        return self.point_at(self.get_length() / 2)

    def direct_at_start(self) -> P2:
        return self.direct_at(0.0)

    def direct_at_end(self) -> P2:
        return self.direct_at(self.get_length())

    def direct_at_middle(self) -> P2:
        return self.direct_at(self.get_length() / 2)

    def disperse2d(self, step: float = 1 * MM) -> List[P2]:
        n = int(math.ceil(self.get_length() / step)) + 1
        return [self.point_at(s) for s in BaseUtils.linspace(0, self.get_length(), n)]

    def disperse2d_with_distance(self, step: float = 1 * MM) -> List[ValueWithDistance[P2]]:
        n = int(math.ceil(self.get_length() / step)) + 1
        return [ValueWithDistance(self.point_at(s), s) for s in BaseUtils.linspace(0, self.get_length(), n)]

    def disperse3d(self, p2_t0_p3=lambda p2: P3(p2.x, p2.y, 0.0), step: float = 1 * MM) -> List[P3]:
        return [p.to_p3(p2_t0_p3) for p in self.disperse2d(step=step)]

    def __str__(self) -> str:
        return f"Line2[start {self.point_at_start()}, length {self.get_length()}]"


class StraightLine2(Line2):
    """Straight segment: length, unit direction, start point."""

    def __init__(self, length: float, direct: P2, start_point: P2):
        self.length = float(length)
        self.direct = direct.normalize()
        self.start_point = start_point

    def get_length(self) -> float:
        return self.length

    def point_at(self, s: float) -> P2:
        return self.start_point + self.direct.copy().change_length(s)

    def direct_at(self, s: float) -> P2:
        return self.direct

    def __str__(self) -> str:
        return f"StraightLine2[start {self.start_point}, direct {self.direct}, length {self.length}]"

    __repr__ = __str__

    def position_of(self, p: P2) -> int:
        """1 if p is right of the directed line, -1 if left, 0 if on it (line2s.py:303-342)."""
        rel = p - self.start_point
        proj = (self.direct * rel) * self.direct
        vert = rel - proj
        if vert == P2.zeros():
            return 0
        vert = vert.normalize()
        right = self.direct.rotate(BaseUtils.angle_to_radian(-90)).normalize()
        return 1 if vert == right else -1

    @staticmethod
    def create_by(start_point: P2, end_point: P2) -> "StraightLine2":
        return StraightLine2(length=(end_point - start_point).length(), direct=(end_point - start_point),
                             start_point=start_point)


class ArcLine2(Line2):
    """Circular arc described in polar form around its centre (line2s.py:431-560)."""

    def __init__(self, starting_phi: float, center: P2, radius: float, total_phi: float, clockwise: bool):
        self.starting_phi = starting_phi
        self.center = center
        self.radius = radius
        self.total_phi = total_phi
        self.clockwise = clockwise
        self.length = radius * total_phi

    def get_length(self) -> float:
        return self.length

    def _phi(self, s: float) -> float:
        phi = s / self.radius
        return self.starting_phi - phi if self.clockwise else self.starting_phi + phi

    def point_at(self, s: float) -> P2:
        return ArcLine2.unit_circle(self._phi(s)).change_length(self.radius) + self.center

    def direct_at(self, s: float) -> P2:
        return ArcLine2.unit_circle(self._phi(s)).rotate(-math.pi / 2 if self.clockwise else math.pi / 2)

    @staticmethod
    def create(start_point: P2, start_direct: P2, radius: float, clockwise: bool, total_deg: float) -> "ArcLine2":
        center = start_point + start_direct.copy().rotate(-math.pi / 2 if clockwise else math.pi / 2).change_length(
            radius)
        return ArcLine2((start_point - center).angle_to_x_axis(), center, radius,
                        BaseUtils.angle_to_radian(total_deg), clockwise)

    @staticmethod
    def unit_circle(phi: float) -> P2:
        return P2(math.cos(phi), math.sin(phi))

    def __str__(self) -> str:
        return (f"ArcLine2[start {self.point_at_start()}, direct {self.direct_at_start()}, "
                f"{'clockwise' if self.clockwise else 'anticlockwise'}, radius {self.radius}, angle {self.total_phi}]")

    __repr__ = __str__
