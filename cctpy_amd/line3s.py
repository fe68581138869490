"""Three-dimensional curves (final_code/packages/line3s.py:20-131, 266-360): the
interface used by ``Magnet.multipole_field_along_line3`` to slice a tracked
orbit, with a straight segment and a polyline of tracked points."""
from __future__ import annotations

import math
from typing import List

from .base_utils import BaseUtils
from .constants import MM
from .point import P3, ValueWithDistance


class Line3:
    def get_length(self) -> float:
        raise NotImplementedError

    def point_at(self, s: float) -> P3:
        raise NotImplementedError

    def direct_at(self, s: float) -> P3:
        """Forward difference with a 1e-7 step (line3s.py:57-73)."""
        delta = 1e-7
        return (self.point_at(s + delta) - self.point_at(s)).normalize()

    def right_hand_side_point(self, s: float, d: float, plane_direct: P3) -> P3:
        """Point at distance d to the right of s, in the plane normal to ``plane_direct``."""
        return self.point_at(s) + (self.direct_at(s) @ plane_direct).change_length(d)

    def left_hand_side_point(self, s: float, d: float, plane_direct: P3) -> P3:
        return self.right_hand_side_point(s, -d, plane_direct)

    def disperse3d(self, step: float = 1.0 * MM) -> List[P3]:
        number = int(math.ceil(self.get_length() / step)) + 1
        return [self.point_at(s) for s in BaseUtils.linspace(0, self.get_length(), number)]

    def disperse3d_with_distance(self, step: float = 1.0 * MM) -> List[ValueWithDistance[P3]]:
        number = int(math.ceil(self.get_length() / step)) + 1
        return [ValueWithDistance(self.point_at(s), s) for s in BaseUtils.linspace(0, self.get_length(), number)]


class TwoPointLine3(Line3):
    """Straight segment p0 -> p1 (immutable)."""

    def __init__(self, p0: P3, p1: P3) -> None:
        self._p0, self._p1 = p0, p1
        self._length = (p1 - p0).length()
        self._direct = (p1 - p0).normalize()

    def to_vector(self) -> P3:
        return self._p1 - self._p0

    def get_length(self) -> float:
        return self._length

    def point_at(self, s: float) -> P3:
        return self._p0 + self._direct * s

    def direct_at(self, s: float) -> P3:
        return self._direct


class DiscretePointLine3(Line3):
    """Polyline through tracked points, walked piece by piece like the reference."""

    def __init__(self, track: List[P3]) -> None:
        self._parts = [TwoPointLine3(track[i], track[i + 1]) for i in range(len(track) - 1)]
        length = 0.0
        for p in self._parts:
            length += p.get_length()
        self._length = length

    def get_length(self) -> float:
        return self._length

    def _part(self, s: float):
        part = None
        for part in self._parts:
            if s < part.get_length():
                break
            s -= part.get_length()
        return part, s

    def point_at(self, s: float) -> P3:
        part, r = self._part(s)
        return part.point_at(r)

    def direct_at(self, s: float) -> P3:
        part, r = self._part(s)
        return part.direct_at(r)
