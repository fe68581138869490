"""Design orbit made of straight lines and arcs (final_code/packages/trajectory.py:27-215)."""
from __future__ import annotations

from typing import List

from .line2s import ArcLine2, Line2, StraightLine2
from .point import P2


class Trajectory(Line2):
    def __init__(self, first_line: Line2):
        self._lines: List[Line2] = [first_line]
        self._length = first_line.get_length()
        self._reported = False

    def add_line2(self, line2: Line2) -> "Trajectory":
        self._lines.append(line2)
        self._length += line2.get_length()
        return self

    def add_strait_line(self, length: float) -> "Trajectory":
        last = self._lines[-1]
        self._lines.append(StraightLine2(length, last.direct_at_end(), last.point_at_end()))
        self._length += length
        return self

    def add_arc_line(self, radius: float, clockwise: bool, angle_deg: float) -> "Trajectory":
        last = self._lines[-1]
        arc = ArcLine2.create(last.point_at_end(), last.direct_at_end(), radius, clockwise, angle_deg)
        self._lines.append(arc)
        self._length += arc.get_length()
        return self

    def get_length(self) -> float:
        return self._length

    def _walk(self, s: float):
        for line in self._lines:
            if s <= line.get_length():
                return line, s, False
            s -= line.get_length()
        return self._lines[-1], s, True

    def point_at(self, s: float) -> P2:
        line, r, beyond = self._walk(s)
        if beyond and abs(r) <= 1e-8:
            return line.point_at_end()
        if beyond and not self._reported:
            self._reported = True
            print(f"ERROR Trajectory::point_at{s}")
        return line.point_at(r)

    def direct_at(self, s: float) -> P2:
        line, r, beyond = self._walk(s)
        if beyond and abs(r) <= 1e-8:
            return line.direct_at_end()
        if beyond and not self._reported:
            self._reported = True
            print(f"ERROR Trajectory::direct_at{s}")
        return line.direct_at(r)

    def __str__(self) -> str:
        return "Trajectory:\t\n" + "\t\n".join(f"# {i + 1:02d} {l}" for i, l in enumerate(self._lines))

    __repr__ = __str__

    class _Builder:
        def __init__(self, start_point: P2):
            self.start_point = start_point

        def first_line(self, direct: P2 = None, length: float = 1.0) -> "Trajectory":
            direct = P2.x_direct() if direct is None else direct
            return Trajectory(StraightLine2(length, direct, self.start_point))

    @staticmethod
    def set_start_point(start_point: P2 = None) -> "Trajectory._Builder":
        return Trajectory._Builder(P2.origin() if start_point is None else start_point)

    def get_line2_list(self) -> List[Line2]:
        return self._lines
