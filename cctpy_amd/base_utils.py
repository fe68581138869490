"""Host utilities used by the field/tracking path (subset of final_code/packages/base_utils.py)."""
from __future__ import annotations

import math
from typing import Dict, Iterable, List, Optional, Tuple, Union

import numpy

from .point import P2, P3


class BaseUtils:
    @staticmethod
    def equal(a, b, err: float = 1e-6, msg: Optional[str] = None) -> bool:
        """Approximate equality or assertion (base_utils.py:35-78)."""
        if isinstance(a, (float, int)) and isinstance(b, (float, int)):
            ok = a == b or abs(a - b) <= err or ((a + b != 0.0) and (2 * abs((a - b) / (a + b)) <= err))
        elif (isinstance(a, P2) and isinstance(b, P2)) or (isinstance(a, P3) and isinstance(b, P3)):
            ok = a.__eq__(b, err=err, msg=msg)
        else:
            ok = a == b
        if ok:
            return True
        if msg is None:
            return False
        raise AssertionError(msg)

    @staticmethod
    def linspace(start, end, number: int) -> list:
        """start + d*i with d = (end-start)*(1/(number-1)) (base_utils.py:80-90)."""
        d = (end - start) * (1 / (number - 1))
        return [start + d * float(i) for i in range(number)]

    @staticmethod
    def angle_to_radian(deg):
        if isinstance(deg, (float, int)):
            return deg / 180.0 * math.pi
        if isinstance(deg, list):
            return [BaseUtils.angle_to_radian(d) for d in deg]
        raise NotImplementedError

    @staticmethod
    def radian_to_angle(rad):
        if isinstance(rad, (float, int)):
            return rad * 180.0 / math.pi
        if isinstance(rad, list):
            return [BaseUtils.radian_to_angle(d) for d in rad]
        if isinstance(rad, numpy.ndarray):
            return numpy.array([BaseUtils.radian_to_angle(float(d)) for d in rad])
        raise NotImplementedError

    @staticmethod
    def circle_center_and_radius(p1: P2, p2: P2, p3: P2) -> Tuple[P2, float]:
        """Least-squares centre of three points and mean radius (base_utils.py:124-151)."""
        xs = (p1.x, p2.x, p3.x)
        ys = (p1.y, p2.y, p3.y)
        rows = [(0, 1), (0, 2), (1, 2)]
        # rhs keeps the reference's term order x_j^2 + y_j^2 - x_i^2 - y_i^2
        rhs = [xs[j] ** 2 + ys[j] ** 2 - xs[i] ** 2 - ys[i] ** 2 for i, j in rows]
        a = numpy.array([[xs[j] - xs[i], ys[j] - ys[i]] for i, j in rows])
        b = 0.5 * numpy.array([[r] for r in rhs])
        c = numpy.linalg.inv(a.T @ a) @ a.T @ b
        cx, cy = float(c[0][0]), float(c[1][0])
        r = [math.sqrt((cx - xs[i]) ** 2 + (cy - ys[i]) ** 2) for i in range(3)]
        return P2(cx, cy), (r[0] + r[1] + r[2]) / 3

    @staticmethod
    def list_multiply(li: list, number) -> list:
        return [e * number for e in li]

    @staticmethod
    def is_sorted(li: list) -> bool:
        return all(li[i] <= li[i + 1] for i in range(len(li) - 1))

    @staticmethod
    def i_am_sure_my_code_closed_in_if_name_equal_main() -> None:
        """Kept for call compatibility; tracking here never forks processes."""

    class Ellipse:
        """Ellipse A x^2 + B xy + C y^2 = D (base_utils.py:446-601)."""

        CIRCUMFERENCE_CACHE: Dict["BaseUtils.Ellipse", float] = {}

        def __init__(self, A: float, B: float, C: float, D: float):
            self.A, self.B, self.C, self.D = float(A), float(B), float(C), float(D)

        def __eq__(self, o) -> bool:
            return (self.A, self.B, self.C, self.D) == (o.A, o.B, o.C, o.D)

        def __hash__(self) -> int:
            return hash((self.A, self.B, self.C, self.D))

        def point_at(self, theta: float) -> P2:
            """Intersection with the ray from the origin at angle theta."""
            d = P2()
            while theta < 0:
                theta += 2 * math.pi
            while theta > 2 * math.pi:
                theta -= 2 * math.pi
            if BaseUtils.equal(theta, 0) or BaseUtils.equal(theta, 2 * math.pi):
                d.x, d.y = math.sqrt(self.D / self.A), 0
            if BaseUtils.equal(theta, math.pi):
                d.x, d.y = -math.sqrt(self.D / self.A), 0
            if 0 < theta < math.pi:
                t = 1 / math.tan(theta)
                d.y = math.sqrt(self.D / (self.A * t * t + self.B * t + self.C))
                d.x = t * d.y
            if math.pi < theta < 2 * math.pi:
                t = 1 / math.tan(theta - math.pi)
                d.y = -math.sqrt(self.D / (self.A * t * t + self.B * t + self.C))
                d.x = t * d.y
            return d

        @property
        def circumference(self) -> float:
            c = BaseUtils.Ellipse.CIRCUMFERENCE_CACHE.get(self)
            if c is None:
                num = 3600 * 4
                c = 0.0
                for i in range(num):
                    c += (self.point_at(2.0 * math.pi / float(num) * (i + 1))
                          - self.point_at(2.0 * math.pi / float(num) * i)).length()
                BaseUtils.Ellipse.CIRCUMFERENCE_CACHE[self] = c
            return c

        def point_after(self, length: float) -> P2:
            step = BaseUtils.angle_to_radian(0.05)
            theta = 0.0
            while length > 0.0:
                length -= (self.point_at(theta + step) - self.point_at(theta)).length()
                theta += step
            return self.point_at(theta)

        def uniform_distribution_points_along_edge(self, num: int) -> List[P2]:
            c = self.circumference
            return [self.point_after(c / num * i) for i in range(num)]

        @classmethod
        def create_standard_ellipse(cls, a: float, b: float) -> "BaseUtils.Ellipse":
            return cls(A=b ** 2, B=0.0, C=a ** 2, D=(a ** 2) * (b ** 2))

    class Statistic:
        """Running statistics of a list of numbers (base_utils.py:603-707)."""

        def __init__(self):
            self._data: List[float] = []

        def data(self) -> List[float]:
            return self._data

        def add(self, v: float) -> "BaseUtils.Statistic":
            self._data.append(v)
            return self

        def add_all(self, vs: Iterable[float]) -> "BaseUtils.Statistic":
            self._data.extend(vs)
            return self

        def max(self) -> float:
            return numpy.max(self._data)

        def absolute_max(self) -> float:
            return numpy.max(numpy.abs(self._data))

        def min(self) -> float:
            return numpy.min(self._data)

        def var(self) -> float:
            return numpy.var(self._data)

        def average(self) -> float:
            return sum(self._data) / len(self._data)

        def width(self) -> float:
            return self.max() - self.min()

        def half_width(self) -> float:
            return self.width() / 2

        def undulate(self) -> float:
            return abs((self.max() - self.min()) / self.average())

        def clear(self) -> "BaseUtils.Statistic":
            self._data = []
            return self
