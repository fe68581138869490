"""2-D and 3-D value types with the reference's operators (final_code/packages/point.py).

``*`` is scaling or dot product, ``@`` the cross product; component
arithmetic follows the reference's order so host-side geometry rounds the
same way.
"""
from __future__ import annotations

import math
from typing import Callable, Generic, List, Optional, Tuple, TypeVar, Union

import numpy

T = TypeVar("T")


def _num(x) -> bool:
    return isinstance(x, (int, float)) and not isinstance(x, bool) or isinstance(x, numpy.floating)


class P2:
    """2-D point / vector (point.py:22-315)."""

    __slots__ = ("x", "y")

    def __init__(self, x: float = 0.0, y: float = 0.0):
        self.x = float(x)
        self.y = float(y)

    def length(self) -> float:
        return math.sqrt(self.x ** 2 + self.y ** 2)

    def normalize(self) -> "P2":
        return self * (1 / self.length())

    def change_length(self, new_length: float) -> "P2":
        return self.normalize() * float(new_length)

    def copy(self) -> "P2":
        return P2(self.x, self.y)

    def __add__(self, o: "P2") -> "P2":
        return P2(self.x + o.x, self.y + o.y)

    def __neg__(self) -> "P2":
        return P2(-self.x, -self.y)

    def __sub__(self, o: "P2") -> "P2":
        return self + (-o)

    def __iadd__(self, o: "P2") -> "P2":
        self.x += o.x
        self.y += o.y
        return self

    def __isub__(self, o: "P2") -> "P2":
        self.x -= o.x
        self.y -= o.y
        return self

    def rotate(self, phi: float) -> "P2":
        c, s = math.cos(phi), math.sin(phi)
        return P2(c * self.x + (-s) * self.y, s * self.x + c * self.y)

    def angle_to_x_axis(self) -> float:
        a = float(math.atan2(self.y, self.x))
        return a if a >= 0 else math.pi * 2 + a

    def __mul__(self, o):
        if _num(o):
            return P2(self.x * o, self.y * o)
        return self.x * o.x + self.y * o.y

    def __rmul__(self, o):
        return self.__mul__(o)

    def __truediv__(self, n):
        if _num(n):
            return self * (1 / n)
        raise ValueError(f"P2{self} can only be divided by a number")

    def angle_to(self, o: "P2") -> float:
        to_x = self.angle_to_x_axis()
        return o.rotate(-to_x).angle_to_x_axis()

    def to_p3(self, transformation: Callable[["P2"], "P3"] = lambda p2: P3(p2.x, p2.y, 0.0)) -> "P3":
        return transformation(self)

    def __str__(self) -> str:
        return f"({self.x}, {self.y})"

    __repr__ = __str__

    def __eq__(self, o: "P2", err: float = 1e-6, msg: Optional[str] = None) -> bool:
        if not isinstance(o, P2):
            raise ValueError(f"{o} is not a P2")
        if abs(self.x - o.x) <= err and abs(self.y - o.y) <= err:
            return True
        if msg is None:
            return False
        raise AssertionError(msg)

    __hash__ = None

    @staticmethod
    def x_direct(x: float = 1.0) -> "P2":
        return P2(x=x)

    @staticmethod
    def y_direct(y: float = 1.0) -> "P2":
        return P2(y=y)

    @staticmethod
    def origin() -> "P2":
        return P2()

    @staticmethod
    def zeros() -> "P2":
        return P2()

    def to_list(self) -> List[float]:
        return [self.x, self.y]

    @staticmethod
    def from_numpy_ndarry(a: numpy.ndarray):
        a = numpy.asarray(a)
        if a.shape in ((2,), (2, 1)):
            return P2(float(a[0]), float(a[1]))
        if a.ndim == 2 and a.shape[1] == 2:
            return [P2(float(r[0]), float(r[1])) for r in a]
        raise ValueError(f"cannot convert array of shape {a.shape} to P2")

    @classmethod
    def from_list(cls, li: List):
        if len(li) == 2 and _num(li[0]):
            return P2(li[0], li[1])
        return [cls.from_list(e) for e in li]

    @staticmethod
    def extract(ps: List["P2"]) -> Tuple[List[float], List[float]]:
        return [p.x for p in ps], [p.y for p in ps]

    @staticmethod
    def extract_x(ps: List["P2"]) -> List[float]:
        return [p.x for p in ps]

    @staticmethod
    def extract_y(ps: List["P2"]) -> List[float]:
        return [p.y for p in ps]


class P3:
    """3-D point / vector (point.py:317-588)."""

    __slots__ = ("x", "y", "z")

    def __init__(self, x: float = 0.0, y: float = 0.0, z: float = 0.0):
        self.x = float(x)
        self.y = float(y)
        self.z = float(z)

    def length(self) -> float:
        return math.sqrt(self.x ** 2 + self.y ** 2 + self.z ** 2)

    def normalize(self) -> "P3":
        return self * (1 / self.length())

    def change_length(self, new_length: float) -> "P3":
        return self.normalize() * new_length

    def copy(self) -> "P3":
        return P3(self.x, self.y, self.z)

    def __add__(self, o: "P3") -> "P3":
        return P3(self.x + o.x, self.y + o.y, self.z + o.z)

    def __neg__(self) -> "P3":
        return P3(-self.x, -self.y, -self.z)

    def __sub__(self, o: "P3") -> "P3":
        return self + (-o)

    def __iadd__(self, o: "P3") -> "P3":
        self.x += o.x
        self.y += o.y
        self.z += o.z
        return self

    def __isub__(self, o: "P3") -> "P3":
        self.x -= o.x
        self.y -= o.y
        self.z -= o.z
        return self

    def __mul__(self, o):
        if _num(o):
            return P3(self.x * o, self.y * o, self.z * o)
        if isinstance(o, P3):
            return self.x * o.x + self.y * o.y + self.z * o.z
        raise ValueError(f"{self} and {o} cannot be multiplied")

    def __rmul__(self, o):
        return self.__mul__(o)

    def __truediv__(self, n):
        if _num(n):
            if n == 0:
                raise ValueError(f"{self}/{n}: division by zero")
            return self * (1 / n)
        raise ValueError("P3 can only be divided by a number")

    def __matmul__(self, o: "P3") -> "P3":
        return P3(self.y * o.z - self.z * o.y, -self.x * o.z + self.z * o.x, self.x * o.y - self.y * o.x)

    def __str__(self) -> str:
        return f"({self.x}, {self.y}, {self.z})"

    __repr__ = __str__

    def __eq__(self, o: "P3", err: float = 1e-6, msg: Optional[str] = None) -> bool:
        if not isinstance(o, P3):
            raise ValueError(f"{o} is not a P3")
        if abs(self.x - o.x) <= err and abs(self.y - o.y) <= err and abs(self.z - o.z) <= err:
            return True
        if msg is None:
            return False
        raise AssertionError(msg)

    __hash__ = None

    @staticmethod
    def x_direct(x: float = 1.0) -> "P3":
        return P3(x=x)

    @staticmethod
    def y_direct(y: float = 1.0) -> "P3":
        return P3(y=y)

    @staticmethod
    def z_direct(z: float = 1.0) -> "P3":
        return P3(z=z)

    @staticmethod
    def origin() -> "P3":
        return P3()

    @staticmethod
    def zeros() -> "P3":
        return P3()

    def to_list(self) -> List[float]:
        return [self.x, self.y, self.z]

    @staticmethod
    def from_numpy_ndarry(a: numpy.ndarray):
        a = numpy.asarray(a)
        if a.shape in ((3,), (3, 1)):
            return P3(float(a[0]), float(a[1]), float(a[2]))
        if a.ndim == 2 and a.shape[1] == 3:
            return [P3(float(r[0]), float(r[1]), float(r[2])) for r in a]
        raise ValueError(f"cannot convert array of shape {a.shape} to P3")

    def to_numpy_ndarry3(self, numpy_dtype=numpy.float64) -> numpy.ndarray:
        return numpy.array(self.to_list(), dtype=numpy_dtype)

    def to_p2(self, transformation: Callable[["P3"], P2] = lambda p3: P2(p3.x, p3.y)) -> P2:
        return transformation(self)

    def populate(self, o: "P3") -> None:
        self.x, self.y, self.z = o.x, o.y, o.z

    @staticmethod
    def random() -> "P3":
        return P3(*numpy.random.random(3))

    @staticmethod
    def as_p3(anything) -> "P3":
        return anything

    @staticmethod
    def extract(ps: List["P3"]) -> Tuple[List[float], List[float], List[float]]:
        return [p.x for p in ps], [p.y for p in ps], [p.z for p in ps]

    @staticmethod
    def extract_x(ps: List["P3"]) -> List[float]:
        return [p.x for p in ps]

    @staticmethod
    def extract_y(ps: List["P3"]) -> List[float]:
        return [p.y for p in ps]

    @staticmethod
    def extract_z(ps: List["P3"]) -> List[float]:
        return [p.z for p in ps]


class ValueWithDistance(Generic[T]):
    """A value tagged with a distance along a line (point.py:590-630)."""

    def __init__(self, value: T, distance: float) -> None:
        self.value = value
        self.distance = distance

    def __str__(self) -> str:
        return f"({self.distance}:{self.value})"

    __repr__ = __str__

    @staticmethod
    def convert_to_p2(data, convertor: Callable[[T], float]):
        if isinstance(data, ValueWithDistance):
            return P2(data.distance, convertor(data.value))
        return [P2(d.distance, convertor(d.value)) for d in data]
