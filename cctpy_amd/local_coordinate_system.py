"""Rigid frames (final_code/packages/local_coordinate_system.py:25-199)."""
from __future__ import annotations

from typing import Optional

from .base_utils import BaseUtils
from .point import P3


class LocalCoordinateSystem:
    """Origin plus orthonormal axes XI, YI, ZI with YI = ZI x XI."""

    def __init__(self, location: P3 = None, x_direction: P3 = None, z_direction: P3 = None):
        location = P3.origin() if location is None else location
        x_direction = P3.x_direct() if x_direction is None else x_direction
        z_direction = P3.z_direct() if z_direction is None else z_direction
        BaseUtils.equal(x_direction.copy().normalize() * z_direction.copy().normalize(), 0.0, err=1e-4,
                        msg=f"LocalCoordinateSystem: x_direction {x_direction} and z_direction {z_direction} "
                            f"are not orthogonal")
        self.location = location.copy()
        self.ZI = z_direction.copy().normalize()
        self.XI = x_direction.copy().normalize()
        self.YI = self.ZI @ self.XI

    def point_to_local_coordinate(self, p: P3) -> P3:
        d = p - self.location
        return P3(self.XI * d, self.YI * d, self.ZI * d)

    def point_to_global_coordinate(self, p: P3) -> P3:
        return self.location + (self.XI * p.x + self.YI * p.y + self.ZI * p.z)

    def vector_to_local_coordinate(self, v: P3) -> P3:
        return P3(self.XI * v, self.YI * v, self.ZI * v)

    def vector_to_global_coordinate(self, v: P3) -> P3:
        return self.XI * v.x + self.YI * v.y + self.ZI * v.z

    def __str__(self) -> str:
        return f"LOCATION={self.location}, xi={self.XI}, yi={self.YI}, zi={self.ZI}"

    __repr__ = __str__

    def __eq__(self, o: "LocalCoordinateSystem", err: float = 1e-6, msg: Optional[str] = None) -> bool:
        return (BaseUtils.equal(self.location, o.location, err, msg) and BaseUtils.equal(self.XI, o.XI, err, msg)
                and BaseUtils.equal(self.YI, o.YI, err, msg) and BaseUtils.equal(self.ZI, o.ZI, err, msg))

    __hash__ = None

    @staticmethod
    def create_by_y_and_z_direction(location: P3, y_direction: P3, z_direction: P3) -> "LocalCoordinateSystem":
        BaseUtils.equal(y_direction * z_direction, 0.0,
                        msg=f"LocalCoordinateSystem: y_direction {y_direction} and z_direction {z_direction} "
                            f"are not orthogonal")
        return LocalCoordinateSystem(location=location, x_direction=y_direction @ z_direction,
                                     z_direction=z_direction)

    @staticmethod
    def global_coordinate_system() -> "LocalCoordinateSystem":
        return LocalCoordinateSystem()

    def copy(self) -> "LocalCoordinateSystem":
        return LocalCoordinateSystem(location=self.location.copy(), x_direction=self.XI.copy(),
                                     z_direction=self.ZI.copy())

    def to_array(self) -> list:
        """[O, XI, YI, ZI] as 12 floats (the frame part of the device records)."""
        return self.location.to_list() + self.XI.to_list() + self.YI.to_list() + self.ZI.to_list()
