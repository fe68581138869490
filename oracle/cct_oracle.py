"""FP64 CPU oracle for the CCT Biot-Savart + RK4 tracking path.

TEST INFRASTRUCTURE ONLY. This module is the checker: it is imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py``, never by the ``cctpy_amd`` product package.

It restates, from scratch and in plain Python floats plus numpy, the
arithmetic of the reference CPU path (``final_code/packages``) in the same
operation order, so that the restated field and tracking results agree with
the reference to the last bits wherever the same libm/numpy are used:

* vectors are 3-tuples of Python floats; every helper below evaluates its
  components in the order the reference's ``P2``/``P3`` operators do
  (``final_code/packages/point.py:27-135, 317-421``);
* the Biot-Savart sum uses the same numpy calls as
  ``CCT.magnetic_field_at`` (``final_code/packages/cct.py:249-285``):
  ``norm(axis=1) ** -3``, ``numpy.cross`` and ``numpy.sum(axis=0)``;
* RK4 follows ``BaseUtils.runge_kutta4`` (``base_utils.py:263-311``) with the
  callback of ``ParticleRunner`` (``particles.py:514-537``).

Parity pins (checked in ``tests/test_oracle_golden.py``): every function here
is compared with vectors printed by the reference's own demos and unit tests
(``final_code/demos/A07, A09, A11, A13, A14, C05, C06``,
``codes/tests/sole_layer_cct_test.py``, ``codes/tests/cuda_test.py``) and with
closed forms (circular motion radius, on-axis solenoid field, linspace end
points, rotation invariants).
"""

from __future__ import annotations

import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy

Vec3 = Tuple[float, float, float]
Vec2 = Tuple[float, float]

# ---------------------------------------------------------------- constants
# final_code/packages/constants.py:9-17 and particles.py:29-50
MM = 0.001
MRAD = 1.0 * MM * 1.0
LIGHT_SPEED = 299792458.0 * 1.0
EV = 1.6021766208e-19 * 1.0
MEV = 1000 * 1000 * EV
PROTON_MASS_KG = 1.672621898e-27
PROTON_ENERGY_MEV = 938.2720813
PROTON_CHARGE = 1.6021766208e-19


# ------------------------------------------------------------ vector helpers
# Each mirrors one P2/P3 operator of final_code/packages/point.py.

def v_add(a: Vec3, b: Vec3) -> Vec3:  # point.py:364-368
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def v_sub(a: Vec3, b: Vec3) -> Vec3:  # point.py:376-380 (a + (-b))
    return (a[0] + (-b[0]), a[1] + (-b[1]), a[2] + (-b[2]))


def v_scale(a: Vec3, k: float) -> Vec3:  # point.py:387-393
    return (a[0] * k, a[1] * k, a[2] * k)


def v_dot(a: Vec3, b: Vec3) -> float:  # point.py:394-395
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def v_cross(a: Vec3, b: Vec3) -> Vec3:  # point.py:413-421
    return (a[1] * b[2] - a[2] * b[1],
            -a[0] * b[2] + a[2] * b[0],
            a[0] * b[1] - a[1] * b[0])


def v_len(a: Vec3) -> float:  # point.py:327-331
    return math.sqrt(a[0] ** 2 + a[1] ** 2 + a[2] ** 2)


def v_normalize(a: Vec3) -> Vec3:  # point.py:333-337
    return v_scale(a, 1 / v_len(a))


def v_change_length(a: Vec3, n: float) -> Vec3:  # point.py:339-343
    return v_scale(v_normalize(a), n)


def p2_add(a: Vec2, b: Vec2) -> Vec2:
    return (a[0] + b[0], a[1] + b[1])


def p2_sub(a: Vec2, b: Vec2) -> Vec2:  # point.py:67-71
    return (a[0] + (-b[0]), a[1] + (-b[1]))


def p2_scale(a: Vec2, k: float) -> Vec2:
    return (a[0] * k, a[1] * k)


def p2_dot(a: Vec2, b: Vec2) -> float:
    return a[0] * b[0] + a[1] * b[1]


def p2_len(a: Vec2) -> float:  # point.py:31-35
    return math.sqrt(a[0] ** 2 + a[1] ** 2)


def p2_normalize(a: Vec2) -> Vec2:
    return p2_scale(a, 1 / p2_len(a))


def p2_change_length(a: Vec2, n: float) -> Vec2:  # point.py:43-47
    return p2_scale(p2_normalize(a), float(n))


def p2_rotate(a: Vec2, phi: float) -> Vec2:  # point.py:89-112
    c, s = math.cos(phi), math.sin(phi)
    return (c * a[0] + (-s) * a[1], s * a[0] + c * a[1])


def p2_angle_to_x_axis(a: Vec2) -> float:  # point.py:114-119
    t = float(math.atan2(a[1], a[0]))
    return t if t >= 0 else math.pi * 2 + t


def p2_eq(a: Vec2, b: Vec2, err: float = 1e-6) -> bool:  # point.py:179-191
    return abs(a[0] - b[0]) <= err and abs(a[1] - b[1]) <= err


def deg_to_rad(d: float) -> float:  # base_utils.py:92-105
    return d / 180.0 * math.pi


def equal(a: float, b: float, err: float = 1e-6) -> bool:  # base_utils.py:47-55
    return a == b or abs(a - b) <= err or ((a + b != 0.0) and (2 * abs((a - b) / (a + b)) <= err))


def linspace(start: float, end: float, number: int) -> numpy.ndarray:
    """base_utils.py:80-90: d=(end-start)*(1/(number-1)); x_i=start+d*i."""
    d = (end - start) * (1 / (number - 1))
    return start + d * numpy.arange(number, dtype=numpy.float64)


def circle_center_and_radius(p1: Vec2, p2: Vec2, p3: Vec2) -> Tuple[Vec2, float]:
    """base_utils.py:124-151: least-squares circle through three points."""
    x1, y1 = p1
    x2, y2 = p2
    x3, y3 = p3
    z1 = x2 ** 2 + y2 ** 2 - x1 ** 2 - y1 ** 2
    z2 = x3 ** 2 + y3 ** 2 - x1 ** 2 - y1 ** 2
    z3 = x3 ** 2 + y3 ** 2 - x2 ** 2 - y2 ** 2
    a = numpy.array([[(x2 - x1), (y2 - y1)], [(x3 - x1), (y3 - y1)], [(x3 - x2), (y3 - y2)]])
    b = 0.5 * numpy.array([[z1], [z2], [z3]])
    c = numpy.linalg.inv(a.T @ a) @ a.T @ b
    cx, cy = float(c[0][0]), float(c[1][0])
    r1 = math.sqrt((cx - x1) ** 2 + (cy - y1) ** 2)
    r2 = math.sqrt((cx - x2) ** 2 + (cy - y2) ** 2)
    r3 = math.sqrt((cx - x3) ** 2 + (cy - y3) ** 2)
    return (cx, cy), (r1 + r2 + r3) / 3


# ------------------------------------------------------------------ frames
class Frame:
    """Rigid frame, local_coordinate_system.py:25-125 (YI = ZI x XI)."""

    def __init__(self, location: Vec3, x_dir: Vec3, z_dir: Vec3):
        if not equal(v_dot(v_normalize(x_dir), v_normalize(z_dir)), 0.0, err=1e-4):
            raise AssertionError("frame axes are not orthogonal")
        self.O = tuple(location)
        self.ZI = v_normalize(z_dir)
        self.XI = v_normalize(x_dir)
        self.YI = v_cross(self.ZI, self.XI)

    @staticmethod
    def from_y_z(location: Vec3, y_dir: Vec3, z_dir: Vec3) -> "Frame":
        """local_coordinate_system.py:152-177: x = y @ z."""
        if not equal(v_dot(y_dir, z_dir), 0.0):
            raise AssertionError("frame axes are not orthogonal")
        return Frame(location, v_cross(y_dir, z_dir), z_dir)

    def to_local(self, p: Vec3) -> Vec3:  # local_coordinate_system.py:60-74
        d = v_sub(p, self.O)
        return (v_dot(self.XI, d), v_dot(self.YI, d), v_dot(self.ZI, d))

    def vec_to_global(self, v: Vec3) -> Vec3:  # local_coordinate_system.py:109-125
        return v_add(v_add(v_scale(self.XI, v[0]), v_scale(self.YI, v[1])), v_scale(self.ZI, v[2]))

    def point_to_global(self, p: Vec3) -> Vec3:  # local_coordinate_system.py:76-91
        return v_add(self.O, self.vec_to_global(p))

    def as_array(self) -> numpy.ndarray:
        return numpy.array(list(self.O) + list(self.XI) + list(self.YI) + list(self.ZI))


GLOBAL_FRAME = Frame((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0))


# -------------------------------------------------------------- trajectory
class _Line:
    """line2s.py:257-295 StraightLine2."""

    def __init__(self, length: float, direct: Vec2, start: Vec2):
        self.length = float(length)
        self.direct = p2_normalize(direct)
        self.start = start

    def point_at(self, s: float) -> Vec2:
        return p2_add(self.start, p2_change_length(self.direct, s))

    def direct_at(self, s: float) -> Vec2:
        return self.direct


class _Arc:
    """line2s.py:431-547 ArcLine2 (polar description around the centre)."""

    def __init__(self, start: Vec2, start_dir: Vec2, radius: float, clockwise: bool, deg: float):
        turn = -math.pi / 2 if clockwise else math.pi / 2
        self.center = p2_add(start, p2_change_length(p2_rotate(start_dir, turn), radius))
        self.phi0 = p2_angle_to_x_axis(p2_sub(start, self.center))
        self.radius = radius
        self.total_phi = deg_to_rad(deg)
        self.clockwise = clockwise
        self.length = radius * self.total_phi

    def _phi(self, s: float) -> float:
        d = s / self.radius
        return self.phi0 - d if self.clockwise else self.phi0 + d

    def point_at(self, s: float) -> Vec2:
        ph = self._phi(s)
        return p2_add(p2_change_length((math.cos(ph), math.sin(ph)), self.radius), self.center)

    def direct_at(self, s: float) -> Vec2:
        ph = self._phi(s)
        return p2_rotate((math.cos(ph), math.sin(ph)), -math.pi / 2 if self.clockwise else math.pi / 2)


class Trajectory:
    """trajectory.py:27-192: chain of straight lines and arcs."""

    def __init__(self, start: Vec2, direct: Vec2, length: float):
        self.parts: list = [_Line(length, direct, start)]
        self.length = self.parts[0].length

    def _end(self) -> Tuple[Vec2, Vec2]:
        last = self.parts[-1]
        return last.point_at(last.length), last.direct_at(last.length)

    def add_line(self, length: float) -> "Trajectory":
        sp, sd = self._end()
        self.parts.append(_Line(length, sd, sp))
        self.length += length
        return self

    def add_arc(self, radius: float, clockwise: bool, deg: float) -> "Trajectory":
        sp, sd = self._end()
        arc = _Arc(sp, sd, radius, clockwise, deg)
        self.parts.append(arc)
        self.length += arc.length
        return self

    def _locate(self, s: float):
        for part in self.parts:
            if s <= part.length:
                return part, s, False
            s -= part.length
        return self.parts[-1], s, True

    def point_at(self, s: float) -> Vec2:  # trajectory.py:110-134
        part, s, over = self._locate(s)
        if over and abs(s) <= 1e-8:
            return part.point_at(part.length)
        return part.point_at(s)

    def direct_at(self, s: float) -> Vec2:  # trajectory.py:136-160
        part, s, over = self._locate(s)
        if over and abs(s) <= 1e-8:
            return part.direct_at(part.length)
        return part.direct_at(s)


# ----------------------------------------------------------------- magnets
class CCTMagnet:
    """One CCT layer: geometry of cct.py:36-188/213-224 and field of 249-285."""

    def __init__(self, frame: Frame, big_r: float, small_r: float, bending_angle: float,
                 tilt_angles: Sequence[float], winding_number: int, current: float,
                 start: Vec2, end: Vec2, disperse_number_per_winding: int = 120):
        if bending_angle < 0 or big_r < 0 or small_r < 0 or small_r >= big_r:
            raise ValueError("invalid CCT geometry")
        self.frame = frame
        self.big_r = float(big_r)
        self.small_r = float(small_r)
        self.bending_angle = float(bending_angle)
        self.tilt_angles = [float(t) for t in tilt_angles]
        self.winding_number = int(winding_number)
        self.current = float(current)
        self.start = (float(start[0]), float(start[1]))
        self.end = (float(end[0]), float(end[1]))
        self.n = int(disperse_number_per_winding)
        self.a = math.sqrt(self.big_r ** 2 - self.small_r ** 2)
        self.eta = 0.5 * math.log((self.big_r + self.a) / (self.big_r - self.a))
        self.path3 = self.local_points(linspace(self.start[0], self.end[0], self.winding_number * self.n + 1))
        self.kl = 1e-7 * self.current * (self.path3[1:] - self.path3[:-1])
        self.mid = 0.5 * (self.path3[1:] + self.path3[:-1])

    def phi_of_ksi(self, ksi):
        """cct.py:166-188: phi = k ksi + b + sum_j cot(t_j)/((j+1) sinh eta) sin((j+1) ksi)."""
        x1, y1 = self.start
        x2, y2 = self.end
        k = (y2 - y1) / (x2 - x1)
        b = -k * x1 + y1
        phi = k * ksi + b
        for j, t in enumerate(self.tilt_angles):
            if equal(t, 90.0):
                continue
            phi = phi + (1 / math.tan(deg_to_rad(t))) / ((j + 1) * math.sinh(self.eta)) * numpy.sin((j + 1) * ksi)
        return phi

    def local_points(self, ksi: numpy.ndarray) -> numpy.ndarray:
        """cct.py:213-224 bipolar-toroidal map (ksi, phi) -> (x, y, z)."""
        phi = self.phi_of_ksi(ksi)
        sh, ch = math.sinh(self.eta), math.cosh(self.eta)
        t = self.a / (ch - numpy.cos(ksi))
        return numpy.stack([t * sh * numpy.cos(phi), t * sh * numpy.sin(phi), t * numpy.sin(ksi)], axis=1)

    def field_local(self, p_local: Vec3) -> numpy.ndarray:
        """cct.py:258-276 with the same numpy calls (sequential row sum)."""
        return bs_sum(self.kl, self.mid, p_local)

    def field(self, p: Vec3) -> Vec3:
        if equal(self.current, 0, err=1e-6):  # cct.py:254-255
            return (0.0, 0.0, 0.0)
        b = self.field_local(self.frame.to_local(p))
        return self.frame.vec_to_global((float(b[0]), float(b[1]), float(b[2])))

    def global_segments(self) -> Tuple[numpy.ndarray, numpy.ndarray]:
        """Segments in the global frame (kl rotated, midpoints moved), FP64 (N,3)."""
        f = self.frame
        R = numpy.array([f.XI, f.YI, f.ZI])  # rows are the local axes
        return self.kl @ R, self.mid @ R + numpy.array(f.O)

    def conductor_length(self, line_number: int, disperse_number_per_winding: int) -> float:
        """cct.py:475-491."""
        pts = self.local_points(linspace(self.start[0], self.end[0],
                                         disperse_number_per_winding * self.winding_number + 1))
        total = 0.0
        for i in range(len(pts) - 1):
            total += v_len(tuple(float(c) for c in (pts[i + 1] - pts[i])))
        return total * line_number


def segments_from_path(path3: numpy.ndarray, current: float) -> Tuple[numpy.ndarray, numpy.ndarray]:
    """cct.py:157-164: kl_i = 1e-7 I (P_{i+1} - P_i), m_i = (P_{i+1} + P_i) / 2 for a polyline (n+1, 3)."""
    return 1e-7 * current * (path3[1:] - path3[:-1]), 0.5 * (path3[1:] + path3[:-1])


def bs_sum(kl: numpy.ndarray, mid: numpy.ndarray, p: Vec3) -> numpy.ndarray:
    """cct.py:261-276: sum_i kl_i x (p - m_i) |p - m_i|^-3, same numpy calls and summation order."""
    r = numpy.array(p, dtype=numpy.float64) - mid
    rr = (numpy.linalg.norm(r, ord=2, axis=1) ** (-3)).reshape((r.shape[0], 1))
    return numpy.sum(numpy.cross(kl, r) * rr, axis=0)


class QSMagnet:
    """Hard-edge quadrupole+sextupole, magnets.py:589-668."""

    def __init__(self, frame: Frame, length: float, gradient: float, second_gradient: float,
                 aperture_radius: float):
        self.frame = frame
        self.length = float(length)
        self.gradient = float(gradient)
        self.second_gradient = float(second_gradient)
        self.aperture_radius = float(aperture_radius)

    def field(self, p: Vec3) -> Vec3:
        x, y, z = self.frame.to_local(p)
        if z < 0 or z > self.length:
            return (0.0, 0.0, 0.0)
        a = self.aperture_radius
        if abs(x) > a or abs(y) > a or math.sqrt(x ** 2 + y ** 2) > a:
            return (0.0, 0.0, 0.0)
        bx = self.gradient * y + self.second_gradient * (x * y)
        by = self.gradient * x + 0.5 * self.second_gradient * (x ** 2 - y ** 2)
        return v_add(v_scale(self.frame.XI, bx), v_scale(self.frame.YI, by))

    def packed(self) -> numpy.ndarray:
        """magnets.py:726-748: [O, XI, YI, ZI, L, g, s, a]."""
        return numpy.concatenate([self.frame.as_array(),
                                  [self.length, self.gradient, self.second_gradient, self.aperture_radius]])


class UniformMagnet:
    """magnets.py:417-437."""

    def __init__(self, b: Vec3):
        self.b = tuple(b)

    def field(self, p: Vec3) -> Vec3:
        return self.b


def total_field(magnets: Sequence, p: Vec3) -> Vec3:
    """beamline.py:46-53: sum in append order starting from zero."""
    b = (0.0, 0.0, 0.0)
    for m in magnets:
        b = v_add(b, m.field(p))
    return b


# ----------------------------------------------------------- placement
def position_of(start: Vec2, direct: Vec2, p: Vec2) -> int:
    """line2s.py:303-342 StraightLine2(1, direct, start).position_of(p)."""
    d = p2_normalize(direct)
    rel = p2_sub(p, start)
    proj = p2_scale(d, p2_dot(d, rel))
    vert = p2_sub(rel, proj)
    if p2_eq(vert, (0.0, 0.0)):
        return 0
    vert = p2_normalize(vert)
    right = p2_normalize(p2_rotate(d, deg_to_rad(-90)))
    return 1 if p2_eq(vert, right) else -1


def create_cct_along(traj: Trajectory, s: float, big_r: float, small_r: float, bending_angle: float,
                     tilt_angles: Sequence[float], winding_number: int, current: float,
                     start: Vec2, end: Vec2, disperse_number_per_winding: int = 120) -> CCTMagnet:
    """cct.py:341-421: centre from two three-point circle fits, y axis along +/- start direction."""
    sp = traj.point_at(s)
    arc_length = big_r * deg_to_rad(bending_angle)
    ep = traj.point_at(s + arc_length)
    m0 = traj.point_at(s + arc_length / 3 * 1)
    m1 = traj.point_at(s + arc_length / 3 * 2)
    c1, _ = circle_center_and_radius(sp, m0, m1)
    c2, _ = circle_center_and_radius(m0, m1, ep)
    center = p2_scale(p2_add(c1, c2), 0.5)
    sd = traj.direct_at(s)
    pos = position_of(sp, sd, center)
    if pos == 0:
        raise ValueError("CCT centre lies on the design orbit")
    ydir = (-sd[0], -sd[1], -0.0) if pos == 1 else (sd[0], sd[1], 0.0)
    frame = Frame.from_y_z((center[0], center[1], 0.0), ydir, (0.0, 0.0, 1.0))
    return CCTMagnet(frame, big_r, small_r, bending_angle, tilt_angles, winding_number, current,
                     start, end, disperse_number_per_winding)


def create_qs_along(traj: Trajectory, s: float, length: float, gradient: float,
                    second_gradient: float, aperture_radius: float) -> QSMagnet:
    """magnets.py:687-724: z along the orbit, x = z rotated by +90 degrees."""
    o = traj.point_at(s)
    zd = traj.direct_at(s)
    xd = p2_rotate(zd, deg_to_rad(90))
    return QSMagnet(Frame((o[0], o[1], 0.0), (xd[0], xd[1], 0.0), (zd[0], zd[1], 0.0)),
                    length, gradient, second_gradient, aperture_radius)


class Beamline:
    """beamline.py:31-793 builder restated: trajectory + magnets in append order."""

    def __init__(self, start: Vec2 = (0.0, 0.0), direct: Vec2 = (1.0, 0.0), length: float = 1.0):
        self.traj = Trajectory(start, direct, length)
        self.magnets: list = []

    @property
    def length(self) -> float:
        return self.traj.length

    def append_drift(self, length: float) -> "Beamline":
        self.traj.add_line(length)
        return self

    def append_qs(self, length, gradient, second_gradient, aperture_radius) -> "Beamline":
        s0 = self.traj.length
        self.traj.add_line(length)
        self.magnets.append(create_qs_along(self.traj, s0, length, gradient, second_gradient, aperture_radius))
        return self

    def append_dipole_cct(self, big_r, small_r_inner, small_r_outer, bending_angle, tilt_angles,
                          winding_number, current, disperse_number_per_winding=120) -> "Beamline":
        """beamline.py:507-574: inner layer then outer layer (tilts and ksi end negated)."""
        s0 = self.traj.length
        self.traj.add_arc(big_r, bending_angle < 0, abs(bending_angle))
        end_phi = deg_to_rad(bending_angle)
        for small_r, sign in ((small_r_inner, 1), (small_r_outer, -1)):
            tilts = list(tilt_angles) if sign == 1 else [t * -1 for t in tilt_angles]
            self.magnets.append(create_cct_along(
                self.traj, s0, big_r, small_r, abs(bending_angle), tilts, winding_number, current,
                (0.0, 0.0), (sign * 2 * math.pi * winding_number, end_phi), disperse_number_per_winding))
        return self

    def append_agcct(self, big_r, small_rs, bending_angles, tilt_angles, winding_numbers, currents,
                     disperse_number_per_winding=120) -> "Beamline":
        """beamline.py:576-793: dipole outer, dipole inner, then each AG part inner/outer."""
        if len(small_rs) != 4:
            raise ValueError("small_rs must hold 4 radii")
        if not all(small_rs[::-1][i] <= small_rs[::-1][i + 1] for i in range(3)):
            raise ValueError("small_rs must be sorted from large to small")
        total = sum(bending_angles)
        s0 = self.traj.length
        self.traj.add_arc(big_r, total < 0, abs(total))
        n = disperse_number_per_winding
        wd = winding_numbers[0][0]
        total_rad = deg_to_rad(total)
        self.magnets.append(create_cct_along(
            self.traj, s0, big_r, small_rs[0], abs(total), [t * -1 for t in tilt_angles[0]], wd,
            currents[0], (0.0, 0.0), (-2 * math.pi * wd, total_rad), n))
        self.magnets.append(create_cct_along(
            self.traj, s0, big_r, small_rs[1], abs(total), list(tilt_angles[0]), wd,
            currents[0], (0.0, 0.0), (2 * math.pi * wd, total_rad), n))
        ws = winding_numbers[1]
        rads = [deg_to_rad(a) for a in bending_angles]
        ag_tilts = list(tilt_angles[1])
        start_in = start_out = (0.0, 0.0)
        end_in = end_out = (0.0, 0.0)
        for i in range(len(bending_angles)):
            if i == 0:
                start_in, start_out = (0.0, 0.0), (0.0, 0.0)
                end_in = (((-1.0) ** i) * 2 * math.pi * ws[i], rads[i])
                end_out = (((-1.0) ** (i + 1)) * 2 * math.pi * ws[i], rads[i])
            else:
                shift = (0, rads[i - 1] / ws[i - 1])
                start_in = p2_add(end_in, shift)
                start_out = p2_add(end_out, shift)
                end_in = p2_add(start_in, (((-1) ** i) * 2 * math.pi * ws[i], rads[i]))
                end_out = p2_add(start_out, (((-1) ** (i + 1)) * 2 * math.pi * ws[i], rads[i]))
            self.magnets.append(create_cct_along(
                self.traj, s0, big_r, small_rs[3], abs(bending_angles[i]), [t * -1 for t in ag_tilts],
                ws[i], currents[1], start_in, end_in, n))
            self.magnets.append(create_cct_along(
                self.traj, s0, big_r, small_rs[2], abs(bending_angles[i]), ag_tilts,
                ws[i], currents[1], start_out, end_out, n))
        return self

    def field(self, p: Vec3) -> Vec3:
        return total_field(self.magnets, p)


def hust_gantry_section(which: str, **kw) -> Beamline:
    """final_code/hust_sc_gantry.py:19-385 (defaults at 19-89); which in {first, second, total}."""
    P = dict(
        DL1=0.9007765, GAP1=0.4301517, GAP2=0.370816,
        qs1_length=0.2340128, qs1_aperture_radius=60 * MM, qs1_gradient=3.1, qs1_second_gradient=7.9,
        qs2_length=0.200139, qs2_aperture_radius=60 * MM, qs2_gradient=3.8, qs2_second_gradient=-34.5,
        cct12_big_r=0.95, agcct12_inner_small_r=72.5 * MM, agcct12_outer_small_r=88.5 * MM,
        dicct12_inner_small_r=104.5 * MM, dicct12_outer_small_r=120.5 * MM,
        agcct1_winding_number=21, agcct2_winding_number=21, dicct12_winding_number=42,
        dicct12_bending_angle=22.5, agcct1_bending_angle=22.5 / 2, agcct2_bending_angle=22.5 / 2,
        dicct12_tilt_angles=[30, 102.8, 95.6, 87.3], agcct12_tilt_angles=[93.5, 30, 78.3, 97.5],
        dicct12_current=-9780.7, agcct12_current=9736.8,
        DL2=2.1162209, GAP3=0.1978111,
        qs3_length=0.2382791, qs3_aperture_radius=60 * MM, qs3_gradient=-7.3733,
        qs3_second_gradient=-45.31 * 2,
        cct345_big_r=0.95, agcct345_inner_small_r=83 * MM, agcct345_outer_small_r=83 * MM + 15 * MM,
        dicct345_inner_small_r=83 * MM + 15 * MM * 2, dicct345_outer_small_r=83 * MM + 15 * MM * 3,
        agcct3_winding_number=21, agcct4_winding_number=50, agcct5_winding_number=50,
        dicct345_winding_number=128, dicct345_bending_angle=-67.5,
        agcct3_bending_angle=-(8 + 3.716404), agcct4_bending_angle=-(8 + 19.93897),
        agcct5_bending_angle=-(8 + 19.844626),
        dicct345_tilt_angles=[30, 80], agcct345_tilt_angles=[90, 30],
        dicct345_current=9664, agcct345_current=-6000, part_per_winding=120,
    )
    for k in kw:
        if k not in P:
            raise KeyError(k)
    P.update(kw)
    n = P["part_per_winding"]

    def first_half(bl: Beamline) -> Beamline:
        r12 = [P["dicct12_outer_small_r"], P["dicct12_inner_small_r"],
               P["agcct12_outer_small_r"], P["agcct12_inner_small_r"]]
        tl = [P["dicct12_tilt_angles"], P["agcct12_tilt_angles"]]
        cur = [P["dicct12_current"], P["agcct12_current"]]
        a1, a2 = P["agcct1_bending_angle"], P["agcct2_bending_angle"]
        w1, w2, wd = P["agcct1_winding_number"], P["agcct2_winding_number"], P["dicct12_winding_number"]
        bl.append_agcct(P["cct12_big_r"], r12, [a1, a2], tl, [[wd], [w1, w2]], cur, n)
        bl.append_drift(P["GAP1"])
        bl.append_qs(P["qs1_length"], P["qs1_gradient"], P["qs1_second_gradient"], P["qs1_aperture_radius"])
        bl.append_drift(P["GAP2"])
        bl.append_qs(P["qs2_length"], P["qs2_gradient"], P["qs2_second_gradient"], P["qs2_aperture_radius"])
        bl.append_drift(P["GAP2"])
        bl.append_qs(P["qs1_length"], P["qs1_gradient"], P["qs1_second_gradient"], P["qs1_aperture_radius"])
        bl.append_drift(P["GAP1"])
        bl.append_agcct(P["cct12_big_r"], r12, [a2, a1], tl, [[wd], [w2, w1]], cur, n)
        bl.append_drift(P["DL1"])
        return bl

    def second_half(bl: Beamline) -> Beamline:
        r345 = [P["dicct345_outer_small_r"], P["dicct345_inner_small_r"],
                P["agcct345_outer_small_r"], P["agcct345_inner_small_r"]]
        tl = [P["dicct345_tilt_angles"], P["agcct345_tilt_angles"]]
        cur = [P["dicct345_current"], P["agcct345_current"]]
        a3, a4, a5 = P["agcct3_bending_angle"], P["agcct4_bending_angle"], P["agcct5_bending_angle"]
        w3, w4, w5 = P["agcct3_winding_number"], P["agcct4_winding_number"], P["agcct5_winding_number"]
        wd = P["dicct345_winding_number"]
        bl.append_agcct(P["cct345_big_r"], r345, [a3, a4, a5], tl, [[wd], [w3, w4, w5]], cur, n)
        bl.append_drift(P["GAP3"])
        bl.append_qs(P["qs3_length"], P["qs3_gradient"], P["qs3_second_gradient"], P["qs3_aperture_radius"])
        bl.append_drift(P["GAP3"])
        bl.append_agcct(P["cct345_big_r"], r345, [a5, a4, a3], tl, [[wd], [w5, w4, w3]], cur, n)
        bl.append_drift(P["DL2"])
        return bl

    if which == "first":
        return first_half(Beamline((0.0, 0.0), (1.0, 0.0), P["DL1"]))
    if which == "second":
        first = first_half(Beamline((0.0, 0.0), (1.0, 0.0), P["DL1"]))
        sp = first.traj.point_at(first.length)
        sd = first.traj.direct_at(first.length)
        return second_half(Beamline(sp, sd, P["DL2"]))
    if which == "total":
        bl = first_half(Beamline((0.0, 0.0), (1.0, 0.0), P["DL1"]))
        bl.append_drift(P["DL2"])
        return second_half(bl)
    raise ValueError(which)


# --------------------------------------------------------------- particles
class Particle:
    """RunningParticle state, particles.py:286-332 (position, velocity, mass, charge, speed, distance)."""

    def __init__(self, position: Vec3, velocity: Vec3, rm: float, e: float, speed: float, distance: float = 0.0):
        self.position = tuple(position)
        self.velocity = tuple(velocity)
        self.rm = rm
        self.e = e
        self.speed = speed
        self.distance = distance

    def copy(self) -> "Particle":
        return Particle(self.position, self.velocity, self.rm, self.e, self.speed, self.distance)

    def natural_frame(self, y_dir: Vec3 = (0.0, 0.0, 1.0)) -> Frame:  # particles.py:424-435
        return Frame.from_y_z(self.position, y_dir, self.velocity)

    def momentum(self) -> float:  # particles.py:372-378
        return self.speed * self.rm

    def as_array(self) -> numpy.ndarray:  # particles.py:443-457 packing order
        return numpy.array(list(self.position) + list(self.velocity) + [self.rm, self.e, self.speed, self.distance])


def protons_speed(kinetic_mev: float) -> float:
    """particles.py:63-106 Protons.get_speed_m_per_s (through the relativistic mass)."""
    rm = protons_relativistic_mass(kinetic_mev)
    return LIGHT_SPEED * math.sqrt(1 - (PROTON_MASS_KG / rm) ** 2)


def protons_relativistic_mass(kinetic_mev: float) -> float:  # particles.py:78-91
    return (PROTON_ENERGY_MEV + kinetic_mev) * MEV / LIGHT_SPEED / LIGHT_SPEED


def protons_momentum(kinetic_mev: float) -> float:  # particles.py:108-120
    return protons_relativistic_mass(kinetic_mev) * protons_speed(kinetic_mev)


def create_proton(position: Vec3, direct: Vec3, kinetic_mev: float = 250) -> Particle:
    """particles.py:1324-1353 (speed from the energy ratio)."""
    speed = LIGHT_SPEED * math.sqrt(1.0 - (PROTON_ENERGY_MEV / (PROTON_ENERGY_MEV + kinetic_mev)) ** 2)
    rm = PROTON_MASS_KG / math.sqrt(1.0 - (speed ** 2) / (LIGHT_SPEED ** 2))
    return Particle(tuple(position), v_change_length(direct, speed), rm, PROTON_CHARGE, speed)


def create_proton_along(traj, s: float = 0.0, kinetic_mev: float = 250) -> Particle:
    """particles.py:1381-1395."""
    p = traj.point_at(s)
    d = traj.direct_at(s)
    return create_proton((p[0], p[1], 0.0), (d[0], d[1], 0.0), kinetic_mev)


def change_scalar_momentum(p: Particle, momentum: float) -> None:
    """particles.py:380-422."""
    m0 = p.rm * math.sqrt(1 - (p.speed ** 2) / (LIGHT_SPEED ** 2))
    new_speed = momentum / math.sqrt(m0 ** 2 + (momentum / LIGHT_SPEED) ** 2)
    p.rm = m0 / math.sqrt(1 - (new_speed / LIGHT_SPEED) ** 2)
    p.speed = new_speed
    p.velocity = v_change_length(p.velocity, new_speed)


def from_phase_space(ideal: Particle, frame: Frame, x: float, xp: float, y: float, yp: float,
                     z: float, delta: float) -> Particle:
    """particles.py:1397-1437 ParticleFactory.create_from_phase_space_particle."""
    p = ideal.copy()
    p.position = v_add(p.position, v_scale(frame.XI, x))
    p.position = v_add(p.position, v_scale(frame.YI, y))
    p.position = v_add(p.position, v_scale(frame.ZI, z))
    if delta != 0.0:
        change_scalar_momentum(p, p.momentum() * (1.0 + delta))
    p.velocity = v_add(p.velocity, v_scale(frame.XI, xp * p.speed))
    p.velocity = v_add(p.velocity, v_scale(frame.YI, yp * p.speed))
    p.velocity = v_change_length(p.velocity, p.speed)
    return p


def to_phase_space(ideal: Particle, frame: Frame, rp: Particle) -> Tuple[float, float, float, float, float, float]:
    """particles.py:1077-1121 PhaseSpaceParticle.create_from_running_particle."""
    rel = v_sub(rp.position, ideal.position)
    x, y, z = v_dot(frame.XI, rel), v_dot(frame.YI, rel), v_dot(frame.ZI, rel)
    dv = v_sub(rp.velocity, ideal.velocity)
    xv, yv = v_dot(frame.XI, dv), v_dot(frame.YI, dv)
    xp = xv / math.sqrt(rp.speed ** 2 - xv ** 2)
    yp = yv / math.sqrt(rp.speed ** 2 - yv ** 2)
    rm, im = rp.momentum(), ideal.momentum()
    return (x, xp, y, yp, z, (rm - im) / im)


# --------------------------------------------------------------------- RK4
def rk4_steps(length: float, footstep: float, speed: float) -> Tuple[int, float]:
    """particles.py:585-587 + base_utils.py:284-285: step count and time step of one particle."""
    dt = footstep / speed
    t_end = length / speed
    number = math.ceil((t_end - 0.0) / dt)
    return number, (t_end - 0.0) / float(number)


def run_only(p: Particle, field_fn, length: float, footstep: float = 20 * MM,
             record: Optional[List] = None) -> Particle:
    """particles.py:561-596 single-particle branch: RK4 over [0, length/speed], distance += length.

    ``field_fn(point) -> Vec3`` is the magnet's field.  The RK4 arithmetic
    follows base_utils.py:303-309 on Y=[position, velocity] with
    callback Y'=[v, (e/m) v x B(p)] (particles.py:526-535).  Mutates and
    returns ``p``; with ``record`` a list receives (position, velocity) after
    every step (record=True branch, base_utils.py:287-300).
    """
    number, dt = rk4_steps(length, footstep, p.speed)
    kq = p.e / p.rm
    pos, vel = p.position, p.velocity
    h2, h6 = dt / 2, dt / 6
    if record is not None:
        record.append((pos, vel))
    for _ in range(number):
        a1 = v_scale(v_cross(vel, field_fn(pos)), kq)
        p2, v2 = v_add(pos, v_scale(vel, h2)), v_add(vel, v_scale(a1, h2))
        a2 = v_scale(v_cross(v2, field_fn(p2)), kq)
        p3, v3 = v_add(pos, v_scale(v2, h2)), v_add(vel, v_scale(a2, h2))
        a3 = v_scale(v_cross(v3, field_fn(p3)), kq)
        p4, v4 = v_add(pos, v_scale(v3, dt)), v_add(vel, v_scale(a3, dt))
        a4 = v_scale(v_cross(v4, field_fn(p4)), kq)
        sp = v_add(v_add(v_add(vel, v_scale(v2, 2)), v_scale(v3, 2)), v4)
        sv = v_add(v_add(v_add(a1, v_scale(a2, 2)), v_scale(a3, 2)), a4)
        pos = v_add(pos, v_scale(sp, h6))
        vel = v_add(vel, v_scale(sv, h6))
        if record is not None:
            record.append((pos, vel))
    p.position, p.velocity = pos, vel
    p.distance += length
    return p


# ------------------------------------------------------- packed geometry
def beamline_arrays(magnets: Sequence) -> Dict[str, numpy.ndarray]:
    """Flatten a magnet list for the C oracle: per-CCT frames and local segments, QS records,
    uniform fields, and the append order (kind 0 CCT, 1 QS, 2 uniform; index within its kind)."""
    ccts, qss, unis, order = [], [], [], []
    for m in magnets:
        if isinstance(m, CCTMagnet):
            if not equal(m.current, 0, err=1e-6):
                order.append((0, len(ccts)))
                ccts.append(m)
        elif isinstance(m, QSMagnet):
            order.append((1, len(qss)))
            qss.append(m)
        elif isinstance(m, UniformMagnet):
            order.append((2, len(unis)))
            unis.append(m)
        else:
            raise ValueError(f"unsupported magnet {m}")
    off = numpy.zeros(len(ccts) + 1, dtype=numpy.int64)
    for i, c in enumerate(ccts):
        off[i + 1] = off[i] + c.kl.shape[0]
    cat = (lambda xs: numpy.ascontiguousarray(numpy.concatenate(xs))) if ccts else (lambda xs: numpy.zeros((0, 3)))
    return dict(
        cct_frame=numpy.ascontiguousarray(numpy.array([c.frame.as_array() for c in ccts]).reshape(len(ccts), 12)),
        cct_off=off,
        kl=cat([c.kl for c in ccts]),
        mid=cat([c.mid for c in ccts]),
        qs=numpy.ascontiguousarray(numpy.array([q.packed() for q in qss]).reshape(len(qss), 16)),
        uni=numpy.ascontiguousarray(numpy.array([u.b for u in unis], dtype=numpy.float64).reshape(len(unis), 3)),
        order=numpy.ascontiguousarray(numpy.array(order, dtype=numpy.int64).reshape(len(order), 2)),
    )


# ------------------------------------------------------ phase ellipses
# BaseUtils.Ellipse (final_code/packages/base_utils.py:446-601): the ellipse
# A x^2 + B x y + C y^2 = D, walked counter-clockwise from its +x axis point in
# angle steps of 0.05 deg, chord lengths summed.
def ellipse_point_at(A: float, B: float, C: float, D: float, theta: float) -> Vec2:
    """base_utils.py:480-524: the ray from the origin at angle theta meets the ellipse."""
    x, y = 0.0, 0.0
    while theta < 0:
        theta += 2 * math.pi
    while theta > 2 * math.pi:
        theta -= 2 * math.pi
    if equal(theta, 0) or equal(theta, 2 * math.pi):
        x, y = math.sqrt(D / A), 0
    if equal(theta, math.pi):
        x, y = -math.sqrt(D / A), 0
    if 0 < theta < math.pi:
        t = 1 / math.tan(theta)
        y = math.sqrt(D / (A * t * t + B * t + C))
        x = t * y
    if math.pi < theta < 2 * math.pi:
        theta -= math.pi
        t = 1 / math.tan(theta)
        y = -math.sqrt(D / (A * t * t + B * t + C))
        x = t * y
    return (x, y)


def ellipse_circumference(A: float, B: float, C: float, D: float) -> float:
    """base_utils.py:529-551: sum of 14,400 chords at equal angle steps."""
    num = 3600 * 4
    c = 0.0
    for i in range(num):
        c += p2_len(p2_sub(ellipse_point_at(A, B, C, D, 2.0 * math.pi / float(num) * (i + 1)),
                           ellipse_point_at(A, B, C, D, 2.0 * math.pi / float(num) * i)))
    return c


def ellipse_point_after(A: float, B: float, C: float, D: float, length: float) -> Vec2:
    """base_utils.py:553-573: walk `length` along the edge in 0.05 deg angle steps."""
    step = deg_to_rad(0.05)
    theta = 0.0
    while length > 0.0:
        length -= p2_len(p2_sub(ellipse_point_at(A, B, C, D, theta + step), ellipse_point_at(A, B, C, D, theta)))
        theta += step
    return ellipse_point_at(A, B, C, D, theta)


def ellipse_edge_points(A: float, B: float, C: float, D: float, num: int) -> List[Vec2]:
    """base_utils.py:575-584 uniform_distribution_points_along_edge."""
    c = ellipse_circumference(A, B, C, D)
    return [ellipse_point_after(A, B, C, D, c / num * i) for i in range(num)]


def phase_ellipse(plane: str, amax: float, apmax: float, delta: float, number: int) -> List[Tuple]:
    """(x, xp, y, yp, z, delta) on the upright phase ellipse of one plane (particles.py:925-983)."""
    pts = ellipse_edge_points(1 / (amax ** 2), 0.0, 1 / (apmax ** 2), 1.0, number)
    if plane == "x":
        return [(p[0], p[1], 0.0, 0.0, 0.0, delta) for p in pts]
    return [(0.0, 0.0, p[0], p[1], 0.0, delta) for p in pts]


# ----------------------------------------------- gantry design search
def a05_gantry_kwargs(param: Sequence[float], part_per_winding: int = 60) -> Dict:
    """hust_gantry_section keyword arguments of one 27-value design vector, as
    final_code/work/optim/A05run.py:26-180 builds HUST_SC_GANTRY from it (no vector: its
    defaults)."""
    p = list(param)
    have = len(p) > 0

    def v(i, d):
        return p[i] if have else d

    ag1, ag2 = (int(p[12]), int(p[13])) if have else (20, 20)
    ag3, ag4, ag5 = (int(p[24]), int(p[25]), int(p[26])) if have else (20, 20, 20)
    ag345 = ag3 + ag4 + ag5
    return dict(
        DL1=0.8541973, GAP1=0.45, GAP2=0.45,
        qs1_length=0.27, qs1_aperture_radius=60 * MM, qs1_gradient=v(0, 0), qs1_second_gradient=v(2, 0),
        qs2_length=0.27, qs2_aperture_radius=60 * MM, qs2_gradient=v(1, 0), qs2_second_gradient=v(3, 0),
        cct12_big_r=0.95,
        agcct12_inner_small_r=92.5 * MM - 20 * MM + 17 * MM, agcct12_outer_small_r=108.5 * MM - 20 * MM + 17 * MM,
        dicct12_inner_small_r=124.5 * MM - 20 * MM + 17 * MM, dicct12_outer_small_r=140.5 * MM - 20 * MM + 17 * MM,
        agcct1_winding_number=ag1, agcct2_winding_number=ag2, dicct12_winding_number=42,
        dicct12_bending_angle=22.5,
        agcct1_bending_angle=22.5 * (ag1 / (ag1 + ag2)), agcct2_bending_angle=22.5 * (ag2 / (ag1 + ag2)),
        dicct12_tilt_angles=[30, v(4, 90), v(5, 90), v(6, 90)],
        agcct12_tilt_angles=[v(7, 90), 30, v(8, 90), v(9, 90)],
        dicct12_current=v(10, 10000), agcct12_current=v(11, 10000),
        DL2=2.6, GAP3=0.45,
        qs3_length=0.27, qs3_aperture_radius=60 * MM, qs3_gradient=v(14, 0), qs3_second_gradient=v(15, 0),
        cct345_big_r=0.95,
        agcct345_inner_small_r=92.5 * MM + 0.1 * MM + 17 * MM, agcct345_outer_small_r=108.5 * MM + 0.1 * MM + 17 * MM,
        dicct345_inner_small_r=124.5 * MM + 0.1 * MM + 17 * MM,
        dicct345_outer_small_r=140.5 * MM + 0.1 * MM + 17 * MM,
        agcct3_winding_number=ag3, agcct4_winding_number=ag4, agcct5_winding_number=ag5,
        dicct345_winding_number=128, dicct345_bending_angle=-67.5,
        agcct3_bending_angle=-67.5 * (ag3 / ag345), agcct4_bending_angle=-67.5 * (ag4 / ag345),
        agcct5_bending_angle=-67.5 * (ag5 / ag345),
        dicct345_tilt_angles=[30, v(16, 90), v(17, 90), v(18, 90)],
        agcct345_tilt_angles=[v(19, 90), 30, v(20, 90), v(21, 90)],
        dicct345_current=v(22, 10000), agcct345_current=v(23, 10000),
        part_per_winding=part_per_winding,
    )


def a05_objectives(pps: Sequence[Tuple], group: int) -> List[float]:
    """The 11 objectives of one design from its end-plane phase-space particles
    (final_code/work/optim/A05run.py:248-330): x, y in mm and xp, yp in mrad; absolute maxima,
    |averages|, |max|x| - max|y||, and the spread (max - min) of the half widths (max - min)/2
    of x and of y over consecutive groups of `group` particles."""
    x = [pp[0] / MM for pp in pps]
    xp = [pp[1] / MRAD for pp in pps]
    y = [pp[2] / MM for pp in pps]
    yp = [pp[3] / MRAD for pp in pps]
    hwx, hwy = [], []
    for g in range(0, len(pps), group):
        gx, gy = x[g:g + group], y[g:g + group]
        hwx.append((numpy.max(gx) - numpy.min(gx)) / 2)
        hwy.append((numpy.max(gy) - numpy.min(gy)) / 2)
    amax = [float(numpy.max(numpy.abs(a))) for a in (x, xp, y, yp)]
    cen = [abs(sum(a) / len(a)) for a in (x, xp, y, yp)]
    return amax + cen + [abs(amax[0] - amax[2]), float(numpy.max(hwx) - numpy.min(hwx)),
                         float(numpy.max(hwy) - numpy.min(hwy))]
