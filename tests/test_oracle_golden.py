"""The CPU oracle against the reference's own known-answer vectors and closed forms."""
import math

import numpy as np
import pytest

from oracle import cct_oracle as O
from oracle import c_oracle as C
from tests.golden.params import SECTION1_DEMO


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_protons_constants(golden):
    g = golden["protons_250"]
    assert O.PROTON_MASS_KG * O.LIGHT_SPEED * O.LIGHT_SPEED == g["static_energy_J"]
    assert O.protons_relativistic_mass(250) == g["relativistic_mass"]
    assert O.protons_speed(250) == g["speed"]
    assert O.protons_momentum(250) == g["momentum"]
    assert O.protons_momentum(250) / O.PROTON_CHARGE == pytest.approx(g["magnetic_stiffness"], rel=1e-15)


def test_factory_and_phase_space(golden):
    g = golden["factory"]
    p = O.create_proton((0, 0, 0), (1, 0, 0), 250)
    assert p.speed == g["proton_250_speed"] and p.rm == g["proton_250_rm"]
    traj = O.Trajectory((0, 0), (1, 0), 1)
    p = O.create_proton_along(traj, 0.5, 215)
    assert list(p.position) == g["along_215"]["position"]
    assert list(p.velocity) == g["along_215"]["velocity"] and p.rm == g["along_215"]["rm"]
    traj.add_arc(1, False, 45)
    p = O.create_proton_along(traj, traj.length, 200)
    assert rel(p.position, g["arc_end_200"]["position"]) < 1e-15
    assert rel(p.velocity, g["arc_end_200"]["velocity"]) < 1e-15
    assert p.speed == g["arc_end_200"]["speed"]
    ip = O.create_proton((0, 0, 0), (1, 0, 0), 215)
    cs = ip.natural_frame((0, 0, 1))
    ps = g["phase_space_in"]
    rp = O.from_phase_space(ip, cs, ps["x"], ps["xp"], ps["y"], ps["yp"], ps["z"], ps["delta"])
    assert list(rp.position) == g["phase_space_particle"]["position"]
    assert rel(rp.velocity, g["phase_space_particle"]["velocity"]) < 1e-15
    assert rp.speed == g["phase_space_particle"]["speed"]
    back = O.to_phase_space(ip, cs, rp)
    assert np.allclose(back, g["phase_space_back"], rtol=1e-14, atol=1e-18)


def test_rk4_uniform_field(golden):
    g = golden["rk4_uniform"]
    u = O.UniformMagnet((0, 0, g["bz"]))
    p = O.create_proton((0, 0, 0), (1, 0, 0), 250)
    O.run_only(p, u.field, math.pi / 2)
    assert list(p.position) == g["quarter"]["position"] and list(p.velocity) == g["quarter"]["velocity"]
    O.run_only(p, u.field, math.pi / 2)
    assert list(p.position) == g["half"]["position"] and list(p.velocity) == g["half"]["velocity"]
    O.run_only(p, u.field, math.pi / 2)
    assert list(p.position) == g["three_quarter"]["position"]
    assert list(p.velocity) == g["three_quarter"]["velocity"]
    rec = []
    O.run_only(O.create_proton((0, 0, 0), (1, 0, 0), 250), u.field, 1, 0.1, record=rec)
    assert len(rec) == len(g["all_info_1m_100mm"])
    for (pos, vel), (wp, wv) in zip(rec, g["all_info_1m_100mm"]):
        assert list(pos) == wp and list(vel) == wv


def test_rk4_closed_form_circle():
    """Gyration radius p/(qB): one full turn returns to the start within RK4 truncation error."""
    b = 2.0
    p = O.create_proton((0, 0, 0), (1, 0, 0), 250)
    radius = p.momentum() / (p.e * b)
    u = O.UniformMagnet((0, 0, b))
    O.run_only(p, u.field, 2 * math.pi * radius, 1e-3)
    assert abs(p.position[0]) < 1e-9 and abs(p.position[1]) < 1e-9
    assert abs(O.v_len(p.velocity) - p.speed) / p.speed < 1e-12


def test_qs_fields(golden):
    for key in ("qs_a13", "qs_c05"):
        g = golden[key]
        fr = O.Frame(tuple(g["frame"]["location"]), tuple(g["frame"]["x"]), tuple(g["frame"]["z"]))
        q = O.QSMagnet(fr, g["length"], g["gradient"], g["second_gradient"], g["aperture"])
        pts = g["points"] if "points" in g else [g["point"]]
        fs = g["fields"] if "fields" in g else [g["field"]]
        for pt, f in zip(pts, fs):
            assert list(q.field(tuple(pt))) == f


def test_cct_geometry(golden):
    g = golden["cct_a14"]
    cct = O.CCTMagnet(O.GLOBAL_FRAME, 1.0, 0.03, 30, [30], 30, 1000, (0, 0), (2 * 30 * math.pi, 30 / 180 * math.pi))
    assert cct.a == g["a"] and cct.eta == g["eta"]
    assert float(cct.phi_of_ksi(np.array([2 * math.pi]))[0]) == g["phi_at_2pi"]
    assert rel(cct.local_points(np.array([0.0]))[0], g["convert_origin"]) < 1e-15
    traj = O.Trajectory((0, 0), (1, 0), 1).add_arc(1, True, 30).add_line(1)
    c2 = O.create_cct_along(traj, 1, 1, 0.06, 30, [90, 30], 30, 1000, (0, 0), (30 * 2 * math.pi, -30 / 180 * math.pi))
    assert c2.conductor_length(1, 3600) == pytest.approx(g["conductor_length_3600"], rel=1e-14)
    assert c2.conductor_length(1, 360) == pytest.approx(g["conductor_length_360"], rel=1e-14)
    # invariants of the bipolar-toroidal map (cct.py:201-211): R = sqrt(a^2/(1-1/cosh^2 eta)), r = R/cosh eta
    assert cct.big_r == pytest.approx(math.sqrt(cct.a ** 2 / (1 - 1 / math.cosh(cct.eta) ** 2)), rel=1e-12)
    assert cct.small_r == pytest.approx(cct.big_r / math.cosh(cct.eta), rel=1e-12)
    # every path point lies on the torus of radii (R, r) in the local frame
    rho = np.hypot(cct.path3[:, 0], cct.path3[:, 1])
    assert np.max(np.abs(np.hypot(rho - cct.big_r, cct.path3[:, 2]) - cct.small_r)) < 1e-12
    # phi advances by bend/W per 2*pi of ksi when tilts vanish at multiples of 2*pi
    w = np.array([2 * math.pi * k for k in range(31)])
    assert np.allclose(np.diff(cct.phi_of_ksi(w)), math.radians(30) / 30, rtol=1e-12)


def test_bs_known_answers(golden):
    g = golden["solenoid"]
    t = np.linspace(0, g["turns"] * 2 * np.pi, g["turns"] * 360)
    path = np.array([[g["radius"] * np.sin(x), g["radius"] * np.cos(x), x / (g["turns"] * 2 * np.pi) * g["length"]] for x in t])
    kl, mid = O.segments_from_path(path, g["current"])
    for pt, f in zip(g["points"], g["fields"]):
        assert np.linalg.norm(O.bs_sum(kl, mid, tuple(pt)) - np.array(f)) < g["tolerance"]
    s = golden["single_segment"]
    kl, mid = O.segments_from_path(np.array([s["p0"], s["p1"]]), 1e7)
    assert rel(O.bs_sum(kl, mid, tuple(s["point"])), s["field"]) < 1e-14
    h = golden["helix"]
    t = np.linspace(0, h["turns"] * 2 * np.pi, h["turns"] * h["per_turn"])
    path = np.array([[h["radius"] * np.cos(x), h["radius"] * np.sin(x), x / (h["turns"] * 2 * np.pi) * h["length"]] for x in t])
    kl, mid = O.segments_from_path(path, h["current"])
    # the reference test builds this helix in float32 and asserts 1e-6 (cuda_test.py:146-156)
    assert np.linalg.norm(O.bs_sum(kl, mid, (0.0, 0.0, 0.0)) - np.array(h["field_at_origin"])) < 1e-6


def test_bs_closed_form_loop():
    """On-axis field of a circular loop, mu0 I R^2 / (2 (R^2+z^2)^1.5); midpoint rule error O(1/n^2)."""
    R, I, z = 0.05, 100.0, 0.03
    exact = 4e-7 * math.pi * I * R * R / (2 * (R * R + z * z) ** 1.5)
    errs = []
    for n in (720, 2880):
        t = np.linspace(0, 2 * np.pi, n + 1)
        path = np.stack([R * np.cos(t), R * np.sin(t), 0 * t], axis=1)
        kl, mid = O.segments_from_path(path, I)
        b = O.bs_sum(kl, mid, (0.0, 0.0, z))
        errs.append(abs(b[2] - exact) / exact)
        assert abs(b[0]) < 1e-12 and abs(b[1]) < 1e-12
    assert errs[1] < 1e-6 and 14.0 < errs[0] / errs[1] < 18.0  # second order: x16 for n x4


def test_beamline_field_c05(golden):
    g = golden["beamline_c05"]
    bl = (O.Beamline((0, 0), (1, 0), 1).append_qs(0.27, 5, 20, 0.1).append_drift(1)
          .append_dipole_cct(1, 0.1, 0.12, 45, [30], 60, 10000).append_drift(1))
    assert rel(bl.field(tuple(g["point_cpu"])), g["field_cpu"]) < 1e-12
    assert rel(bl.field(tuple(g["point_gpu64"])), g["field_gpu64"]) < 1e-12
    m = O.beamline_arrays(bl.magnets)
    cf = C.field(m, np.array([g["point_cpu"], g["point_gpu64"]]))
    assert rel(cf[0], g["field_cpu"]) < 1e-12 and rel(cf[1], g["field_gpu64"]) < 1e-12


def _track_c(bl, ip, footstep):
    m = O.beamline_arrays(bl.magnets)
    n, dt = O.rk4_steps(bl.length, footstep, ip.speed)
    st = np.array([list(ip.position) + list(ip.velocity)])
    return C.track(m, st, np.array([ip.e / ip.rm]), np.array([dt]), np.array([n]))[0]


@pytest.mark.parametrize("variant,kin,key", [
    ({}, 215, "cpu"), ({}, 220, "cpu_220MeV"),
    ({"agcct12_current": -3000}, 215, "cpu_agcct12_m3000_215MeV"),
    ({"agcct12_current": -3000}, 220, "cpu_agcct12_m3000_220MeV")])
def test_gantry_section1_tracking(golden, variant, kin, key):
    g = golden["gantry_section1_215MeV_100mm"]
    bl = O.hust_gantry_section("first", **{**SECTION1_DEMO, **variant})
    assert bl.length == g["length"]
    ip = O.create_proton_along(bl.traj, 0.0, kin)
    st = _track_c(bl, ip, 0.1)
    assert np.max(np.abs(st[:3] - np.array(g[key]["position"]))) < 1e-9
    assert rel(st[3:], g[key]["velocity"]) < 1e-9


def test_gantry_section2_tracking(golden):
    g = golden["gantry_section2_215MeV_100mm"]
    bl = O.hust_gantry_section("second", **SECTION1_DEMO)
    assert bl.length == g["length"]
    ip = O.create_proton_along(bl.traj, 0.0, 215)
    st = _track_c(bl, ip, 0.1)
    assert np.max(np.abs(st[:3] - np.array(g["cpu"]["position"]))) < 1e-9
    assert rel(st[3:], g["cpu"]["velocity"]) < 1e-9


def test_c_oracle_matches_numpy_oracle():
    """The C restatement and the numpy restatement agree on fields and on short tracks."""
    bl = O.hust_gantry_section("first", **SECTION1_DEMO)
    m = O.beamline_arrays(bl.magnets)
    rng = np.random.default_rng(7)
    pts = [(*bl.traj.point_at(s), z) for s, z in zip(rng.uniform(0, bl.length, 6), rng.uniform(-0.01, 0.01, 6))]
    cf = C.field(m, np.array(pts))
    for p, b in zip(pts, cf):
        assert rel(b, bl.field(p)) < 1e-13
    ip = O.create_proton_along(bl.traj, 0.0, 250)
    n, dt = O.rk4_steps(0.05, 0.01, ip.speed)
    st = C.track(m, np.array([list(ip.position) + list(ip.velocity)]), np.array([ip.e / ip.rm]), np.array([dt]), np.array([n]))[0]
    q = O.run_only(ip.copy(), bl.field, 0.05, 0.01)
    assert rel(st[:3], q.position) < 1e-14 and rel(st[3:], q.velocity) < 1e-14


def _ellipse_arc(a: float, b: float, theta: float, n: int = 400001) -> float:
    """Arc length of the ellipse x^2/a^2 + y^2/b^2 = 1 from its +x point to polar angle theta
    (trapezoid rule on the polar form r(t) = ab / sqrt(b^2 cos^2 t + a^2 sin^2 t))."""
    t = np.linspace(0.0, theta, n)
    r = a * b / np.sqrt((b * np.cos(t)) ** 2 + (a * np.sin(t)) ** 2)
    dr = np.gradient(r, t)
    f = np.sqrt(r * r + dr * dr)
    return float(np.sum((f[1:] + f[:-1]) * np.diff(t)) / 2)


@pytest.mark.parametrize("a,b,num", [(3.5e-3, 7.5e-3, 12), (7.5e-3, 3.5e-3, 8), (2e-3, 2e-3, 6)])
def test_phase_ellipse_points(a, b, num):
    """The edge walk of BaseUtils.Ellipse (base_utils.py:446-584) against closed forms: its
    chord-sum circumference against the exact perimeter 4 a E(1 - b^2/a^2) (complete elliptic
    integral of the second kind), every point on the ellipse, the first at (a, 0), the points
    counter-clockwise at arc lengths c i / num, each reached within one 0.05 deg angle step."""
    from scipy.special import ellipe
    A, B, C, D = 1 / a ** 2, 0.0, 1 / b ** 2, 1.0
    c = O.ellipse_circumference(A, B, C, D)
    big, small = max(a, b), min(a, b)
    assert c == pytest.approx(4 * big * ellipe(1 - small ** 2 / big ** 2), rel=1e-7)
    pts = O.ellipse_edge_points(A, B, C, D, num)
    assert pts[0] == (math.sqrt(D / A), 0)
    th = [math.atan2(p[1], p[0]) % (2 * math.pi) for p in pts]
    assert all(t1 > t0 for t0, t1 in zip(th, th[1:]))
    step_chord = max(a, b) * math.radians(0.05) * 1.01
    for i, (p, t) in enumerate(zip(pts, th)):
        assert A * p[0] ** 2 + C * p[1] ** 2 == pytest.approx(D, rel=1e-12)
        if i:
            s = _ellipse_arc(a, b, t)
            assert c / num * i - 1e-9 <= s <= c / num * i + step_chord, (i, s, c / num * i)
    # the package's ellipse (host code of the product) gives the same bits
    from cctpy_amd.base_utils import BaseUtils
    ours = BaseUtils.Ellipse(A, B, C, D).uniform_distribution_points_along_edge(num)
    assert [(q.x, q.y) for q in ours] == pts


def test_a05_objectives_closed_form():
    """The 11 design-search objectives (A05run.py:248-330) on a hand-made end plane: two groups
    of two particles with known x, xp, y, yp (m, rad)."""
    MM = 1e-3
    pps = [(1 * MM, 2e-3, -4 * MM, 1e-3, 0, 0), (-3 * MM, -1e-3, 2 * MM, 0.5e-3, 0, 0),
           (2 * MM, 0.5e-3, 1 * MM, -3e-3, 0, 0), (0.5 * MM, 1e-3, -1 * MM, 2e-3, 0, 0)]
    got = O.a05_objectives(pps, 2)
    want = [3.0, 2.0, 4.0, 3.0,                 # max |x| mm, |xp| mrad, |y| mm, |yp| mrad
            abs(0.5 / 4), abs(2.5 / 4), abs(-2.0 / 4), abs(0.5 / 4),   # |mean|
            1.0,                                # | max|x| - max|y| |
            abs((4.0 / 2) - (1.5 / 2)),         # half widths of x per group: 2.0, 0.75
            abs((6.0 / 2) - (2.0 / 2))]         # half widths of y per group: 3.0, 1.0
    assert got == pytest.approx(want, rel=1e-12, abs=1e-12)
