"""GPU_ACCELERATOR facade methods (final_code/packages/gpu_accelerator.py:1147-2152) against
the reference's printed results and the oracle."""
import numpy as np
import pytest

from tests._helpers import ensure_built, gpu_available, rel

pytestmark = pytest.mark.gpu

ensure_built()

from cctpy_amd.cctpy import (GPU_ACCELERATOR, HUST_SC_GANTRY, MM, P2, P3, QS, Beamline,  # noqa: E402
                             LocalCoordinateSystem, ParticleFactory, PhaseSpaceParticle)
from tests.golden.params import SECTION1_DEMO  # noqa: E402

if not gpu_available():
    pytest.skip("no GPU visible", allow_module_level=True)

GA64 = GPU_ACCELERATOR(GPU_ACCELERATOR.FLOAT64)
GA32 = GPU_ACCELERATOR(GPU_ACCELERATOR.FLOAT32)


def test_current_element_b_single_segment(golden):
    g = golden["single_segment"]
    p0, p1 = np.array(g["p0"]), np.array(g["p1"])
    kls, p0s = (p1 - p0), (p0 + p1) / 2
    b = GA64.current_element_B(kls, p0s, 1, P3(*g["point"]))
    assert rel(b.to_list(), g["field"]) < 1e-13
    b32 = GA32.current_element_B(np.tile(kls, 3), np.tile(p0s, 3), 1, P3(*g["point"]))
    assert rel(b32.to_list(), g["field"]) < 1e-6


def test_qs_methods(golden):
    g = golden["qs_c05"]
    fr = g["frame"]
    qs = QS(LocalCoordinateSystem(P3(*fr["location"]), P3(*fr["x"]), P3(*fr["z"])), g["length"], g["gradient"],
            g["second_gradient"], g["aperture"])
    p = P3(*g["point"])
    for ga in (GA64, GA32):
        assert ga.magnet_at_qs(qs.to_numpy_array(), p).to_list() == g["field"]
        assert ga.magnet_at_qs_date(qs.to_numpy_array(), p).to_list() == g["field"]
        assert ga.magnet_at_qs_dates([qs, qs], p).to_list() == [0.0, 0.0, 2 * g["field"][2]]


def test_beamline_field_methods(golden):
    g = golden["beamline_c05"]
    bl = (Beamline.set_start_point(P2.origin()).first_drift(P2.x_direct(), 1).append_qs(0.27, 5, 20, 100 * MM)
          .append_drift(1).append_dipole_cct(1, 100 * MM, 120 * MM, 45, [30], 60, 10000).append_drift(1))
    for m in (GA64.magnet_at_beamline_with_single_qs, GA64.magnet_at_beamline_with_multi_qs):
        assert rel(m(bl, P3(*g["point_cpu"])).to_list(), g["field_cpu"]) < 1e-12
        assert rel(m(bl, P3(*g["point_gpu64"])).to_list(), g["field_gpu64"]) < 1e-12
    assert rel(GA32.magnet_at_beamline_with_multi_qs(bl, P3(*g["point_cpu"])).to_list(), g["field_cpu"]) < 1e-6


def test_track_multi_particle_for_magnet(golden):
    g = golden["gantry_section2_215MeV_100mm"]
    bl = HUST_SC_GANTRY(**SECTION1_DEMO).create_second_bending_part_beamline()
    for method in (GA64.track_multi_particle_for_magnet_with_single_qs,
                   GA64.track_multi_particle_for_magnet_with_multi_qs):
        ps = [ParticleFactory.create_proton_along(bl, kinetic_MeV=215) for _ in range(3)]
        method(bl, ps, bl.get_length(), 100 * MM)  # mutates the particles
        for p in ps:
            assert np.max(np.abs(np.array(p.position.to_list()) - g["cpu"]["position"])) < 1e-9
            assert rel(p.velocity.to_list(), g["cpu"]["velocity"]) < 1e-9
            assert p.distance == pytest.approx(g["length"], abs=1e-12)
    p = ParticleFactory.create_proton_along(bl, kinetic_MeV=215)
    GA32.track_multi_particle_for_magnet_with_multi_qs(bl, [p], bl.get_length(), 100 * MM)
    assert np.max(np.abs(np.array(p.position.to_list()) - g["cpu"]["position"])) < 1e-6


def test_track_multi_particle_beamline_single_equals_multi(golden):
    g = golden["gantry_section1_215MeV_100mm"]
    bls = [HUST_SC_GANTRY(**SECTION1_DEMO).create_first_bending_part_beamline(),
           HUST_SC_GANTRY(**{**SECTION1_DEMO, "agcct12_current": -3000}).create_first_bending_part_beamline()]
    ps = [ParticleFactory.create_proton_along(bls[0], kinetic_MeV=e) for e in (215, 220)]
    a = GA64.track_multi_particle_beamline_for_magnet_with_single_qs(bls, ps, bls[0].get_length(), 100 * MM)
    b = GA64.track_multi_particle_beamline_for_magnet_with_multi_qs(bls, ps, bls[0].get_length(), 100 * MM)
    for d in range(2):
        for i in range(2):
            assert a[d][i].position.to_list() == b[d][i].position.to_list()
    assert np.max(np.abs(np.array(a[0][0].position.to_list()) - g["cpu"]["position"])) < 1e-9


def test_track_phase_ellipse_in_multi_beamline():
    bls = [HUST_SC_GANTRY(**SECTION1_DEMO).create_first_bending_part_beamline(),
           HUST_SC_GANTRY(**{**SECTION1_DEMO, "agcct12_current": -3000}).create_first_bending_part_beamline()]
    n, L = 6, bls[0].get_length()
    got = GA64.track_phase_ellipse_in_multi_beamline(bls, 3.5, 7.5, 3.5, 7.5, 0.02, n, 215, 0.0, L, 50 * MM,
                                                     report=False)
    assert len(got) == 2 and all(len(x) == n and len(y) == n for x, y in got)
    # the same through the public building blocks
    ip0 = ParticleFactory.create_proton_along(bls[0].trajectory, 0.0, 215)
    ip1 = ParticleFactory.create_proton_along(bls[0].trajectory, L, 215)
    cs0, cs1 = ip0.get_natural_coordinate_system(), ip1.get_natural_coordinate_system()
    pps = (PhaseSpaceParticle.phase_space_particles_along_positive_ellipse_in_xxp_plane(3.5 * MM, 7.5e-3, 0.02, n)
           + PhaseSpaceParticle.phase_space_particles_along_positive_ellipse_in_yyp_plane(3.5 * MM, 7.5e-3, 0.02, n))
    rps = ParticleFactory.create_from_phase_space_particles(ip0, cs0, pps)
    ends = GA64.track_multi_particle_beamline_for_magnet_with_multi_qs(bls, rps, L, 50 * MM)
    for d in range(2):
        for q in ends[d]:
            q.speed = q.velocity.length()
        back = PhaseSpaceParticle.create_from_running_particles(ip1, cs1, ends[d])
        for k in range(n):
            assert got[d][0][k].to_list() == pytest.approx([back[k].x / MM, back[k].xp / 1e-3], rel=1e-12, abs=1e-12)
            assert got[d][1][k].to_list() == pytest.approx([back[n + k].y / MM, back[n + k].yp / 1e-3], rel=1e-12,
                                                           abs=1e-12)
