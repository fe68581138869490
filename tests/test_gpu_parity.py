"""GPU parity: the HIP path (through the C ABI) against the oracle and the reference's golden vectors."""
import math

import numpy as np
import pytest

from tests._helpers import ensure_built, gpu_available, rel

pytestmark = pytest.mark.gpu

ensure_built()

from cctpy_amd import engine  # noqa: E402
from cctpy_amd.cctpy import (CCT, GPU_ACCELERATOR, HUST_SC_GANTRY, MM, P2, P3, QS, Beamline,  # noqa: E402
                             LocalCoordinateSystem, Magnet, ParticleFactory, ParticleRunner, Trajectory)
from cctpy_amd.magnets import device_parts  # noqa: E402
from cctpy_amd.particles import track_particles  # noqa: E402
from oracle import c_oracle as C  # noqa: E402
from oracle import cct_oracle as O  # noqa: E402
from tests.golden.params import SECTION1_DEMO  # noqa: E402

if not gpu_available():
    pytest.skip("no GPU visible", allow_module_level=True)


def _c05_pair():
    ours = (Beamline.set_start_point(P2.origin()).first_drift(P2.x_direct(), 1)
            .append_qs(0.27, 5, 20, 100 * MM).append_drift(1)
            .append_dipole_cct(1, 100 * MM, 120 * MM, 45, [30], 60, 10000).append_drift(1))
    ref = (O.Beamline((0, 0), (1, 0), 1).append_qs(0.27, 5, 20, 0.1).append_drift(1)
           .append_dipole_cct(1, 0.1, 0.12, 45, [30], 60, 10000).append_drift(1))
    return ours, ref


def test_winding_generator_matches_oracle():
    ours, ref = _c05_pair()
    model = engine.DeviceModel([device_parts(ours).spec()], fp32=False)
    seg = model.segments64()
    kl = np.concatenate([m.global_segments()[0] for m in ref.magnets if isinstance(m, O.CCTMagnet)])
    mid = np.concatenate([m.global_segments()[1] for m in ref.magnets if isinstance(m, O.CCTMagnet)])
    n = kl.shape[0]
    assert model.nseg == n and model.nseg_padded % 32 == 0
    assert np.max(np.abs(seg[:n, 4:7] - mid)) < 1e-13
    assert np.max(np.abs(seg[:n, 0:3] - kl)) < 1e-13 * np.max(np.abs(kl))
    assert np.all(seg[n:, 0:3] == 0.0)


def test_winding_points_match_oracle(golden):
    g = golden["cct_a14"]
    cct = CCT(LocalCoordinateSystem.global_coordinate_system(), 1.0, 0.03, 30, [30], 30, 1000, P2(0, 0),
              P2(2 * 30 * math.pi, 30 / 180 * math.pi))
    ref = O.CCTMagnet(O.GLOBAL_FRAME, 1.0, 0.03, 30, [30], 30, 1000, (0, 0), (2 * 30 * math.pi, 30 / 180 * math.pi))
    assert np.max(np.abs(cct.dispersed_path3 - ref.path3)) < 1e-14
    assert rel(cct.dispersed_path3[0], g["convert_origin"]) < 1e-15
    traj = Trajectory.set_start_point(P2.origin()).first_line(P2.x_direct(), 1).add_arc_line(1, True, 30).add_strait_line(1)
    c2 = CCT.create_cct_along(traj, 1, 1, 60 * MM, 30, [90, 30], 30, 1000, P2(0, 0),
                              P2(30 * 2 * math.pi, -30 / 180 * math.pi))
    assert c2.conductor_length(1, 3600) == pytest.approx(g["conductor_length_3600"], rel=1e-13)
    assert c2.conductor_length(1, 360) == pytest.approx(g["conductor_length_360"], rel=1e-13)


@pytest.mark.parametrize("ppl", [1, 2, 4])
def test_field_fp64_matches_golden_and_oracle(golden, ppl):
    g = golden["beamline_c05"]
    ours, ref = _c05_pair()
    model = engine.model_for([device_parts(ours).spec()])
    rng = np.random.default_rng(3)
    pts = [(*ref.traj.point_at(float(s)), float(z)) for s, z in zip(rng.uniform(0, ref.length, 200),
                                                                     rng.uniform(-0.05, 0.05, 200))]
    pts = np.array([g["point_cpu"], g["point_gpu64"]] + pts)
    b = model.field(pts, [len(pts)], precision=64, ppl=ppl)
    assert rel(b[0], g["field_cpu"]) < 1e-12
    assert rel(b[1], g["field_gpu64"]) < 1e-12
    want = C.field(O.beamline_arrays(ref.magnets), pts)
    err = np.linalg.norm(b - want, axis=1) / np.maximum(np.linalg.norm(want, axis=1), 1e-3)
    assert err.max() < 1e-9, err.max()


@pytest.mark.parametrize("target", [1, 64, 4000, 20000])
def test_field_fp64_any_chunk_count(target):
    """Plans from a few chunks per point block to more than 256 (several rounds of the
    fixed-order chunk sum, whose lanes load four chunks at a time, the last round partly
    empty): FP64 field at in-bore points of a 368,640-segment AG-CCT section vs the oracle."""
    import torch
    from cctpy_amd import workloads as W
    from oracle import workload_builder as WB
    wl = W.agcct_section(n_particles=8)
    bl = W.build_beamline(wl.designs[0])
    ref = WB.build_beamline(wl.designs[0])
    rng = np.random.default_rng(13)
    pts = []
    for s in rng.uniform(0.3, bl.get_length() - 0.3, 100):
        q, d = ref.traj.point_at(float(s)), ref.traj.direct_at(float(s))
        dx, dz = rng.uniform(-0.03, 0.03), rng.uniform(-0.03, 0.03)
        pts.append((q[0] - d[1] * dx, q[1] + d[0] * dx, dz))
    pts = np.array(pts)
    model = engine.model_for([device_parts(bl).spec()])
    plan = model.plan([len(pts)], 1, target)
    nc = int(plan.nchunk.max())
    assert nc <= 8 if target == 1 else (nc > 256 if target >= 4000 else 8 < nc <= 256), nc
    _, partial = model.workspace(plan)
    p = torch.tensor(pts.T.copy(), dtype=torch.float64, device=engine.device())
    out = torch.empty_like(p)
    model.field_tensors(plan, 64, p, partial, out)
    got = out.cpu().numpy().T
    want = C.field(O.beamline_arrays(ref.magnets), pts)
    err = np.linalg.norm(got - want, axis=1) / np.linalg.norm(want, axis=1)
    assert err.max() < 1e-12, (nc, err.max())


def test_field_fp32_recentred_accuracy():
    ours, ref = _c05_pair()
    model = engine.model_for([device_parts(ours).spec()])
    rng = np.random.default_rng(4)
    pts = np.array([(*ref.traj.point_at(float(s)), float(z)) for s, z in zip(rng.uniform(0, ref.length, 300),
                                                                              rng.uniform(-0.03, 0.03, 300))])
    b32 = model.field(pts, [len(pts)], precision=32)
    want = C.field(O.beamline_arrays(ref.magnets), pts)
    err = np.linalg.norm(b32 - want, axis=1) / np.maximum(np.linalg.norm(want, axis=1), 1e-3)
    assert err.max() < 1e-6, err.max()


def test_field_fp32_kc_in_bore_accuracy():
    """Precision 33 (tile sums K x q - C) inside the bore of a 720-segment/winding AG-CCT section."""
    from cctpy_amd import workloads as W
    from oracle import workload_builder as WB
    wl = W.agcct_section(n_particles=8)
    bl = W.build_beamline(wl.designs[0])
    ref = WB.build_beamline(wl.designs[0])
    rng = np.random.default_rng(5)
    pts = []
    for s in rng.uniform(0.3, bl.get_length() - 0.3, 200):
        p, d = ref.traj.point_at(float(s)), ref.traj.direct_at(float(s))
        dx, dz = rng.uniform(-0.03, 0.03), rng.uniform(-0.03, 0.03)
        pts.append((p[0] - d[1] * dx, p[1] + d[0] * dx, dz))
    pts = np.array(pts)
    model = engine.model_for([device_parts(bl).spec()])
    want = C.field(O.beamline_arrays(ref.magnets), pts)
    for prec in (32, 33):
        b = model.field(pts, [len(pts)], precision=prec)
        err = np.linalg.norm(b - want, axis=1) / np.linalg.norm(want, axis=1)
        assert err.max() < 1e-6, (prec, err.max())


@pytest.mark.parametrize("precision,ppl", [(33, 1), (33, 2), (33, -2), (33, -4), (32, 1)])
def test_field_fp32_next_to_windings(precision, ppl):
    """FP32 paths at points 2-6 mm from the conductors (a beam that leaves the bore): the tiles
    near a point are summed in FP64, so the field stays within 1e-6 of the oracle."""
    from cctpy_amd import workloads as W
    from oracle import workload_builder as WB
    wl = W.agcct_section(n_particles=8, disperse_number_per_winding=180)
    bl = W.build_beamline(wl.designs[0])
    ref = WB.build_beamline(wl.designs[0])
    ccts = [m for m in ref.magnets if isinstance(m, O.CCTMagnet)]
    rng = np.random.default_rng(9)
    pts = []
    for c in ccts:
        _, mids = c.global_segments()
        for i in rng.integers(0, len(mids), 12):
            d = rng.normal(size=3)
            pts.append(mids[i] + d / np.linalg.norm(d) * rng.uniform(2e-3, 6e-3))
    pts = np.array(pts)
    model = engine.model_for([device_parts(bl).spec()])
    want = C.field(O.beamline_arrays(ref.magnets), pts)
    b = model.field(pts, [len(pts)], precision=precision, ppl=ppl)
    err = np.linalg.norm(b - want, axis=1) / np.linalg.norm(want, axis=1)
    assert err.max() < 1e-6, err.max()


def _two_uneven_designs():
    """Two population designs with 45 and 13 in-bore points, and the oracle's field at them."""
    from cctpy_amd import workloads as W
    from oracle import workload_builder as WB
    wl = W.population(n_designs=2, n_particles=8, seed=3)
    rng = np.random.default_rng(6)
    pts_all, want_all, npt = [], [], []
    for d, n in zip(wl.designs, (45, 13)):
        ref = WB.build_beamline(d)
        pts = []
        for s in rng.uniform(0.3, ref.length - 0.3, n):
            p, t = ref.traj.point_at(float(s)), ref.traj.direct_at(float(s))
            dx, dz = rng.uniform(-0.02, 0.02), rng.uniform(-0.02, 0.02)
            pts.append((p[0] - t[1] * dx, p[1] + t[0] * dx, dz))
        pts = np.array(pts)
        pts_all.append(pts)
        want_all.append(C.field(O.beamline_arrays(ref.magnets), pts))
        npt.append(n)
    model = engine.model_for([device_parts(W.build_beamline(d)).spec() for d in wl.designs])
    assert model.design_nseg[0] != model.design_nseg[1]
    return model, np.concatenate(pts_all), npt, np.concatenate(want_all)


@pytest.mark.parametrize("ppl", [-2, -4])
def test_field_fp32_small_blocks_batched(ppl):
    """Split-wave kernel (32- or 16-point blocks) on two designs with uneven point and segment counts."""
    model, pts, npt, want = _two_uneven_designs()
    got = model.field(pts, npt, precision=33, ppl=ppl)
    err = np.linalg.norm(got - want, axis=1) / np.linalg.norm(want, axis=1)
    assert err.max() < 1e-6, err.max()
    ref2 = model.field(pts, npt, precision=33, ppl=2)
    assert np.max(np.linalg.norm(got - ref2, axis=1) / np.linalg.norm(want, axis=1)) < 1e-6
    with pytest.raises(Exception):
        model.field(pts, npt, precision=32, ppl=ppl)


@pytest.mark.parametrize("ppl", [-2, -4])
def test_field_fp64_small_blocks_batched(ppl):
    """FP64 split-wave kernel: same pair terms as the 64-point kernel, summed in another tile
    grouping, so it agrees with the oracle to 1e-12 and with the 64-point kernel to rounding."""
    model, pts, npt, want = _two_uneven_designs()
    got = model.field(pts, npt, precision=64, ppl=ppl)
    err = np.linalg.norm(got - want, axis=1) / np.linalg.norm(want, axis=1)
    assert err.max() < 1e-12, err.max()
    ref2 = model.field(pts, npt, precision=64, ppl=2)
    assert np.max(np.linalg.norm(got - ref2, axis=1) / np.linalg.norm(want, axis=1)) < 1e-13


def test_qs_fields_bitwise(golden):
    for key in ("qs_a13", "qs_c05"):
        g = golden[key]
        fr = g["frame"]
        qs = QS(LocalCoordinateSystem(P3(*fr["location"]), P3(*fr["x"]), P3(*fr["z"])), g["length"], g["gradient"],
                g["second_gradient"], g["aperture"])
        pts = g["points"] if "points" in g else [g["point"]]
        fs = g["fields"] if "fields" in g else [g["field"]]
        got = qs.magnetic_field_at_points([P3(*p) for p in pts])
        for a, b in zip(got, fs):
            assert a.to_list() == b


def test_rk4_uniform_field_bitwise(golden):
    g = golden["rk4_uniform"]
    m = Magnet.uniform_magnet(P3(z=g["bz"]))
    p = ParticleFactory.create_proton(P3.origin(), P3.x_direct(), 250)
    ParticleRunner.run_only(p, m, math.pi / 2)
    assert p.position.to_list() == g["quarter"]["position"]
    assert p.velocity.to_list() == g["quarter"]["velocity"]
    ParticleRunner.run_only(p, m, math.pi / 2)
    assert p.position.to_list() == g["half"]["position"] and p.velocity.to_list() == g["half"]["velocity"]
    info = ParticleRunner.run_get_all_info(ParticleFactory.create_proton(P3.origin(), P3.x_direct(), 250), m, 1, 0.1)
    assert len(info) == len(g["all_info_1m_100mm"])
    for q, (wp, wv) in zip(info, g["all_info_1m_100mm"]):
        assert q.position.to_list() == wp and q.velocity.to_list() == wv


@pytest.mark.parametrize("kin,key", [(215, "cpu"), (220, "cpu_220MeV")])
def test_gantry_section1_tracking_fp64(golden, kin, key):
    g = golden["gantry_section1_215MeV_100mm"]
    bl = HUST_SC_GANTRY(**SECTION1_DEMO).create_first_bending_part_beamline()
    assert bl.get_length() == g["length"]
    p = ParticleFactory.create_proton_along(bl, kinetic_MeV=kin)
    GPU_ACCELERATOR(GPU_ACCELERATOR.FLOAT64).track_one_particle_with_multi_qs(bl, p, bl.get_length(), 100 * MM)
    assert np.max(np.abs(np.array(p.position.to_list()) - g[key]["position"])) < 1e-9
    assert rel(p.velocity.to_list(), g[key]["velocity"]) < 1e-9
    assert p.distance == g["length"]


def test_gantry_section2_tracking_fp64_and_fp32(golden):
    g = golden["gantry_section2_215MeV_100mm"]
    bl = HUST_SC_GANTRY(**SECTION1_DEMO).create_second_bending_part_beamline()
    p64 = ParticleFactory.create_proton_along(bl, kinetic_MeV=215)
    p32 = p64.copy()
    GPU_ACCELERATOR(GPU_ACCELERATOR.FLOAT64).track_one_particle_with_single_qs(bl, p64, bl.get_length(), 100 * MM)
    GPU_ACCELERATOR(GPU_ACCELERATOR.FLOAT32).track_one_particle_with_single_qs(bl, p32, bl.get_length(), 100 * MM)
    assert np.max(np.abs(np.array(p64.position.to_list()) - g["cpu"]["position"])) < 1e-9
    assert rel(p64.velocity.to_list(), g["cpu"]["velocity"]) < 1e-9
    # FP32 Biot-Savart with FP64 state: <= 1 um after 7.1 m (the reference's own FP32 path: 0.77 um)
    assert (p32.position - p64.position).length() < 1e-6


def test_batched_designs_match_golden(golden):
    g = golden["gantry_section1_215MeV_100mm"]
    bl1 = HUST_SC_GANTRY(**SECTION1_DEMO).create_first_bending_part_beamline()
    bl2 = HUST_SC_GANTRY(**{**SECTION1_DEMO, "agcct12_current": -3000}).create_first_bending_part_beamline()
    p1 = ParticleFactory.create_proton_along(bl1, kinetic_MeV=215)
    p2 = ParticleFactory.create_proton_along(bl1, kinetic_MeV=220)
    ga = GPU_ACCELERATOR(GPU_ACCELERATOR.FLOAT64, block_dim_x=512)
    out = ga.track_multi_particle_beamline_for_magnet_with_multi_qs([bl1, bl2], [p1, p2], bl1.get_length(), 100 * MM)
    keys = [["cpu", "cpu_220MeV"], ["cpu_agcct12_m3000_215MeV", "cpu_agcct12_m3000_220MeV"]]
    for d in range(2):
        for i in range(2):
            w = g[keys[d][i]]
            assert np.max(np.abs(np.array(out[d][i].position.to_list()) - w["position"])) < 1e-9
            assert rel(out[d][i].velocity.to_list(), w["velocity"]) < 1e-9
    assert p1.distance == 0.0  # inputs untouched


@pytest.mark.parametrize("ppl", [1, 2, 4, -2, -4])
def test_tracking_beam_fp64_vs_c_oracle(ppl):
    """70 protons (not a multiple of 64), 1 mm steps through a 720-segment/winding AG-CCT section slice."""
    bl = HUST_SC_GANTRY(part_per_winding=720).create_second_bending_part_beamline()
    ref = O.hust_gantry_section("second", part_per_winding=720)
    rng = np.random.default_rng(11)
    ip = ParticleFactory.create_proton_along(bl, 2.0, 250)
    cs = ip.get_natural_coordinate_system()
    oip = O.create_proton_along(ref.traj, 2.0, 250)
    ofr = oip.natural_frame()
    ps, ops = [], []
    for _ in range(70):
        x, xp, y, yp = rng.normal(0, [1.75e-3, 2e-3, 1.75e-3, 2e-3])
        d = rng.uniform(-0.02, 0.02)
        from cctpy_amd.particles import PhaseSpaceParticle
        ps.append(ParticleFactory.create_from_phase_space_particle(ip, cs, PhaseSpaceParticle(x, xp, y, yp, 0, d)))
        ops.append(O.from_phase_space(oip, ofr, x, xp, y, yp, 0, d))
    length, step = 0.03, 1e-3
    out, _, ns, dts = track_particles(ps, [bl], [0] * len(ps), length, step, ppl=ppl)
    st = np.array([list(q.position) + list(q.velocity) for q in ops])
    want = C.track(O.beamline_arrays(ref.magnets), st, np.array([q.e / q.rm for q in ops]), dts, ns)
    assert np.max(np.abs(out[:, :3] - want[:, :3])) < 1e-12
    assert np.max(np.abs(out[:, 3:] - want[:, 3:])) / 1.8e8 < 1e-12


@pytest.mark.parametrize("ppl", [-2, -4])
def test_tracking_beam_fp32_small_blocks_vs_fp64(ppl):
    """FP32 tracking with the split-wave kernels (32- and 16-point blocks, 70 protons: ragged last
    block) against the FP64 path over 30 steps of 1 mm through an AG-CCT section slice."""
    from cctpy_amd.particles import PhaseSpaceParticle
    bl = HUST_SC_GANTRY(part_per_winding=720).create_second_bending_part_beamline()
    rng = np.random.default_rng(17)
    ip = ParticleFactory.create_proton_along(bl, 2.0, 250)
    cs = ip.get_natural_coordinate_system()
    ps = []
    for _ in range(70):
        x, xp, y, yp = rng.normal(0, [1.75e-3, 2e-3, 1.75e-3, 2e-3])
        ps.append(ParticleFactory.create_from_phase_space_particle(ip, cs, PhaseSpaceParticle(x, xp, y, yp, 0,
                                                                                           rng.uniform(-0.02, 0.02))))
    out64, _, _, _ = track_particles(ps, [bl], [0] * len(ps), 0.03, 1e-3, precision=64, ppl=2)
    out32, _, _, _ = track_particles(ps, [bl], [0] * len(ps), 0.03, 1e-3, precision=33, ppl=ppl)
    assert np.max(np.abs(out32[:, :3] - out64[:, :3])) < 1e-9
    assert np.max(np.abs(out32[:, 3:] - out64[:, 3:])) / 1.8e8 < 1e-9


def test_edge_cases():
    bl = HUST_SC_GANTRY(**SECTION1_DEMO).create_first_bending_part_beamline()
    assert ParticleRunner.run_only([], bl, 1.0) == []
    # a design made of QS only (no segments) and a zero-current CCT
    qs_only = Beamline.set_start_point().first_drift(length=0.5).append_qs(0.3, 10, 0, 0.05).append_drift(0.5)
    zero = Beamline.set_start_point().first_drift(length=0.1).append_dipole_cct(0.95, 0.08, 0.09, 30, [30], 10, 0.0)
    p = ParticleFactory.create_proton_along(qs_only, 0.0, 250)
    q = p.copy()
    ParticleRunner.run_only(p, qs_only, 1.3, 0.01)
    oq = O.Beamline((0, 0), (1, 0), 0.5).append_qs(0.3, 10, 0, 0.05).append_drift(0.5)
    op = O.create_proton_along(oq.traj, 0.0, 250)
    O.run_only(op, oq.field, 1.3, 0.01)
    assert rel(p.position.to_list(), op.position) < 1e-13
    b = zero.magnetic_field_at(P3(0.5, 0.0, 0.0))
    assert b.to_list() == [0.0, 0.0, 0.0]
    # the reference GPU facade refuses magnets it does not know (gpu_accelerator.py:1970)
    with pytest.raises(ValueError):
        class Foreign(Magnet):
            def magnetic_field_at(self, point):
                return P3()
        GPU_ACCELERATOR().track_one_particle_with_multi_qs(Foreign(), q, 0.1, 0.01)


def test_run_only_through_a_user_magnet():
    """ParticleRunner.run_only tracks through any Magnet subclass (the reference's callback
    contract, particles.py:514-537): a beamline whose uniform field comes from a user class
    tracks like the same beamline with the device's own UniformMagnet, and a user-magnet-only
    design matches the oracle's RK4 in that field."""
    from cctpy_amd.magnets import UniformMagnet

    class UserUniform(Magnet):
        def __init__(self, b):
            self.b = b

        def magnetic_field_at(self, point):
            return self.b.copy()

    b0 = P3(0.0, 0.0, 0.05)
    rng = np.random.default_rng(19)

    def section(extra):
        bl = HUST_SC_GANTRY(part_per_winding=60).create_second_bending_part_beamline()
        bl.magnets.append(extra)
        return bl
    bl_user, bl_dev = section(UserUniform(b0)), section(UniformMagnet(b0))
    ip = ParticleFactory.create_proton_along(bl_user, 2.0, 250)
    cs = ip.get_natural_coordinate_system()
    from cctpy_amd.particles import PhaseSpaceParticle
    ps = [ParticleFactory.create_from_phase_space_particle(
        ip, cs, PhaseSpaceParticle(*rng.normal(0, [1e-3, 1e-3, 1e-3, 1e-3]), 0, rng.uniform(-0.02, 0.02)))
        for _ in range(5)]
    got, rec, ns, _ = track_particles(ps, [bl_user], [0] * 5, 0.05, 2e-3, record=True)
    want, rec2, _, _ = track_particles(ps, [bl_dev], [0] * 5, 0.05, 2e-3, record=True)
    assert np.max(np.abs(got[:, :3] - want[:, :3])) < 1e-13
    for i in range(5):  # rows past a particle's own step count carry nothing
        assert np.max(np.abs(rec[:ns[i] + 1, i, :3] - rec2[:ns[i] + 1, i, :3])) < 1e-13
    # a design with only a user magnet: the oracle's RK4 in the same uniform field
    q = ParticleFactory.create_proton(P3(), P3(1, 0, 0), 250)
    q0 = q.copy()
    ParticleRunner.run_only(q, UserUniform(P3(0, 0, 2.4321282996)), np.pi / 2, 20 * MM)
    op = O.create_proton((0, 0, 0), (1, 0, 0), 250)
    O.run_only(op, lambda p: (0.0, 0.0, 2.4321282996), np.pi / 2, 20 * MM)
    assert rel(q.position.to_list(), op.position) < 1e-15 and rel(q.velocity.to_list(), op.velocity) < 1e-15
    assert q.distance == q0.distance + np.pi / 2
