"""CPU tests: C ABI exports, host planner, and host-side geometry against the oracle."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from tests._helpers import ROOT, ensure_built, rel

ensure_built()

from cctpy_amd import _native as N  # noqa: E402
from cctpy_amd.cctpy import (CCT, P2, P3, Beamline, HUST_SC_GANTRY, LocalCoordinateSystem,  # noqa: E402
                             ParticleFactory, PhaseSpaceParticle, Trajectory)
from oracle import cct_oracle as O  # noqa: E402
from tests.golden.params import SECTION1_DEMO  # noqa: E402


def test_library_exports_every_declared_symbol():
    header = open(f"{ROOT}/include/cctamd.h").read()
    declared = set(re.findall(r"\b(cctamd_[a-z0-9_]+)\s*\(", header))
    assert declared == set(N.EXPORTED)
    for name in declared:
        assert hasattr(N.lib, name)
    assert N.lib.cctamd_version() >= 100
    assert N.lib.cctamd_error_string(-22) == b"cctamd: invalid argument"


def _plan(ntile, npt, ppl, target, min_tiles=4):
    ntile = np.asarray(ntile, np.int64)
    npt = np.asarray(npt, np.int64)
    toff = np.concatenate([[0], np.cumsum(ntile)[:-1]]).astype(np.int64)
    poff = np.concatenate([[0], np.cumsum(npt)]).astype(np.int64)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    n = N.lib.cctamd_plan(len(ntile), p(ntile), p(npt), p(toff), p(poff), ppl, target, min_tiles, None, None, None)
    work = np.zeros((max(n, 1), 8), np.int32)
    wbase = np.zeros(len(ntile), np.int64)
    nch = np.zeros(len(ntile), np.int32)
    n2 = N.lib.cctamd_plan(len(ntile), p(ntile), p(npt), p(toff), p(poff), ppl, target, min_tiles, p(work),
                           p(wbase), p(nch))
    assert n2 == n
    return work[:n], wbase, nch, toff, poff


@pytest.mark.parametrize("ppl", [1, 2, 4, -2, -4])
def test_plan_covers_every_point_tile_pair_once(ppl):
    """Every (tile, point) pair of a design lies in exactly one work item, and every point
    block has at least one item (a design without tiles gets one item of no tiles per block:
    the fused RK4 stages update a block's particles at the end of its last item)."""
    ntile = [37, 0, 5, 1200, 3]
    npt = [300, 17, 0, 64, 1]
    work, wbase, nch, toff, poff = _plan(ntile, npt, ppl, target=200)
    pb = 64 * ppl if ppl > 0 else 64 // -ppl
    cover = {}
    blocks = set()
    for d in range(len(ntile)):
        nb = -(-npt[d] // pb)
        if npt[d] == 0:
            assert nch[d] == 0
            continue
        if ntile[d] == 0:
            assert nch[d] == 1
        for b in range(nb):
            for c in range(nch[d]):
                w = work[wbase[d] + b * nch[d] + c]
                blocks.add(int(w[1]))
        for b in range(nb):
            for c in range(nch[d]):
                w = work[wbase[d] + b * nch[d] + c]
                assert w[0] == d and w[1] == poff[d] + b * pb and w[2] == min(pb, npt[d] - b * pb)
                assert (w[5], w[6], w[7]) == (c, nch[d], b)
                for t in range(w[3], w[3] + w[4]):
                    for q in range(w[1], w[1] + w[2]):
                        cover[(t, q)] = cover.get((t, q), 0) + 1
    expect = {(toff[d] + t, poff[d] + q) for d in range(len(ntile)) for t in range(ntile[d]) for q in range(npt[d])}
    assert set(cover) == expect and all(v == 1 for v in cover.values())
    assert blocks == {poff[d] + b * pb for d in range(len(ntile)) for b in range(-(-npt[d] // pb))}



def test_plan_chunks_of_one_tile_range_share_an_xcd():
    """A design's chunk count is rounded up to a multiple of the 8 XCDs (1,000 tiles x 512
    points in 128-point blocks at 10 tiles per item: 100 -> 104 chunks), so that the items of one
    chunk (launched as wbase + b nchunk + c, dispatched round-robin) land on one XCD; for the
    split-wave codes the tiles per item are a multiple of the 16 tile streams."""
    work, wbase, nch, toff, poff = _plan([11520], [64], -2, target=600)  # 39 -> 48 tiles per item
    assert nch[0] == 240 and all(w[4] == 48 for w in work)
    work, wbase, nch, toff, poff = _plan([1000], [512], 2, target=400)  # 10 tiles per item: 100 -> 104
    assert nch[0] == 104
    for c in range(nch[0]):
        xcd = {(wbase[0] + b * nch[0] + c) % 8 for b in range(4)}
        assert len(xcd) == 1
    tiles = [w[4] for w in work[:nch[0]]]
    assert sum(tiles) == 1000 and min(tiles) > 0
def test_plan_rejects_bad_arguments():
    z = np.zeros(1, np.int64)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    assert N.lib.cctamd_plan(1, p(z), p(z), p(z), p(z), 3, 10, 4, None, None, None) == -22
    assert N.lib.cctamd_plan(0, p(z), p(z), p(z), p(z), 2, 10, 4, None, None, None) == -22
    neg = np.array([-1], np.int64)
    assert N.lib.cctamd_plan(1, p(neg), p(z), p(z), p(z), 2, 10, 4, None, None, None) == -22
    # split-wave blocks exist for precisions 33 and 64, not for the re-centred FP32 kernel (32)
    assert N.lib.cctamd_bs_partial(32, -2, p(z), p(z), p(z), p(z), 1, p(z), 1, p(z), None) == -22
    # the FP32 paths read FP64 rows for near tiles: a NULL seg_near is rejected before any launch
    assert N.lib.cctamd_bs_partial(33, 1, p(z), p(z), None, p(z), 1, p(z), 1, p(z), None) == -22


def test_adjoint_requires_fp32_tile_origins():
    """The backward kernel reads the FP32 tile origins at every forward precision: a NULL
    tile_org is rejected on the host (before any launch) also for precision 64."""
    buf = np.zeros(64)
    q = buf.ctypes.data_as(ctypes.c_void_p)
    args = [64, 2, q, None, None, q, q, q, q, 1, q, 1, q, q, q, 1, q, q, q, q, 4, 1, q, q, q, 1,
            q, q, q, q, q, q, q, q, q, q, None]
    assert N.lib.cctamd_track_rk4_adjoint(*args) == -22


def _frame_eq(lcs: LocalCoordinateSystem, fr: O.Frame):
    for a, b in ((lcs.location, fr.O), (lcs.XI, fr.XI), (lcs.YI, fr.YI), (lcs.ZI, fr.ZI)):
        assert a.to_list() == list(b)


def test_host_placement_matches_oracle_bitwise():
    ours = HUST_SC_GANTRY(**SECTION1_DEMO).create_second_bending_part_beamline()
    ref = O.hust_gantry_section("second", **SECTION1_DEMO)
    assert ours.get_length() == ref.length
    assert len(ours.magnets) == len(ref.magnets)
    for a, b in zip(ours.magnets, ref.magnets):
        _frame_eq(a.local_coordinate_system, b.frame)
    for s in np.linspace(0, ref.length, 37):
        assert ours.point_at(float(s)).to_list() == list(ref.traj.point_at(float(s)))
        assert ours.direct_at(float(s)).to_list() == list(ref.traj.direct_at(float(s)))


def test_cct_device_record_reproduces_phi():
    traj = Trajectory.set_start_point(P2.origin()).first_line(P2.x_direct(), 1).add_arc_line(1, True, 30)
    cct = CCT.create_cct_along(traj, 1, 1, 0.06, 30, [90, 30, 75.5], 30, 1000, P2(0, 0),
                               P2(30 * 2 * math.pi, -30 / 180 * math.pi))
    ref = O.create_cct_along(O.Trajectory((0, 0), (1, 0), 1).add_arc(1, True, 30), 1, 1, 0.06, 30, [90, 30, 75.5],
                             30, 1000, (0, 0), (30 * 2 * math.pi, -30 / 180 * math.pi))
    rec = cct.device_record()
    ksi = rec[15] + rec[16] * np.arange(0, 3601, 97, dtype=np.float64)
    phi = rec[17] * ksi + rec[18]
    for j in range(int(rec[20])):
        if not (int(rec[29]) >> j) & 1:
            phi = phi + rec[21 + j] * np.sin((j + 1) * ksi)
    assert np.max(np.abs(phi - ref.phi_of_ksi(ksi))) < 1e-15
    assert int(rec[29]) == 1  # the 90 degree term is skipped
    assert rec[19] == 1e-7 * 1000


def test_host_particles_match_oracle(golden):
    g = golden["factory"]
    ip = ParticleFactory.create_proton(P3.origin(), P3.x_direct(), 215)
    cs = ip.get_natural_coordinate_system(y_direction=P3.z_direct())
    pp = PhaseSpaceParticle(x=3.5e-3, xp=7.5e-3, y=0, yp=0, z=0, delta=0.1)
    p = ParticleFactory.create_from_phase_space_particle(ip, cs, pp)
    assert p.position.to_list() == g["phase_space_particle"]["position"]
    assert rel(p.velocity.to_list(), g["phase_space_particle"]["velocity"]) < 1e-15
    back = PhaseSpaceParticle.create_from_running_particle(ip, cs, p)
    assert np.allclose([back.x, back.xp, back.y, back.yp, back.z, back.delta], g["phase_space_back"], rtol=1e-14,
                       atol=1e-18)


def test_beamline_builder_errors():
    bl = Beamline.set_start_point().first_drift(length=1)
    with pytest.raises(ValueError):
        bl.append_agcct(1, [0.1, 0.2, 0.3, 0.4], [10], [[30], [30]], [[10], [5]], [1, 1])
    with pytest.raises(ValueError):
        bl.append_agcct(1, [0.1, 0.2, 0.3], [10], [[30], [30]], [[10], [5]], [1, 1])
    with pytest.raises(ValueError):
        CCT(LocalCoordinateSystem(), 0.1, 0.2, 30, [30], 10, 1, P2(), P2(1, 1))


def test_single_hip_runtime_in_process():
    """The library shares torch's HIP runtime (one libamdhip64 mapped), so torch buffers are valid for it."""
    assert len(N.loaded_runtimes()) == 1


def test_population_adapter_host_parts():
    """geatpy is optional: the problem description and the design builder work without it."""
    from cctpy_amd import optim
    assert len(optim.LB) == len(optim.UB) == len(optim.VAR_TYPES) == optim.DIM == 27
    assert all(lo <= hi for lo, hi in zip(optim.LB, optim.UB))
    bl = optim.create_gantry_beamline(part_per_winding=60)
    assert bl.get_length() > 7.0 and len(bl.magnets) > 10
    try:
        import geatpy  # noqa: F401
    except ImportError:
        with pytest.raises(ImportError):
            optim.geatpy_problem()


def test_integration_binding_matches_the_loader():
    """The ctypes binding shown in the integration guide declares the argument types the package's own
    loader declares for the same entry points."""
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    native = open(os.path.join(ROOT, "cctpy_amd", "_native.py")).read()
    found = re.findall(r"lib\.(cctamd_\w+)\.argtypes = (\[[^\]]*\])", doc)
    assert len(found) >= 4
    for name, args in found:
        m = re.search(r"lib\." + name + r"\.argtypes = (\[[^\]]*\])", native)
        assert m, name
        assert re.sub(r"\s", "", args) == re.sub(r"\s", "", m.group(1)), name


def test_autotune_pick_breaks_near_ties_toward_more_items():
    from cctpy_amd.autotune import pick
    cost = {(1, 4): 0.0797, (1, 8): 0.0804, (2, 8): 0.0830, (1, 16): 0.0850}
    assert pick(cost) == ((1, 8), 0.0804)            # 0.9 % slower: a tie, more items win
    assert pick(cost, tie=0.005) == ((1, 4), 0.0797)  # outside the tie band the fastest wins
    assert pick({(2, 4): 1.0}) == ((2, 4), 1.0)
    # equal items per CU: the faster shape
    assert pick({(1, 8): 0.0805, (2, 8): 0.0800}) == ((2, 8), 0.0800)


def test_block_shape_for_particle_counts():
    from cctpy_amd.engine import block_shape_for, points_per_block
    assert block_shape_for([8, 16], 64) == -4 and points_per_block(-4) == 16
    assert block_shape_for([32, 5], 33) == -2 and points_per_block(-2) == 32
    assert block_shape_for([33], 64) == 1
    assert block_shape_for([16], 32) == 1      # no split-wave kernel for the re-centred FP32 field path
    assert block_shape_for([65, 1], 64) == 2
    assert block_shape_for([], 64) == -4       # nothing to track: any valid code


def _nested_design(th):
    """C05-style section whose two CCT layers sit in a CombinedMagnet placed before the QS; the
    CCTs carry five tilt angles, th[1] moves the second and th[2] (if given) the fifth."""
    from cctpy_amd.magnets import CombinedMagnet
    tilts = [30.0, th[1], 90.0, 90.0, th[2] if len(th) > 2 else 60.0]
    bl = (Beamline.set_start_point(P2.origin()).first_drift(P2.x_direct(), 1.0)
          .append_qs(0.27, th[0], 0.0, 0.1).append_drift(1.0)
          .append_dipole_cct(1.0, 0.1, 0.12, 45.0, tilts, 60, 10000.0).append_drift(1.0))
    ccts = [m for m in bl.magnets if isinstance(m, CCT)]
    bl.magnets = [CombinedMagnet(*ccts)] + [m for m in bl.magnets if not isinstance(m, CCT)]
    return bl


def test_design_parameter_map_follows_the_device_order_and_guards_tilts():
    """DesignProblem maps parameters onto the magnets in the order the device holds them (CCTs
    nested in a CombinedMagnet included), and refuses a parameter that moves a tilt angle beyond
    the four the adjoint carries."""
    from cctpy_amd.adjoint import MAX_GRAD_TILTS, DesignProblem
    from cctpy_amd.magnets import device_parts
    prob = DesignProblem(_nested_design, lambda bl: [], 1.0, 0.01)
    parts = device_parts(_nested_design([5.0, 80.0]))
    assert [type(m).__name__ for m in parts.cct_objects] == ["CCT", "CCT"] and len(parts.qs_objects) == 1
    pmap, ccts = prob.parameter_map(np.array([5.0, 80.0]))
    (c0, q0), (c1, q1) = pmap
    assert c0 == [] and q0 == [(0, pytest.approx(1.0, rel=1e-9), 0.0)]
    assert q1 == [] and [i for i, _, _ in c1] == [0, 1]
    for i, dI, dc in c1:
        assert dI == 0.0 and dc[0] == 0.0 and dc[2] == 0.0 and dc[3] == 0.0
        # dc/dalpha = -(pi/180) / sin^2(alpha) / (2 sinh(eta)) for the second tilt
        a = ccts[i].tilt_angles[1]
        want = -(math.pi / 180.0) / math.sin(math.radians(a)) ** 2 / (2 * math.sinh(ccts[i].eta))
        assert dc[1] == pytest.approx(want * (1.0 if a == 80.0 else -1.0), rel=1e-6)
    assert MAX_GRAD_TILTS == 4
    with pytest.raises(ValueError, match="tilt angle 5"):
        prob.parameter_map(np.array([5.0, 80.0, 60.0]))


def test_adjoint_tracker_refuses_unknown_block_shapes_before_any_device_work():
    """Points-per-lane codes outside 1, 2, 4, -2, -4 (the shapes the backward kernel has) are
    refused on the host, in the constructor and in tune()."""
    from cctpy_amd.adjoint import AdjointTracker
    with pytest.raises(ValueError, match="not one of"):
        AdjointTracker([], ppl=8)
    tr = AdjointTracker.__new__(AdjointTracker)
    with pytest.raises(ValueError, match="not among"):
        tr.tune([1], None, None, None, None, candidates=(1, 3))


@pytest.mark.parametrize("ppl", [0, 3, 8, -1, -8])
def test_plan_refuses_unknown_block_shapes(ppl):
    """Points-per-lane codes other than 1, 2, 4, -2, -4 have no kernel: the planner returns
    CCTAMD_EINVAL = -22 (a host-only call)."""
    one = np.ones(1, np.int64)
    zero = np.zeros(1, np.int64)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    assert N.lib.cctamd_plan(1, p(one), p(one), p(zero), p(zero), ppl, 16, 1, None, None, None) == -22
