"""cctamd_bs_field_noalloc (partial sums in the library's static device array) against
cctamd_bs_field (caller-allocated partials): the same kernels, so the fields agree bit for
bit, on two calls in a row whose second has fewer points (the array keeps the first
call's partials)."""
import numpy as np
import pytest
import torch

from tests._helpers import ensure_built, gpu_available

pytestmark = pytest.mark.gpu

ensure_built()

from cctpy_amd import _native as N  # noqa: E402
from cctpy_amd import engine, workloads as W  # noqa: E402
from cctpy_amd.engine import _p, _stream  # noqa: E402
from cctpy_amd.magnets import device_parts  # noqa: E402

if not gpu_available():
    pytest.skip("no GPU visible", allow_module_level=True)


@pytest.fixture(scope="module")
def section():
    wl = W.agcct_section(n_particles=8, disperse_number_per_winding=180)
    bl = W.build_beamline(wl.designs[0])
    model = engine.DeviceModel([device_parts(bl).spec()], fp32=True)
    rng = np.random.default_rng(21)
    pts = []
    for s in rng.uniform(0.1, bl.get_length() - 0.1, 1000):
        q = bl.point_at(float(s))
        pts.append((q.x + rng.uniform(-0.09, 0.09), q.y + rng.uniform(-0.09, 0.09), rng.uniform(-0.09, 0.09)))
    return model, np.array(pts)


def _noalloc(model, plan, precision, p, out):
    seg, org, near = model._seg(precision)
    return N.lib.cctamd_bs_field_noalloc(precision, plan.ppl, _p(seg), _p(org), _p(near), _p(plan.work), plan.nwork,
                                         _p(plan.wbase), _p(plan.nchunk), _p(plan.design_pt_off), model.ndesign,
                                         _p(model.qs), _p(model.design_qs_off), _p(model.bconst), _p(p), plan.npt,
                                         _p(out), _stream())


@pytest.mark.parametrize("precision,ppl", [(64, 1), (33, 1), (33, 2), (33, -2), (32, 1)])
def test_noalloc_field_matches_bs_field(section, precision, ppl):
    model, pts = section
    dev = engine.device()
    for n in (1000, 137):  # second call smaller: the static array still holds the first call's partials
        plan = model.plan([n], ppl)
        assert plan.nwork * 3 * engine.points_per_block(ppl) <= N.NOALLOC_PARTIAL_DOUBLES
        p = torch.from_numpy(np.ascontiguousarray(pts[:n].T)).to(dev)
        want = torch.empty_like(p)
        _, partial = model.workspace(plan)
        model.field_tensors(plan, precision, p, partial, want)
        got = torch.full_like(p, float("nan"))
        N.check(_noalloc(model, plan, precision, p, got), "cctamd_bs_field_noalloc")
        torch.cuda.synchronize()
        assert torch.equal(got, want), (n, float((got - want).abs().max()))
        assert bool(torch.isfinite(got).all())


def test_noalloc_rejects_plans_beyond_its_array(section):
    """A plan whose partials exceed the static array is refused on the host (no launch)."""
    model, pts = section
    plan = model.plan([256], 4)
    too_many = N.NOALLOC_PARTIAL_DOUBLES // (3 * 256) + 1
    p = torch.zeros((3, 256), dtype=torch.float64, device=engine.device())
    out = torch.empty_like(p)
    seg, org, near = model._seg(33)
    rc = N.lib.cctamd_bs_field_noalloc(33, 4, _p(seg), _p(org), _p(near), _p(plan.work), too_many, _p(plan.wbase),
                                       _p(plan.nchunk), _p(plan.design_pt_off), model.ndesign, _p(model.qs),
                                       _p(model.design_qs_off), _p(model.bconst), _p(p), 256, _p(out), _stream())
    assert rc == -34
    assert "size out of range" in N.lib.cctamd_error_string(rc).decode()
