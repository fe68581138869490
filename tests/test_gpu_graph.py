"""A tracking pass captured as a HIP graph gives the eager result on every replay."""
import numpy as np
import pytest
import torch

from tests._helpers import ensure_built, gpu_available

pytestmark = pytest.mark.gpu

ensure_built()

from cctpy_amd import engine, workloads as W  # noqa: E402
from cctpy_amd.magnets import device_parts  # noqa: E402

if not gpu_available():
    pytest.skip("no GPU visible", allow_module_level=True)


@pytest.mark.parametrize("precision,ppl", [(64, 2), (64, -2), (33, 2), (33, -2)])
def test_graph_replays_match_eager(precision, ppl):
    wl = W.population(n_designs=3, n_particles=16, seed=11)
    bls = [W.build_beamline(d) for d in wl.designs]
    ps, npt, lengths = [], [], []
    for i, (bl, beam) in enumerate(zip(bls, wl.beams)):
        q = W.beam_particles(beam, bl)
        ps += q
        npt.append(len(q))
        lengths += [0.05] * len(q)
    state = np.array([p.position.to_list() + p.velocity.to_list() for p in ps])
    sd = [engine.rk4_steps(L, 1e-3, p.speed) for L, p in zip(lengths, ps)]
    dev = engine.device()
    ns = torch.tensor([s for s, _ in sd], dtype=torch.int64, device=dev)
    dt = torch.tensor([d for _, d in sd], dtype=torch.float64, device=dev)
    kq = torch.tensor([p.e / p.relativistic_mass for p in ps], dtype=torch.float64, device=dev)
    model = engine.DeviceModel([device_parts(bl).spec() for bl in bls], fp32=True)
    plan = model.plan(npt, ppl)
    scratch, partial = model.workspace(plan)
    s0 = torch.from_numpy(np.ascontiguousarray(state.T)).to(dev)
    max_steps = int(ns.max())

    st = s0.clone()
    model.track_tensors(plan, precision, st, kq, dt, ns, max_steps, scratch, partial)
    eager = st.cpu().numpy()
    assert np.all(np.isfinite(eager)) and not np.array_equal(eager, state.T)

    g_state = torch.empty_like(s0)
    g_state.copy_(s0)
    graph = engine.TrackGraph(model, plan, precision, g_state, kq, dt, ns, max_steps, scratch, partial)
    for rep in range(3):
        g_state.copy_(s0)
        graph.replay()
        torch.cuda.synchronize()
        assert np.array_equal(g_state.cpu().numpy(), eager), rep
        # eager work on the same buffers between replays
        st.copy_(s0)
        model.track_tensors(plan, precision, st, kq, dt, ns, max_steps, scratch, partial)
        torch.cuda.synchronize()
        assert np.array_equal(st.cpu().numpy(), eager), rep
