"""RK4 stages fused into the Biot-Savart launch against the two-launch form.

With CCTAMD_FUSED_STAGES=1 and points-per-lane codes 1, 2, 4 each RK4 stage runs at the end of
the Biot-Savart launch: the last work item of a point block to finish sums the block's chunk
partials and updates its particles.  By default a separate stage kernel does.  Both must give
the same bits: final states, recorded states, the adjoint's tape and gradients, for ragged point
blocks, unequal step counts, several designs and a design without CCT tiles (one empty work
item per block)."""
import os

import numpy as np
import pytest
import torch

from tests._helpers import ensure_built, gpu_available

pytestmark = pytest.mark.gpu

ensure_built()

from cctpy_amd import adjoint, engine, workloads as W  # noqa: E402
from cctpy_amd.cctpy import MM, P2, Beamline, ParticleFactory  # noqa: E402
from cctpy_amd.magnets import device_parts  # noqa: E402

if not gpu_available():
    pytest.skip("no GPU visible", allow_module_level=True)


class _fused:
    """CCTAMD_FUSED_STAGES=1 for the duration of a with-block (read at every tracking call)."""

    def __enter__(self):
        self.old = os.environ.get("CCTAMD_FUSED_STAGES")
        os.environ["CCTAMD_FUSED_STAGES"] = "1"

    def __exit__(self, *exc):
        if self.old is None:
            del os.environ["CCTAMD_FUSED_STAGES"]
        else:
            os.environ["CCTAMD_FUSED_STAGES"] = self.old


def _qs_only_beamline():
    return (Beamline.set_start_point(P2.origin()).first_drift(P2.x_direct(), 0.5)
            .append_qs(0.27, 5, 20, 100 * MM).append_drift(1.0))


def _case(n_cct_designs=2, n_particles=70):
    wl = W.population(n_designs=n_cct_designs, n_particles=n_particles, seed=21)
    bls = [W.build_beamline(d) for d in wl.designs]
    ps, npt, lengths = [], [], []
    for i, (bl, beam) in enumerate(zip(bls, wl.beams)):
        q = W.beam_particles(beam, bl)
        ps += q
        npt.append(len(q))
        lengths += [0.04 + 0.01 * i + 1e-4 * j for j in range(len(q))]  # unequal step counts
    qbl = _qs_only_beamline()
    ip = ParticleFactory.create_proton_along(qbl, 0.4, 250)
    qps = [ip.copy() for _ in range(5)]
    for j, p in enumerate(qps):
        p.position = p.position + P2(0, 1e-3 * (j - 2)).to_p3()
    bls.append(qbl)
    ps += qps
    npt.append(len(qps))
    lengths += [0.3] * len(qps)
    state = np.array([p.position.to_list() + p.velocity.to_list() for p in ps])
    sd = [engine.rk4_steps(L, 1e-3, p.speed) for L, p in zip(lengths, ps)]
    ns = np.array([s for s, _ in sd], np.int64)
    dt = np.array([d for _, d in sd])
    kq = np.array([p.e / p.relativistic_mass for p in ps])
    return bls, npt, state, kq, dt, ns


@pytest.mark.parametrize("precision,ppl", [(64, 1), (64, 2), (33, 1), (33, 2), (33, 4), (32, 1)])
def test_fused_stages_give_the_two_launch_bits(precision, ppl):
    bls, npt, state, kq, dt, ns = _case()
    dev = engine.device()
    model = engine.DeviceModel([device_parts(bl).spec() for bl in bls], fp32=True)
    plan = model.plan(npt, ppl, 300)
    scratch, partial = model.workspace(plan)
    s0 = torch.from_numpy(np.ascontiguousarray(state.T)).to(dev)
    kq_t = torch.from_numpy(kq).to(dev)
    dt_t = torch.from_numpy(dt).to(dev)
    ns_t = torch.from_numpy(ns).to(dev)
    max_steps = int(ns.max())
    every = 7
    out = []
    for fused in (True, False, True):
        st = s0.clone()
        rec = torch.full((max_steps // every + 1, 6, s0.shape[1]), float("nan"), dtype=torch.float64, device=dev)
        if fused:
            with _fused():
                model.track_tensors(plan, precision, st, kq_t, dt_t, ns_t, max_steps, scratch, partial, record=rec,
                                    record_every=every)
        else:
            model.track_tensors(plan, precision, st, kq_t, dt_t, ns_t, max_steps, scratch, partial, record=rec,
                                record_every=every)
        torch.cuda.synchronize()
        out.append((st.cpu().numpy(), rec.cpu().numpy()))
    final = out[0][0]
    assert np.all(np.isfinite(final)) and not np.array_equal(final, state.T)
    # the QS-only design's particles moved too (its block has one item of no tiles)
    assert not np.array_equal(final[:, -5:], state.T[:, -5:])
    for f, r in out[1:]:
        assert np.array_equal(f, final)
        assert np.array_equal(np.isnan(r), np.isnan(out[0][1]))
        assert np.array_equal(np.nan_to_num(r), np.nan_to_num(out[0][1]))


@pytest.mark.parametrize("ckpt_every", [0, 16])
def test_fused_stages_keep_the_adjoint_bits(ckpt_every):
    """The taped forward and the checkpoint recompute of the reverse pass run the fused stages
    too: gradients equal those of the two-launch form bit for bit."""
    bls, npt, state, kq, dt, ns = _case(n_particles=40)
    w = np.random.default_rng(5).normal(size=(len(state), 6)) * np.array([1.0, 1.0, 1.0, 1e-8, 1e-8, 1e-8])
    res = []
    for fused in (True, False):
        tr = adjoint.AdjointTracker(bls, precision=33, ckpt_every=ckpt_every, ppl=1)
        if fused:
            with _fused():
                final = tr.forward(state, kq, dt, ns, npt)
                a0, cg, qg = tr.backward(w)
        else:
            final = tr.forward(state, kq, dt, ns, npt)
            a0, cg, qg = tr.backward(w)
        res.append((final, a0, cg, qg))
    for x, y in zip(res[0], res[1]):
        assert np.array_equal(x, y)
    assert np.all(np.isfinite(res[0][1])) and np.any(res[0][2] != 0.0)


def test_fused_stages_replay_in_a_graph():
    """A tracking pass with fused stages captured as a HIP graph (the arrival counters are zeroed
    by a kernel inside the graph and reset by each last arrival) gives the eager result on every
    replay, with eager work on the same buffers between replays."""
    bls, npt, state, kq, dt, ns = _case(n_particles=40)
    dev = engine.device()
    model = engine.DeviceModel([device_parts(bl).spec() for bl in bls], fp32=True)
    plan = model.plan(npt, 1, 300)
    scratch, partial = model.workspace(plan)
    s0 = torch.from_numpy(np.ascontiguousarray(state.T)).to(dev)
    kq_t, dt_t, ns_t = (torch.from_numpy(a).to(dev) for a in (kq, dt, ns))
    max_steps = int(ns.max())
    with _fused():
        st = s0.clone()
        model.track_tensors(plan, 33, st, kq_t, dt_t, ns_t, max_steps, scratch, partial)
        torch.cuda.synchronize()
        eager = st.cpu().numpy()
        g_state = s0.clone()
        graph = engine.TrackGraph(model, plan, 33, g_state, kq_t, dt_t, ns_t, max_steps, scratch, partial)
        for rep in range(3):
            g_state.copy_(s0)
            graph.replay()
            torch.cuda.synchronize()
            assert np.array_equal(g_state.cpu().numpy(), eager), rep
            st.copy_(s0)
            model.track_tensors(plan, 33, st, kq_t, dt_t, ns_t, max_steps, scratch, partial)
            torch.cuda.synchronize()
            assert np.array_equal(st.cpu().numpy(), eager), rep
    # and the two-launch form gives the same bits
    st.copy_(s0)
    model.track_tensors(plan, 33, st, kq_t, dt_t, ns_t, max_steps, scratch, partial)
    torch.cuda.synchronize()
    assert np.array_equal(st.cpu().numpy(), eager)
