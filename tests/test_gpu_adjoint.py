"""GPU adjoint: gradients of tracked states and of the beam-spot objective against central
finite differences of the FP64 forward path."""
import math

import numpy as np
import pytest
import torch

from tests._helpers import ensure_built, gpu_available

pytestmark = pytest.mark.gpu

ensure_built()

from cctpy_amd import adjoint, workloads as W  # noqa: E402
from cctpy_amd.cctpy import ParticleFactory  # noqa: E402

if not gpu_available():
    pytest.skip("no GPU visible", allow_module_level=True)


def _problem(n_particles, n, length, step, precision, ckpt):
    beam = W.gaussian_beam(np.random.default_rng(17), n_particles, 0.02)

    def particles(bl):
        return W.beam_particles(W.Beam(beam, 0.4, 250.0), bl)

    prob = adjoint.DesignProblem(lambda th: W.build_beamline(W.section_ops(th, n, 0.0)), particles, length, step,
                                 precision=precision, ckpt_every=ckpt)
    return prob, beam


def per_parameter_error(grad, fd, floor_frac=1e-3):
    """|g_k - fd_k| / max(|fd_k|, floor) per parameter, floor = floor_frac * max |fd|: a small
    component is held to its own size (down to 1/1000 of the largest), not to the largest one's.
    The parameters mix T/m, degrees and amperes, so a single max-normalised figure would hide a
    wrong small component."""
    grad, fd = np.asarray(grad), np.asarray(fd)
    floor = floor_frac * np.abs(fd).max()
    return np.abs(grad - fd) / np.maximum(np.abs(fd), floor)


def _fd(prob, theta, J, rel_h=1e-6):
    g = np.zeros_like(theta)
    for k in range(theta.size):
        h = rel_h * max(abs(theta[k]), 1.0)
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        g[k] = (J(prob.forward(tp)) - J(prob.forward(tm))) / (2 * h)
    return g


def test_adjoint_linear_functional_vs_fd():
    """J = w . final states (a linear functional): the adjoint's dJ/dtheta against central FD."""
    prob, _ = _problem(24, 90, 1.2, 5e-3, 64, 16)
    theta = W.section_theta0()
    final = prob.forward(theta)
    rng = np.random.default_rng(3)
    w = rng.normal(size=final.shape) * np.array([1.0, 1.0, 1.0, 1e-8, 1e-8, 1e-8])
    grad = prob.backward(w)
    fd = _fd(prob, theta, lambda f: float(np.sum(w * f)))
    err = per_parameter_error(grad, fd)
    print("per-parameter relative error:", np.array2string(err, precision=2))
    assert err.max() < 1e-4, (grad, fd, err)


def test_adjoint_beam_spot_objective_torch():
    """Config-4 style objective through torch autograd, 1.1 m drift, checkpoints every 50 steps."""
    beam = W.gaussian_beam(np.random.default_rng(5), 32, 0.08)
    n = 90

    def particles(bl):
        return W.beam_particles(W.Beam(beam, 0.0, 250.0), bl)

    bl0 = W.build_beamline(W.section_ops(W.section_theta0(), n))
    L = bl0.get_length()
    ip = ParticleFactory.create_proton_along(bl0, L, 250.0)
    fr = ip.get_natural_coordinate_system()
    prob = adjoint.DesignProblem(lambda th: W.build_beamline(W.section_ops(th, n)), particles, L, 1e-2,
                                 precision=64, ckpt_every=50)

    def J_np(final):
        return float(W.beam_spot_objective(torch.from_numpy(final), ip.position.to_list(), fr.XI.to_list(),
                                           fr.YI.to_list(), beam[:, 4]))

    theta = torch.tensor(W.section_theta0(), requires_grad=True)
    final = adjoint.track_fn(theta, prob)
    J = W.beam_spot_objective(final, ip.position.to_list(), fr.XI.to_list(), fr.YI.to_list(), beam[:, 4])
    J.backward()
    fd = _fd(prob, theta.detach().numpy().copy(), J_np)
    g = theta.grad.numpy()
    err = per_parameter_error(g, fd)
    print("per-parameter relative error:", np.array2string(err, precision=2))
    assert err.max() < 1e-4, (g, fd, err)


@pytest.mark.parametrize("ppl", [1, -2])
def test_adjoint_population_current_gradients_vs_fd(ppl):
    """Two designs of different arc lengths (particles finish at different steps) whose
    windings do not fill their last tile: every gradient is finite and dJ/d(current) of each
    CCT matches central differences of the FP64 forward path (FP64 forward with 64- or
    32-point blocks)."""
    import copy
    from cctpy_amd import engine
    from cctpy_amd.magnets import device_parts
    wl = W.population(n_designs=2, n_particles=8, seed=6)
    bls = [W.build_beamline(d) for d in wl.designs]
    specs = [device_parts(bl).spec() for bl in bls]
    assert any(c % 32 for s in specs for c in s.cct_counts)
    ps, npt, lengths = [], [], []
    for i, (bl, beam) in enumerate(zip(bls, wl.beams)):
        q = W.beam_particles(beam, bl)
        ps += q
        npt.append(len(q))
        lengths += [wl.beam_length(i)] * len(q)
    state = np.array([p.position.to_list() + p.velocity.to_list() for p in ps])
    kq = np.array([p.e / p.relativistic_mass for p in ps])
    sd = [engine.rk4_steps(L, 2e-3, p.speed) for L, p in zip(lengths, ps)]
    ns = np.array([s for s, _ in sd], np.int64)
    dt = np.array([d for _, d in sd])
    assert len(set(ns.tolist())) > 1
    tr = adjoint.AdjointTracker(bls, precision=64, ckpt_every=16, ppl=ppl)
    final = tr.forward(state, kq, dt, ns, npt)
    rng = np.random.default_rng(8)
    w = rng.normal(size=final.shape) * np.array([1.0, 1.0, 1.0, 1e-8, 1e-8, 1e-8])
    _, cct_grad, _ = tr.backward(w)
    assert np.all(np.isfinite(cct_grad)) and np.all(np.isfinite(tr.tgrad.cpu().numpy()))
    got, fd = [], []
    for c, (d, k) in enumerate(tr.model.cct_design):
        cur = specs[d].cct_params[k, 19] / 1e-7
        h = 1e-6 * abs(cur)
        J = []
        for sgn in (1.0, -1.0):
            sp = copy.deepcopy(specs)
            sp[d].cct_params[k, 19] = 1e-7 * (cur + sgn * h)
            m = engine.DeviceModel(sp, fp32=True, align_cct=True)
            out, _ = m.track(state, kq, dt, ns, npt, precision=64)
            J.append(float(np.sum(w * out)))
        fd.append((J[0] - J[1]) / (2 * h))
        got.append(cct_grad[c, 0] / cur)
    got, fd = np.array(got), np.array(fd)
    err = per_parameter_error(got, fd)
    assert err.max() < 1e-4, (got, fd, err)


def test_adjoint_vs_fp64_cpu_oracle_gradient():
    """The GPU adjoint (FP32 Biot-Savart, precision 33) against central differences of the FP64
    CPU oracle (C restatement of the reference's CPU tracking) on the 10-parameter vector."""
    from oracle import c_oracle as C, cct_oracle as O, workload_builder as WB
    n, length, step = 60, 1.0, 5e-3
    beam = W.gaussian_beam(np.random.default_rng(23), 12, 0.02)
    prob, _ = _problem(12, n, length, step, 33, 20)
    prob.particles = lambda bl: W.beam_particles(W.Beam(beam, 0.4, 250.0), bl)
    theta = W.section_theta0()
    final = prob.forward(theta)
    rng = np.random.default_rng(4)
    w = rng.normal(size=final.shape) * np.array([1.0, 1.0, 1.0, 1e-8, 1e-8, 1e-8])
    grad = prob.backward(w)

    def J_cpu(th):
        bl = WB.build_beamline(W.section_ops(th, n, 0.0))
        st, kq, speed = WB.beam_state(beam, 0.4, 250.0, bl)
        sd = [O.rk4_steps(length, step, s) for s in speed]
        out = C.track(O.beamline_arrays(bl.magnets), st, kq, np.array([d for _, d in sd]),
                      np.array([s for s, _ in sd], np.int64))
        return float(np.sum(w * out))

    fd = np.zeros_like(theta)
    for k in range(theta.size):
        h = 1e-6 * max(abs(theta[k]), 1.0)
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd[k] = (J_cpu(tp) - J_cpu(tm)) / (2 * h)
    err = per_parameter_error(grad, fd)
    print("per-parameter relative error:", np.array2string(err, precision=2))
    assert err.max() < 1e-4, (grad, fd, err)


def test_adjoint_beam_spot_objective_vs_cpu_oracle():
    """The beam-spot objective J = rms x + rms y + |D| (D = least-squares slope of x against
    delta) at the end plane, differentiated by the GPU adjoint (FP32 Biot-Savart, checkpoints
    every 20 steps, J and its seed through torch autograd) against central differences of J
    evaluated on final states of the FP64 CPU oracle, per parameter of the 10-parameter vector
    (reduced size: 60 segments per winding, 12 protons, 1 m in 5 mm steps)."""
    from oracle import c_oracle as C, cct_oracle as O, workload_builder as WB
    n, length, step, s0 = 60, 1.0, 5e-3, 0.4
    beam = W.gaussian_beam(np.random.default_rng(29), 12, 0.04)
    prob, _ = _problem(12, n, length, step, 33, 20)
    prob.particles = lambda bl: W.beam_particles(W.Beam(beam, s0, 250.0), bl)
    bl0 = W.build_beamline(W.section_ops(W.section_theta0(), n, 0.0))
    ip = ParticleFactory.create_proton_along(bl0, s0 + length, 250.0)
    fr = ip.get_natural_coordinate_system()
    args = (ip.position.to_list(), fr.XI.to_list(), fr.YI.to_list(), beam[:, 4])
    theta = torch.tensor(W.section_theta0(), requires_grad=True)
    J = W.beam_spot_objective(adjoint.track_fn(theta, prob), *args)
    J.backward()
    g = theta.grad.numpy().copy()
    th0 = theta.detach().numpy().copy()
    # the end plane of the oracle's own ideal particle (same orbit: theta moves no geometry)
    ref0 = WB.build_beamline(W.section_ops(th0, n, 0.0))
    oip = O.create_proton_along(ref0.traj, s0 + length, 250.0)
    ofr = oip.natural_frame()
    assert np.max(np.abs(np.array(oip.position) - np.array(args[0]))) < 1e-12

    def J_cpu(th):
        bl = WB.build_beamline(W.section_ops(th, n, 0.0))
        st, kq, speed = WB.beam_state(beam, s0, 250.0, bl)
        sd = [O.rk4_steps(length, step, s) for s in speed]
        out = C.track(O.beamline_arrays(bl.magnets), st, kq, np.array([d for _, d in sd]),
                      np.array([s for s, _ in sd], np.int64))
        return float(W.beam_spot_objective(torch.from_numpy(out), list(oip.position), list(ofr.XI), list(ofr.YI),
                                           beam[:, 4]))

    fd = np.zeros_like(th0)
    for k in range(th0.size):
        h = 1e-6 * max(abs(th0[k]), 1.0)
        tp, tm = th0.copy(), th0.copy()
        tp[k] += h
        tm[k] -= h
        fd[k] = (J_cpu(tp) - J_cpu(tm)) / (2 * h)
    assert abs(J_cpu(th0) - float(J)) <= 1e-5 * abs(float(J))
    err = per_parameter_error(g, fd)
    print("per-parameter relative error:", np.array2string(err, precision=2))
    assert err.max() < 1e-4, (g, fd, err)


def test_adjoint_split_wave_blocks_match_full_blocks():
    """32-point split-wave blocks (forward recompute and backward kernel) give the gradients of
    64-point blocks, on a population with 24 and 9 protons per design."""
    from cctpy_amd import engine
    wl = W.population(n_designs=2, n_particles=24, seed=6)
    bls = [W.build_beamline(d) for d in wl.designs]
    ps, npt, lengths = [], [], []
    for i, (bl, beam) in enumerate(zip(bls, wl.beams)):
        q = W.beam_particles(beam, bl)[: (24 if i == 0 else 9)]
        ps += q
        npt.append(len(q))
        lengths += [0.6] * len(q)
    state = np.array([p.position.to_list() + p.velocity.to_list() for p in ps])
    kq = np.array([p.e / p.relativistic_mass for p in ps])
    sd = [engine.rk4_steps(L, 2e-3, p.speed) for L, p in zip(lengths, ps)]
    ns = np.array([s for s, _ in sd], np.int64)
    dt = np.array([d for _, d in sd])
    w = np.random.default_rng(2).normal(size=(len(ps), 6)) * np.array([1.0, 1.0, 1.0, 1e-8, 1e-8, 1e-8])
    res = []
    for ppl in (1, -2, -4):
        tr = adjoint.AdjointTracker(bls, precision=33, ckpt_every=16, ppl=ppl)
        final = tr.forward(state, kq, dt, ns, npt)
        a0, cg, _ = tr.backward(w)
        res.append((final, a0, cg))
    for final, a0, cg in res[1:]:
        # the shapes decide near tiles (summed in FP64) over different lane sets (a wave or a
        # 16-lane group), so a few tiles are FP32 K/C sums in one shape and FP64 in the other
        assert np.max(np.abs(final[:, :3] - res[0][0][:, :3])) < 1e-8
        assert np.max(np.abs(a0 - res[0][1])) <= 1e-6 * np.max(np.abs(res[0][1]))
        assert np.max(np.abs(cg - res[0][2])) <= 1e-5 * np.max(np.abs(res[0][2]))


def test_adjoint_full_tape_matches_checkpoints():
    """ckpt_every = 0 (the forward keeps every step's tape, no recompute) gives the bits of the
    checkpointed reverse pass, for 64-point and split-wave blocks."""
    from cctpy_amd import engine
    wl = W.population(n_designs=2, n_particles=16, seed=6)
    bls = [W.build_beamline(d) for d in wl.designs]
    ps, npt, lengths = [], [], []
    for i, (bl, beam) in enumerate(zip(bls, wl.beams)):
        q = W.beam_particles(beam, bl)
        ps += q
        npt.append(len(q))
        lengths += [0.5 + 0.1 * i] * len(q)
    state = np.array([p.position.to_list() + p.velocity.to_list() for p in ps])
    kq = np.array([p.e / p.relativistic_mass for p in ps])
    sd = [engine.rk4_steps(L, 2e-3, p.speed) for L, p in zip(lengths, ps)]
    ns = np.array([s for s, _ in sd], np.int64)
    dt = np.array([d for _, d in sd])
    w = np.random.default_rng(3).normal(size=(len(ps), 6)) * np.array([1.0, 1.0, 1.0, 1e-8, 1e-8, 1e-8])
    for ppl in (1, -2):
        out = []
        for k in (16, 0):
            tr = adjoint.AdjointTracker(bls, precision=33, ckpt_every=k, ppl=ppl)
            final = tr.forward(state, kq, dt, ns, npt)
            a0, cg, qg = tr.backward(w)
            out.append((final, a0, cg))
        assert np.array_equal(out[0][0], out[1][0])
        assert np.array_equal(out[0][1], out[1][1])
        assert np.array_equal(out[0][2], out[1][2])


def test_adjoint_tune_picks_a_candidate_and_keeps_results():
    """tune() times short forward + adjoint passes, sets one of the candidate block shapes and
    work-item counts, and a tracker prepared after it gives the bits of a fresh tracker with
    that shape."""
    from cctpy_amd import engine
    wl = W.population(n_designs=2, n_particles=16, seed=6)
    bls = [W.build_beamline(d) for d in wl.designs]
    ps, npt, lengths = [], [], []
    for i, (bl, beam) in enumerate(zip(bls, wl.beams)):
        q = W.beam_particles(beam, bl)
        ps += q
        npt.append(len(q))
        lengths += [0.4] * len(q)
    state = np.array([p.position.to_list() + p.velocity.to_list() for p in ps])
    kq = np.array([p.e / p.relativistic_mass for p in ps])
    sd = [engine.rk4_steps(L, 2e-3, p.speed) for L, p in zip(lengths, ps)]
    ns = np.array([s for s, _ in sd], np.int64)
    dt = np.array([d for _, d in sd])
    dev = engine.device()
    s0 = torch.from_numpy(np.ascontiguousarray(state.T)).to(dev)
    kq_t, dt_t, ns_t = (torch.from_numpy(a).to(dev) for a in (kq, dt, ns))
    w = np.random.default_rng(5).normal(size=(len(ps), 6)) * np.array([1.0, 1.0, 1.0, 1e-8, 1e-8, 1e-8])
    tr = adjoint.AdjointTracker(bls, precision=33, ckpt_every=16, ppl=2)
    best = tr.tune(npt, s0, kq_t, dt_t, ns_t, candidates=(1, 2, -2), probe_steps=6, items_per_cu=(4, 16))
    assert best in (1, 2, -2) and tr.ppl == best and tr.items_per_cu in (4, 16)
    out = []
    for t in (tr, adjoint.AdjointTracker(bls, precision=33, ckpt_every=16, ppl=best, items_per_cu=tr.items_per_cu)):
        final = t.forward(state, kq, dt, ns, npt)
        a0, cg, qg = t.backward(w)
        out.append((final, a0, cg))
    for x, y in zip(out[0], out[1]):
        assert np.array_equal(x, y)
