"""Gradients of tracking results with respect to magnet parameters (reverse mode on the GPU).

``AdjointTracker`` runs the forward RK4 tracking with checkpoints every K
steps and the discrete adjoint (``cctamd_track_rk4_adjoint``: forward
recompute per checkpoint interval, stage adjoints, backward Biot-Savart
kernel).  It returns dJ/d(initial state), per-CCT dJ/d(current) and
dJ/d(tilt coefficients), and per-QS dJ/d(gradient), dJ/d(second gradient).

``DesignProblem`` chains these to a parameter vector theta of a beamline
builder (tilt angles, currents, QS gradients, ...) and ``track_fn`` exposes
theta -> final particle states as a ``torch.autograd.Function``.
"""
from __future__ import annotations

import math
from typing import Callable, List, Optional, Sequence, Tuple

import numpy
import torch

from . import _native as N
from . import engine
from .engine import _p, _stream
from .magnets import Magnet, device_parts

# tilt angles per CCT whose gradients the adjoint carries (kGradTilts in csrc/adjoint.hip)
MAX_GRAD_TILTS = 4


class AdjointTracker:
    """Forward tracking with checkpoints and its adjoint for one set of designs."""

    def __init__(self, magnets: Sequence[Magnet], precision: int = 33, ckpt_every: int = 50, ppl: int = 2,
                 items_per_cu: Optional[int] = None):
        """``ckpt_every`` K > 0: checkpoints every K steps and recompute in the reverse pass;
        K = 0: the forward keeps the tape of every step (288 B per particle-step of HBM) and the
        reverse pass recomputes nothing.  ``ppl`` is the points-per-lane code of the work plan,
        ``items_per_cu`` its work items per CU (None: the engine's default)."""
        if ckpt_every < 0:
            raise ValueError("ckpt_every must be >= 0")
        if ppl not in (1, 2, 4, -2, -4):
            raise ValueError(f"points-per-lane code {ppl} is not one of 1, 2, 4, -2, -4")
        specs = []
        for m in magnets:
            parts = device_parts(m)
            if parts.foreign:
                raise ValueError(f"magnet {parts.foreign[0]} cannot be tracked on the GPU")
            specs.append(parts.spec())
        self.specs = specs
        self.model = engine.DeviceModel(specs, fp32=True, align_cct=True)
        self.precision = precision
        self.ckpt_every = int(ckpt_every)
        self.ppl = ppl
        self.items_per_cu = items_per_cu

    def prepare(self, design_npt: Sequence[int], kq: torch.Tensor, dt: torch.Tensor, nsteps: torch.Tensor,
                max_steps: int) -> None:
        """Plan and device workspaces for particles grouped by design; kq, dt, nsteps are
        device tensors (P,) that stay bound to this tracker."""
        dev = engine.device()
        m = self.model
        target = self.items_per_cu * engine.cu_count() if self.items_per_cu else None
        self.plan = plan = m.plan(design_npt, self.ppl, target)
        self.P = P = plan.npt
        self.kq, self.dt, self.ns = kq, dt, nsteps
        self.max_steps = int(max_steps)
        if self.ckpt_every > 0:
            nck = -(-self.max_steps // self.ckpt_every)
            self.ckpt = torch.zeros((nck + 1, 6, P), dtype=torch.float64, device=dev)
        else:  # full tape, no recompute: 288 B per particle-step
            self.ckpt = None
        self.scratch, self.partial = m.workspace(plan)
        self.ntile = max(m.nseg_padded // N.TILE, 1)
        pb = engine.points_per_block(plan.ppl)
        nblock = max(-(-int(n) // pb) for n in numpy.diff(plan.design_pt_off.cpu().numpy()))
        self.tgrad = torch.zeros((max(nblock, 1), self.ntile, 5), dtype=torch.float64, device=dev)
        self.nq = nq = int(m.design_qs_off_host[-1])
        self.qgrad = torch.zeros((max(2 * nq, 1), P), dtype=torch.float64, device=dev)
        self.state_ws = torch.empty((6, P), dtype=torch.float64, device=dev)
        self.tape = torch.zeros((self.ckpt_every if self.ckpt_every > 0 else max(self.max_steps, 1), 36, P),
                                dtype=torch.float64, device=dev)
        self.adj_ws = torch.zeros((15, P), dtype=torch.float64, device=dev)
        self.bs_pts = torch.zeros((3, P), dtype=torch.float64, device=dev)
        self.bs_wts = torch.zeros((3, P), dtype=torch.float64, device=dev)

    def tune(self, design_npt: Sequence[int], state: torch.Tensor, kq: torch.Tensor, dt: torch.Tensor,
             nsteps: torch.Tensor, candidates: Sequence[int] = (1, 2, -2), probe_steps: int = 8,
             items_per_cu: Sequence[int] = (4, 8, 16)) -> int:
        """Sets the points-per-lane code and the work items per CU whose forward + adjoint pass
        over the first ``probe_steps`` steps is fastest, and returns the code.  The backward
        kernel dominates the adjoint and prefers other block shapes than the forward alone (2
        points per lane against 1 on a 256-proton section), so the shape is chosen on both.  The
        tracker is left prepared for ``nsteps`` with the chosen shape."""
        bad = [c for c in candidates if c not in (1, 2, 4, -2, -4)]
        if bad:
            raise ValueError(f"points-per-lane codes {bad} are not among 1, 2, 4, -2, -4")
        ns = torch.clamp(nsteps, max=probe_steps)
        best, best_ms = (self.ppl, self.items_per_cu), math.inf
        for ppl, ipc in [(a, b) for a in candidates for b in items_per_cu]:
            self.ppl, self.items_per_cu = ppl, ipc
            self.prepare(design_npt, kq, dt, ns, probe_steps)
            seed = torch.zeros((6, self.P), dtype=torch.float64, device=state.device)
            seed[:3] = 1.0
            times = []
            for _ in range(2):  # the first pass warms up
                st = state.clone()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                self.forward_tensors(st)
                self.backward_tensors(seed.clone())
                e1.record()
                torch.cuda.synchronize()
                times.append(e0.elapsed_time(e1))
            if times[-1] < best_ms:
                best, best_ms = (ppl, ipc), times[-1]
        self.ppl, self.items_per_cu = best
        # prepared for the caller's own steps again, not the probe's
        self.prepare(design_npt, kq, dt, nsteps, int(nsteps.max().item()) if nsteps.numel() else 0)
        return self.ppl

    def forward_tensors(self, state: torch.Tensor) -> None:
        """In-place tracking of device state (6, P), keeping checkpoints every K steps (or, with
        K = 0, the tape of every step)."""
        if self.ckpt_every > 0:
            self.model.track_tensors(self.plan, self.precision, state, self.kq, self.dt, self.ns, self.max_steps,
                                     self.scratch, self.partial, record=self.ckpt, record_every=self.ckpt_every)
            return
        m, plan = self.model, self.plan
        seg, org, near = m._seg(self.precision)
        N.check(N.lib.cctamd_track_rk4_taped(
            self.precision, plan.ppl, _p(seg), _p(org), _p(near), _p(plan.work), plan.nwork, _p(plan.wbase), _p(plan.nchunk),
            _p(plan.design_pt_off), m.ndesign, _p(m.qs), _p(m.design_qs_off), _p(m.bconst), _p(state), self.P,
            _p(self.kq), _p(self.dt), _p(self.ns), self.max_steps, _p(self.scratch), _p(self.partial), _p(self.tape),
            _stream()), "cctamd_track_rk4_taped")

    def backward_tensors(self, adj: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
        """In-place reverse pass: adj (6, P) dJ/d(final state) -> dJ/d(initial state).
        Returns the device accumulators (tgrad (nblock, ntile, 5), qgrad (2Q, P))."""
        m, plan = self.model, self.plan
        seg, _, near = m._seg(self.precision)
        org = m.tile_org  # FP32 tile origins: the backward kernel needs them at every precision
        tan, tile_cct, coef = m.adjoint_data()
        self.tgrad.zero_()
        self.qgrad.zero_()
        N.check(N.lib.cctamd_track_rk4_adjoint(
            self.precision, plan.ppl, _p(seg), _p(org), _p(near), _p(m.seg32), _p(tan), _p(tile_cct), _p(coef), self.ntile,
            _p(plan.work), plan.nwork, _p(plan.wbase), _p(plan.nchunk), _p(plan.design_pt_off), m.ndesign,
            _p(m.qs), _p(m.design_qs_off), _p(m.bconst), _p(self.ckpt), self.ckpt_every, self.P, _p(self.kq),
            _p(self.dt), _p(self.ns), self.max_steps, _p(adj), _p(self.tgrad), _p(self.qgrad), _p(self.state_ws),
            _p(self.scratch), _p(self.partial), _p(self.tape), _p(self.adj_ws), _p(self.bs_pts), _p(self.bs_wts),
            _stream()), "cctamd_track_rk4_adjoint")
        return self.tgrad, self.qgrad

    def reduce_gradients(self) -> Tuple[numpy.ndarray, numpy.ndarray]:
        """(cct_grad (ncct, 5) per global CCT: [sum B-bar.dB, dJ/dc_0..3], qs_grad (Q, 2))."""
        m = self.model
        _, tile_cct, _ = m.adjoint_data()
        tg = self.tgrad.sum(0).cpu().numpy()
        tc = tile_cct.cpu().numpy()
        cct_grad = numpy.zeros((m.ncct, 5))
        for c in range(m.ncct):
            cct_grad[c] = tg[tc == c].sum(0)
        q = self.qgrad.cpu().numpy()[: 2 * self.nq].reshape(self.nq, 2, self.P).sum(2)
        return cct_grad, q

    def forward(self, state: numpy.ndarray, kq: numpy.ndarray, dt: numpy.ndarray, nsteps: numpy.ndarray,
                design_npt: Sequence[int]) -> numpy.ndarray:
        """Track (P, 6) states; keeps the checkpoints for ``backward``."""
        dev = engine.device()
        ns = numpy.asarray(nsteps, numpy.int64)
        self.prepare(design_npt, torch.from_numpy(numpy.ascontiguousarray(kq, numpy.float64)).to(dev),
                     torch.from_numpy(numpy.ascontiguousarray(dt, numpy.float64)).to(dev),
                     torch.from_numpy(numpy.ascontiguousarray(ns)).to(dev), int(ns.max()))
        st = torch.from_numpy(numpy.ascontiguousarray(numpy.asarray(state, numpy.float64).reshape(-1, 6).T)).to(dev)
        self.forward_tensors(st)
        return st.cpu().numpy().T.copy()

    def backward(self, adj_final: numpy.ndarray):
        """Reverse pass for dJ/d(final state) (P, 6).

        Returns (dJ/d(initial state) (P, 6), cct_grad (ncct, 5) per global CCT:
        [sum B-bar.dB, dJ/dc_0..3], qs_grad (Q, 2)).
        """
        dev = engine.device()
        adj = torch.from_numpy(numpy.ascontiguousarray(numpy.asarray(adj_final, numpy.float64).reshape(-1, 6).T)).to(dev)
        self.backward_tensors(adj)
        cct_grad, q = self.reduce_gradients()
        return adj.cpu().numpy().T.copy(), cct_grad, q


class DesignProblem:
    """theta -> beamline -> tracked beam, with dJ/dtheta by the adjoint.

    ``builder(theta) -> Beamline`` must keep the winding counts and the orbit
    fixed (tilts, currents and QS gradients may vary).  ``particles(beamline)``
    gives the initial RunningParticles (independent of theta).
    """

    def __init__(self, builder: Callable, particles: Callable, length: float, footstep: float,
                 precision: int = 33, ckpt_every: int = 50, fd_step: float = 1e-6):
        self.builder = builder
        self.particles = particles
        self.length = length
        self.footstep = footstep
        self.precision = precision
        self.ckpt_every = ckpt_every
        self.fd_step = fd_step

    def _setup(self, theta: numpy.ndarray):
        bl = self.builder(theta)
        ps = self.particles(bl)
        state = numpy.array([p.position.to_list() + p.velocity.to_list() for p in ps])
        kq = numpy.array([p.e / p.relativistic_mass for p in ps])
        sd = [engine.rk4_steps(self.length, self.footstep, p.speed) for p in ps]
        return bl, state, kq, numpy.array([d for _, d in sd]), numpy.array([s for s, _ in sd], numpy.int64)

    def forward(self, theta: numpy.ndarray) -> numpy.ndarray:
        self.theta = numpy.asarray(theta, numpy.float64).copy()
        bl, state, kq, dt, ns = self._setup(self.theta)
        self.tracker = AdjointTracker([bl], self.precision, self.ckpt_every)
        self.beamline = bl
        return self.tracker.forward(state, kq, dt, ns, [len(state)])

    def _magnet_params(self, theta: numpy.ndarray):
        """The CCTs and QS of the design at theta, flattened as the device holds them (nested
        CombinedMagnets included, append order)."""
        parts = device_parts(self.builder(theta))
        return parts.cct_objects, parts.qs_objects

    def parameter_map(self, theta: numpy.ndarray):
        """How each parameter moves the device's magnet parameters at theta (central
        differences of the builder, step fd_step relative): per parameter k a list of
        (CCT index, dI/dtheta_k, [dc_j/dtheta_k for the first four tilt coefficients]) and of
        (QS index, dg/dtheta_k, ds/dtheta_k), indices in the device's flattened order.

        Parameters enter through the CCT currents, the first four tilt angles of each CCT
        (the tilt slots the adjoint carries) and the QS gradients; a parameter that moves
        anything else the adjoint cannot see (a fifth tilt, the number of magnets or tilts)
        raises ValueError.  Host only."""
        theta = numpy.asarray(theta, numpy.float64)
        ccts0, qss0 = self._magnet_params(theta)
        out = []
        for kth in range(theta.size):
            h = self.fd_step * max(1.0, abs(theta[kth]))
            tp, tm = theta.copy(), theta.copy()
            tp[kth] += h
            tm[kth] -= h
            cp, qp = self._magnet_params(tp)
            cm, qm = self._magnet_params(tm)
            if len(cp) != len(ccts0) or len(cm) != len(ccts0) or len(qp) != len(qss0) or len(qm) != len(qss0):
                raise ValueError(f"parameter {kth} changes the number of CCT or QS magnets")
            cct_terms, qs_terms = [], []
            for i, c0 in enumerate(ccts0):
                nt = len(c0.tilt_angles)
                if len(cp[i].tilt_angles) != nt or len(cm[i].tilt_angles) != nt:
                    raise ValueError(f"parameter {kth} changes the number of tilt angles of CCT {i}")
                for j in range(MAX_GRAD_TILTS, nt):
                    if cp[i].tilt_angles[j] != cm[i].tilt_angles[j]:
                        raise ValueError(f"parameter {kth} moves tilt angle {j + 1} of CCT {i}: gradients exist "
                                         f"for the first {MAX_GRAD_TILTS} tilt angles only")
                dI = (cp[i].current - cm[i].current) / (2 * h)
                sh = math.sinh(c0.eta)
                dc = [0.0] * MAX_GRAD_TILTS
                for j, a0 in enumerate(c0.tilt_angles[:MAX_GRAD_TILTS]):
                    da = (cp[i].tilt_angles[j] - cm[i].tilt_angles[j]) / (2 * h)
                    if da != 0.0:
                        dc[j] = -(math.pi / 180.0) / math.sin(math.radians(a0)) ** 2 / ((j + 1) * sh) * da
                if dI != 0.0 or any(dc):
                    cct_terms.append((i, dI, dc))
            for i in range(len(qss0)):
                dg = (qp[i].gradient - qm[i].gradient) / (2 * h)
                ds = (qp[i].second_gradient - qm[i].second_gradient) / (2 * h)
                if dg != 0.0 or ds != 0.0:
                    qs_terms.append((i, dg, ds))
            out.append((cct_terms, qs_terms))
        return out, ccts0

    def backward(self, adj_final: numpy.ndarray) -> numpy.ndarray:
        """dJ/dtheta for dJ/d(final state) (parameters as in parameter_map)."""
        pmap, ccts0 = self.parameter_map(self.theta)  # raises before any device work
        _, cct_grad, qs_grad = self.tracker.backward(adj_final)
        # design CCT index -> global CCT index (zero-current CCTs carry no segments)
        gl = {k: c for c, (d, k) in enumerate(self.tracker.model.cct_design)}
        grad = numpy.zeros_like(self.theta)
        for kth, (cct_terms, qs_terms) in enumerate(pmap):
            g = 0.0
            for i, dI, dc in cct_terms:
                if i not in gl:
                    continue
                cg = cct_grad[gl[i]]
                if dI != 0.0:
                    g += cg[0] / ccts0[i].current * dI
                for j in range(MAX_GRAD_TILTS):
                    if dc[j] != 0.0:
                        g += cg[1 + j] * dc[j]
            for i, dg, ds in qs_terms:
                g += qs_grad[i, 0] * dg + qs_grad[i, 1] * ds
            grad[kth] = g
        return grad


class _TrackFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta: torch.Tensor, problem: DesignProblem) -> torch.Tensor:
        out = problem.forward(theta.detach().cpu().numpy())
        ctx.problem = problem
        return torch.from_numpy(out)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        g = ctx.problem.backward(grad_out.detach().cpu().numpy())
        return torch.from_numpy(g), None


def track_fn(theta: torch.Tensor, problem: DesignProblem) -> torch.Tensor:
    """Final states (P, 6) [x, y, z, vx, vy, vz] as a differentiable function of theta."""
    return _TrackFn.apply(theta, problem)
