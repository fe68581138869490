"""Beam statistics of tracked populations on the device.

``BeamGroups`` projects tracked states into the natural frame of each design's
ideal particle (``PhaseSpaceParticle.create_from_running_particle``,
final_code/packages/particles.py:1077-1121), reduces per group of particles
(count, x/y extent, means, rms sizes, dispersion moments; the beam-size loop of
cctpy3/run.py:72-103 uses (max - min) / 2 per group) and writes the adjoint seed
dJ/d(final state) of the beam-spot objective J = w0 rms x + w1 rms y + w2 |D|.
Everything stays in device memory; only the (G, 12) statistics are small enough
to be worth copying back.
"""
from __future__ import annotations

from typing import Sequence, Tuple

import numpy
import torch

from . import _native as N
from .engine import _p, _stream, device
from .local_coordinate_system import LocalCoordinateSystem
from .particles import RunningParticle

STATS = ("n", "min_x", "max_x", "min_y", "max_y", "mean_x", "mean_y", "mean_delta", "rms_x", "rms_y",
         "sdd", "sxd")


def ideal_record(ideal: RunningParticle, frame: LocalCoordinateSystem) -> numpy.ndarray:
    """16 doubles {p, v, P = speed * mass, XI, YI, ZI} of one design's ideal particle."""
    return numpy.array(ideal.position.to_list() + ideal.velocity.to_list() + [ideal.compute_scalar_momentum()]
                       + frame.XI.to_list() + frame.YI.to_list() + frame.ZI.to_list(), numpy.float64)


class BeamGroups:
    """Groups of consecutive particles of a population (each group within one design)."""

    def __init__(self, ideals: Sequence[RunningParticle], frames: Sequence[LocalCoordinateSystem],
                 design_npt: Sequence[int], speeds: Sequence[float], momenta: Sequence[float], group_size: int = 0):
        """``group_size`` 0: one group per design; otherwise consecutive groups of that many
        particles inside each design (the last one may be shorter)."""
        if len(ideals) != len(design_npt) or len(frames) != len(design_npt):
            raise ValueError("one ideal particle and frame per design is required")
        if group_size < 0:
            raise ValueError("group_size must be >= 0")
        dev = device()
        self.ndesign = len(design_npt)
        self.P = int(sum(design_npt))
        if len(speeds) != self.P or len(momenta) != self.P:
            raise ValueError("one speed and momentum per particle is required")
        off, gdes = [0], []
        start = 0
        for d, n in enumerate(design_npt):
            n = int(n)
            step = n if group_size == 0 else group_size
            k = 0
            while k < n:
                m = min(step, n - k)
                off.append(start + k + m)
                gdes.append(d)
                k += m
            start += n
        self.ngroups = len(gdes)
        self.group_off_host = numpy.array(off, numpy.int64)
        self.group_design_host = numpy.array(gdes, numpy.int32)
        self.group_off = torch.from_numpy(self.group_off_host).to(dev)
        self.group_design = torch.from_numpy(self.group_design_host).to(dev)
        rec = numpy.stack([ideal_record(i, f) for i, f in zip(ideals, frames)]) if self.ndesign else numpy.zeros((1, 16))
        self.ideal = torch.from_numpy(numpy.ascontiguousarray(rec)).to(dev)
        self.speed = torch.tensor(list(speeds), dtype=torch.float64, device=dev)
        self.momentum = torch.tensor(list(momenta), dtype=torch.float64, device=dev)

    def project(self, state: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
        """(pps (6, P) rows x, xp, y, yp, z, delta; stats (G, 12) in the order of ``STATS``)."""
        if tuple(state.shape) != (6, self.P) or state.dtype != torch.float64 or not state.is_contiguous():
            raise ValueError("state must be a contiguous float64 (6, P) device tensor")
        pps = torch.empty_like(state)
        stats = torch.empty((max(self.ngroups, 1), 12), dtype=torch.float64, device=state.device)
        N.check(N.lib.cctamd_phase_space(_p(self.ideal), self.ndesign, _p(state), _p(self.speed),
                                         _p(self.momentum), self.P, _p(self.group_off), _p(self.group_design),
                                         self.ngroups, _p(pps), _p(stats), _stream()), "cctamd_phase_space")
        return pps, stats[: self.ngroups]

    @staticmethod
    def objective(stats: torch.Tensor, weights=(1.0, 1.0, 1.0)) -> numpy.ndarray:
        """J per group = w0 rms x + w1 rms y + w2 |D|, D = sum (x - mx)(d - md) / sum (d - md)^2."""
        s = stats.cpu().numpy()
        d = numpy.where(s[:, 10] > 0, s[:, 11] / numpy.where(s[:, 10] > 0, s[:, 10], 1.0), 0.0)
        return weights[0] * s[:, 8] + weights[1] * s[:, 9] + weights[2] * numpy.abs(d)

    @staticmethod
    def half_sizes(stats: torch.Tensor) -> numpy.ndarray:
        """((max x - min x) / 2, (max y - min y) / 2) per group (cctpy3/run.py:88-89)."""
        s = stats.cpu().numpy()
        return numpy.stack([(s[:, 2] - s[:, 1]) / 2, (s[:, 4] - s[:, 3]) / 2], axis=1)

    def seed(self, pps: torch.Tensor, stats: torch.Tensor, weights=(1.0, 1.0, 1.0)) -> torch.Tensor:
        """dJ/d(final state) (6, P) of the objective of every group."""
        adj = torch.zeros_like(pps)
        w = numpy.ascontiguousarray(weights, numpy.float64)
        if w.shape != (3,):
            raise ValueError("three weights are required")
        N.check(N.lib.cctamd_beam_objective_seed(_p(self.ideal), self.ndesign, _p(pps), _p(stats.contiguous()),
                                                 self.P, _p(self.group_off), _p(self.group_design), self.ngroups,
                                                 w.ctypes.data_as(N.ctypes.c_void_p), _p(adj), _stream()),
                "cctamd_beam_objective_seed")
        return adj
