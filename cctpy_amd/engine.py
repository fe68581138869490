"""Device engine: HBM-resident segment sets, work plans and calls into libcctamd.

One *design* is one magnet arrangement (a Beamline, a CCT, a CombinedMagnet);
a ``DeviceModel`` holds D designs at once:

* CCT windings are generated on the GPU (``cctamd_winding_generate``) straight
  into an FP64 AoS segment array in the global frame, each design padded to a
  multiple of 32 segments (one *tile*), and packed once into FP32 tiles
  re-centred on an FP64 tile origin, the mean midpoint of the tile
  (``cctamd_pack_fp32``);
* QS magnets travel as the reference's 16-double records, uniform magnets as a
  per-design constant field.

Field and tracking calls take points/particles per design in SoA FP64 and run
on torch's current HIP stream.  There is no CPU path: without a GPU every
entry point raises.
"""
from __future__ import annotations

import collections
import ctypes
import hashlib
import math
from dataclasses import dataclass, field as dc_field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy
import torch

from . import _native as N

_TARGET_WG_PER_CU = 4
_MIN_TILES_PER_ITEM = 4


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("cctpy_amd runs on an AMD Instinct GPU (gfx950); no GPU is visible")
    return torch.device("cuda", torch.cuda.current_device())


def _p(t: Optional[torch.Tensor]):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else ctypes.c_void_p(0)


def _np(a: numpy.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def points_per_block(ppl: int) -> int:
    """Point-block size of a points-per-lane code (1, 2, 4 -> 64 ppl; -2, -4 -> 32, 16)."""
    return 64 * ppl if ppl > 0 else 64 // (-ppl)


def block_shape_for(design_npt: Sequence[int], precision: int) -> int:
    """Points-per-lane code whose blocks waste the fewest lanes for these per-design particle
    counts, for callers that do not tune: the smallest of 16, 32, 64 points per block that
    holds the largest design's particles (16 and 32 only where split-wave kernels exist:
    precisions 33 and 64), else 128 (2 points per lane)."""
    m = int(numpy.max(design_npt)) if len(design_npt) else 0
    if precision != 32 and m <= 16:
        return -4
    if precision != 32 and m <= 32:
        return -2
    return 1 if m <= 64 else 2


def cu_count() -> int:
    return int(torch.cuda.get_device_properties(device()).multi_processor_count)


@dataclass
class DesignSpec:
    """Host description of one design: CCT records, QS records, uniform field."""
    cct_params: numpy.ndarray = dc_field(default_factory=lambda: numpy.zeros((0, N.CCT_PARAMS)))
    cct_counts: numpy.ndarray = dc_field(default_factory=lambda: numpy.zeros((0,), numpy.int64))
    qs: numpy.ndarray = dc_field(default_factory=lambda: numpy.zeros((0, 16)))
    bconst: numpy.ndarray = dc_field(default_factory=lambda: numpy.zeros(3))
    # explicit global-frame current elements (kl, midpoint), placed before the CCT windings
    raw_kl: numpy.ndarray = dc_field(default_factory=lambda: numpy.zeros((0, 3)))
    raw_m: numpy.ndarray = dc_field(default_factory=lambda: numpy.zeros((0, 3)))

    @property
    def nseg(self) -> int:
        return int(self.cct_counts.sum()) + int(self.raw_kl.shape[0])

    def key(self) -> bytes:
        h = hashlib.sha1()
        for a in (self.cct_params, self.cct_counts, self.qs, self.bconst, self.raw_kl, self.raw_m):
            h.update(numpy.ascontiguousarray(a).tobytes())
            h.update(str(a.shape).encode())
        return h.digest()


@dataclass
class Plan:
    ppl: int
    nwork: int
    work: torch.Tensor
    wbase: torch.Tensor
    nchunk: torch.Tensor
    design_pt_off: torch.Tensor
    npt: int


class DeviceModel:
    """D designs resident in HBM (FP64 segments, FP32 tiles, QS records, uniform fields)."""

    def __init__(self, specs: Sequence[DesignSpec], fp32: bool = True, align_cct: bool = False):
        """``align_cct`` starts every CCT on a tile boundary (needed by the adjoint,
        which attributes each tile to one CCT)."""
        dev = device()
        self.specs = list(specs)
        self.ndesign = D = len(self.specs)
        if D == 0:
            raise ValueError("no design given")
        self.align_cct = align_cct
        if align_cct:
            nseg = numpy.array([-(-int(s.raw_kl.shape[0]) // N.TILE) * N.TILE
                                + sum(-(-int(c) // N.TILE) * N.TILE for c in s.cct_counts)
                                for s in self.specs], numpy.int64)
        else:
            nseg = numpy.array([s.nseg for s in self.specs], numpy.int64)
        self.design_ntile = (nseg + N.TILE - 1) // N.TILE
        self.design_tile_off = numpy.concatenate([[0], numpy.cumsum(self.design_ntile)]).astype(numpy.int64)
        self.design_nseg = nseg
        S = int(self.design_tile_off[-1]) * N.TILE
        self.nseg_padded = S
        self.nseg = int(nseg.sum())
        # pad segments: kl = 0 and a midpoint far from every field point
        seg64 = torch.zeros((max(S, 1), 8), dtype=torch.float64, device=dev)
        seg64[:, 4:7] = 1.0e3
        params, counts, dst, cct_design = [], [], [], []
        for d, s in enumerate(self.specs):
            row = int(self.design_tile_off[d]) * N.TILE
            nraw = int(s.raw_kl.shape[0])
            if nraw:
                rows = numpy.zeros((nraw, 8))
                rows[:, 0:3] = s.raw_kl
                rows[:, 4:7] = s.raw_m
                seg64[row:row + nraw] = torch.from_numpy(rows).to(dev)
                row += nraw
            for k in range(s.cct_params.shape[0]):
                c = int(s.cct_counts[k])
                if c == 0:
                    continue
                if align_cct:
                    row = -(-row // N.TILE) * N.TILE
                params.append(s.cct_params[k])
                counts.append(c)
                dst.append(row)
                cct_design.append((d, k))
                row += c
        self.ncct = len(params)
        self.cct_design = cct_design          # global CCT index -> (design, CCT index within the design)
        self.cct_dst = numpy.array(dst, numpy.int64)
        self.cct_counts = numpy.array(counts, numpy.int64)
        self.cct_params = numpy.array(params, numpy.float64).reshape(len(params), N.CCT_PARAMS)
        self._gen = None
        if self.ncct:
            prm = torch.from_numpy(numpy.ascontiguousarray(self.cct_params)).to(dev)
            off = torch.from_numpy(numpy.concatenate([[0], numpy.cumsum(counts)]).astype(numpy.int64)).to(dev)
            dsts = torch.from_numpy(self.cct_dst).to(dev)
            self._gen = (prm, off, dsts)
            N.check(N.lib.cctamd_winding_generate(_p(prm), _p(off), _p(dsts), self.ncct, int(sum(counts)),
                                                  _p(seg64), _stream()), "cctamd_winding_generate")
        self.seg64 = seg64
        self.seg32 = self.seg32k = self.tile_org = None
        if fp32:
            self.seg32 = torch.empty((max(S, 1), 8), dtype=torch.float32, device=dev)
            self.seg32k = torch.empty((max(S, 1), 12), dtype=torch.float32, device=dev)
            self.tile_org = torch.empty((max(S // N.TILE, 1), 4), dtype=torch.float64, device=dev)
            N.check(N.lib.cctamd_pack_fp32(_p(seg64), S, _p(self.seg32), _p(self.tile_org), _p(self.seg32k),
                                           _stream()), "cctamd_pack_fp32")
        qs = [s.qs for s in self.specs]
        nq = numpy.array([q.shape[0] for q in qs], numpy.int64)
        self.design_qs_off_host = numpy.concatenate([[0], numpy.cumsum(nq)]).astype(numpy.int64)
        allq = numpy.concatenate(qs).reshape(-1, 16) if nq.sum() else numpy.zeros((1, 16))
        self.qs = torch.from_numpy(numpy.ascontiguousarray(allq, dtype=numpy.float64)).to(dev)
        self.design_qs_off = torch.from_numpy(self.design_qs_off_host).to(dev)
        b = numpy.zeros((D, 4))
        for d, s in enumerate(self.specs):
            b[d, :3] = s.bconst
        self.bconst = torch.from_numpy(b).to(dev)
        self._plans: Dict[tuple, Plan] = {}

    # --------------------------------------------------------- adjoint data
    def adjoint_data(self):
        """(seg_tan [T][33][8] f32, tile_cct [T] i32, cct_coef [ncct] f64) for cctamd_track_rk4_adjoint."""
        if not self.align_cct:
            raise ValueError("the adjoint needs a model built with align_cct=True")
        if getattr(self, "_adj", None) is None:
            dev = device()
            S = self.nseg_padded
            nt = max(S // N.TILE, 1)
            tan = torch.zeros((nt, N.TILE + 1, 8), dtype=torch.float32, device=dev)
            if self.ncct:
                prm, off, dsts = self._gen
                N.check(N.lib.cctamd_winding_tangents(_p(prm), _p(off), _p(dsts), self.ncct,
                                                      int(self.cct_counts.sum()), _p(tan), _stream()),
                        "cctamd_winding_tangents")
            tc = numpy.full(nt, -1, numpy.int32)
            for c in range(self.ncct):
                t0 = int(self.cct_dst[c]) // N.TILE
                t1 = -(-(int(self.cct_dst[c]) + int(self.cct_counts[c])) // N.TILE)
                tc[t0:t1] = c
            coef = self.cct_params[:, 19].copy() if self.ncct else numpy.zeros(1)
            self._adj = (tan, torch.from_numpy(tc).to(dev), torch.from_numpy(numpy.ascontiguousarray(coef)).to(dev))
        return self._adj

    # ------------------------------------------------------------- plans
    def plan(self, design_npt: Sequence[int], ppl: int = 2, target_items: Optional[int] = None) -> "Plan":
        npt = numpy.asarray(design_npt, numpy.int64)
        if npt.shape != (self.ndesign,):
            raise ValueError("one point count per design is required")
        if target_items is None:
            target_items = _TARGET_WG_PER_CU * cu_count()
        key = (tuple(npt.tolist()), ppl, target_items)
        if key in self._plans:
            return self._plans[key]
        pt_off = numpy.concatenate([[0], numpy.cumsum(npt)]).astype(numpy.int64)
        nt = numpy.ascontiguousarray(self.design_ntile, numpy.int64)
        toff = numpy.ascontiguousarray(self.design_tile_off[:-1], numpy.int64)
        n = N.lib.cctamd_plan(self.ndesign, _np(nt), _np(npt), _np(toff), _np(pt_off), ppl, target_items,
                              _MIN_TILES_PER_ITEM, None, None, None)
        if n < 0:
            N.check(int(n), "cctamd_plan")
        work = numpy.zeros((max(n, 1), N.WORK_INTS), numpy.int32)
        wbase = numpy.zeros(self.ndesign, numpy.int64)
        nchunk = numpy.zeros(self.ndesign, numpy.int32)
        n2 = N.lib.cctamd_plan(self.ndesign, _np(nt), _np(npt), _np(toff), _np(pt_off), ppl, target_items,
                               _MIN_TILES_PER_ITEM, _np(work), _np(wbase), _np(nchunk))
        if n2 != n:
            raise N.CctamdError(f"cctamd_plan: inconsistent work count {n2} != {n}")
        dev = device()
        p = Plan(ppl=ppl, nwork=int(n), work=torch.from_numpy(work).to(dev), wbase=torch.from_numpy(wbase).to(dev),
                 nchunk=torch.from_numpy(nchunk).to(dev), design_pt_off=torch.from_numpy(pt_off).to(dev),
                 npt=int(pt_off[-1]))
        self._plans[key] = p
        return p

    def _seg(self, precision: int):
        """(segment rows, tile origins, FP64 rows for near tiles) of a precision."""
        if precision == 64:
            return self.seg64, None, None
        if precision in (32, 33):
            if self.seg32 is None:
                raise ValueError("model was built without FP32 tiles")
            return (self.seg32 if precision == 32 else self.seg32k), self.tile_org, self.seg64
        raise ValueError("precision must be 32, 33 or 64")

    # ------------------------------------------------------------- field
    def field(self, points: numpy.ndarray, design_npt: Sequence[int], precision: int = 64,
              ppl: Optional[int] = None) -> numpy.ndarray:
        """B at points (P, 3) grouped by design (design_npt[d] consecutive rows each); ``ppl``
        defaults to the block shape that fits the largest design's point count."""
        pts = numpy.asarray(points, numpy.float64).reshape(-1, 3)
        if ppl is None:
            ppl = block_shape_for(design_npt, precision)
        plan = self.plan(design_npt, ppl)
        if plan.npt != pts.shape[0]:
            raise ValueError("point count does not match design_npt")
        if plan.npt == 0:
            return numpy.zeros((0, 3))
        dev = device()
        soa = torch.from_numpy(numpy.ascontiguousarray(pts.T)).to(dev)
        out = torch.empty_like(soa)
        partial = torch.empty(max(plan.nwork * 3 * points_per_block(ppl), 1), dtype=torch.float64, device=dev)
        seg, org, near = self._seg(precision)
        N.check(N.lib.cctamd_bs_field(precision, ppl, _p(seg), _p(org), _p(near), _p(plan.work), plan.nwork, _p(plan.wbase),
                                      _p(plan.nchunk), _p(plan.design_pt_off), self.ndesign, _p(self.qs),
                                      _p(self.design_qs_off), _p(self.bconst), _p(soa), plan.npt, _p(partial),
                                      _p(out), _stream()), "cctamd_bs_field")
        return out.cpu().numpy().T.copy()

    # ---------------------------------------------------------- tracking
    def track(self, state: numpy.ndarray, kq: numpy.ndarray, dt: numpy.ndarray, nsteps: numpy.ndarray,
              design_npt: Sequence[int], precision: int = 64, ppl: Optional[int] = None,
              record: bool = False) -> Tuple[numpy.ndarray, Optional[numpy.ndarray]]:
        """RK4 tracking; state (P, 6) rows [x, y, z, vx, vy, vz] grouped by design.

        Returns the final state (P, 6) and, with ``record``, the state after
        every step as (max_steps+1, P, 6) (rows past a particle's own step
        count repeat nothing meaningful and are masked by the caller).  ``ppl``
        defaults to the block shape that fits the largest design's particle count.
        """
        st = numpy.asarray(state, numpy.float64).reshape(-1, 6)
        if ppl is None:
            ppl = block_shape_for(design_npt, precision)
        plan = self.plan(design_npt, ppl)
        if plan.npt != st.shape[0]:
            raise ValueError("particle count does not match design_npt")
        if plan.npt == 0:
            return st.copy(), None
        dev = device()
        ns = numpy.asarray(nsteps, numpy.int64)
        max_steps = int(ns.max()) if ns.size else 0
        soa = torch.from_numpy(numpy.ascontiguousarray(st.T)).to(dev)
        kq_t = torch.from_numpy(numpy.ascontiguousarray(kq, numpy.float64)).to(dev)
        dt_t = torch.from_numpy(numpy.ascontiguousarray(dt, numpy.float64)).to(dev)
        ns_t = torch.from_numpy(numpy.ascontiguousarray(ns)).to(dev)
        scratch = torch.empty((15, plan.npt), dtype=torch.float64, device=dev)
        partial = torch.empty(max(plan.nwork * 3 * points_per_block(ppl), 1), dtype=torch.float64, device=dev)
        rec = torch.empty((max_steps + 1, 6, plan.npt), dtype=torch.float64, device=dev) if record else None
        seg, org, near = self._seg(precision)
        N.check(N.lib.cctamd_track_rk4(precision, ppl, _p(seg), _p(org), _p(near), _p(plan.work), plan.nwork, _p(plan.wbase),
                                       _p(plan.nchunk), _p(plan.design_pt_off), self.ndesign, _p(self.qs),
                                       _p(self.design_qs_off), _p(self.bconst), _p(soa), plan.npt, _p(kq_t),
                                       _p(dt_t), _p(ns_t), max_steps, _p(scratch), _p(partial), _p(rec), 1,
                                       _stream()), "cctamd_track_rk4")
        out = soa.cpu().numpy().T.copy()
        return out, (rec.cpu().numpy().transpose(0, 2, 1).copy() if rec is not None else None)

    # ------------------------------------------- device-resident variants
    def workspace(self, plan: "Plan") -> Tuple[torch.Tensor, torch.Tensor]:
        """(scratch [15][P], partial) device buffers for a plan."""
        dev = device()
        return (torch.empty((15, max(plan.npt, 1)), dtype=torch.float64, device=dev),
                torch.empty(max(plan.nwork * 3 * points_per_block(plan.ppl), 1), dtype=torch.float64, device=dev))

    def track_tensors(self, plan: "Plan", precision: int, state: torch.Tensor, kq: torch.Tensor, dt: torch.Tensor,
                      nsteps: torch.Tensor, max_steps: int, scratch: torch.Tensor, partial: torch.Tensor,
                      record: Optional[torch.Tensor] = None, record_every: int = 1) -> None:
        """In-place RK4 tracking of device-resident SoA state (6, P); nothing crosses PCIe."""
        seg, org, near = self._seg(precision)
        N.check(N.lib.cctamd_track_rk4(precision, plan.ppl, _p(seg), _p(org), _p(near), _p(plan.work), plan.nwork,
                                       _p(plan.wbase), _p(plan.nchunk), _p(plan.design_pt_off), self.ndesign,
                                       _p(self.qs), _p(self.design_qs_off), _p(self.bconst), _p(state), plan.npt,
                                       _p(kq), _p(dt), _p(nsteps), int(max_steps), _p(scratch), _p(partial),
                                       _p(record), int(record_every), _stream()), "cctamd_track_rk4")

    def field_tensors(self, plan: "Plan", precision: int, points: torch.Tensor, partial: torch.Tensor,
                      out: torch.Tensor) -> None:
        """B (3, P) at device points (3, P): Biot-Savart pass + fixed-order reduction + QS + uniform."""
        seg, org, near = self._seg(precision)
        N.check(N.lib.cctamd_bs_field(precision, plan.ppl, _p(seg), _p(org), _p(near), _p(plan.work), plan.nwork,
                                      _p(plan.wbase), _p(plan.nchunk), _p(plan.design_pt_off), self.ndesign,
                                      _p(self.qs), _p(self.design_qs_off), _p(self.bconst), _p(points), plan.npt,
                                      _p(partial), _p(out), _stream()), "cctamd_bs_field")

    def bs_partial_tensors(self, plan: "Plan", precision: int, points: torch.Tensor, partial: torch.Tensor) -> None:
        """One launch of the split-K Biot-Savart kernel on device points (3, P)."""
        seg, org, near = self._seg(precision)
        N.check(N.lib.cctamd_bs_partial(precision, plan.ppl, _p(seg), _p(org), _p(near), _p(plan.work), plan.nwork,
                                        _p(points), plan.npt, _p(partial), _stream()), "cctamd_bs_partial")

    def segments64(self) -> numpy.ndarray:
        """The FP64 segment rows (S, 8) copied to the host (inspection and tests)."""
        return self.seg64.cpu().numpy()[: self.nseg_padded]


class TrackGraph:
    """One tracking pass (every RK4 step of a plan: 8 launches per step) captured
    once as a HIP graph and replayed.

    The graph binds the device buffers given here: refresh ``state`` in place
    (``state.copy_(initial)``) before each replay.  The captured launches hold
    no memset nodes and no host synchronisation.
    """

    def __init__(self, model: DeviceModel, plan: Plan, precision: int, state: torch.Tensor, kq: torch.Tensor,
                 dt: torch.Tensor, nsteps: torch.Tensor, max_steps: int, scratch: torch.Tensor,
                 partial: torch.Tensor):
        self.buffers = (state, kq, dt, nsteps, scratch, partial)  # kept alive with the graph
        self.graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()  # default priority
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.graph(self.graph, stream=side):
            model.track_tensors(plan, precision, state, kq, dt, nsteps, max_steps, scratch, partial)
        torch.cuda.current_stream().wait_stream(side)

    def replay(self) -> None:
        self.graph.replay()


_CACHE: "collections.OrderedDict[bytes, DeviceModel]" = collections.OrderedDict()
_CACHE_SIZE = 8


def model_for(specs: Sequence[DesignSpec], fp32: bool = True, align_cct: bool = False) -> DeviceModel:
    """DeviceModel for the given designs, reused while the magnet parameters are unchanged."""
    h = hashlib.sha1()
    for s in specs:
        h.update(s.key())
    h.update(b"fp32" if fp32 else b"fp64")
    h.update(b"aligned" if align_cct else b"packed")
    k = h.digest()
    m = _CACHE.get(k)
    if m is not None:
        _CACHE.move_to_end(k)
        return m
    m = DeviceModel(specs, fp32=fp32, align_cct=align_cct)
    _CACHE[k] = m
    while len(_CACHE) > _CACHE_SIZE:
        _CACHE.popitem(last=False)
    return m


def model_from_segments(kl: numpy.ndarray, m: numpy.ndarray, fp32: bool = True) -> DeviceModel:
    """A one-design model from raw global-frame current elements kl (n, 3) and midpoints m (n, 3)."""
    spec = DesignSpec(raw_kl=numpy.ascontiguousarray(kl, numpy.float64).reshape(-1, 3),
                      raw_m=numpy.ascontiguousarray(m, numpy.float64).reshape(-1, 3))
    return model_for([spec], fp32=fp32)


def rk4_steps(length: float, footstep: float, speed: float) -> Tuple[int, float]:
    """Step count and time step of one particle, exactly as the reference computes them
    (final_code/packages/particles.py:585-587 and base_utils.py:284-285)."""
    dt = footstep / speed
    t_end = length / speed
    number = math.ceil((t_end - 0.0) / dt)
    return number, (t_end - 0.0) / float(number)


def winding_points(params: numpy.ndarray, npoints: int) -> numpy.ndarray:
    """Local-frame path points of one CCT (npoints, 3) generated on the GPU."""
    dev = device()
    prm = torch.from_numpy(numpy.ascontiguousarray(params, numpy.float64)).to(dev)
    out = torch.empty((max(npoints, 1), 3), dtype=torch.float64, device=dev)
    N.check(N.lib.cctamd_winding_points(_p(prm), npoints, _p(out), _stream()), "cctamd_winding_points")
    return out.cpu().numpy()[:npoints]
