"""Drop-in for the reference's GPU_ACCELERATOR facade (final_code/packages/gpu_accelerator.py:47-2152).

Same constructor and method signatures; every method runs on libcctamd:
``FLOAT64`` selects the plain FP64 path, ``FLOAT32`` the FP32 Biot-Savart
path (per-tile re-centred FP32 terms, FP64 accumulation) with FP64 particle
state.  Tracking follows the reference CPU semantics (per-particle step count
ceil(length/footstep) in time units, distance += length), magnets of any
number per kind, segment sets without a fixed capacity.  ``block_dim_x`` is
validated as before but work shapes come from the planner/autotuner.
"""
from __future__ import annotations

import math
from typing import List, Optional, Tuple

import numpy

from . import engine
from .beamline import Beamline
from .cct import CCT
from .constants import MM, MRAD
from .local_coordinate_system import LocalCoordinateSystem
from .magnets import QS, CombinedMagnet, DeviceParts, Magnet, device_parts
from .particles import ParticleFactory, PhaseSpaceParticle, RunningParticle, track_particles
from .point import P2, P3


class GPU_ACCELERATOR:
    FLOAT32: str = "FLOAT32"
    FLOAT64: str = "FLOAT64"
    QS_DATA_LENGTH = 16

    def __init__(self, float_number_type: str = FLOAT32, block_dim_x: int = 1024,
                 max_current_element_number: int = 2000 * 120, max_qs_datas_length: int = 10000,
                 cpu_mode: bool = False) -> None:
        if block_dim_x > 1024 or block_dim_x < 0:
            raise ValueError(f"block_dim_x must be in [1, 1024], not {block_dim_x}")
        if block_dim_x & (block_dim_x - 1) != 0:
            raise ValueError(f"block_dim_x must be a power of two, not {block_dim_x}")
        if float_number_type == GPU_ACCELERATOR.FLOAT32:
            self.precision, self.numpy_dtype = 32, numpy.float32
        elif float_number_type == GPU_ACCELERATOR.FLOAT64:
            self.precision, self.numpy_dtype = 64, numpy.float64
        else:
            raise ValueError("float_number_type must be GPU_ACCELERATOR.FLOAT32 or GPU_ACCELERATOR.FLOAT64")
        self.float_number_type = float_number_type
        self.block_dim_x = int(block_dim_x)
        self.max_current_element_number = int(max_current_element_number)
        self.max_qs_datas_length = int(max_qs_datas_length)
        self.cpu_mode = bool(cpu_mode)
        if self.cpu_mode:
            # the reference's CPU mode is its FP64 path; here that is the FP64 GPU path
            print("GPU_ACCELERATOR: cpu_mode requested, running the FP64 path")
            self.precision = 64
        engine.device()

    # ------------------------------------------------------------ fields
    def _field(self, magnet_or_parts, p: P3) -> P3:
        parts = magnet_or_parts if isinstance(magnet_or_parts, DeviceParts) else device_parts(magnet_or_parts)
        if parts.foreign:
            raise ValueError(f"magnet {parts.foreign[0]} cannot be accelerated on the GPU")
        model = engine.model_for([parts.spec()], fp32=self.precision == 32)
        return P3(*model.field(numpy.array([p.to_list()]), [1], precision=self.precision)[0])

    def current_element_B(self, kls: numpy.ndarray, p0s: numpy.ndarray, number: int, p: P3) -> P3:
        """Field of raw global-frame current elements (gpu_accelerator.py:1147-1196)."""
        kl = numpy.asarray(kls, numpy.float64).reshape(-1, 3)[:number]
        m = numpy.asarray(p0s, numpy.float64).reshape(-1, 3)[:number]
        model = engine.model_from_segments(kl, m, fp32=self.precision == 32)
        return P3(*model.field(numpy.array([p.to_list()]), [1], precision=self.precision)[0])

    def magnet_at_qs(self, qs_data, p3: P3) -> P3:
        parts = DeviceParts()
        parts.qs.append(numpy.asarray(qs_data, numpy.float64).reshape(16))
        return self._field(parts, p3)

    def magnet_at_qs_date(self, qs_data, p3: P3) -> P3:
        return self.magnet_at_qs(qs_data, p3)

    def magnet_at_qs_dates(self, qss: List[QS], p: P3) -> P3:
        return self._field(CombinedMagnet(*qss), p)

    def magnet_at_beamline_with_single_qs(self, bl: Beamline, p: P3) -> P3:
        return self._field(bl, p)

    def magnet_at_beamline_with_multi_qs(self, bl: Beamline, p: P3) -> P3:
        return self._field(bl, p)

    # ---------------------------------------------------------- tracking
    @property
    def _track_precision(self) -> int:
        # FP32 tracking uses the K x q - C tile sums (precision 33 of the C ABI)
        return 33 if self.precision == 32 else 64

    @staticmethod
    def _device_only(bls) -> None:
        """The reference facade tracks CCTs and QS only and raises ValueError for any other magnet
        (gpu_accelerator.py:1970); ParticleRunner is the path for user magnets."""
        for bl in bls:
            parts = device_parts(bl)
            if parts.foreign:
                raise ValueError(f"magnet {parts.foreign[0]} cannot be accelerated on the GPU")

    def _track_inplace(self, bl: Magnet, ps: List[RunningParticle], distance: float, footstep: float) -> None:
        self._device_only([bl])
        out, _, _, _ = track_particles(ps, [bl], [0] * len(ps), distance, footstep, precision=self._track_precision)
        for q, row in zip(ps, out):
            q.position = P3(*row[:3])
            q.velocity = P3(*row[3:])
            q.distance += distance

    def track_one_particle_with_single_qs(self, bl: Beamline, p: RunningParticle, distance: float,
                                          footstep: float) -> None:
        self._track_inplace(bl, [p], distance, footstep)

    def track_one_particle_with_multi_qs(self, bl: Beamline, p: RunningParticle, distance: float,
                                         footstep: float) -> None:
        self._track_inplace(bl, [p], distance, footstep)

    def track_multi_particle_for_magnet_with_single_qs(self, bl: Beamline, ps: List[RunningParticle],
                                                       distance: float, footstep: float) -> None:
        self._track_inplace(bl, list(ps), distance, footstep)

    def track_multi_particle_for_magnet_with_multi_qs(self, bl: Beamline, ps: List[RunningParticle],
                                                      distance: float, footstep: float) -> None:
        self._track_inplace(bl, list(ps), distance, footstep)

    def track_multi_particle_beamline_for_magnet_with_single_qs(
            self, bls: List[Beamline], ps: List[RunningParticle], distance: float,
            footstep: float) -> List[List[RunningParticle]]:
        return self.track_multi_particle_beamline_for_magnet_with_multi_qs(bls, ps, distance, footstep)

    def track_multi_particle_beamline_for_magnet_with_multi_qs(
            self, bls: List[Beamline], ps: List[RunningParticle], distance: float,
            footstep: float) -> List[List[RunningParticle]]:
        """Every particle through every beamline in one batched launch; returns new particles
        [beamline][particle], inputs untouched (gpu_accelerator.py:1912-2044)."""
        self._device_only(bls)
        allp = [q for _ in bls for q in ps]
        design_of = [d for d in range(len(bls)) for _ in ps]
        out, _, _, _ = track_particles(allp, list(bls), design_of, distance, footstep,
                                       precision=self._track_precision)
        ret = []
        for d in range(len(bls)):
            row = []
            for i, q in enumerate(ps):
                s = out[d * len(ps) + i]
                row.append(RunningParticle(P3(*s[:3]), P3(*s[3:]), q.relativistic_mass, q.e, q.speed,
                                           q.distance + distance))
            ret.append(row)
        return ret

    def track_phase_ellipse_in_multi_beamline(self, beamlines: List[Beamline], x_sigma_mm: float,
                                              xp_sigma_mrad: float, y_sigma_mm: float, yp_sigma_mrad,
                                              delta: float, particle_number: int, kinetic_MeV: float,
                                              s: float = 0.0, length: Optional[float] = None,
                                              footstep: float = 10 * MM,
                                              report: bool = True) -> List[Tuple[List[P2], List[P2]]]:
        """Phase-ellipse edge particles through several beamlines sharing one orbit (gpu_accelerator.py:2047-2152)."""
        if length is None:
            length = beamlines[0].trajectory.get_length() - s
        ip_start = ParticleFactory.create_proton_along(beamlines[0].trajectory, s, kinetic_MeV)
        ip_end = ParticleFactory.create_proton_along(beamlines[0].trajectory, s + length, kinetic_MeV)
        pp_x = PhaseSpaceParticle.phase_space_particles_along_positive_ellipse_in_xxp_plane(
            xMax=x_sigma_mm * MM, xpMax=xp_sigma_mrad * MRAD, delta=delta, number=particle_number)
        pp_y = PhaseSpaceParticle.phase_space_particles_along_positive_ellipse_in_yyp_plane(
            yMax=y_sigma_mm * MM, ypMax=yp_sigma_mrad * MRAD, delta=delta, number=particle_number)
        cs = ip_start.get_natural_coordinate_system()
        all_rp = (ParticleFactory.create_from_phase_space_particles(ip_start, cs, pp_x)
                  + ParticleFactory.create_from_phase_space_particles(ip_start, cs, pp_y))
        ends = self.track_multi_particle_beamline_for_magnet_with_multi_qs(beamlines, all_rp, length, footstep)
        ce = ip_end.get_natural_coordinate_system()
        ret = []
        for row in ends:
            for q in row:
                q.speed = q.velocity.length()
            pps = PhaseSpaceParticle.create_from_running_particles(ip_end, ce, row)
            ret.append((PhaseSpaceParticle.phase_space_particles_project_to_xxp_plane(pps[:particle_number], True),
                        PhaseSpaceParticle.phase_space_particles_project_to_yyp_plane(pps[particle_number:], True)))
        return ret
