"""Parametric HUST superconducting gantry: two bending sections built from
drifts, QS magnets and AG-CCT pairs (final_code/hust_sc_gantry.py:18-385)."""
from __future__ import annotations

from typing import Dict

from .beamline import Beamline
from .constants import MM
from .point import P2

# Constructor defaults of final_code/hust_sc_gantry.py:19-89.
DEFAULTS: Dict[str, object] = dict(
    DL1=0.9007765, GAP1=0.4301517, GAP2=0.370816,
    qs1_length=0.2340128, qs1_aperture_radius=60 * MM, qs1_gradient=3.1, qs1_second_gradient=7.9,
    qs2_length=0.200139, qs2_aperture_radius=60 * MM, qs2_gradient=3.8, qs2_second_gradient=-34.5,
    cct12_big_r=0.95,
    agcct12_inner_small_r=72.5 * MM, agcct12_outer_small_r=88.5 * MM,
    dicct12_inner_small_r=104.5 * MM, dicct12_outer_small_r=120.5 * MM,
    agcct1_winding_number=21, agcct2_winding_number=21, dicct12_winding_number=42,
    dicct12_bending_angle=22.5, agcct1_bending_angle=22.5 / 2, agcct2_bending_angle=22.5 / 2,
    dicct12_tilt_angles=[30, 102.8, 95.6, 87.3], agcct12_tilt_angles=[93.5, 30, 78.3, 97.5],
    dicct12_current=-9780.7, agcct12_current=9736.8,
    DL2=2.1162209, GAP3=0.1978111,
    qs3_length=0.2382791, qs3_aperture_radius=60 * MM, qs3_gradient=-7.3733, qs3_second_gradient=-45.31 * 2,
    cct345_big_r=0.95,
    agcct345_inner_small_r=83 * MM, agcct345_outer_small_r=83 * MM + 15 * MM,
    dicct345_inner_small_r=83 * MM + 15 * MM * 2, dicct345_outer_small_r=83 * MM + 15 * MM * 3,
    agcct3_winding_number=21, agcct4_winding_number=50, agcct5_winding_number=50, dicct345_winding_number=128,
    dicct345_bending_angle=-67.5,
    agcct3_bending_angle=-(8 + 3.716404), agcct4_bending_angle=-(8 + 19.93897), agcct5_bending_angle=-(8 + 19.844626),
    dicct345_tilt_angles=[30, 80], agcct345_tilt_angles=[90, 30],
    dicct345_current=9664, agcct345_current=-6000,
    part_per_winding=120,
)


class HUST_SC_GANTRY:
    def __init__(self, **kwargs) -> None:
        for k in kwargs:
            if k not in DEFAULTS:
                raise TypeError(f"HUST_SC_GANTRY got an unexpected keyword argument '{k}'")
        for k, v in DEFAULTS.items():
            setattr(self, k, kwargs.get(k, v))
        self._total = self._first = self._second = None

    def _first_half(self, bl: Beamline) -> Beamline:
        r = [self.dicct12_outer_small_r, self.dicct12_inner_small_r, self.agcct12_outer_small_r,
             self.agcct12_inner_small_r]
        tilts = [self.dicct12_tilt_angles, self.agcct12_tilt_angles]
        cur = [self.dicct12_current, self.agcct12_current]
        a1, a2 = self.agcct1_bending_angle, self.agcct2_bending_angle
        w1, w2, wd = self.agcct1_winding_number, self.agcct2_winding_number, self.dicct12_winding_number
        n = self.part_per_winding
        return (bl.append_agcct(self.cct12_big_r, r, [a1, a2], tilts, [[wd], [w1, w2]], cur, n)
                .append_drift(self.GAP1)
                .append_qs(self.qs1_length, self.qs1_gradient, self.qs1_second_gradient, self.qs1_aperture_radius)
                .append_drift(self.GAP2)
                .append_qs(self.qs2_length, self.qs2_gradient, self.qs2_second_gradient, self.qs2_aperture_radius)
                .append_drift(self.GAP2)
                .append_qs(self.qs1_length, self.qs1_gradient, self.qs1_second_gradient, self.qs1_aperture_radius)
                .append_drift(self.GAP1)
                .append_agcct(self.cct12_big_r, r, [a2, a1], tilts, [[wd], [w2, w1]], cur, n)
                .append_drift(self.DL1))

    def _second_half(self, bl: Beamline) -> Beamline:
        r = [self.dicct345_outer_small_r, self.dicct345_inner_small_r, self.agcct345_outer_small_r,
             self.agcct345_inner_small_r]
        tilts = [self.dicct345_tilt_angles, self.agcct345_tilt_angles]
        cur = [self.dicct345_current, self.agcct345_current]
        a3, a4, a5 = self.agcct3_bending_angle, self.agcct4_bending_angle, self.agcct5_bending_angle
        w3, w4, w5 = self.agcct3_winding_number, self.agcct4_winding_number, self.agcct5_winding_number
        wd = self.dicct345_winding_number
        n = self.part_per_winding
        return (bl.append_agcct(self.cct345_big_r, r, [a3, a4, a5], tilts, [[wd], [w3, w4, w5]], cur, n)
                .append_drift(self.GAP3)
                .append_qs(self.qs3_length, self.qs3_gradient, self.qs3_second_gradient, self.qs3_aperture_radius)
                .append_drift(self.GAP3)
                .append_agcct(self.cct345_big_r, r, [a5, a4, a3], tilts, [[wd], [w5, w4, w3]], cur, n)
                .append_drift(self.DL2))

    def create_first_bending_part_beamline(self) -> Beamline:
        if self._first is None:
            self._first = self._first_half(
                Beamline.set_start_point(P2.origin()).first_drift(direct=P2.x_direct(), length=self.DL1))
        return self._first

    def create_second_bending_part_beamline(self) -> Beamline:
        if self._second is None:
            first = self.create_first_bending_part_beamline()
            self._second = self._second_half(
                Beamline.set_start_point(first.point_at_end()).first_drift(direct=first.direct_at_end(),
                                                                          length=self.DL2))
        return self._second

    def create_total_beamline(self) -> Beamline:
        if self._total is None:
            bl = self._first_half(Beamline.set_start_point(P2.origin()).first_drift(direct=P2.x_direct(),
                                                                                   length=self.DL1))
            self._total = self._second_half(bl.append_drift(self.DL2))
        return self._total
