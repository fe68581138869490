"""Parameter sets the reference's demo outputs were produced with (data only)."""

MM = 0.001

# Section-1 parameters of the HUST gantry used for the printed demo outputs
# (final_code/demos/A15束线beamline示例.py:176-208; the same drift/QS values are
# the defaults at cctpy3/cctpy.py:4856-4888). Section 2 uses the defaults of
# final_code/hust_sc_gantry.py:55-89.
SECTION1_DEMO = dict(
    DL1=0.8001322, GAP1=0.1765959, GAP2=0.2960518,
    qs1_length=0.2997797, qs1_aperture_radius=30 * MM, qs1_gradient=28.33, qs1_second_gradient=-140.44 * 2.0,
    qs2_length=0.2585548, qs2_aperture_radius=30 * MM, qs2_gradient=-12.12, qs2_second_gradient=316.22 * 2.0,
    cct12_big_r=0.95,
    agcct12_inner_small_r=35 * MM, agcct12_outer_small_r=35 * MM + 15 * MM,
    dicct12_inner_small_r=35 * MM + 15 * MM * 2, dicct12_outer_small_r=35 * MM + 15 * MM * 3,
    agcct1_winding_number=30, agcct2_winding_number=39, dicct12_winding_number=71,
    dicct12_bending_angle=22.5, agcct1_bending_angle=9.782608695652174, agcct2_bending_angle=12.717391304347826,
    dicct12_tilt_angles=[30, 80], agcct12_tilt_angles=[90, 30],
    dicct12_current=-6192, agcct12_current=-3319,
)
