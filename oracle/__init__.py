"""CPU oracle (test infrastructure only): numpy restatement and C restatement of the
reference's FP64 field + RK4 tracking path. Not imported by the cctpy_amd product."""
