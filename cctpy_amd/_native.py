"""ctypes binding of the C ABI declared in include/cctamd.h.

The library is the only compute path of this package: if it is missing the
import of this module raises, there is no fallback.
"""
from __future__ import annotations

import ctypes
import os

from ._build import LIB

P = ctypes.c_void_p
I64 = ctypes.c_int64
INT = ctypes.c_int

TILE = 32
NOALLOC_PARTIAL_DOUBLES = 2 * 1024 * 1024
CCT_PARAMS = 32
MAX_TILTS = 8
WORK_INTS = 8


class CctamdError(RuntimeError):
    pass


def _load() -> ctypes.CDLL:
    # torch brings its own HIP runtime; loading it first makes the library's
    # libamdhip64/libhsa-runtime64 dependencies resolve (by soname) to that same
    # runtime, so device pointers and streams are shared by one runtime.
    import torch  # noqa: F401
    if not os.path.exists(LIB):
        raise ImportError(
            f"cctpy_amd: native library {LIB} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(LIB)
    lib.cctamd_version.restype = INT
    lib.cctamd_error_string.restype = ctypes.c_char_p
    lib.cctamd_error_string.argtypes = [INT]
    lib.cctamd_winding_generate.argtypes = [P, P, P, I64, I64, P, P]
    lib.cctamd_winding_generate.restype = INT
    lib.cctamd_winding_points.argtypes = [P, I64, P, P]
    lib.cctamd_winding_points.restype = INT
    lib.cctamd_pack_fp32.argtypes = [P, I64, P, P, P, P]
    lib.cctamd_pack_fp32.restype = INT
    lib.cctamd_plan.argtypes = [I64, P, P, P, P, INT, I64, I64, P, P, P]
    lib.cctamd_plan.restype = I64
    lib.cctamd_bs_partial.argtypes = [INT, INT, P, P, P, P, I64, P, I64, P, P]
    lib.cctamd_bs_partial.restype = INT
    lib.cctamd_bs_field.argtypes = [INT, INT, P, P, P, P, I64, P, P, P, I64, P, P, P, P, I64, P, P, P]
    lib.cctamd_bs_field.restype = INT
    lib.cctamd_bs_field_noalloc.argtypes = [INT, INT, P, P, P, P, I64, P, P, P, I64, P, P, P, P, I64, P, P]
    lib.cctamd_bs_field_noalloc.restype = INT
    lib.cctamd_track_rk4.argtypes = [INT, INT, P, P, P, P, I64, P, P, P, I64, P, P, P, P, I64, P, P, P, I64, P, P,
                                     P, I64, P]
    lib.cctamd_track_rk4.restype = INT
    lib.cctamd_winding_tangents.argtypes = [P, P, P, I64, I64, P, P]
    lib.cctamd_winding_tangents.restype = INT
    lib.cctamd_track_rk4_adjoint.argtypes = [INT, INT, P, P, P, P, P, P, P, I64, P, I64, P, P, P, I64, P, P, P,
                                             P, I64, I64, P, P, P, I64, P, P, P, P, P, P, P, P, P, P, P]
    lib.cctamd_track_rk4_adjoint.restype = INT
    lib.cctamd_track_rk4_taped.argtypes = [INT, INT, P, P, P, P, I64, P, P, P, I64, P, P, P, P, I64, P, P, P, I64,
                                           P, P, P, P]
    lib.cctamd_track_rk4_taped.restype = INT
    lib.cctamd_phase_space.argtypes = [P, I64, P, P, P, I64, P, P, I64, P, P, P]
    lib.cctamd_phase_space.restype = INT
    lib.cctamd_beam_objective_seed.argtypes = [P, I64, P, P, I64, P, P, I64, P, P, P]
    lib.cctamd_beam_objective_seed.restype = INT
    return lib


lib = _load()


def loaded_runtimes() -> list:
    """Paths of the HIP runtime libraries mapped in this process."""
    with open("/proc/self/maps") as f:
        return sorted({line.split()[-1] for line in f if "libamdhip64" in line})


def check(code: int, what: str) -> None:
    if code != 0:
        msg = lib.cctamd_error_string(int(code)).decode()
        raise CctamdError(f"{what}: {msg} (code {code})")


EXPORTED = (
    "cctamd_version", "cctamd_error_string", "cctamd_winding_generate", "cctamd_winding_points",
    "cctamd_pack_fp32", "cctamd_plan", "cctamd_bs_partial", "cctamd_bs_field", "cctamd_track_rk4",
    "cctamd_winding_tangents", "cctamd_track_rk4_taped", "cctamd_track_rk4_adjoint", "cctamd_phase_space",
    "cctamd_beam_objective_seed", "cctamd_bs_field_noalloc",
)
