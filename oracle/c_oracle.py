"""ctypes front-end of the C oracle (test infrastructure only).

Wraps ``liboracle_c.so`` built from ``cct_oracle_c.c`` by ``oracle/Makefile``.
The magnet description is the dictionary produced by
``cct_oracle.beamline_arrays``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Dict

import numpy

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle_c.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle_c.so"], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = ctypes.CDLL(_LIB)
        P = ctypes.c_void_p
        I64 = ctypes.c_int64
        _lib.oracle_field.argtypes = [P, I64, P, P, P, P, P, P, P, I64, P, ctypes.c_int]
        _lib.oracle_track.argtypes = [P, I64, P, P, P, P, P, P, P, P, P, P, I64, ctypes.c_int]
        _lib.oracle_threads.restype = ctypes.c_int
    return _lib


def _ptr(a: numpy.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _magnet_args(m: Dict[str, numpy.ndarray]):
    keep = [numpy.ascontiguousarray(m[k]) for k in ("order", "cct_frame", "cct_off", "kl", "mid", "qs", "uni")]
    order, frame, off, kl, mid, qs, uni = keep
    return keep, [_ptr(order), order.shape[0], _ptr(frame), _ptr(off), _ptr(kl), _ptr(mid), _ptr(qs), _ptr(uni)]


def field(m: Dict[str, numpy.ndarray], points: numpy.ndarray, threads: int = 0) -> numpy.ndarray:
    """B (n,3) at global points (n,3)."""
    pts = numpy.ascontiguousarray(points, dtype=numpy.float64).reshape(-1, 3)
    out = numpy.zeros_like(pts)
    keep, args = _magnet_args(m)
    lib().oracle_field(*args, _ptr(pts), pts.shape[0], _ptr(out), int(threads))
    del keep
    return out


def track(m: Dict[str, numpy.ndarray], state: numpy.ndarray, kq: numpy.ndarray, dt: numpy.ndarray,
          nsteps: numpy.ndarray, threads: int = 0) -> numpy.ndarray:
    """RK4-track particles; state (n,6) position+velocity, returns the new state."""
    st = numpy.array(state, dtype=numpy.float64, order="C").reshape(-1, 6)
    kq = numpy.ascontiguousarray(kq, dtype=numpy.float64)
    dt = numpy.ascontiguousarray(dt, dtype=numpy.float64)
    ns = numpy.ascontiguousarray(nsteps, dtype=numpy.int64)
    keep, args = _magnet_args(m)
    lib().oracle_track(*args, _ptr(st), _ptr(kq), _ptr(dt), _ptr(ns), st.shape[0], int(threads))
    del keep
    return st


def max_threads() -> int:
    return int(lib().oracle_threads())
