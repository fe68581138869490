"""Shared helpers for the test-suite (builds the native library when absent)."""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def ensure_built():
    # cctpy_amd/_build.py from its file: importing the package loads the library
    import importlib.util
    spec = importlib.util.spec_from_file_location("cctpy_amd_build", os.path.join(ROOT, "cctpy_amd", "_build.py"))
    _build = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(_build)
    if not os.path.exists(_build.LIB):
        _build.build()
    from oracle import c_oracle
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle_c.so")):
        c_oracle.build()


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
