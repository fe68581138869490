"""Units and physical constants (values of final_code/packages/constants.py:9-17)."""
from typing import TypeVar

M: float = 1.0
MM: float = 0.001
LIGHT_SPEED: float = 299792458.0 * M
RAD: float = 1.0
MRAD: float = 1.0 * MM * RAD
J: float = 1.0
eV = 1.6021766208e-19 * J
MeV = 1000 * 1000 * eV
MeV_PER_C = 5.3442857792e-22

T = TypeVar("T")
V = TypeVar("V")
