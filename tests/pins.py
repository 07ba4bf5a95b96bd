"""Hand-derived known answers shared by the oracle pins (CPU) and the HIP
pins (GPU).  Nothing here is computed with the formulas under test.

Bucket keys (Warp's HashGrid convention as restated for main.py:66,73,85,87):
cell width w = 0.25 (1/w = 4 exactly in FP32), grid dims (3, 5, 7), so the
2^20 origin does not vanish in the fold: 2^20 = 1048576 leaves 1 mod 3,
1 mod 5 and 4 mod 7.  Per axis i = trunc(p * 4), c = (i + 2^20) mod dim
(0 if i + 2^20 < 0), key = cz*15 + cy*3 + cx.

  A ( 0.1,   0.1,   0.1)  i = ( 0,  0,  0)  c = (1, 1, 4)  key 60+3+1 = 64
  B (-0.1,  -0.1,  -0.1)  trunc(-0.4) = 0   c = (1, 1, 4)  key 64
                          (floor would give i = -1, c = (0, 0, 3), key 45)
  C ( 0.3,   0.55,  1.9)  i = ( 1,  2,  7)  c = (2, 3, 4)  key 60+9+2 = 71
  D (-0.3,  -0.55, -1.9)  i = (-1, -2, -7)  c = (0, 4, 4)  key 60+12+0 = 72
                          (floor: i = (-2, -3, -8), c = (2, 3, 3), key 56)
  E ( 2.0,   0.0,   0.0)  i = ( 8,  0,  0)  c = (0, 1, 4)  key 63 (x folds)
  F (-3e5,   0.0,   0.0)  i + 2^20 = -151424 < 0, c = (0, 1, 4), key 63

Points in the order [C, A, D, B, E, F] (indices 0..5) sort stably to
[E(4), F(5), A(1), B(3), C(0), D(2)]; buckets 63: [0, 2), 64: [2, 4),
71: [4, 5), 72: [5, 6), all others empty.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np

KEY_DIMS = (3, 5, 7)
KEY_WIDTH = 0.25
KEY_POINTS = np.array([[0.3, 0.55, 1.9], [0.1, 0.1, 0.1], [-0.3, -0.55, -1.9],
                       [-0.1, -0.1, -0.1], [2.0, 0.0, 0.0], [-3.0e5, 0.0, 0.0]], np.float32)
KEY_OF_POINT = np.array([71, 64, 72, 64, 63, 63], np.int32)
KEY_SORTED_IDS = np.array([4, 5, 1, 3, 0, 2], np.int32)
KEY_BOUNDS = {63: (0, 2), 64: (2, 4), 71: (4, 5), 72: (5, 6)}


def key_bounds_arrays():
    n = KEY_DIMS[0] * KEY_DIMS[1] * KEY_DIMS[2]
    s, e = np.zeros(n, np.int32), np.zeros(n, np.int32)
    for c, (a, b) in KEY_BOUNDS.items():
        s[c], e[c] = a, b
    return s, e


# Tait EOS, kernel/KERNEL_pres.py:21-23: p = B max((rho/rho0)^gamma - 1, 0),
# B = rho0 c0^2 / gamma.  Densities are exact binary fractions, so the
# closed form is evaluated exactly with rationals (integer gamma).
EOS_RHO0, EOS_C0 = 1000.0, 31.3209


def eos_densities() -> np.ndarray:
    """rho0 exactly, the clamped side below it, and [0.98, 1.05] rho0 on a
    fine grid of exact binary fractions."""
    grid = 1000.0 * (0.98 + np.arange(0, 1793) * (0.07 / 1792))
    grid = (np.round(grid * 1024.0) / 1024.0)
    return np.concatenate([[1000.0, 999.0, 980.0, 1000.0 + 1.0 / 1024.0, 1050.0], grid])


def eos_exact(rho: np.ndarray, gamma: int, rho0=EOS_RHO0, c0=EOS_C0) -> np.ndarray:
    """Closed form with exact rationals (rho and rho0 as given in FP64)."""
    B = Fraction(rho0) * Fraction(c0) ** 2 / gamma
    out = []
    for r in rho:
        q = (Fraction(float(r)) / Fraction(rho0)) ** gamma - 1
        out.append(float(B * max(q, Fraction(0))))
    return np.array(out)


# Known values of the closed form at gamma = 7, written out by hand:
# 1.01^7 = 1.0721353521070100 (binomial sum), so p(1010) = B * 0.07213535210701
# with B = 1000 * 31.3209^2 / 7 = 140143.7...; p(rho0) = 0; p(999) = 0.
EOS_KNOWN = [(1000.0, 0.0), (999.0, 0.0), (1010.0, 1000.0 * 31.3209 ** 2 / 7 * 0.07213535210701)]
