"""Pins of the oracle against hand-derived answers (CPU only; see pins.py).

* Bucket keys, stable order and bucket bounds of the restated Warp hash grid
  (main.py:66,73,85,87): a table derived by hand for points at +-0.4 and
  +-1.2 cells, past the fold and past the 2^20 origin, on a grid whose dims
  do not divide 2^20.  A floor in place of the truncation toward zero, a
  missing or different origin offset, a missing clamp, or a key with the
  axes transposed changes at least one entry.
* The Tait EOS (kernel/KERNEL_pres.py:21-23, 43-45) against its closed form
  evaluated exactly with rationals: rho = rho0 gives 0 exactly, the clamped
  side gives 0, and [0.98, 1.05] rho0 for gamma = 7 and gamma = 1.
"""
from __future__ import annotations

import numpy as np
import pytest

import pins
from oracle import oracle as O


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_bucket_keys_order_and_bounds_match_the_hand_table(dtype):
    g = O.OracleGrid(*pins.KEY_DIMS, dtype=dtype).build(pins.KEY_POINTS.astype(dtype),
                                                         pins.KEY_WIDTH)
    np.testing.assert_array_equal(g.point_ids, pins.KEY_SORTED_IDS)
    # the oracle keeps keys in sorted order
    np.testing.assert_array_equal(g.point_cells, pins.KEY_OF_POINT[pins.KEY_SORTED_IDS])
    s, e = pins.key_bounds_arrays()
    np.testing.assert_array_equal(g.cell_starts, s)
    np.testing.assert_array_equal(g.cell_ends, e)


class _S:
    def __init__(self, rho):
        self.rho = rho
        self.pres = np.zeros_like(rho)
        self.pos = np.zeros((len(rho), 3), rho.dtype)

    @property
    def n(self):
        return len(self.rho)


@pytest.mark.parametrize("gamma", [7, 1])
def test_tait_eos_matches_the_exact_closed_form(gamma):
    rho = pins.eos_densities()
    exact = pins.eos_exact(rho, gamma)
    s64 = _S(rho.astype(np.float64))
    O.Kernel_pres_sph(s64, pins.EOS_RHO0, pins.EOS_C0, float(gamma))
    pos = exact > 0
    assert pos.any() and (~pos).any()
    assert np.all(s64.pres[~pos] == 0.0)
    # B, the ratio and the power each round once; the difference
    # (rho/rho0)^gamma - 1 keeps the power's rounding, up to gamma units of
    # the ratio's: |error| <= 4 u (p + gamma B) with u the unit roundoff
    B = pins.EOS_RHO0 * pins.EOS_C0 ** 2 / gamma
    err = np.abs(s64.pres - exact)
    assert np.all(err <= 4 * 2.0 ** -53 * (exact + gamma * B)), err.max()
    assert np.max(np.abs(s64.pres - exact)[exact > 0.01 * B] / exact[exact > 0.01 * B]) < 1e-12
    s32 = _S(rho.astype(np.float32))
    O.Kernel_pres_sph(s32, pins.EOS_RHO0, pins.EOS_C0, float(gamma))
    assert np.all(s32.pres[~pos] == 0.0)
    err = np.abs(s32.pres.astype(np.float64) - exact)
    assert np.all(err <= 4 * 2.0 ** -24 * (exact + gamma * B)), err.max()
    # rho = rho0 exactly: zero pressure in both precisions
    assert s64.pres[0] == 0.0 and s32.pres[0] == 0.0


def test_tait_eos_known_values():
    rho = np.array([r for r, _ in pins.EOS_KNOWN])
    s = _S(rho.astype(np.float64))
    O.Kernel_pres_sph(s, pins.EOS_RHO0, pins.EOS_C0, 7.0)
    for (r, p), got in zip(pins.EOS_KNOWN, s.pres):
        assert got == pytest.approx(p, rel=1e-12, abs=0.0), r
