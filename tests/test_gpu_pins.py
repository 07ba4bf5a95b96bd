"""The HIP hash grid and EOS against the hand-derived answers of pins.py
(the same table and closed form that pin the oracle)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import pins
from helpers import np_
from warp_sph_gpt_amd import kernels as K
from warp_sph_gpt_amd.hashgrid import HashGrid
from warp_sph_gpt_amd.struct import SPHptl

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
def test_hip_bucket_keys_order_and_bounds_match_the_hand_table(dtype):
    pts = torch.tensor(pins.KEY_POINTS, dtype=dtype, device=DEV)
    g = HashGrid(*pins.KEY_DIMS, device=DEV).build(pts, pins.KEY_WIDTH)
    np.testing.assert_array_equal(np_(g.point_cells), pins.KEY_OF_POINT)
    np.testing.assert_array_equal(np_(g.point_ids), pins.KEY_SORTED_IDS)
    s, e = g.warp_bounds()
    ref_s, ref_e = pins.key_bounds_arrays()
    np.testing.assert_array_equal(np_(s), ref_s)
    np.testing.assert_array_equal(np_(e), ref_e)


@pytest.mark.parametrize("gamma", [7, 1])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
def test_hip_tait_eos_matches_the_exact_closed_form(gamma, dtype):
    rho = pins.eos_densities()
    exact = pins.eos_exact(rho, gamma)
    P = SPHptl.allocate(len(rho), device=DEV, dtype=dtype)
    P.rho.copy_(torch.tensor(rho, dtype=dtype))
    K.Kernel_pres_sph(P, pins.EOS_RHO0, pins.EOS_C0, float(gamma))
    got = np_(P.pres).astype(np.float64)
    u = 2.0 ** -24 if dtype == torch.float32 else 2.0 ** -53
    B = pins.EOS_RHO0 * pins.EOS_C0 ** 2 / gamma
    assert np.all(got[exact == 0] == 0.0)
    assert got[0] == 0.0                  # rho = rho0 exactly
    err = np.abs(got - exact)
    assert np.all(err <= 4 * u * (exact + gamma * B)), err.max()
