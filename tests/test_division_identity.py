"""The identity behind the SPH passes' divisions by pass constants.

For FP32 a and b, (float)((double)a * RN64(1/b)) equals the correctly
rounded FP32 quotient a/b whenever it is normal: an exact quotient of two
floats is never a midpoint between floats and lies at least 2^-50 (relative)
from one, while the double product is within 2^-52 of a/b.  The device code
divides by h, 16 pi h^3 and 16 pi h^4 this way (common.hpp WendlandK); the
GPU test of W and dW/dr checks the kernels themselves bit for bit.  Here the
identity is checked on numpy's IEEE FP32 division, with dividends placed at
and beside the midpoints, where a faulty shortcut would round the other way.
"""
from __future__ import annotations

import numpy as np
import pytest


def _divisors(h):
    h = np.float32(h)
    c3 = np.float32(16.0) * np.float32(np.pi) * h * h * h
    return (h, c3, c3 * h)


@pytest.mark.parametrize("h", [0.052, 0.026, 0.013, 0.0052, 1.0 / 3.0])
def test_double_reciprocal_product_is_the_ieee_quotient(h):
    rng = np.random.default_rng(int(h * 1e6))
    for b in _divisors(h):
        inv = 1.0 / np.float64(b)
        q = (rng.random(400_000) * 2).astype(np.float32)
        q = q[q > 0]
        mids = (q.astype(np.float64) + np.spacing(q).astype(np.float64) / 2) * np.float64(b)
        m32 = mids.astype(np.float32)
        a = np.concatenate([
            (rng.random(400_000) * 64 * float(b)).astype(np.float32),
            m32, np.nextafter(m32, np.float32(0)), np.nextafter(m32, np.float32(np.inf)),
        ])
        ref = a / b
        got = (a.astype(np.float64) * inv).astype(np.float32)
        assert np.all(np.isfinite(ref)) and float(np.abs(ref[ref != 0]).min()) > 1.2e-38
        np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_a_plain_float_reciprocal_product_is_not():
    """The same shortcut in FP32 (a * RN32(1/b)) does differ from a / b, so
    the check above can see a wrong rounding."""
    rng = np.random.default_rng(7)
    b = np.float32(0.026)
    a = (rng.random(200_000) * 0.052).astype(np.float32)
    got = a * (np.float32(1) / b)
    assert int(np.count_nonzero(got != a / b)) > 0
