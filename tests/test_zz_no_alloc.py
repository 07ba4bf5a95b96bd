"""sphdem_exclusive_scan_noalloc (the scan with library-held tile totals)
against sphdem_exclusive_scan (the caller's scratch) through the C-ABI:
two calls in a row, the second smaller, bit-exact with each other and with
numpy's prefix sum; a length beyond SPHDEM_SCAN_NOALLOC_MAX is rejected
before any work is enqueued."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from warp_sph_gpt_amd import _lib

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
NOALLOC_MAX = 1 << 26   # SPHDEM_SCAN_NOALLOC_MAX in include/sphdem.h


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def test_noalloc_scan_matches_the_scratch_scan_on_two_calls():
    lib = _lib.load()
    rng = np.random.default_rng(17)
    # the second call is smaller and leaves the first call's tile totals in
    # the shared array beyond its own tiles
    for n in (3_000_001, 70_001, 2048, 1):
        a = rng.integers(0, 60, n).astype(np.int32)
        ref = np.concatenate([[0], np.cumsum(a, dtype=np.int64)[:-1]]).astype(np.int32)
        t_scratch = torch.tensor(a, device=DEV)
        t_noalloc = t_scratch.clone()
        scratch = torch.zeros(lib.sphdem_scan_scratch(n), dtype=torch.int32, device=DEV)
        assert lib.sphdem_exclusive_scan(ctypes.c_void_p(t_scratch.data_ptr()), n,
                                         ctypes.c_void_p(scratch.data_ptr()), _stream()) == 0
        assert lib.sphdem_exclusive_scan_noalloc(ctypes.c_void_p(t_noalloc.data_ptr()), n,
                                                 _stream()) == 0
        torch.cuda.synchronize()
        np.testing.assert_array_equal(t_noalloc.cpu().numpy(), t_scratch.cpu().numpy())
        np.testing.assert_array_equal(t_noalloc.cpu().numpy(), ref)


def test_noalloc_scan_rejects_lengths_beyond_its_maximum():
    lib = _lib.load()
    t = torch.arange(8, dtype=torch.int32, device=DEV)
    rc = lib.sphdem_exclusive_scan_noalloc(ctypes.c_void_p(t.data_ptr()), NOALLOC_MAX + 1,
                                           _stream())
    assert rc == _lib.SPHDEM_E_ARG
    assert b"SPHDEM_SCAN_NOALLOC_MAX" in lib.sphdem_last_error()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(t.cpu().numpy(), np.arange(8))   # untouched
    assert lib.sphdem_exclusive_scan_noalloc(ctypes.c_void_p(t.data_ptr()), 0, _stream()) == 0


def test_noalloc_scan_after_device_init_and_at_the_maximum_tile_count():
    """sphdem_device_init resolves the array first; a scan of exactly
    SPHDEM_SCAN_NOALLOC_MAX entries uses every tile total it holds."""
    lib = _lib.load()
    assert lib.sphdem_device_init(_stream()) == 0
    n = NOALLOC_MAX
    t = torch.ones(n, dtype=torch.int32, device=DEV)
    assert lib.sphdem_exclusive_scan_noalloc(ctypes.c_void_p(t.data_ptr()), n, _stream()) == 0
    torch.cuda.synchronize()
    assert torch.equal(t, torch.arange(n, dtype=torch.int32, device=DEV))
