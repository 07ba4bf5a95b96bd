"""Host side of the batched ensemble (CPU only): struct concatenation with
re-based contact CSR, instance numbering, and the C3 scene generator."""
from __future__ import annotations

import numpy as np
import torch

from warp_sph_gpt_amd.ensemble import Layout, _offsets, concat_structs
from warp_sph_gpt_amd.struct import DEMptl
from warp_sph_gpt_amd.workloads import ensemble_scene


def dem_with_rows(n, rows, E_wall_ids):
    D = DEMptl.allocate(n)
    off = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    D.contact_dem_offset_old = torch.tensor(off)
    D.contact_dem_id_old = torch.tensor([i for r in rows for i in r], dtype=torch.int32)
    D.tang_dem_old = torch.arange(3 * int(off[-1]), dtype=torch.float32).reshape(-1, 3)
    D.contact_bnd_offset_old = torch.tensor([0] * n + [len(E_wall_ids)], dtype=torch.int32)
    D.contact_bnd_id_old = torch.tensor(E_wall_ids, dtype=torch.int32)
    D.tang_bnd_old = torch.ones((len(E_wall_ids), 3))
    return D


def test_concat_rebases_csr_offsets_and_ids():
    a = dem_with_rows(3, [[1], [0, 2], [1]], [5, 6])
    b = dem_with_rows(2, [[], [0]], [0])
    empty = DEMptl.allocate(0)
    lay = {"dem": _offsets([3, 0, 2]), "bnd": [0, 10, 10, 30]}
    D = concat_structs(DEMptl, [a, empty, b], lay)
    assert D.n == 5
    np.testing.assert_array_equal(D.contact_dem_offset_old.numpy(), [0, 1, 3, 4, 4, 5])
    np.testing.assert_array_equal(D.contact_dem_id_old.numpy(), [1, 0, 2, 1, 3])
    np.testing.assert_array_equal(D.contact_bnd_offset_old.numpy(), [0, 0, 0, 2, 2, 3])
    np.testing.assert_array_equal(D.contact_bnd_id_old.numpy(), [5, 6, 10])
    assert D.tang_dem_old.shape == (5, 3)


def test_instance_numbering_is_nondecreasing_ranges():
    lay = Layout({k: _offsets(c) for k, c in
                  {"fluid": [3, 0, 2], "wall": [1, 1, 1], "dem": [0, 2, 0],
                   "dem_wall": [0, 4, 0]}.items()})
    assert lay.n == 3
    np.testing.assert_array_equal(lay.inst_array("fluid", "cpu").numpy(), [0, 0, 0, 2, 2])
    np.testing.assert_array_equal(lay.inst_array("dem", "cpu").numpy(), [1, 1])


def test_ensemble_scene_is_seeded_and_inside_the_tank():
    g1, (P1, B1, D1, W1), p1 = ensemble_scene(20263, dx=0.02, dt=1e-4)
    g2, (P2, B2, D2, W2), p2 = ensemble_scene(20263, dx=0.02, dt=1e-4)
    assert p1 == p2 and torch.equal(P1.pos, P2.pos)
    assert 0.3 <= p1["Lx"] <= 0.7 and 0.3 <= p1["Lz"] <= 0.6
    assert P1.n == round(p1["Lx"] / 0.02) * 20 * round(p1["Lz"] / 0.02)
    if D1 is not None:
        assert torch.equal(D1.pos, D2.pos)
        assert float(D1.radius.min()) >= 0.015 - 1e-7 and float(D1.radius.max()) <= 0.03 + 1e-7
        z = D1.pos[:, 2] - D1.radius
        assert float(z.min()) > p1["Lz"]          # above the free surface
        d = torch.cdist(D1.pos.double(), D1.pos.double())
        rr = (D1.radius[:, None] + D1.radius[None, :]).double()
        d.fill_diagonal_(1e9)
        assert bool((d >= rr - 1e-6).all())       # no overlap
    _, (P3, B3, D3, W3), p3 = ensemble_scene(5, dx=0.02, dt=1e-4, n_dem=0)
    assert D3 is None and p3["n_dem"] == 0
