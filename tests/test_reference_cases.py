"""Known-answer cases of the reference's own tests, on both backends.

Each case rebuilds the physical setup of a test in the reference's
tests/test_dem.py, tests/test_3d.py or tests/test_input_reader.py (cited
per test) and checks the same closed-form answers with the same
tolerances.  `oracle` runs the CPU restatement (pins the oracle);
`hip` runs libsphdem on the GPU through the package's reference-shaped API.
"""
from __future__ import annotations

import numpy as np
import pytest

from backends import HipBackend, OracleBackend

BACKENDS = [pytest.param("oracle", id="oracle"),
            pytest.param("hip", id="hip", marks=pytest.mark.gpu)]


@pytest.fixture(params=BACKENDS)
def bk(request):
    return OracleBackend() if request.param == "oracle" else HipBackend()


PAIR = np.array([[0.200, 0.16, 0.60], [0.249, 0.16, 0.60], [0.600, 0.16, 0.60]], np.float32)
PAIR_VEL = np.array([[0, 0, 0], [0, 0.01, 0], [0, 0, 0]], np.float32)


def history(bk, D, kind, a, b):
    off = bk.np(getattr(D, f"contact_{kind}_offset_old"))
    ids = bk.np(getattr(D, f"contact_{kind}_id_old"))
    tang = bk.np(getattr(D, f"tang_{kind}_old")).reshape(-1, 3)
    for k in range(int(off[a]), int(off[a + 1])):
        if int(ids[k]) == b:
            return tang[k]
    raise KeyError((kind, a, b))


def active(bk, D, kind):
    return int(bk.np(getattr(D, f"contact_{kind}_offset_old"))[-1])


def rebuild_dem(bk, S):
    bk.build(S.grids[2], S.D.pos, S.gen.dem_support)


# ------------------------------------------------------------------ kernel function


def test_wendland_unit_integral_slope_and_ends(bk):
    """tests/test_3d.py:36-53, plus the expanded polynomial form in FP64."""
    h = 0.08
    r = np.linspace(0.0, 2.0 * h, 4001, dtype=np.float32)
    w, dw = bk.wendland(r, h)
    rd = r.astype(np.float64)
    integrand = 4.0 * np.pi * rd ** 2 * w
    integral = np.sum(0.5 * (integrand[:-1] + integrand[1:]) * np.diff(rd))
    assert abs(integral - 1.0) < 5e-7
    numerical = np.gradient(w.astype(np.float64), rd)
    np.testing.assert_allclose(dw[50:-50:100], numerical[50:-50:100], rtol=0.002)
    assert w[-1] == 0.0 and dw[0] == 0.0 and dw[-1] == 0.0
    # W = C (1 - 5/2 q^2 + 5/2 q^3 - 15/16 q^4 + 1/8 q^5), dW/dr = C/h (-5 q + 15/2 q^2
    # - 15/4 q^3 + 5/8 q^4), C = 21 / (16 pi h^3): the expanded product.
    q = rd / h
    C = 21.0 / (16.0 * np.pi * h ** 3)
    w_ref = C * (1 - 2.5 * q ** 2 + 2.5 * q ** 3 - 0.9375 * q ** 4 + 0.125 * q ** 5)
    dw_ref = C / h * (-5 * q + 7.5 * q ** 2 - 3.75 * q ** 3 + 0.625 * q ** 4)
    np.testing.assert_allclose(w, w_ref, rtol=2e-5, atol=C * 2e-6)
    np.testing.assert_allclose(dw, dw_ref, rtol=2e-5, atol=C / h * 2e-6)


def test_drag_branches_and_stokes_limit(bk):
    """tests/test_dem.py:494-513."""
    eps = np.array([0.5, 0.8, 0.9, 1.0, 1.0], np.float32)
    speed = np.array([0.2, 0.1, 0.2, 2.0, 0.0], np.float32)
    got = bk.drag_beta(eps, speed)
    expected = []
    for e, v in zip(eps.astype(float), speed.astype(float)):
        if e <= 0.80000002:
            expected.append(150 * (1 - e) / e * 0.05 / 0.05 ** 2 + 1.75 * 1000 * v / 0.05)
        elif v == 0:
            expected.append(18 * 0.05 / 0.05 ** 2)
        else:
            re = 1000 * v
            cd = 24 / re * (1 + 0.15 * re ** 0.687) if re <= 1000 else 0.44
            expected.append(0.75 * cd * 1000 * v / 0.05 * e ** -1.65)
    np.testing.assert_allclose(got, expected, rtol=2e-6)


# ------------------------------------------------------------------ gravity / initial fields


def test_gravity_points_along_negative_z(bk):
    """tests/test_3d.py:87-102."""
    S = bk.scene()
    s, g = S.gen, S.grids
    bk.K.Kernel_force_sph(S.P, S.B, g[0], g[1], s.support, s.h, s.mu, s.g)
    acc = bk.np(S.P.acc)
    expected = np.zeros_like(acc)
    expected[:, 2] = -s.g
    np.testing.assert_allclose(acc, expected, rtol=1e-7)
    bk.run_force_dem(S.D, g[2], s.dem_radius, s.g, s.dt)
    f = bk.np(S.D.force)
    ef = np.zeros_like(f)
    ef[:, 2] = -s.dem_mass * s.g
    np.testing.assert_allclose(f, ef, rtol=1e-7)


def test_coupling_fields_initialised(bk):
    """tests/test_dem.py:160-177 and :179-199: porosity 1, pgf/acc_dem 0,
    empty CSR sides with zero offsets."""
    S = bk.scene()
    np.testing.assert_array_equal(bk.np(S.P.porosity), 1.0)
    np.testing.assert_array_equal(bk.np(S.P.pgf), 0.0)
    np.testing.assert_array_equal(bk.np(S.P.acc_dem), 0.0)
    for kind in ("dem", "bnd"):
        for side in ("old", "new"):
            np.testing.assert_array_equal(bk.np(getattr(S.D, f"contact_{kind}_offset_{side}")), 0)
            assert bk.np(getattr(S.D, f"contact_{kind}_id_{side}")).shape == (0,)


# ------------------------------------------------------------------ CSR lifecycle


def test_compact_rows_hold_every_directed_id(bk):
    """tests/test_dem.py:201-218: three coincident spheres, six directed edges."""
    S = bk.scene(dem_nx=3)
    bk.set(S.D, "pos", [0.3, 0.16, 0.6])
    rebuild_dem(bk, S)
    count = bk.run_force_dem(S.D, S.grids[2], S.gen.dem_radius, 0.0, S.gen.dt)
    off = bk.np(S.D.contact_dem_offset_old)
    ids = bk.np(S.D.contact_dem_id_old)
    assert count == 6 and int(off[-1]) == 6 and ids.shape == (6,)
    np.testing.assert_array_equal(np.diff(off), [2, 2, 2])
    for a in range(3):
        assert set(map(int, ids[off[a]:off[a + 1]])) == {0, 1, 2} - {a}


def test_csr_grows_shrinks_and_recontact_starts_fresh(bk):
    """tests/test_dem.py:220-286: E = 0, 2, 6, 2, 0, 2 with survivor history."""
    S = bk.scene(dem_nx=3)
    s, D, g = S.gen, S.D, S.grids
    run = lambda: bk.run_force_dem(D, g[2], s.dem_radius, 0.0, s.dt)  # noqa: E731
    assert run() == 0
    bk.set(D, "pos", PAIR)
    bk.set(D, "vel", PAIR_VEL)
    rebuild_dem(bk, S)
    assert run() == 2
    first = history(bk, D, "dem", 0, 1).copy()
    bk.set(D, "pos", [0.3, 0.16, 0.6])
    rebuild_dem(bk, S)
    assert run() == 6
    bk.set(D, "pos", PAIR)
    rebuild_dem(bk, S)
    assert run() == 2
    np.testing.assert_allclose(history(bk, D, "dem", 0, 1), 3.0 * first, rtol=1e-6, atol=1e-9)
    bk.set(D, "pos", [[0.1, 0.16, 0.6], [0.4, 0.16, 0.6], [0.7, 0.16, 0.6]])
    rebuild_dem(bk, S)
    assert run() == 0
    bk.set(D, "pos", PAIR)
    rebuild_dem(bk, S)
    assert run() == 2
    np.testing.assert_allclose(history(bk, D, "dem", 0, 1), first, rtol=1e-6, atol=1e-9)


def test_coupled_step_keeps_history_across_steps(bk):
    """tests/test_dem.py:288-318."""
    S = bk.scene(dem_nx=3, g=0.0)
    s, D, g = S.gen, S.D, S.grids
    solv = S.solv
    solv.g = 0.0
    bk.set(D, "pos", PAIR)
    bk.set(D, "vel", PAIR_VEL)
    rebuild_dem(bk, S)
    bk.SPHDEM_OneStep(solv, S.P, S.B, D, S.W, *g, 0)
    assert active(bk, D, "dem") == 2
    first = history(bk, D, "dem", 0, 1).copy()
    bk.set(D, "pos", PAIR)
    bk.set(D, "vel", PAIR_VEL)
    bk.set(D, "omega", [0.0, 0.0, 0.0])
    bk.build(g[0], S.P.pos, s.support)
    rebuild_dem(bk, S)
    bk.SPHDEM_OneStep(solv, S.P, S.B, D, S.W, *g, 1)
    assert active(bk, D, "dem") == 2
    np.testing.assert_allclose(history(bk, D, "dem", 0, 1), 2.0 * first, rtol=1e-6, atol=1e-9)


def test_row_replaces_one_neighbour_and_keeps_survivor(bk):
    """tests/test_dem.py:320-356."""
    S = bk.scene(dem_nx=4)
    s, D, g = S.gen, S.D, S.grids
    p1 = np.array([[0.300, 0.16, 0.600], [0.349, 0.16, 0.600], [0.251, 0.16, 0.600],
                   [0.600, 0.16, 0.700]], np.float32)
    bk.set(D, "pos", p1)
    bk.set(D, "vel", [[0, 0, 0], [0, 0, 0], [0, 0.01, 0], [0.01, 0, 0]])
    rebuild_dem(bk, S)
    assert bk.run_force_dem(D, g[2], s.dem_radius, 0.0, s.dt) == 4
    first = history(bk, D, "dem", 0, 2).copy()
    off, ids = bk.np(D.contact_dem_offset_old), bk.np(D.contact_dem_id_old)
    assert set(ids[off[0]:off[1]]) == {1, 2}
    p2 = p1.copy()
    p2[1] = [0.600, 0.16, 0.600]
    p2[3] = [0.300, 0.16, 0.649]
    bk.set(D, "pos", p2)
    rebuild_dem(bk, S)
    assert bk.run_force_dem(D, g[2], s.dem_radius, 0.0, s.dt) == 4
    off, ids = bk.np(D.contact_dem_offset_old), bk.np(D.contact_dem_id_old)
    row = ids[off[0]:off[1]]
    assert set(row) == {2, 3} and len(row[row == 2]) == 1
    np.testing.assert_allclose(history(bk, D, "dem", 0, 2), 2.0 * first, rtol=1e-6, atol=1e-9)


def test_history_survives_grid_rebuild(bk):
    """tests/test_dem.py:358-384."""
    S = bk.scene(dem_nx=3)
    s, D, g = S.gen, S.D, S.grids
    bk.set(D, "pos", PAIR)
    bk.set(D, "vel", PAIR_VEL)
    rebuild_dem(bk, S)
    assert bk.run_force_dem(D, g[2], s.dem_radius, 0.0, s.dt) == 2
    first = history(bk, D, "dem", 0, 1).copy()
    bk.set(D, "pos", [[0.401, 0.16, 0.60], [0.450, 0.16, 0.60], [0.100, 0.16, 0.72]])
    rebuild_dem(bk, S)
    assert bk.run_force_dem(D, g[2], s.dem_radius, 0.0, s.dt) == 2
    np.testing.assert_allclose(history(bk, D, "dem", 0, 1), 2.0 * first, rtol=1e-6, atol=1e-9)


# ------------------------------------------------------------------ contact mechanics


def test_normal_contact_and_history_release(bk):
    """tests/test_dem.py:386-416."""
    S = bk.scene(dem_nx=2, dem_ny=1)
    s, D, g = S.gen, S.D, S.grids
    bk.set(D, "pos", [[0.2, 0.16, 0.6], [0.249, 0.16, 0.6]])
    bk.set(D, "vel", [[1, 0, 0], [-1, 0, 0]])
    rebuild_dem(bk, S)
    bk.run_force_dem(D, g[2], s.dem_radius, 0.0, s.dt)
    expected = s.dem_K * 0.001 + s.dem_eta * 2  # 1 mm overlap, 2 m/s approach
    f = bk.np(D.force)
    np.testing.assert_allclose(f, [[-expected, 0, 0], [expected, 0, 0]], rtol=2e-5, atol=0.01)
    np.testing.assert_allclose(f.sum(axis=0), 0, atol=1e-6)
    seeded = bk.np(D.tang_dem_old).copy()
    seeded[:] = [0.0, 0.1, 0.0]
    bk.set(D, "tang_dem_old", seeded)
    bk.set(D, "pos", [[0.2, 0.16, 0.6], [1.5, 0.16, 0.6]])
    rebuild_dem(bk, S)
    bk.run_force_dem(D, g[2], s.dem_radius, 0.0, s.dt)
    assert active(bk, D, "dem") == 0
    np.testing.assert_array_equal(bk.np(D.force), 0)
    bk.set(D, "pos", [[0.2, 0.16, 0.6], [0.249, 0.16, 0.6]])
    bk.set(D, "vel", [[0, 0, 0], [0, 0.01, 0]])
    rebuild_dem(bk, S)
    bk.run_force_dem(D, g[2], s.dem_radius, 0.0, s.dt)
    assert abs(float(history(bk, D, "dem", 0, 1)[1]) - 0.01 * s.dt) < 1e-8


def test_friction_history_torque_spin_and_coulomb_limit(bk):
    """tests/test_dem.py:418-444."""
    S = bk.scene(dem_nx=2, dem_ny=1)
    s, D, g = S.gen, S.D, S.grids
    bk.set(D, "pos", [[0.2, 0.16, 0.6], [0.249, 0.16, 0.6]])
    bk.set(D, "vel", [[0, 0, 0], [0, 0.01, 0]])
    rebuild_dem(bk, S)
    for count in (1, 2):
        bk.run_force_dem(D, g[2], s.dem_radius, 0.0, s.dt)
        tangential = s.dem_K * 0.01 * s.dt * count + s.dem_eta * 0.01
        assert abs(float(bk.np(D.force)[0, 1]) - tangential) < 1e-5
        assert abs(float(history(bk, D, "dem", 0, 1)[1]) - count * 0.01 * s.dt) < 1e-8
    torque = bk.np(D.torque)
    np.testing.assert_allclose(torque[:, 2], 0.0245 * tangential, rtol=1e-5)
    bk.K.Kernel_step_dem(D, s.dt)
    np.testing.assert_allclose(bk.np(D.omega)[:, 2], torque[:, 2] / s.dem_inertia * s.dt, rtol=1e-5)
    bk.set(D, "vel", [[0, 0, 0], [0, 10, 0]])
    bk.set(D, "omega", [0.0, 0.0, 0.0])
    bk.set(D, "pos", [[0.2, 0.16, 0.6], [0.249, 0.16, 0.6]])
    rebuild_dem(bk, S)
    bk.run_force_dem(D, g[2], s.dem_radius, 0.0, s.dt)
    assert abs(float(bk.np(D.force)[0, 1]) - s.dem_mu * s.dem_K * 0.001) < 0.01


def test_oblique_contact_and_three_axis_translation(bk):
    """tests/test_3d.py:104-121."""
    S = bk.scene(dem_nx=2)
    s, D, g = S.gen, S.D, S.grids
    normal = np.array([1.0, 2.0, 3.0]) / np.sqrt(14.0)
    pos = np.array([[0.2, 0.12, 0.6], [0.2, 0.12, 0.6] + 0.049 * normal], np.float32)
    bk.set(D, "pos", pos)
    bk.set(D, "vel", np.array([normal, -normal], np.float32))
    rebuild_dem(bk, S)
    bk.run_force_dem(D, g[2], s.dem_radius, 0.0, s.dt)
    f = bk.np(D.force)
    np.testing.assert_allclose(f[0], -(s.dem_K * 0.001 + 2 * s.dem_eta) * normal, rtol=5e-5, atol=0.002)
    np.testing.assert_allclose(f.sum(axis=0), 0.0, atol=1e-5)
    vel = bk.np(D.vel) + f / s.dem_mass * s.dt
    bk.K.Kernel_step_dem(D, s.dt)
    np.testing.assert_allclose(bk.np(D.pos), pos + vel * s.dt, atol=1e-7)
    assert np.all(bk.np(D.pos)[0] != pos[0])


def test_transverse_slip_gives_x_torque(bk):
    """tests/test_3d.py:123-138."""
    S = bk.scene(dem_nx=2, dem_omega_x=1.0, dem_omega_y=2.0, dem_omega_z=3.0)
    s, D, g = S.gen, S.D, S.grids
    np.testing.assert_allclose(bk.np(D.omega), [[1, 2, 3], [1, 2, 3]])
    bk.set(D, "omega", [0.0, 0.0, 0.0])
    bk.set(D, "pos", [[0.2, 0.10, 0.6], [0.2, 0.149, 0.6]])
    bk.set(D, "vel", [[0, 0, 0], [0, 0, 0.01]])
    rebuild_dem(bk, S)
    bk.run_force_dem(D, g[2], s.dem_radius, 0.0, s.dt)
    tangential = s.dem_K * 0.01 * s.dt + s.dem_eta * 0.01
    np.testing.assert_allclose(bk.np(D.torque)[:, 0], 0.0245 * tangential, rtol=1e-5)
    bk.K.Kernel_step_dem(D, s.dt)
    np.testing.assert_allclose(bk.np(D.omega)[:, 0], 0.0245 * tangential / s.dem_inertia * s.dt,
                               rtol=1e-5)


def test_fixed_wall_contact_history_and_release(bk):
    """tests/test_dem.py:446-492."""
    S = bk.scene(dem_nx=1, dem_ny=1)
    s, D, W, g = S.gen, S.D, S.W, S.grids
    cpos, cvel = [[0.4, 0.16, 0.020]], [0.1, 0.0, -0.2]
    bk.set(D, "pos", cpos)
    bk.set(D, "vel", cvel)
    rebuild_dem(bk, S)
    bk.run_force_dem(D, g[2], s.dem_radius, 0.0, s.dt)
    fixed = bk.np(W.pos).copy()
    n = bk.run_bc_dem(D, W, g[3], s.dem_bnd_radius, s.dt)
    ids = bk.np(D.contact_bnd_id_old)
    assert n > 0 and ids.shape == (n,) and int(bk.np(D.contact_bnd_offset_old)[-1]) == n
    first_nb = int(ids[0])
    first = history(bk, D, "bnd", 0, first_nb).copy()
    assert np.linalg.norm(first) > 0
    assert bk.np(D.force)[0, 2] > 0 and np.linalg.norm(bk.np(D.torque)[0]) > 0
    bk.K.Kernel_step_dem(D, s.dt)
    np.testing.assert_array_equal(bk.np(W.pos), fixed)
    np.testing.assert_array_equal(bk.np(W.vel), 0)
    np.testing.assert_array_equal(bk.np(W.omega), 0)
    for field in ("acc", "force", "torque"):
        assert not hasattr(W, field)
    bk.set(D, "pos", [0.4, 0.16, 0.8])
    rebuild_dem(bk, S)
    bk.run_bc_dem(D, W, g[3], s.dem_bnd_radius, s.dt)
    assert active(bk, D, "bnd") == 0
    bk.set(D, "pos", cpos)
    bk.set(D, "vel", cvel)
    bk.set(D, "omega", [0.0, 0.0, 0.0])
    rebuild_dem(bk, S)
    bk.run_bc_dem(D, W, g[3], s.dem_bnd_radius, s.dt)
    np.testing.assert_allclose(history(bk, D, "bnd", 0, first_nb), first, rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("back", [False, True])
def test_front_and_back_wall_repel(bk, back):
    """tests/test_3d.py:140-158."""
    S = bk.scene(dem_nx=1)
    s, D, W, g = S.gen, S.D, S.W, S.grids
    sign = -1.0 if back else 1.0
    y = s.tank_depth - 0.02 if back else 0.02
    bk.set(D, "pos", [0.4, y, 0.4])
    bk.set(D, "vel", [0.1, -sign * 0.2, 0.0])
    rebuild_dem(bk, S)
    wall = bk.np(W.pos).copy()
    bk.run_bc_dem(D, W, g[3], s.dem_bnd_radius, s.dt)
    assert sign * bk.np(D.force)[0, 1] > 0.0
    assert np.linalg.norm(bk.np(D.torque)[0]) > 0.0
    np.testing.assert_array_equal(bk.np(W.pos), wall)


# ------------------------------------------------------------------ coupling


def test_buoyancy_bounded_drag_reaction_balance_and_dry_reset(bk):
    """tests/test_dem.py:515-556."""
    S = bk.scene(dem_nx=1, dem_ny=1)
    s, P, D, g = S.gen, S.P, S.D, S.grids
    bk.set(D, "pos", [0.24, 0.24, 0.16])
    rebuild_dem(bk, S)
    bk.set(P, "vel", [0.2, 0, 0])
    bk.set(P, "pgf", [0, 0, s.rho0 * s.g])
    bk.set(P, "porosity", 0.7)
    bk.K.Kernel_interaction_dem(P, D, g[0], g[2], s.mu, s.dem_porosity_min, s.dem_porosity_max, s.dt)
    np.testing.assert_allclose(bk.np(D.pressure_force)[0], [0, 0, s.dem_volume * s.rho0 * s.g],
                               rtol=2e-6)
    drag = bk.np(D.drag)
    assert drag[0, 0] > 0 and drag[0, 0] / s.dem_mass * s.dt < 0.2
    bk.set(P, "acc", [1.0, 0.0, -s.g])
    acc0 = bk.np(P.acc).copy()
    bk.K.Kernel_interaction_sph(P, D, g[0], g[2], s.dem_support)
    m, eps, acc_dem = bk.np(P.m), bk.np(P.porosity), bk.np(P.acc_dem)
    reaction = np.sum(m[:, None] * eps[:, None] * acc_dem, axis=0)
    np.testing.assert_allclose(reaction, -drag.sum(axis=0), rtol=3e-6, atol=1e-5)
    np.testing.assert_allclose(bk.np(P.acc), acc0 + acc_dem, rtol=1e-6)
    bk.set(D, "pos", [0.24, 0.16, 2.0])
    rebuild_dem(bk, S)
    bk.K.Kernel_interaction_dem(P, D, g[0], g[2], s.mu, s.dem_porosity_min, s.dem_porosity_max, s.dt)
    for f in ("drag", "pressure_force", "flt_s"):
        np.testing.assert_array_equal(bk.np(getattr(D, f)), 0)
    prev = bk.np(P.acc).copy()
    bk.K.Kernel_interaction_sph(P, D, g[0], g[2], s.dem_support)
    np.testing.assert_array_equal(bk.np(P.acc_dem), 0.0)
    np.testing.assert_array_equal(bk.np(P.acc), prev)


def test_hydrostatic_pressure_gradient_points_up(bk):
    """tests/test_dem.py:558-572."""
    S = bk.scene()
    s, P, B, D, g = S.gen, S.P, S.B, S.D, S.grids
    for part in (P, B):
        bk.set(part, "pres", (10000 - s.rho0 * s.g * bk.np(part.pos)[:, 2]).astype(np.float32))
    bk.K.Kernel_prep_sphdem(P, B, D, *g[:3], s.dem_support, s.dem_h, s.support, s.h,
                            s.dem_porosity_min, s.dem_porosity_max)
    pos = bk.np(P.pos)
    c = np.argmin(np.linalg.norm(pos - [0.24, 0.16, 0.24], axis=1))
    np.testing.assert_allclose(bk.np(P.pgf)[c], [0, 0, s.rho0 * s.g], rtol=0.06, atol=0.02)


def test_linear_pressure_gradient_and_drag_in_xyz(bk):
    """tests/test_3d.py:160-188."""
    S = bk.scene(dem_nx=1)
    s, P, B, D, g = S.gen, S.P, S.B, S.D, S.grids
    grad = np.array([400.0, -900.0, 700.0])
    for part in (P, B):
        bk.set(part, "pres", (10000.0 + bk.np(part.pos) @ grad).astype(np.float32))
    bk.K.Kernel_prep_sphdem(P, B, D, *g[:3], s.dem_support, s.dem_h, s.support, s.h,
                            s.dem_porosity_min, s.dem_porosity_max)
    c = np.argmin(np.linalg.norm(bk.np(P.pos) - [0.2, 0.2, 0.16], axis=1))
    np.testing.assert_allclose(bk.np(P.pgf)[c], -grad, rtol=0.06)
    bk.set(D, "pos", [0.2, 0.2, 0.16])
    rebuild_dem(bk, S)
    bk.set(P, "pgf", -grad)
    fv = np.array([0.2, -0.1, 0.3])
    bk.set(P, "vel", fv)
    bk.set(P, "porosity", 0.8)
    bk.K.Kernel_interaction_dem(P, D, g[0], g[2], s.mu, s.dem_porosity_min, s.dem_porosity_max, s.dt)
    np.testing.assert_allclose(bk.np(D.pressure_force)[0], -s.dem_volume * grad, rtol=3e-6)
    drag = bk.np(D.drag)[0]
    np.testing.assert_allclose(drag / np.linalg.norm(drag), fv / np.linalg.norm(fv), rtol=3e-6)
    bk.K.Kernel_interaction_sph(P, D, g[0], g[2], s.dem_support)
    reaction = np.sum(bk.np(P.m)[:, None] * bk.np(P.porosity)[:, None] * bk.np(P.acc_dem), axis=0)
    np.testing.assert_allclose(reaction, -drag, rtol=3e-6, atol=1e-6)


def test_dry_free_fall_leaves_sph_trajectory_bit_identical(bk):
    """tests/test_dem.py:574-597."""
    S = bk.scene(dem_origin_z=1.5)
    s, solv, g = S.gen, S.solv, S.grids
    from helpers import particles
    from warp_sph_gpt_amd.input_gen import DamPtlGeneration
    P0, B0 = bk.from_structs(s, DamPtlGeneration(s).build())
    gs, gb = bk.grid(32), bk.grid(32)
    bk.build(gb, B0.pos, s.support)
    z0 = bk.np(S.D.pos).copy()
    for step in range(5):
        bk.build(g[0], S.P.pos, s.support)
        bk.build(g[2], S.D.pos, s.dem_support)
        bk.build(gs, P0.pos, s.support)
        bk.SPH_OneStep(solv, P0, B0, gs, gb, step)
        bk.SPHDEM_OneStep(solv, S.P, S.B, S.D, S.W, *g, step)
    np.testing.assert_array_equal(bk.np(S.P.pos), bk.np(P0.pos))
    np.testing.assert_array_equal(bk.np(S.P.vel), bk.np(P0.vel))
    v = bk.np(S.D.vel)
    np.testing.assert_array_equal(v[:, :2], 0.0)
    np.testing.assert_allclose(v[:, 2], -5 * s.g * s.dt, rtol=1e-6)
    np.testing.assert_allclose(bk.np(S.D.pos)[:, 2], z0[:, 2] - s.g * s.dt ** 2 * 15, atol=5e-7)


def test_variable_dem_h_porosity_brute_force_and_reaction(bk):
    """tests/test_input_reader.py:205-253 (DEM inside the fluid, h = 0.04 and 0.10)."""
    S = bk.scene(dem_origin_z=0.2)
    s, P, B, D = S.gen, S.P, S.B, S.D
    bk.set(D, "h", [0.04, 0.10])
    bk.set(P, "vel", [0.5, -0.2, 0.3])
    grids = [bk.grid(32) for _ in range(3)]
    for grid, part in zip(grids, (P, B, D)):
        bk.build(grid, part.pos, 0.2)
    bk.K.Kernel_prep_sphdem(P, B, D, *grids, 0.2, 0.0, s.support, s.h, s.dem_porosity_min,
                            s.dem_porosity_max)
    pos, centers = bk.np(P.pos).astype(np.float64), bk.np(D.pos).astype(np.float64)
    vol = (bk.np(P.m) / bk.np(P.rho)).astype(np.float64)
    sample = np.arange(0, len(pos), 17)
    solid = np.zeros(len(sample))

    def weight(r, h):
        q = r / h
        return np.where(q < 2, 21 / (16 * np.pi * h ** 3) * (1 - q / 2) ** 4 * (1 + 2 * q), 0)

    dvol = bk.np(D.volume).astype(np.float64)
    for b, h in enumerate(bk.np(D.h).astype(np.float64)):
        norm = (weight(np.linalg.norm(pos[sample, None] - pos[None], axis=2), h) * vol[None]).sum(1)
        solid += dvol[b] * weight(np.linalg.norm(pos[sample] - centers[b], axis=1), h) / (norm + 1e-20)
    np.testing.assert_allclose(bk.np(P.porosity)[sample],
                               np.clip(1 - solid, s.dem_porosity_min, s.dem_porosity_max),
                               rtol=2e-6, atol=2e-6)
    bk.K.Kernel_interaction_dem(P, D, grids[0], grids[2], s.mu, s.dem_porosity_min,
                                s.dem_porosity_max, s.dt)
    bk.K.Kernel_interaction_sph(P, D, grids[0], grids[2], 0.2)
    assert np.all(bk.np(D.flt_s) > 0)
    reaction = np.sum((bk.np(P.m) * bk.np(P.porosity))[:, None] * bk.np(P.acc_dem), axis=0)
    np.testing.assert_allclose(reaction, -bk.np(D.drag).sum(axis=0), rtol=2e-5, atol=1e-8)
    prev = bk.np(D.flt_s).copy()
    bk.set(D, "h", 0.07)
    bk.K.Kernel_interaction_dem(P, D, grids[0], grids[2], s.mu, s.dem_porosity_min,
                                s.dem_porosity_max, s.dt)
    assert float(np.max(np.abs(prev - bk.np(D.flt_s)))) > 1e-4
