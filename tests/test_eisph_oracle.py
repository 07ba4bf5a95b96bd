"""The CPU oracle of the Eulerian ISPH cavity, pinned (CPU only).

Mirrors the checks of the reference's tests/test_eisph.py (kernel
normalisation, PDF signs, -4 nu diffusion, Dirichlet ghosts, the one-pass
pressure update, the cavity run, Ghia et al. Re=100) and adds the
first-order consistency of the kernel-gradient correction.  The Re=100 run
is also held to the figures the reference publishes for that run
(README.md:79-83): centre-line RMSE 0.00840 and maximum 0.01550 against Ghia,
centre velocity (-0.1906, 0.0563).
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import eisph as E
from warp_sph_gpt_amd.config import EISPHSolv
from warp_sph_gpt_amd.input_gen import CavityPtlGeneration, GenerationConfig

GHIA_Y = np.array([0.9766, 0.9688, 0.9609, 0.9531, 0.8516, 0.7344, 0.6172, 0.5000, 0.4531,
                   0.2813, 0.1719, 0.1016, 0.0703, 0.0625, 0.0547])
GHIA_U = np.array([0.84123, 0.78871, 0.73722, 0.68717, 0.23151, 0.00332, -0.13641, -0.20581,
                   -0.21090, -0.15662, -0.10150, -0.06434, -0.04775, -0.04192, -0.03717])
GHIA_X = np.array([0.9688, 0.9609, 0.9531, 0.9453, 0.9063, 0.8594, 0.8047, 0.5000, 0.2344,
                   0.2266, 0.1563, 0.0938, 0.0781, 0.0703, 0.0625])
GHIA_W = np.array([-0.05906, -0.07391, -0.08864, -0.10313, -0.16914, -0.22445, -0.24533,
                   0.05454, 0.17527, 0.17507, 0.16077, 0.12317, 0.10890, 0.10091, 0.09233])


def ghia_errors(vel, gen):
    """Centre-line velocity minus Ghia et al. (1982) Tables I-II, Re=100."""
    n = gen.cells
    v = np.asarray(vel, np.float64).reshape(n, n, 3)
    c = n // 2
    axis = (np.arange(n) + 0.5) * gen.dx
    return np.concatenate((np.interp(GHIA_Y, axis, v[:, c, 0]) - GHIA_U,
                           np.interp(GHIA_X, axis, v[c, :, 2]) - GHIA_W))


def cavity(dtype=np.float32, **kw):
    gen = GenerationConfig(device="cpu", **kw)
    P, B = CavityPtlGeneration(gen).build()
    solv = EISPHSolv(device="cpu", h=gen.h, rho0=gen.rho0, nu=gen.nu, dt=gen.dt,
                     n_steps=gen.n_steps, output_step=gen.output_step,
                     grid_slice=gen.grid_slice)
    return gen, solv, E.from_struct(P, dtype), E.from_struct(B, dtype)


def test_wendland_2d_is_normalised_with_matching_slope():
    h = 0.13
    r = np.linspace(0.0, 2.0 * h, 4001)
    w, dw = E.wendland2(r, h, np.float64)
    integrand = 2.0 * np.pi * r * w
    integral = np.sum(0.5 * (integrand[:-1] + integrand[1:]) * np.diff(r))
    assert integral == pytest.approx(1.0, abs=1e-6)
    np.testing.assert_allclose(dw[50:-50:100], np.gradient(w, r)[50:-50:100], rtol=3e-3)
    # closed form at q = 0 and q = 1, zero at and beyond 2h
    assert w[0] == pytest.approx(7.0 / (4.0 * np.pi * h * h), rel=1e-12)
    w1, dw1 = E.wendland2([h, 2.0 * h, 3.0 * h], h, np.float64)
    assert w1[0] == pytest.approx(7.0 * 0.5 ** 4 * 3.0 / (4.0 * np.pi * h * h), rel=1e-12)
    assert dw1[0] == pytest.approx(-35.0 * 0.5 ** 3 / (4.0 * np.pi * h ** 3), rel=1e-12)
    assert w1[1] == w1[2] == dw1[1] == dw1[2] == 0.0


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_corrected_gradient_is_exact_for_linear_pressure(dtype):
    """First-order consistency of the KGC: with ghosts completing the support,
    the projection of a linear pressure field returns its exact gradient at
    every fluid point, walls and corners included."""
    gen, solv, P, B = cavity(dtype, dx=0.1, grid_slice=16)
    gs, gb = E.build_grids(solv, P, B)
    a, b, c = 3.0, -2.0, 0.5
    P.pres[:] = a * P.pos[:, 0] + b * P.pos[:, 2] + c
    B.pres[:] = a * B.pos[:, 0] + b * B.pos[:, 2] + c
    P.vel_star[:] = 0.0
    E.Kernel_correct_eisph(P, B, gs, gb, solv.support, solv.h, 1.0)
    tol = 2e-4 if dtype == np.float32 else 1e-9
    np.testing.assert_allclose(P.vel[:, 0], -a, rtol=tol)
    np.testing.assert_allclose(P.vel[:, 2], -b, rtol=tol)
    assert np.all(P.vel[:, 1] == 0.0)
    # without the correction the wall rows miss part of the gradient
    P2 = P.copy()
    P2.kgc[:] = np.array([1.0, 0.0, 0.0, 1.0], dtype)
    E.Kernel_correct_eisph(P2, B, gs, gb, solv.support, solv.h, 1.0)
    assert np.abs(P2.vel[:, 0] + a).max() > 1e-2


def test_ppe_assembly_uses_pdf_rhs_sign():
    """test_eisph.py:112-142: b = rho0 div(u*)/dt, A_ij sums negative, and a
    constant pressure gives Aijpj = p Aij."""
    gen, solv, P, B = cavity(dx=0.2, n_steps=1, output_step=1, grid_slice=16)
    P.pres[:] = 3.25
    gs, gb = E.build_grids(solv, P, B)
    E.Kernel_Dirichlet_BC(P, B, 0)
    E.Kernel_advec_vis_eisph(P, B, gs, gb, solv.support, solv.h, solv.nu, solv.dt)
    E.Kernel_Dirichlet_BC(P, B, 1)
    E.Kernel_build_ppe_eisph(P, B, gs, gb, solv.support, solv.h, solv.rho0, solv.dt)
    np.testing.assert_allclose(P.bi, solv.rho0 * P.divergence / solv.dt, rtol=2e-6, atol=2e-5)
    assert np.all(P.Aij < 0.0)
    np.testing.assert_allclose(P.Aijpj, 3.25 * P.Aij, rtol=2e-6, atol=2e-5)


def test_laplacian_diffuses_a_centre_maximum():
    """test_eisph.py:144-181: u = -r^2 has Laplacian -4 and no convection at
    the centre, so acc = -4 nu and vel_star = -4 nu dt there."""
    gen, solv, P, B = cavity(dx=0.2, n_steps=1, output_step=1, grid_slice=16)
    r2 = (P.pos[:, 0] - 0.5) ** 2 + (P.pos[:, 2] - 0.5) ** 2
    P.vel[:, 0] = -r2
    gs, gb = E.build_grids(solv, P, B)
    E.Kernel_Dirichlet_BC(P, B, 0)
    E.Kernel_advec_vis_eisph(P, B, gs, gb, solv.support, solv.h, solv.nu, solv.dt)
    c = int(np.argmin(r2))
    assert P.vel[c, 0] == pytest.approx(0.0, abs=1e-12)
    np.testing.assert_allclose(P.acc[c, 0], -4.0 * solv.nu, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(P.vel_star[c, 0], -4.0 * solv.nu * solv.dt, rtol=1e-5, atol=1e-7)


def test_general_velocity_dirichlet_ghosts():
    """test_eisph.py:183-212: u_b = 2 u_D - u_mirror for vel and vel_star."""
    gen, solv, P, B = cavity(dx=0.2, n_steps=1, output_step=1)
    fid = np.arange(P.n, dtype=np.float32)
    bid = np.arange(B.n, dtype=np.float32)
    P.vel[:] = np.stack((0.01 * fid, -0.02 * fid, 0.03 * fid), 1)
    P.vel_star[:] = -0.5 * P.vel
    B.vel_bc[:] = np.stack((0.2 + 0.001 * bid, -0.1 + 0.002 * bid, 0.05 - 0.001 * bid), 1)
    E.Kernel_Dirichlet_BC(P, B, 0)
    E.Kernel_Dirichlet_BC(P, B, 1)
    np.testing.assert_array_equal(B.vel, 2.0 * B.vel_bc - P.vel[B.mirror])
    np.testing.assert_array_equal(B.vel_star, 2.0 * B.vel_bc - P.vel_star[B.mirror])


def test_cavity_step_is_fixed_finite_and_mirrored():
    """test_eisph.py:214-249 on the oracle."""
    gen, solv, P, B = cavity(dx=0.1, n_steps=20, output_step=20, grid_slice=16)
    pos0 = P.pos.copy()
    positions, frames, times = E.run_eisph(solv, P, B)
    np.testing.assert_array_equal(P.pos, pos0)
    np.testing.assert_array_equal(positions, pos0)
    assert times == [0.0, solv.n_steps * solv.dt] and len(frames) == 2
    for name in ("vel", "vel_star", "acc", "pres", "Aij", "bi", "Aijpj", "divergence",
                 "residual", "kgc"):
        assert np.isfinite(getattr(P, name)).all(), name
    assert np.all(P.vel[:, 1] == 0.0)
    np.testing.assert_allclose(B.vel, 2.0 * B.vel_bc - P.vel[B.mirror], rtol=1e-6, atol=1e-7)
    assert float(P.bi.mean()) == pytest.approx(0.0, abs=1e-4)
    assert float(P.pres[0]) == 0.0
    assert np.linalg.norm(P.divergence) < np.linalg.norm(P.bi * solv.dt / solv.rho0)
    u = P.vel[:, 0].reshape(gen.cells, gen.cells)
    assert u[-1, gen.cells // 2] > 0.0
    assert u[gen.cells // 2, gen.cells // 2] < 0.0


def test_pressure_update_is_one_implicit_diagonal_pass():
    """test_eisph.py:266-320: p = (b + Aijpj)/Aij on active rows, shifted by
    row 0's value; a constant shift of p^t changes neither p nor u."""
    gen, solv, _, _ = cavity(dx=0.2, n_steps=1, output_step=1, grid_slice=16)
    old = np.linspace(-0.25, 0.25, gen.cells ** 2, dtype=np.float32)

    def one_step(p0):
        _, s, P, B = cavity(dx=0.2, n_steps=1, output_step=1, grid_slice=16)
        P.pres[:] = p0
        E.run_eisph(s, P, B)
        return P

    P = one_step(old)
    active = np.abs(P.Aij) > 1e-12
    assert np.all(P.Aij[active] < 0.0)
    expected = old.copy()
    expected[active] = (P.bi[active] + P.Aijpj[active]) / P.Aij[active]
    expected -= expected[0] if active[0] else 0.0
    np.testing.assert_allclose(P.pres, expected, rtol=2e-6, atol=2e-6)
    shifted = one_step(old + 7.5)
    np.testing.assert_allclose(shifted.pres, P.pres, rtol=2e-5, atol=2e-5)
    np.testing.assert_allclose(shifted.vel, P.vel, rtol=2e-5, atol=2e-6)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_re100_cavity_matches_ghia_and_the_published_run(dtype):
    """test_eisph.py:322-359 (RMSE < 0.015, max < 0.030 against Ghia) and the
    reference's own figures for this run, README.md:79-83."""
    gen, solv, P, B = cavity(dtype)
    E.run_eisph(solv, P, B)
    err = ghia_errors(P.vel, gen)
    rmse, emax = float(np.sqrt(np.mean(err ** 2))), float(np.abs(err).max())
    assert rmse < 0.015 and emax < 0.030
    assert round(rmse, 5) == 0.00840
    assert round(emax, 5) == 0.01550
    c = gen.cells // 2
    v = P.vel.reshape(gen.cells, gen.cells, 3)
    assert round(float(v[c, c, 0]), 4) == -0.1906
    assert round(float(v[c, c, 2]), 4) == 0.0563
    assert np.linalg.norm(P.residual) / np.linalg.norm(P.bi) < 1e-5
