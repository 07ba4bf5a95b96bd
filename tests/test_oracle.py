"""Pins of the CPU oracle that do not depend on its own formulas (CPU only).

* hash grid: stable permutation and bucket partition; neighbour sets of the
  Warp-order query against an O(N^2) search on folded grids with negative
  coordinates;
* neighbour sums (Shepard, fluid/wall density, force, -grad p, porosity,
  DEM interpolation) against O(N^2) FP64 sums written from the textbook
  forms (factored Wendland C2 kernel, pairwise pressure and Morris terms),
  for the FP32 oracle (1e-5) and the FP64 oracle (1e-11);
* invariants: Shepard-corrected density of a uniform lattice equals rho0;
  fluid-only pressure and viscous forces conserve momentum.
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import COMPACT, generation, particles, rel_err, solv_for
from oracle import oracle as O


def W64(r, h):
    q = r / h
    return np.where(q < 2.0, 21.0 / (16.0 * np.pi * h ** 3) * (1 - q / 2) ** 4 * (1 + 2 * q), 0.0)


def dW64(r, h):
    q = r / h
    return np.where(q < 2.0, 21.0 / (16.0 * np.pi * h ** 4) * (-5.0 * q) * (1 - q / 2) ** 3, 0.0)


def pairs(xa, xb, chunk=512):
    for s in range(0, len(xa), chunk):
        d = xa[s:s + chunk, None, :] - xb[None, :, :]
        yield s, d, np.sqrt((d ** 2).sum(-1))


@pytest.fixture(scope="module", params=[np.float32, np.float64], ids=["f32", "f64"])
def scene(request):
    dt = request.param
    gen = generation(**dict(COMPACT, dem_origin_z=0.2))
    structs = particles(gen)
    P, B, D, W = (O.OStruct.from_struct(s, dt) for s in structs)
    rng = np.random.default_rng(5)
    P.pos = (P.pos + rng.uniform(-0.1, 0.1, P.pos.shape) * gen.dx).astype(dt)
    P.vel = rng.normal(0, 0.1, P.pos.shape).astype(dt)
    P.rho = (1000.0 * (1 + rng.normal(0, 0.003, P.n))).astype(dt)
    B.rho = (1000.0 * (1 + rng.normal(0, 0.003, B.n))).astype(dt)
    P.pres = (3000.0 * (0.45 - P.pos[:, 2])).astype(dt)
    B.pres = (3000.0 * (0.45 - B.pos[:, 2])).astype(dt)
    grids = [O.OracleGrid(32, 32, 32, dt).build(s.pos, r)
             for s, r in zip((P, B, D), (gen.support, gen.support, 0.16))]
    tol = 1e-5 if dt == np.float32 else 1e-11
    return gen, solv_for(gen, "cpu"), P, B, D, W, grids, tol


def test_grid_is_a_stable_partition_and_queries_are_complete():
    rng = np.random.default_rng(3)
    pts = rng.uniform(-0.7, 0.9, (3000, 3)).astype(np.float32)
    g = O.OracleGrid(8, 8, 8).build(pts, 0.07)
    ids, cells = g.point_ids, g.point_cells
    assert sorted(ids) == list(range(len(pts)))
    assert np.all(np.diff(cells) >= 0)
    same = np.diff(cells) == 0
    assert np.all(np.diff(ids)[same] > 0)  # stable: ascending index inside a bucket
    for c in np.unique(cells):
        sl = np.nonzero(cells == c)[0]
        assert g.cell_starts[c] == sl[0] and g.cell_ends[c] == sl[-1] + 1
    empty = np.setdiff1d(np.arange(8 ** 3), cells)
    assert np.all(g.cell_starts[empty] == 0) and np.all(g.cell_ends[empty] == 0)
    for r in (0.07, 0.15):
        got = O.neighbour_counts(g, pts, g, pts, r)
        d = pts[:, None, :] - pts[None, :, :]
        r2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
        np.testing.assert_array_equal(got, (r2 < np.float32(r) * np.float32(r)).sum(1))


def test_shepard_and_densities_match_fp64_sums(scene):
    gen, solv, P, B, D, W, g, tol = scene
    P2, B2 = P.copy(), B.copy()
    O.Kernel_shepard_sph(P2, B2, g[0], g[1], gen.support, gen.h)
    O.Kernel_density_sph(P2, B2, g[0], g[1], gen.support, gen.h)
    O.Kernel_density_bnd(P2, B2, g[0], g[1], gen.support, gen.h)
    xp, xb = P.pos.astype(np.float64), B.pos.astype(np.float64)
    h = float(np.asarray(gen.h, P.pos.dtype))
    flt, rho = np.zeros(P.n), np.zeros(P.n)
    for s, _, r in pairs(xp, xp):
        w = W64(r, h)
        flt[s:s + 512] += (w * (P.m / P.rho)).sum(1)
        rho[s:s + 512] += (w * P.m).sum(1)
    for s, _, r in pairs(xp, xb):
        w = W64(r, h)
        flt[s:s + 512] += (w * (B.m / B.rho)).sum(1)
        rho[s:s + 512] += (w * B.m).sum(1)
    assert rel_err(P2.flt, flt) < tol
    assert rel_err(P2.rho, rho / P2.flt.astype(np.float64)) < tol
    rb = np.zeros(B.n)
    for s, _, r in pairs(xb, xp):
        rb[s:s + 512] += (W64(r, h) * P.m).sum(1)
    for s, _, r in pairs(xb, xb):
        rb[s:s + 512] += (W64(r, h) * B.m).sum(1)
    assert rel_err(B2.rho, rb) < tol


def test_force_and_pressure_gradient_match_fp64_sums(scene):
    gen, solv, P, B, D, W, g, tol = scene
    P2 = P.copy()
    O.Kernel_force_sph(P2, B, g[0], g[1], gen.support, gen.h, gen.mu, gen.g)
    P3 = P.copy()
    O.Kernel_prep_sphdem(P3, B, D, *g, 0.16, 0.08, gen.support, gen.h, 0.05, 1.0)
    h = float(np.asarray(gen.h, P.pos.dtype))
    mu = float(np.asarray(gen.mu, P.pos.dtype))
    acc = np.zeros((P.n, 3))
    pgf = np.zeros((P.n, 3))
    xp = P.pos.astype(np.float64)
    for other in (P, B):
        xo = other.pos.astype(np.float64)
        for s, d, r in pairs(xp, xo):
            e = slice(s, s + 512)
            ok = (r ** 2 > 1e-12) & (r < 2 * h)
            rs = np.where(ok, r, 1.0)
            gw = np.where(ok, dW64(rs, h), 0.0)[..., None] * d / rs[..., None]
            pi, ri = P.pres[e][:, None], P.rho[e][:, None]
            pj, rj, mj = other.pres[None, :], other.rho[None, :], other.m[None, :]
            acc[e] -= ((mj * (pi / ri ** 2 + pj / rj ** 2))[..., None] * gw).sum(1)
            visc = 2 * mu * (d * gw).sum(-1) / (ri * rj * (r ** 2 + 0.01 * h * h))
            dv = P.vel[e][:, None, :] - other.vel[None, :, :]
            acc[e] += ((mj * visc)[..., None] * dv).sum(1)
            pgf[e] -= ((mj / rj * (pj - pi))[..., None] * gw).sum(1)
    acc[:, 2] -= gen.g
    assert rel_err(P2.acc, acc) < 3 * tol
    assert rel_err(P3.pgf, pgf) < 3 * tol


def test_porosity_and_dem_interpolation_match_fp64_sums(scene):
    gen, solv, P, B, D, W, g, tol = scene
    P3, D3 = P.copy(), D.copy()
    O.Kernel_prep_sphdem(P3, B, D3, *g, 0.16, 0.08, gen.support, gen.h, 0.05, 1.0)
    O.Kernel_interaction_dem(P3, D3, g[0], g[2], gen.mu, 0.05, 1.0, gen.dt)
    xp, xd = P.pos.astype(np.float64), D.pos.astype(np.float64)
    V = P.m.astype(np.float64) / P.rho
    hu = float(np.asarray(0.08, P.pos.dtype))       # the uniform h passed to prep
    hb = float(D.h[0])                               # the DEM's stored h (all equal)
    flt = np.zeros(P.n)
    for s, _, r in pairs(xp, xp):
        flt[s:s + 512] += (W64(r, hu) * V).sum(1)
    rd = np.sqrt(((xp[:, None] - xd[None]) ** 2).sum(-1))
    solid = (D.volume[None] * W64(rd, hb)).sum(1) / (flt + 1e-20)
    assert rel_err(P3.porosity, np.clip(1 - solid, 0.05, 1.0)) < tol
    assert P3.porosity.min() < 0.99
    w = W64(rd, hb) * V[:, None]          # (fluid, dem)
    fs = w.sum(0)
    vel_f = (w[..., None] * P.vel[:, None, :]).sum(0)
    assert rel_err(D3.flt_s, fs) < tol
    urel = vel_f / fs[:, None] - D.vel
    assert np.all(np.linalg.norm(urel, axis=1) > 0)
    dd = np.sqrt(((xd[:, None] - xd[None]) ** 2).sum(-1))
    sol = (D.volume[None] * W64(dd, hb)).sum(1)
    eps = np.clip(1 - sol / fs, 0.05, 1.0)
    assert rel_err(D3.porosity, eps) < 10 * tol


def test_uniform_lattice_density_after_shepard_is_rho0():
    gen = generation(**COMPACT)
    P, B, D, W = (O.OStruct.from_struct(s) for s in particles(gen))
    g = [O.OracleGrid(32, 32, 32).build(s.pos, gen.support) for s in (P, B)]
    O.Kernel_shepard_sph(P, B, *g, gen.support, gen.h)
    O.Kernel_density_sph(P, B, *g, gen.support, gen.h)
    np.testing.assert_allclose(P.rho, gen.rho0, rtol=2e-6)


def test_fluid_only_forces_conserve_momentum():
    gen = generation(**COMPACT)
    P, B, D, W = (O.OStruct.from_struct(s, np.float64) for s in particles(gen))
    rng = np.random.default_rng(9)
    P.vel = rng.normal(0, 0.2, P.pos.shape)
    P.pres = rng.uniform(0, 2000, P.n)
    P.rho = 1000 * (1 + rng.normal(0, 0.01, P.n))
    E = O.OStruct(**{k: (v[:0] if isinstance(v, np.ndarray) else v) for k, v in B.__dict__.items()})
    g = [O.OracleGrid(32, 32, 32, np.float64).build(s.pos, gen.support) for s in (P, E)]
    O.Kernel_force_sph(P, E, *g, gen.support, gen.h, gen.mu, 0.0)
    f = P.m[:, None] * P.acc
    assert np.abs(f.sum(0)).max() < 1e-12 * np.abs(f).sum(0).max()
