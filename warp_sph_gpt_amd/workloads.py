"""The dam-break workloads used for parity runs and measurement.

  c0  the reference's shipped case: dx=0.02, 12,500 fluid, 48,336 wall,
      20 DEM on a 5x2x2 lattice, 14,688 DEM-wall spheres, dt=1e-4, 100 steps
  c1  dx=0.01: 100,000 fluid, 180,456 wall, 200 DEM by seeded random
      sequential addition in [0.05,0.45]x[0.05,0.35]x[0.55,0.9], DEM-wall
      spacing 0.01 (58,017), dt=5e-5, 1,000 steps
  c2  dx=0.004: 1,562,500 fluid, 1,080,816 wall, 2,000 DEM (R=0.025) by
      seeded random sequential addition starting in
      [0.05,0.6]x[0.03,0.37]x[0.55,0.95]; that box cannot hold 2,000
      non-overlapping spheres (124 % packing), so whenever 20,000 draws do
      not place them all the box top is raised by 0.05 m and drawing
      continues (the top ends at z = 2.35 m for the default seed).
      DEM-wall spacing 0.02 (14,688), dt=2.5e-5, 500 steps.
  c3  ensemble of 64 independent dam breaks at dx=0.01, dt=5e-5, 500 steps:
      instance k draws (numpy default_rng(seed + k)) a fluid block length
      Lx ~ U[0.3, 0.7] and height Lz ~ U[0.3, 0.6] (rounded to dx; full
      0.4 m depth), N_dem ~ U{0..100} spheres with R ~ U[0.015, 0.03] placed
      by random sequential addition above its free surface in
      [0.04, Lx-0.04] x [0.04, 0.36] x [Lz+0.06, Lz+0.36], its own 3-layer
      tank wall (180,456) and DEM-wall spheres (spacing 0.02, 14,688).
All: tank 2 x 0.4 x 1 m, fluid block 0.5 x 0.4 x 0.5 m at the origin (c0-c2),
3-layer dummy wall, h = 1.3 dx, rho0 = 1000, DEM h = 0.08, K=2e4, eta=4,
mu=0.3, rho_s=2500, DEM-wall radius 0.01.

The Eulerian lid-driven cavity (Re = 100, 1 x 1 m, lid speed 1 m/s,
h = 1.3 dx, 3 ghost layers):
  e0  the reference's shipped cavity: 25 x 25 fluid points, 336 ghosts,
      dt = 2e-3, 5,000 steps
  e1  1024 x 1024 fluid points (1,048,576), 12,324 ghosts, dt = 1e-5
      (the predictor's viscosity limit 0.125 dx^2/nu is 1.19e-5), 500 steps
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .config import Solv
from .input_gen import (CoupledGenerationConfig, DamPtlGeneration, DEMPtlGeneration,
                        check_dem_inside, rsa_spheres)


@dataclass
class Workload:
    name: str
    gen: CoupledGenerationConfig
    solv: Solv
    structs: tuple  # (SPHptl, BNDptl, DEMptl, DEMBNDptl) on CPU
    steps: int
    note: str = ""

    @property
    def n_fluid(self) -> int:
        return self.structs[0].n


def runtime_solv(gen: CoupledGenerationConfig, device="cuda:0") -> Solv:
    s = Solv(device=device)
    for name in ("h", "rho0", "c0", "gamma", "mu", "g", "shepard_step", "grid_slice", "dt",
                 "n_steps", "dem_porosity_min", "dem_porosity_max", "dem_dt_safety"):
        setattr(s, name, getattr(gen, name))
    return s


def _build(name, gen, steps, dem_pos=None, note="", device="cuda:0"):
    P, B = DamPtlGeneration(gen).build()
    D, W = DEMPtlGeneration(gen).build(dem_pos=dem_pos)
    solv = runtime_solv(gen, device)
    solv.n_steps = steps
    return Workload(name, gen, solv, (P, B, D, W), steps, note)


def c0(device="cuda:0") -> Workload:
    gen = CoupledGenerationConfig(device="cpu")
    return _build("c0", gen, 100, note="reference input files (dx=0.02)", device=device)


def c1(device="cuda:0", seed=20261) -> Workload:
    gen = CoupledGenerationConfig(device="cpu", dx=0.01, dt=5.0e-5, n_steps=1000,
                                  dem_bnd_spacing=0.01)
    pos, _, _ = rsa_spheres(200, gen.dem_radius, (0.05, 0.05, 0.55), (0.45, 0.35, 0.9), seed)
    check_dem_inside(pos, gen.dem_radius, gen)
    return _build("c1", gen, 1000, dem_pos=pos, note=f"200 DEM, RSA seed {seed}", device=device)


def c2(device="cuda:0", seed=20262, n_dem=2000) -> Workload:
    gen = CoupledGenerationConfig(device="cpu", dx=0.004, dt=2.5e-5, n_steps=500)
    pos, _, hi = rsa_spheres(n_dem, gen.dem_radius, (0.05, 0.03, 0.55), (0.6, 0.37, 0.95), seed,
                             attempts=20_000, grow_z=0.05)
    check_dem_inside(pos, gen.dem_radius, gen)
    note = (f"{len(pos)} DEM, RSA seed {seed}, centre box z in [0.55, {hi[2]:.2f}] "
            "(raised from 0.95 to fit 2,000 spheres)")
    return _build("c2", gen, 500, dem_pos=pos, note=note, device=device)


@dataclass
class Ensemble:
    name: str
    solv: Solv
    scenes: list     # [(SPHptl, BNDptl, DEMptl | None, DEMBNDptl | None)] on CPU
    steps: int
    params: list     # per instance: dict(Lx, Lz, n_dem, seed)
    note: str = ""

    @property
    def n_fluid(self) -> int:
        return sum(sc[0].n for sc in self.scenes)


def ensemble_scene(seed: int, dx: float = 0.01, dt: float = 5.0e-5, max_dem: int = 100,
                   n_dem: int | None = None):
    """One instance of the ensemble: (gen, (P, B, D, W), params); `n_dem`
    overrides the drawn sphere count."""
    rng = np.random.default_rng(seed)
    lx = round(float(rng.uniform(0.3, 0.7)) / dx) * dx
    lz = round(float(rng.uniform(0.3, 0.6)) / dx) * dx
    drawn = int(rng.integers(0, max_dem + 1))
    n_dem = drawn if n_dem is None else int(n_dem)
    gen = CoupledGenerationConfig(device="cpu", dx=dx, dt=dt, fluid_width=lx, fluid_height=lz)
    P, B = DamPtlGeneration(gen).build()
    D = W = None
    if n_dem:
        lo, hi = (0.04, 0.04, lz + 0.06), (lx - 0.04, 0.36, lz + 0.36)
        pos, rad, _ = rsa_spheres(n_dem, lambda r, m: r.uniform(0.015, 0.03, m), lo, hi,
                                  seed + 7919, attempts=50_000, grow_z=0.02, max_grow=10,
                                  r_max=0.03)
        check_dem_inside(pos, rad, gen)
        D, W = DEMPtlGeneration(gen).build(dem_pos=pos, dem_radius=rad)
    return gen, (P, B, D, W), {"Lx": lx, "Lz": lz, "n_dem": n_dem, "seed": seed}


def c3(device="cuda:0", seed=20263, n_inst=64, steps=500, dx=0.01) -> Ensemble:
    scenes, params, gen = [], [], None
    for k in range(n_inst):
        gen, sc, prm = ensemble_scene(seed + k, dx=dx)
        scenes.append(sc)
        params.append(prm)
    solv = runtime_solv(gen, device)
    solv.n_steps = steps
    nf = sum(sc[0].n for sc in scenes)
    nd = sum(p["n_dem"] for p in params)
    note = f"{n_inst} instances, {nf} fluid, {nd} DEM, seeds {seed}..{seed + n_inst - 1}"
    return Ensemble("c3", solv, scenes, steps, params, note)


WORKLOADS = {"c0": c0, "c1": c1, "c2": c2}
ENSEMBLES = {"c3": c3}


@dataclass
class Cavity:
    name: str
    cells: int
    dt: float
    steps: int
    note: str


CAVITIES = {
    "e0": Cavity("e0", 25, 2.0e-3, 5000, "shipped Re=100 cavity, 25x25 fluid points"),
    "e1": Cavity("e1", 1024, 1.0e-5, 500, "Re=100 cavity, 1024x1024 fluid points"),
}


def cavity_case(name: str, device="cuda:0"):
    """(generator config, EISPHSolv, fluid, ghosts) of a cavity workload."""
    from .config import EISPHSolv
    from .input_gen import CavityPtlGeneration, GenerationConfig

    c = CAVITIES[name]
    gen = GenerationConfig(device="cpu", dx=1.0 / c.cells, dt=c.dt, n_steps=c.steps,
                           output_step=c.steps)
    P, B = CavityPtlGeneration(gen).build()
    solv = EISPHSolv(device=device, h=gen.h, rho0=gen.rho0, nu=gen.nu, dt=gen.dt,
                     n_steps=c.steps, output_step=c.steps, gif_save=False)
    return gen, solv, P.to(device), B.to(device)
