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
All: tank 2 x 0.4 x 1 m, fluid block 0.5 x 0.4 x 0.5 m at the origin,
3-layer dummy wall, h = 1.3 dx, rho0 = 1000, DEM h = 0.08, K=2e4, eta=4,
mu=0.3, rho_s=2500, DEM-wall radius 0.01.
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


WORKLOADS = {"c0": c0, "c1": c1, "c2": c2}
