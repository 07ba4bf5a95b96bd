"""Shared fixtures of the test suite: scenes, conversions and error metrics."""
from __future__ import annotations

import numpy as np
import torch

from oracle import oracle as O
from warp_sph_gpt_amd.config import Solv
from warp_sph_gpt_amd.input_gen import (CoupledGenerationConfig, DamPtlGeneration,
                                        DEMPtlGeneration)

# The compact dam break of the reference tests (tests/test_dem.py:40-66).
COMPACT = dict(dx=0.04, h=0.052, support=0.104, tank_width=0.8, tank_height=0.8,
               tank_depth=0.32, fluid_width=0.4, fluid_height=0.4, fluid_depth=0.32,
               dem_nx=2, dem_ny=1, dem_nz=1, dem_origin_y=0.16, dem_origin_z=0.60,
               grid_slice=32)


def generation(device="cpu", **kw) -> CoupledGenerationConfig:
    return CoupledGenerationConfig(device=device, **kw)


def particles(gen: CoupledGenerationConfig):
    P, B = DamPtlGeneration(gen).build()
    D, W = DEMPtlGeneration(gen).build()
    return P, B, D, W


def solv_for(gen: CoupledGenerationConfig, device="cuda:0") -> Solv:
    """Runtime Solv with the generator's solver controls (as test_input_reader.runtime)."""
    s = Solv(device=device)
    for name in ("h", "rho0", "c0", "gamma", "mu", "g", "shepard_step", "grid_slice", "dt",
                 "n_steps", "dem_porosity_min", "dem_porosity_max", "dem_dt_safety"):
        setattr(s, name, getattr(gen, name))
    return s


def to_oracle(structs, dtype=np.float32):
    return [O.OStruct.from_struct(s, dtype) for s in structs]


def np_(t) -> np.ndarray:
    return t.detach().cpu().numpy()


def rel_err(a, b, floor: float = 0.0) -> float:
    """max_i |a_i - b_i| / max(max_i |b_i|, floor), rows as vectors."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.shape != b.shape:
        raise AssertionError(f"shape {a.shape} != {b.shape}")
    if a.size == 0:
        return 0.0
    if a.ndim == 2:
        d = np.linalg.norm(a - b, axis=1).max()
        s = np.linalg.norm(b, axis=1).max()
    else:
        d = np.abs(a - b).max()
        s = np.abs(b).max()
    return float(d / max(s, floor, 1e-300))


def assert_rel(a, b, tol: float, floor: float = 0.0, what: str = ""):
    e = rel_err(a, b, floor)
    assert e <= tol, f"{what}: relative error {e:.3e} > {tol:.1e}"
    return e
