"""Solver controls of the coupled WCSPH-DEM step.

Mirrors input/Config.py:11-93 (`Solv` fields, `support`, validation; the
EISPH cavity defaults, here `EISPHSolv`) and input/Config_SPH_DEM.py:6-15
(the WCSPH-DEM defaults, here `Solv`).  Geometry and particle properties
come from the input files; only solver controls live here.
"""
from __future__ import annotations

import math
from dataclasses import dataclass


@dataclass
class BaseSolv:
    """Shared solver controls with the generic defaults (input/Config.py:11-45)."""

    input_dir: str = "input_file"
    device: str = "cuda:0"

    h: float = 0.052
    rho0: float = 1.0
    c0: float = 31.3209
    gamma: float = 7.0
    mu: float = 0.05
    g: float = 9.81
    nu: float = 0.01
    shepard_step: int = 20
    grid_slice: int = 64

    dt: float = 2.0e-3
    n_steps: int = 5000

    dem_enable: bool = False
    dem_porosity_min: float = 0.05
    dem_porosity_max: float = 1.0
    dem_dt_safety: float = 0.1

    output_step: int = 100
    gif_save: bool = True
    gif_fps: int = 20
    output_dir: str = "result/3d"
    animation_dir: str = "animation/3d"
    animation_path: str = "animation/lid_driven_cavity.gif"

    @property
    def support(self) -> float:
        """Wendland C2 support radius, 2*h (input/Config.py:47-50)."""
        return 2.0 * self.h

    def _validate_common(self) -> None:
        for name in ("h", "rho0", "dt"):
            value = getattr(self, name)
            if not math.isfinite(value) or value <= 0.0:
                raise ValueError(f"{name} must be finite and positive")
        for name in ("n_steps", "output_step", "shepard_step"):
            value = getattr(self, name)
            if type(value) is not int or value < 0:
                raise ValueError(f"{name} must be a nonnegative integer")
        for name in ("grid_slice", "gif_fps"):
            value = getattr(self, name)
            if type(value) is not int or value < 1:
                raise ValueError(f"{name} must be a positive integer")

    def validate_sph(self) -> None:
        """input/Config.py:65-78: solver settings and the acoustic dt limit."""
        self._validate_common()
        for name in ("c0", "gamma"):
            value = getattr(self, name)
            if not math.isfinite(value) or value <= 0.0:
                raise ValueError(f"{name} must be finite and positive")
        for name in ("mu", "g"):
            value = getattr(self, name)
            if not math.isfinite(value) or value < 0.0:
                raise ValueError(f"{name} must be finite and nonnegative")
        if self.dt > 0.25 * self.h / self.c0:
            raise ValueError("dt exceeds the initial SPH acoustic limit 0.25*h/c0")
        self.validate_dem()

    def validate_eisph(self) -> None:
        """input/Config.py:89-93: EISPH controls; the loaded points supply the
        geometry and the wall speeds."""
        self._validate_common()
        if not math.isfinite(self.nu) or self.nu < 0.0:
            raise ValueError("nu must be finite and nonnegative")

    def validate_dem(self) -> None:
        """input/Config.py:80-87: coupling controls."""
        if not self.dem_enable:
            return
        if not 0.0 < self.dem_porosity_min <= self.dem_porosity_max <= 1.0:
            raise ValueError("DEM porosity bounds must satisfy 0 < min <= max <= 1")
        if not math.isfinite(self.dem_dt_safety) or self.dem_dt_safety <= 0.0:
            raise ValueError("dem_dt_safety must be finite and positive")


@dataclass
class Solv(BaseSolv):
    """WCSPH-DEM defaults (input/Config_SPH_DEM.py:6-15)."""

    h: float = 0.026
    rho0: float = 1000.0
    dt: float = 1.0e-4
    n_steps: int = 9000
    output_step: int = 90
    dem_enable: bool = True


@dataclass
class EISPHSolv(BaseSolv):
    """The Eulerian ISPH lid-driven cavity defaults (input/Config.py:11-45):
    h = 0.052, rho0 = 1, nu = 0.01, dt = 2e-3, 5000 steps, a frame every 100."""
