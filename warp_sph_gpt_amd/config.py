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

    # ---- validation: each rule in the reference's order (the first broken
    # one names itself); `exact` integers exclude bool, as type(v) is int does
    def _validate_common(self) -> None:
        positive(self, "h", "rho0", "dt")
        integers(self, 0, "a nonnegative integer", "n_steps", "output_step", "shepard_step",
                 exact=True)
        integers(self, 1, "a positive integer", "grid_slice", "gif_fps", exact=True)

    def validate_sph(self) -> None:
        """input/Config.py:65-78: solver settings and the acoustic dt limit."""
        self._validate_common()
        positive(self, "c0", "gamma")
        nonnegative(self, "mu", "g")
        rule(self.dt <= 0.25 * self.h / self.c0,
             "dt exceeds the initial SPH acoustic limit 0.25*h/c0")
        self.validate_dem()

    def validate_eisph(self) -> None:
        """input/Config.py:89-93: EISPH controls; the loaded points supply the
        geometry and the wall speeds."""
        self._validate_common()
        nonnegative(self, "nu")

    def validate_dem(self) -> None:
        """input/Config.py:80-87: coupling controls."""
        if not self.dem_enable:
            return
        rule(0.0 < self.dem_porosity_min <= self.dem_porosity_max <= 1.0,
             "DEM porosity bounds must satisfy 0 < min <= max <= 1")
        positive(self, "dem_dt_safety")


def rule(holds: bool, message: str) -> None:
    """Raise ValueError(message) unless the rule holds."""
    if not holds:
        raise ValueError(message)


def positive(cfg, *names) -> None:
    for name in names:
        v = getattr(cfg, name)
        rule(math.isfinite(v) and v > 0.0, f"{name} must be finite and positive")


def nonnegative(cfg, *names) -> None:
    for name in names:
        v = getattr(cfg, name)
        rule(math.isfinite(v) and v >= 0.0, f"{name} must be finite and nonnegative")


def integers(cfg, minimum: int, what: str, *names, exact: bool = False) -> None:
    for name in names:
        v = getattr(cfg, name)
        is_int = type(v) is int if exact else isinstance(v, int)
        rule(is_int and v >= minimum, f"{name} must be {what}")


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
