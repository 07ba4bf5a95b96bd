"""Initial-condition generators and the particle-file writer.

Mirrors input_gen/config.py:10-363 (GenerationConfig, CoupledGenerationConfig),
input_gen/gen_ptl.py:8-101 (DamPtlGeneration), input_gen/gen_dem.py:9-125
(DEMPtlGeneration), input_gen/gen_eisph.py:8-114 (CavityPtlGeneration) and
input_gen/generate.py:20-57 (write_particle_file, generate_files, main).  The lattice formulas, float32 rounding and the "%.9g"
tab-separated format are kept, so the reference's shipped input files are
reproduced byte for byte.  Seeded random sequential addition of DEM
spheres (`rsa_spheres`) is an addition for the larger workloads.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from .config import integers, positive, rule
from .input_reader import INPUT_FIELDS, column_indices
from .struct import BNDptl, DEMBNDptl, DEMptl, EISPHptl, SPHptl


@dataclass
class GenerationConfig:
    """Preprocessor settings (input_gen/config.py:10-175)."""

    device: str = "cuda:0"
    dx: float = 0.04
    tank_width: float = 2.0
    tank_height: float = 1.0
    tank_depth: float = 0.4
    fluid_width: float = 0.5
    fluid_height: float = 0.5
    fluid_depth: float = 0.4
    fluid_origin_x: float = 0.0
    fluid_origin_y: float = 0.0
    fluid_origin_z: float = 0.0
    bnd_layer: int = 3
    h: float | None = None
    support: float | None = None
    rho0: float = 1.0
    gamma: float = 7.0
    c0: float = 31.3209
    mu: float = 0.05
    g: float = 9.81
    h_factor: float = 1.3
    shepard_step: int = 20
    dt: float = 2.0e-3
    n_steps: int = 5000
    output_step: int = 100
    gif_save: bool = True
    output_dir: str = "result/3d"
    animation_dir: str = "animation/3d"
    grid_slice: int = 64
    dem_enable: bool = False
    dem_radius: float = 0.025
    dem_rho: float = 2500.0
    dem_K: float = 2.0e4
    dem_eta: float = 4.0
    dem_mu: float = 0.3
    dem_inertia_factor: float = 0.4
    dem_nx: int = 5
    dem_ny: int = 2
    dem_nz: int = 2
    dem_spacing: float = 0.065
    dem_origin_x: float = 0.12
    dem_origin_y: float = 0.1675
    dem_origin_z: float = 0.62
    dem_vel_x: float = 0.0
    dem_vel_y: float = 0.0
    dem_vel_z: float = 0.0
    dem_omega_x: float = 0.0
    dem_omega_y: float = 0.0
    dem_omega_z: float = 0.0
    dem_bnd_radius: float = 0.01
    dem_bnd_rho: float = 2500.0
    dem_bnd_spacing: float = 0.02
    dem_h: float = 0.08
    dem_porosity_min: float = 0.05
    dem_porosity_max: float = 1.0
    dem_dt_safety: float = 0.1
    length: float = 1.0
    lid_velocity: float = 1.0
    reynolds: float = 100.0
    gif_fps: int = 20
    animation_path: str = "animation/lid_driven_cavity.gif"

    def __post_init__(self) -> None:
        if self.h is None:
            self.h = self.h_factor * self.dx
        if self.support is None:
            self.support = 2.0 * self.h

    @property
    def dem_volume(self) -> float:
        return (4.0 / 3.0) * math.pi * self.dem_radius ** 3

    @property
    def dem_mass(self) -> float:
        return self.dem_rho * self.dem_volume

    @property
    def dem_inertia(self) -> float:
        return self.dem_inertia_factor * self.dem_mass * self.dem_radius ** 2

    @property
    def dem_bnd_volume(self) -> float:
        return (4.0 / 3.0) * math.pi * self.dem_bnd_radius ** 3

    @property
    def dem_bnd_mass(self) -> float:
        return self.dem_bnd_rho * self.dem_bnd_volume

    @property
    def dem_bnd_inertia(self) -> float:
        return self.dem_inertia_factor * self.dem_bnd_mass * self.dem_bnd_radius ** 2

    @property
    def dem_support(self) -> float:
        return 2.0 * self.dem_h

    @property
    def nu(self) -> float:
        """Cavity kinematic viscosity U_lid L / Re (input_gen/config.py:220-223)."""
        return self.lid_velocity * self.length / self.reynolds

    @property
    def cells(self) -> int:
        """Cell-centred fluid points along one cavity side."""
        return int(round(self.length / self.dx))

    # ---- validation (input_gen/config.py:227-353): the same rules, checked
    # in the same order, so that the first broken rule names itself
    def validate_eisph(self) -> None:
        """The cavity lattice, ghost layers covering the 2h support, and the
        explicit predictor's advection and viscosity dt limits
        (input_gen/config.py:269-305)."""
        positive(self, "length", "dx", "h_factor", "h", "support", "rho0", "lid_velocity",
                  "reynolds", "dt")
        integers(self, 1, "a positive integer", "bnd_layer", "n_steps", "output_step",
                  "gif_fps", "grid_slice")
        rule(math.isclose(self.h, self.h_factor * self.dx),
              "EISPH smoothing length must equal h_factor*dx")
        rule(math.isclose(self.support, 2.0 * self.h), "Wendland support must equal 2*h")
        rule(self.cells >= 5 and math.isclose(self.cells * self.dx, self.length, rel_tol=0.0,
                                               abs_tol=1.0e-10),
              "length must be an integer multiple of dx with at least 5 cells")
        rule(self.bnd_layer >= math.ceil(self.support / self.dx),
              "bnd_layer must cover the complete kernel support")
        rule(self.bnd_layer <= self.cells, "bnd_layer cannot exceed the cavity cell count")
        rule(self.dt <= 0.25 * self.dx / self.lid_velocity,
              "dt exceeds the explicit advection limit")
        rule(self.dt <= 0.125 * self.dx * self.dx / self.nu,
              "dt exceeds the explicit viscosity limit")

    def validate_sph(self) -> None:
        """The dam-break domain, wall lattice, material and the initial
        acoustic dt limit (input_gen/config.py:227-263)."""
        positive(self, "dx", "h", "support", "tank_width", "tank_height", "tank_depth",
                  "fluid_width", "fluid_height", "fluid_depth", "rho0", "c0", "gamma", "dt")
        rule(math.isclose(self.support, 2.0 * self.h), "Wendland support must equal 2*h")
        rule(self.dt <= 0.25 * self.h / self.c0,
              "dt exceeds the initial SPH acoustic limit 0.25*h/c0")
        for name in ("g", "mu"):
            v = getattr(self, name)
            rule(math.isfinite(v) and v >= 0.0, f"{name} must be finite and nonnegative")
        integers(self, 1, "a positive integer", "bnd_layer", "grid_slice")
        integers(self, 0, "a nonnegative integer", "n_steps", "output_step", "shepard_step")
        # the wall lattice is integer: tank extents are whole multiples of dx
        for name in ("tank_width", "tank_height", "tank_depth"):
            cells = getattr(self, name) / self.dx
            rule(math.isclose(cells, round(cells), abs_tol=1.0e-8),
                  f"{name} must be an integer multiple of dx")
        blocks = {"x": (self.fluid_width, self.tank_width),
                  "y": (self.fluid_depth, self.tank_depth),
                  "z": (self.fluid_height, self.tank_height)}
        for axis, (extent, tank) in blocks.items():
            lo = getattr(self, f"fluid_origin_{axis}")
            fits = (math.isfinite(lo) and lo >= 0.0 and lo + extent <= tank + 1.0e-12
                    and round(extent / self.dx) >= 1)
            rule(fits, f"initial fluid block must fit the tank along {axis}")

    def validate_dem(self) -> None:
        """DEM material, initial spacing, wall pores, porosity bounds and the
        contact dt limit of two identical moving spheres
        (input_gen/config.py:307-353); nothing to check without DEM."""
        if not self.dem_enable:
            return
        positive(self, "dem_radius", "dem_rho", "dem_K", "dem_inertia_factor",
                  "dem_bnd_radius", "dem_bnd_rho", "dem_bnd_spacing", "dem_h", "dem_dt_safety",
                  "dt")
        eta, mu = self.dem_eta, self.dem_mu
        rule(math.isfinite(eta) and math.isfinite(mu) and eta >= 0.0 and mu >= 0.0,
              "dem_eta and dem_mu must be nonnegative")
        counts = (self.dem_nx, self.dem_ny, self.dem_nz)
        rule(all(isinstance(c, int) and c >= 1 for c in counts),
              "dem_nx, dem_ny and dem_nz must be positive integers")
        rule(math.isfinite(self.dem_spacing) and self.dem_spacing >= 2.0 * self.dem_radius,
              "initial DEM particles must not overlap")
        rule(self.dem_bnd_spacing <= 2.0 * self.dem_bnd_radius,
              "DEM boundary spacing must not exceed the boundary diameter")
        # a moving centre must not pass a square pore of the wall lattice
        rule(self.dem_bnd_spacing / math.sqrt(2.0) < self.dem_radius + self.dem_bnd_radius,
              "DEM boundary pores must be smaller than the moving spheres")
        for name in ("dem_origin_x", "dem_origin_y", "dem_origin_z", "dem_vel_x", "dem_vel_y",
                     "dem_vel_z", "dem_omega_x", "dem_omega_y", "dem_omega_z"):
            rule(math.isfinite(getattr(self, name)), f"{name} must be finite")
        rule(0.0 < self.dem_porosity_min <= self.dem_porosity_max <= 1.0,
              "DEM porosity bounds must satisfy 0 < min <= max <= 1")
        m_star = 0.5 * self.dem_mass
        limit = self.dem_dt_safety * min(math.sqrt(m_star / self.dem_K),
                                         m_star / max(self.dem_eta, 1.0e-20))
        rule(self.dt <= limit, f"dt exceeds DEM contact limit {limit:.6g} s")


@dataclass
class CoupledGenerationConfig(GenerationConfig):
    """The 3-D SPH-DEM dam-break example (input_gen/config.py:355-363)."""

    dx: float = 0.02
    rho0: float = 1000.0
    dt: float = 1.0e-4
    n_steps: int = 9000
    output_step: int = 90
    dem_enable: bool = True


def _f32(a) -> torch.Tensor:
    """float64 numpy -> float32 tensor, rounding to nearest as Warp arrays do."""
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64)).astype(np.float32))


class DamPtlGeneration:
    """Fluid block and 3-layer dummy walls of the dam break (gen_ptl.py:8-101)."""

    def __init__(self, solv: GenerationConfig) -> None:
        self.solv = solv
        self.dx = solv.dx
        self.bnd_layer = solv.bnd_layer

    def fluid_particle(self) -> np.ndarray:
        s, dx = self.solv, self.dx
        nx, ny, nz = (int(round(s.fluid_width / dx)), int(round(s.fluid_depth / dx)),
                      int(round(s.fluid_height / dx)))
        gx, gy, gz = np.meshgrid(s.fluid_origin_x + np.arange(nx) * dx,
                                 s.fluid_origin_y + np.arange(ny) * dx,
                                 s.fluid_origin_z + np.arange(nz) * dx, indexing="ij")
        return np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)

    def boundary_particle(self) -> np.ndarray:
        """Bottom and four side walls, bnd_layer thick, open top; each lattice
        site once (gen_ptl.py:39-61)."""
        s, dx, L = self.solv, self.dx, self.bnd_layer
        nx, ny, nz = (int(round(s.tank_width / dx)), int(round(s.tank_depth / dx)),
                      int(round(s.tank_height / dx)))
        gx, gy, gz = np.meshgrid(np.arange(-L, nx + L), np.arange(-L, ny + L),
                                 np.arange(-L, nz + L), indexing="ij")
        wall = (gz < 0) | (gx < 0) | (gx >= nx) | (gy < 0) | (gy >= ny)
        return np.stack([gx[wall], gy[wall], gz[wall]], axis=1) * dx

    def build(self) -> tuple[SPHptl, BNDptl]:
        s = self.solv
        s.validate_sph()
        fpos, bpos = self.fluid_particle(), self.boundary_particle()
        mass = s.rho0 * s.dx ** 3
        dev = s.device
        P = SPHptl.allocate(len(fpos), device="cpu")
        P.pos = _f32(fpos)
        P.rho_raw.fill_(s.rho0)
        P.rho.fill_(s.rho0)
        P.m.fill_(mass)
        P.porosity.fill_(1.0)
        B = BNDptl.allocate(len(bpos), device="cpu")
        B.pos = _f32(bpos)
        B.rho_raw.fill_(s.rho0)
        B.rho.fill_(s.rho0)
        B.m.fill_(mass)
        return P.to(dev), B.to(dev)


class DEMPtlGeneration:
    """Moving DEM lattice and fixed DEM wall spheres (gen_dem.py:9-125)."""

    def __init__(self, solv: GenerationConfig) -> None:
        self.solv = solv

    def dem_particle(self) -> np.ndarray:
        s = self.solv
        gx, gy, gz = np.meshgrid(s.dem_origin_x + np.arange(s.dem_nx) * s.dem_spacing,
                                 s.dem_origin_y + np.arange(s.dem_ny) * s.dem_spacing,
                                 s.dem_origin_z + np.arange(s.dem_nz) * s.dem_spacing,
                                 indexing="ij")
        pos = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)
        check_dem_inside(pos, s.dem_radius, s)
        return pos

    def boundary_particle(self) -> np.ndarray:
        """Wall spheres one radius outside the tank, spacing <= dem_bnd_spacing,
        shell only, no ceiling (gen_dem.py:45-67)."""
        s = self.solv
        r = s.dem_bnd_radius
        nx = int(np.ceil((s.tank_width + 2.0 * r) / s.dem_bnd_spacing))
        ny = int(np.ceil((s.tank_depth + 2.0 * r) / s.dem_bnd_spacing))
        nz = int(np.ceil((s.tank_height + r) / s.dem_bnd_spacing))
        x = np.linspace(-r, s.tank_width + r, nx + 1)
        y = np.linspace(-r, s.tank_depth + r, ny + 1)
        z = np.linspace(-r, s.tank_height, nz + 1)
        gx, gy, gz = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1),
                                 indexing="ij")
        wall = (gz == 0) | (gx == 0) | (gx == nx) | (gy == 0) | (gy == ny)
        return np.stack([x[gx[wall]], y[gy[wall]], z[gz[wall]]], axis=1)

    def build(self, dem_pos: np.ndarray | None = None,
              dem_radius: np.ndarray | None = None) -> tuple[DEMptl, DEMBNDptl]:
        """Allocate DEM and DEM-wall structs.  `dem_pos`/`dem_radius` replace
        the lattice with given centres and (per-sphere) radii."""
        s = self.solv
        s.validate_dem()
        dpos = self.dem_particle() if dem_pos is None else np.asarray(dem_pos, np.float64)
        bpos = self.boundary_particle()
        nd, nb = len(dpos), len(bpos)
        D = DEMptl.allocate(nd, device="cpu")
        W = DEMBNDptl.allocate(nb, device="cpu")
        D.pos, W.pos = _f32(dpos), _f32(bpos)
        if dem_radius is None:
            D.radius.fill_(s.dem_radius)
            D.rho.fill_(s.dem_rho)
            D.m.fill_(s.dem_mass)
            D.volume.fill_(s.dem_volume)
            D.inertia.fill_(s.dem_inertia)
        else:
            r = np.broadcast_to(np.asarray(dem_radius, np.float64), (nd,))
            vol = (4.0 / 3.0) * math.pi * r ** 3
            mass = s.dem_rho * vol
            D.radius, D.rho = _f32(r), _f32(np.full(nd, s.dem_rho))
            D.m, D.volume = _f32(mass), _f32(vol)
            D.inertia = _f32(s.dem_inertia_factor * mass * r ** 2)
        for T in (D, W):
            T.K.fill_(s.dem_K)
            T.eta.fill_(s.dem_eta)
            T.mu.fill_(s.dem_mu)
        W.radius.fill_(s.dem_bnd_radius)
        W.rho.fill_(s.dem_bnd_rho)
        W.m.fill_(s.dem_bnd_mass)
        W.volume.fill_(s.dem_bnd_volume)
        W.inertia.fill_(s.dem_bnd_inertia)
        D.h.fill_(s.dem_h)
        D.vel[:] = torch.tensor([s.dem_vel_x, s.dem_vel_y, s.dem_vel_z], dtype=torch.float32)
        D.omega[:] = torch.tensor([s.dem_omega_x, s.dem_omega_y, s.dem_omega_z],
                                  dtype=torch.float32)
        D.porosity.fill_(1.0)
        return D.to(s.device), W.to(s.device)


class CavityPtlGeneration:
    """Fixed points of the lid-driven cavity (input_gen/gen_eisph.py:8-114).

    The square cavity lies in the x-z plane: cell-centred fluid points, and
    ghost points continuing the lattice bnd_layer cells beyond the four
    walls.  Each ghost reflects through its nearest wall onto a fluid row
    (`mirror`); ghosts above the cavity (the lid) prescribe u = lid_velocity,
    the others no slip."""

    def __init__(self, solv: GenerationConfig) -> None:
        self.solv = solv

    def fluid_particle(self) -> np.ndarray:
        """Cell-centred fluid positions [cells^2, 3], x fastest."""
        n = self.solv.cells
        axis = (np.arange(n, dtype=np.float32) + 0.5) * self.solv.dx
        gx, gz = np.meshgrid(axis, axis, indexing="xy")
        return np.stack((gx.ravel(), np.zeros(gx.size, np.float32), gz.ravel()), axis=1)

    def boundary_particle(self) -> tuple:
        """Ghost positions [N_bnd, 3], prescribed velocity [N_bnd, 3] and
        mirrored fluid row [N_bnd]."""
        n, layers, dx = self.solv.cells, self.solv.bnd_layer, self.solv.dx
        ids = np.arange(-layers, n + layers, dtype=np.int32)
        gx, gz = np.meshgrid(ids, ids, indexing="xy")
        shell = (gx < 0) | (gx >= n) | (gz < 0) | (gz >= n)
        ix, iz = gx[shell], gz[shell]
        pos = np.stack(((ix + 0.5) * dx, np.zeros(ix.size), (iz + 0.5) * dx),
                       axis=1).astype(np.float32)
        vel_bc = np.zeros_like(pos)
        vel_bc[(iz >= n) & (ix >= 0) & (ix < n), 0] = self.solv.lid_velocity
        mx = np.where(ix < 0, -ix - 1, np.where(ix >= n, 2 * n - ix - 1, ix))
        mz = np.where(iz < 0, -iz - 1, np.where(iz >= n, 2 * n - iz - 1, iz))
        mirror = (mz * n + mx).astype(np.int32)
        if np.any((mirror < 0) | (mirror >= n * n)):
            raise ValueError("bnd_layer cannot exceed the cavity cell count")
        return pos, vel_bc, mirror

    def _build(self, pos, vel, vel_bc, mirror) -> EISPHptl:
        solv = self.solv
        n = len(pos)
        p = EISPHptl.allocate(n)
        p.pos = _f32(pos)
        p.vel = _f32(vel)
        p.vel_star = _f32(vel)
        p.vel_bc = _f32(vel_bc)
        p.rho = torch.full((n,), solv.rho0, dtype=torch.float32)
        p.m = torch.full((n,), solv.rho0 * solv.dx ** 2, dtype=torch.float32)
        p.mirror = torch.from_numpy(np.asarray(mirror, dtype=np.int32))
        return p.to(solv.device) if str(solv.device) != "cpu" else p

    def build(self) -> tuple:
        """(fluid, ghost) EISPHptl structs on solv.device; the fluid starts at
        rest, so the ghost velocity starts at 2 vel_bc."""
        self.solv.validate_eisph()
        fluid_pos = self.fluid_particle()
        bpos, vel_bc, mirror = self.boundary_particle()
        zero = np.zeros_like(fluid_pos)
        return (self._build(fluid_pos, zero, zero, np.arange(len(fluid_pos), dtype=np.int32)),
                self._build(bpos, 2.0 * vel_bc, vel_bc, mirror))


def check_dem_inside(pos: np.ndarray, radius, s: GenerationConfig) -> None:
    """Sphere extents inside the five closed tank sides (gen_dem.py:32-40)."""
    r = np.broadcast_to(np.asarray(radius, np.float64), (len(pos),))
    if len(pos) and (np.min(pos[:, 0] - r) < 0 or np.max(pos[:, 0] + r) > s.tank_width
                     or np.min(pos[:, 1] - r) < 0 or np.max(pos[:, 1] + r) > s.tank_depth
                     or np.min(pos[:, 2] - r) < 0):
        raise ValueError("initial DEM spheres must lie inside the tank walls")


def rsa_spheres(n: int, radius, box_lo, box_hi, seed: int, attempts: int = 200_000,
                grow_z: float = 0.0, max_grow: int = 200, r_max: float | None = None,
                batch: int = 4096):
    """Seeded random sequential addition of non-overlapping spheres.

    Centres are drawn uniformly in [box_lo, box_hi] (numpy default_rng(seed),
    three doubles per draw); a draw is kept when its distance to every kept
    centre is at least the sum of the radii.  `radius` is a scalar or a
    sampler (rng, size) -> radii bounded by r_max.  If `attempts` draws do
    not place n spheres and grow_z > 0, the box top is raised by grow_z and
    drawing continues (at most max_grow times).  Draws are tested in blocks
    against the kept spheres through a bucket grid, then in draw order
    against each other, which keeps exactly the sequential outcome.
    Returns (centres, radii, box_hi_used) in placement order.
    """
    rng = np.random.default_rng(seed)
    lo = np.asarray(box_lo, np.float64)
    hi = np.asarray(box_hi, np.float64).copy()
    sampler = callable(radius)
    if sampler and r_max is None:
        raise ValueError("r_max is required with a radius sampler")
    rmax = float(r_max) if sampler else float(radius)
    cell = 2.0 * rmax
    top = hi[2] + (grow_z * max_grow if grow_z > 0 else 0.0)
    dims = np.ceil((np.array([hi[0], hi[1], top]) - lo) / cell).astype(int) + 3
    slots = 8  # centres at least `cell` apart: at most 8 per cubic bucket
    table = np.full((*dims, slots), -1, np.int64)
    fill = np.zeros(dims, np.int64)
    pos = np.zeros((n, 3))
    rad = np.zeros(n)
    k = 0
    offs = np.array([(i, j, l) for i in (-1, 0, 1) for j in (-1, 0, 1) for l in (-1, 0, 1)])

    def bucket(c):
        return np.floor((c - lo) / cell).astype(int) + 1

    for _ in range(max_grow + 1):
        left = attempts
        while left > 0 and k < n:
            m = min(batch, left)
            left -= m
            c = lo + rng.random((m, 3)) * (hi - lo)
            r = radius(rng, m) if sampler else np.full(m, rmax)
            bc = bucket(c)
            nb = bc[:, None, :] + offs[None]
            nb = np.clip(nb, 0, dims - 1)
            ids = table[nb[..., 0], nb[..., 1], nb[..., 2]].reshape(m, -1)
            valid = ids >= 0
            idx = np.where(valid, ids, 0)
            d2 = np.sum((pos[idx] - c[:, None, :]) ** 2, axis=2)
            clash = valid & (d2 < (rad[idx] + r[:, None]) ** 2)
            ok = ~clash.any(axis=1)
            start = k
            for j in np.nonzero(ok)[0]:
                if k == n:
                    break
                if k > start:
                    dd = np.sum((pos[start:k] - c[j]) ** 2, axis=1)
                    if np.any(dd < (rad[start:k] + r[j]) ** 2):
                        continue
                pos[k], rad[k] = c[j], r[j]
                b = tuple(bc[j])
                table[b][fill[b]] = k
                fill[b] += 1
                k += 1
        if k == n or grow_z <= 0.0:
            break
        hi[2] += grow_z
    return pos[:k], rad[:k], hi


def write_particle_file(path: str | Path, kind: str, particle) -> Path:
    """Write the input columns of one struct with integer column IDs
    (input_gen/generate.py:20-31): '%.9g', tab separated, no comment mark."""
    destination = Path(path)
    destination.parent.mkdir(parents=True, exist_ok=True)
    n = particle.pos.shape[0]
    cols = [getattr(particle, name).detach().cpu().numpy().reshape(n, width)
            for name, width in INPUT_FIELDS[kind]]
    header = "\t".join(str(i) for ids in column_indices(kind).values() for i in ids)
    np.savetxt(destination, np.concatenate(cols, axis=1), fmt="%.9g", delimiter="\t",
               header=header, comments="", encoding="utf-8")
    return destination


def generate_files(output_dir: str | Path,
                   coupled: CoupledGenerationConfig | None = None,
                   cavity: GenerationConfig | None = None) -> dict:
    """Write the six input files: input_{SPH,BND,DEM,DEMBND}.txt of the dam
    break and input_{EISPH,EISPHBND}.txt of the cavity
    (input_gen/generate.py:34-44)."""
    coupled = CoupledGenerationConfig(device="cpu") if coupled is None else coupled
    cavity = GenerationConfig(device="cpu") if cavity is None else cavity
    particles = dict(zip(("SPH", "BND"), DamPtlGeneration(coupled).build()))
    particles.update(zip(("DEM", "DEMBND"), DEMPtlGeneration(coupled).build()))
    particles.update(zip(("EISPH", "EISPHBND"), CavityPtlGeneration(cavity).build()))
    return {kind: write_particle_file(Path(output_dir) / f"input_{kind}.txt", kind, p)
            for kind, p in particles.items()}


def write_structs(output_dir: str | Path, structs) -> dict:
    """input_{SPH,BND,DEM,DEMBND}.txt of given (P, B, D, W) structs."""
    return {kind: write_particle_file(Path(output_dir) / f"input_{kind}.txt", kind, p)
            for kind, p in zip(("SPH", "BND", "DEM", "DEMBND"), structs)}


def main(argv=None) -> None:
    """python -m warp_sph_gpt_amd.input_gen [--output-dir DIR] [--workload c0|c1|c2]

    c0 (default) writes the reference's shipped case; c1 and c2 write the
    larger dam breaks with their seeded random sequential DEM placement."""
    import argparse
    parser = argparse.ArgumentParser(description="Generate the SPH/DEM dam-break input files")
    parser.add_argument("--output-dir", type=Path, default=Path("input_file"))
    parser.add_argument("--workload", choices=("c0", "c1", "c2"), default="c0")
    args = parser.parse_args(argv)
    if args.workload == "c0":
        out = generate_files(args.output_dir)
    else:
        from .workloads import WORKLOADS
        out = write_structs(args.output_dir, WORKLOADS[args.workload](device="cpu").structs)
    for kind, path in out.items():
        print(f"[{kind}] {path}")


if __name__ == "__main__":
    main()
