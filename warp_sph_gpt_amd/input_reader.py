"""Particle input files: numeric-header column-ID text format.

Mirrors input/input_reader.py:1-220: the column IDs of each kind, the
validation and error messages of `_read_table`, `input_parser`,
`validate_dem_particles`, `dem_search_scales`, `load_sph_particles`,
`validate_eisph_particles` and `load_eisph_particles`.  Structs are returned
as torch tensors on the requested device.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import torch

from .struct import BNDptl, DEMBNDptl, DEMptl, EISPHptl, SPHptl

INPUT_FIELDS = {
    "SPH": (("pos", 3), ("vel", 3), ("rho", 1), ("m", 1)),
    "BND": (("pos", 3), ("vel", 3), ("rho", 1), ("m", 1)),
    "DEM": (("pos", 3), ("vel", 3), ("omega", 3), ("radius", 1), ("rho", 1), ("m", 1),
            ("inertia", 1), ("K", 1), ("eta", 1), ("mu", 1), ("h", 1)),
    "DEMBND": (("pos", 3), ("vel", 3), ("omega", 3), ("radius", 1), ("rho", 1), ("m", 1),
               ("inertia", 1), ("K", 1), ("eta", 1), ("mu", 1)),
    "EISPH": (("pos", 3), ("vel", 3), ("rho", 1), ("m", 1)),
    "EISPHBND": (("pos", 3), ("vel_bc", 3), ("rho", 1), ("m", 1), ("mirror", 1)),
}
STRUCTURES = {"SPH": SPHptl, "BND": BNDptl, "DEM": DEMptl, "DEMBND": DEMBNDptl,
              "EISPH": EISPHptl, "EISPHBND": EISPHptl}
PROJECT_ROOT = Path(__file__).resolve().parents[1]
_F32_MAX = float(np.finfo(np.float32).max)


def column_indices(kind: str) -> dict:
    """Documented numeric IDs of each input field; IDs restart at 1 per kind."""
    if kind not in INPUT_FIELDS:
        raise ValueError(f"unknown particle kind {kind!r}; choose {tuple(INPUT_FIELDS)}")
    index, result = 1, {}
    for field, width in INPUT_FIELDS[kind]:
        result[field] = tuple(range(index, index + width))
        index += width
    return result


def _read_table(path: Path, kind: str) -> dict:
    columns = column_indices(kind)
    required = {i for ids in columns.values() for i in ids}
    with path.open(encoding="utf-8-sig") as stream:
        try:
            header = [int(token) for token in stream.readline().split()]
        except ValueError as exc:
            raise ValueError(f"{path}:1: header must contain integer column IDs") from exc
        if not header or len(header) != len(set(header)):
            raise ValueError(f"{path}:1: empty header or duplicate column IDs")
        if set(header) != required:
            raise ValueError(f"{path}:1: missing IDs {sorted(required - set(header))}; "
                             f"unknown IDs {sorted(set(header) - required)}")
        rows = []
        for line_number, line in enumerate(stream, 2):
            tokens = line.split()
            if len(tokens) != len(header):
                raise ValueError(f"{path}:{line_number}: expected {len(header)} values, "
                                 f"got {len(tokens)}")
            try:
                row = [float(token) for token in tokens]
            except ValueError as exc:
                raise ValueError(f"{path}:{line_number}: invalid numeric value") from exc
            if not all(np.isfinite(row)) or any(abs(v) > _F32_MAX for v in row):
                raise ValueError(f"{path}:{line_number}: values must be finite float32 numbers")
            rows.append(row)
    data = np.asarray(rows, dtype=np.float64).reshape(-1, len(header))
    result = {}
    for field, ids in columns.items():
        values = data[:, [header.index(i) for i in ids]]
        result[field] = values[:, 0] if len(ids) == 1 else values
    for field in ("rho", "m", "radius", "inertia", "K", "h"):
        if field in result and np.any(result[field].astype(np.float32) <= 0.0):
            raise ValueError(f"{path}: {field} must be positive in float32")
    for field in ("eta", "mu"):
        if field in result and np.any(result[field] < 0.0):
            raise ValueError(f"{path}: {field} must be nonnegative")
    if "mirror" in result:
        ids = result["mirror"]
        if np.any((ids != np.floor(ids)) | (ids < 0) | (ids > np.iinfo(np.int32).max)):
            raise ValueError(f"{path}: mirror must be a nonnegative int32 fluid row index")
    if "radius" in result:
        volume = (4.0 / 3.0) * np.pi * result["radius"] ** 3
        if not np.allclose(result["m"], result["rho"] * volume, rtol=1e-5, atol=0.0):
            raise ValueError(f"{path}: sphere mass must equal rho*(4/3)*pi*radius^3")
        v32 = volume.astype(np.float32)
        if np.any(v32 <= 0) or not np.isfinite(v32).all():
            raise ValueError(f"{path}: derived sphere volume is outside float32 range")
    return result


def input_parser(file_path, kind: str | None = None, device="cpu"):
    """Read input_<kind>.txt into a fully allocated struct (input_reader.py:98-134).

    Computed fields start at zero, rho_raw copies rho, fluid and DEM porosity
    start at one, DEM volume is (4/3)*pi*radius^3, and both contact CSR sides
    are empty with zero offsets.  EISPH fluid rows get vel_star = vel and the
    identity mirror; ghost mirror rows are read as int32."""
    path = Path(file_path)
    kind = path.stem.removeprefix("input_") if kind is None else kind
    if kind not in STRUCTURES:
        raise ValueError(f"unknown particle kind {kind!r}; choose {tuple(STRUCTURES)}")
    data = _read_table(path, kind)
    n = len(data["pos"])
    particle = STRUCTURES[kind].allocate(n, device="cpu")
    for name, values in data.items():
        if name == "mirror":
            setattr(particle, name, torch.from_numpy(np.asarray(values).astype(np.int32)))
            continue
        # column selection leaves numpy in column-major order; the kernels
        # take packed row-major (N, 3) arrays
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(values, np.float64),
                                                  dtype=np.float32))
        setattr(particle, name, t.reshape(n, 3) if t.ndim == 2 else t)
    if kind in ("SPH", "BND"):
        particle.rho_raw = particle.rho.clone()
    if kind in ("SPH", "DEM"):
        particle.porosity.fill_(1.0)
    if kind in ("DEM", "DEMBND"):
        volume = (4.0 / 3.0) * np.pi * data["radius"] ** 3
        particle.volume = torch.from_numpy(np.asarray(volume, np.float64).astype(np.float32))
    if kind == "EISPH":
        particle.vel_star = particle.vel.clone()
        particle.mirror = torch.arange(n, dtype=torch.int32)
    return particle.to(device) if str(device) != "cpu" else particle


def input_directory(solv) -> Path:
    directory = Path(solv.input_dir)
    return directory if directory.is_absolute() else PROJECT_ROOT / directory


def dem_search_scales(dem, boundary) -> tuple:
    """(max DEM radius, max DEM-wall radius or 0, max h, common h or 0)
    (input_reader.py:163-168)."""
    h = dem.h.detach().cpu().numpy()
    rb = boundary.radius.detach().cpu().numpy()
    return (float(dem.radius.detach().cpu().numpy().max()),
            float(rb.max()) if boundary.pos.shape[0] else 0.0,
            float(h.max()), float(h[0]) if np.all(h == h[0]) else 0.0)


def validate_dem_particles(solv, dem, boundary) -> tuple:
    """Contact dt limit 0.1*min(sqrt(m_eff/K_max), m_eff/eta_max), m_eff = min(m)/2
    (input_reader.py:142-160); returns dem_search_scales."""
    solv.validate_dem()
    if dem.pos.shape[0] == 0:
        raise ValueError("DEM enabled but input_DEM.txt contains no particles")
    mass = dem.m.detach().cpu().numpy().astype(float)
    stiffness = dem.K.detach().cpu().numpy().astype(float)
    damping = dem.eta.detach().cpu().numpy().astype(float)
    m_eff = 0.5 * mass.min()
    limit = solv.dem_dt_safety * min(np.sqrt(m_eff / stiffness.max()),
                                     m_eff / max(float(damping.max()), 1e-20))
    if solv.dt > limit:
        raise ValueError(f"dt exceeds loaded DEM contact limit {limit:.6g} s")
    return dem_search_scales(dem, boundary)


def load_sph_particles(solv):
    """Read the two or four WCSPH particle files selected by dem_enable
    (input_reader.py:171-184)."""
    solv.validate_sph()
    directory = input_directory(solv)
    sph, bnd = (input_parser(directory / f"input_{k}.txt", k, solv.device) for k in ("SPH", "BND"))
    if sph.pos.shape[0] == 0:
        raise ValueError("input_SPH.txt must contain fluid particles")
    dem = dem_bnd = None
    if solv.dem_enable:
        dem, dem_bnd = (input_parser(directory / f"input_{k}.txt", k, solv.device)
                        for k in ("DEM", "DEMBND"))
        validate_dem_particles(solv, dem, dem_bnd)
    return sph, bnd, dem, dem_bnd


def validate_eisph_particles(solv, fluid, boundary) -> None:
    """A fixed planar EISPH cloud and its mirror map (input_reader.py:187-208):
    nonempty sets, mirror rows inside the fluid file, uniform rho = rho0,
    points in the x-z plane, and the explicit predictor's advection and
    viscosity dt limits for the loaded speeds and spacing."""
    solv.validate_eisph()
    n = fluid.pos.shape[0]
    if n == 0 or boundary.pos.shape[0] == 0:
        raise ValueError("EISPH needs nonempty fluid and ghost files")
    if np.any(boundary.mirror.detach().cpu().numpy() >= n):
        raise ValueError("EISPHBND mirror is outside the input_EISPH fluid row range")
    for particle in (fluid, boundary):
        if not np.allclose(particle.rho.detach().cpu().numpy(), solv.rho0, rtol=1e-6, atol=0):
            raise ValueError("EISPH currently requires uniform rho equal to rho0")
        if not np.allclose(particle.pos.detach().cpu().numpy()[:, 1], 0.0, atol=1e-8):
            raise ValueError("EISPH points must lie in the x-z plane (y=0)")
    m = fluid.m.detach().cpu().numpy()
    rho = fluid.rho.detach().cpu().numpy()
    dx = float(np.sqrt(m / rho).min())
    speed = max(float(np.linalg.norm(fluid.vel.detach().cpu().numpy(), axis=1).max()),
                float(np.linalg.norm(boundary.vel_bc.detach().cpu().numpy(), axis=1).max()))
    if speed > 0 and solv.dt > 0.25 * dx / speed:
        raise ValueError("dt exceeds the loaded EISPH initial advection limit")
    if solv.nu > 0 and solv.dt > 0.125 * dx * dx / solv.nu:
        raise ValueError("dt exceeds the loaded EISPH viscosity limit")


def load_eisph_particles(solv):
    """input_EISPH.txt and input_EISPHBND.txt, validated, with the initial
    ghost velocity 2 vel_bc - vel[mirror] in vel and vel_star
    (input_reader.py:211-220)."""
    directory = input_directory(solv)
    fluid, boundary = (input_parser(directory / f"input_{k}.txt", k, solv.device)
                       for k in ("EISPH", "EISPHBND"))
    validate_eisph_particles(solv, fluid, boundary)
    mirror = boundary.mirror.long()
    boundary.vel = 2.0 * boundary.vel_bc - fluid.vel[mirror]
    boundary.vel_star = boundary.vel.clone()
    return fluid, boundary
