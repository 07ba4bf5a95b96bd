"""Particle input files: numeric-header column-ID text format.

Interface of input/input_reader.py:1-220, kept so that the reference's files
and callers work unchanged: one file per particle kind, a first line of
integer column IDs (IDs restart at 1 per kind, vec3 fields take three
consecutive IDs, columns may come in any order), then one particle per row.
The column IDs, the inputs that are rejected and the error messages follow
the reference; the reader itself is organised differently:

  * a `Layout` per kind maps each ID to (field, component);
  * the body is tokenised once and converted by numpy in one pass, and only
    when that fails is the first offending line searched for, so that the
    error names it;
  * value checks are a table of (fields, predicate, message);
  * fields the file does not carry are filled by a per-kind table of
    initialisers when the struct is built.

Structs are returned as torch tensors on the requested device.
"""
from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from .struct import BNDptl, DEMBNDptl, DEMptl, EISPHptl, SPHptl

# (field, components) in ID order, per kind (input/input_reader.py:18-31).
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
_F32 = np.finfo(np.float32)
_I32_MAX = np.iinfo(np.int32).max


@dataclass(frozen=True)
class Layout:
    """Column IDs of one particle kind."""
    kind: str
    ids: dict            # field -> tuple of IDs

    @classmethod
    def of(cls, kind: str) -> "Layout":
        if kind not in INPUT_FIELDS:
            raise ValueError(f"unknown particle kind {kind!r}; choose {tuple(INPUT_FIELDS)}")
        ids, start = {}, 1
        for field, width in INPUT_FIELDS[kind]:
            ids[field] = tuple(range(start, start + width))
            start += width
        return cls(kind, ids)

    @property
    def required(self) -> set:
        return {i for group in self.ids.values() for i in group}


def column_indices(kind: str) -> dict:
    """Documented numeric IDs of each input field; IDs restart at 1 per kind."""
    return dict(Layout.of(kind).ids)


def _header(path: Path, line: str, layout: Layout) -> list:
    try:
        header = [int(tok) for tok in line.split()]
    except ValueError as exc:
        raise ValueError(f"{path}:1: header must contain integer column IDs") from exc
    present = set(header)
    if not header or len(present) < len(header):
        raise ValueError(f"{path}:1: empty header or duplicate column IDs")
    need = layout.required
    if present != need:
        raise ValueError(f"{path}:1: missing IDs {sorted(need - present)}; "
                         f"unknown IDs {sorted(present - need)}")
    return header


def _body(path: Path, lines: list, width: int) -> np.ndarray:
    """The rows as an (n, width) float64 array; errors name the first bad line."""
    tokens = [ln.split() for ln in lines]
    short = next((k for k, t in enumerate(tokens) if len(t) != width), None)
    if short is not None:
        raise ValueError(f"{path}:{short + 2}: expected {width} values, "
                         f"got {len(tokens[short])}")
    try:
        data = np.array(tokens, dtype=np.float64).reshape(len(lines), width)
    except ValueError:
        for k, t in enumerate(tokens):
            try:
                [float(v) for v in t]
            except ValueError as exc:
                raise ValueError(f"{path}:{k + 2}: invalid numeric value") from exc
        raise
    bad = ~(np.isfinite(data) & (np.abs(data) <= _F32.max)).all(axis=1)
    if bad.any():
        raise ValueError(f"{path}:{int(np.argmax(bad)) + 2}: values must be finite float32 "
                         "numbers")
    return data


def _sphere_volume(radius: np.ndarray) -> np.ndarray:
    return (4.0 / 3.0) * np.pi * radius ** 3


# Value checks after parsing: (fields, test that must hold for every entry, message).
_CHECKS = (
    (("rho", "m", "radius", "inertia", "K", "h"), lambda v: v.astype(np.float32) > 0.0,
     "{f} must be positive in float32"),
    (("eta", "mu"), lambda v: v >= 0.0, "{f} must be nonnegative"),
    (("mirror",), lambda v: (v == np.floor(v)) & (v >= 0) & (v <= _I32_MAX),
     "mirror must be a nonnegative int32 fluid row index"),
)


def _check_values(path: Path, fields: dict) -> None:
    for names, ok, message in _CHECKS:
        for f in names:
            if f in fields and not np.all(ok(fields[f])):
                raise ValueError(f"{path}: " + message.format(f=f))
    if "radius" in fields:
        vol = _sphere_volume(fields["radius"])
        if not np.allclose(fields["m"], fields["rho"] * vol, rtol=1e-5, atol=0.0):
            raise ValueError(f"{path}: sphere mass must equal rho*(4/3)*pi*radius^3")
        vol32 = vol.astype(np.float32)
        if not (np.isfinite(vol32).all() and np.all(vol32 > 0)):
            raise ValueError(f"{path}: derived sphere volume is outside float32 range")


def read_fields(path, kind: str) -> dict:
    """Parse one input file into {field: float64 array} (vec3 fields (n, 3))."""
    path = Path(path)
    layout = Layout.of(kind)
    with path.open(encoding="utf-8-sig") as stream:
        first = stream.readline()
        header = _header(path, first, layout)
        data = _body(path, stream.read().splitlines(), len(header))
    where = {cid: k for k, cid in enumerate(header)}
    fields = {}
    for f, ids in layout.ids.items():
        cols = data[:, [where[i] for i in ids]]
        fields[f] = cols[:, 0] if len(ids) == 1 else cols
    _check_values(path, fields)
    return fields


# Fields a file does not carry, per kind: name -> value from (fields, n).
_DERIVED = {
    "SPH": {"rho_raw": lambda f, n: f["rho"], "porosity": lambda f, n: np.ones(n)},
    "BND": {"rho_raw": lambda f, n: f["rho"]},
    "DEM": {"porosity": lambda f, n: np.ones(n),
            "volume": lambda f, n: _sphere_volume(f["radius"])},
    "DEMBND": {"volume": lambda f, n: _sphere_volume(f["radius"])},
    "EISPH": {"vel_star": lambda f, n: f["vel"], "mirror": lambda f, n: np.arange(n)},
    "EISPHBND": {},
}
_INTEGER_FIELDS = ("mirror",)


def _tensor(name: str, values: np.ndarray, n: int) -> torch.Tensor:
    if name in _INTEGER_FIELDS:
        return torch.from_numpy(np.asarray(values).astype(np.int32))
    t = torch.from_numpy(np.ascontiguousarray(values, dtype=np.float32))
    return t.reshape(n, 3) if t.ndim == 2 else t


def input_parser(file_path, kind: str | None = None, device="cpu"):
    """Read input_<kind>.txt into a fully allocated struct (input_reader.py:98-134).

    The kind defaults to the file name's.  Fields the file does not carry
    start at zero except: rho_raw = rho, fluid and DEM porosity = 1, DEM
    volume = (4/3) pi radius^3, EISPH vel_star = vel and the identity mirror.
    Both contact CSR sides are empty with zero offsets; a header-only file
    gives an empty struct."""
    path = Path(file_path)
    kind = path.stem.removeprefix("input_") if kind is None else kind
    if kind not in STRUCTURES:
        raise ValueError(f"unknown particle kind {kind!r}; choose {tuple(STRUCTURES)}")
    fields = read_fields(path, kind)
    n = int(fields["pos"].shape[0])
    fields.update({name: make(fields, n) for name, make in _DERIVED[kind].items()})
    particle = STRUCTURES[kind].allocate(n, device="cpu")
    for name, values in fields.items():
        setattr(particle, name, _tensor(name, values, n))
    return particle if str(device) == "cpu" else particle.to(device)


def input_directory(solv) -> Path:
    path = Path(solv.input_dir)
    return path if path.is_absolute() else PROJECT_ROOT / path


def _host(t) -> np.ndarray:
    return t.detach().cpu().numpy()


def dem_search_scales(dem, boundary) -> tuple:
    """(max DEM radius, max DEM-wall radius or 0, max h, common h or 0)
    (input_reader.py:163-168)."""
    h = _host(dem.h)
    wall_r = float(_host(boundary.radius).max()) if boundary.pos.shape[0] else 0.0
    common = float(h[0]) if (h == h[0]).all() else 0.0
    return float(_host(dem.radius).max()), wall_r, float(h.max()), common


def dem_contact_dt_limit(solv, dem) -> float:
    """solv.dem_dt_safety x min(sqrt(m*/K_max), m*/eta_max) with the
    conservative reduced mass m* = min(m)/2 (input_reader.py:142-160)."""
    m_star = 0.5 * float(_host(dem.m).astype(np.float64).min())
    k_max = float(_host(dem.K).astype(np.float64).max())
    eta_max = max(float(_host(dem.eta).astype(np.float64).max()), 1e-20)
    return solv.dem_dt_safety * min(np.sqrt(m_star / k_max), m_star / eta_max)


def validate_dem_particles(solv, dem, boundary) -> tuple:
    """Reject a time step above the loaded spheres' contact limit; returns
    dem_search_scales (input_reader.py:142-160)."""
    solv.validate_dem()
    if dem.pos.shape[0] == 0:
        raise ValueError("DEM enabled but input_DEM.txt contains no particles")
    limit = dem_contact_dt_limit(solv, dem)
    if solv.dt > limit:
        raise ValueError(f"dt exceeds loaded DEM contact limit {limit:.6g} s")
    return dem_search_scales(dem, boundary)


def _load(solv, kinds):
    folder = input_directory(solv)
    return [input_parser(folder / f"input_{k}.txt", k, solv.device) for k in kinds]


def load_sph_particles(solv):
    """The WCSPH files: fluid and wall, plus DEM and DEM wall when
    dem_enable (input_reader.py:171-184)."""
    solv.validate_sph()
    sph, bnd = _load(solv, ("SPH", "BND"))
    if sph.pos.shape[0] == 0:
        raise ValueError("input_SPH.txt must contain fluid particles")
    if not solv.dem_enable:
        return sph, bnd, None, None
    dem, dem_bnd = _load(solv, ("DEM", "DEMBND"))
    validate_dem_particles(solv, dem, dem_bnd)
    return sph, bnd, dem, dem_bnd


def validate_eisph_particles(solv, fluid, boundary) -> None:
    """A fixed planar cavity cloud and its mirror map (input_reader.py:187-208):
    nonempty sets, mirror rows inside the fluid file, uniform rho = rho0,
    points in the x-z plane, and the explicit predictor's advection and
    viscosity dt limits for the loaded speeds and spacing."""
    solv.validate_eisph()
    n_fluid = fluid.pos.shape[0]
    if n_fluid == 0 or boundary.pos.shape[0] == 0:
        raise ValueError("EISPH needs nonempty fluid and ghost files")
    if (_host(boundary.mirror) >= n_fluid).any():
        raise ValueError("EISPHBND mirror is outside the input_EISPH fluid row range")
    for cloud in (fluid, boundary):
        if not np.allclose(_host(cloud.rho), solv.rho0, rtol=1e-6, atol=0):
            raise ValueError("EISPH currently requires uniform rho equal to rho0")
        if not np.allclose(_host(cloud.pos)[:, 1], 0.0, atol=1e-8):
            raise ValueError("EISPH points must lie in the x-z plane (y=0)")
    # planar mass m = rho dx^2: the smallest equivalent spacing
    spacing = float(np.sqrt(_host(fluid.m) / _host(fluid.rho)).min())
    speeds = [np.linalg.norm(_host(v), axis=1).max() for v in (fluid.vel, boundary.vel_bc)]
    speed = float(max(speeds))
    if speed > 0 and solv.dt > 0.25 * spacing / speed:
        raise ValueError("dt exceeds the loaded EISPH initial advection limit")
    if solv.nu > 0 and solv.dt > 0.125 * spacing ** 2 / solv.nu:
        raise ValueError("dt exceeds the loaded EISPH viscosity limit")


def load_eisph_particles(solv):
    """The cavity files, validated, with each ghost's initial velocity
    2 vel_bc - vel[mirror] in vel and vel_star (input_reader.py:211-220)."""
    fluid, ghosts = _load(solv, ("EISPH", "EISPHBND"))
    validate_eisph_particles(solv, fluid, ghosts)
    ghosts.vel = 2.0 * ghosts.vel_bc - fluid.vel[ghosts.mirror.long()]
    ghosts.vel_star = ghosts.vel.clone()
    return fluid, ghosts
