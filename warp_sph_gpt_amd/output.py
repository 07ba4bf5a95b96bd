"""Frame output: VTK XML PolyData (.vtp) frames and ParaView collections (.pvd).

Keeps the reference's on-disk format (output/output.py:1-207): one vertex
cell per particle, ASCII data arrays with "%.6g" floats six per line and
"%d" integers twenty per line, fluid and dummy wall in one "sph" dataset
(type 1 / 0), moving DEM and fixed DEM wall in one "dem" dataset (type 2 /
3) with zero placeholders for fields the fixed wall does not have, and a
.pvd collection carrying each frame's physical time.  `FrameWriter`
(output/frame_writer.py:9-87) checks every frame for non-finite values
before writing any of it and raises FloatingPointError.

The numbers are formatted by libsphdem's host formatter (C "%.6g" / "%d",
the same text as Python's for finite values), which is what keeps frames of
millions of particles affordable.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib

OUT_DIR = "result/3d"

SPH_TYPE = 1        # fluid particle
BND_TYPE = 0        # dummy boundary particle
DEM_TYPE = 2        # moving DEM sphere
DEM_BND_TYPE = 3    # fixed DEM wall sphere


def _host(t) -> np.ndarray:
    return t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)


def _format(values: np.ndarray, per_line: int, integer: bool) -> str:
    flat = np.ascontiguousarray(np.asarray(values).ravel(),
                                dtype=np.int32 if integer else np.float64)
    n = flat.size
    cap = n * (12 if integer else 16) + 16
    buf = ctypes.create_string_buffer(cap)
    fn = "sphdem_format_int" if integer else "sphdem_format_g6"
    got = getattr(_lib.load(), fn)(flat.ctypes.data_as(ctypes.c_void_p), n, per_line, buf, cap)
    if got < 0:
        raise RuntimeError(f"{fn}: output buffer too small")
    return buf.raw[:got].decode("ascii")


def fmt_float(a, per_line: int = 6) -> str:
    """ASCII floats, "%.6g", per_line on a line (output/output.py:14-17)."""
    return _format(a, per_line, False)


def fmt_int(a, per_line: int = 20) -> str:
    """ASCII integers, per_line on a line (output/output.py:20-23)."""
    return _format(a, per_line, True)


def gather_state(P_sph, P_bnd) -> dict:
    """Fluid and dummy wall stacked into one host point set (fluid first),
    with the fluid-only coupling fields padded for the wall
    (output/output.py:27-60)."""
    n_sph, n_bnd = P_sph.n, P_bnd.n
    return {
        "pos": np.vstack([_host(P_sph.pos), _host(P_bnd.pos)]),
        "vel": np.vstack([_host(P_sph.vel), _host(P_bnd.vel)]),
        "rho": np.concatenate([_host(P_sph.rho), _host(P_bnd.rho)]),
        "pres": np.concatenate([_host(P_sph.pres), _host(P_bnd.pres)]),
        "m": np.concatenate([_host(P_sph.m), _host(P_bnd.m)]),
        "porosity": np.concatenate([_host(P_sph.porosity), np.ones(n_bnd)]),
        "pgf": np.vstack([_host(P_sph.pgf), np.zeros((n_bnd, 3))]),
        "acc_dem": np.vstack([_host(P_sph.acc_dem), np.zeros((n_bnd, 3))]),
        "type": np.concatenate([np.full(n_sph, SPH_TYPE, np.int32),
                                np.full(n_bnd, BND_TYPE, np.int32)]),
    }


def gather_dem_state(P_dem, P_dem_bnd) -> dict:
    """Moving DEM and fixed DEM wall in one point set (output/output.py:88-127)."""
    n_bnd = P_dem_bnd.n
    d = {}
    for key in ("pos", "vel", "omega", "radius", "rho", "m", "volume", "inertia"):
        d[key] = np.concatenate([_host(getattr(P_dem, key)), _host(getattr(P_dem_bnd, key))])
    for key in ("acc", "force", "torque"):
        d[key] = np.vstack([_host(getattr(P_dem, key)), np.zeros((n_bnd, 3))])
    d["type"] = np.concatenate([np.full(P_dem.n, DEM_TYPE, np.int32),
                                np.full(n_bnd, DEM_BND_TYPE, np.int32)])
    for key in ("drag", "pressure_force"):
        d[key] = np.vstack([_host(getattr(P_dem, key)), np.zeros((n_bnd, 3))])
    d["porosity"] = np.concatenate([_host(P_dem.porosity), np.ones(n_bnd)])
    return d


def write_vtp(d: dict, step: int, t: float, out_dir: str, name: str, scalar: str):
    """One point set as VTK XML PolyData `{name}_{step:06d}.vtp` in out_dir;
    returns (file name, t) for the collection (output/output.py:130-180)."""
    os.makedirs(out_dir, exist_ok=True)
    n = d["pos"].shape[0]
    fname = f"{name}_{step:06d}.vtp"
    parts = ['<?xml version="1.0"?>\n',
             '<VTKFile type="PolyData" version="0.1" byte_order="LittleEndian">\n',
             '  <PolyData>\n',
             f'    <Piece NumberOfPoints="{n}" NumberOfVerts="{n}" '
             'NumberOfLines="0" NumberOfStrips="0" NumberOfPolys="0">\n',
             '      <Points>\n',
             '        <DataArray type="Float32" Name="Points" NumberOfComponents="3" '
             'format="ascii">\n',
             fmt_float(d["pos"]), "\n",
             '        </DataArray>\n',
             '      </Points>\n',
             '      <Verts>\n',
             '        <DataArray type="Int32" Name="connectivity" format="ascii">\n',
             fmt_int(np.arange(n)), "\n",
             '        </DataArray>\n',
             '        <DataArray type="Int32" Name="offsets" format="ascii">\n',
             fmt_int(np.arange(1, n + 1)), "\n",
             '        </DataArray>\n',
             '      </Verts>\n',
             f'      <PointData Scalars="{scalar}" Vectors="vel">\n']
    for key, values in d.items():
        if key == "pos":
            continue
        integer = key == "type"
        comps = ' NumberOfComponents="3"' if np.ndim(values) == 2 else ""
        parts += [f'        <DataArray type="{"Int32" if integer else "Float32"}" '
                  f'Name="{key}"{comps} format="ascii">\n',
                  fmt_int(values) if integer else fmt_float(values), "\n",
                  '        </DataArray>\n']
    parts += ['      </PointData>\n', '    </Piece>\n', '  </PolyData>\n', '</VTKFile>\n']
    with open(os.path.join(out_dir, fname), "w") as f:
        f.write("".join(parts))
    return fname, t


def save_vtk(P_sph, P_bnd, step: int, t: float, out_dir: str = OUT_DIR, name: str = "ptl"):
    """Fluid + dummy wall frame, default scalar "pres" (output/output.py:63-85)."""
    return write_vtp(gather_state(P_sph, P_bnd), step, t, out_dir, name, "pres")


def save_dem_vtk(P_dem, P_dem_bnd, step: int, t: float, out_dir: str = OUT_DIR,
                 name: str = "dem"):
    """Moving DEM + fixed DEM wall frame, default scalar "radius"."""
    return write_vtp(gather_dem_state(P_dem, P_dem_bnd), step, t, out_dir, name, "radius")


def save_pvd(frames: list, out_dir: str = OUT_DIR, name: str = "ptl") -> str:
    """The .pvd collection of (file, t) frames (output/output.py:183-207)."""
    os.makedirs(out_dir, exist_ok=True)
    path = os.path.join(out_dir, f"{name}.pvd")
    lines = ['<?xml version="1.0"?>\n',
             '<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">\n',
             '  <Collection>\n']
    lines += [f'    <DataSet timestep="{t:.9g}" group="" part="0" file="{fname}"/>\n'
              for fname, t in frames]
    lines += ['  </Collection>\n', '</VTKFile>\n']
    with open(path, "w") as f:
        f.write("".join(lines))
    return path


class FrameWriter:
    """Synchronised SPH / DEM output of a run (output/frame_writer.py:9-87)."""

    def __init__(self, solv, P_sph, P_bnd, P_dem=None, P_dem_bnd=None):
        if (P_dem is None) != (P_dem_bnd is None):
            raise ValueError("provide both moving and fixed DEM particles")
        self.solv = solv
        self.P_sph, self.P_bnd = P_sph, P_bnd
        self.P_dem, self.P_dem_bnd = P_dem, P_dem_bnd
        self.coupled = P_dem is not None
        self.frames: list = []
        self.dem_frames: list = []
        self.gif_frames: list = []

    def write_frame(self, step: int) -> None:
        """Write the state after `step` steps (0 = initial); raise
        FloatingPointError, before writing anything, if pos/vel/acc/rho of
        any particle set is not finite."""
        t = step * self.solv.dt
        checks = [(self.P_sph, ("pos", "vel", "acc", "rho")),
                  (self.P_bnd, ("pos", "vel", "acc", "rho"))]
        if self.coupled:
            checks += [(self.P_dem, ("pos", "vel", "acc", "rho")),
                       (self.P_dem_bnd, ("pos", "vel", "rho"))]
        for S, fields in checks:
            for key in fields:
                if not np.isfinite(_host(getattr(S, key))).all():
                    raise FloatingPointError(f"non-finite {key} at step {step}")
        self.frames.append(save_vtk(self.P_sph, self.P_bnd, step, t,
                                    out_dir=self.solv.output_dir, name="sph"))
        if self.coupled:
            self.dem_frames.append(save_dem_vtk(self.P_dem, self.P_dem_bnd, step, t,
                                                out_dir=self.solv.output_dir))
        if self.solv.gif_save:
            from .gif import collect_frame
            self.gif_frames.append(collect_frame(self.P_sph, self.P_bnd, t, self.P_dem,
                                                 self.P_dem_bnd))

    def finalize(self) -> None:
        """Write the GIF (if frames were collected) and the .pvd collections."""
        if self.gif_frames:
            from .gif import save_gif
            name = "dam_break_sph_dem_3d" if self.coupled else "dam_break_3d"
            gif = save_gif(self.gif_frames, self.solv, ani_dir=self.solv.animation_dir, name=name)
            print(f"\n[output] {len(self.gif_frames)} frames -> {gif}")
        if self.frames:
            path = save_pvd(self.frames, out_dir=self.solv.output_dir, name="sph")
            print(f"\n[output] {len(self.frames)} frames -> {path}\n")
        if self.dem_frames:
            path = save_pvd(self.dem_frames, out_dir=self.solv.output_dir, name="dem")
            print(f"[output] {len(self.dem_frames)} frames -> {path}\n")
