"""GIF rendering of a run (host only, matplotlib): the reference's
output/gif_gen.py:19-204 interface -- collect_frame, sphere_faces, save_gif
with field pres/vel/rho, one colour scale for all frames (clipped at a
percentile), fixed axis limits (given or covering every particle), fluid as
translucent points, a sampled dummy wall, DEM spheres as shaded surfaces.
"""
from __future__ import annotations

import os

import numpy as np

from .output import SPH_TYPE, _host, gather_state

ANI_DIR = "animation/3d"


def collect_frame(P_sph, P_bnd, t: float, P_dem=None, P_dem_bnd=None) -> dict:
    """Host copy of one state for later rendering: positions [N, 3], fluid
    |v|, pressure and density, DEM centres and radii."""
    d = gather_state(P_sph, P_bnd)
    fluid = d["type"] == SPH_TYPE
    frame = {"t": t, "pos_sph": d["pos"][fluid], "pos_bnd": d["pos"][~fluid],
             "vel_sph": np.linalg.norm(d["vel"][fluid], axis=1),
             "pres_sph": d["pres"][fluid], "rho_sph": d["rho"][fluid]}
    if P_dem is not None:
        frame["pos_dem"], frame["radius_dem"] = _host(P_dem.pos), _host(P_dem.radius)
    if P_dem_bnd is not None:
        frame["pos_dem_bnd"] = _host(P_dem_bnd.pos)
        frame["radius_dem_bnd"] = _host(P_dem_bnd.radius)
    return frame


def sphere_faces(n_lat: int = 10, n_lon: int = 16) -> np.ndarray:
    """Quadrilaterals [n_lat * n_lon, 4, 3] of the unit sphere (z polar)."""
    th = np.linspace(0.0, np.pi, n_lat + 1)
    ph = np.linspace(0.0, 2.0 * np.pi, n_lon + 1)
    T, Pp = np.meshgrid(th, ph, indexing="ij")
    v = np.stack([np.sin(T) * np.cos(Pp), np.sin(T) * np.sin(Pp), np.cos(T)], axis=-1)
    quads = np.stack([v[:-1, :-1], v[1:, :-1], v[1:, 1:], v[:-1, 1:]], axis=2)
    return quads.reshape(-1, 4, 3)


def _limits(frames, pad):
    lo, hi = np.full(3, np.inf), np.full(3, -np.inf)
    for fr in frames:
        for key in ("pos_sph", "pos_bnd", "pos_dem", "pos_dem_bnd"):
            pts = fr.get(key)
            if pts is None or not len(pts):
                continue
            r = fr.get(key.replace("pos_", "radius_"))
            r = 0.0 if r is None else r[:, None]
            lo = np.minimum(lo, (pts - r).min(0) - pad)
            hi = np.maximum(hi, (pts + r).max(0) + pad)
    return lo, hi


def save_gif(frames: list, solv, field: str = "pres", fps: int = 20, clip: float = 99.0,
             ani_dir: str = ANI_DIR, name: str = "dam_break_3d", axis_limits=None) -> str:
    """Render `frames` (collect_frame) to `{ani_dir}/{name}.gif`; returns the path."""
    if not frames:
        raise ValueError("no frames to write")
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    from matplotlib.animation import FuncAnimation, PillowWriter
    from matplotlib.lines import Line2D
    from mpl_toolkits.mplot3d.art3d import Poly3DCollection

    key = {"vel": "vel_sph", "pres": "pres_sph", "rho": "rho_sph"}[field]
    label = {"vel": "|v| [m/s]", "pres": "pressure [Pa]", "rho": "density [kg/m^3]"}[field]
    os.makedirs(ani_dir, exist_ok=True)
    path = os.path.join(ani_dir, f"{name}.gif")
    allv = np.concatenate([f[key] for f in frames])
    vmin = float(allv.min())
    vmax = max(float(np.percentile(allv, clip)), vmin + 1.0)
    if axis_limits is None:
        lo, hi = _limits(frames, 0.5 * solv.h)
    else:
        lo, hi = (np.asarray(b, dtype=float) for b in axis_limits)
        if (lo.shape != (3,) or hi.shape != (3,) or not np.isfinite(lo).all()
                or not np.isfinite(hi).all() or np.any(hi <= lo)):
            raise ValueError("axis_limits must contain finite increasing xyz bounds")
    fig = plt.figure(figsize=(10, 6.4), dpi=100)
    ax = fig.add_subplot(111, projection="3d", computed_zorder=False)
    ax.set_xlim(lo[0], hi[0])
    ax.set_ylim(lo[1], hi[1])
    ax.set_zlim(lo[2], hi[2])
    ax.set_box_aspect(hi - lo)
    ax.view_init(elev=24, azim=-65)
    ax.set_proj_type("ortho")
    ax.set_xlabel("x [m] (length)")
    ax.set_ylabel("y [m] (depth)")
    ax.set_zlabel("z [m] (height)")
    ax.grid(False)
    for a in (ax.xaxis, ax.yaxis, ax.zaxis):
        a.pane.fill = False
    wall = frames[0]["pos_bnd"]
    if len(wall):
        wall = wall[::max(1, len(wall) // 4000)]
        ax.scatter(wall[:, 0], wall[:, 1], wall[:, 2], color="0.55", s=1, alpha=0.12,
                   linewidths=0, zorder=1)
    p0 = frames[0]["pos_sph"]
    sc = ax.scatter(p0[:, 0], p0[:, 1], p0[:, 2], c=frames[0][key], s=5, cmap="viridis",
                    vmin=vmin, vmax=vmax, alpha=0.30, depthshade=False, linewidths=0, zorder=2)
    dem, faces = None, None
    if "pos_dem" in frames[0]:
        faces = sphere_faces()
        nrm = faces.mean(axis=1)
        nrm /= np.linalg.norm(nrm, axis=1)[:, None]
        light = np.array([-0.5, -0.8, 1.0]) / np.linalg.norm([-0.5, -0.8, 1.0])
        shade = 0.45 + 0.55 * np.maximum(nrm @ light, 0.0)
        colours = shade[:, None] * np.array([1.0, 0.55, 0.05])

        def verts(fr):
            v = fr["pos_dem"][:, None, None, :] + fr["radius_dem"][:, None, None, None] * faces
            return v.reshape(-1, 4, 3)

        dem = Poly3DCollection(verts(frames[0]),
                               facecolors=np.tile(colours, (len(frames[0]["pos_dem"]), 1)),
                               edgecolors="none", zsort="average", zorder=3)
        ax.add_collection3d(dem)
    handles = [Line2D([], [], marker="o", linestyle="none", color="teal", label="SPH fluid")]
    if dem is not None:
        handles.append(Line2D([], [], marker="o", linestyle="none", color="darkorange",
                              label="DEM spheres"))
    ax.legend(handles=handles, loc="upper left", framealpha=0.85)
    fig.colorbar(sc, ax=ax, label=label, fraction=0.025, pad=0.06, shrink=0.68,
                 extend="max" if allv.max() > vmax else "neither")
    title = ax.set_title(f"3D Dam Break | t = {frames[0]['t']:.4f} s")
    fig.subplots_adjust(left=0.02, right=0.88, bottom=0.06, top=0.92)

    def draw(i):
        fr = frames[i]
        p = fr["pos_sph"]
        sc._offsets3d = (p[:, 0], p[:, 1], p[:, 2])
        sc.set_array(fr[key])
        if dem is not None:
            dem.set_verts(verts(fr))
        title.set_text(f"3D Dam Break | t = {fr['t']:.4f} s")
        return sc, title

    FuncAnimation(fig, draw, frames=len(frames), blit=False).save(path,
                                                                   writer=PillowWriter(fps=fps))
    plt.close(fig)
    return path


def save_cavity_gif(positions, velocity_frames: list, times: list, solv, path=None) -> str:
    """GIF of the Eulerian cavity (output/gif_eisph.py:11-61): fixed x-z
    points coloured by speed |(u, w)| [m/s] on one scale for all frames,
    velocity arrows on a sample of about 150 points, bounds from the point
    cloud.  Writes `path` (default solv.animation_path); returns its
    absolute path."""
    if len(velocity_frames) != len(times) or not velocity_frames:
        raise ValueError("velocity_frames and times must have the same non-zero length")
    positions = np.asarray(positions)
    if positions.ndim != 2 or positions.shape[1] != 3 or not len(positions):
        raise ValueError("positions must be a nonempty [N,3] array")
    if not np.isfinite(positions).all():
        raise ValueError("positions must be finite")
    frames = [np.asarray(v) for v in velocity_frames]
    for v in frames:
        if v.shape != positions.shape or not np.isfinite(v).all():
            raise ValueError("every velocity frame must be finite and match positions")
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    from PIL import Image

    dest = os.path.abspath(solv.animation_path if path is None else str(path))
    os.makedirs(os.path.dirname(dest), exist_ok=True)
    speed = [np.linalg.norm(v[:, (0, 2)], axis=1) for v in frames]
    vmax = max(max(float(s.max()) for s in speed), 1e-8)
    x, z = positions[:, 0], positions[:, 2]
    span = max(float(np.ptp(x)), float(np.ptp(z)), solv.h)
    pad = 0.04 * span
    pick = slice(None, None, max(1, len(positions) // 150))
    fig, ax = plt.subplots(figsize=(6.1, 5.5), dpi=100)
    dots = ax.scatter(x, z, c=speed[0], cmap="turbo", vmin=0.0, vmax=vmax, marker="s",
                      s=max(3.0, min(70.0, 25000.0 / len(positions))))
    arrows = ax.quiver(x[pick], z[pick], frames[0][pick, 0], frames[0][pick, 2], color="white",
                       angles="xy", scale_units="xy", scale=15.0 * vmax / span, width=0.003)
    ax.set(xlabel="x [m]", ylabel="z [m]", aspect="equal",
           xlim=(x.min() - pad, x.max() + pad), ylim=(z.min() - pad, z.max() + pad))
    title = ax.set_title(f"Eulerian ISPH   t = {times[0]:.3f} s")
    fig.colorbar(dots, ax=ax, pad=0.03).set_label("speed [m/s]")
    fig.tight_layout()
    images = []
    for s, v, t in zip(speed, frames, times):
        dots.set_array(s)
        arrows.set_UVC(v[pick, 0], v[pick, 2])
        title.set_text(f"Eulerian ISPH   t = {t:.3f} s")
        fig.canvas.draw()
        rgba = np.asarray(fig.canvas.buffer_rgba())
        images.append(Image.fromarray(rgba[..., :3].copy()).convert("P", palette=Image.ADAPTIVE))
    plt.close(fig)
    images[0].save(dest, save_all=True, append_images=images[1:], loop=0,
                   duration=int(round(1000.0 / solv.gif_fps)))
    return dest
