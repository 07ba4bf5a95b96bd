"""Command line of the WCSPH-DEM dam break: the reference's main.py:1-130.

    python -m warp_sph_gpt_amd.main [--device D] [--input-dir DIR] [--steps N]
        [--dt DT] [--output-step K] [--no-gif] [--no-dem] [--output-dir DIR]
        [--animation-dir DIR] [--engine]

Same flags, defaults (Solv), input files and outputs as the reference:
omitted flags keep the Solv defaults; --no-dem runs the SPH-only step;
frames are written at step 0, every output_step steps and at the last step.
--engine (not in the reference) runs the same kernels through
engine.CoupledEngine, which keeps the contact CSR on the device instead of
reading its size back every step; frames are identical in content.
"""
from __future__ import annotations

import argparse
from typing import Any

from .config import Solv
from .engine import CoupledEngine
from .hashgrid import HashGrid
from .input_reader import load_sph_particles, validate_dem_particles
from .output import FrameWriter
from .simulation import SPH_OneStep, SPHDEM_OneStep


def parsing(argv=None) -> dict[str, Any]:
    """Command-line overrides for Solv (main.py:13-37); omitted options are
    left out so that Solv keeps its defaults."""
    p = argparse.ArgumentParser(description="MI355X 3D SPH-DEM Dam Break Solver")
    S = argparse.SUPPRESS
    p.add_argument("--device", type=str, help="device to use", default=S)
    p.add_argument("--input-dir", type=str, default=S)
    p.add_argument("--steps", dest="n_steps", type=int, default=S)
    p.add_argument("--dt", type=float, default=S)
    p.add_argument("--output-step", type=int, default=S)
    p.add_argument("--no-gif", dest="gif_save", action="store_false", default=S)
    p.add_argument("--no-dem", dest="dem_enable", action="store_false", default=S)
    p.add_argument("--output-dir", type=str, default=S)
    p.add_argument("--animation-dir", type=str, default=S)
    p.add_argument("--engine", action="store_true", default=S,
                   help="device-resident contact CSR, no per-step host round trip")
    return vars(p.parse_args(argv))


def run_forward(solv, P_sph, P_bnd, grid_sph, grid_bnd, P_dem=None, P_dem_bnd=None,
                grid_dem=None, grid_dem_bnd=None, engine: bool = False) -> None:
    """The step loop of main.py:40-98: fixed grids once, moving grids before
    every step, frames at step 0, every output_step and the last step."""
    print(f"\n[forward] {solv.n_steps} step \t\t (t={solv.n_steps * solv.dt:.3f} s\t"
          f"dt={solv.dt:.1e} s)\n")
    grid_bnd.build(P_bnd.pos, solv.support)
    coupled = solv.dem_enable and P_dem is not None
    if solv.dem_enable and not coupled:
        raise ValueError("DEM is enabled: provide DEM particles, boundaries and coupling grids")
    scales = None
    if coupled:
        scales = validate_dem_particles(solv, P_dem, P_dem_bnd)
        grid_dem_bnd.build(P_dem_bnd.pos, max(2.0 * scales[1], solv.support))
    writer = FrameWriter(solv, P_sph, P_bnd, P_dem if coupled else None,
                         P_dem_bnd if coupled else None)
    if solv.output_step > 0:
        writer.write_frame(0)
    eng = None
    if engine:
        eng = CoupledEngine(solv, P_sph, P_bnd, P_dem if coupled else None,
                            P_dem_bnd if coupled else None)
    for step in range(solv.n_steps):
        if eng is not None:
            eng.step()
        else:
            grid_sph.build(P_sph.pos, solv.support)
            if coupled:
                grid_dem.build(P_dem.pos, max(2.0 * scales[0], 2.0 * scales[2]))
                SPHDEM_OneStep(solv, P_sph, P_bnd, P_dem, P_dem_bnd, grid_sph, grid_bnd,
                               grid_dem, grid_dem_bnd, step, scales)
            else:
                SPH_OneStep(solv, P_sph, P_bnd, grid_sph, grid_bnd, step)
        if solv.output_step > 0 and ((step + 1) % solv.output_step == 0
                                     or step + 1 == solv.n_steps):
            if eng is not None:
                eng.check()
            writer.write_frame(step + 1)
            print(f"[output] step {step + 1:>6d} / {solv.n_steps}\t t={(step + 1) * solv.dt:.4f} s")
    if eng is not None:
        eng.check()
    writer.finalize()


def main(argv=None) -> None:
    """Solv from the command line, particles from input_dir, grids of
    grid_slice^3 buckets on solv.device, then run_forward (main.py:101-126)."""
    args = parsing(argv)
    engine = bool(args.pop("engine", False))
    solv = Solv(**args)
    P_sph, P_bnd, P_dem, P_dem_bnd = load_sph_particles(solv)
    g = solv.grid_slice
    grid_sph = HashGrid(g, g, g, device=solv.device)
    grid_bnd = HashGrid(g, g, g, device=solv.device)
    if solv.dem_enable:
        grid_dem = HashGrid(g, g, g, device=solv.device)
        grid_dem_bnd = HashGrid(g, g, g, device=solv.device)
        run_forward(solv, P_sph, P_bnd, grid_sph, grid_bnd, P_dem, P_dem_bnd, grid_dem,
                    grid_dem_bnd, engine=engine)
    else:
        run_forward(solv, P_sph, P_bnd, grid_sph, grid_bnd, engine=engine)


if __name__ == "__main__":
    main()
