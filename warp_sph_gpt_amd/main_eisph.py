"""Command line of the Eulerian ISPH lid-driven cavity: the reference's
main_EISPH.py:1-65.

    python -m warp_sph_gpt_amd.main_eisph [--device D] [--steps N] [--dt DT]
        [--input-dir DIR] [--h H] [--nu NU] [--no-gif] [--output-step K]
        [--fps F] [--animation PATH]

Same flags and defaults (EISPHSolv): omitted flags keep the cavity
defaults; input_EISPH.txt and input_EISPHBND.txt are read from input_dir;
the GIF of the stored velocity frames goes to animation_path.
"""
from __future__ import annotations

import argparse
from typing import Any

from .config import EISPHSolv
from .eisph import run_eisph
from .input_reader import load_eisph_particles


def parsing(argv=None) -> dict[str, Any]:
    """Command-line overrides for EISPHSolv (main_EISPH.py:11-35)."""
    p = argparse.ArgumentParser(description="Eulerian ISPH lid-driven cavity")
    S = argparse.SUPPRESS
    p.add_argument("--device", type=str, help="device to use", default=S)
    p.add_argument("--steps", dest="n_steps", type=int, default=S)
    p.add_argument("--dt", type=float, default=S)
    p.add_argument("--input-dir", type=str, default=S)
    p.add_argument("--h", type=float, default=S)
    p.add_argument("--nu", type=float, default=S)
    p.add_argument("--no-gif", dest="gif_save", action="store_false", default=S)
    p.add_argument("--output-step", type=int, default=S)
    p.add_argument("--fps", dest="gif_fps", type=int, default=S)
    p.add_argument("--animation", dest="animation_path", type=str, default=S)
    return vars(p.parse_args(argv))


def main(argv=None) -> None:
    """Load the cavity, run it on solv.device and save its GIF
    (main_EISPH.py:38-61)."""
    solv = EISPHSolv(**parsing(argv))
    P_sph, P_bnd = load_eisph_particles(solv)
    positions, velocities, times = run_eisph(
        solv, P_sph, P_bnd,
        progress=lambda step, total: print(f"[EISPH] {step:>5d} / {total}"))
    if solv.gif_save:
        from .gif import save_cavity_gif
        print(f"[EISPH] animation: {save_cavity_gif(positions, velocities, times, solv)}")


if __name__ == "__main__":
    main()
