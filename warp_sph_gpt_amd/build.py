"""Build recipe for libsphdem.so (hand-written HIP for gfx950).

Run as ``python -m warp_sph_gpt_amd.build``.  The library is built in-tree
next to this file so that it travels with the repository to the GPU host.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libsphdem.so"
SOURCES = ("grid.hip", "sph.hip", "dem.hip", "porosity.hip", "mpnn.hip", "eisph.hip",
           "format.hip")

# -ffp-contract=off: every a*b+c rounds twice, as the reference's FP32
# expressions are written.  Packed FP32 instructions are disabled.  Matrix
# instructions take and write ordinary VGPRs (no accumulation-register
# copies around each product; the porosity tiles read their results at once).
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found; ROCm is required to build libsphdem.so")
    return path


def _run(cmd, verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}: {cmd[-1]}")
    if verbose and res.stderr.strip():
        sys.stderr.write(res.stderr)


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every HIP source into warp_sph_gpt_amd/libsphdem.so: one
    object per source, compiled in parallel, then one link."""
    from concurrent.futures import ThreadPoolExecutor

    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    objs = [objdir / (Path(s).stem + ".o") for s in SOURCES]
    cmds = [[hipcc(), *compile_flags, "-c", "-o", str(o), str(CSRC / s)]
            for s, o in zip(SOURCES, objs)]
    workers = jobs or min(len(cmds), max(1, (os.cpu_count() or 2) // 2), 8)
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp),
          *[str(o) for o in objs]], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
