"""FP64 shadow run: the same HIP kernels in double precision beside the FP32 path.

Every kernel of libsphdem is a template over the real type, and a struct's
`precision` (4 or 8) selects the instantiation.  `ShadowRun` advances an FP32
state and an FP64 copy of the same initial state step by step through the
same `CoupledEngine` loop (same kernels, same order, same grids), and reports
for each field the relative error of the FP32 result against the FP64 one:

    err(field) = max_i |a_i - b_i| / max_i |b_i|      (rows of vec3 fields as vectors)

and the absolute error max_i |a_i - b_i| beside it (a field that is ~0 in
FP64, e.g. the pressure of a lattice at rest, has a meaningless ratio),
together with the FP64 totals (mass, kinetic energy, surge front) of both runs.

    python -m warp_sph_gpt_amd.shadow --workload c0 --steps 100
"""
from __future__ import annotations

import argparse
import json
import sys

import torch

from .engine import CoupledEngine

SPH_FIELDS = ("pos", "vel", "rho", "pres", "acc", "flt", "porosity", "pgf", "acc_dem")
BND_FIELDS = ("rho", "pres")
DEM_FIELDS = ("pos", "vel", "omega", "force", "torque", "flt_s", "porosity", "drag",
              "pressure_force")


def field_error(a: torch.Tensor, b: torch.Tensor, absolute: bool = False) -> float:
    """max-norm relative (or absolute) error of a against b, vec3 rows as vectors, in FP64."""
    a = a.to(torch.float64)
    b = b.to(torch.float64)
    if a.shape != b.shape:
        raise ValueError(f"shape {tuple(a.shape)} != {tuple(b.shape)}")
    if a.numel() == 0:
        return 0.0
    if a.dim() == 2:
        d = torch.linalg.vector_norm(a - b, dim=1).max()
        s = torch.linalg.vector_norm(b, dim=1).max()
    else:
        d = (a - b).abs().max()
        s = b.abs().max()
    return float(d) if absolute else float(d / s.clamp_min(1e-300))


class ShadowRun:
    """An FP32 engine and its FP64 shadow over one initial state."""

    def __init__(self, solv, P, B, D=None, W=None, fast_math: bool | None = None,
                 csr_per_dem: int = 64):
        structs = [s for s in (P, B, D, W)]
        f32 = [None if s is None else s.to(dtype=torch.float32) for s in structs]
        f64 = [None if s is None else s.to(dtype=torch.float64) for s in structs]
        self.e32 = CoupledEngine(solv, *f32, csr_per_dem=csr_per_dem, fast_math=fast_math)
        self.e64 = CoupledEngine(solv, *f64, csr_per_dem=csr_per_dem, fast_math=fast_math)

    @property
    def steps(self) -> int:
        return self.e32.step_index

    def run(self, n_steps: int) -> None:
        for _ in range(int(n_steps)):
            self.e32.step()
            self.e64.step()
        self.e32.check()
        self.e64.check()

    def report(self) -> dict:
        """Per-field relative error of FP32 against FP64, and both runs' totals."""
        a, b = self.e32, self.e64
        out = {"steps": self.steps}
        groups = [("fluid", a.P, b.P, SPH_FIELDS), ("wall", a.B, b.B, BND_FIELDS)]
        if a.coupled:
            groups.append(("dem", a.D, b.D, DEM_FIELDS))
        for name, sa, sb, fields in groups:
            out[name] = {f: field_error(getattr(sa, f), getattr(sb, f)) for f in fields}
            out[name + "_abs"] = {f: field_error(getattr(sa, f), getattr(sb, f), True)
                                  for f in fields}
        d32, d64 = a.diagnostics(), b.diagnostics()
        out["totals_fp32"], out["totals_fp64"] = d32, d64
        out["totals_rel"] = {k: abs(d32[k] - d64[k]) / max(abs(d64[k]), 1e-300) for k in d32}
        return out


def main(argv=None) -> int:
    from .workloads import WORKLOADS
    ap = argparse.ArgumentParser(description="FP32 vs FP64 run of the same kernels")
    ap.add_argument("--workload", default="c0", choices=sorted(WORKLOADS))
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--math", choices=("exact", "fast"), default="fast")
    ap.add_argument("--every", type=int, default=0, help="also report every k steps")
    args = ap.parse_args(argv)
    wl = WORKLOADS[args.workload](device="cuda:0")
    run = ShadowRun(wl.solv, *(s.to("cuda:0") for s in wl.structs),
                    fast_math=args.math == "fast")
    done = 0
    while done < args.steps:
        k = min(args.every or args.steps, args.steps - done)
        run.run(k)
        done += k
        rep = run.report()
        rep.update(workload=wl.name, math=args.math)
        print(json.dumps(rep), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
