"""Batched ensemble: many independent dam breaks advanced together.

The instances' particles are stored back to back in one struct of each
kind (instance k owns one contiguous range of every kind), every grid is
built batched (each instance has its own buckets, see HashGrid.instances),
and each pass of the coupled step is ONE launch over all instances: a
subject queries only its own instance's buckets, so it adds exactly the
terms, in the same order, that it adds when its instance runs alone.  The
whole step is captured in a hipGraph and replayed, i.e. one graph launch
per step for the ensemble.

Contact CSR ids of the batched DEM refer to the batched arrays (instance
offsets added); `instance_state(k)` takes them back to instance numbering.

Instances share the solver settings (dx, h, dt, ...).  Search scales
(largest DEM radius and h) are the maxima over the ensemble; they size the
query boxes only, which does not change any instance's terms.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .engine import CoupledEngine
from .struct import BNDptl, DEMBNDptl, DEMptl, SPHptl, IDS, OFFSET

KINDS = ("fluid", "wall", "dem", "dem_wall")


def _offsets(counts) -> list:
    out = [0]
    for c in counts:
        out.append(out[-1] + int(c))
    return out


def concat_structs(cls, parts: list, id_offsets: dict | None = None):
    """One struct holding `parts` back to back.  CSR rows are re-based:
    offsets by the edges before, ids by `id_offsets[kind][k]` (the first
    batched index of instance k's neighbour set)."""
    out = {}
    for name, layout in cls.FIELDS:
        ts = [getattr(p, name) for p in parts]
        if layout == OFFSET:
            idname = name.replace("offset", "id")
            acc, rows = 0, []
            for p, t in zip(parts, ts):
                rows.append(t[:-1].to(torch.int64) + acc)
                acc += int(getattr(p, idname).shape[0])
            rows.append(torch.tensor([acc], dtype=torch.int64, device=ts[0].device))
            out[name] = torch.cat(rows).to(torch.int32)
        elif layout == IDS:
            kind = "dem" if "_dem_" in name else "bnd"
            base = id_offsets[kind]
            out[name] = torch.cat([t + int(base[k]) for k, t in enumerate(ts)])
        else:
            out[name] = torch.cat(ts)
    return cls(**out)


@dataclass
class Layout:
    """Per-kind instance ranges: offsets[kind][k] .. offsets[kind][k+1]."""

    offsets: dict

    @property
    def n(self) -> int:
        return len(self.offsets["fluid"]) - 1

    def inst_array(self, kind: str, device) -> torch.Tensor:
        off = self.offsets[kind]
        counts = np.diff(np.asarray(off))
        return torch.from_numpy(np.repeat(np.arange(len(counts), dtype=np.int32),
                                          counts)).to(device)


class EnsembleEngine:
    """N independent scenes ((P, B, D, W) each; D and W may be None or empty)
    advanced by one CoupledEngine over the batched state."""

    def __init__(self, solv, scenes: list, device="cuda:0", csr_per_dem: int = 64,
                 fast_math: bool | None = None):
        if not scenes:
            raise ValueError("the ensemble needs at least one scene")
        dev = torch.device(device)
        P0 = scenes[0][0]
        dtype = P0.dtype
        Ps, Bs, Ds, Ws = [], [], [], []
        for P, B, D, W in scenes:
            if D is None or D.n == 0:
                D = DEMptl.allocate(0, dtype=dtype)
                W = DEMBNDptl.allocate(0, dtype=dtype)
            elif W is None:
                W = DEMBNDptl.allocate(0, dtype=dtype)
            Ps.append(P.to(dev))
            Bs.append(B.to(dev))
            Ds.append(D.to(dev))
            Ws.append(W.to(dev))
        self.layout = Layout({
            "fluid": _offsets(p.n for p in Ps), "wall": _offsets(b.n for b in Bs),
            "dem": _offsets(d.n for d in Ds), "dem_wall": _offsets(w.n for w in Ws)})
        lay = self.layout.offsets
        self.P = concat_structs(SPHptl, Ps)
        self.B = concat_structs(BNDptl, Bs)
        nd = lay["dem"][-1]
        self.D = concat_structs(DEMptl, Ds, {"dem": lay["dem"], "bnd": lay["dem_wall"]}) \
            if nd else None
        self.W = concat_structs(DEMBNDptl, Ws) if nd else None
        inst = {k: self.layout.inst_array(k, dev) for k in KINDS}
        inst["n"] = self.layout.n
        self.engine = CoupledEngine(solv, self.P, self.B, self.D, self.W,
                                    csr_per_dem=csr_per_dem, fast_math=fast_math, instances=inst)
        # the engine may replace CSR tensors (capacity buffers): read them from it
        self.D = self.engine.D

    # ---------------------------------------------------------------- stepping
    @property
    def n(self) -> int:
        return self.layout.n

    @property
    def n_fluid(self) -> int:
        return self.P.n

    @property
    def step_index(self) -> int:
        return self.engine.step_index

    def step(self) -> None:
        self.engine.step()

    def run(self, n_steps: int) -> None:
        self.engine.run(n_steps)

    def capture(self) -> None:
        self.engine.capture()

    def check(self) -> None:
        self.engine.check()

    # ---------------------------------------------------------------- views
    def _slice(self, kind: str, k: int) -> slice:
        off = self.layout.offsets[kind]
        return slice(off[k], off[k + 1])

    def fluid(self, k: int, field: str) -> torch.Tensor:
        """Instance k's rows of a fluid field (a view of the batched array)."""
        return getattr(self.P, field)[self._slice("fluid", k)]

    def instance_state(self, k: int):
        """Copies of instance k's (P, B, D, W), CSR in instance numbering
        (D, W are None for an instance without DEM)."""
        P = SPHptl(**{n: getattr(self.P, n)[self._slice("fluid", k)].clone()
                      for n in SPHptl.field_names()})
        B = BNDptl(**{n: getattr(self.B, n)[self._slice("wall", k)].clone()
                      for n in BNDptl.field_names()})
        sd = self._slice("dem", k)
        if self.D is None or sd.stop == sd.start:
            return P, B, None, None
        W = DEMBNDptl(**{n: getattr(self.W, n)[self._slice("dem_wall", k)].clone()
                         for n in DEMBNDptl.field_names()})
        vals = {}
        base = {"dem": self.layout.offsets["dem"][k], "bnd": self.layout.offsets["dem_wall"][k]}
        for name, layout in DEMptl.FIELDS:
            t = getattr(self.D, name)
            if layout == OFFSET:
                rows = t[sd.start:sd.stop + 1]
                vals[name] = (rows - rows[0]).clone()
            elif layout in (IDS,) or name.startswith("tang_"):
                continue
            else:
                vals[name] = t[sd].clone()
        for kind in ("dem", "bnd"):
            for side in ("old", "new"):
                off = getattr(self.D, f"contact_{kind}_offset_{side}")
                a, b = int(off[sd.start]), int(off[sd.stop])
                vals[f"contact_{kind}_id_{side}"] = (
                    getattr(self.D, f"contact_{kind}_id_{side}")[a:b] - base[kind]).clone()
                vals[f"tang_{kind}_{side}"] = getattr(self.D, f"tang_{kind}_{side}")[a:b].clone()
        return P, B, DEMptl(**vals), W
