"""Two backends behind one interface, so each known-answer case runs on the
CPU oracle (no GPU needed) and on the HIP path (GPU), the way the
reference's tests loop over their devices.

The kernel functions of both backends carry the reference's kernel names
and argument lists; structs carry the reference's field names.
"""
from __future__ import annotations

import numpy as np
import torch

from helpers import COMPACT, generation, particles, solv_for
from oracle import oracle as O


class Scene:
    def __init__(self, bk, gen, solv, structs, grids):
        self.bk, self.gen, self.solv = bk, gen, solv
        self.P, self.B, self.D, self.W = structs
        self.grids = grids


class OracleBackend:
    name = "oracle"
    K = O

    def scene(self, **kw):
        settings = dict(COMPACT)
        settings.update(kw)
        gen = generation(**settings)
        structs = [O.OStruct.from_struct(s) for s in particles(gen)]
        grids = [self.grid(32) for _ in range(4)]
        for g, s, r in zip(grids, structs, (gen.support, gen.support, gen.dem_support,
                                            2 * gen.dem_bnd_radius)):
            g.build(s.pos, r)
        return Scene(self, gen, solv_for(gen, "cpu"), structs, grids)

    def from_structs(self, gen, structs):
        return [O.OStruct.from_struct(s) for s in structs]

    def grid(self, d):
        return O.OracleGrid(d, d, d)

    def array(self, values, dtype=np.float32):
        return np.ascontiguousarray(np.asarray(values, dtype=dtype))

    def np(self, a):
        return np.asarray(a)

    def set(self, S, name, values):
        arr = getattr(S, name)
        arr[...] = np.asarray(values, dtype=arr.dtype)

    def build(self, grid, pos, radius):
        grid.build(pos, radius)

    def run_force_dem(self, D, grid, rmax, g, dt):
        return O.run_force_dem(D, grid, rmax, g, dt)

    def run_bc_dem(self, D, W, grid, rmax, dt):
        return O.run_bc_dem(D, W, grid, rmax, dt)

    def SPH_OneStep(self, *a, **kw):
        return O.SPH_OneStep(*a, **kw)

    def SPHDEM_OneStep(self, *a, **kw):
        return O.SPHDEM_OneStep(*a, **kw)

    def drag_beta(self, eps, speed, rho=1000.0, mu=0.05, d=0.05):
        return np.array([O.drag_beta(float(e), rho, mu, d, float(v)) for e, v in zip(eps, speed)])

    def wendland(self, r, h):
        return O.wendland(r, h)

    def sync(self):
        pass


class HipBackend:
    name = "hip"

    def __init__(self):
        from warp_sph_gpt_amd import kernels
        self.K = kernels
        self.dev = torch.device("cuda:0")

    def scene(self, **kw):
        from warp_sph_gpt_amd.hashgrid import HashGrid  # noqa: F401
        settings = dict(COMPACT)
        settings.update(kw)
        gen = generation(**settings)
        structs = [s.to(self.dev) for s in particles(gen)]
        grids = [self.grid(32) for _ in range(4)]
        for g, s, r in zip(grids, structs, (gen.support, gen.support, gen.dem_support,
                                            2 * gen.dem_bnd_radius)):
            g.build(s.pos, r)
        return Scene(self, gen, solv_for(gen), structs, grids)

    def from_structs(self, gen, structs):
        return [s.to(self.dev) for s in structs]

    def grid(self, d):
        from warp_sph_gpt_amd.hashgrid import HashGrid
        return HashGrid(d, d, d, device=self.dev)

    def array(self, values, dtype=np.float32):
        t = {np.float32: torch.float32, np.int32: torch.int32}[dtype]
        return torch.tensor(np.asarray(values), dtype=t, device=self.dev)

    def np(self, a):
        return a.detach().cpu().numpy()

    def set(self, S, name, values):
        arr = getattr(S, name)
        arr.copy_(torch.as_tensor(np.asarray(values), dtype=arr.dtype).expand_as(arr))

    def build(self, grid, pos, radius):
        grid.build(pos, radius)

    def run_force_dem(self, D, grid, rmax, g, dt):
        K = self.K
        D.contact_dem_offset_new.zero_()
        K.Kernel_count_dem_contacts(D, grid, rmax)
        K.array_scan(D.contact_dem_offset_new, D.contact_dem_offset_new)
        E = int(D.contact_dem_offset_new[-1].item())
        if E < 0:
            raise OverflowError("directed DEM contact count exceeded int32 CSR capacity")
        D.contact_dem_id_new = torch.empty(E, dtype=torch.int32, device=self.dev)
        D.tang_dem_new = torch.empty((E, 3), dtype=D.dtype, device=self.dev)
        K.Kernel_force_dem(D, grid, rmax, g, dt)
        for a in ("contact_dem_offset", "contact_dem_id", "tang_dem"):
            x, y = getattr(D, a + "_old"), getattr(D, a + "_new")
            setattr(D, a + "_old", y)
            setattr(D, a + "_new", x)
        return E

    def run_bc_dem(self, D, W, grid, rmax, dt):
        K = self.K
        D.contact_bnd_offset_new.zero_()
        K.Kernel_count_bnd_contacts(D, W, grid, rmax)
        K.array_scan(D.contact_bnd_offset_new, D.contact_bnd_offset_new)
        E = int(D.contact_bnd_offset_new[-1].item())
        D.contact_bnd_id_new = torch.empty(E, dtype=torch.int32, device=self.dev)
        D.tang_bnd_new = torch.empty((E, 3), dtype=D.dtype, device=self.dev)
        K.Kernel_bc_dem(D, W, grid, rmax, dt)
        for a in ("contact_bnd_offset", "contact_bnd_id", "tang_bnd"):
            x, y = getattr(D, a + "_old"), getattr(D, a + "_new")
            setattr(D, a + "_old", y)
            setattr(D, a + "_new", x)
        return E

    def SPH_OneStep(self, *a, **kw):
        from warp_sph_gpt_amd.simulation import SPH_OneStep
        return SPH_OneStep(*a, **kw)

    def SPHDEM_OneStep(self, *a, **kw):
        from warp_sph_gpt_amd.simulation import SPHDEM_OneStep
        return SPHDEM_OneStep(*a, **kw)

    def drag_beta(self, eps, speed, rho=1000.0, mu=0.05, d=0.05):
        e = torch.tensor(np.asarray(eps, np.float32), device=self.dev)
        v = torch.tensor(np.asarray(speed, np.float32), device=self.dev)
        return self.K.DEM_drag_beta(e, rho, mu, d, v).cpu().numpy()

    def wendland(self, r, h):
        w, dw = self.K.wendland(torch.tensor(np.asarray(r, np.float32), device=self.dev), h)
        return w.cpu().numpy(), dw.cpu().numpy()

    def sync(self):
        torch.cuda.synchronize()
