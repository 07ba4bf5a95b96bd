"""One time step: the drop-in boundary of the coupled WCSPH-DEM path.

`SPH_OneStep` and `SPHDEM_OneStep` keep the signatures, argument meaning,
side effects and exceptions of source/Simulation.py:17-99 and :102-221.
Each is a few calls into libsphdem (fused HIP passes), enqueued on the
current stream:

  SPH_OneStep:     P1 fluid Shepard+density+EOS, P1b wall density+EOS,
                   P2 force, [integration]
  SPHDEM_OneStep:  stage 0  SPH passes (P2 also forms -grad p), DEM-DEM
                            counts + scan          -> read E_dem, allocate
                   stage 1  DEM-DEM contacts, DEM-wall counts + scan
                                                   -> swap, read E_bnd, allocate
                   stage 2  DEM-wall contacts, porosity, DEM interpolation
                            and drag, drag reaction + integration
                                                   -> swap

As in the reference, the two CSR totals are read back to the host and the
*_new id/tang arrays are allocated with exactly E entries; the retired
arrays stay in *_new after the swap.  (`engine.CoupledEngine` runs the same
kernels with capacity-managed CSR buffers and no host round trip.)
"""
from __future__ import annotations

import torch

from . import _lib
from .hashgrid import HashGrid
from .input_reader import dem_search_scales
from .kernels import workspace


def solver_params(solv, dtype=torch.float32, search_scales=None) -> _lib.Params:
    """sphdem_params_t from a Solv and the cached DEM search scales."""
    p = _lib.params(_lib.precision_of(dtype), h=float(solv.h), support=float(solv.support),
                    rho0=float(solv.rho0), c0=float(solv.c0), gamma=float(solv.gamma),
                    mu=float(solv.mu), g=float(solv.g), dt=float(solv.dt),
                    eps_min=float(solv.dem_porosity_min), eps_max=float(solv.dem_porosity_max),
                    shepard_step=int(solv.shepard_step))
    if search_scales is not None:
        p.radius_max, p.boundary_radius_max, p.h_max, p.h_uniform = (float(v) for v in search_scales)
    return p


def SPH_OneStep(solv, P_sph, P_bnd, grid_sph: HashGrid, grid_bnd: HashGrid, step: int,
                integrate: bool = True) -> None:
    """Shepard filter (step 0 and every shepard_step) -> density -> pressure
    -> force -> [semi-implicit Euler] (source/Simulation.py:17-99).

    Outputs: P_sph.flt, rho, pres, acc; P_bnd.rho, pres; vel and pos when
    integrate is True."""
    grid_sph.ensure_weights(P_sph.m)
    grid_bnd.ensure_weights(P_bnd.m)
    ws = workspace(P_sph.device, P_sph.dtype, P_sph.n, P_bnd.n)
    p = solver_params(solv, P_sph.dtype)
    _lib.call("sphdem_sph_step", p, _lib.sph_view(P_sph), _lib.bnd_view(P_bnd),
              grid_sph.struct(), grid_bnd.struct(), int(step), 1 if integrate else 0,
              ws.struct(), _lib.stream_handle())


def _swap(P_dem, kind: str) -> None:
    for a, b in ((f"contact_{kind}_offset_old", f"contact_{kind}_offset_new"),
                 (f"contact_{kind}_id_old", f"contact_{kind}_id_new"),
                 (f"tang_{kind}_old", f"tang_{kind}_new")):
        x, y = getattr(P_dem, a), getattr(P_dem, b)
        setattr(P_dem, a, y)
        setattr(P_dem, b, x)


def _allocate_new(P_dem, kind: str) -> int:
    n = P_dem.n
    count = int(getattr(P_dem, f"contact_{kind}_offset_new")[n].item())
    if count < 0:
        label = "directed DEM" if kind == "dem" else "DEM-boundary"
        raise OverflowError(f"{label} contact count exceeded int32 CSR capacity")
    setattr(P_dem, f"contact_{kind}_id_new",
            torch.empty(count, dtype=torch.int32, device=P_dem.device))
    setattr(P_dem, f"tang_{kind}_new", torch.empty((count, 3), dtype=P_dem.dtype,
                                                   device=P_dem.device))
    return count


def SPHDEM_OneStep(solv, P_sph, P_bnd, P_dem, P_dem_bnd, grid_sph: HashGrid,
                   grid_bnd: HashGrid, grid_dem: HashGrid, grid_dem_bnd: HashGrid, step: int,
                   search_scales=None) -> None:
    """Coupled step at a common time level (source/Simulation.py:102-221):
    SPH fields/forces -> DEM contacts (CSR rebuilt) -> porosity, -grad p,
    interpolation, drag and its reaction -> SPH and DEM integration."""
    scales = dem_search_scales(P_dem, P_dem_bnd) if search_scales is None else search_scales
    grid_sph.ensure_weights(P_sph.m)
    grid_bnd.ensure_weights(P_bnd.m)
    ws = workspace(P_sph.device, P_sph.dtype, P_sph.n, P_bnd.n, P_dem.n)
    p = solver_params(solv, P_sph.dtype, scales)
    grids = (grid_sph.struct(), grid_bnd.struct(), grid_dem.struct(), grid_dem_bnd.struct())
    sph, bnd, wall = _lib.sph_view(P_sph), _lib.bnd_view(P_bnd), _lib.dembnd_view(P_dem_bnd)
    st = _lib.stream_handle()

    def stage(k):
        _lib.call("sphdem_coupled_step", p, sph, bnd, _lib.dem_view(P_dem), wall, *grids,
                  int(step), k, ws.struct(), st)

    stage(0)
    _allocate_new(P_dem, "dem")
    stage(1)
    _swap(P_dem, "dem")
    _allocate_new(P_dem, "bnd")
    stage(2)
    _swap(P_dem, "bnd")
