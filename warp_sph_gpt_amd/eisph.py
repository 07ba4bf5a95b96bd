"""Eulerian ISPH lid-driven cavity on the GPU.

Mirrors source/EISPH.py:26-141 (EISPH_OneStep, run_eisph) and the kernel
launches of kernel/KERNEL_EISPH_{KNL,BC,force,PPE,step}.py, with the same
names, arguments and outputs; the work runs in libsphdem.so (csrc/eisph.hip).

The points are fixed, so the HashGrid walks are done once: after the
kernel-gradient correction (Kernel_prepare_kgc), a `NeighbourList` records,
per fluid point and in the grid's visiting order, every neighbour inside the
support with its corrected gradient, Laplacian weight and volume.  Every pass
then streams that list and adds the reference's terms in the reference's
order.  The list is cached on the fluid grid and rebuilt when either grid is
rebuilt, the KGC is recomputed, or m or rho change.

`EISPH_OneStep` is five launches (sphdem_eisph_step); the kernel-level
functions run one reference launch each and read the ghost fields as stored,
as the reference's kernels do.  `EISPH_Steps` runs the steps between two
output steps (sphdem_eisph_steps): one launch of one workgroup for a small
cloud, with the results of as many `EISPH_OneStep` calls.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .hashgrid import HashGrid
from .input_reader import validate_eisph_particles


def _params(S, support, h, rho0=1.0, nu=0.0, dt=1.0) -> _lib.EisphParams:
    p = _lib.eisph_params(_lib.precision_of(S.pos.dtype), h, rho0, nu, dt)
    p.support = float(support)
    return p


def _stamp(S) -> tuple:
    """Identity and in-place version of the fields a list depends on."""
    return tuple((t.data_ptr(), t._version) for t in (S.pos, S.m, S.rho))


class NeighbourList:
    """The fixed neighbour list of the fluid points (sphdem_eisph_nbr_t):
    slices of 64 rows, entry k of row i in slot offset[i // 64] + 64 k + i % 64."""

    SLICE = 64

    def __init__(self, P_sph, P_bnd, grid_sph: HashGrid, grid_bnd: HashGrid, support: float,
                 h: float):
        dev = P_sph.pos.device
        n, nb = P_sph.n, P_bnd.n
        lib = _lib.load()
        self.n = n
        self.slices = lib.sphdem_eisph_slices(n)
        self.len = torch.zeros(n + 1, dtype=torch.int32, device=dev)
        self.offset = torch.zeros(self.slices + 1, dtype=torch.int32, device=dev)
        self.width = torch.zeros(self.slices + 1, dtype=torch.int32, device=dev)
        self.scan_tmp = torch.zeros(lib.sphdem_scan_scratch(self.slices + 1), dtype=torch.int32,
                                    device=dev)
        self.part = torch.zeros(lib.sphdem_eisph_part_size(n, nb), dtype=torch.float64,
                                device=dev)
        self.id = torch.zeros(0, dtype=torch.int32, device=dev)
        self.geo = torch.zeros((0, 4), dtype=P_sph.pos.dtype, device=dev)
        p = _params(P_sph, support, h)
        Pv, Bv = _lib.eisph_view(P_sph), _lib.eisph_view(P_bnd, "P_bnd")
        gs, gb = grid_sph.struct(), grid_bnd.struct()
        st = _lib.stream_handle()
        _lib.call("sphdem_eisph_nbr_count", p, Pv, Bv, gs, gb, self.view(), st)
        slots = int(self.offset[self.slices].item())
        self.id = torch.empty(slots, dtype=torch.int32, device=dev)
        self.geo = torch.empty((slots, 4), dtype=P_sph.pos.dtype, device=dev)
        _lib.call("sphdem_eisph_nbr_fill", p, Pv, Bv, gs, gb, self.view(), st)
        self.entries = int(self.len.sum().item())
        self.key = None

    @property
    def slots(self) -> int:
        return int(self.id.shape[0])

    def rows(self):
        """Host copy as CSR: (offsets [n+1], ids [entries], geo [entries, 4]),
        row i's entries in their visiting order."""
        ln = self.len[:self.n].cpu().numpy().astype("int64")
        off = self.offset.cpu().numpy().astype("int64")
        i = np.repeat(np.arange(self.n), ln)
        k = np.arange(i.size) - np.repeat(np.cumsum(ln) - ln, ln)
        slot = off[i // self.SLICE] + self.SLICE * k + i % self.SLICE
        csr = np.concatenate(([0], np.cumsum(ln)))
        return csr, self.id.cpu().numpy()[slot], self.geo.cpu().numpy()[slot]

    def view(self) -> _lib.EisphNbr:
        v = _lib.EisphNbr()
        v.n = self.n
        v.len = self.len.data_ptr()
        v.offset = self.offset.data_ptr()
        v.width = self.width.data_ptr()
        v.scan_tmp = self.scan_tmp.data_ptr()
        v.part = self.part.data_ptr()
        v.id = self.id.data_ptr() if self.id.numel() else None
        v.geo = self.geo.data_ptr() if self.geo.numel() else None
        v.cap = self.slots
        return v


def _key(P_sph, P_bnd, grid_sph, grid_bnd, support, h) -> tuple:
    return (id(grid_bnd), grid_sph.builds, grid_bnd.builds, float(support), float(h),
            getattr(P_sph, "_kgc_epoch", 0), P_sph.kgc.data_ptr(), P_sph.kgc._version,
            _stamp(P_sph), _stamp(P_bnd))


def neighbours(P_sph, P_bnd, grid_sph, grid_bnd, support, h) -> NeighbourList:
    """The cached list for these points and grids, built when missing or stale."""
    key = _key(P_sph, P_bnd, grid_sph, grid_bnd, support, h)
    L = getattr(grid_sph, "_eisph_list", None)
    if L is None or L.key != key:
        L = NeighbourList(P_sph, P_bnd, grid_sph, grid_bnd, support, h)
        L.key = key
        grid_sph._eisph_list = L
    return L


# ------------------------------------------------------------------ kernels

def Kernel_prepare_kgc(P_sph, P_bnd, grid_sph, grid_bnd, support, h) -> None:
    """KERNEL_EISPH_KNL.py:74-132: P_sph.kgc, the inverse first-order moment
    matrix in x-z (identity where it is singular)."""
    _lib.call("sphdem_eisph_kgc", _params(P_sph, support, h), _lib.eisph_view(P_sph),
              _lib.eisph_view(P_bnd, "P_bnd"), grid_sph.struct(), grid_bnd.struct(),
              _lib.stream_handle())
    P_sph._kgc_epoch = getattr(P_sph, "_kgc_epoch", 0) + 1


def Kernel_Dirichlet_BC(P_sph, P_bnd, predicted: int) -> None:
    """KERNEL_EISPH_BC.py:6-31: ghost vel (predicted == 0) or vel_star =
    2 vel_bc - fluid[mirror]."""
    _lib.call("sphdem_eisph_dirichlet", _lib.precision_of(P_sph.pos.dtype),
              _lib.eisph_view(P_sph), _lib.eisph_view(P_bnd, "P_bnd"), int(predicted),
              _lib.stream_handle())


def Kernel_pressure_boundary_eisph(P_sph, P_bnd) -> None:
    """KERNEL_EISPH_BC.py:34-47: ghost pres = fluid pres[mirror]."""
    _lib.call("sphdem_eisph_pressure_boundary", _lib.precision_of(P_sph.pos.dtype),
              _lib.eisph_view(P_sph), _lib.eisph_view(P_bnd, "P_bnd"), _lib.stream_handle())


def _pass(name, P_sph, P_bnd, grid_sph, grid_bnd, support, h, rho0=1.0, nu=0.0, dt=1.0):
    L = neighbours(P_sph, P_bnd, grid_sph, grid_bnd, support, h)
    _lib.call(name, _params(P_sph, support, h, rho0, nu, dt), _lib.eisph_view(P_sph),
              _lib.eisph_view(P_bnd, "P_bnd"), L.view(), _lib.stream_handle())
    return L


def Kernel_advec_vis_eisph(P_sph, P_bnd, grid_sph, grid_bnd, support, h, nu, dt) -> None:
    """KERNEL_EISPH_force.py:12-72: acc = -(u.grad)u + nu lap u, vel_star."""
    _pass("sphdem_eisph_advec_vis", P_sph, P_bnd, grid_sph, grid_bnd, support, h, nu=nu, dt=dt)


def Kernel_build_ppe_eisph(P_sph, P_bnd, grid_sph, grid_bnd, support, h, rho0, dt) -> None:
    """KERNEL_EISPH_PPE.py:12-79: divergence, Aij, bi, Aijpj; also the FP64
    sum of bi into residual[0] (EISPH.py:72)."""
    _pass("sphdem_eisph_build_ppe", P_sph, P_bnd, grid_sph, grid_bnd, support, h, rho0=rho0,
          dt=dt)


def solve_ppe(P_sph, P_bnd, grid_sph, grid_bnd, support, h) -> None:
    """KERNEL_EISPH_PPE.py:82-125: Kernel_remove_bi_mean and
    Kernel_solve_ppe_eisph after Kernel_build_ppe_eisph on the same list."""
    L = neighbours(P_sph, P_bnd, grid_sph, grid_bnd, support, h)
    _lib.call("sphdem_eisph_solve_ppe", _params(P_sph, support, h), _lib.eisph_view(P_sph),
              L.view(), _lib.stream_handle())


def Kernel_correct_eisph(P_sph, P_bnd, grid_sph, grid_bnd, support, h, dt) -> None:
    """KERNEL_EISPH_step.py:12-61: u = u* - dt grad p / rho, acc."""
    _pass("sphdem_eisph_correct", P_sph, P_bnd, grid_sph, grid_bnd, support, h, dt=dt)


def Kernel_diagnostics_eisph(P_sph, P_bnd, grid_sph, grid_bnd, support, h) -> None:
    """KERNEL_EISPH_step.py:64-121: divergence and the PPE residual."""
    _pass("sphdem_eisph_diagnostics", P_sph, P_bnd, grid_sph, grid_bnd, support, h)


# ------------------------------------------------------------------ driver

def EISPH_OneStep(solv, P_sph, P_bnd, grid_sph, grid_bnd, diagnostics: bool = True) -> None:
    """source/EISPH.py:26-90: ghost velocity, predictor, pressure equation,
    projection, optional diagnostics; positions never change."""
    L = neighbours(P_sph, P_bnd, grid_sph, grid_bnd, solv.support, solv.h)
    p = _lib.eisph_params(_lib.precision_of(P_sph.pos.dtype), solv.h, solv.rho0, solv.nu,
                          solv.dt)
    p.support = float(solv.support)
    _lib.call("sphdem_eisph_step", p, _lib.eisph_view(P_sph), _lib.eisph_view(P_bnd, "P_bnd"),
              L.view(), int(bool(diagnostics)), _lib.stream_handle())


def EISPH_Steps(solv, P_sph, P_bnd, grid_sph, grid_bnd, steps: int,
                diagnostics: bool = True) -> None:
    """`steps` calls of EISPH_OneStep, diagnostics (if asked) on the last
    only: the run_eisph loop between output steps (source/EISPH.py:131-139).
    Clouds that fit one workgroup's LDS (`one_group`) run in one launch."""
    steps = int(steps)
    if steps < 0:
        raise ValueError("steps must be >= 0")
    L = neighbours(P_sph, P_bnd, grid_sph, grid_bnd, solv.support, solv.h)
    p = _lib.eisph_params(_lib.precision_of(P_sph.pos.dtype), solv.h, solv.rho0, solv.nu,
                          solv.dt)
    p.support = float(solv.support)
    _lib.call("sphdem_eisph_steps", p, _lib.eisph_view(P_sph), _lib.eisph_view(P_bnd, "P_bnd"),
              L.view(), steps, int(bool(diagnostics)), _lib.stream_handle())


def one_group(P_sph, P_bnd) -> bool:
    """True when EISPH_Steps runs this cloud in one workgroup."""
    return bool(_lib.load().sphdem_eisph_one_group(_lib.precision_of(P_sph.pos.dtype),
                                                  int(P_sph.pos.shape[0]),
                                                  int(P_bnd.pos.shape[0])))


def prepare(solv, P_sph, P_bnd, dims=None):
    """source/EISPH.py:117-126: both fixed grids (grid_slice^3 buckets, or
    `dims`, radius = support), the KGC and the neighbour list; returns
    (grid_sph, grid_bnd).  The buckets only order the one-time walk: any
    dims give the same list."""
    g = solv.grid_slice
    dims = (g, g, g) if dims is None else tuple(int(d) for d in dims)
    dev = P_sph.pos.device
    grid_sph = HashGrid(*dims, device=dev)
    grid_bnd = HashGrid(*dims, device=dev)
    grid_sph.build(P_sph.pos, solv.support)
    grid_bnd.build(P_bnd.pos, solv.support)
    Kernel_prepare_kgc(P_sph, P_bnd, grid_sph, grid_bnd, solv.support, solv.h)
    neighbours(P_sph, P_bnd, grid_sph, grid_bnd, solv.support, solv.h)
    return grid_sph, grid_bnd


def run_eisph(solv, P_sph, P_bnd, progress=None):
    """source/EISPH.py:93-141: validate, build the fixed grids and KGC once,
    run n_steps steps with diagnostics at output steps and the last step, and
    collect host velocity frames.  Returns (positions [N, 3], velocity frames,
    frame times), numpy on the host."""
    validate_eisph_particles(solv, P_sph, P_bnd)
    grid_sph, grid_bnd = prepare(solv, P_sph, P_bnd)
    positions = P_sph.pos.detach().cpu().numpy().copy()
    frames, times = [P_sph.vel.detach().cpu().numpy().copy()], [0.0]
    step, n, every = 0, int(solv.n_steps), int(solv.output_step)
    while step < n:
        # the next step the reference stores at (a multiple of output_step, or the last)
        store = min(n, (step // every + 1) * every) if every > 0 else n
        EISPH_Steps(solv, P_sph, P_bnd, grid_sph, grid_bnd, store - step, diagnostics=True)
        step = store
        frames.append(P_sph.vel.detach().cpu().numpy().copy())
        times.append(step * solv.dt)
        if progress is not None:
            progress(step, n)
    return positions, frames, times
