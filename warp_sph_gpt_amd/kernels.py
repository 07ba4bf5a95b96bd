"""One Python entry per reference kernel, and a wp.launch-shaped `launch`.

Each function takes the arguments of the @wp.kernel of the same name in the
reference's kernel/ directory, in the same order and with the same meaning,
and runs the HIP implementation in libsphdem on the current stream.  Grids
are HashGrid objects (grid.id returns the grid).  Example, as in the
reference tests:

    launch(Kernel_interaction_dem, dim=n_dem,
           inputs=[P, D, grid_sph.id, grid_dem.id, mu, eps_min, eps_max, dt])
"""
from __future__ import annotations

import torch

from . import _lib
from .hashgrid import HashGrid

# ------------------------------------------------------------------ workspace


class Workspace:
    """Sorted working arrays of the fused passes (sphdem_work_t) plus the
    DEM scan scratch and the overflow flags, grown on demand."""

    def __init__(self, device, dtype=torch.float32):
        self.device = torch.device(device)
        self.dtype = dtype
        self.cap_f = self.cap_b = self.cap_scan = self.cap_d = -1
        self.wl_static = None   # see static_walls()
        self.flags = torch.zeros(4, dtype=torch.int32, device=self.device)
        # fluid bounding box scratch (sphdem_work_t.fl_box), refilled every
        # step; None turns the wall pass's far-wall skip off
        self.fl_box = torch.zeros(6 * 257, dtype=dtype, device=self.device)
        self.reserve(1, 1, 1)

    def reserve(self, nf: int, nb: int, nd: int) -> None:
        d, t = self.device, self.dtype
        if nf > self.cap_f:
            c = max(nf, 1)
            self.fl_rho = torch.zeros(c, dtype=t, device=d)
            self.fl_pres = torch.zeros(c, dtype=t, device=d)
            self.fl_V = torch.zeros(c, dtype=t, device=d)
            self.fl_vel = torch.zeros((c, 4), dtype=t, device=d)
            self.fl_pgf = torch.zeros((c, 4), dtype=t, device=d)
            self.fl_flag = torch.zeros(c, dtype=torch.int32, device=d)
            # the symmetric porosity's per-slot own sums and (as c uint64)
            # fixed-point accumulators of the mirrored terms, zero between steps
            self.fl_sum = torch.zeros(c, dtype=t, device=d) if t == torch.float32 else None
            self.fl_list = torch.zeros(2 * c, dtype=torch.int32, device=d)
            self.fl_chunk = torch.zeros(7 * c, dtype=torch.int32, device=d)
            self.fl_count = torch.zeros(4, dtype=torch.int32, device=d)
            self.fl_scan = torch.zeros(_lib.load().sphdem_scan_scratch(c), dtype=torch.int32,
                                       device=d)
            self.cap_f = c
        if nb > self.cap_b:
            c = max(nb, 1)
            self.wl_rho = torch.zeros(c, dtype=t, device=d)
            self.wl_pres = torch.zeros(c, dtype=t, device=d)
            self.wl_vel = torch.zeros((c, 4), dtype=t, device=d)
            self.cap_b = c
            if self.wl_static is not None:
                self.static_walls()
        need = _lib.load().sphdem_scan_scratch(max(nd, 1) + 1)
        if need > self.cap_scan:
            self.scan_tmp = torch.zeros(need, dtype=torch.int32, device=d)
            self.cap_scan = need
        if nd > self.cap_d:
            # zero between steps: the coupled step sets and clears the flags it uses
            self.dem_near = torch.zeros(max(nd, 1), dtype=torch.int32, device=d)
            self.cap_d = max(nd, 1)

    def static_walls(self) -> None:
        """Keep the wall-only density sums of wall particles with no fluid
        near (sphdem_work_t.wl_static) for walls that never move: only for
        an owner that builds its wall grid once (an engine).  Call again
        after changing the wall grid, positions or masses."""
        self.wl_static = torch.full((max(self.cap_b, 1),), float("nan"), dtype=self.dtype,
                                    device=self.device)

    def struct(self) -> _lib.WorkStruct:
        w = _lib.WorkStruct()
        for name in ("fl_rho", "fl_pres", "fl_V", "fl_vel", "fl_pgf", "wl_rho", "wl_pres",
                     "wl_vel", "scan_tmp", "flags", "fl_flag", "fl_list", "fl_count", "fl_scan",
                     "dem_near"):
            setattr(w, name, getattr(self, name).data_ptr())
        w.wl_static = self.wl_static.data_ptr() if self.wl_static is not None else None
        w.fl_box = self.fl_box.data_ptr() if self.fl_box is not None else None
        w.fl_sum = self.fl_sum.data_ptr() if self.fl_sum is not None else None
        w.fl_chunk = self.fl_chunk.data_ptr() if self.fl_chunk is not None else None
        w.cap_f, w.cap_b, w.cap_d = self.cap_f, self.cap_b, self.cap_d
        return w


_WORKSPACES: dict = {}


def workspace(device, dtype, nf=1, nb=1, nd=1) -> Workspace:
    key = (str(torch.device(device)), dtype)
    ws = _WORKSPACES.get(key)
    if ws is None:
        ws = _WORKSPACES[key] = Workspace(device, dtype)
    ws.reserve(nf, nb, nd)
    return ws


def _prec(struct) -> int:
    return _lib.precision_of(struct.pos.dtype)


def _grid(g) -> HashGrid:
    if not isinstance(g, HashGrid):
        raise TypeError("expected a HashGrid (or grid.id)")
    return g


# ------------------------------------------------------------------ helpers


def wendland(r: torch.Tensor, h: float):
    """Kernel_w_Wendland and Kernel_dw_Wendland (kernel/Kernel_KNL.py:7-41)
    sampled at every r; returns (W, dW/dr)."""
    w, dw = torch.empty_like(r), torch.empty_like(r)
    _lib.call("sphdem_wendland", _lib.precision_of(r.dtype), _lib.ptr(r, r.dtype, "r"),
              r.numel(), float(h), _lib.ptr(w), _lib.ptr(dw), _lib.stream_handle())
    return w, dw


def DEM_drag_beta(eps, rho, mu, diameter, speed) -> torch.Tensor:
    """beta/(1-eps) of kernel/KERNEL_SPHDEM_interaction.py:11-40, elementwise
    over tensors (scalars broadcast)."""
    eps = torch.as_tensor(eps)
    ref = eps if eps.is_cuda else None
    args = [torch.as_tensor(a) for a in (eps, rho, mu, diameter, speed)]
    dev = ref.device if ref is not None else next((a.device for a in args if a.is_cuda), None)
    if dev is None:
        raise ValueError("DEM_drag_beta needs at least one GPU tensor")
    dtype = next((a.dtype for a in args if a.is_floating_point()), torch.float32)
    shape = torch.broadcast_shapes(*(a.shape for a in args))
    args = [a.to(device=dev, dtype=dtype).expand(shape).contiguous() for a in args]
    out = torch.empty(shape, dtype=dtype, device=dev)
    _lib.call("sphdem_drag_beta", _lib.precision_of(dtype), *[_lib.ptr(a) for a in args],
              out.numel(), _lib.ptr(out), _lib.stream_handle())
    return out


def array_scan(src: torch.Tensor, dst: torch.Tensor, inclusive: bool = False) -> None:
    """wp.utils.array_scan for int32 tensors (exclusive or inclusive)."""
    if src.dtype != torch.int32 or dst.dtype != torch.int32:
        raise ValueError("array_scan supports int32 arrays")
    if dst.data_ptr() != src.data_ptr():
        dst.copy_(src)
    n = dst.numel()
    tmp = torch.zeros(_lib.load().sphdem_scan_scratch(max(n, 1)), dtype=torch.int32,
                      device=dst.device)
    if inclusive:
        last = dst.clone()
    _lib.call("sphdem_exclusive_scan", _lib.ptr(dst, torch.int32, "dst"), n, _lib.ptr(tmp),
              _lib.stream_handle())
    if inclusive:
        dst.add_(last)


def _sph_params(P, **kw):
    return _lib.params(_prec(P), **kw)


# ------------------------------------------------------------------ SPH kernels


def Kernel_shepard_sph(P_sph, P_bnd, grid_sph, grid_bnd, support, h):
    """kernel/KERNEL_rho.py:6-40."""
    gs, gb = _grid(grid_sph), _grid(grid_bnd)
    gs.ensure_weights(P_sph.m)
    gb.ensure_weights(P_bnd.m)
    p = _sph_params(P_sph, support=float(support), h=float(h))
    _lib.call("sphdem_shepard_sph", p, _lib.sph_view(P_sph), _lib.bnd_view(P_bnd), gs.struct(),
              gb.struct(), _lib.stream_handle())


def Kernel_density_sph(P_sph, P_bnd, grid_sph, grid_bnd, support, h):
    """kernel/KERNEL_rho.py:45-82."""
    gs, gb = _grid(grid_sph), _grid(grid_bnd)
    gs.ensure_weights(P_sph.m)
    gb.ensure_weights(P_bnd.m)
    p = _sph_params(P_sph, support=float(support), h=float(h))
    _lib.call("sphdem_density_sph", p, _lib.sph_view(P_sph), _lib.bnd_view(P_bnd), gs.struct(),
              gb.struct(), _lib.stream_handle())


def Kernel_density_bnd(P_sph, P_bnd, grid_sph, grid_bnd, support, h):
    """kernel/KERNEL_rho.py:87-126."""
    gs, gb = _grid(grid_sph), _grid(grid_bnd)
    gs.ensure_weights(P_sph.m)
    gb.ensure_weights(P_bnd.m)
    p = _sph_params(P_sph, support=float(support), h=float(h))
    _lib.call("sphdem_density_bnd", p, _lib.sph_view(P_sph), _lib.bnd_view(P_bnd), gs.struct(),
              gb.struct(), _lib.stream_handle())


def _pres(P, rho0, c0, gamma):
    p = _sph_params(P, rho0=float(rho0), c0=float(c0), gamma=float(gamma))
    _lib.call("sphdem_pres", p, _lib.ptr(P.rho, P.pos.dtype, "rho"),
              _lib.ptr(P.pres, P.pos.dtype, "pres"), P.n, _lib.stream_handle())


def Kernel_pres_sph(P_sph, rho0, c0, gamma):
    """kernel/KERNEL_pres.py:6-23."""
    _pres(P_sph, rho0, c0, gamma)


def Kernel_pres_bnd(P_bnd, rho0, c0, gamma):
    """kernel/KERNEL_pres.py:26-45."""
    _pres(P_bnd, rho0, c0, gamma)


def Kernel_force_sph(P_sph, P_bnd, grid_sph, grid_bnd, support, h, mu, g):
    """kernel/KERNEL_force.py:7-71."""
    gs, gb = _grid(grid_sph), _grid(grid_bnd)
    gs.ensure_weights(P_sph.m)
    gb.ensure_weights(P_bnd.m)
    ws = workspace(P_sph.device, P_sph.dtype, P_sph.n, P_bnd.n)
    p = _sph_params(P_sph, support=float(support), h=float(h), mu=float(mu), g=float(g))
    _lib.call("sphdem_force_sph", p, _lib.sph_view(P_sph), _lib.bnd_view(P_bnd), gs.struct(),
              gb.struct(), ws.struct(), _lib.stream_handle())


def Kernel_step_sph(P_sph, dt):
    """kernel/KERNEL_step.py:6-22."""
    p = _sph_params(P_sph, dt=float(dt))
    _lib.call("sphdem_step_sph", p, _lib.sph_view(P_sph), _lib.stream_handle())


# ------------------------------------------------------------------ DEM kernels


def _check_flags(ws: Workspace, what: str) -> None:
    if int(ws.flags[0].item()) & 1:
        ws.flags.zero_()
        raise OverflowError(f"{what}: contact CSR row exceeds the allocated *_new arrays")


def Kernel_count_dem_contacts(P_dem, grid_dem, radius_max):
    """kernel/KERNEL_DEM_force.py:9-22 (also zeroes offset_new[n])."""
    p = _lib.params(_prec(P_dem))
    _lib.call("sphdem_count_dem_contacts", p, _lib.dem_view(P_dem), _grid(grid_dem).struct(),
              float(radius_max), _lib.stream_handle())


def Kernel_force_dem(P_dem, grid_dem, radius_max, g, dt):
    """kernel/KERNEL_DEM_force.py:77-135."""
    ws = workspace(P_dem.device, P_dem.dtype, nd=P_dem.n)
    p = _lib.params(_prec(P_dem), dt=float(dt))
    _lib.call("sphdem_force_dem", p, _lib.dem_view(P_dem), _grid(grid_dem).struct(),
              float(radius_max), float(g), ws.flags.data_ptr(), _lib.stream_handle())
    _check_flags(ws, "Kernel_force_dem")


def Kernel_count_bnd_contacts(P_dem, P_dem_bnd, grid_dem_bnd, radius_max):
    """kernel/KERNEL_DEM_BC.py:10-24."""
    p = _lib.params(_prec(P_dem))
    _lib.call("sphdem_count_bnd_contacts", p, _lib.dem_view(P_dem), _lib.dembnd_view(P_dem_bnd),
              _grid(grid_dem_bnd).struct(), float(radius_max), _lib.stream_handle())


def Kernel_bc_dem(P_dem, P_dem_bnd, grid_dem_bnd, radius_max, dt):
    """kernel/KERNEL_DEM_BC.py:27-84."""
    ws = workspace(P_dem.device, P_dem.dtype, nd=P_dem.n)
    p = _lib.params(_prec(P_dem), dt=float(dt))
    _lib.call("sphdem_bc_dem", p, _lib.dem_view(P_dem), _lib.dembnd_view(P_dem_bnd),
              _grid(grid_dem_bnd).struct(), float(radius_max), ws.flags.data_ptr(),
              _lib.stream_handle())
    _check_flags(ws, "Kernel_bc_dem")


def Kernel_step_dem(P_dem, dt):
    """kernel/KERNEL_DEM_step.py:6-30."""
    p = _lib.params(_prec(P_dem), dt=float(dt))
    _lib.call("sphdem_step_dem", p, _lib.dem_view(P_dem), _lib.stream_handle())


# ------------------------------------------------------------------ coupling kernels


def Kernel_prep_sphdem(P_sph, P_bnd, P_dem, grid_sph, grid_bnd, grid_dem, support, h,
                       sph_support, sph_h, eps_min, eps_max):
    """kernel/KERNEL_SPHDEM_interaction.py:43-115."""
    gs, gb, gd = _grid(grid_sph), _grid(grid_bnd), _grid(grid_dem)
    gs.ensure_weights(P_sph.m)
    gb.ensure_weights(P_bnd.m)
    ws = workspace(P_sph.device, P_sph.dtype, P_sph.n, P_bnd.n, P_dem.n)
    p = _sph_params(P_sph, support=float(sph_support), h=float(sph_h), eps_min=float(eps_min),
                    eps_max=float(eps_max))
    _lib.call("sphdem_prep_sphdem", p, _lib.sph_view(P_sph), _lib.bnd_view(P_bnd),
              _lib.dem_view(P_dem), gs.struct(), gb.struct(), gd.struct(), float(support),
              float(h), ws.struct(), _lib.stream_handle())


def Kernel_interaction_dem(P_sph, P_dem, grid_sph, grid_dem, mu, eps_min, eps_max, dt):
    """kernel/KERNEL_SPHDEM_interaction.py:118-187."""
    gs, gd = _grid(grid_sph), _grid(grid_dem)
    ws = workspace(P_sph.device, P_sph.dtype, P_sph.n, 1, P_dem.n)
    p = _sph_params(P_sph, mu=float(mu), eps_min=float(eps_min), eps_max=float(eps_max),
                    dt=float(dt))
    _lib.call("sphdem_interaction_dem", p, _lib.sph_view(P_sph), _lib.dem_view(P_dem),
              gs.struct(), gd.struct(), ws.struct(), _lib.stream_handle())


def Kernel_interaction_sph(P_sph, P_dem, grid_sph, grid_dem, support):
    """kernel/KERNEL_SPHDEM_interaction.py:190-226."""
    p = _sph_params(P_sph)
    _lib.call("sphdem_interaction_sph", p, _lib.sph_view(P_sph), _lib.dem_view(P_dem),
              _grid(grid_sph).struct(), _grid(grid_dem).struct(), float(support),
              _lib.stream_handle())


# ------------------------------------------------------------------ launch

_SUBJECT = {
    Kernel_shepard_sph: 0, Kernel_density_sph: 0, Kernel_force_sph: 0, Kernel_step_sph: 0,
    Kernel_pres_sph: 0, Kernel_prep_sphdem: 0, Kernel_interaction_sph: 0,
    Kernel_density_bnd: 1, Kernel_pres_bnd: 0,
    Kernel_count_dem_contacts: 0, Kernel_force_dem: 0, Kernel_count_bnd_contacts: 0,
    Kernel_bc_dem: 0, Kernel_step_dem: 0, Kernel_interaction_dem: 1,
}


def launch(kernel, dim: int, inputs) -> None:
    """wp.launch(kernel, dim=..., inputs=[...]): dim must be the kernel's
    subject count (one thread per subject)."""
    inputs = list(inputs)
    idx = _SUBJECT.get(kernel)
    if idx is None:
        raise ValueError(f"{getattr(kernel, '__name__', kernel)} is not a launchable kernel")
    n = inputs[idx].n
    if int(dim) != n:
        raise ValueError(f"dim={dim} does not match the {n} subjects of "
                         f"{kernel.__name__}")
    kernel(*inputs)
