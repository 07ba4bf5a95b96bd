"""Hash grid: the stand-in for wp.HashGrid.

Mirrors the reference's use of Warp's grid: HashGrid(dim_x, dim_y, dim_z)
(main.py:117-122), build(points=..., radius=...) (main.py:66,73,85,87) and
the handle passed to kernels as grid.id.  Buckets, point order and queries
follow Warp's hash grid (bucket key of int(p/radius) folded into
dim_x*dim_y*dim_z buckets, stable sort by key, query order z-outer/y/x).
The build runs on the GPU (libsphdem: keys, atomic counting sort, device
scan, per-bucket ordering); all buffers are torch tensors owned here.

A grid can also batch independent scenes ("instances"): with
`instances(inst, n_inst)` every point carries its instance number (a
nondecreasing int32 array), each instance gets its own dim_x*dim_y*dim_z
buckets, and a query by a point of instance k sees instance k only.

Beyond Warp's interface the grid keeps a sorted copy of the points with a
fourth column (`weights`, e.g. particle mass), which the neighbour kernels
read contiguously.  Kernels read neighbour positions from that copy, i.e.
the positions at build time, as every caller in the step loop builds the
grid from the current positions first.
"""
from __future__ import annotations

import torch

from . import _lib


def _tag(t):
    return None if t is None else (t.data_ptr(), t._version, t.numel(), t.dtype)


class HashGrid:
    """A bucket grid of dim_x*dim_y*dim_z cells on one GPU."""

    def __init__(self, dim_x: int, dim_y: int, dim_z: int, device=None):
        for d in (dim_x, dim_y, dim_z):
            if int(d) < 1:
                raise ValueError("hash grid dimensions must be positive")
        self.dim_x, self.dim_y, self.dim_z = int(dim_x), int(dim_y), int(dim_z)
        self.device = torch.device(device if device is not None else "cuda")
        if self.device.type != "cuda":
            raise ValueError("HashGrid lives on a GPU device")
        self.n_inst, self.inst = 1, None
        self._alloc_buckets()
        self._cap = 0
        self._cells = self._ids = self._rank = self._sorted = None
        self.n = 0
        self.radius = None
        self.dtype = torch.float32
        self.points = None
        self._w_tag = None
        self.builds = 0  # number of builds so far (caches derived from the buckets key on it)

    def _alloc_buckets(self) -> None:
        ncell = self.dim_x * self.dim_y * self.dim_z * self.n_inst
        self.ncell = ncell
        self.cell_lo = torch.zeros(ncell, dtype=torch.int32, device=self.device)
        self.cell_hi = torch.zeros(ncell, dtype=torch.int32, device=self.device)
        self._scan_tmp = torch.zeros(_lib.load().sphdem_scan_scratch(ncell), dtype=torch.int32,
                                     device=self.device)

    def instances(self, inst: torch.Tensor, n_inst: int):
        """Batch n_inst scenes: inst[i] (int32, nondecreasing, on the grid's
        device) is the instance of point i in every later build."""
        n_inst = int(n_inst)
        if n_inst < 1:
            raise ValueError("n_inst must be positive")
        if inst.dtype != torch.int32 or inst.ndim != 1 or inst.device != self.device:
            raise ValueError("inst must be a 1-D int32 tensor on the grid's device")
        if inst.numel() and (int(inst.min()) < 0 or int(inst.max()) >= n_inst
                             or bool((inst[1:] < inst[:-1]).any())):
            raise ValueError("inst must be nondecreasing values in [0, n_inst)")
        self.inst = inst.contiguous()
        if n_inst != self.n_inst:
            self.n_inst = n_inst
            self._alloc_buckets()
        return self

    @property
    def id(self):
        """Kernels take the grid object itself where Warp takes grid.id."""
        return self

    def reserve(self, n: int, dtype=torch.float32) -> None:
        if n <= self._cap and self._sorted is not None and self._sorted.dtype == dtype:
            return
        cap = max(n, 1)
        self._cells = torch.zeros(cap, dtype=torch.int32, device=self.device)
        self._ids = torch.zeros(cap, dtype=torch.int32, device=self.device)
        self._rank = torch.zeros(cap, dtype=torch.int32, device=self.device)
        self._sorted = torch.zeros((cap, 4), dtype=dtype, device=self.device)
        self._cap = cap

    def struct(self) -> _lib.GridStruct:
        g = _lib.GridStruct()
        g.precision = _lib.precision_of(self.dtype)
        g.dim_x, g.dim_y, g.dim_z = self.dim_x, self.dim_y, self.dim_z
        g.cell_width = float(self.radius) if self.radius is not None else 1.0
        g.n = self.n
        g.point_cells = self._cells.data_ptr() if self._cells is not None else None
        g.point_ids = self._ids.data_ptr() if self._ids is not None else None
        g.rank = self._rank.data_ptr() if self._rank is not None else None
        g.cell_lo = self.cell_lo.data_ptr()
        g.cell_hi = self.cell_hi.data_ptr()
        g.sorted = self._sorted.data_ptr() if self._sorted is not None else None
        g.scan_tmp = self._scan_tmp.data_ptr()
        g.n_inst = self.n_inst
        g.inst = self.inst.data_ptr() if self.inst is not None and self.inst.numel() else None
        return g

    def build(self, points: torch.Tensor, radius: float, weights: torch.Tensor | None = None):
        """Bucket the (N, 3) points with cell width `radius`; `weights` (N,)
        fills the fourth column of the sorted copy."""
        if points.ndim != 2 or points.shape[1] != 3:
            raise ValueError("points must be an (N, 3) tensor")
        if points.device != self.device and not (points.is_cuda and self.device.index is None):
            raise ValueError(f"points live on {points.device}, the grid on {self.device}")
        radius = float(radius)
        if not radius > 0.0:
            raise ValueError("grid radius must be positive")
        n = int(points.shape[0])
        if self.inst is not None and self.inst.numel() != n:
            raise ValueError(f"inst has {self.inst.numel()} entries for {n} points")
        self.dtype = points.dtype
        _lib.precision_of(self.dtype)
        self.reserve(n, self.dtype)
        self.n = n
        self.radius = radius
        self.points = points
        g = self.struct()
        _lib.call("sphdem_grid_build", g, _lib.ptr(points, self.dtype, "points"), n, radius,
                  _lib.ptr(weights, self.dtype, "weights"), _lib.stream_handle())
        self._w_tag = _tag(weights)
        self.builds += 1
        return self

    def ensure_weights(self, w: torch.Tensor) -> None:
        """Make the sorted copy's fourth column equal w[point_ids]."""
        if w.shape[0] != self.n:
            raise ValueError("weights must have one value per grid point")
        tag = _tag(w)
        if tag != self._w_tag:
            _lib.call("sphdem_grid_set_w", self.struct(), _lib.ptr(w, self.dtype, "weights"),
                      _lib.stream_handle())
            self._w_tag = tag

    # ---- parity views -------------------------------------------------
    @property
    def point_ids(self) -> torch.Tensor:
        """Sorted slot -> original index (hash_grid_point_id)."""
        return self._ids[: self.n]

    @property
    def point_cells(self) -> torch.Tensor:
        """Bucket key of each original point."""
        return self._cells[: self.n]

    @property
    def sorted_points(self) -> torch.Tensor:
        return self._sorted[: self.n]

    def warp_bounds(self):
        """(cell_starts, cell_ends) with Warp's convention: 0, 0 when empty."""
        s = torch.empty(self.ncell, dtype=torch.int32, device=self.device)
        e = torch.empty(self.ncell, dtype=torch.int32, device=self.device)
        _lib.call("sphdem_grid_warp_bounds", self.struct(), s.data_ptr(), e.data_ptr(),
                  _lib.stream_handle())
        return s, e


def neighbour_counts(grid_subject: HashGrid, grid_neighbour: HashGrid, radius: float) -> torch.Tensor:
    """True neighbours (r^2 < radius^2, self included) of every subject point
    among the neighbour grid's points, at the subject's original index."""
    out = torch.zeros(grid_subject.n, dtype=torch.int32, device=grid_subject.device)
    _lib.call("sphdem_neighbour_counts", grid_subject.struct(), grid_neighbour.struct(),
              float(radius), _lib.ptr(out, torch.int32, "counts"), _lib.stream_handle())
    return out
