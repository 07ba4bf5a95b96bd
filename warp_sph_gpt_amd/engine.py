"""Step loop without host round trips, for long runs and measurement.

`CoupledEngine` runs what main.py:84-91 runs every step -- rebuild the
moving grids from the current positions, then one SPH_OneStep /
SPHDEM_OneStep -- with the same kernels as simulation.py, but:
  * the contact CSR lives in capacity-managed buffers (csr_per_dem edges
    per sphere and kind); the logical edge counts stay on the device, the
    finished rows are copied into the read side instead of swapping, and a
    device flag records overflow;
  * steps run in guarded windows of `check_every` steps: the state the steps
    change is copied aside on the device, the window's steps are enqueued,
    and the overflow flag is read once at its end.  If a row did not fit,
    the state is put back, the capacity grows and the window runs again, so
    no step ever continues on a truncated contact history (the reference
    reads the totals and raises before filling, source/Simulation.py:151-154,
    184-187; here the step is redone with room for every edge instead);
  * every buffer address is fixed between capacity changes, so one step can
    be captured in a hipGraph and replayed (`capture()`): two graphs, with
    and without the Shepard refresh, captured again after a capacity change.
With `instances` the structs hold several independent scenes back to back
(ensemble.py): every grid is built batched, and each kernel queries the
buckets of its subject's own scene only.
"""
from __future__ import annotations

import torch

from . import _lib
from .hashgrid import HashGrid
from .input_reader import validate_dem_particles
from .kernels import Workspace
from .simulation import solver_params


class CoupledEngine:
    # fields no step writes (masses, static walls, DEM material): not copied
    # aside by the overflow guard
    _STATIC = {"SPHptl": ("m",), "BNDptl": ("pos", "vel", "m"),
               "DEMptl": ("radius", "rho", "m", "volume", "inertia", "K", "eta", "mu", "h")}

    def __init__(self, solv, P_sph, P_bnd, P_dem=None, P_dem_bnd=None, csr_per_dem: int = 64,
                 grid_slice: int | None = None, fast_math: bool | None = None,
                 instances: dict | None = None, check_every: int = 25,
                 csr_capacity: int | None = None):
        """instances: {"n": scenes, "fluid"/"wall"/"dem"/"dem_wall": int32
        nondecreasing scene number of each particle} for a batched state.
        check_every: steps per guarded window (see the module docstring).
        csr_capacity: initial edges per contact kind (default csr_per_dem
        per sphere, at least 64); the guard grows it when a step needs more."""
        self.solv = solv
        self.P, self.B, self.D, self.W = P_sph, P_bnd, P_dem, P_dem_bnd
        dev, dtype = P_sph.device, P_sph.dtype
        self.device, self.dtype = dev, dtype
        gsl = int(grid_slice or solv.grid_slice)
        self.coupled = bool(solv.dem_enable and P_dem is not None)
        self.gs = HashGrid(gsl, gsl, gsl, device=dev)
        self.gb = HashGrid(gsl, gsl, gsl, device=dev)
        self.instances = instances
        if instances is not None:
            self.gs.instances(instances["fluid"], instances["n"])
            self.gb.instances(instances["wall"], instances["n"])
        self.gs.reserve(P_sph.n, dtype)
        self.gb.build(P_bnd.pos, solv.support, weights=P_bnd.m)
        self.ws = Workspace(dev, dtype)
        nd = P_dem.n if self.coupled else 0
        self.ws.reserve(P_sph.n, P_bnd.n, nd)
        # the wall grid is built once above: walls never move in an engine
        self.ws.static_walls()
        self.scales = None
        if self.coupled:
            self.scales = validate_dem_particles(solv, P_dem, P_dem_bnd)
            self.gd = HashGrid(gsl, gsl, gsl, device=dev)
            self.gw = HashGrid(gsl, gsl, gsl, device=dev)
            if instances is not None:
                self.gd.instances(instances["dem"], instances["n"])
                self.gw.instances(instances["dem_wall"], instances["n"])
            self.gd.reserve(nd, dtype)
            self.gw.build(P_dem_bnd.pos, max(2.0 * self.scales[1], solv.support))
            self.r_dem = max(2.0 * self.scales[0], 2.0 * self.scales[2])
            cap = max(int(csr_per_dem) * nd, 64) if csr_capacity is None else int(csr_capacity)
            if cap < 1:
                raise ValueError("csr_capacity must be positive")
            self._capacity_csr(cap)
        self.params = solver_params(solv, dtype, self.scales)
        if fast_math is not None:
            self.params.fast_math = 1 if fast_math else 0
        self.step_index = 0
        self.check_every = max(1, int(check_every))
        self.regrowths = 0          # capacity increases after an overflow
        self._graphs = None
        self._closed = False
        self._snap = None

    def _capacity_csr(self, cap: int) -> None:
        D = self.D
        for kind in ("dem", "bnd"):
            ids_old = getattr(D, f"contact_{kind}_id_old")
            tang_old = getattr(D, f"tang_{kind}_old")
            E = int(ids_old.shape[0])
            cap_k = max(cap, E)
            new_ids = torch.zeros(cap_k, dtype=torch.int32, device=self.device)
            new_tang = torch.zeros((cap_k, 3), dtype=self.dtype, device=self.device)
            new_ids[:E] = ids_old
            new_tang[:E] = tang_old
            setattr(D, f"contact_{kind}_id_old", new_ids)
            setattr(D, f"tang_{kind}_old", new_tang)
            setattr(D, f"contact_{kind}_id_new", torch.zeros_like(new_ids))
            setattr(D, f"tang_{kind}_new", torch.zeros_like(new_tang))
        self.cap = cap

    # ------------------------------------------------------------------ stepping
    def _enqueue(self, step: int) -> None:
        s = self.solv
        self.gs.build(self.P.pos, s.support, weights=self.P.m)
        st = _lib.stream_handle()
        if self.coupled:
            self.gd.build(self.D.pos, self.r_dem)
            _lib.call("sphdem_coupled_step", self.params, _lib.sph_view(self.P),
                      _lib.bnd_view(self.B), _lib.dem_view(self.D), _lib.dembnd_view(self.W),
                      self.gs.struct(), self.gb.struct(), self.gd.struct(), self.gw.struct(),
                      int(step), -1, self.ws.struct(), st)
        else:
            _lib.call("sphdem_sph_step", self.params, _lib.sph_view(self.P), _lib.bnd_view(self.B),
                      self.gs.struct(), self.gb.struct(), int(step), 1, self.ws.struct(), st)

    def shepard_due(self, step: int) -> bool:
        k = self.solv.shepard_step
        return step == 0 or (k > 0 and step % k == 0)

    def _one(self) -> None:
        if self._graphs is not None:
            self._graphs[self.shepard_due(self.step_index)].replay()
        else:
            self._enqueue(self.step_index)
        self.step_index += 1

    def _live(self) -> None:
        if self._closed:
            raise RuntimeError("publish_csr() has trimmed the contact CSR to its exact size; "
                               "this engine can no longer step")

    # ---------------------------------------------------------- overflow guard
    def _mutable(self):
        """(struct, field) of every tensor a step may write."""
        out = []
        for S in (self.P, self.B, self.D if self.coupled else None):
            if S is None:
                continue
            static = self._STATIC.get(type(S).__name__, ())
            for name in S.field_names():
                t = getattr(S, name)
                if name not in static and t is not None and t.numel():
                    out.append((S, name))
        return out

    def _snapshot(self) -> None:
        keys = self._mutable()
        srcs = [getattr(S, n) for S, n in keys] + [self.ws.flags]
        if self._snap is None or [b.shape for b in self._snap[1]] != [t.shape for t in srcs]:
            self._snap = (keys, [torch.empty_like(t) for t in srcs], 0)
        bufs = self._snap[1]
        torch._foreach_copy_(bufs, srcs)
        self._snap = (keys, bufs, self.step_index)

    def _restore(self) -> None:
        keys, bufs, k = self._snap
        torch._foreach_copy_([getattr(S, n) for S, n in keys] + [self.ws.flags], bufs)
        self.step_index = k

    def _overflowed(self) -> bool:
        return bool(int(self.ws.flags[0].item()) & 1)

    def _grow(self) -> None:
        n = self.D.n
        need = max(int(getattr(self.D, f"contact_{k}_offset_new")[n].item())
                   for k in ("dem", "bnd"))
        self._capacity_csr(max(2 * self.cap, (3 * need) // 2 + 64))
        self.regrowths += 1
        if self._graphs is not None:
            self._graphs = None
            self.capture()

    def _window(self, k: int) -> None:
        self._live()
        if not self.coupled:
            for _ in range(k):
                self._one()
            return
        # bit 0 of the flags is the engine's: it starts every window clear
        if self._overflowed():
            self.ws.flags[0:1].bitwise_and_(~1)
        while True:
            self._snapshot()
            for _ in range(k):
                self._one()
            if not self._overflowed():
                return
            self._restore()
            self._grow()

    def step(self) -> None:
        """One step, guarded: it synchronises to read the overflow flag."""
        self._window(1)

    def run(self, n_steps: int) -> None:
        """n_steps steps in guarded windows of check_every steps."""
        left = int(n_steps)
        while left > 0:
            k = min(left, self.check_every)
            self._window(k)
            left -= k

    def capture(self) -> None:
        """Capture one step with and one without the Shepard refresh.  The
        steps they stand for must run on buffers of the same addresses (the
        engine captures again after a capacity change).  Capturing records
        work without running it; the side stream the DEM contact work runs
        on is created first (sphdem_device_init), so the graphs keep the
        overlap of an eager step."""
        self._live()
        _lib.call("sphdem_device_init", _lib.stream_handle())
        side = torch.cuda.Stream(device=self.device)
        graphs = {}
        for shep in (True, False):
            probe = 0 if shep else 1
            g = torch.cuda.CUDAGraph()
            torch.cuda.synchronize()
            with torch.cuda.stream(side):
                with torch.cuda.graph(g, stream=side):
                    self._enqueue(probe)
            graphs[shep] = g
        torch.cuda.synchronize()
        self._graphs = graphs

    def walls_changed(self) -> None:
        """After editing the wall positions or masses in place: rebuild the
        wall grid and forget the cached wall-only densities (in place, so
        captured graphs stay valid)."""
        self.gb.build(self.B.pos, self.solv.support, weights=self.B.m)
        self.ws.wl_static.fill_(float("nan"))

    def check(self) -> None:
        """Raise OverflowError if a contact row did not fit the capacity.
        The guarded windows redo such steps, so this fires only for work
        enqueued around them (e.g. a caller's own kernels on this state)."""
        if self.coupled and int(self.ws.flags[0].item()) & 1:
            raise OverflowError(f"contact CSR exceeded {self.cap} edges per kind")

    def publish_csr(self) -> None:
        """Trim the read-side CSR to exactly E edges, as SPHDEM_OneStep
        leaves it.  This ends the engine: the trimmed arrays no longer have
        the capacity its kernels write, so step(), run() and capture() raise
        afterwards."""
        if not self.coupled or self._closed:
            self._closed = True
            return
        D = self.D
        n = D.n
        for kind in ("dem", "bnd"):
            E = int(getattr(D, f"contact_{kind}_offset_old")[n].item())
            setattr(D, f"contact_{kind}_id_old", getattr(D, f"contact_{kind}_id_old")[:E].clone())
            setattr(D, f"tang_{kind}_old", getattr(D, f"tang_{kind}_old")[:E].clone())
        self._graphs = None
        self._snap = None
        self._closed = True

    def diagnostics(self) -> dict:
        """FP64 totals: fluid mass and KE, DEM mass and KE (incl. spin), front."""
        out = torch.zeros(5, dtype=torch.float64, device=self.device)
        P, D = self.P, self.D
        nd = D.n if self.coupled else 0
        _lib.call("sphdem_reduce_f64", _lib.precision_of(self.dtype), _lib.ptr(P.pos),
                  _lib.ptr(P.vel), _lib.ptr(P.m), P.n,
                  _lib.ptr(D.vel) if nd else None, _lib.ptr(D.omega) if nd else None,
                  _lib.ptr(D.m) if nd else None, _lib.ptr(D.inertia) if nd else None, nd,
                  out.data_ptr(), _lib.stream_handle())
        v = out.cpu().tolist()
        return {"fluid_mass": v[0], "fluid_ke": v[1], "dem_ke": v[2], "dem_mass": v[3],
                "front": v[4]}
