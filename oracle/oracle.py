"""CPU oracle of the coupled WCSPH-DEM step -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
use this module, always as the checker and never as the product path.

The heavy loops are the C restatement in oracle_impl.h (liboracle.so, FP32
and FP64); this module restates the orchestration of
source/Simulation.py:17-221 and the grid radii of main.py:66-87 around
them, with numpy arrays.  The CSR lifecycle follows the reference exactly:
count -> exclusive scan -> read E -> allocate exactly E -> fill -> swap.

Parity pins (see tests/test_oracle.py):
  * Wendland: closed form at sample radii, unit integral, FD slope, ends.
  * hash grid: stable-permutation / partition invariants, and complete
    neighbour sets against an O(N^2) brute force.  Bucket keys and order
    restate Warp's hash grid; against Warp itself they are parity UNPINNED.
  * SPH sums: O(N^2) brute force in FP64; lattice density, hydrostatic and
    gravity-at-rest invariants.
  * DEM/CSR/coupling: the closed-form cases of the reference's
    tests/test_dem.py, tests/test_3d.py and tests/test_input_reader.py.
"""
from __future__ import annotations

import ctypes
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

SPH_FIELDS = ("pos", "vel", "rho_raw", "rho", "pres", "acc", "m", "flt", "porosity", "pgf",
              "acc_dem")
BND_FIELDS = ("pos", "vel", "rho_raw", "rho", "pres", "acc", "m")
DEM_REAL = ("pos", "vel", "acc", "force", "omega", "torque", "radius", "rho", "m", "volume",
            "inertia", "K", "eta", "mu", "h", "flt_s", "porosity", "drag", "pressure_force")
DEM_CSR = ("contact_dem_offset_old", "contact_dem_id_old", "tang_dem_old",
           "contact_dem_offset_new", "contact_dem_id_new", "tang_dem_new",
           "contact_bnd_offset_old", "contact_bnd_id_old", "tang_bnd_old",
           "contact_bnd_offset_new", "contact_bnd_id_new", "tang_bnd_new")
WALL_FIELDS = ("pos", "vel", "omega", "radius", "rho", "m", "volume", "inertia", "K", "eta",
               "mu")
VEC_FIELDS = {"pos", "vel", "acc", "pgf", "acc_dem", "force", "omega", "torque", "drag",
              "pressure_force"}


def build() -> Path:
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


_LIBS: dict = {}


def _structs(real):
    def rs(names):
        return [(n, c_void_p) for n in names]

    class OSph(ctypes.Structure):
        _fields_ = rs(SPH_FIELDS) + [("n", c_int)]

    class OBnd(ctypes.Structure):
        _fields_ = rs(BND_FIELDS) + [("n", c_int)]

    class ODem(ctypes.Structure):
        _fields_ = rs(DEM_REAL) + rs(DEM_CSR) + [("n", c_int)]

    class OWall(ctypes.Structure):
        _fields_ = rs(WALL_FIELDS) + [("n", c_int)]

    class OGrid(ctypes.Structure):
        _fields_ = [("dim_x", c_int), ("dim_y", c_int), ("dim_z", c_int),
                    ("cell_width_inv", real), ("n", c_int), ("point_ids", c_void_p),
                    ("cell_starts", c_void_p), ("cell_ends", c_void_p)]

    return OSph, OBnd, ODem, OWall, OGrid


class _Lib:
    def __init__(self, suffix, real, npdt):
        if not LIB.exists():
            build()
        self.c = ctypes.CDLL(str(LIB))
        self.sfx, self.real, self.np = suffix, real, npdt
        self.OSph, self.OBnd, self.ODem, self.OWall, self.OGrid = _structs(real)

    def fn(self, name):
        return getattr(self.c, name + self.sfx)


def lib(dtype=np.float32) -> _Lib:
    dtype = np.dtype(dtype)
    if dtype not in _LIBS:
        _LIBS[dtype] = (_Lib("_f32", c_float, np.float32) if dtype == np.float32
                        else _Lib("_f64", c_double, np.float64))
    return _LIBS[dtype]


def num_threads() -> int:
    return int(lib().c.ora_num_threads())


def set_num_threads(n: int) -> None:
    """Threads of the oracle's OpenMP loops from here on."""
    lib().c.ora_set_num_threads(int(n))


class OStruct:
    """numpy arrays under the reference's field names."""

    def __init__(self, **fields):
        self.__dict__.update(fields)

    @property
    def n(self):
        return int(self.pos.shape[0])

    def copy(self):
        return OStruct(**{k: (v.copy() if isinstance(v, np.ndarray) else v)
                          for k, v in self.__dict__.items()})

    @classmethod
    def from_struct(cls, s, dtype=np.float32):
        """From a torch-backed particle struct (or any object with the fields)."""
        out = {}
        for name, _ in s.FIELDS:
            t = getattr(s, name)
            a = t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)
            if a.dtype.kind == "f":
                a = a.astype(dtype)
            out[name] = np.ascontiguousarray(a)
        return cls(**out)


def _p(a):
    return None if a is None or a.size == 0 else a.ctypes.data


def _view(L, S, kind):
    if kind == "sph":
        v = L.OSph()
        names = SPH_FIELDS
    elif kind == "bnd":
        v = L.OBnd()
        names = BND_FIELDS
    elif kind == "dem":
        v = L.ODem()
        names = DEM_REAL + DEM_CSR
    else:
        v = L.OWall()
        names = WALL_FIELDS
    for name in names:
        a = getattr(S, name)
        want = np.int32 if name.startswith("contact_") else L.np
        if a.dtype != want or not a.flags.c_contiguous:
            a = np.ascontiguousarray(a, dtype=want)
            setattr(S, name, a)
        setattr(v, name, _p(a))
    v.n = S.n
    return v


class OracleGrid:
    """Warp hash grid restated (oracle_impl.h ora_grid_build)."""

    def __init__(self, dim_x, dim_y, dim_z, dtype=np.float32):
        self.dims = (int(dim_x), int(dim_y), int(dim_z))
        self.dtype = np.dtype(dtype)
        ncell = self.dims[0] * self.dims[1] * self.dims[2]
        self.cell_starts = np.zeros(ncell, np.int32)
        self.cell_ends = np.zeros(ncell, np.int32)
        self.point_ids = np.zeros(0, np.int32)
        self.point_cells = np.zeros(0, np.int32)
        self.inv = 1.0
        self.n = 0

    @property
    def id(self):
        return self

    def build(self, points, radius):
        L = lib(self.dtype)
        pts = np.ascontiguousarray(points, dtype=L.np)
        n = len(pts)
        self.point_ids = np.zeros(n, np.int32)
        self.point_cells = np.zeros(n, np.int32)
        inv = L.real()
        L.fn("ora_grid_build")(c_int(self.dims[0]), c_int(self.dims[1]), c_int(self.dims[2]),
                               L.real(radius), c_void_p(_p(pts)), c_int(n), ctypes.byref(inv),
                               c_void_p(_p(self.point_cells)), c_void_p(_p(self.point_ids)),
                               c_void_p(self.cell_starts.ctypes.data),
                               c_void_p(self.cell_ends.ctypes.data))
        self.inv = inv.value
        self.n = n
        return self

    def struct(self):
        L = lib(self.dtype)
        g = L.OGrid()
        g.dim_x, g.dim_y, g.dim_z = self.dims
        g.cell_width_inv = self.inv
        g.n = self.n
        g.point_ids = _p(self.point_ids)
        g.cell_starts = self.cell_starts.ctypes.data
        g.cell_ends = self.cell_ends.ctypes.data
        return g

    @property
    def sorted_cells(self):
        return self.point_cells


def _dt(S):
    return S.pos.dtype


def _call(name, dtype, *args):
    L = lib(dtype)
    f = L.fn(name)
    conv = []
    for a in args:
        if isinstance(a, ctypes.Structure):
            conv.append(ctypes.byref(a))
        elif isinstance(a, float):
            conv.append(L.real(a))
        elif isinstance(a, (int, np.integer)):
            conv.append(c_int(int(a)))
        else:
            conv.append(a)
    return f(*conv)


# ------------------------------------------------------------------ kernels


def wendland(r, h, dtype=np.float32):
    L = lib(dtype)
    fw, fd = L.fn("wendland_w"), L.fn("wendland_dw")
    fw.restype = fd.restype = L.real
    fw.argtypes = fd.argtypes = [L.real, L.real]
    r = np.asarray(r, dtype=L.np)
    return (np.array([fw(x, h) for x in r], dtype=L.np),
            np.array([fd(x, h) for x in r], dtype=L.np))


def drag_beta(eps, rho, mu, d, speed, dtype=np.float32):
    L = lib(dtype)
    f = L.fn("ora_drag_beta")
    f.restype = L.real
    f.argtypes = [L.real] * 5
    return L.np(f(eps, rho, mu, d, speed))


def neighbour_counts(gs, spos, gn, npos, radius):
    dt = gs.dtype
    L = lib(dt)
    counts = np.zeros(gs.n, np.int32)
    sp = np.ascontiguousarray(spos, dtype=L.np)
    nps = np.ascontiguousarray(npos, dtype=L.np)
    _call("ora_neighbour_counts", dt, gs.struct(), c_void_p(_p(sp)), gn.struct(),
          c_void_p(_p(nps)), float(radius), c_void_p(_p(counts)))
    return counts


def Kernel_shepard_sph(P, B, gs, gb, support, h):
    L = lib(_dt(P))
    _call("ora_shepard_sph", _dt(P), _view(L, P, "sph"), _view(L, B, "bnd"), gs.struct(),
          gb.struct(), float(support), float(h))


def Kernel_density_sph(P, B, gs, gb, support, h):
    L = lib(_dt(P))
    _call("ora_density_sph", _dt(P), _view(L, P, "sph"), _view(L, B, "bnd"), gs.struct(),
          gb.struct(), float(support), float(h))


def Kernel_density_bnd(P, B, gs, gb, support, h):
    L = lib(_dt(P))
    _call("ora_density_bnd", _dt(P), _view(L, P, "sph"), _view(L, B, "bnd"), gs.struct(),
          gb.struct(), float(support), float(h))


def _pres(S, rho0, c0, gamma):
    dt = _dt(S)
    _call("ora_pres", dt, c_void_p(_p(S.rho)), c_void_p(_p(S.pres)), S.n, float(rho0),
          float(c0), float(gamma))


Kernel_pres_sph = _pres
Kernel_pres_bnd = _pres


def Kernel_force_sph(P, B, gs, gb, support, h, mu, g):
    L = lib(_dt(P))
    _call("ora_force_sph", _dt(P), _view(L, P, "sph"), _view(L, B, "bnd"), gs.struct(),
          gb.struct(), float(support), float(h), float(mu), float(g))


def Kernel_step_sph(P, dt):
    L = lib(_dt(P))
    _call("ora_step_sph", _dt(P), _view(L, P, "sph"), float(dt))


def Kernel_count_dem_contacts(D, gd, radius_max):
    L = lib(_dt(D))
    _call("ora_count_dem_contacts", _dt(D), _view(L, D, "dem"), gd.struct(), float(radius_max))


def Kernel_force_dem(D, gd, radius_max, g, dt):
    L = lib(_dt(D))
    _call("ora_force_dem", _dt(D), _view(L, D, "dem"), gd.struct(), float(radius_max),
          float(g), float(dt))


def Kernel_count_bnd_contacts(D, W, gw, radius_max):
    L = lib(_dt(D))
    _call("ora_count_bnd_contacts", _dt(D), _view(L, D, "dem"), _view(L, W, "wall"),
          gw.struct(), float(radius_max))


def Kernel_bc_dem(D, W, gw, radius_max, dt):
    L = lib(_dt(D))
    _call("ora_bc_dem", _dt(D), _view(L, D, "dem"), _view(L, W, "wall"), gw.struct(),
          float(radius_max), float(dt))


def Kernel_step_dem(D, dt):
    L = lib(_dt(D))
    _call("ora_step_dem", _dt(D), _view(L, D, "dem"), float(dt))


def Kernel_prep_sphdem(P, B, D, gs, gb, gd, support, h, sph_support, sph_h, eps_min, eps_max,
                       skip_fluid_sum=1):
    L = lib(_dt(P))
    _call("ora_prep_sphdem", _dt(P), _view(L, P, "sph"), _view(L, B, "bnd"), _view(L, D, "dem"),
          gs.struct(), gb.struct(), gd.struct(), float(support), float(h), float(sph_support),
          float(sph_h), float(eps_min), float(eps_max), int(skip_fluid_sum))


def Kernel_interaction_dem(P, D, gs, gd, mu, eps_min, eps_max, dt):
    L = lib(_dt(P))
    _call("ora_interaction_dem", _dt(P), _view(L, P, "sph"), _view(L, D, "dem"), gs.struct(),
          gd.struct(), float(mu), float(eps_min), float(eps_max), float(dt))


def Kernel_interaction_sph(P, D, gs, gd, support):
    L = lib(_dt(P))
    _call("ora_interaction_sph", _dt(P), _view(L, P, "sph"), _view(L, D, "dem"), gs.struct(),
          gd.struct(), float(support))


def array_scan_exclusive(a):
    _call("ora_exclusive_scan", np.float32, c_void_p(_p(a)), len(a))


# ------------------------------------------------------------------ steps


def dem_search_scales(D, W):
    """input/input_reader.py:163-168."""
    h = D.h
    return (float(D.radius.max()), float(W.radius.max()) if W.n else 0.0, float(h.max()),
            float(h[0]) if np.all(h == h[0]) else 0.0)


def SPH_OneStep(solv, P, B, gs, gb, step, integrate=True):
    """source/Simulation.py:17-99."""
    use_shepard = step == 0 or (solv.shepard_step > 0 and step % solv.shepard_step == 0)
    if use_shepard:
        Kernel_shepard_sph(P, B, gs, gb, solv.support, solv.h)
    Kernel_density_sph(P, B, gs, gb, solv.support, solv.h)
    Kernel_density_bnd(P, B, gs, gb, solv.support, solv.h)
    Kernel_pres_sph(P, solv.rho0, solv.c0, solv.gamma)
    Kernel_pres_bnd(B, solv.rho0, solv.c0, solv.gamma)
    Kernel_force_sph(P, B, gs, gb, solv.support, solv.h, solv.mu, solv.g)
    if integrate:
        Kernel_step_sph(P, solv.dt)


def _swap(D, kind):
    for a, b in ((f"contact_{kind}_offset_old", f"contact_{kind}_offset_new"),
                 (f"contact_{kind}_id_old", f"contact_{kind}_id_new"),
                 (f"tang_{kind}_old", f"tang_{kind}_new")):
        x, y = getattr(D, a), getattr(D, b)
        setattr(D, a, y)
        setattr(D, b, x)


def _alloc(D, kind):
    E = int(getattr(D, f"contact_{kind}_offset_new")[D.n])
    if E < 0:
        raise OverflowError("contact count exceeded int32 CSR capacity")
    setattr(D, f"contact_{kind}_id_new", np.zeros(E, np.int32))
    setattr(D, f"tang_{kind}_new", np.zeros((E, 3), _dt(D)))
    return E


def run_force_dem(D, gd, radius_max, g, dt):
    """Simulation.py:142-171: zero, count, scan, read E, allocate, fill, swap."""
    D.contact_dem_offset_new[:] = 0
    Kernel_count_dem_contacts(D, gd, radius_max)
    array_scan_exclusive(D.contact_dem_offset_new)
    E = _alloc(D, "dem")
    Kernel_force_dem(D, gd, radius_max, g, dt)
    _swap(D, "dem")
    return E


def run_bc_dem(D, W, gw, radius_max, dt):
    """Simulation.py:175-203."""
    D.contact_bnd_offset_new[:] = 0
    Kernel_count_bnd_contacts(D, W, gw, radius_max)
    array_scan_exclusive(D.contact_bnd_offset_new)
    E = _alloc(D, "bnd")
    Kernel_bc_dem(D, W, gw, radius_max, dt)
    _swap(D, "bnd")
    return E


def SPHDEM_OneStep(solv, P, B, D, W, gs, gb, gd, gw, step, search_scales=None,
                   skip_fluid_sum=1):
    """source/Simulation.py:102-221."""
    rmax, rbmax, hmax, huni = dem_search_scales(D, W) if search_scales is None else search_scales
    SPH_OneStep(solv, P, B, gs, gb, step, integrate=False)
    run_force_dem(D, gd, rmax, solv.g, solv.dt)
    run_bc_dem(D, W, gw, rbmax, solv.dt)
    Kernel_prep_sphdem(P, B, D, gs, gb, gd, 2.0 * hmax, huni, solv.support, solv.h,
                       solv.dem_porosity_min, solv.dem_porosity_max, skip_fluid_sum)
    Kernel_interaction_dem(P, D, gs, gd, solv.mu, solv.dem_porosity_min, solv.dem_porosity_max,
                           solv.dt)
    Kernel_interaction_sph(P, D, gs, gd, 2.0 * hmax)
    Kernel_step_sph(P, solv.dt)
    Kernel_step_dem(D, solv.dt)


class Run:
    """The step loop of main.py:38-98 (without output): fixed grids once,
    moving grids every step, with the reference's grid radii."""

    def __init__(self, solv, P, B, D=None, W=None, grid_slice=None):
        self.solv = solv
        self.P, self.B, self.D, self.W = P, B, D, W
        gsl = solv.grid_slice if grid_slice is None else grid_slice
        dt = _dt(P)
        self.gs, self.gb = OracleGrid(gsl, gsl, gsl, dt), OracleGrid(gsl, gsl, gsl, dt)
        self.gb.build(B.pos, solv.support)
        self.coupled = D is not None
        if self.coupled:
            self.scales = dem_search_scales(D, W)
            self.gd, self.gw = OracleGrid(gsl, gsl, gsl, dt), OracleGrid(gsl, gsl, gsl, dt)
            self.gw.build(W.pos, max(2.0 * self.scales[1], solv.support))
        self.step_index = 0

    def step(self):
        s = self.solv
        self.gs.build(self.P.pos, s.support)
        if self.coupled:
            self.gd.build(self.D.pos, max(2.0 * self.scales[0], 2.0 * self.scales[2]))
            SPHDEM_OneStep(s, self.P, self.B, self.D, self.W, self.gs, self.gb, self.gd,
                           self.gw, self.step_index, self.scales)
        else:
            SPH_OneStep(s, self.P, self.B, self.gs, self.gb, self.step_index)
        self.step_index += 1


def set_stride(k: int) -> None:
    """Compute only every k-th fluid/wall subject (timing samples only)."""
    lib().c.ora_set_stride(int(k))
