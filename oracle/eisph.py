"""CPU oracle of the Eulerian ISPH lid-driven cavity -- TEST INFRASTRUCTURE ONLY.

Only tests/ and the cpu_baseline leg of bench.py use this module, as the
checker; the product path never imports it.

The per-point loops are the C restatement in eisph_impl.h (liboracle.so,
FP32 and FP64): every neighbour sum walks the HashGrid query, as the
reference's kernels do (kernel/KERNEL_EISPH_*.py).  This module restates the
orchestration of source/EISPH.py:26-141 (EISPH_OneStep, run_eisph) around
them on numpy arrays.  The right-hand-side sum of EISPH.py:72 is taken in
FP64 in index order and rounded to the particle precision in residual[0].

How the oracle is pinned is listed in eisph_impl.h; tests/test_eisph_oracle.py
holds the checks.
"""
from __future__ import annotations

import ctypes
from ctypes import c_int, c_void_p

import numpy as np

from .oracle import OStruct, OracleGrid, _call, _p, lib

EISPH_REAL = ("pos", "vel", "vel_star", "vel_bc", "acc", "kgc", "rho", "m", "pres", "Aij",
              "bi", "Aijpj", "divergence", "residual")
_WIDTH = {"pos": 3, "vel": 3, "vel_star": 3, "vel_bc": 3, "acc": 3, "kgc": 4}

_VIEW: dict = {}


def _view_cls(L):
    if L.sfx not in _VIEW:
        class OEisph(ctypes.Structure):
            _fields_ = [(n, c_void_p) for n in EISPH_REAL] + [("mirror", c_void_p),
                                                              ("n", c_int)]
        _VIEW[L.sfx] = OEisph
    return _VIEW[L.sfx]


def _view(L, S):
    v = _view_cls(L)()
    for name in EISPH_REAL:
        a = getattr(S, name)
        if a.dtype != L.np or not a.flags.c_contiguous:
            a = np.ascontiguousarray(a, dtype=L.np)
            setattr(S, name, a)
        setattr(v, name, _p(a))
    if S.mirror.dtype != np.int32 or not S.mirror.flags.c_contiguous:
        S.mirror = np.ascontiguousarray(S.mirror, dtype=np.int32)
    v.mirror = _p(S.mirror)
    v.n = S.n
    return v


def eisph_struct(pos, vel=None, vel_bc=None, rho=None, m=None, mirror=None,
                 dtype=np.float32) -> OStruct:
    """An EISPHptl (input/struct_eisph.py:4-28) of numpy arrays: computed
    fields zero, kgc zero, vel_star = vel, mirror = arange(n) unless given."""
    pos = np.asarray(pos, dtype=dtype).reshape(-1, 3)
    n = len(pos)
    out = {}
    for name in EISPH_REAL:
        w = _WIDTH.get(name)
        out[name] = np.zeros((n, w) if w else n, dtype=dtype)
    out["pos"] = pos.copy()
    if vel is not None:
        out["vel"] = np.asarray(vel, dtype=dtype).reshape(n, 3).copy()
    out["vel_star"] = out["vel"].copy()
    if vel_bc is not None:
        out["vel_bc"] = np.asarray(vel_bc, dtype=dtype).reshape(n, 3).copy()
    if rho is not None:
        out["rho"] = np.broadcast_to(np.asarray(rho, dtype=dtype), (n,)).copy()
    if m is not None:
        out["m"] = np.broadcast_to(np.asarray(m, dtype=dtype), (n,)).copy()
    out["mirror"] = (np.arange(n, dtype=np.int32) if mirror is None
                     else np.asarray(mirror, dtype=np.int32).copy())
    return OStruct(**out)


def from_struct(s, dtype=np.float32) -> OStruct:
    """From a torch-backed EISPH struct (or any object with the fields)."""
    out = {}
    for name in EISPH_REAL + ("mirror",):
        t = getattr(s, name)
        a = t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)
        if name == "kgc":
            a = a.reshape(-1, 4)
        out[name] = np.ascontiguousarray(a, dtype=np.int32 if name == "mirror" else dtype)
    return OStruct(**out)


def _dt(S):
    return S.pos.dtype


# ------------------------------------------------------------------ kernels

def wendland2(r, h, dtype=np.float32):
    """kernel/KERNEL_EISPH_KNL.py:14-49 at radii r: (W, dW/dr)."""
    L = lib(dtype)
    fw, fd = L.fn("wendland2_w"), L.fn("wendland2_dw")
    fw.restype = fd.restype = L.real
    fw.argtypes = fd.argtypes = [L.real, L.real]
    r = np.asarray(r, dtype=L.np)
    return (np.array([fw(x, h) for x in r], dtype=L.np),
            np.array([fd(x, h) for x in r], dtype=L.np))


def Kernel_prepare_kgc(P, B, gs, gb, support, h):
    L = lib(_dt(P))
    _call("ora_eisph_kgc", _dt(P), _view(L, P), _view(L, B), gs.struct(), gb.struct(),
          float(support), float(h))


def Kernel_Dirichlet_BC(P, B, predicted):
    L = lib(_dt(P))
    _call("ora_eisph_dirichlet", _dt(P), _view(L, P), _view(L, B), int(predicted))


def Kernel_pressure_boundary_eisph(P, B):
    L = lib(_dt(P))
    _call("ora_eisph_pressure_boundary", _dt(P), _view(L, P), _view(L, B))


def Kernel_advec_vis_eisph(P, B, gs, gb, support, h, nu, dt):
    L = lib(_dt(P))
    _call("ora_eisph_advec_vis", _dt(P), _view(L, P), _view(L, B), gs.struct(), gb.struct(),
          float(support), float(h), float(nu), float(dt))


def Kernel_build_ppe_eisph(P, B, gs, gb, support, h, rho0, dt):
    L = lib(_dt(P))
    _call("ora_eisph_build_ppe", _dt(P), _view(L, P), _view(L, B), gs.struct(), gb.struct(),
          float(support), float(h), float(rho0), float(dt))


def remove_bi_mean(P):
    """EISPH.py:70-73: residual[0] = sum(bi), then Kernel_remove_bi_mean."""
    L = lib(_dt(P))
    _call("ora_eisph_remove_bi_mean", _dt(P), _view(L, P))


def Kernel_solve_ppe_eisph(P):
    L = lib(_dt(P))
    _call("ora_eisph_solve_ppe", _dt(P), _view(L, P))


def Kernel_correct_eisph(P, B, gs, gb, support, h, dt):
    L = lib(_dt(P))
    _call("ora_eisph_correct", _dt(P), _view(L, P), _view(L, B), gs.struct(), gb.struct(),
          float(support), float(h), float(dt))


def Kernel_diagnostics_eisph(P, B, gs, gb, support, h):
    L = lib(_dt(P))
    _call("ora_eisph_diagnostics", _dt(P), _view(L, P), _view(L, B), gs.struct(), gb.struct(),
          float(support), float(h))


# ------------------------------------------------------------------ driver

def build_grids(solv, P, B, dims=None):
    """EISPH.py:117-122: both fixed grids (grid_slice^3 buckets, or `dims`)
    and the KGC."""
    dtype = _dt(P)
    g = solv.grid_slice
    dims = (g, g, g) if dims is None else dims
    gs = OracleGrid(*dims, dtype).build(P.pos, solv.support)
    gb = OracleGrid(*dims, dtype).build(B.pos, solv.support)
    Kernel_prepare_kgc(P, B, gs, gb, solv.support, solv.h)
    return gs, gb


def EISPH_OneStep(solv, P, B, gs, gb, diagnostics: bool = True) -> None:
    """source/EISPH.py:26-90, one C call."""
    L = lib(_dt(P))
    _call("ora_eisph_step", _dt(P), _view(L, P), _view(L, B), gs.struct(), gb.struct(),
          float(solv.support), float(solv.h), float(solv.rho0), float(solv.nu),
          float(solv.dt), int(bool(diagnostics)))


def run_eisph(solv, P, B, progress=None):
    """source/EISPH.py:93-141: fixed grids and KGC once, then n_steps steps;
    diagnostics and a host velocity frame at output steps and the last step.
    Returns (positions, velocity frames, frame times)."""
    gs, gb = build_grids(solv, P, B)
    positions = P.pos.copy()
    frames, times = [P.vel.copy()], [0.0]
    for step in range(1, solv.n_steps + 1):
        store = (solv.output_step > 0 and step % solv.output_step == 0) or step == solv.n_steps
        EISPH_OneStep(solv, P, B, gs, gb, diagnostics=store)
        if store:
            frames.append(P.vel.copy())
            times.append(step * solv.dt)
            if progress is not None:
                progress(step, solv.n_steps)
    return positions, frames, times
