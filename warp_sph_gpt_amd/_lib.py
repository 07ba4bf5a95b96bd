"""ctypes binding of libsphdem.so (declared in include/sphdem.h).

Structures below mirror the C structs field for field.  Tensors are passed
as raw device pointers; every pointer is checked for device, dtype and
contiguity before a call.  Error codes map to the reference's exception
types: SPHDEM_E_ARG -> ValueError, SPHDEM_E_OVERFLOW -> OverflowError,
SPHDEM_E_HIP -> RuntimeError.  There is no fallback: if the library is not
built, or no GPU is present, calls raise.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int, c_void_p
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "libsphdem.so"
SPHDEM_OK, SPHDEM_E_ARG, SPHDEM_E_HIP, SPHDEM_E_OVERFLOW = 0, 1, 2, 3


class Params(ctypes.Structure):
    _fields_ = [("precision", c_int), ("h", c_double), ("support", c_double),
                ("rho0", c_double), ("c0", c_double), ("gamma", c_double), ("mu", c_double),
                ("g", c_double), ("dt", c_double), ("eps_min", c_double), ("eps_max", c_double),
                ("shepard_step", c_int), ("radius_max", c_double),
                ("boundary_radius_max", c_double), ("h_max", c_double), ("h_uniform", c_double),
                ("fast_math", c_int)]


def _ptr_fields(names):
    return [(n, c_void_p) for n in names]


class SphView(ctypes.Structure):
    _fields_ = _ptr_fields(("pos", "vel", "rho_raw", "rho", "pres", "acc", "m", "flt",
                            "porosity", "pgf", "acc_dem")) + [("n", c_int)]


class BndView(ctypes.Structure):
    _fields_ = _ptr_fields(("pos", "vel", "rho_raw", "rho", "pres", "acc", "m")) + [("n", c_int)]


DEM_REAL = ("pos", "vel", "acc", "force", "omega", "torque", "radius", "rho", "m", "volume",
            "inertia", "K", "eta", "mu", "h", "flt_s", "porosity", "drag", "pressure_force")
DEM_CSR = ("contact_dem_offset_old", "contact_dem_id_old", "tang_dem_old",
           "contact_dem_offset_new", "contact_dem_id_new", "tang_dem_new",
           "contact_bnd_offset_old", "contact_bnd_id_old", "tang_bnd_old",
           "contact_bnd_offset_new", "contact_bnd_id_new", "tang_bnd_new")


class DemView(ctypes.Structure):
    _fields_ = (_ptr_fields(DEM_REAL) + _ptr_fields(DEM_CSR)
                + [("cap_dem_new", c_int), ("cap_bnd_new", c_int), ("n", c_int)])


class DemBndView(ctypes.Structure):
    _fields_ = _ptr_fields(("pos", "vel", "omega", "radius", "rho", "m", "volume", "inertia",
                            "K", "eta", "mu")) + [("n", c_int)]


class GridStruct(ctypes.Structure):
    _fields_ = [("precision", c_int), ("dim_x", c_int), ("dim_y", c_int), ("dim_z", c_int),
                ("cell_width", c_double), ("n", c_int), ("point_cells", c_void_p),
                ("point_ids", c_void_p), ("rank", c_void_p), ("cell_lo", c_void_p),
                ("cell_hi", c_void_p), ("sorted", c_void_p), ("scan_tmp", c_void_p),
                ("n_inst", c_int), ("inst", c_void_p)]


class WorkStruct(ctypes.Structure):
    _fields_ = _ptr_fields(("fl_rho", "fl_pres", "fl_V", "fl_vel", "fl_pgf", "wl_rho",
                            "wl_pres", "wl_vel", "scan_tmp", "flags", "fl_flag", "fl_list",
                            "fl_count", "fl_scan", "dem_near", "wl_static", "fl_box",
                            "fl_sum", "fl_chunk")) + [
        ("cap_f", c_int), ("cap_b", c_int), ("cap_d", c_int)]


NUM_PASSES = 11


EISPH_FIELDS = ("pos", "vel", "vel_star", "vel_bc", "acc", "kgc", "rho", "m", "pres", "Aij",
                "bi", "Aijpj", "divergence", "residual")


class EisphView(ctypes.Structure):
    _fields_ = _ptr_fields(EISPH_FIELDS) + [("mirror", c_void_p), ("n", c_int)]


class EisphParams(ctypes.Structure):
    _fields_ = [("precision", c_int), ("h", c_double), ("support", c_double),
                ("rho0", c_double), ("nu", c_double), ("dt", c_double)]


class EisphNbr(ctypes.Structure):
    _fields_ = [("n", c_int), ("len", c_void_p), ("offset", c_void_p), ("id", c_void_p),
                ("geo", c_void_p), ("cap", c_int), ("width", c_void_p), ("scan_tmp", c_void_p),
                ("part", c_void_p)]


class TuningStruct(ctypes.Structure):
    _fields_ = [("block", c_int * NUM_PASSES), ("xcd_remap", c_int * NUM_PASSES),
                ("por_x_tile", c_int), ("dem_beside", c_int), ("por_parts", c_int),
                ("por_queue", c_int), ("wall_beside", c_int), ("force_beside", c_int)]


_SIGNATURES = {
    "sphdem_version": ([], c_int),
    "sphdem_last_error": ([], ctypes.c_char_p),
    "sphdem_device_count": ([], c_int),
    "sphdem_scan_scratch": ([c_int], c_int),
    "sphdem_tuning_get": ([POINTER(TuningStruct)], c_int),
    "sphdem_mpnn_pack": ([c_void_p, c_int, c_void_p, c_void_p, c_void_p], c_int),
    "sphdem_mpnn_to_bf16": ([c_void_p, c_int, c_void_p, c_void_p], c_int),
    "sphdem_mpnn_dense_fwd": ([c_void_p, c_int, c_void_p, c_int, c_int, c_void_p, c_void_p, c_int,
                               c_void_p, c_void_p, c_void_p, c_void_p], c_int),
    "sphdem_mpnn_dense_bwd_x": ([c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p,
                                 c_int, c_int, c_void_p], c_int),
    "sphdem_mpnn_dense_bwd_w": ([c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int, c_int,
                                 c_void_p, c_void_p, c_void_p], c_int),
    "sphdem_mpnn_edge_fwd": ([POINTER(Params), POINTER(GridStruct), POINTER(GridStruct), c_void_p,
                              c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                              c_void_p, c_void_p, c_void_p, c_void_p], c_int),
    "sphdem_mpnn_edge_bwd": ([POINTER(Params), POINTER(GridStruct), POINTER(GridStruct), c_void_p,
                              c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                              c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                              c_void_p, c_void_p, c_void_p, c_void_p, c_void_p], c_int),
    "sphdem_mpnn_features": ([POINTER(Params), POINTER(SphView), POINTER(GridStruct), c_void_p,
                              c_void_p, c_void_p], c_int),
    "sphdem_mpnn_targets": ([POINTER(GridStruct), c_void_p, c_void_p, c_void_p, c_void_p,
                             c_void_p, c_double, c_void_p, c_void_p], c_int),
    "sphdem_mpnn_explicit_out": ([POINTER(GridStruct), c_void_p, c_void_p, c_void_p], c_int),
    "sphdem_mpnn_loss": ([c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p], c_int),
    "sphdem_mpnn_apply": ([POINTER(Params), POINTER(SphView), POINTER(GridStruct), c_void_p,
                           c_double, c_void_p], c_int),
    "sphdem_surrogate_prep": ([POINTER(Params), POINTER(SphView), POINTER(BndView),
                               POINTER(DemView), POINTER(GridStruct), POINTER(GridStruct),
                               POINTER(GridStruct), c_int, POINTER(WorkStruct), c_void_p], c_int),
    "sphdem_dem_substeps_scratch": ([c_int, c_int], ctypes.c_long),
    "sphdem_dem_substeps": ([POINTER(Params), POINTER(DemView), POINTER(DemBndView),
                             POINTER(GridStruct), POINTER(GridStruct), c_int, c_double, c_void_p,
                             POINTER(WorkStruct), c_void_p], c_int),
    "sphdem_eisph_part_size": ([c_int, c_int], ctypes.c_long),
    "sphdem_eisph_slices": ([c_int], c_int),
    "sphdem_eisph_kgc": ([POINTER(EisphParams), POINTER(EisphView), POINTER(EisphView), POINTER(GridStruct),
                          POINTER(GridStruct), c_void_p], c_int),
    "sphdem_eisph_nbr_count": ([POINTER(EisphParams), POINTER(EisphView), POINTER(EisphView),
                                POINTER(GridStruct), POINTER(GridStruct), POINTER(EisphNbr),
                                c_void_p], c_int),
    "sphdem_eisph_nbr_fill": ([POINTER(EisphParams), POINTER(EisphView), POINTER(EisphView),
                               POINTER(GridStruct), POINTER(GridStruct), POINTER(EisphNbr),
                               c_void_p], c_int),
    "sphdem_eisph_dirichlet": ([c_int, POINTER(EisphView), POINTER(EisphView), c_int, c_void_p],
                               c_int),
    "sphdem_eisph_pressure_boundary": ([c_int, POINTER(EisphView), POINTER(EisphView),
                                        c_void_p], c_int),
    "sphdem_eisph_advec_vis": ([POINTER(EisphParams), POINTER(EisphView), POINTER(EisphView), POINTER(EisphNbr),
                                c_void_p], c_int),
    "sphdem_eisph_build_ppe": ([POINTER(EisphParams), POINTER(EisphView), POINTER(EisphView), POINTER(EisphNbr),
                                c_void_p], c_int),
    "sphdem_eisph_solve_ppe": ([POINTER(EisphParams), POINTER(EisphView), POINTER(EisphNbr), c_void_p], c_int),
    "sphdem_eisph_correct": ([POINTER(EisphParams), POINTER(EisphView), POINTER(EisphView), POINTER(EisphNbr),
                              c_void_p], c_int),
    "sphdem_eisph_diagnostics": ([POINTER(EisphParams), POINTER(EisphView), POINTER(EisphView),
                                  POINTER(EisphNbr), c_void_p], c_int),
    "sphdem_eisph_step": ([POINTER(EisphParams), POINTER(EisphView), POINTER(EisphView), POINTER(EisphNbr), c_int,
                           c_void_p], c_int),
    "sphdem_eisph_steps": ([POINTER(EisphParams), POINTER(EisphView), POINTER(EisphView), POINTER(EisphNbr),
                            c_int, c_int, c_void_p], c_int),
    "sphdem_eisph_one_group": ([c_int, c_int, c_int], c_int),
    "sphdem_format_g6": ([c_void_p, ctypes.c_long, c_int, c_void_p, ctypes.c_long], ctypes.c_long),
    "sphdem_format_int": ([c_void_p, ctypes.c_long, c_int, c_void_p, ctypes.c_long], ctypes.c_long),
    "sphdem_adam": ([c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_double, c_double, c_double,
                     c_double, c_int, c_void_p], c_int),
    "sphdem_tuning_set": ([POINTER(TuningStruct)], c_int),
    "sphdem_exclusive_scan": ([c_void_p, c_int, c_void_p, c_void_p], c_int),
    "sphdem_device_init": ([c_void_p], c_int),
    "sphdem_exclusive_scan_noalloc": ([c_void_p, c_int, c_void_p], c_int),
    "sphdem_grid_build": ([POINTER(GridStruct), c_void_p, c_int, c_double, c_void_p, c_void_p], c_int),
    "sphdem_grid_set_w": ([POINTER(GridStruct), c_void_p, c_void_p], c_int),
    "sphdem_grid_warp_bounds": ([POINTER(GridStruct), c_void_p, c_void_p, c_void_p], c_int),
    "sphdem_neighbour_counts": ([POINTER(GridStruct), POINTER(GridStruct), c_double, c_void_p,
                                 c_void_p], c_int),
    "sphdem_wendland": ([c_int, c_void_p, c_int, c_double, c_void_p, c_void_p, c_void_p], c_int),
    "sphdem_drag_beta": ([c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                          c_void_p, c_void_p], c_int),
    "sphdem_shepard_sph": ([POINTER(Params), POINTER(SphView), POINTER(BndView),
                            POINTER(GridStruct), POINTER(GridStruct), c_void_p], c_int),
    "sphdem_density_sph": ([POINTER(Params), POINTER(SphView), POINTER(BndView),
                            POINTER(GridStruct), POINTER(GridStruct), c_void_p], c_int),
    "sphdem_density_bnd": ([POINTER(Params), POINTER(SphView), POINTER(BndView),
                            POINTER(GridStruct), POINTER(GridStruct), c_void_p], c_int),
    "sphdem_pres": ([POINTER(Params), c_void_p, c_void_p, c_int, c_void_p], c_int),
    "sphdem_force_sph": ([POINTER(Params), POINTER(SphView), POINTER(BndView),
                          POINTER(GridStruct), POINTER(GridStruct), POINTER(WorkStruct),
                          c_void_p], c_int),
    "sphdem_step_sph": ([POINTER(Params), POINTER(SphView), c_void_p], c_int),
    "sphdem_count_dem_contacts": ([POINTER(Params), POINTER(DemView), POINTER(GridStruct),
                                   c_double, c_void_p], c_int),
    "sphdem_force_dem": ([POINTER(Params), POINTER(DemView), POINTER(GridStruct), c_double,
                          c_double, c_void_p, c_void_p], c_int),
    "sphdem_count_bnd_contacts": ([POINTER(Params), POINTER(DemView), POINTER(DemBndView),
                                   POINTER(GridStruct), c_double, c_void_p], c_int),
    "sphdem_bc_dem": ([POINTER(Params), POINTER(DemView), POINTER(DemBndView),
                       POINTER(GridStruct), c_double, c_void_p, c_void_p], c_int),
    "sphdem_step_dem": ([POINTER(Params), POINTER(DemView), c_void_p], c_int),
    "sphdem_prep_sphdem": ([POINTER(Params), POINTER(SphView), POINTER(BndView),
                            POINTER(DemView), POINTER(GridStruct), POINTER(GridStruct),
                            POINTER(GridStruct), c_double, c_double, POINTER(WorkStruct),
                            c_void_p], c_int),
    "sphdem_interaction_dem": ([POINTER(Params), POINTER(SphView), POINTER(DemView),
                                POINTER(GridStruct), POINTER(GridStruct), POINTER(WorkStruct),
                                c_void_p], c_int),
    "sphdem_interaction_sph": ([POINTER(Params), POINTER(SphView), POINTER(DemView),
                                POINTER(GridStruct), POINTER(GridStruct), c_double, c_void_p],
                               c_int),
    "sphdem_sph_step": ([POINTER(Params), POINTER(SphView), POINTER(BndView),
                         POINTER(GridStruct), POINTER(GridStruct), c_int, c_int,
                         POINTER(WorkStruct), c_void_p], c_int),
    "sphdem_coupled_step": ([POINTER(Params), POINTER(SphView), POINTER(BndView),
                             POINTER(DemView), POINTER(DemBndView), POINTER(GridStruct),
                             POINTER(GridStruct), POINTER(GridStruct), POINTER(GridStruct),
                             c_int, c_int, POINTER(WorkStruct), c_void_p], c_int),
    "sphdem_reduce_f64": ([c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                           c_void_p, c_void_p, c_int, c_void_p, c_void_p], c_int),
    "sphdem_profile_enable": ([c_int], c_int),
    "sphdem_profile_read": ([c_void_p, c_void_p, c_int], c_int),
}

PASS_NAMES = ("grid_build", "fluid_density", "wall_density", "fluid_force", "dem_csr",
              "dem_contact", "porosity", "interaction_dem", "fluid_final", "step_dem", "other")
PROFILED_PASSES = PASS_NAMES + ("eisph_advec", "eisph_ppe", "eisph_solve", "eisph_correct",
                                "eisph_finish", "eisph_steps")


class PassTimer:
    """Per-pass HIP-event timing of everything enqueued while active."""

    def __enter__(self):
        torch.cuda.synchronize()
        call("sphdem_profile_enable", 1)
        return self

    def __exit__(self, *exc):
        n = len(PROFILED_PASSES)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int * n)()
        call("sphdem_profile_read", ctypes.addressof(ms), ctypes.addressof(cnt), n)
        call("sphdem_profile_enable", 0)
        self.ms = {PROFILED_PASSES[i]: ms[i] for i in range(n)}
        self.count = {PROFILED_PASSES[i]: cnt[i] for i in range(n)}
        return False

_lib = None


def load() -> ctypes.CDLL:
    """Load libsphdem.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m warp_sph_gpt_amd.build`")
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def exported_symbols() -> tuple:
    return tuple(_SIGNATURES)


def check(rc: int, what: str) -> None:
    if rc == SPHDEM_OK:
        return
    msg = load().sphdem_last_error().decode(errors="replace")
    text = f"{what}: {msg}"
    if rc == SPHDEM_E_ARG:
        raise ValueError(text)
    if rc == SPHDEM_E_OVERFLOW:
        raise OverflowError(text)
    raise RuntimeError(text)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)


def stream_handle():
    return c_void_p(torch.cuda.current_stream().cuda_stream)


def precision_of(dtype: torch.dtype) -> int:
    if dtype == torch.float32:
        return 4
    if dtype == torch.float64:
        return 8
    raise ValueError(f"particle arrays must be float32 or float64, not {dtype}")


def ptr(t, dtype=None, what="array"):
    """Device pointer of a contiguous CUDA tensor (None for missing/empty)."""
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{what} must be a torch tensor")
    if t.numel() == 0:
        return None
    if not t.is_cuda:
        raise ValueError(f"{what} must live on a GPU device (got {t.device}); "
                         "the SPH-DEM kernels run only on the GPU")
    if dtype is not None and t.dtype != dtype:
        raise ValueError(f"{what} must be {dtype}, not {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{what} must be contiguous")
    return t.data_ptr()


def _real_dtype(struct):
    return struct.pos.dtype


def sph_view(P) -> SphView:
    dt = _real_dtype(P)
    v = SphView()
    for name, _ in P.FIELDS:
        setattr(v, name, ptr(getattr(P, name), dt, f"P_sph.{name}"))
    v.n = P.n
    return v


def bnd_view(B) -> BndView:
    dt = _real_dtype(B)
    v = BndView()
    for name, _ in B.FIELDS:
        setattr(v, name, ptr(getattr(B, name), dt, f"P_bnd.{name}"))
    v.n = B.n
    return v


def eisph_view(S, what="P_sph") -> EisphView:
    dt = _real_dtype(S)
    v = EisphView()
    for name in EISPH_FIELDS:
        setattr(v, name, ptr(getattr(S, name), dt, f"{what}.{name}"))
    v.mirror = ptr(S.mirror, torch.int32, f"{what}.mirror")
    v.n = S.n
    return v


def eisph_params(precision: int, h: float, rho0: float, nu: float, dt: float) -> EisphParams:
    return EisphParams(precision, float(h), 2.0 * float(h), float(rho0), float(nu), float(dt))


def dem_view(D) -> DemView:
    dt = _real_dtype(D)
    v = DemView()
    for name in DEM_REAL:
        setattr(v, name, ptr(getattr(D, name), dt, f"P_dem.{name}"))
    for name in DEM_CSR:
        t = getattr(D, name)
        want = dt if name.startswith("tang_") else torch.int32
        setattr(v, name, ptr(t, want, f"P_dem.{name}"))
    v.cap_dem_new = int(D.contact_dem_id_new.shape[0])
    v.cap_bnd_new = int(D.contact_bnd_id_new.shape[0])
    if D.tang_dem_new.shape[0] < v.cap_dem_new:
        v.cap_dem_new = int(D.tang_dem_new.shape[0])
    if D.tang_bnd_new.shape[0] < v.cap_bnd_new:
        v.cap_bnd_new = int(D.tang_bnd_new.shape[0])
    v.n = D.n
    return v


def dembnd_view(W) -> DemBndView:
    dt = _real_dtype(W)
    v = DemBndView()
    for name, _ in W.FIELDS:
        setattr(v, name, ptr(getattr(W, name), dt, f"P_dem_bnd.{name}"))
    v.n = W.n
    return v


_FAST_MATH = [False]


def set_fast_math(flag: bool) -> None:
    """Default arithmetic of the neighbour sums: False evaluates W and dW/dr
    with IEEE division and square root exactly as the reference writes them;
    True uses the hardware square root and reciprocal and multiplies by 1/h
    in the porosity term and the force pass (densities stay exact)."""
    _FAST_MATH[0] = bool(flag)


def fast_math() -> bool:
    return _FAST_MATH[0]


def params(precision: int, **kw) -> Params:
    p = Params()
    p.precision = precision
    p.fast_math = 1 if _FAST_MATH[0] else 0
    for name, value in kw.items():
        setattr(p, name, value)
    return p


# ---------------------------------------------------------------- launch shapes
TUNABLE_PASSES = ("grid_build", "fluid_density", "wall_density", "fluid_force", "porosity",
                  "fluid_final")


def get_tuning() -> dict:
    """The library's current launch shapes, by pass name."""
    t = TuningStruct()
    call("sphdem_tuning_get", t)
    return {"block": {n: int(t.block[i]) for i, n in enumerate(PASS_NAMES)},
            "xcd_remap": {n: int(t.xcd_remap[i]) for i, n in enumerate(PASS_NAMES)},
            "por_x_tile": int(t.por_x_tile), "dem_beside": int(t.dem_beside),
            "por_parts": int(t.por_parts), "por_queue": int(t.por_queue),
            "wall_beside": int(t.wall_beside), "force_beside": int(t.force_beside)}


def set_tuning(cfg: dict) -> None:
    """Set launch shapes; keys missing from `cfg` keep their current value."""
    cur = get_tuning()
    t = TuningStruct()
    for i, n in enumerate(PASS_NAMES):
        t.block[i] = int(cfg.get("block", {}).get(n, cur["block"][n]))
        t.xcd_remap[i] = int(cfg.get("xcd_remap", {}).get(n, cur["xcd_remap"][n]))
    t.por_x_tile = int(cfg.get("por_x_tile", cur["por_x_tile"]))
    t.dem_beside = int(cfg.get("dem_beside", cur["dem_beside"]))
    t.por_parts = int(cfg.get("por_parts", cur["por_parts"]))
    t.por_queue = int(cfg.get("por_queue", cur["por_queue"]))
    t.wall_beside = int(cfg.get("wall_beside", cur["wall_beside"]))
    t.force_beside = int(cfg.get("force_beside", cur["force_beside"]))
    call("sphdem_tuning_set", t)


def fptr(t):
    """Raw device pointer of a contiguous CUDA tensor of any dtype (None
    passes through): for the surrogate's bf16 (int16) and FP32 buffers."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError(f"buffer must live on a GPU device (got {t.device})")
    if not t.is_contiguous():
        raise ValueError("buffer must be contiguous")
    return c_void_p(t.data_ptr())
