"""Particle containers: structure-of-arrays with the reference's field names.

Mirrors input/struct.py:3-104 (SPHptl, BNDptl, DEMptl, DEMBNDptl) and
input/struct_eisph.py:4-28 (EISPHptl).  Each field is one torch tensor: vec3
fields are (N, 3), scalars (N,), 2x2 matrices (N, 2, 2) row-major, point
indices int32 (N,), contact offsets int32 (N+1,), contact ids int32 (E,),
tangential histories (E, 3).
The tensors are plain buffers; the HIP kernels read and write them in place
through raw device pointers.
"""
from __future__ import annotations

import torch

VEC3, SCALAR, OFFSET, IDS, TANG = "vec3", "scalar", "offset", "ids", "tang"
MAT22, INDEX = "mat22", "index"


class ParticleStruct:
    """Base of the particle kinds; FIELDS lists (name, layout)."""

    FIELDS: tuple = ()

    def __init__(self, **values):
        for name, _ in self.FIELDS:
            setattr(self, name, values.get(name))

    @classmethod
    def field_names(cls) -> tuple:
        return tuple(name for name, _ in cls.FIELDS)

    @classmethod
    def allocate(cls, n: int, device="cpu", dtype=torch.float32):
        """Zero-initialised struct of n particles with empty contact CSR."""
        values = {}
        for name, layout in cls.FIELDS:
            if layout == VEC3:
                values[name] = torch.zeros((n, 3), dtype=dtype, device=device)
            elif layout == SCALAR:
                values[name] = torch.zeros(n, dtype=dtype, device=device)
            elif layout == MAT22:
                values[name] = torch.zeros((n, 2, 2), dtype=dtype, device=device)
            elif layout == INDEX:
                values[name] = torch.zeros(n, dtype=torch.int32, device=device)
            elif layout == OFFSET:
                values[name] = torch.zeros(n + 1, dtype=torch.int32, device=device)
            elif layout == IDS:
                values[name] = torch.zeros(0, dtype=torch.int32, device=device)
            elif layout == TANG:
                values[name] = torch.zeros((0, 3), dtype=dtype, device=device)
        return cls(**values)

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    @property
    def device(self) -> torch.device:
        return self.pos.device

    @property
    def dtype(self) -> torch.dtype:
        return self.pos.dtype

    def to(self, device=None, dtype=None):
        """Copy of this struct on another device and/or real type."""
        out = {}
        for name, layout in self.FIELDS:
            t = getattr(self, name)
            if t is None:
                out[name] = None
                continue
            kw = {}
            if device is not None:
                kw["device"] = device
            if dtype is not None and t.is_floating_point():
                kw["dtype"] = dtype
            out[name] = t.to(**kw).clone() if kw else t.clone()
        return type(self)(**out)

    def clone(self):
        return self.to()


class SPHptl(ParticleStruct):
    """Fluid particles (input/struct.py:3-21)."""

    FIELDS = (("pos", VEC3), ("vel", VEC3), ("rho_raw", SCALAR), ("rho", SCALAR),
              ("pres", SCALAR), ("acc", VEC3), ("m", SCALAR), ("flt", SCALAR),
              ("porosity", SCALAR), ("pgf", VEC3), ("acc_dem", VEC3))


class BNDptl(ParticleStruct):
    """Dummy wall particles (input/struct.py:23-36)."""

    FIELDS = (("pos", VEC3), ("vel", VEC3), ("rho_raw", SCALAR), ("rho", SCALAR),
              ("pres", SCALAR), ("acc", VEC3), ("m", SCALAR))


class DEMptl(ParticleStruct):
    """Moving DEM spheres with old/new contact CSR (input/struct.py:39-82)."""

    FIELDS = (("pos", VEC3), ("vel", VEC3), ("acc", VEC3), ("force", VEC3), ("omega", VEC3),
              ("torque", VEC3), ("radius", SCALAR), ("rho", SCALAR), ("m", SCALAR),
              ("volume", SCALAR), ("inertia", SCALAR), ("K", SCALAR), ("eta", SCALAR),
              ("mu", SCALAR), ("h", SCALAR),
              ("contact_dem_offset_old", OFFSET), ("contact_dem_id_old", IDS),
              ("tang_dem_old", TANG),
              ("contact_dem_offset_new", OFFSET), ("contact_dem_id_new", IDS),
              ("tang_dem_new", TANG),
              ("contact_bnd_offset_old", OFFSET), ("contact_bnd_id_old", IDS),
              ("tang_bnd_old", TANG),
              ("contact_bnd_offset_new", OFFSET), ("contact_bnd_id_new", IDS),
              ("tang_bnd_new", TANG),
              ("flt_s", SCALAR), ("porosity", SCALAR), ("drag", VEC3),
              ("pressure_force", VEC3))


class DEMBNDptl(ParticleStruct):
    """Fixed DEM wall spheres: geometry and material only (input/struct.py:85-104)."""

    FIELDS = (("pos", VEC3), ("vel", VEC3), ("omega", VEC3), ("radius", SCALAR),
              ("rho", SCALAR), ("m", SCALAR), ("volume", SCALAR), ("inertia", SCALAR),
              ("K", SCALAR), ("eta", SCALAR), ("mu", SCALAR))


class EISPHptl(ParticleStruct):
    """Fixed Eulerian ISPH points, fluid or ghost (input/struct_eisph.py:4-28).

    One layout for both sets: `mirror` is the stable fluid row a ghost
    reflects (an identity map on the fluid set), `residual[0]` doubles as the
    right-hand-side sum after the pressure assembly, as in the reference."""

    FIELDS = (("pos", VEC3), ("vel", VEC3), ("vel_star", VEC3), ("vel_bc", VEC3),
              ("acc", VEC3), ("kgc", MAT22), ("rho", SCALAR), ("m", SCALAR),
              ("pres", SCALAR), ("Aij", SCALAR), ("bi", SCALAR), ("Aijpj", SCALAR),
              ("divergence", SCALAR), ("residual", SCALAR), ("mirror", INDEX))
