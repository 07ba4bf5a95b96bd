"""Multi-step runs of the no-round-trip engine against the CPU oracle, the
FP64 totals, and the FP64 shadow run of the same kernels.

Bars:
  * 100-step fluid kinetic energy and surge front (max x of the fluid) of the
    FP32 engine within 0.5 % of the FP32 oracle's, on the reference's own case
    (shipped input files, dt = 1e-4, 100 steps) and on a wet scene with the
    DEM inside the fluid and in contact;
  * FP64 totals equal numpy's FP64 sums of the same state (1e-12 relative);
  * the FP64 instantiation of the kernels matches the FP64 oracle over
    coupled steps (1e-10 relative per field);
  * every field of the FP32 run is reported against the FP64 run.
"""
from __future__ import annotations

import json

import numpy as np
import pytest
import torch

from helpers import COMPACT, assert_rel, generation, np_, particles, solv_for, to_oracle
from oracle import oracle as O
from warp_sph_gpt_amd.engine import CoupledEngine
from warp_sph_gpt_amd.shadow import DEM_FIELDS, SPH_FIELDS, ShadowRun
from warp_sph_gpt_amd.struct import BNDptl
from warp_sph_gpt_amd.workloads import c0

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def totals(P, D):
    """FP64 mass, kinetic energy and front of a state held in numpy arrays."""
    m, v = np.asarray(P.m, np.float64), np.asarray(P.vel, np.float64)
    out = {"fluid_mass": m.sum(), "fluid_ke": 0.5 * (m * (v * v).sum(1)).sum(),
           "front": float(np.asarray(P.pos, np.float64)[:, 0].max())}
    if D is not None:
        md, vd = np.asarray(D.m, np.float64), np.asarray(D.vel, np.float64)
        w, I = np.asarray(D.omega, np.float64), np.asarray(D.inertia, np.float64)
        out["dem_mass"] = md.sum()
        out["dem_ke"] = 0.5 * (md * (vd * vd).sum(1)).sum() + 0.5 * (I * (w * w).sum(1)).sum()
    return out


def wet_structs(seed=4):
    """The C0 tank with the DEM lattice lowered into the fluid, one sphere
    resting on the floor spheres, and seeded velocities."""
    gen = generation()
    P, B, D, W = particles(gen)
    rng = np.random.default_rng(seed)
    dpos = np_(D.pos).astype(np.float64)
    dpos[:, 2] -= 0.32
    dpos[19] = [0.45, 0.2, 0.02]
    D.pos = torch.tensor(dpos, dtype=torch.float32)
    D.vel = torch.tensor(rng.normal(0, 0.05, (D.n, 3)), dtype=torch.float32)
    P.vel = torch.tensor(rng.normal(0, 0.05, (P.n, 3)), dtype=torch.float32)
    return gen, (P, B, D, W)


def rollout_pair(solv, structs, steps, fast_math):
    o = to_oracle(structs)
    run = O.Run(solv, *o)
    eng = CoupledEngine(solv, *(s.to(DEV) for s in structs), fast_math=fast_math)
    for _ in range(steps):
        run.step()
    eng.run(steps)
    eng.check()
    return run, eng


@pytest.mark.parametrize("fast_math", [False, True], ids=["exact", "fast"])
@pytest.mark.parametrize("case", ["c0", "wet"])
def test_100_steps_ke_and_front_within_half_percent(case, fast_math):
    if case == "c0":
        wl = c0(device=DEV)
        solv, structs = wl.solv, wl.structs
    else:
        gen, structs = wet_structs()
        solv = solv_for(gen)
    run, eng = rollout_pair(solv, structs, 100, fast_math)
    ref = totals(run.P, run.D)
    got = eng.diagnostics()
    print(json.dumps({"case": case, "fast_math": fast_math, "oracle": ref, "hip": got}))
    assert ref["fluid_ke"] > 0
    for k in ("fluid_ke", "front", "dem_ke"):
        assert abs(got[k] - ref[k]) <= 5e-3 * abs(ref[k]), (k, got[k], ref[k])
    assert got["fluid_mass"] == pytest.approx(ref["fluid_mass"], rel=1e-12)
    # positions stay close particle by particle, not only in the totals
    assert_rel(np_(eng.P.pos), run.P.pos, 1e-3, what="pos after 100 steps")


def test_fp64_totals_match_numpy_sums():
    gen, structs = wet_structs(seed=11)
    P, B, D, W = structs
    rng = np.random.default_rng(12)
    D.omega = torch.tensor(rng.normal(0, 2.0, (D.n, 3)), dtype=torch.float32)
    eng = CoupledEngine(solv_for(gen), *(s.to(DEV) for s in structs))
    got = eng.diagnostics()
    ref = totals(*(to_oracle([P, D])))
    for k, v in ref.items():
        assert got[k] == pytest.approx(v, rel=1e-12), k
    # the totals of a state built from a closed form: unit masses moving at (3,4,0)
    P2 = P.clone()
    P2.m = torch.ones(P.n)
    P2.vel = torch.tensor([[3.0, 4.0, 0.0]]).repeat(P.n, 1)
    eng2 = CoupledEngine(solv_for(gen), P2.to(DEV), B.to(DEV))
    d = eng2.diagnostics()
    assert d["fluid_mass"] == P.n and d["fluid_ke"] == 12.5 * P.n
    assert d["front"] == pytest.approx(float(np_(P.pos)[:, 0].max()), rel=0, abs=0)


def test_fp64_kernels_match_fp64_oracle():
    gen = generation(**dict(COMPACT, dem_origin_z=0.3))
    structs = particles(gen)
    rng = np.random.default_rng(6)
    structs[0].vel = torch.tensor(rng.normal(0, 0.05, (structs[0].n, 3)), dtype=torch.float32)
    solv = solv_for(gen)
    o = to_oracle(structs, np.float64)
    run = O.Run(solv, *o)
    eng = CoupledEngine(solv, *(s.to(DEV, torch.float64) for s in structs), fast_math=False)
    for _ in range(6):
        run.step()
    eng.run(6)
    eng.check()
    for f in ("pos", "vel", "rho", "pres", "acc", "porosity", "pgf", "acc_dem"):
        assert_rel(np_(getattr(eng.P, f)), getattr(run.P, f), 1e-10, 1e-6, what=f"fluid {f}")
    for f in ("pos", "vel", "omega", "force", "torque", "flt_s", "drag", "pressure_force"):
        assert_rel(np_(getattr(eng.D, f)), getattr(run.D, f), 1e-10, 1e-12, what=f"dem {f}")


def test_shadow_reports_every_field():
    gen, structs = wet_structs(seed=8)
    sh = ShadowRun(solv_for(gen), *(s.to(DEV) for s in structs), fast_math=True)
    sh.run(20)
    rep = sh.report()
    print(json.dumps(rep))
    assert set(rep["fluid"]) == set(SPH_FIELDS) and set(rep["dem"]) == set(DEM_FIELDS)
    for group in ("fluid", "wall", "dem"):
        for f, e in rep[group].items():
            assert np.isfinite(e), (group, f)
    # after 20 steps the FP32 run is still close to FP64 in the totals
    assert rep["totals_rel"]["fluid_ke"] < 5e-3
    assert rep["totals_rel"]["front"] < 1e-5
    assert rep["fluid"]["pos"] < 1e-5


@pytest.mark.parametrize("fast_math", [False, True], ids=["exact", "fast"])
def test_dem_near_flags_skip_only_dry_dems(fast_math):
    """The porosity mark's per-DEM flags let the DEM interpolation skip DEMs
    with no fluid within 2 h.  DEMs inside the fluid, DEMs far above it (whose
    search boxes fold onto fluid buckets) and one exactly on the 2 h edge:
    five coupled steps with the flags and without them leave every fluid and
    DEM field equal bit for bit, and the flags all zero after each step."""
    gen, structs = wet_structs(seed=6)
    P, B, D, W = structs
    dpos = np_(D.pos).astype(np.float64)
    dpos[::2, 2] += 1.1   # every other sphere well above the free surface
    top = float(np_(P.pos)[:, 2].max())
    dpos[1] = [0.2, 0.2, top + 2.0 * float(np_(D.h)[1])]
    D.pos = torch.tensor(dpos, dtype=torch.float32)
    solv = solv_for(gen)
    flagged = CoupledEngine(solv, *(s.to(DEV) for s in structs), fast_math=fast_math)
    plain = CoupledEngine(solv, *(s.to(DEV) for s in structs), fast_math=fast_math)
    plain.ws.cap_d = 0   # the library then runs every DEM's search
    assert flagged.ws.cap_d >= D.n
    for _ in range(5):
        flagged.step()
        plain.step()
        torch.cuda.synchronize()
        assert int(flagged.ws.dem_near.abs().sum()) == 0
    for name in ("pos", "vel", "acc", "rho", "pres", "porosity"):
        assert torch.equal(getattr(flagged.P, name), getattr(plain.P, name)), name
    for name in ("pos", "vel", "omega", "force", "torque", "flt_s", "porosity", "drag",
                 "pressure_force"):
        assert torch.equal(getattr(flagged.D, name), getattr(plain.D, name)), name
    dry = np_(flagged.D.flt_s) == 0
    assert dry.any() and not dry.all()


def test_static_wall_densities_change_nothing():
    """Wall particles with no fluid in their support take their density from
    the engine's cache of wall-only sums: 40 steps of the shipped dam break
    with the cache and without it leave every field equal bit for bit, wall
    densities and pressures included; walls_changed() forgets the cache."""
    wl = c0(device=DEV)
    cached = CoupledEngine(wl.solv, *[s.to(DEV).clone() for s in wl.structs])
    plain = CoupledEngine(wl.solv, *[s.to(DEV).clone() for s in wl.structs])
    plain.ws.wl_static = None
    for _ in range(40):
        cached.step()
        plain.step()
    torch.cuda.synchronize()
    for name in ("pos", "vel", "acc", "rho", "pres", "flt"):
        assert torch.equal(getattr(cached.P, name), getattr(plain.P, name)), name
    for name in ("rho", "pres"):
        assert torch.equal(getattr(cached.B, name), getattr(plain.B, name)), f"wall {name}"
    for name in ("pos", "vel", "omega", "force", "drag"):
        assert torch.equal(getattr(cached.D, name), getattr(plain.D, name)), name
    known = torch.isfinite(cached.ws.wl_static[:cached.B.n])
    assert 0 < int(known.sum()) < cached.B.n
    cached.walls_changed()
    assert not bool(torch.isfinite(cached.ws.wl_static).any())


def test_fluid_box_skip_changes_nothing():
    """On wall sets of 4M particles or more the wall pass skips the fluid
    query of a wall particle farther than the support from the fluid's
    bounding box (that sum is +0).  The shipped dam break with 4.2M extra
    wall particles on a lattice above the tank: two coupled steps with the
    skip and without it leave every field equal bit for bit."""
    wl = c0(device=DEV)
    P, B, D, W = [s.to(DEV) for s in wl.structs]
    n = 162
    ax = torch.arange(n, dtype=torch.float32) * 0.02
    gx, gy, gz = torch.meshgrid(ax - 1.0, ax - 1.0, ax[:161] + 1.5, indexing="ij")
    extra = torch.stack([gx.reshape(-1), gy.reshape(-1), gz.reshape(-1)], 1).to(DEV)
    m = extra.shape[0]
    B2 = BNDptl.allocate(B.n + m, device=DEV)
    for name in ("pos", "vel", "rho_raw", "rho", "pres", "acc", "m"):
        getattr(B2, name)[:B.n] = getattr(B, name)
    B2.pos[B.n:] = extra
    B2.rho_raw[B.n:] = B.rho_raw[0]
    B2.rho[B.n:] = B.rho[0]
    B2.m[B.n:] = B.m[0]
    assert B2.n >= 1 << 22
    skip = CoupledEngine(wl.solv, *[s.clone() for s in (P, B2, D, W)])
    plain = CoupledEngine(wl.solv, *[s.clone() for s in (P, B2, D, W)])
    plain.ws.fl_box = None
    for _ in range(2):
        skip.step()
        plain.step()
    torch.cuda.synchronize()
    for name in ("pos", "vel", "acc", "rho", "pres", "flt"):
        assert torch.equal(getattr(skip.P, name), getattr(plain.P, name)), name
    for name in ("rho", "pres"):
        assert torch.equal(getattr(skip.B, name), getattr(plain.B, name)), f"wall {name}"
    for name in ("pos", "vel", "omega", "force", "drag"):
        assert torch.equal(getattr(skip.D, name), getattr(plain.D, name)), name
    # the far lattice has wall neighbours but no fluid term
    assert float(skip.B.rho[B.n:].min()) > 0.0
