"""Parity at the sizes the benchmarks run (the workloads c1, c2, c3 and the
surrogate rollout on c1), through the fused coupled step the engine runs.

* c2 (1,562,500 fluid, 1,080,816 wall, 2,000 DEM): fluid and wall grids
  bit-exact against the oracle's at full size; then two coupled steps (the
  Shepard refresh at step 0 and a held-filter step), each checked pass by
  pass against the oracle.  The fluid and wall neighbour passes are checked
  on every 97th sorted subject (the oracle's stride mode), each pass fed
  the GPU's outputs of the passes before it (densities and pressures for
  the force and coupling sums, porosity and -grad p for the DEM
  interpolation, DEM drag for the reaction); the DEM contacts, the CSR and
  the DEM interpolation on every DEM; the integration on every particle.
  Bar: 1e-5 relative per field (pressure against 1e-2 of the Tait
  stiffness), CSR offsets and ids bit-exact, in both math modes.
* c1 (100,000 fluid, 180,456 wall, 200 DEM): the engine against the
  oracle's own step loop (main.py:84-91): every field after 3 steps at
  5e-5, and after 100 steps the fluid kinetic energy and the surge front
  within 0.5 % and positions within 1e-3.
* c3 (64 dam breaks, 5.67 M fluid, 11.5 M wall): 25 batched steps of the
  whole ensemble; several instances (the smallest, the largest, some with
  and without DEM) bit-equal to their solo runs with the reference's
  arithmetic, and within 5e-5 after 5 steps with the fast forms.
* the surrogate rollout on c1 for 0.5 s simulated (1,000 calls, judged
  every 50): a healthy end state, fluid and DEM mass conserved, and kinetic
  energy and surge front within the stated bounds of the plain run's.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from helpers import assert_rel, np_, rel_err, to_oracle
from oracle import oracle as O
from warp_sph_gpt_amd import _lib
from warp_sph_gpt_amd.engine import CoupledEngine
from warp_sph_gpt_amd.hashgrid import HashGrid
from warp_sph_gpt_amd.workloads import c1, c2

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
TOL = 1.0e-5
STRIDE = 97


@pytest.fixture(scope="module")
def c2_workload():
    return c2(device=DEV)


def test_c2_grids_bit_exact_at_full_size(c2_workload):
    wl = c2_workload
    P, B = wl.structs[:2]
    gsl = wl.solv.grid_slice
    for pts in (np_(P.pos), np_(B.pos)):
        og = O.OracleGrid(gsl, gsl, gsl).build(pts, wl.solv.support)
        gg = HashGrid(gsl, gsl, gsl, device=DEV).build(torch.tensor(pts, device=DEV),
                                                       wl.solv.support)
        np.testing.assert_array_equal(np_(gg.point_ids), og.point_ids)
        np.testing.assert_array_equal(np_(gg.point_cells),
                                      og.point_cells[np.argsort(og.point_ids)])
        s, e = gg.warp_bounds()
        np.testing.assert_array_equal(np_(s), og.cell_starts)
        np.testing.assert_array_equal(np_(e), og.cell_ends)
        # the folded x axis: 195 cells onto 64 buckets at dx = 0.004
        assert og.cell_starts.size == gsl ** 3


def _ostate(eng):
    return [O.OStruct.from_struct(s) for s in (eng.P, eng.B, eng.D, eng.W)]


def _host(t):
    return np_(t) if isinstance(t, torch.Tensor) else np.asarray(t)


def _gpu(t):
    return _host(t).astype(np.float32)


def _as64(o):
    return O.OStruct(**{k: (v.astype(np.float64) if isinstance(v, np.ndarray) and
                            v.dtype.kind == "f" else v) for k, v in o.__dict__.items()})


def _check_step_passes(solv, before, eng, step):
    """The step just run by `eng` from the state `before`, pass by pass."""
    P0, B0, D0, W0 = before
    rmax, rbmax, hmax, hu = O.dem_search_scales(D0, W0)
    gsl = solv.grid_slice
    og = [O.OracleGrid(gsl, gsl, gsl).build(s.pos, r)
          for s, r in ((P0, solv.support), (B0, solv.support),
                       (D0, max(2 * rmax, 2 * hmax)), (W0, max(2 * rbmax, solv.support)))]
    gP, gB, gD = eng.P, eng.B, eng.D
    sf = og[0].point_ids[::STRIDE]
    sb = og[1].point_ids[::STRIDE]
    stiff = solv.rho0 * solv.c0 ** 2 / solv.gamma
    shepard = step == 0 or step % solv.shepard_step == 0

    def cmp(ref, got, rows, what, floor=0.0):
        return assert_rel(_gpu(got)[rows], np.asarray(ref)[rows], TOL, floor,
                          f"step {step} {what}")

    O.set_stride(STRIDE)
    try:
        # densities and EOS
        P, B = P0.copy(), B0.copy()
        if shepard:
            O.Kernel_shepard_sph(P, B, og[0], og[1], solv.support, solv.h)
            cmp(P.flt, gP.flt, sf, "fluid flt")
        O.Kernel_density_sph(P, B, og[0], og[1], solv.support, solv.h)
        O.Kernel_density_bnd(P, B, og[0], og[1], solv.support, solv.h)
        O.Kernel_pres_sph(P, solv.rho0, solv.c0, solv.gamma)
        O.Kernel_pres_bnd(B, solv.rho0, solv.c0, solv.gamma)
        cmp(P.rho, gP.rho, sf, "fluid rho")
        cmp(B.rho, gB.rho, sb, "wall rho")
        cmp(P.pres, gP.pres, sf, "fluid pres", 1e-2 * stiff)
        cmp(B.pres, gB.pres, sb, "wall pres", 1e-2 * stiff)
        # the force and coupling sums from the GPU's densities and pressures
        Pf, Bf = P0.copy(), B0.copy()
        for name in ("rho", "pres", "flt"):
            setattr(Pf, name, _gpu(getattr(gP, name)))
        for name in ("rho", "pres"):
            setattr(Bf, name, _gpu(getattr(gB, name)))
        O.Kernel_force_sph(Pf, Bf, og[0], og[1], solv.support, solv.h, solv.mu, solv.g)
        O.Kernel_prep_sphdem(Pf, Bf, D0, og[0], og[1], og[2], 2 * hmax, hu, solv.support,
                             solv.h, solv.dem_porosity_min, solv.dem_porosity_max)
        cmp(Pf.porosity, gP.porosity, sf, "fluid porosity")
        cmp(Pf.pgf, gP.pgf, sf, "fluid -grad p")
        assert float(Pf.porosity[sf].min()) < 1.0 - 1e-6, "no sampled subject near a DEM"
        # the drag reaction from the GPU's DEM drag and interpolation
        Dr = D0.copy()
        for name in ("drag", "flt_s"):
            setattr(Dr, name, _gpu(getattr(gD, name)))
        Pf.porosity = _gpu(gP.porosity)
        O.Kernel_interaction_sph(Pf, Dr, og[0], og[2], 2 * hmax)
        cmp(Pf.acc_dem, gP.acc_dem, sf, "fluid acc_dem")
        cmp(Pf.acc, gP.acc, sf, "fluid acc")
    finally:
        O.set_stride(1)
    # DEM contacts and the CSR, every DEM, from the step's initial state
    D = D0.copy()
    O.run_force_dem(D, og[2], rmax, solv.g, solv.dt)
    O.run_bc_dem(D, W0, og[3], rbmax, solv.dt)
    for kind in ("dem", "bnd"):
        off = getattr(D, f"contact_{kind}_offset_old")
        E = int(off[-1])
        np.testing.assert_array_equal(_host(getattr(gD, f"contact_{kind}_offset_old")), off)
        np.testing.assert_array_equal(_host(getattr(gD, f"contact_{kind}_id_old"))[:E],
                                      getattr(D, f"contact_{kind}_id_old"))
        assert_rel(_gpu(getattr(gD, f"tang_{kind}_old"))[:E], getattr(D, f"tang_{kind}_old"),
                   TOL, what=f"step {step} tang_{kind}")
    # the DEM interpolation, every DEM, from the GPU's fluid fields, in FP32
    # and in FP64 (the same oracle code).  A wet sphere's sums run over ~2.7e5
    # fluid particles at this resolution, where the reference's sequential
    # FP32 sum itself drifts by ~1.2e-5 from the exact value (flt_s, and the
    # drag and pressure force divided by it): the GPU's sums (per-wave
    # partials, then a fixed tree) are held to 1e-5 of the FP64 evaluation
    # of the same code, and to 3e-5 of the sequential FP32 sums
    Pi = P0.copy()
    for name in ("rho", "pgf", "porosity"):
        setattr(Pi, name, _gpu(getattr(gP, name)))
    D64 = _as64(D)
    O.Kernel_interaction_dem(Pi, D, og[0], og[2], solv.mu, solv.dem_porosity_min,
                             solv.dem_porosity_max, solv.dt)
    g64 = [O.OracleGrid(gsl, gsl, gsl, np.float64).build(x.pos.astype(np.float64), r)
           for x, r in ((P0, solv.support), (D0, max(2 * rmax, 2 * hmax)))]
    O.Kernel_interaction_dem(_as64(Pi), D64, g64[0], g64[1], solv.mu, solv.dem_porosity_min,
                             solv.dem_porosity_max, solv.dt)
    every = np.arange(D.n)
    for name in ("force", "torque", "drag", "pressure_force", "flt_s", "porosity"):
        got = _gpu(getattr(gD, name)).astype(np.float64)
        e64, e32 = rel_err(got, getattr(D64, name)), rel_err(getattr(D, name), getattr(D64, name))
        print(f"step {step} dem {name}: hip vs fp64 {e64:.2e}, oracle fp32 vs fp64 {e32:.2e}")
        assert e64 <= TOL, f"step {step} dem {name} vs FP64: {e64:.3e}"
        assert_rel(got, getattr(D, name), 3 * TOL, what=f"step {step} dem {name} vs FP32")
    assert float(np.abs(D.flt_s).max()) > 0, "no wet DEM"
    # integration, every particle, from the GPU's accelerations and loads
    Ps = P0.copy()
    Ps.acc = _gpu(gP.acc)
    O.Kernel_step_sph(Ps, solv.dt)
    allf = np.arange(Ps.n)
    cmp(Ps.pos, gP.pos, allf, "fluid pos")
    cmp(Ps.vel, gP.vel, allf, "fluid vel")
    Ds = D0.copy()
    Ds.force, Ds.torque = _gpu(gD.force), _gpu(gD.torque)
    O.Kernel_step_dem(Ds, solv.dt)
    for name in ("pos", "vel", "omega"):
        cmp(getattr(Ds, name), getattr(gD, name), every, f"dem {name}")


@pytest.mark.parametrize("fast_math", [False, True], ids=["exact", "fast"])
def test_c2_coupled_steps_pass_by_pass(c2_workload, fast_math):
    wl = c2_workload
    eng = CoupledEngine(wl.solv, *(s.to(DEV) for s in wl.structs), fast_math=fast_math)
    assert eng.P.n == 1_562_500 and eng.B.n == 1_080_816 and eng.D.n == 2000
    for step in range(2):
        before = _ostate(eng)
        eng.step()
        torch.cuda.synchronize()
        _check_step_passes(wl.solv, before, eng, step)


# ---------------------------------------------------------------- c1


@pytest.fixture(scope="module")
def c1_oracle():
    """The oracle's step loop on c1: states after 3 and after 100 steps."""
    wl = c1(device=DEV)
    run = O.Run(wl.solv, *to_oracle(wl.structs))
    states = {}
    for k in range(1, 101):
        run.step()
        if k in (3, 100):
            states[k] = [s.copy() for s in (run.P, run.B, run.D)]
    return wl, states


def _totals(P):
    m, v = np.asarray(P.m, np.float64), np.asarray(P.vel, np.float64)
    return {"fluid_ke": 0.5 * float((m * (v * v).sum(1)).sum()),
            "front": float(np.asarray(P.pos, np.float64)[:, 0].max())}


@pytest.mark.parametrize("fast_math", [False, True], ids=["exact", "fast"])
def test_c1_engine_follows_the_oracle_loop(c1_oracle, fast_math):
    wl, states = c1_oracle
    eng = CoupledEngine(wl.solv, *(s.to(DEV) for s in wl.structs), fast_math=fast_math)
    assert eng.P.n == 100_000 and eng.D.n == 200
    eng.run(3)
    P, B, D = states[3]
    stiff = wl.solv.rho0 * wl.solv.c0 ** 2 / wl.solv.gamma
    for name in ("pos", "vel", "rho", "acc", "flt", "porosity", "pgf", "acc_dem"):
        assert_rel(_gpu(getattr(eng.P, name)), getattr(P, name), 5 * TOL, what=f"fluid {name}")
    # after a few steps the pressures differ by what a few ulp of density
    # move them (one ulp of rho near rho0, 6.1e-5 kg/m^3, is gamma B ulp/rho0
    # = 0.06 Pa): the floor 5 % of B allows 0.35 Pa, about six ulp
    assert_rel(_gpu(eng.P.pres), P.pres, 5 * TOL, 5e-2 * stiff, "fluid pres")
    assert_rel(_gpu(eng.B.rho), B.rho, 5 * TOL, what="wall rho")
    for name in ("pos", "vel", "omega", "force", "torque", "drag", "pressure_force", "flt_s"):
        assert_rel(_gpu(getattr(eng.D, name)), getattr(D, name), 5 * TOL, what=f"dem {name}")
    eng.run(97)
    eng.check()
    P = states[100][0]
    ref, got = _totals(P), eng.diagnostics()
    assert ref["fluid_ke"] > 0
    for k in ("fluid_ke", "front"):
        assert abs(got[k] - ref[k]) <= 5e-3 * abs(ref[k]), (k, got[k], ref[k])
    assert_rel(_gpu(eng.P.pos), P.pos, 1e-3, what="pos after 100 steps")


# ---------------------------------------------------------------- c3


@pytest.fixture(scope="module")
def c3_ensemble():
    from warp_sph_gpt_amd.workloads import c3
    return c3(device=DEV)


def _pick_instances(wl):
    sizes = [sc[0].n for sc in wl.scenes]
    dem = [sc[2] is not None for sc in wl.scenes]
    pick = {int(np.argmin(sizes)), int(np.argmax(sizes))}
    pick.add(next(k for k, d in enumerate(dem) if d))
    pick.add(next((k for k, d in enumerate(dem) if not d), 1))
    pick.add(len(sizes) - 1)
    return sorted(pick)


@pytest.mark.parametrize("fast_math", [False, True], ids=["exact", "fast"])
def test_c3_instances_match_their_solo_runs(c3_ensemble, fast_math):
    from test_gpu_ensemble import DEM as DEM_FIELDS
    from test_gpu_ensemble import FLUID, COUPLED, compare
    from warp_sph_gpt_amd.ensemble import EnsembleEngine

    wl = c3_ensemble
    steps = 5 if fast_math else 25
    ens = EnsembleEngine(wl.solv, wl.scenes, device=DEV, fast_math=fast_math)
    assert ens.n == 64 and ens.P.n == wl.n_fluid
    ens.capture()
    ens.run(steps)
    ens.check()
    for k in _pick_instances(wl):
        P, B, D, W = wl.scenes[k]
        alone = CoupledEngine(wl.solv, P.to(DEV), B.to(DEV), None if D is None else D.to(DEV),
                              None if W is None else W.to(DEV), fast_math=fast_math)
        alone.run(steps)
        alone.check()
        if not fast_math:
            compare(ens, k, alone)
            continue
        Pk, Bk, Dk, _ = ens.instance_state(k)
        for f in FLUID + (COUPLED if Dk is not None else ()):
            assert_rel(_gpu(getattr(Pk, f)), _gpu(getattr(alone.P, f)), 5 * TOL,
                       1e-2 * wl.solv.rho0 * wl.solv.c0 ** 2 / 7 if f == "pres" else 0.0,
                       f"instance {k} fluid {f}")
        if Dk is not None:
            for f in DEM_FIELDS:
                assert_rel(_gpu(getattr(Dk, f)), _gpu(getattr(alone.D, f)), 5 * TOL,
                           what=f"instance {k} dem {f}")


# ---------------------------------------------------------------- c4


def test_c4_surrogate_rollout_half_a_second():
    """The judged surrogate rollout of the benchmark (c1 scene, judge every
    50 calls at 1 % density) over 0.5 s simulated: it ends healthy, keeps the
    fluid and DEM mass, and stays within 2 % of the plain run's surge front
    and 10 % of its fluid kinetic energy."""
    import os
    from warp_sph_gpt_amd.surrogate import MPNN, WEIGHTS, load_params, rollout_c1
    assert os.path.exists(WEIGHTS)
    net = MPNN(params=load_params(WEIGHTS))
    r = rollout_c1(net, sim_time=0.5, judge_every=50, density_tol=0.01)
    print({k: r[k] for k in ("speedup", "surrogate_calls", "plain_calls", "judged", "rejected",
                             "health_end", "health_plain_end", "final_surrogate",
                             "final_plain")})
    assert r["calls"] == 1000
    assert r["surrogate_calls"] + r["plain_calls"] >= r["calls"]
    fs, fp = r["final_surrogate"], r["final_plain"]
    hs, hp = r["health_end"], r["health_plain_end"]
    assert hs["inside"] and math.isfinite(hs["vmax"]) and hs["vmax"] < 10.44
    # no worse than the plain run's own end state by more than 0.5 % density
    assert hs["density_rms"] <= hp["density_rms"] + 5e-3
    for k in ("fluid_mass", "dem_mass"):
        assert fs[k] == pytest.approx(fp[k], rel=1e-12), k
    assert math.isfinite(fs["fluid_ke"]) and fs["fluid_ke"] > 0
    assert abs(fs["front"] - fp["front"]) <= 0.02 * fp["front"]
    assert abs(fs["fluid_ke"] - fp["fluid_ke"]) <= 0.10 * fp["fluid_ke"]
