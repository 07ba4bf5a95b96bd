"""The engine's contact-capacity contract.

The reference reads each CSR total back and raises OverflowError before it
fills the rows (source/Simulation.py:151-154, 184-187).  The engine keeps
the totals on the device: a step whose rows do not fit its buffers sets a
flag, the guarded window it belongs to is put back and run again with a
larger capacity, so the run ends exactly where the reference's exact-E
lifecycle (SPHDEM_OneStep) ends.  publish_csr() trims the CSR to E edges
and ends the engine."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import np_, solv_for
from test_gpu_rollout import wet_structs
from warp_sph_gpt_amd.engine import CoupledEngine
from warp_sph_gpt_amd.hashgrid import HashGrid
from warp_sph_gpt_amd.input_reader import dem_search_scales
from warp_sph_gpt_amd.simulation import SPHDEM_OneStep

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
FLUID = ("pos", "vel", "rho", "pres", "acc", "flt", "porosity", "pgf", "acc_dem")
DEM = ("pos", "vel", "omega", "force", "torque", "flt_s", "porosity", "drag", "pressure_force")


def contact_scene(seed=7):
    """The wet scene (DEM inside the fluid, one sphere on the floor spheres)
    with two overlapping pairs, so both contact kinds have rows from step 0."""
    gen, (P, B, D, W) = wet_structs(seed=seed)
    D.pos[1] = D.pos[0] + torch.tensor([0.049, 0.0, 0.0])
    D.pos[2] = D.pos[0] + torch.tensor([0.0, 0.049, 0.0])
    return solv_for(gen), (P, B, D, W)


def csr_rows(D, kind):
    off = np_(getattr(D, f"contact_{kind}_offset_old")).astype(np.int64)
    E = int(off[-1])
    return off, np_(getattr(D, f"contact_{kind}_id_old"))[:E], \
        np_(getattr(D, f"tang_{kind}_old"))[:E]


def assert_same_state(a, b, what):
    for f in FLUID:
        np.testing.assert_array_equal(np_(getattr(a.P, f)), np_(getattr(b.P, f)),
                                      err_msg=f"{what}: fluid {f}")
    for f in DEM:
        np.testing.assert_array_equal(np_(getattr(a.D, f)), np_(getattr(b.D, f)),
                                      err_msg=f"{what}: dem {f}")
    for kind in ("dem", "bnd"):
        for x, y in zip(csr_rows(a.D, kind), csr_rows(b.D, kind)):
            np.testing.assert_array_equal(x, y, err_msg=f"{what}: {kind} CSR")


@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_overflow_regrows_and_redoes_the_window(graph):
    solv, st = contact_scene()
    roomy = CoupledEngine(solv, *(s.to(DEV) for s in st))
    tiny = CoupledEngine(solv, *(s.to(DEV) for s in st), csr_capacity=1, check_every=7)
    assert tiny.cap == 1
    if graph:
        tiny.capture()
        roomy.capture()
    roomy.run(30)
    tiny.run(30)
    tiny.check()
    assert tiny.regrowths >= 1 and tiny.cap > 1
    assert tiny.step_index == roomy.step_index == 30
    for kind in ("dem", "bnd"):
        assert int(getattr(tiny.D, f"contact_{kind}_offset_old")[-1]) > 1
    assert_same_state(tiny, roomy, "capacity 1 vs default capacity")
    assert int(tiny.ws.flags[0]) & 1 == 0


def test_overflowing_engine_ends_where_the_exact_e_steps_end():
    """Capacity 1 and single guarded steps against the reference lifecycle:
    SPHDEM_OneStep, which reads E back and allocates exactly E every step."""
    solv, st = contact_scene(seed=9)
    tiny = CoupledEngine(solv, *(s.to(DEV) for s in st), csr_capacity=1)
    P, B, D, W = (s.to(DEV) for s in st)
    scales = dem_search_scales(D, W)
    radii = (solv.support, solv.support, max(2 * scales[0], 2 * scales[2]),
             max(2 * scales[1], solv.support))
    g = [HashGrid(64, 64, 64, device=DEV) for _ in range(4)]
    g[1].build(B.pos, radii[1], weights=B.m)
    g[3].build(W.pos, radii[3])
    for k in range(12):
        tiny.step()
        g[0].build(P.pos, radii[0], weights=P.m)
        g[2].build(D.pos, radii[2])
        SPHDEM_OneStep(solv, P, B, D, W, *g, k, scales)
    torch.cuda.synchronize()
    assert tiny.regrowths >= 1

    class Ref:
        pass
    ref = Ref()
    ref.P, ref.D = P, D
    assert_same_state(tiny, ref, "engine with regrowth vs SPHDEM_OneStep")


def test_publish_csr_trims_to_e_and_ends_the_engine():
    solv, st = contact_scene(seed=3)
    eng = CoupledEngine(solv, *(s.to(DEV) for s in st))
    eng.run(3)
    n = eng.D.n
    E = {k: int(getattr(eng.D, f"contact_{k}_offset_old")[n]) for k in ("dem", "bnd")}
    eng.publish_csr()
    for k in ("dem", "bnd"):
        assert E[k] > 0
        assert getattr(eng.D, f"contact_{k}_id_old").shape == (E[k],)
        assert getattr(eng.D, f"tang_{k}_old").shape == (E[k], 3)
    for call in (eng.step, lambda: eng.run(1), eng.capture):
        with pytest.raises(RuntimeError, match="publish_csr"):
            call()


@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_wall_pass_beside_the_fluid_density_gives_the_same_bits(graph):
    """The wall density pass on its own stream beside the fluid density pass
    (off Shepard steps) against the two in sequence: 25 steps across the
    Shepard refresh at step 20, eager and replayed, every field bit for bit."""
    from warp_sph_gpt_amd import _lib
    solv, st = contact_scene(seed=11)
    saved = _lib.get_tuning()
    runs = {}
    try:
        for b in (0, 1):
            _lib.set_tuning({"wall_beside": b})
            eng = CoupledEngine(solv, *(s.to(DEV) for s in st))
            if graph:
                eng.capture()
            eng.run(25)
            eng.check()
            runs[b] = eng
    finally:
        _lib.set_tuning(saved)
    assert_same_state(runs[1], runs[0], "wall pass beside vs after the fluid density")
    for f in ("rho", "pres"):
        np.testing.assert_array_equal(np_(getattr(runs[1].B, f)), np_(getattr(runs[0].B, f)),
                                      err_msg=f"wall {f}")


@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_force_pass_beside_the_porosity_gives_the_same_bits(graph):
    """The force pass on its own stream beside the porosity passes against
    the force before them on one stream: 25 steps across the Shepard refresh
    at step 20, eager and replayed, every field bit for bit."""
    from warp_sph_gpt_amd import _lib
    solv, st = contact_scene(seed=12)
    saved = _lib.get_tuning()
    runs = {}
    try:
        for b in (0, 1):
            _lib.set_tuning({"force_beside": b})
            eng = CoupledEngine(solv, *(s.to(DEV) for s in st))
            if graph:
                eng.capture()
            eng.run(25)
            eng.check()
            runs[b] = eng
    finally:
        _lib.set_tuning(saved)
    assert_same_state(runs[1], runs[0], "force pass beside vs before the porosity")
