"""Surrogate rollout machinery on the GPU: the DEM substeps (a wave per DEM)
against the oracle's plain DEM stage (contacts, CSR history, integration)
with the coupling force held; the judged call keeps the plain result and a
rejected call reverts to the last judged state and redoes it with the plain
step; a captured surrogate call equals the eager one; training lowers the
loss; and the trained network beats the method without it (one explicit
step of ten substeps) on a held-out scene."""
from __future__ import annotations

import os

import numpy as np
import pytest
import torch

from helpers import assert_rel, np_, solv_for, to_oracle
from oracle import oracle as O
from test_gpu_rollout import wet_structs
from warp_sph_gpt_amd import _lib
from warp_sph_gpt_amd.engine import CoupledEngine
from warp_sph_gpt_amd.ensemble import EnsembleEngine
from warp_sph_gpt_amd.surrogate import (MPNN, SUBSTEPS, SurrogateRollout, Trainer, collect_pairs,
                                        load_params)
from warp_sph_gpt_amd.workloads import ensemble_scene, runtime_solv

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
WEIGHTS = os.path.join(os.path.dirname(__file__), "..", "warp_sph_gpt_amd", "weights",
                       "surrogate_c3.npz")


def wet(seed=4):
    gen, structs = wet_structs(seed)
    D = structs[2]
    D.pos[1] = D.pos[0] + torch.tensor([0.049, 0.0, 0.0])     # an overlapping pair
    return solv_for(gen), structs


def test_dem_substeps_match_the_plain_dem_stage():
    solv, structs = wet()
    rng = np.random.default_rng(3)
    D = structs[2]
    D.drag = torch.tensor(rng.normal(0, 0.05, (D.n, 3)), dtype=torch.float32)
    D.pressure_force = torch.tensor(rng.normal(0, 0.05, (D.n, 3)), dtype=torch.float32)
    eng = CoupledEngine(solv, *(s.to(DEV) for s in structs))
    e = eng
    e.gd.build(e.D.pos, e.r_dem)
    scratch = torch.zeros(int(_lib.load().sphdem_dem_substeps_scratch(e.D.n, 4)),
                          dtype=torch.uint8, device=DEV)
    _lib.call("sphdem_dem_substeps", e.params, _lib.dem_view(e.D), _lib.dembnd_view(e.W),
              e.gd.struct(), e.gw.struct(), SUBSTEPS, 0.01, _lib.fptr(scratch), e.ws.struct(),
              _lib.stream_handle())
    assert int(e.ws.flags[0]) == 0
    # the oracle: ten plain DEM stages (grid rebuilt each), coupling force held
    P, B, Do, Wo = to_oracle(structs)
    scales = O.dem_search_scales(Do, Wo)
    gd = O.OracleGrid(64, 64, 64)
    gw = O.OracleGrid(64, 64, 64).build(Wo.pos, max(2 * scales[1], solv.support))
    drag, pf = Do.drag.copy(), Do.pressure_force.copy()
    for _ in range(SUBSTEPS):
        gd.build(Do.pos, max(2 * scales[0], 2 * scales[2]))
        O.run_force_dem(Do, gd, scales[0], solv.g, solv.dt)
        O.run_bc_dem(Do, Wo, gw, scales[1], solv.dt)
        Do.force = (Do.force + drag) + pf
        O.Kernel_step_dem(Do, solv.dt)
    for f in ("pos", "vel", "omega", "force", "torque"):
        assert_rel(np_(getattr(e.D, f)), getattr(Do, f), 1e-5, what=f"dem {f}")
    for kind in ("dem", "bnd"):
        off = np_(getattr(e.D, f"contact_{kind}_offset_old"))
        np.testing.assert_array_equal(off, getattr(Do, f"contact_{kind}_offset_old"))
        E = int(off[-1])
        assert E > 0
        np.testing.assert_array_equal(np_(getattr(e.D, f"contact_{kind}_id_old"))[:E],
                                      getattr(Do, f"contact_{kind}_id_old"))
        assert_rel(np_(getattr(e.D, f"tang_{kind}_old"))[:E], getattr(Do, f"tang_{kind}_old"),
                   1e-5, what=f"tang {kind}")


def _state(ro):
    return {n: np_(getattr(ro.eng.P, n)).copy() for n in ("pos", "vel")} | \
        {"dpos": np_(ro.eng.D.pos).copy(), "dvel": np_(ro.eng.D.vel).copy()}


def _rollout(solv, structs, net, **kw):
    return SurrogateRollout(solv, *(s.to(DEV) for s in structs), net, **kw)


def test_judged_call_keeps_the_plain_result_and_rejection_reverts():
    solv, structs = wet()
    net = MPNN(DEV, seed=2)
    # keep: 2 surrogate calls, then a judged call == 2 surrogate calls + 10 plain steps
    a = _rollout(solv, structs, net, judge_every=3, density_tol=1e9, velocity_tol=1e9)
    a.run(3)
    b = _rollout(solv, structs, net, judge_every=10**6)
    b.run(2)
    b.plain_call()
    for k, v in _state(a).items():
        np.testing.assert_array_equal(v, _state(b)[k], err_msg=k)
    assert a.stats.judged == 1 and a.stats.rejected == 0 and a.stats.surrogate_calls == 2
    assert np.isfinite(a.stats.density_err[0]) and a.stats.density_err[0] > 0
    # reject: every judged call fails -> the 3 calls are redone with the plain step
    c = _rollout(solv, structs, net, judge_every=3, density_tol=-1.0, fallback_calls=2)
    c.run(5)
    d = CoupledEngine(solv, *(s.to(DEV) for s in structs), fast_math=True)
    d.run(5 * SUBSTEPS)
    np.testing.assert_array_equal(np_(c.eng.P.pos), np_(d.P.pos))
    np.testing.assert_array_equal(np_(c.eng.D.pos), np_(d.D.pos))
    assert c.stats.rejected == 1 and c.stats.surrogate_calls == 0


def test_captured_surrogate_call_equals_eager():
    solv, structs = wet()
    net = MPNN(DEV, seed=3)
    a = _rollout(solv, structs, net, judge_every=10**6)
    b = _rollout(solv, structs, net, judge_every=10**6)
    b.capture()
    for ro in (a, b):
        ro.plain_call()
        ro.run(3)                      # crosses both Shepard variants
    for k, v in _state(a).items():
        np.testing.assert_array_equal(v, _state(b)[k], err_msg=k)


def test_training_lowers_the_loss():
    solv, structs = wet()
    _, sc2, _ = ensemble_scene(41, dx=0.02, dt=5.0e-5, n_dem=0)
    solv.dt = 5.0e-5
    ens = EnsembleEngine(solv, [structs, sc2], device=DEV)
    pairs = collect_pairs(ens, snapshots=3, interval=20, warm=10)
    net = MPNN(DEV, seed=5)
    tr = Trainer(net, solv, structs[1], device=DEV)
    before = tr.evaluate(pairs)["loss"]
    tr.train(pairs, 120, lr=2e-3, log_every=40)
    after = tr.evaluate(pairs)["loss"]
    assert after < 0.7 * before, (before, after)


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="trained weights not present")
def test_trained_network_beats_the_method_without_it():
    """Held-out scene (seed outside the training ensemble), dx = 0.01."""
    gen, sc, _ = ensemble_scene(91001, dx=0.01, dt=5.0e-5)
    solv = runtime_solv(gen, DEV)
    ens = EnsembleEngine(solv, [sc], device=DEV)
    pairs = collect_pairs(ens, snapshots=3, interval=300, warm=40)
    net = MPNN(DEV, params=load_params(WEIGHTS))
    ev = Trainer(net, solv, sc[1], device=DEV).evaluate(pairs)
    print(ev)
    # the gain is reported; the bar is a strict improvement over the explicit step
    assert ev["loss"] < ev["loss_without_network"], ev
