"""Batched ensemble: every instance of one batched run is bit-equal to the
same instance run alone through the plain engine (same kernels, same order),
also under hipGraph replay; launch shapes chosen by the autotuner (or any
valid shapes) change no bit."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import np_
from test_gpu_rollout import wet_structs
from warp_sph_gpt_amd import _lib, autotune
from warp_sph_gpt_amd.engine import CoupledEngine
from warp_sph_gpt_amd.ensemble import EnsembleEngine
from warp_sph_gpt_amd.workloads import ensemble_scene, runtime_solv

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
FLUID = ("pos", "vel", "rho", "pres", "acc", "flt")
COUPLED = ("porosity", "pgf", "acc_dem")
DEM = ("pos", "vel", "omega", "force", "torque", "flt_s", "porosity", "drag", "pressure_force")


def scenes():
    """Uneven instances at dx = 0.02, dt = 5e-5 (the contact limit of the
    smallest spheres): two wet scenes with DEM in the fluid and in contact,
    a dry scene without DEM, a scene with spheres above its free surface."""
    g0, wet0 = wet_structs(seed=4)
    _, wet1 = wet_structs(seed=5)
    wet0[2].pos[1] = wet0[2].pos[0] + torch.tensor([0.049, 0.0, 0.0])   # an overlapping pair
    _, dry, _ = ensemble_scene(31, dx=0.02, dt=5.0e-5, n_dem=0)
    _, above, _ = ensemble_scene(32, dx=0.02, dt=5.0e-5, n_dem=12)
    solv = runtime_solv(g0, DEV)
    solv.dt = 5.0e-5
    return solv, [wet0, dry, above, wet1]


def assert_same(a, b, what):
    a, b = np_(a), np_(b)
    assert a.shape == b.shape, what
    np.testing.assert_array_equal(a, b, err_msg=what)


def compare(ens, k, alone):
    P, B, D, W = ens.instance_state(k)
    fields = FLUID + (COUPLED if D is not None else ())
    for f in fields:
        assert_same(getattr(P, f), getattr(alone.P, f), f"instance {k} fluid {f}")
    for f in ("rho", "pres"):
        assert_same(getattr(B, f), getattr(alone.B, f), f"instance {k} wall {f}")
    if D is None:
        return
    for f in DEM:
        assert_same(getattr(D, f), getattr(alone.D, f), f"instance {k} dem {f}")
    for kind in ("dem", "bnd"):
        off = getattr(alone.D, f"contact_{kind}_offset_old")
        E = int(off[-1])
        assert_same(getattr(D, f"contact_{kind}_offset_old"), off, f"instance {k} {kind} offsets")
        assert_same(getattr(D, f"contact_{kind}_id_old"),
                    getattr(alone.D, f"contact_{kind}_id_old")[:E], f"instance {k} {kind} ids")
        assert_same(getattr(D, f"tang_{kind}_old"), getattr(alone.D, f"tang_{kind}_old")[:E],
                    f"instance {k} {kind} tang")


@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_each_instance_bit_equal_to_its_solo_run(graph):
    autotune.reset()
    solv, sc = scenes()
    ens = EnsembleEngine(solv, sc, device=DEV)
    ens.run(2)
    if graph:
        ens.capture()
    ens.run(23)                               # crosses the Shepard refresh at step 20
    ens.check()
    for k, (P, B, D, W) in enumerate(sc):
        alone = CoupledEngine(solv, P.to(DEV), B.to(DEV), None if D is None else D.to(DEV),
                              None if W is None else W.to(DEV))
        alone.run(25)
        alone.check()
        compare(ens, k, alone)
    # the scenes did interact with their DEM
    _, _, D0, _ = ens.instance_state(0)
    assert int(D0.contact_dem_offset_old[-1]) > 0 and int(D0.contact_bnd_offset_old[-1]) > 0
    assert float(D0.drag.abs().max()) > 0


def test_launch_shapes_change_no_bit():
    solv, sc = scenes()
    runs = []
    for cfg in ({}, {"block": {p: 64 for p in _lib.TUNABLE_PASSES},
                     "xcd_remap": {p: 1 for p in _lib.TUNABLE_PASSES}, "por_x_tile": 2},
                {"block": {p: 128 for p in _lib.TUNABLE_PASSES}, "por_x_tile": 6},
                {"force_beside": 0, "wall_beside": 0, "dem_beside": 0}):
        autotune.reset()
        _lib.set_tuning(cfg)
        ens = EnsembleEngine(solv, sc, device=DEV)
        ens.run(22)
        ens.check()
        runs.append(ens)
    autotune.reset()
    for other in runs[1:]:
        for f in FLUID + COUPLED:
            assert_same(getattr(other.P, f), getattr(runs[0].P, f), f"fluid {f}")
        for f in DEM:
            assert_same(getattr(other.D, f), getattr(runs[0].D, f), f"dem {f}")


def test_autotuner_picks_valid_shapes_and_changes_no_bit():
    solv, sc = scenes()
    autotune.reset()
    ref = EnsembleEngine(solv, sc, device=DEV)
    ref.run(4)
    tuned = EnsembleEngine(solv, sc, device=DEV)
    probe = EnsembleEngine(solv, sc, device=DEV)
    res = autotune.autotune(probe.step, steps=1)
    cfg = _lib.get_tuning()
    for p in _lib.TUNABLE_PASSES:
        assert cfg["block"][p] in (64, 128, 256) and cfg["xcd_remap"][p] in (0, 1)
        assert cfg["block"][p] == res["tuning"]["block"][p]
    tuned.run(4)
    autotune.reset()
    for f in FLUID + COUPLED:
        assert_same(getattr(tuned.P, f), getattr(ref.P, f), f"fluid {f}")
