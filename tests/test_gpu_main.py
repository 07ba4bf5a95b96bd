"""run_forward end to end on the GPU (the reference's
test_separate_output_final_frame_and_animation, tests/test_dem.py:599-623):
frames at step 0, at the output cadence and at the off-cadence last step,
phase-specific VTK types, zero placeholders on the fixed DEM wall, and one
GIF frame per written state -- through the reference's step functions and
through the device-resident engine, whose final state must be the same."""
from __future__ import annotations

import contextlib
import io
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from helpers import COMPACT, generation, np_, particles, solv_for
from warp_sph_gpt_amd.hashgrid import HashGrid
from warp_sph_gpt_amd.main import run_forward

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _run(tmp, engine):
    gen = generation(**COMPACT)
    structs = [s.to(DEV) for s in particles(gen)]
    s = solv_for(gen)
    s.n_steps, s.output_step, s.output_dir, s.animation_dir = 3, 2, tmp, tmp
    s.gif_save = True
    g = [HashGrid(32, 32, 32, device=DEV) for _ in range(4)]
    with contextlib.redirect_stdout(io.StringIO()):
        run_forward(s, structs[0], structs[1], g[0], g[1], structs[2], structs[3], g[2], g[3],
                    engine=engine)
    return s, structs


@pytest.mark.parametrize("engine", [False, True], ids=["steps", "engine"])
def test_run_forward_writes_initial_cadence_and_final_frames(tmp_path, engine):
    from PIL import Image
    tmp = str(tmp_path)
    s, structs = _run(tmp, engine)
    for name, types in (("sph", {0, 1}), ("dem", {2, 3})):
        ds = ET.parse(os.path.join(tmp, name + ".pvd")).findall(".//DataSet")
        assert len(ds) == 3
        assert float(ds[-1].get("timestep")) == pytest.approx(3 * s.dt)
        data = ET.parse(os.path.join(tmp, ds[-1].get("file")))
        t = np.fromstring(data.find(".//PointData/DataArray[@Name='type']").text, sep=" ")
        assert set(t.astype(int)) == types
        if name == "dem":
            fixed = t == 3
            for field in ("acc", "force", "torque"):
                v = np.fromstring(data.find(f".//PointData/DataArray[@Name='{field}']").text,
                                  sep=" ").reshape(-1, 3)
                np.testing.assert_array_equal(v[fixed], 0.0)
    with Image.open(os.path.join(tmp, "dam_break_sph_dem_3d.gif")) as gif:
        assert gif.n_frames == 3


def test_engine_and_step_functions_reach_the_same_state(tmp_path):
    _, a = _run(str(tmp_path / "a"), False)
    _, b = _run(str(tmp_path / "b"), True)
    for k in (0, 2):
        for f in ("pos", "vel"):
            np.testing.assert_array_equal(np_(getattr(a[k], f)), np_(getattr(b[k], f)))


def test_command_line_reads_the_input_files(tmp_path):
    """main.py from files: the generator's shipped dam-break case, read back
    and stepped on the GPU, frames at step 0 and the last step."""
    from warp_sph_gpt_amd.input_gen import CoupledGenerationConfig, generate_files
    from warp_sph_gpt_amd.main import main
    gen = CoupledGenerationConfig(device="cpu")
    generate_files(tmp_path / "in", gen)
    out = tmp_path / "out"
    with contextlib.redirect_stdout(io.StringIO()):
        main(["--device", DEV, "--input-dir", str(tmp_path / "in"), "--steps", "2",
              "--output-step", "5", "--no-gif", "--output-dir", str(out), "--dt", str(gen.dt)])
    ds = ET.parse(str(out / "sph.pvd")).findall(".//DataSet")
    assert [float(d.get("timestep")) for d in ds] == pytest.approx([0.0, 2 * gen.dt])
