"""Output path on the CPU: the native "%.6g" / "%d" formatter against
Python's, the VTK XML frame and collection layout of output/output.py, the
non-finite guard of output/frame_writer.py, the GIF frame count, and the
command-line flags of main.py."""
from __future__ import annotations

import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest
import torch

from warp_sph_gpt_amd import _lib
from warp_sph_gpt_amd.struct import BNDptl, DEMBNDptl, DEMptl, SPHptl

pytestmark = pytest.mark.skipif(not _lib.LIB_PATH.exists(), reason="libsphdem.so is not built")


def py_fmt(a, per_line, spec):
    v = [spec % x for x in np.asarray(a).ravel()]
    return "\n".join(" ".join(v[i:i + per_line]) for i in range(0, len(v), per_line))


def test_native_formatter_matches_python():
    from warp_sph_gpt_amd.output import fmt_float, fmt_int
    rng = np.random.default_rng(0)
    vals = np.concatenate([
        rng.normal(0, 1, 500), rng.normal(0, 1e5, 100), rng.normal(0, 1e-5, 100),
        [0.0, -0.0, 1e-4, 9.99999e-5, 1e-5, 123456.0, 1234567.0, 999999.5, 0.1 + 0.2,
         np.float32(0.1), 1e30, -1e-30, 1e-40]])
    vals32 = vals.astype(np.float32).astype(np.float64)
    for v in (vals, vals32):
        assert fmt_float(v) == py_fmt(v, 6, "%.6g")
        assert fmt_float(v.reshape(-1, 1), per_line=7) == py_fmt(v, 7, "%.6g")
    ints = np.concatenate([np.arange(-5, 45), [2 ** 31 - 1, -2 ** 31]]).astype(np.int32)
    assert fmt_int(ints) == py_fmt(ints, 20, "%d")
    assert fmt_float(np.zeros(0)) == "" and fmt_int(np.zeros(0, np.int32)) == ""


def small_scene(nf=7, nb=5, nd=3, nw=4):
    rng = np.random.default_rng(1)
    P, B = SPHptl.allocate(nf), BNDptl.allocate(nb)
    D, W = DEMptl.allocate(nd), DEMBNDptl.allocate(nw)
    for S in (P, B, D, W):
        for name, layout in S.FIELDS:
            t = getattr(S, name)
            if t.is_floating_point() and t.numel():
                t.copy_(torch.tensor(rng.normal(0, 1, tuple(t.shape)), dtype=t.dtype))
    return P, B, D, W


def test_frames_and_collection_follow_the_vtk_layout(tmp_path):
    from warp_sph_gpt_amd.output import save_dem_vtk, save_pvd, save_vtk
    P, B, D, W = small_scene()
    frames = [save_vtk(P, B, s, s * 1e-4, out_dir=str(tmp_path), name="sph") for s in (0, 2, 3)]
    dem = save_dem_vtk(D, W, 3, 3e-4, out_dir=str(tmp_path))
    pvd = save_pvd(frames, out_dir=str(tmp_path), name="sph")
    root = ET.parse(pvd)
    ds = root.findall(".//DataSet")
    assert [d.get("file") for d in ds] == ["sph_000000.vtp", "sph_000002.vtp", "sph_000003.vtp"]
    assert float(ds[-1].get("timestep")) == pytest.approx(3e-4)
    v = ET.parse(os.path.join(tmp_path, "sph_000003.vtp"))
    piece = v.find(".//Piece")
    assert piece.get("NumberOfPoints") == "12" and piece.get("NumberOfVerts") == "12"
    pts = np.fromstring(v.find(".//Points/DataArray").text, sep=" ").reshape(-1, 3)
    np.testing.assert_allclose(pts[:7], P.pos.numpy(), rtol=1e-5)
    names = [a.get("Name") for a in v.findall(".//PointData/DataArray")]
    assert names == ["vel", "rho", "pres", "m", "porosity", "pgf", "acc_dem", "type"]
    types = np.fromstring(v.find(".//PointData/DataArray[@Name='type']").text, sep=" ")
    assert list(types) == [1] * 7 + [0] * 5
    conn = np.fromstring(v.find(".//Verts/DataArray[@Name='offsets']").text, sep=" ")
    assert list(conn) == list(range(1, 13))
    d = ET.parse(os.path.join(tmp_path, dem[0]))
    assert d.find(".//PointData").get("Scalars") == "radius"
    types = np.fromstring(d.find(".//PointData/DataArray[@Name='type']").text, sep=" ")
    fixed = types == 3
    assert fixed.sum() == 4 and (types[~fixed] == 2).all()
    for field in ("acc", "force", "torque", "drag", "pressure_force"):
        vec = np.fromstring(d.find(f".//PointData/DataArray[@Name='{field}']").text,
                            sep=" ").reshape(-1, 3)
        np.testing.assert_array_equal(vec[fixed], 0.0)
    por = np.fromstring(d.find(".//PointData/DataArray[@Name='porosity']").text, sep=" ")
    np.testing.assert_array_equal(por[fixed], 1.0)


def test_frame_writer_rejects_non_finite_before_writing(tmp_path):
    from warp_sph_gpt_amd.config import Solv
    from warp_sph_gpt_amd.output import FrameWriter
    P, B, D, W = small_scene()
    s = Solv(output_dir=str(tmp_path), gif_save=False)
    with pytest.raises(ValueError):
        FrameWriter(s, P, B, D, None)
    fw = FrameWriter(s, P, B, D, W)
    fw.write_frame(0)
    D.acc[1, 2] = float("nan")
    with pytest.raises(FloatingPointError, match="non-finite acc at step 5"):
        fw.write_frame(5)
    assert sorted(os.listdir(tmp_path)) == ["dem_000000.vtp", "sph_000000.vtp"]
    fw.finalize()
    assert os.path.exists(os.path.join(tmp_path, "sph.pvd"))


def test_gif_has_one_frame_per_state(tmp_path):
    from PIL import Image

    from warp_sph_gpt_amd.config import Solv
    from warp_sph_gpt_amd.gif import collect_frame, save_gif, sphere_faces
    P, B, D, W = small_scene()
    frames = [collect_frame(P, B, 0.1 * k, D, W) for k in range(3)]
    assert frames[0]["pos_sph"].shape == (7, 3) and frames[0]["pos_dem"].shape == (3, 3)
    path = save_gif(frames, Solv(), ani_dir=str(tmp_path), name="g")
    with Image.open(path) as im:
        assert im.n_frames == 3
    f = sphere_faces()
    np.testing.assert_allclose(np.linalg.norm(f.reshape(-1, 3), axis=1), 1.0, atol=1e-14)
    with pytest.raises(ValueError):
        save_gif(frames, Solv(), ani_dir=str(tmp_path), axis_limits=((0, 0, 0), (1, 1, 0)))


def test_command_line_flags_follow_the_reference():
    from warp_sph_gpt_amd.config import Solv
    from warp_sph_gpt_amd.main import parsing
    assert parsing([]) == {}
    a = parsing(["--steps", "12", "--dt", "5e-5", "--output-step", "3", "--no-gif",
                 "--no-dem", "--output-dir", "o", "--animation-dir", "a", "--device", "cuda:0",
                 "--input-dir", "in"])
    assert a == {"n_steps": 12, "dt": 5e-5, "output_step": 3, "gif_save": False,
                 "dem_enable": False, "output_dir": "o", "animation_dir": "a",
                 "device": "cuda:0", "input_dir": "in"}
    s = Solv(**a)
    assert s.n_steps == 12 and not s.dem_enable and s.h == 0.026
