"""Particle files, generators and solver validation (CPU only).

Mirrors the format/validation cases of the reference's
tests/test_input_reader.py:53-164 and the geometry checks of
tests/test_3d.py:55-85; the generator is pinned against the sha256 of the
input files the reference ships (tests/golden/reference_input_sha256.json)
and the parser against their first rows (tests/golden/reference_input_head).
"""
from __future__ import annotations

import hashlib
import json
import math
from dataclasses import fields
from pathlib import Path

import numpy as np
import pytest
import torch

from warp_sph_gpt_amd.config import BaseSolv, Solv
from warp_sph_gpt_amd.input_gen import (CoupledGenerationConfig, DamPtlGeneration,
                                        DEMPtlGeneration, generate_files, rsa_spheres,
                                        write_particle_file)
from warp_sph_gpt_amd.input_reader import (INPUT_FIELDS, column_indices, dem_search_scales,
                                           input_parser, load_sph_particles,
                                           validate_dem_particles)

GOLDEN = Path(__file__).parent / "golden"
COMPACT_GEN = dict(dx=0.04, h=0.052, tank_width=0.8, tank_height=0.8, tank_depth=0.32,
                   fluid_width=0.4, fluid_height=0.4, fluid_depth=0.32, dem_nx=2, dem_ny=1,
                   dem_nz=1, dem_origin_y=0.16, dem_origin_z=0.2)


@pytest.fixture(scope="module")
def compact_files(tmp_path_factory):
    d = tmp_path_factory.mktemp("inputs")
    gen = CoupledGenerationConfig(device="cpu", **COMPACT_GEN)
    return d, gen, generate_files(d, gen)


def runtime(directory, gen) -> Solv:
    s = Solv(device="cpu", input_dir=str(directory))
    for name in ("h", "rho0", "dt", "dem_porosity_min", "dem_porosity_max", "dem_dt_safety"):
        setattr(s, name, getattr(gen, name))
    return s


def test_generator_reproduces_the_reference_input_files(tmp_path):
    golden = json.loads((GOLDEN / "reference_input_sha256.json").read_text())
    paths = generate_files(tmp_path)
    for kind, path in paths.items():
        data = path.read_bytes()
        assert hashlib.sha256(data).hexdigest() == golden[kind]["sha256"], kind
        assert data.count(b"\n") - 1 == golden[kind]["rows"]


def test_parser_reads_reference_rows_like_the_generator():
    gen = CoupledGenerationConfig(device="cpu")
    P, B = DamPtlGeneration(gen).build()
    D, W = DEMPtlGeneration(gen).build()
    for kind, ref in (("SPH", P), ("BND", B), ("DEM", D), ("DEMBND", W)):
        got = input_parser(GOLDEN / "reference_input_head" / f"input_{kind}.txt")
        n = got.n
        for name, _ in INPUT_FIELDS[kind]:
            np.testing.assert_allclose(getattr(got, name).numpy(), getattr(ref, name).numpy()[:n],
                                       rtol=2e-7, atol=0, err_msg=f"{kind}.{name}")


def test_documented_column_ids():
    assert column_indices("SPH") == {"pos": (1, 2, 3), "vel": (4, 5, 6), "rho": (7,), "m": (8,)}
    assert column_indices("DEM")["K"] == (14,)
    assert column_indices("DEM")["h"] == (17,)
    assert column_indices("DEMBND")["mu"] == (16,)
    banned = {"dx", "bnd_layer", "dem_K", "dem_eta", "dem_mu", "dem_h", "dem_radius",
              "tank_width"}
    assert not banned.intersection(f.name for f in fields(Solv))
    for kind in INPUT_FIELDS:
        assert not {"pres", "acc", "volume", "flt", "porosity"}.intersection(column_indices(kind))


def test_files_load_into_the_generated_structs(compact_files):
    d, gen, _ = compact_files
    direct = [*DamPtlGeneration(gen).build(), *DEMPtlGeneration(gen).build()]
    loaded = load_sph_particles(runtime(d, gen))
    for a, b in zip(loaded, direct):
        for name, _ in a.FIELDS:
            np.testing.assert_allclose(getattr(a, name).numpy(), getattr(b, name).numpy(),
                                       rtol=2e-7, atol=0, err_msg=name)


def test_reordered_columns_dispatch_by_id(compact_files, tmp_path):
    _, _, paths = compact_files
    for kind, source in paths.items():
        lines = source.read_text().splitlines()
        path = tmp_path / source.name
        path.write_text(" ".join(reversed(lines[0].split())) + "\n" +
                        " ".join(reversed(lines[1].split())) + "\n")
        got, ref = input_parser(path), input_parser(source)
        for name, _ in INPUT_FIELDS[kind]:
            np.testing.assert_array_equal(getattr(got, name).numpy()[0],
                                          getattr(ref, name).numpy()[0])


@pytest.mark.parametrize("text", [
    "", "1 1 2 3 4 5 6 7\n0 0 0 0 0 0 1000 0.001\n", "1 2 3 4 5 6 7 99\n0 0 0 0 0 0 1000 0.001\n",
    "1.0 2 3 4 5 6 7 8\n0 0 0 0 0 0 1000 0.001\n", "1 2 3 4 5 6 7 8\n0 0\n",
    "1 2 3 4 5 6 7 8\n\n", "1 2 3 4 5 6 7 8\n0 0 0 0 0 0 nan 0.001\n",
    "1 2 3 4 5 6 7 8\n0 0 0 0 0 0 inf 0.001\n", "1 2 3 4 5 6 7 8\n0 0 0 0 0 0 1e100 0.001\n",
    "1 2 3 4 5 6 7 8\n0 0 0 0 0 0 -1 0.001\n", "1 2 3 4 5 6 7 8\n0 0 0 0 0 0 1000 1e-100\n",
    "1 2 3 4 5 6 7 8\n0 0 0 0 0 0 oops 0.001\n"])
def test_invalid_input_is_rejected(tmp_path, text):
    path = tmp_path / "input_SPH.txt"
    path.write_text(text)
    with pytest.raises(ValueError):
        input_parser(path)


def test_empty_wall_file_and_zero_contact_buffers(compact_files, tmp_path):
    d, gen, paths = compact_files
    path = tmp_path / "input_DEMBND.txt"
    path.write_text(" ".join(str(i) for i in range(1, 17)) + "\n")
    wall = input_parser(path)
    assert wall.n == 0
    dem = input_parser(paths["DEM"])
    assert validate_dem_particles(runtime(d, gen), dem, wall)[1] == 0
    for kind in ("dem", "bnd"):
        for side in ("old", "new"):
            assert getattr(dem, f"contact_{kind}_id_{side}").numel() == 0
            np.testing.assert_array_equal(getattr(dem, f"contact_{kind}_offset_{side}").numpy(), 0)


def test_loaded_materials_round_trip_and_dt_limit(compact_files, tmp_path):
    d, gen, paths = compact_files
    _, _, dem, wall = load_sph_particles(runtime(d, gen))
    dem.h = torch.tensor([0.04, 0.12])
    dem.K = torch.tensor([2e4, 8e4])
    dem.eta = torch.tensor([4.0, 7.0])
    dem.mu = torch.tensor([0.1, 0.6])
    path = write_particle_file(tmp_path / "input_DEM.txt", "DEM", dem)
    loaded = input_parser(path)
    for name in ("K", "eta", "mu", "h"):
        np.testing.assert_array_equal(getattr(loaded, name).numpy(), getattr(dem, name).numpy())
    scales = dem_search_scales(loaded, wall)
    assert abs(scales[2] - 0.12) < 1e-7 and scales[3] == 0.0
    s = runtime(d, gen)
    s.dt = 0.5
    with pytest.raises(ValueError, match="loaded DEM contact limit"):
        validate_dem_particles(s, loaded, wall)


def test_solver_validation():
    Solv(device="cpu").validate_sph()
    with pytest.raises(ValueError, match="acoustic"):
        Solv(dt=1e-3).validate_sph()
    with pytest.raises(ValueError):
        Solv(shepard_step=-1).validate_sph()
    with pytest.raises(ValueError):
        Solv(dem_porosity_min=0.0).validate_sph()
    assert Solv().support == 2 * Solv().h
    assert BaseSolv().h == 0.052 and Solv().h == 0.026 and Solv().rho0 == 1000.0


def test_volume_mass_and_five_wall_geometry():
    """tests/test_3d.py:55-85."""
    s = CoupledGenerationConfig(device="cpu")
    fluid = DamPtlGeneration(s).fluid_particle()
    wall = DamPtlGeneration(s).boundary_particle()
    dem = DEMPtlGeneration(s).dem_particle()
    dwall = DEMPtlGeneration(s).boundary_particle()
    assert math.isclose(len(fluid) * s.rho0 * s.dx ** 3, 100.0)
    assert len(dem) == 20 and len(fluid) == 12500 and len(wall) == 48336 and len(dwall) == 14688
    assert math.isclose(s.dem_inertia, 2 * s.dem_mass * s.dem_radius ** 2 / 5)
    for w in (wall, dwall):
        assert len(w) == len(np.unique(w, axis=0))
        assert all(w[:, a].min() < 0.0 for a in range(3))
        assert w[:, 0].max() >= s.tank_width and w[:, 2].max() >= s.tank_height
        interior = (w[:, 0] > 0) & (w[:, 0] < s.tank_width) & (w[:, 1] > 0) & (w[:, 1] < s.tank_depth)
        assert np.all(w[interior, 2] < 0.0)
    with pytest.raises(ValueError):
        DEMPtlGeneration(CoupledGenerationConfig(dem_origin_y=0.39)).dem_particle()
    with pytest.raises(ValueError):
        CoupledGenerationConfig(dt=1.0e-3).validate_dem()


def test_rsa_packing_is_seeded_and_non_overlapping():
    a, ra, _ = rsa_spheres(200, 0.025, (0.05, 0.05, 0.55), (0.45, 0.35, 0.9), 7)
    b, rb, _ = rsa_spheres(200, 0.025, (0.05, 0.05, 0.55), (0.45, 0.35, 0.9), 7)
    np.testing.assert_array_equal(a, b)
    d = np.sqrt(((a[:, None] - a[None]) ** 2).sum(-1))
    np.fill_diagonal(d, 1.0)
    assert d.min() >= 0.05 and len(a) == 200
    c, rc, hi = rsa_spheres(400, lambda rng, n: rng.uniform(0.015, 0.03, n),
                            (0.05, 0.05, 0.5), (0.3, 0.3, 0.6), 3, attempts=5000, grow_z=0.05,
                            r_max=0.03)
    d = np.sqrt(((c[:, None] - c[None]) ** 2).sum(-1))
    np.fill_diagonal(d, 1.0)
    assert len(c) == 400 and np.all(d >= (rc[:, None] + rc[None]) - 1e-12)


def test_generator_command_line_writes_the_shipped_case_and_c1(tmp_path):
    """python -m warp_sph_gpt_amd.input_gen: c0 reproduces the shipped files
    byte for byte; c1 writes 100,000 fluid and 200 DEM rows that read back."""
    import hashlib
    import json
    from pathlib import Path

    from warp_sph_gpt_amd.config import Solv
    from warp_sph_gpt_amd.input_gen import main
    from warp_sph_gpt_amd.input_reader import load_sph_particles
    golden = json.loads((Path(__file__).parent / "golden" / "reference_input_sha256.json")
                        .read_text())
    main(["--output-dir", str(tmp_path / "c0")])
    for kind in ("SPH", "BND", "DEM", "DEMBND"):
        data = (tmp_path / "c0" / f"input_{kind}.txt").read_bytes()
        assert hashlib.sha256(data).hexdigest() == golden[kind]["sha256"]
    main(["--output-dir", str(tmp_path / "c1"), "--workload", "c1"])
    s = Solv(device="cpu", input_dir=str(tmp_path / "c1"), h=0.013, dt=5e-5)
    P, B, D, W = load_sph_particles(s)
    assert (P.n, B.n, D.n, W.n) == (100_000, 180_456, 200, 58_017)


def test_parsed_fields_are_packed_row_major():
    """The kernels take packed (N, 3) arrays: every parsed field is contiguous."""
    for kind in INPUT_FIELDS:
        got = input_parser(GOLDEN / "reference_input_head" / f"input_{kind}.txt")
        for name, _ in got.FIELDS:
            assert getattr(got, name).is_contiguous(), f"{kind}.{name}"


@pytest.mark.parametrize("text,message", [
    ("", r":1: empty header or duplicate column IDs"),
    ("1 1 2 3 4 5 6 7\n", r":1: empty header or duplicate column IDs"),
    ("1 2 3 4 5 6 7 99\n", r":1: missing IDs \[8\]; unknown IDs \[99\]"),
    ("1.0 2 3 4 5 6 7 8\n", r":1: header must contain integer column IDs"),
    ("1 2 3 4 5 6 7 8\n0 0 0 0 0 0 1000 0.001\n0 0\n", r":3: expected 8 values, got 2"),
    ("1 2 3 4 5 6 7 8\n0 0 0 0 0 0 1000 0.001\n0 0 0 0 0 0 oops 0.001\n",
     r":3: invalid numeric value"),
    ("1 2 3 4 5 6 7 8\n0 0 0 0 0 0 1000 0.001\n0 0 0 0 0 0 nan 0.001\n",
     r":3: values must be finite float32 numbers"),
    ("1 2 3 4 5 6 7 8\n0 0 0 0 0 0 1e100 0.001\n", r":2: values must be finite float32 numbers"),
    ("1 2 3 4 5 6 7 8\n0 0 0 0 0 0 -1 0.001\n", r": rho must be positive in float32"),
    ("1 2 3 4 5 6 7 8\n0 0 0 0 0 0 1000 1e-100\n", r": m must be positive in float32"),
])
def test_rejections_name_the_line_and_the_rule(tmp_path, text, message):
    """The messages of input/input_reader.py:48-95: the offending line of a
    row error, the field of a value error."""
    path = tmp_path / "input_SPH.txt"
    path.write_text(text)
    with pytest.raises(ValueError, match=message):
        input_parser(path)


def test_sphere_rules(tmp_path):
    head = " ".join(str(i) for i in range(1, 18))
    ok = "0 0 0 0 0 0 0 0 0 0.025 2500 0.16362461737446837 1e-5 2e4 4 0.3 0.08"
    bad_mass = ok.replace("0.16362461737446837", "0.2")
    bad_eta = ok.replace(" 4 0.3", " -4 0.3")
    path = tmp_path / "input_DEM.txt"
    path.write_text(head + "\n" + ok + "\n")
    assert input_parser(path).n == 1
    for row, message in ((bad_mass, "sphere mass must equal"), (bad_eta, "eta must be nonnegative")):
        path.write_text(head + "\n" + row + "\n")
        with pytest.raises(ValueError, match=message):
            input_parser(path)


def test_generator_rules_name_the_first_broken_one():
    """input_gen/config.py:227-353: rules in order, each with its message."""
    from warp_sph_gpt_amd.input_gen import GenerationConfig
    cases = [
        (CoupledGenerationConfig(dx=-0.02).validate_sph, "dx must be finite and positive"),
        (CoupledGenerationConfig(dt=1.0e-3).validate_sph, "acoustic limit"),
        (CoupledGenerationConfig(tank_width=2.01).validate_sph, "tank_width must be an integer"),
        (CoupledGenerationConfig(fluid_width=2.5).validate_sph,
         "initial fluid block must fit the tank along x"),
        (CoupledGenerationConfig(dem_eta=-1.0).validate_dem, "dem_eta and dem_mu"),
        (CoupledGenerationConfig(dem_spacing=0.04).validate_dem, "must not overlap"),
        (CoupledGenerationConfig(dem_porosity_min=0.0).validate_dem, "porosity bounds"),
        (CoupledGenerationConfig(dt=5.0e-4, c0=1.0, h=10.0).validate_dem, "DEM contact limit"),
        (GenerationConfig(dx=1.0 / 1024, dt=2.0e-5).validate_eisph, "viscosity limit"),
        (GenerationConfig(dx=1.0 / 25, dt=0.02).validate_eisph, "advection limit"),
    ]
    for check, message in cases:
        with pytest.raises(ValueError, match=message):
            check()
    GenerationConfig(dx=1.0 / 1024, dt=1.0e-5).validate_eisph()
    CoupledGenerationConfig().validate_sph()
    CoupledGenerationConfig().validate_dem()
