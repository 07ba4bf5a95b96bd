"""Cavity generator, EISPH input files and validation (CPU only).

Mirrors the host-side cases of the reference's tests/test_eisph.py:47-142
(preprocessor defaults, mirrored ghosts) and tests/test_input_reader.py
(EISPH column IDs, mirror range, initial Dirichlet velocity, file vs
generated structs).  The generated files are pinned byte for byte by the
sha256 of the reference's shipped input_EISPH*.txt in
test_input_files.py::test_generator_reproduces_the_reference_input_files.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
import torch

from warp_sph_gpt_amd.config import EISPHSolv
from warp_sph_gpt_amd.input_gen import (CavityPtlGeneration, CoupledGenerationConfig,
                                        GenerationConfig, write_particle_file)
from warp_sph_gpt_amd.input_reader import (column_indices, input_parser, load_eisph_particles,
                                           validate_eisph_particles)

GOLDEN = Path(__file__).parent / "golden"


def test_preprocessor_keeps_the_example_defaults():
    cav, sd = GenerationConfig(device="cpu"), CoupledGenerationConfig(device="cpu")
    assert (cav.dx, cav.bnd_layer, cav.cells, cav.rho0, cav.dt) == (0.04, 3, 25, 1.0, 2.0e-3)
    assert cav.h == pytest.approx(0.052) and cav.support == pytest.approx(0.104)
    assert cav.nu == pytest.approx(0.01) and not cav.dem_enable
    assert (cav.n_steps, cav.output_step) == (5000, 100)
    assert sd.dx == 0.02 and sd.h == pytest.approx(0.026) and sd.dem_enable
    cav.validate_eisph()
    scaled = GenerationConfig(device="cpu", dx=0.2, h_factor=1.4)
    assert scaled.h == pytest.approx(0.28) and scaled.support == pytest.approx(0.56)
    scaled.validate_eisph()


@pytest.mark.parametrize("kw, msg", [
    (dict(dx=0.3), "integer multiple"),
    (dict(bnd_layer=2), "kernel support"),
    (dict(dx=0.2, bnd_layer=6), "cannot exceed"),
    (dict(dt=0.02), "advection limit"),
    (dict(h=0.06), "h_factor"),
    (dict(reynolds=-1.0), "finite and positive"),
])
def test_cavity_settings_are_validated(kw, msg):
    with pytest.raises(ValueError, match=msg):
        GenerationConfig(device="cpu", **kw).validate_eisph()


def test_generator_builds_mirrored_cavity_ghosts():
    """test_eisph.py:98-121: cell centres, one struct type, lid velocity,
    mirror rows inside the fluid set, initial ghost velocity 2 u_D."""
    solv = GenerationConfig(device="cpu", dx=0.2, bnd_layer=3, n_steps=1, output_step=1)
    g = CavityPtlGeneration(solv)
    fluid = g.fluid_particle()
    bpos, vel_bc, mirror = g.boundary_particle()
    P, B = g.build()
    assert type(P) is type(B)
    assert fluid.shape == (solv.cells ** 2, 3)
    assert fluid[:, 0].min() == pytest.approx(0.5 * solv.dx)
    assert fluid[:, 0].max() == pytest.approx(solv.length - 0.5 * solv.dx)
    assert np.all((mirror >= 0) & (mirror < len(fluid)))
    np.testing.assert_array_equal(B.pos.numpy(), bpos)
    np.testing.assert_array_equal(B.vel_bc.numpy(), vel_bc)
    np.testing.assert_array_equal(B.vel.numpy(), 2.0 * vel_bc)
    np.testing.assert_array_equal(P.mirror.numpy(), np.arange(P.n))
    top = (bpos[:, 2] > solv.length) & (bpos[:, 0] > 0.0) & (bpos[:, 0] < solv.length)
    assert np.all(vel_bc[top, 0] == solv.lid_velocity)
    assert np.all(vel_bc[~top] == 0.0)
    # every ghost reflects through its nearest wall onto a cell centre
    xm = fluid[mirror]
    L = solv.length
    fx = np.where(bpos[:, 0] < 0, -bpos[:, 0], np.where(bpos[:, 0] > L, 2 * L - bpos[:, 0],
                                                         bpos[:, 0]))
    fz = np.where(bpos[:, 2] < 0, -bpos[:, 2], np.where(bpos[:, 2] > L, 2 * L - bpos[:, 2],
                                                         bpos[:, 2]))
    np.testing.assert_allclose(xm[:, 0], fx, atol=1e-6)
    np.testing.assert_allclose(xm[:, 2], fz, atol=1e-6)


def test_column_ids_and_reference_rows():
    assert column_indices("EISPH") == {"pos": (1, 2, 3), "vel": (4, 5, 6), "rho": (7,),
                                       "m": (8,)}
    assert column_indices("EISPHBND")["mirror"] == (9,)
    P, B = CavityPtlGeneration(GenerationConfig(device="cpu")).build()
    for kind, ref in (("EISPH", P), ("EISPHBND", B)):
        got = input_parser(GOLDEN / "reference_input_head" / f"input_{kind}.txt")
        n = got.n
        assert n == 40
        for name in ("pos", "rho", "m"):
            np.testing.assert_array_equal(getattr(got, name).numpy(),
                                          getattr(ref, name).numpy()[:n], err_msg=name)
    got = input_parser(GOLDEN / "reference_input_head" / "input_EISPHBND.txt")
    np.testing.assert_array_equal(got.mirror.numpy(), B.mirror.numpy()[:40])
    np.testing.assert_array_equal(got.vel_bc.numpy(), B.vel_bc.numpy()[:40])
    assert got.mirror.dtype == torch.int32


def _write_case(d: Path, gen: GenerationConfig):
    P, B = CavityPtlGeneration(gen).build()
    write_particle_file(d / "input_EISPH.txt", "EISPH", P)
    write_particle_file(d / "input_EISPHBND.txt", "EISPHBND", B)
    return P, B


def _solv(d, gen) -> EISPHSolv:
    return EISPHSolv(device="cpu", input_dir=str(d), h=gen.h, rho0=gen.rho0, nu=gen.nu,
                     dt=gen.dt, n_steps=gen.n_steps, output_step=gen.output_step)


def test_files_load_into_the_generated_structs(tmp_path):
    gen = GenerationConfig(device="cpu", dx=0.2, n_steps=10, output_step=5)
    P, B = _write_case(tmp_path, gen)
    fl, gh = load_eisph_particles(_solv(tmp_path, gen))
    for a, b in ((fl, P), (gh, B)):
        for name, _ in a.FIELDS:
            np.testing.assert_array_equal(getattr(a, name).numpy(), getattr(b, name).numpy(),
                                          err_msg=name)
    np.testing.assert_array_equal(gh.vel.numpy(),
                                  2 * gh.vel_bc.numpy() - fl.vel.numpy()[gh.mirror.numpy()])
    np.testing.assert_array_equal(fl.vel_star.numpy(), fl.vel.numpy())


@pytest.mark.parametrize("bad", ["-1", "1.5", "25", "2147483648"])
def test_out_of_range_mirror_is_refused(tmp_path, bad):
    gen = GenerationConfig(device="cpu", dx=0.2, n_steps=10, output_step=5)
    _write_case(tmp_path, gen)
    lines = (tmp_path / "input_EISPHBND.txt").read_text().splitlines()
    row = lines[1].split()
    row[-1] = bad
    (tmp_path / "input_EISPHBND.txt").write_text(lines[0] + "\n" + " ".join(row) + "\n")
    with pytest.raises(ValueError):
        load_eisph_particles(_solv(tmp_path, gen))


def test_particle_validation():
    gen = GenerationConfig(device="cpu", dx=0.2, n_steps=10, output_step=5)
    P, B = CavityPtlGeneration(gen).build()
    s = EISPHSolv(device="cpu", h=gen.h, nu=gen.nu, dt=gen.dt)
    validate_eisph_particles(s, P, B)
    for change, msg in [
        (lambda p, b: p.rho.fill_(2.0), "uniform rho"),
        (lambda p, b: p.pos[:, 1].fill_(0.1), "x-z plane"),
        (lambda p, b: p.vel.fill_(100.0), "advection limit"),
    ]:
        p, b = P.clone(), B.clone()
        change(p, b)
        with pytest.raises(ValueError, match=msg):
            validate_eisph_particles(s, p, b)
    with pytest.raises(ValueError, match="viscosity limit"):
        validate_eisph_particles(EISPHSolv(device="cpu", h=gen.h, nu=10.0, dt=gen.dt), P, B)
    with pytest.raises(ValueError, match="nu must be"):
        validate_eisph_particles(EISPHSolv(device="cpu", nu=-1.0), P, B)
    with pytest.raises(ValueError, match="nonempty"):
        validate_eisph_particles(s, P, type(B).allocate(0))


def test_cavity_gif_has_one_frame_per_velocity_field(tmp_path):
    """test_eisph.py:379-397: two frames of a vortex, written where asked."""
    from PIL import Image

    from warp_sph_gpt_amd.gif import save_cavity_gif
    solv = EISPHSolv(device="cpu", gif_fps=4)
    pos = CavityPtlGeneration(GenerationConfig(device="cpu", dx=0.2)).fluid_particle()
    zero = np.zeros_like(pos)
    vortex = zero.copy()
    vortex[:, 0] = -(pos[:, 2] - 0.5)
    vortex[:, 2] = pos[:, 0] - 0.5
    dest = tmp_path / "cavity.gif"
    saved = save_cavity_gif(pos, [zero, vortex], [0.0, 0.1], solv, dest)
    assert Path(saved) == dest.resolve() and dest.is_file()
    with Image.open(saved) as im:
        assert im.n_frames == 2
    with pytest.raises(ValueError, match="same non-zero length"):
        save_cavity_gif(pos, [zero], [0.0, 0.1], solv, dest)
    bad = zero.copy()
    bad[0, 0] = np.nan
    with pytest.raises(ValueError, match="finite"):
        save_cavity_gif(pos, [bad], [0.0], solv, dest)
