"""HIP Eulerian ISPH cavity against the CPU oracle (csrc/eisph.hip).

Bars (stated per test):
  * bit-exact: the KGC matrices, the neighbour list's rows (ids in the
    HashGrid visiting order) and the FP32 fields of the passes whose sums
    follow that order (advection, assembly, projection, diagnostics);
  * the FP64 right-hand-side sum may round differently in its last bit, so
    fields downstream of the pressure solve are held to a relative error
    <= 1e-5 (max-norm over points, rows as vectors) against the oracle;
  * the five-launch step equals the ten reference launches run one by one;
  * the Re=100 run reproduces the reference's published figures.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import eisph as E
from test_eisph_oracle import ghia_errors
from warp_sph_gpt_amd import eisph as G
from warp_sph_gpt_amd.config import EISPHSolv
from warp_sph_gpt_amd.input_gen import CavityPtlGeneration, GenerationConfig

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
TOL = 1.0e-5
FIELDS = ("vel", "vel_star", "acc", "pres", "Aij", "bi", "Aijpj", "divergence", "residual")


def np_(t):
    return t.detach().cpu().numpy()


def rel_err(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    if a.ndim == 2:
        num, den = np.linalg.norm(a - b, axis=1), np.linalg.norm(b, axis=1)
    else:
        num, den = np.abs(a - b), np.abs(b)
    scale = max(float(den.max()) if den.size else 0.0, 1e-30)
    return float(num.max()) / scale if num.size else 0.0


def cavity(dtype=torch.float32, seed=None, **kw):
    gen = GenerationConfig(device="cpu", **kw)
    P, B = CavityPtlGeneration(gen).build()
    if seed is not None:
        rng = np.random.default_rng(seed)
        P.vel = torch.tensor(rng.normal(0, 0.1, (P.n, 3)) * [1, 0, 1], dtype=torch.float32)
        P.vel_star = P.vel.clone()
        P.pres = torch.tensor(rng.normal(0, 0.05, P.n), dtype=torch.float32)
    solv = EISPHSolv(device=DEV, h=gen.h, rho0=gen.rho0, nu=gen.nu, dt=gen.dt,
                     n_steps=gen.n_steps, output_step=gen.output_step, grid_slice=gen.grid_slice)
    odt = np.float32 if dtype == torch.float32 else np.float64
    oP, oB = E.from_struct(P, odt), E.from_struct(B, odt)
    B.vel = 2.0 * B.vel_bc - P.vel[B.mirror.long()]
    B.vel_star = B.vel.clone()
    oB.vel[:] = np_(B.vel)
    oB.vel_star[:] = np_(B.vel)
    return gen, solv, P.to(DEV, dtype), B.to(DEV, dtype), oP, oB


def brute_rows(P, B, support):
    """Neighbour rows by brute force: (fluid ids, ghost ids) sets per point."""
    xp, xb = np_(P.pos).astype(np.float32), np_(B.pos).astype(np.float32)
    s2 = np.float32(support) * np.float32(support)
    rows = []
    for i in range(len(xp)):
        df, db = xp[i] - xp, xp[i] - xb
        rf = (df * df).sum(1, dtype=np.float32)
        rb = (db * db).sum(1, dtype=np.float32)
        rows.append((set(np.nonzero((rf > 1e-12) & (rf < s2))[0]),
                     set(np.nonzero((rb > 1e-12) & (rb < s2))[0])))
    return rows


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_kgc_bit_exact(dtype):
    gen, solv, P, B, oP, oB = cavity(dtype, dx=0.1, grid_slice=16)
    gs, gb = G.prepare(solv, P, B)
    E.build_grids(solv, oP, oB)
    np.testing.assert_array_equal(np_(P.kgc).reshape(-1, 4), oP.kgc)


def test_neighbour_list_rows():
    gen, solv, P, B, oP, oB = cavity(dx=0.1, grid_slice=16)
    gs, gb = G.prepare(solv, P, B)
    L = G.neighbours(P, B, gs, gb, solv.support, solv.h)
    off, ids, geo = L.rows()
    geo = geo.astype(np.float64)
    rows = brute_rows(P, B, solv.support)
    assert off[0] == 0 and off[-1] == L.entries == ids.size
    assert L.slots == int(np_(L.width)[:-1].sum()) >= L.entries
    xp, xb = np_(P.pos).astype(np.float64), np_(B.pos).astype(np.float64)
    for i, (fl, gh) in enumerate(rows):
        r = ids[off[i]:off[i + 1]]
        nf = int((r >= 0).sum())
        assert np.all(r[:nf] >= 0) and np.all(r[nf:] < 0), "fluid entries come first"
        assert set(r[:nf]) == fl and set(-1 - r[nf:]) == gh
        # geometry of each entry against an FP64 restatement
        xj = np.where((r >= 0)[:, None], xp[np.maximum(r, 0)], xb[np.maximum(-1 - r, 0)])
        rij = xp[i] - xj
        rr = np.linalg.norm(rij, axis=1)
        q = rr / solv.h
        dw = -35.0 * q * (1 - 0.5 * q) ** 3 / (4 * np.pi * solv.h ** 3)
        g = dw[:, None] * rij / rr[:, None]
        c = np_(P.kgc)[i].astype(np.float64)
        gc = np.stack((c[0, 0] * g[:, 0] + c[0, 1] * g[:, 2], c[1, 0] * g[:, 0] + c[1, 1] * g[:, 2]), 1)
        vol = gen.dx ** 2
        lap = 2.0 * vol * (rij[:, 0] * gc[:, 0] + rij[:, 2] * gc[:, 1]) / rr ** 2
        e = geo[off[i]:off[i + 1]]
        np.testing.assert_allclose(e[:, :2], gc, rtol=1e-4, atol=1e-4 * np.abs(gc).max())
        np.testing.assert_allclose(e[:, 2], lap, rtol=1e-4)
        np.testing.assert_allclose(e[:, 3], vol, rtol=1e-6)


def _kernel_sequence(solv, P, B, gs, gb, diagnostics=True):
    """The ten launches of source/EISPH.py:57-90, one by one."""
    G.Kernel_Dirichlet_BC(P, B, 0)
    G.Kernel_advec_vis_eisph(P, B, gs, gb, solv.support, solv.h, solv.nu, solv.dt)
    G.Kernel_Dirichlet_BC(P, B, 1)
    G.Kernel_build_ppe_eisph(P, B, gs, gb, solv.support, solv.h, solv.rho0, solv.dt)
    G.solve_ppe(P, B, gs, gb, solv.support, solv.h)
    G.Kernel_pressure_boundary_eisph(P, B)
    G.Kernel_correct_eisph(P, B, gs, gb, solv.support, solv.h, solv.dt)
    G.Kernel_Dirichlet_BC(P, B, 0)
    if diagnostics:
        G.Kernel_diagnostics_eisph(P, B, gs, gb, solv.support, solv.h)


def test_kernel_passes_match_the_oracle_bit_for_bit():
    gen, solv, P, B, oP, oB = cavity(seed=3, dx=0.1, grid_slice=16)
    gs, gb = G.prepare(solv, P, B)
    ogs, ogb = E.build_grids(solv, oP, oB)
    G.Kernel_Dirichlet_BC(P, B, 0)
    E.Kernel_Dirichlet_BC(oP, oB, 0)
    G.Kernel_advec_vis_eisph(P, B, gs, gb, solv.support, solv.h, solv.nu, solv.dt)
    E.Kernel_advec_vis_eisph(oP, oB, ogs, ogb, solv.support, solv.h, solv.nu, solv.dt)
    for name in ("acc", "vel_star"):
        np.testing.assert_array_equal(np_(getattr(P, name)), getattr(oP, name), err_msg=name)
    G.Kernel_Dirichlet_BC(P, B, 1)
    E.Kernel_Dirichlet_BC(oP, oB, 1)
    np.testing.assert_array_equal(np_(B.vel_star), oB.vel_star)
    G.Kernel_build_ppe_eisph(P, B, gs, gb, solv.support, solv.h, solv.rho0, solv.dt)
    E.Kernel_build_ppe_eisph(oP, oB, ogs, ogb, solv.support, solv.h, solv.rho0, solv.dt)
    for name in ("divergence", "Aij", "bi", "Aijpj"):
        np.testing.assert_array_equal(np_(getattr(P, name)), getattr(oP, name), err_msg=name)
    # the FP64 sum of the raw right-hand side, rounded once
    s = float(np.sum(oP.bi.astype(np.float64)))
    assert abs(float(np_(P.residual)[0]) - s) <= 1e-6 * max(np.abs(oP.bi).sum(), 1e-30)
    G.solve_ppe(P, B, gs, gb, solv.support, solv.h)
    E.remove_bi_mean(oP)
    E.Kernel_solve_ppe_eisph(oP)
    for name in ("bi", "pres"):
        assert rel_err(np_(getattr(P, name)), getattr(oP, name)) <= TOL, name
    # the projection and diagnostics from the same pressure are bit-exact
    oP.pres[:] = np_(P.pres)
    oP.bi[:] = np_(P.bi)
    G.Kernel_pressure_boundary_eisph(P, B)
    E.Kernel_pressure_boundary_eisph(oP, oB)
    G.Kernel_correct_eisph(P, B, gs, gb, solv.support, solv.h, solv.dt)
    E.Kernel_correct_eisph(oP, oB, ogs, ogb, solv.support, solv.h, solv.dt)
    for name in ("vel", "acc"):
        np.testing.assert_array_equal(np_(getattr(P, name)), getattr(oP, name), err_msg=name)
    G.Kernel_Dirichlet_BC(P, B, 0)
    E.Kernel_Dirichlet_BC(oP, oB, 0)
    G.Kernel_diagnostics_eisph(P, B, gs, gb, solv.support, solv.h)
    E.Kernel_diagnostics_eisph(oP, oB, ogs, ogb, solv.support, solv.h)
    for name in ("divergence", "residual"):
        np.testing.assert_array_equal(np_(getattr(P, name)), getattr(oP, name), err_msg=name)


@pytest.mark.parametrize("diagnostics", [True, False])
def test_fused_step_equals_the_ten_launches(diagnostics):
    """Ghost values formed where they are read (sphdem_eisph_step) give the
    fields the stored-ghost launches give, bit for bit, ghosts included; a
    non-uniform vel_bc exercises the general Dirichlet condition."""
    gen, solv, P, B, _, _ = cavity(seed=5, dx=0.1, grid_slice=16)
    rng = np.random.default_rng(9)
    B.vel_bc = torch.tensor(rng.normal(0, 0.2, (B.n, 3)), dtype=torch.float32, device=DEV)
    gs, gb = G.prepare(solv, P, B)
    P2, B2 = P.clone(), B.clone()
    P2.kgc = P.kgc.clone()
    for _ in range(3):
        _kernel_sequence(solv, P, B, gs, gb, diagnostics)
        G.EISPH_OneStep(solv, P2, B2, gs, gb, diagnostics)
    torch.cuda.synchronize()
    for name in FIELDS:
        if not diagnostics and name in ("divergence", "residual"):
            continue
        assert torch.equal(getattr(P, name), getattr(P2, name)), name
    for name in ("vel", "vel_star", "pres"):
        assert torch.equal(getattr(B, name), getattr(B2, name)), f"ghost {name}"


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_twenty_steps_match_the_oracle(dtype):
    gen, solv, P, B, oP, oB = cavity(dtype, seed=1, dx=0.1, n_steps=20, output_step=10,
                                     grid_slice=16)
    pos, frames, times = G.run_eisph(solv, P, B)
    opos, oframes, otimes = E.run_eisph(solv, oP, oB)
    assert times == otimes and len(frames) == len(oframes) == 3
    np.testing.assert_array_equal(pos, opos)
    for name in FIELDS:
        assert rel_err(np_(getattr(P, name)), getattr(oP, name)) <= TOL, name
    for name in ("vel", "vel_star", "pres"):
        assert rel_err(np_(getattr(B, name)), getattr(oB, name)) <= TOL, f"ghost {name}"
    for a, b in zip(frames, oframes):
        assert rel_err(a, b) <= TOL


def test_cavity_step_is_fixed_finite_and_mirrored():
    """test_eisph.py:214-249 on the GPU path."""
    gen, solv, P, B, _, _ = cavity(dx=0.1, n_steps=20, output_step=20, grid_slice=16)
    pos0 = P.pos.clone()
    positions, frames, times = G.run_eisph(solv, P, B)
    assert torch.equal(P.pos, pos0)
    assert times == [0.0, solv.n_steps * solv.dt] and len(frames) == 2
    for name in FIELDS + ("kgc",):
        assert torch.isfinite(getattr(P, name)).all(), name
    assert torch.all(P.vel[:, 1] == 0.0)
    mirror = B.mirror.long()
    assert torch.allclose(B.vel, 2.0 * B.vel_bc - P.vel[mirror], rtol=1e-6, atol=1e-7)
    assert abs(float(P.bi.mean())) < 1e-4
    assert float(P.pres[0]) == 0.0
    assert float(P.divergence.norm()) < float((P.bi * solv.dt / solv.rho0).norm())
    u = np_(P.vel)[:, 0].reshape(gen.cells, gen.cells)
    assert u[-1, gen.cells // 2] > 0.0 and u[gen.cells // 2, gen.cells // 2] < 0.0


def test_pressure_update_is_one_implicit_diagonal_pass():
    """test_eisph.py:266-320 on the GPU path."""
    old = np.linspace(-0.25, 0.25, 25, dtype=np.float32)

    def one_step(p0):
        _, s, P, B, _, _ = cavity(dx=0.2, n_steps=1, output_step=1, grid_slice=16)
        P.pres = torch.tensor(p0, device=DEV)
        G.run_eisph(s, P, B)
        return P

    P = one_step(old)
    A, bi, Ap, p = np_(P.Aij), np_(P.bi), np_(P.Aijpj), np_(P.pres)
    active = np.abs(A) > 1e-12
    assert np.all(A[active] < 0.0)
    expected = old.copy()
    expected[active] = (bi[active] + Ap[active]) / A[active]
    expected -= expected[0] if active[0] else 0.0
    np.testing.assert_allclose(p, expected, rtol=2e-6, atol=2e-6)
    shifted = one_step(old + 7.5)
    np.testing.assert_allclose(np_(shifted.pres), p, rtol=2e-5, atol=2e-5)
    np.testing.assert_allclose(np_(shifted.vel), np_(P.vel), rtol=2e-5, atol=2e-6)


def test_laplacian_diffuses_a_centre_maximum():
    """test_eisph.py:144-181 on the GPU path: acc = -4 nu at the centre."""
    gen, solv, P, B, _, _ = cavity(dx=0.2, n_steps=1, output_step=1, grid_slice=16)
    x = np_(P.pos)
    r2 = (x[:, 0] - 0.5) ** 2 + (x[:, 2] - 0.5) ** 2
    v = np.zeros_like(x)
    v[:, 0] = -r2
    P.vel = torch.tensor(v, device=DEV)
    gs, gb = G.prepare(solv, P, B)
    G.Kernel_Dirichlet_BC(P, B, 0)
    G.Kernel_advec_vis_eisph(P, B, gs, gb, solv.support, solv.h, solv.nu, solv.dt)
    c = int(np.argmin(r2))
    np.testing.assert_allclose(np_(P.acc)[c, 0], -4.0 * solv.nu, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(np_(P.vel_star)[c, 0], -4.0 * solv.nu * solv.dt, rtol=1e-5,
                               atol=1e-7)


def test_list_follows_changes_of_mass_and_grids():
    gen, solv, P, B, oP, oB = cavity(dx=0.1, grid_slice=16)
    gs, gb = G.prepare(solv, P, B)
    L1 = G.neighbours(P, B, gs, gb, solv.support, solv.h)
    assert G.neighbours(P, B, gs, gb, solv.support, solv.h) is L1
    _, id1, geo1 = L1.rows()
    P.m.mul_(2.0)
    L2 = G.neighbours(P, B, gs, gb, solv.support, solv.h)
    assert L2 is not L1
    _, id2, geo2 = L2.rows()
    np.testing.assert_array_equal(id1, id2)
    np.testing.assert_allclose(geo2[:, 3][id2 >= 0], 2.0 * geo1[:, 3][id1 >= 0], rtol=1e-6)
    gb.build(B.pos, solv.support)
    assert G.neighbours(P, B, gs, gb, solv.support, solv.h) is not L2


def test_host_tensors_are_refused():
    gen, solv, P, B, _, _ = cavity(dx=0.2, grid_slice=16)
    gs, gb = G.prepare(solv, P, B)
    with pytest.raises(ValueError):
        G.EISPH_OneStep(solv, P.to("cpu"), B, gs, gb)


def test_re100_cavity_reproduces_the_published_run():
    """5000 steps of the shipped cavity: Ghia RMSE 0.00840 and maximum
    0.01550 (README.md:79-83), and the oracle's final field within 1e-4."""
    gen, solv, P, B, oP, oB = cavity()
    G.run_eisph(solv, P, B)
    err = ghia_errors(np_(P.vel), gen)
    rmse, emax = float(np.sqrt(np.mean(err ** 2))), float(np.abs(err).max())
    assert round(rmse, 5) == 0.00840 and round(emax, 5) == 0.01550
    c = gen.cells // 2
    v = np_(P.vel).reshape(gen.cells, gen.cells, 3)
    assert round(float(v[c, c, 0]), 4) == -0.1906 and round(float(v[c, c, 2]), 4) == 0.0563
    E.run_eisph(solv, oP, oB)
    assert rel_err(np_(P.vel), oP.vel) <= 1e-4


def test_command_line_runs_the_cavity_from_files(tmp_path, capsys):
    """main_EISPH.py on the GPU: files in, progress lines, the GIF out."""
    from warp_sph_gpt_amd.input_gen import write_particle_file
    from warp_sph_gpt_amd.main_eisph import main
    gen = GenerationConfig(device="cpu", dx=0.1, n_steps=20, output_step=10)
    P, B = CavityPtlGeneration(gen).build()
    write_particle_file(tmp_path / "input_EISPH.txt", "EISPH", P)
    write_particle_file(tmp_path / "input_EISPHBND.txt", "EISPHBND", B)
    gif = tmp_path / "cavity.gif"
    main(["--device", DEV, "--input-dir", str(tmp_path), "--steps", "20", "--output-step", "10",
          "--h", str(gen.h), "--animation", str(gif), "--fps", "5"])
    out = capsys.readouterr().out
    assert "[EISPH]    10 / 20" in out and "[EISPH]    20 / 20" in out
    assert gif.is_file()


def test_graph_replays_match_eager_steps():
    """A hipGraph of two steps, replayed three times with eager work on the
    same device in between, equals two eager steps each time, bit for bit."""
    gen, solv, P, B, _, _ = cavity(seed=2, dx=0.1, grid_slice=16)
    P2, B2 = P.clone(), B.clone()
    gs, gb = G.prepare(solv, P, B)
    gs2, gb2 = G.prepare(solv, P2, B2)
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(graph, stream=side):
            for _ in range(2):
                G.EISPH_OneStep(solv, P2, B2, gs2, gb2, diagnostics=False)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    for _ in range(3):
        for _ in range(2):
            G.EISPH_OneStep(solv, P, B, gs, gb, diagnostics=False)
        noise = torch.randn(1 << 16, device=DEV).sum()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.isfinite(noise)
        for name in ("vel", "vel_star", "acc", "pres", "Aij", "bi", "Aijpj", "residual"):
            assert torch.equal(getattr(P, name), getattr(P2, name)), name
        for name in ("vel", "vel_star", "pres"):
            assert torch.equal(getattr(B, name), getattr(B2, name)), f"ghost {name}"


def test_many_workgroup_sums_take_the_separate_reduction():
    """160 x 160 points: more than 64 workgroup sums, so the step adds them
    in a one-workgroup launch; still the ten launches' bits, and the oracle
    within the bar after three steps."""
    gen, solv, P, B, oP, oB = cavity(seed=4, dx=0.00625, dt=4e-4, grid_slice=64)
    assert (P.n + B.n + 255) // 256 > 64
    gs, gb = G.prepare(solv, P, B)
    P2, B2 = P.clone(), B.clone()
    gs2, gb2 = G.prepare(solv, P2, B2)
    ogs, ogb = E.build_grids(solv, oP, oB)
    for _ in range(3):
        G.EISPH_OneStep(solv, P, B, gs, gb)
        _kernel_sequence(solv, P2, B2, gs2, gb2)
        E.EISPH_OneStep(solv, oP, oB, ogs, ogb)
    torch.cuda.synchronize()
    for name in FIELDS:
        assert torch.equal(getattr(P, name), getattr(P2, name)), name
        assert rel_err(np_(getattr(P, name)), getattr(oP, name)) <= TOL, name


@pytest.mark.parametrize("dtype,dx,fits", [(torch.float32, 0.04, True),
                                           (torch.float64, 0.1, True),
                                           (torch.float32, 0.025, False)])
@pytest.mark.parametrize("steps,diagnostics", [(1, True), (6, False), (6, True)])
def test_steps_equal_repeated_one_steps(dtype, dx, fits, steps, diagnostics):
    """EISPH_Steps (one workgroup for the shipped 25 x 25 cavity and a 10 x 10
    FP64 one; five-launch steps for 40 x 40, beyond one workgroup's LDS) leaves
    every field, the ghosts and the sum buffer as `steps` EISPH_OneStep calls
    with diagnostics on the last one only, bit for bit."""
    gen, solv, P, B, _, _ = cavity(dtype, seed=7, dx=dx, grid_slice=16)
    rng = np.random.default_rng(11)
    B.vel_bc = torch.tensor(rng.normal(0, 0.2, (B.n, 3)), dtype=dtype, device=DEV)
    P2, B2 = P.clone(), B.clone()
    gs, gb = G.prepare(solv, P, B)
    gs2, gb2 = G.prepare(solv, P2, B2)
    assert G.one_group(P, B) == fits
    for s in range(steps):
        G.EISPH_OneStep(solv, P, B, gs, gb, diagnostics and s == steps - 1)
    G.EISPH_Steps(solv, P2, B2, gs2, gb2, steps, diagnostics)
    torch.cuda.synchronize()
    for name in FIELDS:
        assert torch.equal(getattr(P, name), getattr(P2, name)), name
    for name in ("vel", "vel_star", "pres"):
        assert torch.equal(getattr(B, name), getattr(B2, name)), f"ghost {name}"
    L = G.neighbours(P, B, gs, gb, solv.support, solv.h)
    L2 = G.neighbours(P2, B2, gs2, gb2, solv.support, solv.h)
    assert torch.equal(L.part, L2.part)


def test_steps_with_rows_longer_than_the_register_rows():
    """Support 1.5x the cavity's (about 45 list entries a row): the FP32
    one-point-per-thread kernel takes the entries beyond its register rows
    from the list; 5 steps still equal 5 EISPH_OneStep calls bit for bit."""
    gen, solv, P, B, _, _ = cavity(torch.float32, seed=8, dx=0.04, grid_slice=16)
    solv.h = 1.5 * solv.h
    P2, B2 = P.clone(), B.clone()
    gs, gb = G.prepare(solv, P, B)
    gs2, gb2 = G.prepare(solv, P2, B2)
    L = G.neighbours(P, B, gs, gb, solv.support, solv.h)
    assert G.one_group(P, B) and int(L.len.max()) > 24
    for s in range(5):
        G.EISPH_OneStep(solv, P, B, gs, gb, s == 4)
    G.EISPH_Steps(solv, P2, B2, gs2, gb2, 5, True)
    torch.cuda.synchronize()
    for name in FIELDS:
        assert torch.equal(getattr(P, name), getattr(P2, name)), name
    for name in ("vel", "vel_star", "pres"):
        assert torch.equal(getattr(B, name), getattr(B2, name)), f"ghost {name}"


def test_steps_zero_is_a_no_op_and_negative_is_refused():
    gen, solv, P, B, _, _ = cavity(seed=3, dx=0.1, grid_slice=16)
    gs, gb = G.prepare(solv, P, B)
    before = [getattr(P, name).clone() for name in FIELDS]
    G.EISPH_Steps(solv, P, B, gs, gb, 0)
    torch.cuda.synchronize()
    for name, t in zip(FIELDS, before):
        assert torch.equal(getattr(P, name), t), name
    with pytest.raises(ValueError):
        G.EISPH_Steps(solv, P, B, gs, gb, -1)
