"""The surrogate network's HIP kernels (BF16 MFMA, FP32 accumulation) against
the CPU restatement oracle/mpnn_ref.py: dense layers, the full forward, and
the hand-written backward (every parameter gradient), on a small dam-break
neighbourhood graph.  Bars (relative L2 per tensor): the bf16-emulating
reference within 1e-2 (forward and gradients); the FP64 reference, whose
backward is pinned by finite differences, within 5e-2 (forward and
gradients)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import COMPACT, generation, np_, particles, solv_for
from oracle import mpnn_ref as M
from warp_sph_gpt_amd import _lib
from warp_sph_gpt_amd.hashgrid import HashGrid
from warp_sph_gpt_amd.simulation import solver_params
from warp_sph_gpt_amd.surrogate import MPNN, param_shapes

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_param_layout_matches_the_reference():
    assert param_shapes() == M.param_shapes()


def test_dense_layer_on_mfma():
    """Y = silu(X W + b) and its backward for one layer (checks the MFMA
    operand layouts the packing assumes)."""
    rng = np.random.default_rng(0)
    n, K = 37, 64
    X = M.bf16(rng.normal(0, 1, (n, K)))
    W = rng.normal(0, 0.2, (K, 64))
    b = rng.normal(0, 0.1, 64)
    net = MPNN(DEV)
    net.p["enc1_W"].copy_(torch.tensor(W, dtype=torch.float32))
    net.p["enc1_b"].copy_(torch.tensor(b, dtype=torch.float32))
    net.pack()
    xb = torch.tensor(X.astype(np.float32), device=DEV).view(torch.int32)
    xb = (xb >> 16).to(torch.int16).contiguous()
    Yf = torch.zeros((n, 64), device=DEV)
    Yb = torch.zeros((n, 64), dtype=torch.int16, device=DEV)
    net._fwd(xb, K, None, 0, n, "enc1_W", "enc1_b", True, Yf=Yf, Yb=Yb)
    pre = X @ M.bf16(W) + b
    assert rel_l2(np_(Yf), pre) < 1e-5                     # FP32 output: pre-activation
    yb = (Yb.to(torch.int32) << 16).view(torch.float32)   # bf16 output: activation
    assert rel_l2(np_(yb), M.bf16(M.silu(pre))) < 1e-2
    dY = rng.normal(0, 1, (n, 64))
    dYt = torch.tensor(dY, dtype=torch.float32, device=DEV)
    dX = torch.zeros((n, K), device=DEV)
    net.grad.zero_()
    net._bwd(dYt, Yf, xb, K, None, 0, n, "enc1_W", "enc1_b", dX)
    dZ = dY * M.dsilu(pre)
    assert rel_l2(np_(net.g["enc1_W"]), X.T @ M.bf16(dZ)) < 1e-5
    assert rel_l2(np_(net.g["enc1_b"]), dZ.sum(0)) < 1e-5
    assert rel_l2(np_(dX), M.bf16(dZ) @ M.bf16(W).T) < 1e-5


def scene(seed=0):
    gen = generation(**COMPACT)
    P, B, D, W = particles(gen)
    rng = np.random.default_rng(seed)
    P.pos = P.pos + torch.tensor(rng.uniform(-0.15, 0.15, (P.n, 3)) * gen.dx, dtype=torch.float32)
    P.vel = torch.tensor(rng.normal(0, 0.3, (P.n, 3)), dtype=torch.float32)
    P.acc = torch.tensor(rng.normal(0, 6.0, (P.n, 3)), dtype=torch.float32)
    P.rho = torch.tensor(1000 * (1 + rng.normal(0, 0.01, P.n)), dtype=torch.float32)
    solv = solv_for(gen)
    return gen, solv, P.to(DEV), B.to(DEV)


def random_params(seed):
    p = M.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    for k in p:
        if p[k].ndim == 1 and k != "hw":
            p[k] = rng.normal(0, 0.1, p[k].shape)
    p["dec2_b"][6:] = 0.0
    return p


def test_forward_and_backward_match_the_reference():
    gen, solv, P, B = scene()
    prm = solver_params(solv)
    gs = HashGrid(32, 32, 32, device=DEV).build(P.pos, solv.support)
    gb = HashGrid(32, 32, 32, device=DEV).build(B.pos, solv.support)
    p = random_params(1)
    net = MPNN(DEV, params=p)
    out = net.forward(prm, P, gs, gb)
    ids = np_(gs.point_ids).astype(np.int64)
    got = np.zeros((P.n, 64))
    got[ids] = np_(out)
    xf, vf = np_(P.pos).astype(np.float64), np_(P.vel).astype(np.float64)
    graph = M.build_edges(xf, vf, np_(B.pos).astype(np.float64), np.zeros((B.n, 3)),
                          float(np.float32(solv.h)), float(np.float32(solv.support)))
    nfeat = M.bf16(M.node_features(vf, np_(P.acc), np_(P.rho), 1000.0))
    ref_bf = M.Net(p, bf16_points=True)
    want = ref_bf.forward(nfeat, graph)
    want64 = M.Net(p).forward(nfeat, graph)
    assert np.abs(want[:, :6]).max() > 1e-2
    assert rel_l2(got[:, :6], want[:, :6]) < 1e-2
    assert rel_l2(got[:, :6], want64[:, :6]) < 5e-2
    # backward of a mean-square loss against a random target
    target = np.random.default_rng(3).normal(0, 1, (P.n, 6))
    _, dout = M.loss_and_grad(want, target)
    g_ref = ref_bf.backward(dout)
    dsorted = torch.tensor(dout[ids], dtype=torch.float32, device=DEV)
    net.grad.zero_()
    net.backward(prm, gs, gb, dsorted)
    torch.cuda.synchronize()
    g64 = M.Net(p)
    g64.forward(nfeat, graph)
    g_ref64 = g64.backward(dout)
    bad, errs = {}, {}
    for name, _ in param_shapes():
        if np.linalg.norm(g_ref[name]) == 0:
            continue
        got_g = np_(net.g[name])
        errs[name] = (rel_l2(got_g, g_ref[name]), rel_l2(got_g, g_ref64[name]))
        if errs[name][0] > 1e-2 or errs[name][1] > 5e-2:
            bad[name] = errs[name]
    print({k: (f"{a:.1e}", f"{b:.1e}") for k, (a, b) in errs.items()})
    assert not bad, bad


def test_loss_and_adam_kernels():
    rng = np.random.default_rng(5)
    n = 1000
    out = rng.normal(0, 1, (n, 64)).astype(np.float32)
    tgt = rng.normal(0, 1, (n, 6)).astype(np.float32)
    loss = torch.zeros(1, dtype=torch.float64, device=DEV)
    dout = torch.zeros((n, 64), device=DEV)
    o_t, t_t = torch.tensor(out, device=DEV), torch.tensor(tgt, device=DEV)
    _lib.call("sphdem_mpnn_loss", _lib.fptr(o_t), _lib.fptr(t_t), n, _lib.fptr(loss),
              _lib.fptr(dout), _lib.stream_handle())
    ref, dref = M.loss_and_grad(out.astype(np.float64), tgt.astype(np.float64))
    assert abs(float(loss) - ref) < 1e-6 * ref
    assert rel_l2(np_(dout), dref) < 1e-6
    # Adam: two steps against the textbook update
    p0 = rng.normal(0, 1, 500)
    gr = [rng.normal(0, 1, 500) for _ in range(2)]
    pt = torch.tensor(p0, dtype=torch.float32, device=DEV)
    m = torch.zeros_like(pt)
    v = torch.zeros_like(pt)
    pr, mr, vr = p0.copy(), np.zeros(500), np.zeros(500)
    for k, g in enumerate(gr, 1):
        gt = torch.tensor(g, dtype=torch.float32, device=DEV)
        _lib.call("sphdem_adam", _lib.fptr(pt), _lib.fptr(gt), _lib.fptr(m), _lib.fptr(v), 500,
                  1e-2, 0.9, 0.999, 1e-8, k, _lib.stream_handle())
        mr = 0.9 * mr + 0.1 * g
        vr = 0.999 * vr + 0.001 * g * g
        pr = pr - 1e-2 * (mr / (1 - 0.9 ** k)) / (np.sqrt(vr / (1 - 0.999 ** k)) + 1e-8)
    assert rel_l2(np_(pt), pr) < 1e-6
