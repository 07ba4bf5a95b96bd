"""The CPU restatement of the surrogate network (oracle/mpnn_ref.py) is pinned
by central finite differences in FP64 (every parameter tensor, sampled
entries), and its bf16 mode stays within bf16 rounding of FP64."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mpnn_ref as M


def tiny_graph(seed=0, nf=60, nw=90):
    rng = np.random.default_rng(seed)
    h, sup = 0.026, 0.052
    xf = rng.uniform(0, 0.1, (nf, 3))
    xw = np.c_[rng.uniform(0, 0.1, (nw, 2)), rng.uniform(-0.03, 0.0, nw)]
    vf = rng.normal(0, 0.2, (nf, 3))
    vw = np.zeros((nw, 3))
    graph = M.build_edges(xf, vf, xw, vw, h, sup)
    nfeat = M.node_features(vf, rng.normal(0, 5.0, (nf, 3)), 1000 * (1 + rng.normal(0, .01, nf)),
                            1000.0)
    return nfeat, graph


def test_edges_are_symmetric_and_within_support():
    nfeat, (i, j, wall, e) = tiny_graph()
    fl = wall == 0
    pairs = set(zip(i[fl].tolist(), j[fl].tolist()))
    assert pairs and all((b, a) in pairs for a, b in pairs)
    assert np.all(e[:, 3] < 2.0) and np.all(e[:, 3] > 0)
    np.testing.assert_allclose(np.linalg.norm(e[:, :3], axis=1), e[:, 3], rtol=1e-12)
    assert set(np.unique(e[:, 7])) <= {0.0, 1.0}


def test_backward_matches_central_differences():
    nfeat, graph = tiny_graph(1)
    p = M.init_params(3)
    for k in p:                       # non-zero biases so every term is exercised
        if p[k].ndim == 1:
            p[k] = p[k] + np.random.default_rng(len(k)).normal(0, 0.1, p[k].shape)
    rng = np.random.default_rng(7)
    target = rng.normal(0, 1, (nfeat.shape[0], 6))
    net = M.Net(p)
    out = net.forward(nfeat, graph)
    loss, dout = M.loss_and_grad(out, target)
    g = net.backward(dout)
    eps = 1e-6
    for name in p:
        flat = p[name].reshape(-1)
        gflat = g[name].reshape(-1)
        idx = rng.choice(flat.size, size=min(6, flat.size), replace=False)
        for k in idx:
            old = flat[k]
            flat[k] = old + eps
            lp, _ = M.loss_and_grad(M.Net(p).forward(nfeat, graph), target)
            flat[k] = old - eps
            lm, _ = M.loss_and_grad(M.Net(p).forward(nfeat, graph), target)
            flat[k] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - gflat[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (name, k, fd, gflat[k])


def test_bf16_mode_is_within_bf16_rounding():
    nfeat, graph = tiny_graph(2)
    p = M.init_params(4)
    a = M.Net(p).forward(nfeat, graph)
    b = M.Net(p, bf16_points=True).forward(nfeat, graph)
    err = np.linalg.norm(a - b) / np.linalg.norm(a)
    assert 1e-5 < err < 3e-2
    assert M.bf16(np.float32(1.0 + 2 ** -9)) == 1.0            # ties to even
    assert M.bf16(np.float32(1.0 + 3 * 2 ** -9)) == 1.0 + 2 ** -7
