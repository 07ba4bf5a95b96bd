"""CPU restatement of the learned surrogate's network (test infrastructure only).

The network has no counterpart in the reference; this file states its
arithmetic in numpy so that the HIP forward and backward kernels can be
checked, and its own backward is pinned by central finite differences
(tests/test_mpnn_ref.py).  `bf16=True` rounds at the same points as the HIP
kernels (MFMA operands: node/edge features, weights, z1, Q, the inputs of
every dense layer and, in the backward, the pre-activation gradients that
enter a product; accumulation in wider precision); `bf16=False` is plain
FP64.

Graph: fluid nodes i, neighbours j among fluid and wall points with
1e-12 < |x_j - x_i|^2 < support^2.
Edge features e_ij (16, zero padded): (x_j - x_i)/h, |x_j - x_i|/h,
(v_j - v_i)/U, wall flag, W(r, h) h^3 (Wendland C2 without the 1/h^3).
Node features n_i (16, zero padded): v_i/U, a_i/A, 100 (rho_i/rho0 - 1).
Network (width 64, SiLU act(x) = x / (1 + e^-x)):
  h   = enc(n)                           3 layers 16-64-64-64
  round r = 0, 1:
    P = h Wa_r, Q = h Wb_r, Q_wall = hw Wb_r
    z1_ij = act(P_i + Q_j + e_ij C_r + c1_r)
    z2_ij = act(z1_ij W2_r + c2_r)
    s_i   = (1/64) sum_j z2_ij
    a_i   = s_i W3_r + c3_r
    h     = h + node_r([h, a])           3 layers 128-64-64-64
  out = dec(h)                           3 layers 64-64-64-64 (columns 0..5 used)
"""
from __future__ import annotations

import numpy as np

F = 64
NIN = 16
EIN = 16
SUM_SCALE = 1.0 / 64.0
U_SCALE = 1.0
A_SCALE = 9.81


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def param_shapes():
    """(name, shape) of every parameter, in the flat layout order."""
    out = [("enc0_W", (NIN, F)), ("enc0_b", (F,)), ("enc1_W", (F, F)), ("enc1_b", (F,)),
           ("enc2_W", (F, F)), ("enc2_b", (F,)), ("hw", (F,))]
    for r in range(2):
        out += [(f"r{r}_Wa", (F, F)), (f"r{r}_Wb", (F, F)), (f"r{r}_C", (EIN, F)),
                (f"r{r}_c1", (F,)), (f"r{r}_W2", (F, F)), (f"r{r}_c2", (F,)),
                (f"r{r}_W3", (F, F)), (f"r{r}_c3", (F,)),
                (f"r{r}_n0_W", (2 * F, F)), (f"r{r}_n0_b", (F,)),
                (f"r{r}_n1_W", (F, F)), (f"r{r}_n1_b", (F,)),
                (f"r{r}_n2_W", (F, F)), (f"r{r}_n2_b", (F,))]
    out += [("dec0_W", (F, F)), ("dec0_b", (F,)), ("dec1_W", (F, F)), ("dec1_b", (F,)),
            ("dec2_W", (F, F)), ("dec2_b", (F,))]
    return out


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in param_shapes():
        if len(shape) == 2:
            p[name] = rng.normal(0.0, np.sqrt(2.0 / shape[0]), shape)
        else:
            p[name] = np.zeros(shape)
    p["dec2_W"][:, 6:] = 0.0
    p["dec2_W"] *= 0.1
    p["hw"] = rng.normal(0.0, 0.5, F)
    return p


def silu(x):
    return x / (1.0 + np.exp(-x))


def dsilu(x):
    s = 1.0 / (1.0 + np.exp(-x))
    return s * (1.0 + x * (1.0 - s))


def wendland_h3(r, h):
    q = r / h
    u = 1.0 - 0.5 * q
    return np.where(q < 2.0, 21.0 / (16.0 * np.pi) * u ** 4 * (1.0 + 2.0 * q), 0.0)


def node_features(vel, acc, rho, rho0):
    n = np.zeros((len(vel), NIN))
    n[:, 0:3] = vel / U_SCALE
    n[:, 3:6] = acc / A_SCALE
    n[:, 6] = 100.0 * (rho / rho0 - 1.0)
    return n


def build_edges(xf, vf, xw, vw, h, support):
    """Edges (i, j, is_wall) and features, brute force over all pairs."""
    ss = support * support
    rows, cols, wall, feats = [], [], [], []
    for kind, (xo, vo) in enumerate(((xf, vf), (xw, vw))):
        d = xo[None, :, :] - xf[:, None, :]
        r2 = (d * d).sum(-1)
        ii, jj = np.nonzero((r2 > 1e-12) & (r2 < ss))
        dd = d[ii, jj]
        r = np.sqrt(r2[ii, jj])
        e = np.zeros((len(ii), EIN))
        e[:, 0:3] = dd / h
        e[:, 3] = r / h
        e[:, 4:7] = (vo[jj] - vf[ii]) / U_SCALE
        e[:, 7] = float(kind)
        e[:, 8] = wendland_h3(r, h)
        rows.append(ii)
        cols.append(jj)
        wall.append(np.full(len(ii), kind))
        feats.append(e)
    return (np.concatenate(rows), np.concatenate(cols), np.concatenate(wall),
            np.concatenate(feats))


class Net:
    def __init__(self, params, bf16_points=False):
        self.p = params
        self.q = bf16 if bf16_points else (lambda x: np.asarray(x, np.float64))

    # dense layer y = act(x W + b) (act: SiLU or none); cache for backward
    def _dense(self, x, W, b, act, cache):
        xq, Wq = self.q(x), self.q(W)
        pre = xq @ Wq + b
        y = silu(pre) if act else pre
        cache.append((xq, Wq, pre, act))
        return y

    def _mlp(self, x, names, acts, cache):
        for nm, act in zip(names, acts):
            x = self._dense(x, self.p[nm + "_W"], self.p[nm + "_b"], act, cache)
        return x

    def forward(self, nfeat, graph):
        """graph = (i, j, is_wall, e) from build_edges; returns out (N, 64)."""
        p, q = self.p, self.q
        ii, jj, wall, e = graph
        N = nfeat.shape[0]
        self.N, self.graph = N, graph
        self.c = {"enc": []}
        h = self._mlp(nfeat, ["enc0", "enc1", "enc2"], [True, True, False], self.c["enc"])
        self.h_in = []
        self.rounds = []
        eq = q(e)
        for r in range(2):
            st = {}
            self.h_in.append(h)
            hq = q(h)
            P = hq @ q(p[f"r{r}_Wa"])
            Q = q(hq @ q(p[f"r{r}_Wb"]))
            Qw = q(q(p["hw"]) @ q(p[f"r{r}_Wb"]))
            Qj = np.where(wall[:, None] == 1, Qw[None, :], Q[np.where(wall == 1, 0, jj)])
            pre1 = P[ii] + Qj + eq @ q(p[f"r{r}_C"]) + p[f"r{r}_c1"]
            z1 = q(silu(pre1))
            pre2 = z1 @ q(p[f"r{r}_W2"]) + p[f"r{r}_c2"]
            z2 = silu(pre2)
            s = np.zeros((N, F))
            np.add.at(s, ii, z2)
            s *= SUM_SCALE
            cache_a = []
            a = self._dense(s, p[f"r{r}_W3"], p[f"r{r}_c3"], False, cache_a)
            cache_n = []
            t = self._mlp(np.concatenate([h, a], 1), [f"r{r}_n0", f"r{r}_n1", f"r{r}_n2"],
                          [True, True, False], cache_n)
            st.update(hq=hq, Qj=Qj, pre1=pre1, z1=z1, pre2=pre2, eq=eq, cache_a=cache_a,
                      cache_n=cache_n)
            self.rounds.append(st)
            h = h + t
        self.h_out = h
        self.c["dec"] = []
        out = self._mlp(h, ["dec0", "dec1", "dec2"], [True, True, False], self.c["dec"])
        return out

    def _dense_bwd(self, dy, cache_item, gW, gb):
        xq, Wq, pre, act = cache_item
        dz = dy * dsilu(pre) if act else dy
        dzq = self.q(dz)                  # the MFMA operand of both products
        gW += xq.T @ dzq
        gb += dz.sum(0)
        return dzq @ Wq.T

    def _mlp_bwd(self, dy, names, cache, g):
        for nm, item in zip(reversed(names), reversed(cache)):
            dy = self._dense_bwd(dy, item, g[nm + "_W"], g[nm + "_b"])
        return dy

    def backward(self, dout):
        """Gradients of sum(out * dout) with respect to every parameter."""
        p, q = self.p, self.q
        g = {k: np.zeros_like(v) for k, v in p.items()}
        ii, jj, wall, e = self.graph
        dh = self._mlp_bwd(dout, ["dec0", "dec1", "dec2"], self.c["dec"], g)
        for r in (1, 0):
            st = self.rounds[r]
            dcat = self._mlp_bwd(dh, [f"r{r}_n0", f"r{r}_n1", f"r{r}_n2"], st["cache_n"], g)
            dh = dh + dcat[:, :F]
            da = dcat[:, F:]
            ds = self._dense_bwd(da, st["cache_a"][0], g[f"r{r}_W3"], g[f"r{r}_c3"])
            ds = ds * SUM_SCALE
            dpre2 = ds[ii] * dsilu(st["pre2"])
            g[f"r{r}_W2"] += st["z1"].T @ q(dpre2)
            g[f"r{r}_c2"] += dpre2.sum(0)
            dz1 = q(dpre2) @ q(p[f"r{r}_W2"]).T
            dpre1 = dz1 * dsilu(st["pre1"])
            g[f"r{r}_c1"] += dpre1.sum(0)
            g[f"r{r}_C"] += st["eq"].T @ q(dpre1)
            dP = np.zeros((self.N, F))
            np.add.at(dP, ii, dpre1)
            fl = wall == 0
            dQ = np.zeros((self.N, F))
            np.add.at(dQ, jj[fl], dpre1[fl])
            dQw = dpre1[~fl].sum(0)
            hq = st["hq"]
            dP, dQ, dQw = q(dP), q(dQ), q(dQw)
            g[f"r{r}_Wa"] += hq.T @ dP
            g[f"r{r}_Wb"] += hq.T @ dQ + np.outer(q(p["hw"]), dQw)
            g["hw"] += q(p[f"r{r}_Wb"]) @ dQw
            dh = dh + dP @ q(p[f"r{r}_Wa"]).T + dQ @ q(p[f"r{r}_Wb"]).T
        self._mlp_bwd(dh, ["enc0", "enc1", "enc2"], self.c["enc"], g)
        return g


def loss_and_grad(out, target):
    """Mean squared error over nodes and the six used outputs."""
    d = out[:, :6] - target
    n = d.size
    dout = np.zeros_like(out)
    dout[:, :6] = 2.0 * d / n
    return float((d * d).sum() / n), dout
