"""Learned surrogate: a message-passing network that stands in for ten
WCSPH substeps, with the plain step as judge and fallback.

No reference counterpart.  The network (libsphdem mpnn kernels, BF16 MFMA
with FP32 accumulation) reads, per fluid particle, its velocity, its plain
WCSPH acceleration and its density, and, per neighbour within the SPH
support (fluid and wall, from the same hash grids as the plain step),
relative position, distance, relative velocity, a wall flag and the kernel
weight.  Two message-passing rounds of width 64 with 3-layer SiLU MLPs, then
a decoder to six numbers per particle:

    dv = A T out[0:3]
    dx = T v + A T^2/2 out[3:6]                  (T = 10 dt, A = 9.81 m/s^2)

The method without the network is one explicit step of length T from the
current plain acceleration a, out = (a/A, a/A); the network reads a among
its inputs and is trained to do better.  Training pairs (state, state after ten
plain substeps) come from the batched ensemble; the loss is the mean square
of out - target over particles and the six outputs; Adam on FP32 master
weights, re-packed to bf16 fragments after every update.

Rollout (`SurrogateRollout`): one call = grids + the plain density and
force passes (the acceleration the network reads, DEM coupling included) +
the network + the update; the DEM keeps its own substeps with the coupling
force of the call's start.  Every `judge_every` calls the plain path runs
the same ten substeps from the same state and is kept; the surrogate's
result is judged against it (RMS density error of its positions, RMS
velocity error), and when the density error exceeds `density_tol` the
rollout reverts to plain steps for `fallback_calls` calls.
"""
from __future__ import annotations

import math
import os
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .engine import CoupledEngine
from .hashgrid import HashGrid

F = 64
NIN = 16
EIN = 16
SUBSTEPS = 10


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


def init_params(seed: int = 0) -> dict:
    """He-normal weights, zero biases, a small decoder output layer (the
    untrained network starts near out = 0)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in param_shapes():
        if len(shape) == 2:
            p[name] = rng.normal(0.0, math.sqrt(2.0 / shape[0]), shape)
        else:
            p[name] = np.zeros(shape)
    p["dec2_W"][:, 6:] = 0.0
    p["dec2_W"] *= 0.1
    p["hw"] = rng.normal(0.0, 0.5, F)
    return p


def _i16(shape, dev):
    return torch.zeros(shape, dtype=torch.int16, device=dev)


def _f32(shape, dev):
    return torch.zeros(shape, dtype=torch.float32, device=dev)


class _Act:
    """FP32 and bf16 copies of one (N, 64) activation."""

    def __init__(self, n, dev, bf=True, f32=True):
        self.f = _f32((n, F), dev) if f32 else None
        self.b = _i16((n, F), dev) if bf else None


class MPNN:
    """Parameters, packed weights and activations of the network on one GPU."""

    def __init__(self, device="cuda:0", params: dict | None = None, seed: int = 0):
        self.dev = torch.device(device)
        shapes = param_shapes()
        self.shapes = dict(shapes)
        n = sum(int(np.prod(s)) for _, s in shapes)
        self.theta = _f32(n, self.dev)
        self.grad = _f32(n, self.dev)
        self.m = _f32(n, self.dev)
        self.v = _f32(n, self.dev)
        self.step_count = 0
        self.p, self.g, off = {}, {}, 0
        for name, shape in shapes:
            k = int(np.prod(shape))
            self.p[name] = self.theta[off:off + k].view(shape)
            self.g[name] = self.grad[off:off + k].view(shape)
            off += k
        self.load(params if params is not None else init_params(seed))
        self.cap = 0

    # ---------------------------------------------------------------- weights
    def load(self, params: dict) -> None:
        for name, _ in param_shapes():
            self.p[name].copy_(torch.as_tensor(np.asarray(params[name]), dtype=torch.float32))
        self.pack()

    def state_dict(self) -> dict:
        return {k: v.detach().cpu().numpy().astype(np.float64) for k, v in self.p.items()}

    def pack(self) -> None:
        """bf16 B-operand fragments of every weight matrix (W and W^T)."""
        st = _lib.stream_handle()
        if not hasattr(self, "Bp"):
            self.Bp, self.BTp = {}, {}
            for name, shape in self.shapes.items():
                if len(shape) == 2 and name != "hw":
                    nfr = (shape[0] // 16) * 4 * 64 * 4
                    self.Bp[name] = _i16(nfr, self.dev)
                    self.BTp[name] = _i16(nfr, self.dev)
            self.hw_b = _i16(F, self.dev)
        for name in self.Bp:
            _lib.call("sphdem_mpnn_pack", _lib.fptr(self.p[name]), self.shapes[name][0],
                      _lib.fptr(self.Bp[name]), _lib.fptr(self.BTp[name]), st)
        _lib.call("sphdem_mpnn_to_bf16", _lib.fptr(self.p["hw"]), F, _lib.fptr(self.hw_b), st)

    # ---------------------------------------------------------------- buffers
    def reserve(self, n: int) -> None:
        if n <= self.cap:
            return
        d = self.dev
        self.cap = n
        self.nfeat = _i16((n, NIN), d)
        self.fvel = _f32((n, 4), d)
        self.y0, self.y1 = _Act(n, d), _Act(n, d)
        self.h = [_Act(n, d) for _ in range(3)]
        self.P = [_f32((n, F), d) for _ in range(2)]
        self.Q = [_i16((n, F), d) for _ in range(2)]
        self.Qw = [_i16((1, F), d) for _ in range(2)]
        self.s = [_Act(n, d) for _ in range(2)]
        self.a = [_Act(n, d) for _ in range(2)]
        self.t0 = [_Act(n, d) for _ in range(2)]
        self.t1 = [_Act(n, d) for _ in range(2)]
        self.u0, self.u1 = _Act(n, d), _Act(n, d)
        self.out = _f32((n, F), d)
        # backward scratch
        self.dA, self.dB, self.dh = _f32((n, F), d), _f32((n, F), d), _f32((n, F), d)
        self.da, self.ds = _f32((n, F), d), _f32((n, F), d)
        self.dP, self.dQ = _f32((n, F), d), _f32((n, F), d)
        self.gQw = _f32(F, d)
        self.dout = _f32((n, F), d)

    # ---------------------------------------------------------------- kernels
    def _fwd(self, x0, k0, x1, k1, rows, W, b, act, Yf=None, Yb=None, resid=None):
        """Y = resid + act(X W + b) with X = [x0 | x1] (bf16), W by name; with
        the activation Yf receives the pre-activation and Yb the activation."""
        _lib.call("sphdem_mpnn_dense_fwd", _lib.fptr(x0), k0, _lib.fptr(x1), k1, rows,
                  _lib.fptr(self.Bp[W]), _lib.fptr(self.p[b] if b else None), int(act),
                  _lib.fptr(resid), _lib.fptr(Yf), _lib.fptr(Yb), _lib.stream_handle())

    def _layer(self, x0, k0, x1, k1, rows, W, b, act, out: _Act, resid=None):
        self._fwd(x0, k0, x1, k1, rows, W, b, act, out.f, out.b, resid)

    def _bwd(self, dY, mask, x0, k0, x1, k1, rows, W, b, dX0=None, dX1=None, accumulate=0):
        """dW += X^T dZ, db += sum dZ, dX (+)= dZ W^T with dZ = dY * act'(mask)
        (mask: the layer's pre-activation, or None without activation)."""
        st = _lib.stream_handle()
        _lib.call("sphdem_mpnn_dense_bwd_w", _lib.fptr(dY), _lib.fptr(mask), _lib.fptr(x0), k0,
                  _lib.fptr(x1), k1, rows, _lib.fptr(self.g[W]),
                  _lib.fptr(self.g[b] if b else None), st)
        if dX0 is not None:
            _lib.call("sphdem_mpnn_dense_bwd_x", _lib.fptr(dY), _lib.fptr(mask), rows,
                      _lib.fptr(self.BTp[W]), _lib.fptr(dX0), k0, _lib.fptr(dX1), k1,
                      int(accumulate), st)

    def _edge_args(self, params, gs, gb, wvel, r):
        return (params, gs.struct(), gb.struct(), _lib.fptr(self.fvel), _lib.fptr(wvel),
                _lib.fptr(self.P[r]), _lib.fptr(self.Q[r]), _lib.fptr(self.Qw[r]),
                _lib.fptr(self.Bp[f"r{r}_C"]), _lib.fptr(self.p[f"r{r}_c1"]),
                _lib.fptr(self.Bp[f"r{r}_W2"]))

    def forward(self, params, P, gs: HashGrid, gb: HashGrid, wvel=None) -> torch.Tensor:
        """Network output (N, 64) in the fluid grid's sorted order; columns
        0..5 are used.  P: fluid struct whose vel, acc and rho are current;
        gs/gb: fluid and wall grids built from the current positions; wvel:
        wall velocity (nb, 4) in the wall grid's sorted order, or None (at rest)."""
        n = gs.n
        self.reserve(n)
        st = _lib.stream_handle()
        _lib.call("sphdem_mpnn_features", params, _lib.sph_view(P), gs.struct(),
                  _lib.fptr(self.nfeat), _lib.fptr(self.fvel), st)
        self._layer(self.nfeat, NIN, None, 0, n, "enc0_W", "enc0_b", True, self.y0)
        self._layer(self.y0.b, F, None, 0, n, "enc1_W", "enc1_b", True, self.y1)
        self._layer(self.y1.b, F, None, 0, n, "enc2_W", "enc2_b", False, self.h[0])
        for r in range(2):
            hin = self.h[r]
            self._fwd(hin.b, F, None, 0, n, f"r{r}_Wa", None, False, Yf=self.P[r])
            self._fwd(hin.b, F, None, 0, n, f"r{r}_Wb", None, False, Yb=self.Q[r])
            self._fwd(self.hw_b, F, None, 0, 1, f"r{r}_Wb", None, False, Yb=self.Qw[r])
            _lib.call("sphdem_mpnn_edge_fwd", *self._edge_args(params, gs, gb, wvel, r),
                      _lib.fptr(self.p[f"r{r}_c2"]), _lib.fptr(self.s[r].f),
                      _lib.fptr(self.s[r].b), st)
            self._layer(self.s[r].b, F, None, 0, n, f"r{r}_W3", f"r{r}_c3", False, self.a[r])
            self._layer(hin.b, F, self.a[r].b, F, n, f"r{r}_n0_W", f"r{r}_n0_b", True, self.t0[r])
            self._layer(self.t0[r].b, F, None, 0, n, f"r{r}_n1_W", f"r{r}_n1_b", True, self.t1[r])
            self._layer(self.t1[r].b, F, None, 0, n, f"r{r}_n2_W", f"r{r}_n2_b", False,
                        self.h[r + 1], resid=hin.f)
        self._layer(self.h[2].b, F, None, 0, n, "dec0_W", "dec0_b", True, self.u0)
        self._layer(self.u0.b, F, None, 0, n, "dec1_W", "dec1_b", True, self.u1)
        self._fwd(self.u1.b, F, None, 0, n, "dec2_W", "dec2_b", False, Yf=self.out)
        return self.out[:n]

    def backward(self, params, gs: HashGrid, gb: HashGrid, dout: torch.Tensor, wvel=None) -> None:
        """Accumulate into self.grad the gradient of sum(out * dout), for the
        inputs of the last forward."""
        n = gs.n
        st = _lib.stream_handle()
        self._bwd(dout, None, self.u1.b, F, None, 0, n, "dec2_W", "dec2_b", self.dA)
        self._bwd(self.dA, self.u1.f, self.u0.b, F, None, 0, n, "dec1_W", "dec1_b", self.dB)
        self._bwd(self.dB, self.u0.f, self.h[2].b, F, None, 0, n, "dec0_W", "dec0_b", self.dh)
        for r in (1, 0):
            hin = self.h[r]
            # h_{r+1} = h_r + n2(n1(n0([h_r, a_r]))): dh reaches h_r directly and through n0
            self._bwd(self.dh, None, self.t1[r].b, F, None, 0, n, f"r{r}_n2_W", f"r{r}_n2_b",
                      self.dA)
            self._bwd(self.dA, self.t1[r].f, self.t0[r].b, F, None, 0, n, f"r{r}_n1_W",
                      f"r{r}_n1_b", self.dB)
            self.da[:n].zero_()
            self._bwd(self.dB, self.t0[r].f, hin.b, F, self.a[r].b, F, n, f"r{r}_n0_W",
                      f"r{r}_n0_b", self.dh, self.da, accumulate=1)
            # a = s W3 + c3
            self._bwd(self.da, None, self.s[r].b, F, None, 0, n, f"r{r}_W3", f"r{r}_c3", self.ds)
            self.dQ[:n].zero_()
            self.gQw.zero_()
            _lib.call("sphdem_mpnn_edge_bwd", *self._edge_args(params, gs, gb, wvel, r),
                      _lib.fptr(self.BTp[f"r{r}_W2"]), _lib.fptr(self.p[f"r{r}_c2"]),
                      _lib.fptr(self.ds), _lib.fptr(self.dP), _lib.fptr(self.dQ),
                      _lib.fptr(self.g[f"r{r}_C"]), _lib.fptr(self.g[f"r{r}_c1"]),
                      _lib.fptr(self.g[f"r{r}_W2"]), _lib.fptr(self.g[f"r{r}_c2"]),
                      _lib.fptr(self.gQw), st)
            self._bwd(self.dP, None, hin.b, F, None, 0, n, f"r{r}_Wa", None, self.dh,
                      accumulate=1)
            self._bwd(self.dQ, None, hin.b, F, None, 0, n, f"r{r}_Wb", None, self.dh,
                      accumulate=1)
            # Q_wall = hw Wb (one row)
            self._bwd(self.gQw, None, self.hw_b, F, None, 0, 1, f"r{r}_Wb", None, self.g["hw"],
                      accumulate=1)
        self._bwd(self.dh, None, self.y1.b, F, None, 0, n, "enc2_W", "enc2_b", self.dA)
        self._bwd(self.dA, self.y1.f, self.y0.b, F, None, 0, n, "enc1_W", "enc1_b", self.dB)
        self._bwd(self.dB, self.y0.f, self.nfeat, NIN, None, 0, n, "enc0_W", "enc0_b")

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8) -> None:
        self.step_count += 1
        _lib.call("sphdem_adam", _lib.fptr(self.theta), _lib.fptr(self.grad), _lib.fptr(self.m),
                  _lib.fptr(self.v), self.theta.numel(), float(lr), float(beta1), float(beta2),
                  float(eps), self.step_count, _lib.stream_handle())
        self.pack()


# ---------------------------------------------------------------- training data
@dataclass
class Pair:
    """One training pair of one scene: the state, its plain acceleration and
    density, and the state ten plain substeps later (original index order)."""

    x0: torch.Tensor
    v0: torch.Tensor
    a0: torch.Tensor
    rho0: torch.Tensor
    x1: torch.Tensor
    v1: torch.Tensor
    m: torch.Tensor | None = None      # particle masses
    flt: torch.Tensor | None = None    # Shepard field held at x0
    x2: torch.Tensor | None = None     # the state twenty plain substeps after x0
    v2: torch.Tensor | None = None


def collect_pairs(ensemble, snapshots: int = 8, interval: int = 40, warm: int = 20,
                  log=None, chain: bool = False) -> list:
    """Advance a batched ensemble (ensemble.EnsembleEngine) with the plain
    step and cut (state, state + 10 substeps) pairs for every instance at
    `snapshots` times, `interval` steps apart, after `warm` steps; with
    `chain` also the state 20 substeps on (for pushforward training)."""
    ens = ensemble
    ens.run(warm)
    pairs = []
    for s in range(snapshots):
        x0, v0 = ens.P.pos.clone(), ens.P.vel.clone()
        ens.step()                                   # computes a(x0), rho(x0)
        a0, rho0, flt = ens.P.acc.clone(), ens.P.rho.clone(), ens.P.flt.clone()
        ens.run(SUBSTEPS - 1)
        x1, v1 = ens.P.pos.clone(), ens.P.vel.clone()
        x2 = v2 = None
        if chain:
            ens.run(SUBSTEPS)
            x2, v2 = ens.P.pos.clone(), ens.P.vel.clone()
        for k in range(ens.n):
            sl = ens._slice("fluid", k)
            pairs.append(Pair(x0[sl], v0[sl], a0[sl], rho0[sl], x1[sl], v1[sl],
                              ens.P.m[sl], flt[sl], None if x2 is None else x2[sl],
                              None if v2 is None else v2[sl]))
        if log:
            log(f"[surrogate] snapshot {s + 1}/{snapshots} at step {ens.step_index}")
        ens.run(max(interval - SUBSTEPS * (2 if chain else 1), 0))
    ens.check()
    return pairs


class Trainer:
    """Adam on pairs of one scene layout (all pairs share the wall set).

    Training noise (GNS-style): `position_noise` (m) perturbs x0; the plain
    density and acceleration are then recomputed at the perturbed positions
    (fused SPH passes, fluid and wall), so the network sees the state a
    drifting rollout produces, while the target still points at the clean
    state ten substeps later.  `velocity_noise` (m/s) perturbs v0 likewise."""

    def __init__(self, net: MPNN, solv, walls, device="cuda:0", velocity_noise: float = 0.0,
                 seed: int = 0, position_noise: float = 0.0, pushforward: float = 0.0):
        from .simulation import solver_params
        from .struct import SPHptl
        self.net, self.solv = net, solv
        self.dev = torch.device(device)
        gsl = int(solv.grid_slice)
        self.gs = HashGrid(gsl, gsl, gsl, device=self.dev)
        self.gb = HashGrid(gsl, gsl, gsl, device=self.dev).build(walls.pos.to(self.dev),
                                                                 solv.support)
        self.params = solver_params(solv)
        self.T = SUBSTEPS * float(solv.dt)
        self.loss = torch.zeros(1, dtype=torch.float64, device=self.dev)
        self._P = SPHptl()
        self.target = None
        # training noise on the input velocity (m/s): the target keeps the
        # clean future, so the network learns to pull a perturbed state back
        self.noise = float(velocity_noise)
        self.pos_noise = float(position_noise)
        self.pushforward = float(pushforward)
        self._rng = np.random.default_rng(int(seed) + 1)
        self._gen = torch.Generator(device=self.dev)
        self._gen.manual_seed(int(seed))
        if self.pos_noise > 0 or self.pushforward > 0:
            from .kernels import Workspace
            self.B = walls.to(self.dev)
            self.ws = Workspace(self.dev)
            self.ws.reserve(1, self.B.n, 1)
            self._Pn = SPHptl()
            self._an = self._rn = None

    def _perturbed(self, pair: Pair, v0: torch.Tensor):
        """x0 + noise, and the plain density / acceleration recomputed there."""
        x0 = pair.x0 + self.pos_noise * torch.randn(pair.x0.shape, device=self.dev,
                                                    generator=self._gen)
        return (x0, *self._plain_at(pair, x0, v0))

    def _plain_at(self, pair: Pair, x0: torch.Tensor, v0: torch.Tensor):
        """The plain density and acceleration of the fluid at (x0, v0) (the
        pair's masses and Shepard field; fluid and wall, no DEM)."""
        n = pair.x0.shape[0]
        if self._an is None or self._an.shape[0] < n:
            self._an, self._rn = _f32((n, 3), self.dev), _f32(n, self.dev)
            self._pres, self._pgf = _f32(n, self.dev), _f32((n, 3), self.dev)
            self._por, self._accd = torch.ones(n, device=self.dev), _f32((n, 3), self.dev)
        self.ws.reserve(n, self.B.n, 1)
        Pn = self._Pn
        Pn.pos, Pn.vel, Pn.m, Pn.flt = x0, v0, pair.m, pair.flt
        Pn.rho, Pn.acc = self._rn[:n], self._an[:n]
        Pn.rho.copy_(pair.rho0)
        Pn.pres, Pn.pgf, Pn.porosity, Pn.acc_dem = (self._pres[:n], self._pgf[:n], self._por[:n],
                                                     self._accd[:n])
        self.gs.build(x0, self.solv.support, weights=pair.m)
        self.gb.ensure_weights(self.B.m)
        _lib.call("sphdem_surrogate_prep", self.params, _lib.sph_view(Pn), _lib.bnd_view(self.B),
                  None, self.gs.struct(), self.gb.struct(), None, 1, self.ws.struct(),
                  _lib.stream_handle())
        return Pn.acc, Pn.rho

    def _prepare(self, pair: Pair, noise: bool = False) -> int:
        n = pair.x0.shape[0]
        v0 = pair.v0
        if noise and self.noise > 0:
            v0 = v0 + self.noise * torch.randn(v0.shape, device=self.dev, generator=self._gen)
        if noise and self.pos_noise > 0 and pair.m is not None:
            x0, a0, rho0 = self._perturbed(pair, v0)
        else:
            x0, a0, rho0 = pair.x0, pair.a0, pair.rho0
            self.gs.build(x0, self.solv.support)
        self._P.pos, self._P.vel, self._P.acc, self._P.rho = x0, v0, a0, rho0
        if self.target is None or self.target.shape[0] < n:
            self.target = _f32((n, 6), self.dev)
        _lib.call("sphdem_mpnn_targets", self.gs.struct(), _lib.fptr(x0),
                  _lib.fptr(v0), _lib.fptr(a0), _lib.fptr(pair.x1), _lib.fptr(pair.v1),
                  self.T, _lib.fptr(self.target), _lib.stream_handle())
        return n

    def _pushforward(self, pair: Pair) -> int:
        """Inputs from the network's own prediction of x0 + 10 substeps and
        targets from the plain state 20 substeps after x0: trains the network
        on the states its rollouts visit (no gradient through the first call)."""
        n = self._prepare(pair)
        out = self.net.forward(self.params, self._P, self.gs, self.gb)
        from .struct import SPHptl
        Pt = SPHptl(pos=pair.x0.clone(), vel=pair.v0.clone(), acc=pair.a0)
        _lib.call("sphdem_mpnn_apply", self.params, _lib.sph_view(Pt), self.gs.struct(),
                  _lib.fptr(out), self.T, _lib.stream_handle())
        x1, v1 = Pt.pos, Pt.vel
        a1, rho1 = self._plain_at(pair, x1, v1)
        self._P.pos, self._P.vel, self._P.acc, self._P.rho = x1, v1, a1, rho1
        _lib.call("sphdem_mpnn_targets", self.gs.struct(), _lib.fptr(x1), _lib.fptr(v1),
                  _lib.fptr(a1), _lib.fptr(pair.x2), _lib.fptr(pair.v2), self.T,
                  _lib.fptr(self.target), _lib.stream_handle())
        return n

    def step(self, pair: Pair, lr: float) -> None:
        """One Adam step on one pair; the loss is added to self.loss."""
        if self.pushforward > 0 and pair.x2 is not None and \
                self._rng.random() < self.pushforward:
            n = self._pushforward(pair)
        else:
            n = self._prepare(pair, noise=True)
        out = self.net.forward(self.params, self._P, self.gs, self.gb)
        _lib.call("sphdem_mpnn_loss", _lib.fptr(out), _lib.fptr(self.target[:n]), n,
                  _lib.fptr(self.loss), _lib.fptr(self.net.dout[:n]), _lib.stream_handle())
        self.net.grad.zero_()
        self.net.backward(self.params, self.gs, self.gb, self.net.dout[:n])
        self.net.adam(lr)

    def evaluate(self, pairs: list) -> dict:
        """Mean loss over `pairs` of the network and of the method without it
        (one explicit step of length T from the plain acceleration)."""
        tot, base = 0.0, 0.0
        st = _lib.stream_handle()
        for pr in pairs:
            n = self._prepare(pr)
            out = self.net.forward(self.params, self._P, self.gs, self.gb)
            self.loss.zero_()
            _lib.call("sphdem_mpnn_loss", _lib.fptr(out), _lib.fptr(self.target[:n]), n,
                      _lib.fptr(self.loss), _lib.fptr(self.net.dout[:n]), st)
            tot += float(self.loss.item())
            _lib.call("sphdem_mpnn_explicit_out", self.gs.struct(), _lib.fptr(self._P.acc),
                      _lib.fptr(self.net.dA[:n]), st)
            self.loss.zero_()
            _lib.call("sphdem_mpnn_loss", _lib.fptr(self.net.dA[:n]), _lib.fptr(self.target[:n]),
                      n, _lib.fptr(self.loss), _lib.fptr(self.net.dB[:n]), st)
            base += float(self.loss.item())
        return {"loss": tot / len(pairs), "loss_without_network": base / len(pairs)}

    def train(self, pairs: list, iters: int, lr: float = 1e-3, lr_end: float = 1e-4,
              seed: int = 0, log=None, log_every: int = 50) -> list:
        """Adam over randomly drawn pairs with a cosine learning-rate decay;
        returns the mean training loss of each logging window."""
        rng = np.random.default_rng(seed)
        hist = []
        self.loss.zero_()
        for it in range(iters):
            f = 0.5 * (1 + math.cos(math.pi * it / max(iters, 1)))
            self.step(pairs[int(rng.integers(len(pairs)))], lr_end + (lr - lr_end) * f)
            if (it + 1) % log_every == 0 or it + 1 == iters:
                k = (it % log_every) + 1
                hist.append(float(self.loss.item()) / k)
                self.loss.zero_()
                if log:
                    log(f"[surrogate] iter {it + 1}/{iters}: loss {hist[-1]:.4e}")
        return hist


def save_params(net: MPNN, path: str, meta: dict | None = None) -> None:
    arrays = {k: v.astype(np.float32) for k, v in net.state_dict().items()}
    if meta:
        arrays["_meta"] = np.frombuffer(repr(sorted(meta.items())).encode(), dtype=np.uint8)
    np.savez(path, **arrays)


def load_params(path: str) -> dict:
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k].astype(np.float64) for k in z.files if not k.startswith("_")}


# ---------------------------------------------------------------- rollout
@dataclass
class RolloutStats:
    calls: int = 0
    surrogate_calls: int = 0
    plain_calls: int = 0
    judged: int = 0
    rejected: int = 0
    density_err: list = field(default_factory=list)
    velocity_err: list = field(default_factory=list)
    unhealthy: int = 0
    flag_bits: int = 0
    restarted_plain: bool = False


class SurrogateRollout:
    """Advance a coupled scene by surrogate calls of ten substeps, judged by
    the plain step every `judge_every` calls (see the module docstring)."""

    def __init__(self, solv, P, B, D, W, net: MPNN, judge_every: int = 50,
                 density_tol: float = 0.01, fallback_calls: int = 50,
                 fast_math: bool = True, velocity_tol: float = 0.02):
        from .simulation import solver_params
        self.eng = CoupledEngine(solv, P, B, D, W, fast_math=fast_math)
        self.net, self.solv = net, solv
        self.judge_every, self.density_tol = int(judge_every), float(density_tol)
        self.velocity_tol = float(velocity_tol)
        self.health_tol = 0.01
        wall = self.eng.B.pos
        self._box_lo = wall.min(0).values - float(solv.support)
        self._box_hi = wall.max(0).values + float(solv.support)
        self._box_hi[2] += 1.0
        self.fallback_calls = int(fallback_calls)
        self.T = SUBSTEPS * float(solv.dt)
        self.stats = RolloutStats()
        self._fallback_left = 0
        self._ckpt = None
        self._graphs = None
        e = self.eng
        gsl = int(solv.grid_slice)
        self._gj = HashGrid(gsl, gsl, gsl, device=e.device)
        self._params_sph = solver_params(solv)
        self.net.reserve(e.P.n)
        # DEM substeps: candidates from the call's DEM grid within this skin (m)
        self.skin = 0.01
        nd = e.D.n if e.coupled else 0
        self._dem_scratch = torch.zeros(
            max(int(_lib.load().sphdem_dem_substeps_scratch(nd, 4)), 1), dtype=torch.uint8,
            device=e.device)
        # the DEM substeps' own flags (bit 0 of the engine's belongs to its
        # overflow guard)
        self._flags = torch.zeros(4, dtype=torch.int32, device=e.device)

    # ---- one call of each kind
    def _enqueue_surrogate(self, step: int, mark=None) -> None:
        """Enqueue one call; `mark(stage)` (if given) is called after each
        stage: grids, prep, dem, network, apply."""
        e = self.eng
        s = self.solv
        mark = mark or (lambda stage: None)
        st = _lib.stream_handle()
        e.gs.build(e.P.pos, s.support, weights=e.P.m)
        dem = _lib.dem_view(e.D) if e.coupled else None
        if e.coupled:
            e.gd.build(e.D.pos, e.r_dem)
        mark("grids")
        _lib.call("sphdem_surrogate_prep", e.params, _lib.sph_view(e.P), _lib.bnd_view(e.B),
                  dem, e.gs.struct(), e.gb.struct(), e.gd.struct() if e.coupled else None,
                  int(step), e.ws.struct(), st)
        mark("prep")
        if e.coupled:
            ws = e.ws.struct()
            ws.flags = self._flags.data_ptr()
            _lib.call("sphdem_dem_substeps", e.params, dem, _lib.dembnd_view(e.W), e.gd.struct(),
                      e.gw.struct(), SUBSTEPS, self.skin, _lib.fptr(self._dem_scratch), ws, st)
        mark("dem")
        out = self.net.forward(e.params, e.P, e.gs, e.gb)
        mark("network")
        _lib.call("sphdem_mpnn_apply", e.params, _lib.sph_view(e.P), e.gs.struct(),
                  _lib.fptr(out), self.T, st)
        mark("apply")

    def call_costs(self, reps: int = 20) -> dict:
        """Per-call cost in ms: each stage of an eager call (HIP events on
        the launch stream between the stages), a captured call replayed, and
        ten plain substeps replayed, every repetition from the same state
        (put back between them, outside the timed region)."""
        saved = self._state()
        k0 = self.eng.step_index

        def restore():
            for s, name, t in saved:
                getattr(s, name).copy_(t)
            self.eng.step_index = k0
            torch.cuda.synchronize()
        names = ("grids", "prep", "dem", "network", "apply")
        acc = {n: 0.0 for n in names}
        for r in range(reps):
            restore()
            evs = [torch.cuda.Event(enable_timing=True)]
            evs[0].record()
            marks = []

            def mark(stage):
                ev = torch.cuda.Event(enable_timing=True)
                ev.record()
                marks.append((stage, ev))
            self._enqueue_surrogate(SUBSTEPS, mark)
            torch.cuda.synchronize()
            prev = evs[0]
            for stage, ev in marks:
                if r > 0:
                    acc[stage] += prev.elapsed_time(ev)
                prev = ev
        out = {f"{n}_ms": round(acc[n] / max(reps - 1, 1), 4) for n in names}
        out["eager_call_ms"] = round(sum(acc.values()) / max(reps - 1, 1), 4)

        def replay_ms(fn):
            tot = 0.0
            for r in range(reps + 1):
                restore()
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record()
                fn()
                b.record()
                torch.cuda.synchronize()
                if r > 0:
                    tot += a.elapsed_time(b)
            return tot / reps
        if self._graphs is not None:
            out["graph_call_ms"] = round(replay_ms(lambda: self._graphs[False].replay()), 4)
        out["plain_call_ms"] = round(replay_ms(lambda: self.eng.run(SUBSTEPS)), 4)
        call = out.get("graph_call_ms", out["eager_call_ms"])
        out["speedup_if_every_call_kept"] = round(out["plain_call_ms"] / call, 3)
        restore()
        self._flags.zero_()
        return out

    def capture(self) -> None:
        """hipGraphs of a surrogate call with and without the Shepard refresh
        (and the plain engine's step graphs)."""
        side = torch.cuda.Stream(device=self.eng.device)
        graphs = {}
        for shep in (True, False):
            g = torch.cuda.CUDAGraph()
            torch.cuda.synchronize()
            with torch.cuda.stream(side):
                with torch.cuda.graph(g, stream=side):
                    self._enqueue_surrogate(0 if shep else SUBSTEPS)
            graphs[shep] = g
        torch.cuda.synchronize()
        self._graphs = graphs
        self.eng.capture()

    def surrogate_call(self) -> None:
        e = self.eng
        k = e.step_index
        if self._graphs is not None:
            self._graphs[e.shepard_due(k)].replay()
        else:
            self._enqueue_surrogate(k)
        e.step_index = k + SUBSTEPS

    def plain_call(self) -> None:
        self.eng.run(SUBSTEPS)

    # ---- judge
    def _state(self):
        e = self.eng
        return [(s, name, getattr(s, name).clone()) for s in (e.P, e.B, e.D)
                if s is not None for name, _ in s.FIELDS if getattr(s, name) is not None]

    def _density(self, pos: torch.Tensor, out: torch.Tensor, P=None) -> None:
        """Shepard-normalised density of the fluid at `pos` (FP32) into out,
        with the masses of P (default: this rollout's fluid) and the prior
        densities held in out."""
        from .struct import SPHptl
        e, s = self.eng, self.solv
        src = e.P if P is None else P
        Pj = SPHptl(**{n: getattr(src, n) for n in SPHptl.field_names()})
        Pj.pos, Pj.flt, Pj.rho = pos, torch.empty_like(e.P.flt), out
        self._gj.build(pos, s.support, weights=e.P.m)
        st = _lib.stream_handle()
        _lib.call("sphdem_shepard_sph", self._params_sph, _lib.sph_view(Pj), _lib.bnd_view(e.B),
                  self._gj.struct(), e.gb.struct(), st)
        _lib.call("sphdem_density_sph", self._params_sph, _lib.sph_view(Pj), _lib.bnd_view(e.B),
                  self._gj.struct(), e.gb.struct(), st)

    def judged_call(self) -> dict:
        """Surrogate and plain from the same state S; the plain result is
        kept.  Returns the surrogate's errors against it and whether S (the
        state the unjudged calls since the last checkpoint produced) and the
        plain result are sound (see sound())."""
        e = self.eng
        s_ok = self.sound(self.health())
        saved = self._state()
        k = e.step_index
        self.surrogate_call()
        x_s, v_s = e.P.pos.clone(), e.P.vel.clone()
        for s, name, t in saved:
            getattr(s, name).copy_(t)
        e.step_index = k
        self.plain_call()
        # the Shepard sums of both read the same prior densities m_j / rho_j
        rho_s, rho_p = e.P.rho.clone(), e.P.rho.clone()
        self._density(x_s, rho_s)
        self._density(e.P.pos, rho_p)
        rho0 = float(self.solv.rho0)
        d_err = float(torch.sqrt((((rho_s - rho_p) / rho0).double() ** 2).mean()))
        dv = (v_s - e.P.vel).double()
        v_err = float(torch.sqrt((dv ** 2).sum(1).mean())
                      / torch.sqrt((e.P.vel.double() ** 2).sum(1).mean()).clamp_min(1e-30))
        ok = s_ok and not self._dem_flags() and self.sound(self.health())
        if not ok:
            self.stats.unhealthy += 1
        return {"density_err": d_err, "velocity_err": v_err, "state_ok": ok}

    def _dem_flags(self) -> bool:
        """True (and cleared) when a DEM substep overflowed a contact buffer
        or a DEM outran the skin since the last look: those calls are invalid."""
        f = self._flags
        v = int(f[0].item())
        if v & 7:
            f[0] = 0
            self.stats.flag_bits |= v
        return bool(v & 7)

    def sane(self, h: dict) -> bool:
        """Absolute sanity of a state: finite densities and speeds, every
        fluid particle inside the tank's box (1 m of head room above it), no
        speed above a third of the sound speed."""
        return (math.isfinite(h["density_rms"]) and h["inside"] and math.isfinite(h["vmax"])
                and h["vmax"] < float(self.solv.c0) / 3.0)

    def sound(self, h: dict) -> bool:
        """sane(), and the Shepard density's deviation from rho0 no worse
        than the last checkpoint's (a state the plain step produced and
        the judge accepted) by more than health_tol in RMS and 5 % at most
        (with 15 % allowed in any case: the plain step's own splash reaches
        ~20 % where the front hits the far wall)."""
        ref = self._ref_health
        return (self.sane(h) and h["density_rms"] <= ref["density_rms"] + self.health_tol
                and h["density_max"] <= max(0.15, ref["density_max"] + 0.05))

    def healthy(self) -> bool:
        """The current state is sound() against the last checkpoint."""
        ok = self.sound(self.health())
        if not ok:
            self.stats.unhealthy += 1
        return ok

    def health(self, P=None) -> dict:
        """Density deviation (RMS and maximum, Shepard density at the
        positions), box and speed figures of this rollout's fluid or of
        another state of the same scene (e.g. the plain run's)."""
        P = self.eng.P if P is None else P
        rho = P.rho.clone()
        self._density(P.pos, rho, P)
        d = rho.double() / float(self.solv.rho0) - 1.0
        dev = float(torch.sqrt((d ** 2).mean()))
        dmax = float(d.abs().max())
        pos = P.pos
        inside = bool(((pos >= self._box_lo) & (pos <= self._box_hi)).all())
        vmax = float(P.vel.abs().max())
        h = {"density_rms": dev, "density_max": dmax, "inside": inside, "vmax": vmax}
        h["ok"] = self.sane(h)
        return h

    # ---- loop
    def _checkpoint(self) -> None:
        self._ckpt = (self._state(), self.eng.step_index, self.stats.calls)
        self._ref_health = self.health()
        self._sur_since = 0

    def _restore(self) -> int:
        saved, k, calls = self._ckpt
        for s, name, t in saved:
            getattr(s, name).copy_(t)
        self.eng.step_index = k
        return self.stats.calls - calls

    def run(self, n_calls: int) -> RolloutStats:
        """n_calls calls of ten substeps.  A judged call is accepted when
        its density error is within density_tol, its velocity error within
        velocity_tol and the state it started from and its plain result are
        sound; otherwise the state goes back to the last checkpoint, the
        calls since then are redone with the plain step, and the plain step
        runs fallback_calls more calls.  Each judged call ends in a new
        checkpoint (its plain result)."""
        st = self.stats
        if getattr(self, "_ckpt", None) is None:
            self._start = self._save_start()
            self._checkpoint()
        for _ in range(int(n_calls)):
            st.calls += 1
            if self._fallback_left > 0:
                self._fallback_left -= 1
                self.plain_call()
                st.plain_calls += 1
            elif st.calls % self.judge_every == 0:
                r = self.judged_call()
                st.judged += 1
                st.plain_calls += 1
                st.density_err.append(r["density_err"])
                st.velocity_err.append(r["velocity_err"])
                if not (r["state_ok"] and r["density_err"] <= self.density_tol
                        and r["velocity_err"] <= self.velocity_tol):
                    st.rejected += 1
                    redo = self._restore()
                    for _ in range(redo):
                        self.plain_call()
                    st.plain_calls += redo
                    st.surrogate_calls -= self._sur_since
                    self._fallback_left = self.fallback_calls
                self._checkpoint()
            else:
                self.surrogate_call()
                st.surrogate_calls += 1
                self._sur_since += 1
        return st

    def finish(self) -> None:
        """Validate the calls since the last judged one: if a DEM substep
        flagged them or the state is not sound, revert and redo them with
        the plain step.  If the end state is still not sane (a checkpoint
        the plain step itself cannot carry on from), the whole run is redone
        with the plain step from its start."""
        st = self.stats
        if self._sur_since and (self._dem_flags() or not self.healthy()):
            st.rejected += 1
            redo = self._restore()
            for _ in range(redo):
                self.plain_call()
            st.plain_calls += redo
            st.surrogate_calls -= self._sur_since
            self._checkpoint()
        if not self.sane(self.health()):
            self._restart_plain()
            st.restarted_plain = True
            for _ in range(st.calls):
                self.plain_call()
            st.plain_calls += st.calls
            st.surrogate_calls = 0
            self._checkpoint()

    def _save_start(self):
        return self._state(), self.eng.step_index

    def _restart_plain(self) -> None:
        saved, k = self._start
        for s, name, t in saved:
            getattr(s, name).copy_(t)
        self.eng.step_index = k


# ---------------------------------------------------------------- command line
def train_on_ensemble(n_inst: int = 16, seed: int = 20263, snapshots: int = 8,
                      interval: int = 100, warm: int = 50, iters: int = 2000,
                      lr: float = 1e-3, device="cuda:0", log=None, init_seed: int = 1,
                      velocity_noise: float = 0.0, position_noise: float = 0.0,
                      pushforward: float = 0.0):
    """Collect pairs from a C3-style ensemble (dx = 0.01) and train a new
    network; returns (net, report).  The last snapshot's pairs are held out."""
    from .ensemble import EnsembleEngine
    from .workloads import c3
    wl = c3(device=device, seed=seed, n_inst=n_inst)
    t0 = time.time()
    ens = EnsembleEngine(wl.solv, wl.scenes, device=device, fast_math=True)
    pairs = collect_pairs(ens, snapshots=snapshots, interval=interval, warm=warm, log=log,
                          chain=pushforward > 0)
    torch.cuda.synchronize()
    t_data = time.time() - t0
    del ens
    train, held = pairs[:-n_inst], pairs[-n_inst:]
    net = MPNN(device, seed=init_seed)
    tr = Trainer(net, wl.solv, wl.scenes[0][1], device=device, velocity_noise=velocity_noise,
                 position_noise=position_noise, pushforward=pushforward)
    t0 = time.time()
    hist = tr.train(train, iters, lr=lr, log=log)
    torch.cuda.synchronize()
    t_train = time.time() - t0
    rep = {"instances": n_inst, "pairs": len(pairs), "iters": iters, "data_s": t_data,
           "velocity_noise": velocity_noise, "position_noise": position_noise,
           "pushforward": pushforward,
           "train_s": t_train, "loss_history": hist, "held_out": tr.evaluate(held),
           "train_eval": tr.evaluate(train[:n_inst])}
    return net, rep


WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights",
                       "surrogate_c3.npz")


def rollout_c1(net: MPNN, sim_time: float = 0.5, judge_every: int = 50,
               density_tol: float = 0.01, device="cuda:0", log=None,
               velocity_tol: float = 0.02) -> dict:
    """C1 dam break (100,000 fluid, 200 DEM): plain GPU path and surrogate
    rollout over the same simulated time, both replaying hipGraphs."""
    from .workloads import c1
    wl = c1(device=device)
    dt = float(wl.solv.dt)
    calls = int(round(sim_time / (SUBSTEPS * dt)))
    structs = [s.to(device) for s in wl.structs]
    plain = CoupledEngine(wl.solv, *[s.clone() for s in structs], fast_math=True)
    plain.run(2)
    plain.capture()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain.run(calls * SUBSTEPS - 2)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    plain.check()
    ro = SurrogateRollout(wl.solv, *[s.clone() for s in structs], net, judge_every=judge_every,
                          density_tol=density_tol, velocity_tol=velocity_tol)
    ro.plain_call()                                 # first call plain (eager warm-up)
    ro.capture()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = ro.run(calls - 1)
    ro.finish()
    torch.cuda.synchronize()
    t_sur = time.perf_counter() - t0
    ro.eng.check()
    # the plain path's and the rollout's simulated time per wall second
    sim_plain = (calls * SUBSTEPS - 2) * dt / t_plain
    sim_sur = (calls - 1) * SUBSTEPS * dt / t_sur
    d_p, d_s = plain.diagnostics(), ro.eng.diagnostics()
    out = {"sim_time": calls * SUBSTEPS * dt, "calls": calls, "plain_s": t_plain,
           "surrogate_s": t_sur, "plain_sim_per_wall": sim_plain,
           "surrogate_sim_per_wall": sim_sur, "speedup": sim_sur / sim_plain,
           "surrogate_calls": st.surrogate_calls, "plain_calls": st.plain_calls,
           "judged": st.judged, "rejected": st.rejected, "unhealthy": st.unhealthy,
           "restarted_plain": st.restarted_plain,
           "dem_flag_bits": st.flag_bits,
           "density_err_max": max(st.density_err) if st.density_err else None,
           "density_err_mean": float(np.mean(st.density_err)) if st.density_err else None,
           "velocity_err_max": max(st.velocity_err) if st.velocity_err else None,
           "velocity_err_mean": float(np.mean(st.velocity_err)) if st.velocity_err else None,
           "final_plain": d_p, "final_surrogate": d_s, "health_end": ro.health(),
           "health_plain_end": ro.health(plain.P)}
    # after the timed runs and their figures: the per-call cost breakdown
    out["call_costs"] = ro.call_costs()
    if log:
        log(f"[surrogate] rollout: {out}")
    return out


def main(argv=None) -> int:
    import argparse
    import json
    import sys

    ap = argparse.ArgumentParser(description="train the surrogate / roll it out on C1")
    ap.add_argument("--train", action="store_true")
    ap.add_argument("--instances", type=int, default=16)
    ap.add_argument("--iters", type=int, default=20000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--snapshots", type=int, default=20)
    ap.add_argument("--noise", type=float, default=0.0, help="training velocity noise, m/s")
    ap.add_argument("--pos-noise", type=float, default=0.0, help="training position noise, m")
    ap.add_argument("--pushforward", type=float, default=0.0,
                    help="fraction of steps trained on the network's own first call")
    ap.add_argument("--interval", type=int, default=500)
    ap.add_argument("--out", default=None, help="where to save trained weights (.npz)")
    ap.add_argument("--weights", default=None, help="weights to load instead of training")
    ap.add_argument("--rollout", action="store_true")
    ap.add_argument("--sim-time", type=float, default=0.5)
    args = ap.parse_args(argv)

    def log(m):
        print(m, file=sys.stderr, flush=True)

    rep = {}
    if args.weights:
        net = MPNN(params=load_params(args.weights))
    elif args.train:
        net, rep["train"] = train_on_ensemble(args.instances, iters=args.iters, lr=args.lr,
                                              snapshots=args.snapshots, interval=args.interval,
                                              velocity_noise=args.noise,
                                              position_noise=args.pos_noise,
                                              pushforward=args.pushforward, log=log)
        if args.out:
            save_params(net, args.out, {"instances": args.instances, "iters": args.iters})
    else:
        ap.error("give --weights or --train")
    if args.rollout:
        rep["rollout"] = rollout_c1(net, sim_time=args.sim_time, log=log)
    print(json.dumps(rep), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
