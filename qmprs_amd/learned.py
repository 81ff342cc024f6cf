"""Learned warm start for the environment sweeps (no reference counterpart).

The plain sweeps of ``Sequential._optimize_unitary_layers``
(sequential.py:509-541) converge linearly: after the first few sweeps every
gate keeps moving along a slowly turning direction, and the fidelity creeps
up for tens of sweeps.  A small network, trained on trajectories of the plain
method itself, predicts from a gate slot's last two updates where that gate
will be K sweeps later; the encoder jumps there every few sweeps and keeps a
state's jump only if its fidelity does not drop.  The plain sweeps remain the
fallback and decide the final fidelity.  Everything runs on the GPU through
libqmx:

* features (``qmx_traj_features``): the su(4) coefficients of the updates
  G_{s-2}^H G_{s-1} and G_{s-1}^H G_s of each two-qubit gate slot, in units
  of the last update's norm, plus that norm's log, the sweep index and the
  slot's position; label (``qmx_traj_targets``): c(G_s^H G_{s+K}) in the same
  units, projected on that basis;
* the model is a 3-layer ReLU MLP with BF16 weights and FP32 accumulation
  (``qmx_mlp_forward`` / ``qmx_mlp_backward`` on MFMA, ``qmx_mlp_adam``);
* the network predicts two numbers per slot, the coefficients of the jump on
  an orthonormal basis (e2, e1) of the slot's last two updates (so that every
  jump moves along the sweep's own directions: predictions of all 15 su(4)
  coefficients, with errors in the steep directions, lowered the fidelity);
  the jump is ``qmx_traj_jump``: G_s exp(i |d2| (c_0 e2 + c_1 e1) . l).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from qmprs_amd import _lib, engine

IN_DIM = 64        # feature row width (50 used, see qmx_gate_features)
OUT_DIM = 16       # 15 su(4) coefficients + 1 pad column
VALID_OUT = 15


def _call(name, *args):
    lib, h = _lib.load(), _lib.handle()
    _lib.check(getattr(lib, name)(h, *args), h)


def _f(x: float) -> ctypes.c_float:
    return ctypes.c_float(float(x))


# ------------------------------------------------------------------ MLP ---

class WarmStartMLP:
    """3-layer ReLU MLP, FP32 master parameters + BF16 shadow on the GPU."""

    def __init__(self, width: int = 512, in_dim: int = IN_DIM, out_dim: int = OUT_DIM, seed: int = 0,
                 device=None) -> None:
        dev = device or _lib.require_gpu()
        self.in_dim, self.width, self.out_dim = in_dim, width, out_dim
        n = _lib.load().qmx_mlp_param_count(in_dim, width, out_dim)
        if n == 0:
            raise ValueError("invalid MLP dimensions")
        g = torch.Generator().manual_seed(seed)
        parts = []
        for fan_in, fan_out, gain in ((in_dim, width, 2.0), (width, width, 2.0), (width, out_dim, 0.01)):
            parts.append(torch.randn(fan_in * fan_out, generator=g, dtype=torch.float32)
                         * float(np.sqrt(gain / fan_in)))
            parts.append(torch.zeros(fan_out, dtype=torch.float32))
        p = torch.cat(parts)
        assert p.numel() == n
        self.params = p.to(dev)
        self.params_bf16 = torch.empty(n, dtype=torch.int16, device=dev)
        self.sync_bf16()
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self._ws = None

    @property
    def count(self) -> int:
        return self.params.numel()

    def sync_bf16(self) -> None:
        _call("qmx_mlp_to_bf16", self.count, _lib.ptr(self.params), _lib.ptr(self.params_bf16),
              _lib.stream_ptr())

    def _workspace(self, rows: int) -> torch.Tensor:
        need = _lib.load().qmx_mlp_workspace(rows, self.in_dim, self.width, self.out_dim)
        if self._ws is None or self._ws.numel() < need:
            self._ws = _lib.workspace(need, self.params.device)
        return self._ws

    def forward(self, x_bf16: torch.Tensor) -> torch.Tensor:
        """x: (rows, in_dim) BF16 bits (int16) -> (rows, out_dim) FP32."""
        rows = x_bf16.shape[0]
        ws = self._workspace(rows)
        y = torch.empty((rows, self.out_dim), dtype=torch.float32, device=self.params.device)
        _call("qmx_mlp_forward", rows, self.in_dim, self.width, self.out_dim, _lib.ptr(x_bf16),
              _lib.ptr(self.params), _lib.ptr(self.params_bf16), _lib.ptr(y), _lib.ptr(ws), ws.numel(),
              _lib.stream_ptr())
        return y

    def backward(self, x_bf16: torch.Tensor, dy: torch.Tensor) -> torch.Tensor:
        """Gradient of sum(dy * y) w.r.t. the parameters (after ``forward(x)``)."""
        rows = x_bf16.shape[0]
        ws = self._workspace(rows)
        grads = torch.empty_like(self.params)
        _call("qmx_mlp_backward", rows, self.in_dim, self.width, self.out_dim, _lib.ptr(x_bf16),
              _lib.ptr(self.params), _lib.ptr(self.params_bf16), _lib.ptr(dy.contiguous()),
              _lib.ptr(grads), _lib.ptr(ws), ws.numel(), _lib.stream_ptr())
        return grads

    def adam(self, grads: torch.Tensor, lr: float, beta1=0.9, beta2=0.999, eps=1e-8) -> None:
        self.step += 1
        _call("qmx_mlp_adam", self.count, _lib.ptr(self.params), _lib.ptr(self.params_bf16),
              _lib.ptr(grads), _lib.ptr(self.m), _lib.ptr(self.v), _f(lr), _f(beta1), _f(beta2), _f(eps),
              self.step, _lib.stream_ptr())

    def state_dict(self) -> dict:
        return {"params": self.params.cpu().numpy(), "dims": np.array([self.in_dim, self.width, self.out_dim])}

    @classmethod
    def from_state_dict(cls, sd: dict, device=None) -> "WarmStartMLP":
        in_dim, width, out_dim = (int(v) for v in sd["dims"])
        net = cls(width, in_dim, out_dim, device=device)
        net.params.copy_(torch.from_numpy(np.asarray(sd["params"], dtype=np.float32)))
        net.sync_bf16()
        return net


def mse_loss(y: torch.Tensor, t: torch.Tensor, w: torch.Tensor | None, valid: int = VALID_OUT):
    """(loss, dy) of the weighted MSE over the first ``valid`` columns."""
    rows, ld = y.shape
    nw = float(w.sum().item()) if w is not None else float(rows)
    inv = 1.0 / max(nw * valid, 1.0)
    dy = torch.empty_like(y)
    loss = torch.empty(1, dtype=torch.float32, device=y.device)
    ws = _lib.workspace(_lib.load().qmx_mse_workspace(), y.device)
    _call("qmx_mse_loss", rows, ld, valid, _lib.ptr(y), _lib.ptr(t), _lib.ptr(w), _f(inv), _lib.ptr(dy),
          _lib.ptr(loss), _lib.ptr(ws), ws.numel(), _lib.stream_ptr())
    return loss, dy


# ----------------------------------------------------------- gate maps ---

def gate_features(g_a, g_b, widths, layers_used):
    """(rows, IN_DIM) BF16 features and (rows,) live-row weights."""
    b, nl, n = widths.shape
    rows = b * nl * (n - 1)
    feat = torch.empty((rows, IN_DIM), dtype=torch.int16, device=g_a.device)
    w = torch.empty(rows, dtype=torch.float32, device=g_a.device)
    _call("qmx_gate_features", n, b, nl, _lib.ptr(layers_used), _lib.ptr(g_a), _lib.ptr(g_b),
          _lib.ptr(widths), _lib.ptr(feat), IN_DIM, _lib.ptr(w), _lib.stream_ptr())
    return feat, w


def gate_log(g_a, g_b, widths, layers_used, ld: int = OUT_DIM):
    """(rows, ld) FP32 su(4) coefficients of Ga^H Gb per gate slot."""
    b, nl, n = widths.shape
    rows = b * nl * (n - 1)
    out = torch.empty((rows, ld), dtype=torch.float32, device=g_a.device)
    _call("qmx_gate_log", n, b, nl, _lib.ptr(layers_used), _lib.ptr(g_a), _lib.ptr(g_b), _lib.ptr(widths),
          _lib.ptr(out), ld, _lib.stream_ptr())
    return out


def gate_expmap(gates, widths, layers_used, coef, skip=None) -> None:
    """gates <- gates exp(i sum_k c_k l_k) in place on live two-qubit slots."""
    b, nl, n = widths.shape
    _call("qmx_gate_expmap", n, b, nl, _lib.ptr(layers_used), _lib.ptr(coef.contiguous()), coef.shape[1],
          _lib.ptr(gates), _lib.ptr(widths), _lib.ptr(skip), _lib.stream_ptr())


# ------------------------------------------------- data and training ---

@dataclass
class TrainingSet:
    features: torch.Tensor   # (rows, IN_DIM) BF16 bits
    labels: torch.Tensor     # (rows, OUT_DIM) FP32
    weights: torch.Tensor    # (rows,) 1 for live gate slots


REF_SWEEPS = 30     # sweep count of the reference path (feature sweep_frac = s / REF_SWEEPS)
TRAJ_K = 6          # sweeps ahead the network predicts
JUMP_FIRST = 6      # the first jump after sweep 6, ...
JUMP_EVERY = 4      # ... then after every 4 further sweeps (sweeps to goal at N=20: 6/6 20.97,
                    # 4/4 19.78, 6/4 19.55, 5/5 20.24, 8/8 23.26; 6/4 fastest per step)


def traj_features(ga, gb, gc, widths, layers_used, sweep_frac: float):
    """(rows, IN_DIM) BF16 features, (rows,) live-row weights, (rows,) row scales."""
    b, nl, n = widths.shape
    rows = b * nl * (n - 1)
    dev = ga.device
    feat = torch.empty((rows, IN_DIM), dtype=torch.int16, device=dev)
    w = torch.empty(rows, dtype=torch.float32, device=dev)
    sc = torch.empty(rows, dtype=torch.float32, device=dev)
    _call("qmx_traj_features", n, b, nl, _lib.ptr(layers_used), _lib.ptr(ga), _lib.ptr(gb), _lib.ptr(gc),
          _lib.ptr(widths), _f(sweep_frac), _lib.ptr(feat), IN_DIM, _lib.ptr(w), _lib.ptr(sc), _lib.stream_ptr())
    return feat, w, sc


TRAJ_OUT = 2        # the network's outputs: coefficients on (e2, e1)


def traj_targets(ga, gb, gc, gd, widths, layers_used, ld: int = OUT_DIM):
    """(rows, ld) FP32 labels: the update Gc -> Gd in units of |d2|, on the
    orthonormal basis (e2, e1) of the slot's last two updates."""
    b, nl, n = widths.shape
    rows = b * nl * (n - 1)
    out = torch.empty((rows, ld), dtype=torch.float32, device=gc.device)
    _call("qmx_traj_targets", n, b, nl, _lib.ptr(layers_used), _lib.ptr(ga), _lib.ptr(gb), _lib.ptr(gc),
          _lib.ptr(gd), _lib.ptr(widths), _lib.ptr(out), ld, _lib.stream_ptr())
    return out


def traj_jump(ga, gb, gates, widths, layers_used, coef, skip=None) -> None:
    """gates <- gates exp(i |d2| (c_0 e2 + c_1 e1) . l) in place on live
    two-qubit slots (basis from ga, gb, gates)."""
    b, nl, n = widths.shape
    _call("qmx_traj_jump", n, b, nl, _lib.ptr(layers_used), _lib.ptr(ga), _lib.ptr(gb), _lib.ptr(coef.contiguous()),
          coef.shape[1], _lib.ptr(gates), _lib.ptr(widths), _lib.ptr(skip), _lib.stream_ptr())


def plain_snapshots(statevectors, bond_dimension: int, num_layers: int, num_sweeps: int,
                    sweep_dtype=torch.complex64):
    """The plain method on a batch, one sweep at a time (the same gates as one
    call of num_sweeps sweeps): (prepared, [gates after sweep 0..S], fidelity
    (B, S + 1))."""
    p = engine.encode_prepare(statevectors, bond_dimension, num_layers, sweep_dtype=sweep_dtype)
    gates = p.gates
    snaps = [gates.clone()]
    fids = []
    for _ in range(num_sweeps):
        r = engine.sweep(p.target, gates, p.widths, p.layers_used, 1, psi_ref=p.psi)
        fids.append(r.fidelity[:, :1])
        snaps.append(gates.clone())
    fids.append(r.fidelity[:, -1:])
    return p, snaps, torch.cat(fids, dim=1).abs()


def make_trajectory_set(statevectors, bond_dimension: int, num_layers: int, num_sweeps: int = REF_SWEEPS,
                        sweeps=tuple(range(4, 25, 2)), k: int = TRAJ_K, chunk: int = 64) -> TrainingSet:
    """Training pairs from plain trajectories: for each listed sweep s the
    features of (G_{s-2}, G_{s-1}, G_s) and the label G_s^H G_{s+k}."""
    feats, labels, weights = [], [], []
    for i in range(0, len(statevectors), chunk):
        p, snaps, _ = plain_snapshots(statevectors[i:i + chunk], bond_dimension, num_layers, num_sweeps)
        for s in sweeps:
            f, w, sc = traj_features(snaps[s - 2], snaps[s - 1], snaps[s], p.widths, p.layers_used,
                                     s / REF_SWEEPS)
            feats.append(f)
            weights.append(w)
            labels.append(traj_targets(snaps[s - 2], snaps[s - 1], snaps[s], snaps[min(s + k, num_sweeps)],
                                       p.widths, p.layers_used))
    return TrainingSet(torch.cat(feats), torch.cat(labels), torch.cat(weights))


def train(net: WarmStartMLP, data: TrainingSet, epochs: int = 20, batch: int = 512, lr: float = 1e-3,
          seed: int = 0, valid: int = VALID_OUT) -> list[float]:
    """Minibatch Adam on the weighted MSE over the first ``valid`` output
    columns; returns the mean loss per epoch."""
    live = torch.nonzero(data.weights > 0).flatten()
    g = torch.Generator(device="cpu").manual_seed(seed)
    history = []
    for _ in range(epochs):
        order = live[torch.randperm(live.numel(), generator=g).to(live.device)]
        losses = []
        for s in range(0, order.numel() - batch + 1, batch):
            idx = order[s:s + batch]
            x = data.features[idx].contiguous()
            t = data.labels[idx].contiguous()
            y = net.forward(x)
            loss, dy = mse_loss(y, t, None, valid)
            grads = net.backward(x, dy)
            net.adam(grads, lr)
            losses.append(loss)
        history.append(torch.cat(losses).mean().item() if losses else float("nan"))
    return history


def evaluate_loss(net: WarmStartMLP, data: TrainingSet, valid: int = VALID_OUT) -> float:
    y = net.forward(data.features)
    loss, _ = mse_loss(y, data.labels, data.weights, valid)
    return float(loss.item())


def zero_loss(data: TrainingSet, valid: int = VALID_OUT) -> float:
    """The weighted MSE of predicting no jump (all-zero output)."""
    loss, _ = mse_loss(torch.zeros_like(data.labels), data.labels, data.weights, valid)
    return float(loss.item())


# ------------------------------------------------------------ inference ---

@dataclass
class WarmResult:
    gates: torch.Tensor
    widths: torch.Tensor
    layers_used: torch.Tensor
    fidelity: torch.Tensor      # (B,) final |<psi|C|0>|
    sweeps: torch.Tensor        # (B,) sweeps run in total
    accepted: torch.Tensor      # (B,) jumps kept
    tried: int                  # jump points reached


def encode_warm(statevectors, bond_dimension: int, num_layers: int, num_sweeps: int, net: WarmStartMLP,
                goal: torch.Tensor | None = None, sweep_dtype=torch.complex64, first: int = JUMP_FIRST,
                every: int = JUMP_EVERY) -> WarmResult:
    """prepare_state with learned jumps: plain sweeps, and after sweeps
    first, first + every, ... the network's jump for every state still short of
    its goal, kept where it does not lower the state's fidelity.  Sweeps stop
    per state at |<psi|C|0>| >= goal[b] (when goal is given) or after
    num_sweeps in total."""
    if num_sweeps < 1:
        raise ValueError("the warm start needs num_sweeps >= 1")
    p = engine.encode_prepare(statevectors, bond_dimension, num_layers, sweep_dtype=sweep_dtype)
    psi, tgt, gates, widths, used = p.psi, p.target, p.gates, p.widths, p.layers_used
    b = gates.shape[0]
    done = torch.zeros(b, dtype=torch.int32, device=gates.device)
    accepted = torch.zeros(b, dtype=torch.int32, device=gates.device)

    def run(k, snapshots=None):
        nonlocal done
        r = engine.sweep(tgt, gates, widths, used, k, psi_ref=psi, goal=goal, snapshots=snapshots)
        done = done + (r.sweeps_done if goal is not None else k)
        return r.fidelity[:, -1].abs()

    s, tried, fid = 0, 0, None
    j = first
    # the gates after sweeps j - 2 and j - 1 come from the same call's snapshots
    snaps = torch.empty((2,) + tuple(gates.shape), dtype=gates.dtype, device=gates.device)
    while j <= num_sweeps and j >= 2:
        snaps[0].copy_(gates)
        snaps[1].copy_(gates)
        k = j - s
        fid = run(k, snapshots=(k - 2, snaps))
        ga, gb = snaps[0], snaps[1]
        s = j
        active = (fid < goal) if goal is not None else torch.ones_like(fid, dtype=torch.bool)
        if not bool(active.any()):
            break
        tried += 1
        feat, _, _ = traj_features(ga, gb, gates, widths, used, s / REF_SWEEPS)
        coef = net.forward(feat)
        prop = gates.clone()
        traj_jump(ga, gb, prop, widths, used, coef, skip=(~active).to(torch.int32).contiguous())
        _, fp = engine.circuit_state(prop, widths, used, dtype=sweep_dtype, psi_ref=psi)
        keep = active & (fp.abs() >= fid)
        gates.copy_(torch.where(keep.view(-1, 1, 1, 1, 1), prop, gates))
        accepted += keep.to(torch.int32)
        j += every
    if num_sweeps > s:
        fid = run(num_sweeps - s)
    return WarmResult(gates, widths, used, fid, done, accepted, tried)
