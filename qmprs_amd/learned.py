"""Learned warm start for the environment sweeps (no reference counterpart).

A small network proposes starting gates for the sweeps of
``Sequential._optimize_unitary_layers`` (sequential.py:509-541).  Data,
training and inference all run on the GPU through libqmx:

* training pairs come from the plain method itself: for each state the Ran
  gates G0 (after layer extraction), the gates G1 after one sweep and the
  gates GS after S sweeps; the features of a two-qubit gate slot are
  ``qmx_gate_features(G0, G1)`` (the su(4) coefficients of G0^H G1, G1 itself,
  its position) and the label is ``qmx_gate_log(G1, GS)`` (15 su(4)
  coefficients);
* the model is a 3-layer ReLU MLP with BF16 weights and FP32 accumulation
  (``qmx_mlp_forward`` / ``qmx_mlp_backward`` on MFMA, ``qmx_mlp_adam``);
* at inference the proposal G1 exp(i sum c_k l_k) replaces G1 for a state
  only if its fidelity is at least that of G1; otherwise the state keeps the
  plain gates.  The sweeps then continue as usual and decide the final fidelity.
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


def plain_trajectory(statevectors, bond_dimension: int, num_layers: int, num_sweeps: int,
                     sweep_dtype=torch.complex64):
    """Plain method on a batch: returns (psi, target, widths, used, G0, G1, GS,
    fidelity trace (B, S + 1))."""
    p = engine.encode_prepare(statevectors, bond_dimension, num_layers, sweep_dtype=sweep_dtype)
    psi, tgt, gates, widths, used = p.psi, p.target, p.gates, p.widths, p.layers_used
    g0 = gates.clone()
    r1 = engine.sweep(tgt, gates, widths, used, 1, psi_ref=psi)
    g1 = gates.clone()
    rs = engine.sweep(tgt, gates, widths, used, num_sweeps - 1, psi_ref=psi)
    fid = torch.cat([r1.fidelity[:, :1], rs.fidelity], dim=1).abs()
    return psi, tgt, widths, used, g0, g1, gates, fid


def make_training_set(statevectors, bond_dimension: int, num_layers: int, num_sweeps: int,
                      chunk: int = 64) -> TrainingSet:
    feats, labels, weights = [], [], []
    for i in range(0, len(statevectors), chunk):
        _, _, widths, used, g0, g1, gs, _ = plain_trajectory(statevectors[i:i + chunk], bond_dimension,
                                                             num_layers, num_sweeps)
        f, w = gate_features(g0, g1, widths, used)
        feats.append(f)
        weights.append(w)
        labels.append(gate_log(g1, gs, widths, used))
    return TrainingSet(torch.cat(feats), torch.cat(labels), torch.cat(weights))


def train(net: WarmStartMLP, data: TrainingSet, epochs: int = 20, batch: int = 512, lr: float = 1e-3,
          seed: int = 0) -> list[float]:
    """Minibatch Adam on the weighted MSE; returns the mean loss per epoch."""
    live = torch.nonzero(data.weights > 0).flatten()
    g = torch.Generator(device="cpu").manual_seed(seed)
    history = []
    for _ in range(epochs):
        order = live[torch.randperm(live.numel(), generator=g).to(live.device)]
        total, count = 0.0, 0
        losses = []
        for s in range(0, order.numel() - batch + 1, batch):
            idx = order[s:s + batch]
            x = data.features[idx].contiguous()
            t = data.labels[idx].contiguous()
            y = net.forward(x)
            loss, dy = mse_loss(y, t, None)
            grads = net.backward(x, dy)
            net.adam(grads, lr)
            losses.append(loss)
        if losses:
            total = torch.cat(losses).mean().item()
            count = len(losses)
        history.append(total if count else float("nan"))
    return history


def evaluate_loss(net: WarmStartMLP, data: TrainingSet) -> float:
    y = net.forward(data.features)
    loss, _ = mse_loss(y, data.labels, data.weights)
    return float(loss.item())


# ------------------------------------------------------------ inference ---

@dataclass
class WarmResult:
    gates: torch.Tensor
    widths: torch.Tensor
    layers_used: torch.Tensor
    fidelity: torch.Tensor      # (B,) final |<psi|C|0>|
    sweeps: torch.Tensor        # (B,) sweeps run in total
    accepted: torch.Tensor      # (B,) 1 where the proposal replaced the plain gates
    fid_plain_1: torch.Tensor   # (B,) fidelity after the first (plain) sweep
    fid_proposal: torch.Tensor  # (B,) fidelity of the proposal


def encode_warm(statevectors, bond_dimension: int, num_layers: int, num_sweeps: int, net: WarmStartMLP,
                goal: torch.Tensor | None = None, sweep_dtype=torch.complex64) -> WarmResult:
    """prepare_state with a learned warm start after the first sweep.

    Sweeps stop per state once |<psi|C|0>| >= goal[b] (when goal is given) or
    after num_sweeps in total."""
    if num_sweeps < 1:
        raise ValueError("the warm start needs num_sweeps >= 1")
    p = engine.encode_prepare(statevectors, bond_dimension, num_layers, sweep_dtype=sweep_dtype)
    psi, tgt, gates, widths, used = p.psi, p.target, p.gates, p.widths, p.layers_used
    g0 = gates.clone()
    r1 = engine.sweep(tgt, gates, widths, used, 1, psi_ref=psi)
    f1 = r1.fidelity[:, -1].abs()
    feat, _ = gate_features(g0, gates, widths, used)
    coef = net.forward(feat)
    prop = gates.clone()
    gate_expmap(prop, widths, used, coef)
    rp = engine.sweep(tgt, prop, widths, used, 0, psi_ref=psi)
    fp = rp.fidelity[:, -1].abs()
    accept = fp >= f1
    gates = torch.where(accept.view(-1, 1, 1, 1, 1), prop, gates).contiguous()
    rest = num_sweeps - 1
    if rest > 0:
        res = engine.sweep(tgt, gates, widths, used, rest, psi_ref=psi, goal=goal)
        done = res.sweeps_done
        final = res.fidelity[:, -1].abs()
    else:
        done = torch.zeros_like(used)
        final = torch.where(accept, fp, f1)
    return WarmResult(gates, widths, used, final, done + 1, accept.to(torch.int32), f1, fp)
