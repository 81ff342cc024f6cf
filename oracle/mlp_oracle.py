"""FP64 CPU oracle of the learned warm start (test infrastructure only).

Only ``tests/`` may import this module.  It restates, in NumPy, what the
libqmx warm-start kernels compute, so that their outputs and gradients can be
checked:

* BF16 rounding (round to nearest even on the FP32 bit pattern);
* the 3-layer ReLU MLP with BF16 inputs, weights and stored activations and
  exact (FP64) accumulation, forward and backward, with the same BF16 rounding
  points as the kernels (activations after ReLU, the output gradient, and each
  masked hidden gradient before it enters the next product);
* the su(4) coefficient maps (Gell-Mann basis, tr(l_a l_b) = 2 delta_ab), the
  traceless principal logarithm via ``scipy.linalg.logm`` and the exponential
  via ``scipy.linalg.expm``, and the feature rows of a gate slot.

The learned part has no reference counterpart (the reference has no learned
component); the backward pass is pinned by finite differences of the FP64
forward pass in tests/test_mlp_oracle.py.
"""

from __future__ import annotations

import numpy as np
import scipy.linalg

IN_FEATURES = 50


def bf16_round(x) -> np.ndarray:
    """FP32 -> nearest BF16 (ties to even), returned as float64 values."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def bf16_bits(x) -> np.ndarray:
    """BF16 bit patterns (int16) of FP32 values."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16).view(np.int16)


def bits_to_float(b) -> np.ndarray:
    return (np.asarray(b).view(np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(params: np.ndarray, in_dim: int, width: int, out_dim: int):
    """Flat parameter vector -> (W1, b1, W2, b2, W3, b3), the libqmx layout."""
    o = 0
    out = []
    for shape in ((in_dim, width), (width,), (width, width), (width,), (width, out_dim), (out_dim,)):
        n = int(np.prod(shape))
        out.append(np.asarray(params[o:o + n], dtype=np.float64).reshape(shape))
        o += n
    assert o == params.size
    return out


def forward(x, params, dims, bf16: bool = True):
    """Returns (y, cache).  With bf16=True the rounding points of the kernels
    are applied; with bf16=False everything is exact FP64."""
    in_dim, width, out_dim = dims
    W1, b1, W2, b2, W3, b3 = unpack(params, in_dim, width, out_dim)
    r = bf16_round if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    X = r(x)
    W1r, W2r, W3r = r(W1), r(W2), r(W3)
    # biases stay FP32 in the kernels
    b1, b2, b3 = (np.asarray(b, np.float32).astype(np.float64) if bf16 else b for b in (b1, b2, b3))
    A1 = r(np.maximum(X @ W1r + b1, 0.0))
    A2 = r(np.maximum(A1 @ W2r + b2, 0.0))
    Y = A2 @ W3r + b3
    return Y, (X, W1r, W2r, W3r, A1, A2)


def backward(dy, cache, dims, bf16: bool = True) -> np.ndarray:
    """Gradient of sum(dy * y) w.r.t. the flat parameters."""
    X, W1r, W2r, W3r, A1, A2 = cache
    r = bf16_round if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    dY = r(dy)
    gW3 = A2.T @ dY
    gb3 = dY.sum(axis=0)
    dZ2 = r((dY @ W3r.T) * (A2 > 0))
    gW2 = A1.T @ dZ2
    gb2 = dZ2.sum(axis=0)
    dZ1 = r((dZ2 @ W2r.T) * (A1 > 0))
    gW1 = X.T @ dZ1
    gb1 = dZ1.sum(axis=0)
    return np.concatenate([g.reshape(-1) for g in (gW1, gb1, gW2, gb2, gW3, gb3)])


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v


# ------------------------------------------------------------ su(4) maps --

def gell_mann() -> np.ndarray:
    """The 15 generators in the libqmx order."""
    out = []
    for j, l in ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)):
        s = np.zeros((4, 4), complex)
        s[j, l] = s[l, j] = 1
        a = np.zeros((4, 4), complex)
        a[j, l], a[l, j] = -1j, 1j
        out += [s, a]
    for l in (1, 2, 3):
        d = np.zeros((4, 4), complex)
        d[:l, :l] = np.eye(l)
        d[l, l] = -l
        out.append(d * np.sqrt(2.0 / (l * (l + 1))))
    return np.array(out)


GM = gell_mann()


def coeffs(h: np.ndarray) -> np.ndarray:
    return np.real(np.einsum("kij,ji->k", GM, h)) / 2


def matrix(c) -> np.ndarray:
    return np.einsum("k,kij->ij", np.asarray(c, dtype=np.float64), GM)


def log_coeffs(u: np.ndarray) -> np.ndarray:
    """c with U = e^{i phi} exp(i sum c_k l_k), principal branch."""
    h = -1j * scipy.linalg.logm(u)
    h = (h + h.conj().T) / 2
    h = h - np.trace(h) / 4 * np.eye(4)
    return coeffs(h)


def exp_coeffs(c) -> np.ndarray:
    return scipy.linalg.expm(1j * matrix(c))


def features(g0: np.ndarray, g1: np.ndarray, layer: int, num_layers: int, site: int,
             num_sites: int) -> np.ndarray:
    """Feature row of one two-qubit gate slot (FP64, before BF16 rounding)."""
    f = np.zeros(64)
    f[:15] = log_coeffs(g0.conj().T @ g1)
    f[15:47] = np.stack([g1.real, g1.imag], axis=-1).reshape(-1)
    f[47] = layer / num_layers
    f[48] = site / num_sites
    f[49] = 1.0
    return f


# ---------------------------------------------- trajectory extrapolation --

TRAJ_FEATURES = 35


def traj_features(ga: np.ndarray, gb: np.ndarray, gc: np.ndarray, layer: int, num_layers: int,
                  site: int, num_sites: int, sweep_frac: float):
    """Feature row (FP64 before BF16 rounding) and row scale of one two-qubit
    gate slot from its gates after sweeps s-2, s-1, s: the updates
    d1 = c(Ga^H Gb), d2 = c(Gb^H Gc) in units of |d2| (computed in FP64,
    rounded to FP32 as the kernel keeps it)."""
    d1 = log_coeffs(ga.conj().T @ gb)
    d2 = log_coeffs(gb.conj().T @ gc)
    scale = np.float32(max(np.linalg.norm(d2), 1e-12))
    f = np.zeros(64)
    f[:15] = d1 / float(scale)
    f[15:30] = d2 / float(scale)
    f[30] = np.log10(float(scale)) / 8
    f[31] = sweep_frac
    f[32] = layer / num_layers
    f[33] = site / num_sites
    f[34] = 1.0
    return f, float(scale)


def traj_basis(ga, gb, gc):
    """(scale, e1, e2): e2 = d2 / |d2|, e1 = the part of d1 orthogonal to e2,
    normalised (zero when below 1e-6 of scale)."""
    d1 = log_coeffs(ga.conj().T @ gb)
    d2 = log_coeffs(gb.conj().T @ gc)
    n2 = np.linalg.norm(d2)
    sc = max(n2, 1e-12)
    e2 = d2 / n2 if n2 > 0 else np.zeros(15)
    e1 = d1 - (d1 @ e2) * e2
    n1 = np.linalg.norm(e1)
    e1 = e1 / n1 if n1 > 1e-6 * sc else np.zeros(15)
    return sc, e1, e2


def traj_target(ga, gb, gc, gd) -> np.ndarray:
    """Label: c(Gc^H Gd) / scale on (e2, e1)."""
    sc, e1, e2 = traj_basis(ga, gb, gc)
    t = log_coeffs(gc.conj().T @ gd) / sc
    return np.array([t @ e2, t @ e1])


def traj_jump(ga, gb, g, coef) -> np.ndarray:
    """G exp(i scale (c_0 e2 + c_1 e1) . l) with the basis from (Ga, Gb, G)."""
    sc, e1, e2 = traj_basis(ga, gb, g)
    c = np.asarray(coef, dtype=np.float64)[:2]
    cs = (sc * (c[0] * e2 + c[1] * e1)).astype(np.float32).astype(np.float64)
    return g @ exp_coeffs(cs)
