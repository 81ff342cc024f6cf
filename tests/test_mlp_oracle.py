"""CPU pins of the learned warm start's oracle (oracle/mlp_oracle.py).

* BF16 rounding equals torch's float32 -> bfloat16 conversion (round to
  nearest even), a library routine;
* the FP64 backward pass equals central finite differences of the forward
  pass;
* the su(4) maps: Gell-Mann orthonormality, coefficient round trips, the
  exponential against scipy.linalg.expm of the assembled generator, and the
  logarithm inverting it on the principal branch;
* Adam's first step has the closed form p - lr * g / (|g| + eps).
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.linalg
import torch

from oracle import mlp_oracle as M


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000) * 10 ** rng.uniform(-8, 8, 20000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e38, -1e-30]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    np.testing.assert_array_equal(M.bf16_round(x), ref)
    np.testing.assert_array_equal(M.bits_to_float(M.bf16_bits(x)), ref)


def _tiny(seed=0, in_dim=6, width=8, out_dim=3):
    rng = np.random.default_rng(seed)
    n = in_dim * width + width + width * width + width + width * out_dim + out_dim
    return rng.standard_normal(n) * 0.5, rng.standard_normal((5, in_dim)), (in_dim, width, out_dim)


def test_backward_equals_finite_differences():
    p, x, dims = _tiny()
    rng = np.random.default_rng(1)
    dy = rng.standard_normal((x.shape[0], dims[2]))
    y, cache = M.forward(x, p, dims, bf16=False)
    g = M.backward(dy, cache, dims, bf16=False)
    h = 1e-6
    worst = 0.0
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        yp, _ = M.forward(x, p + e, dims, bf16=False)
        ym, _ = M.forward(x, p - e, dims, bf16=False)
        fd = np.sum(dy * (yp - ym)) / (2 * h)
        worst = max(worst, abs(fd - g[i]))
    assert worst < 1e-6


def test_bf16_forward_is_close_to_fp64():
    p, x, dims = _tiny(2)
    y64, _ = M.forward(x, p, dims, bf16=False)
    y16, _ = M.forward(x, p, dims, bf16=True)
    assert np.linalg.norm(y16 - y64) / np.linalg.norm(y64) < 3e-2


def test_gell_mann_basis():
    g = M.GM
    assert g.shape == (15, 4, 4)
    for a in range(15):
        np.testing.assert_allclose(g[a], g[a].conj().T, atol=0)
        assert abs(np.trace(g[a])) < 1e-15
        for b in range(15):
            assert abs(np.trace(g[a] @ g[b]) - 2.0 * (a == b)) < 1e-14


def test_su4_round_trips():
    rng = np.random.default_rng(3)
    for _ in range(50):
        c = rng.standard_normal(15) * 0.3
        np.testing.assert_allclose(M.coeffs(M.matrix(c)), c, atol=1e-14)
        u = M.exp_coeffs(c)
        np.testing.assert_allclose(u @ u.conj().T, np.eye(4), atol=1e-13)
        np.testing.assert_allclose(u, scipy.linalg.expm(1j * sum(ck * gk for ck, gk in zip(c, M.GM))),
                                   atol=1e-13)
        np.testing.assert_allclose(M.log_coeffs(u), c, atol=1e-10)
        # a global phase does not change the coefficients
        np.testing.assert_allclose(M.log_coeffs(np.exp(0.7j) * u), c, atol=1e-10)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(4)
    p, g = rng.standard_normal(10), rng.standard_normal(10)
    p1, m, v = M.adam(p, g, np.zeros(10), np.zeros(10), 1, lr=0.01)
    np.testing.assert_allclose(p1, p - 0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)


def test_features_layout():
    rng = np.random.default_rng(5)
    q = lambda: np.linalg.qr(rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))[0]
    g0, g1 = q(), q()
    f = M.features(g0, g1, 2, 8, 5, 16)
    # exp(i sum c l) equals G0^H G1 up to a global phase
    ratio = (g0.conj().T @ g1) @ M.exp_coeffs(f[:15]).conj().T
    np.testing.assert_allclose(ratio, ratio[0, 0] * np.eye(4), atol=1e-10)
    assert abs(abs(ratio[0, 0]) - 1) < 1e-10
    np.testing.assert_allclose(f[15:47:2].reshape(4, 4), g1.real)
    np.testing.assert_allclose(f[16:48:2].reshape(4, 4), g1.imag)
    assert f[47] == 2 / 8 and f[48] == 5 / 16 and f[49] == 1.0 and not f[50:].any()
