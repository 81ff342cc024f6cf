"""GPU tests of the learned warm start, through the C-ABI.

Kernels against the CPU oracle (oracle/mlp_oracle.py), which applies the
same BF16 rounding points and accumulates exactly:
  MLP forward           relative error < 5e-3 (FP32 vs exact accumulation;
                        a BF16 activation may round to the neighbouring value)
  MLP backward          per-parameter-block relative error < 2e-2
  Adam                  FP32 vs FP64 update, rtol 1e-5
  su(4) log / exp maps  1e-6 (FP32 coefficients) / 1e-12 (FP64 gates)
  trajectory maps        features to BF16 rounding, targets 1e-5, jumps 1e-10
Then: training on the GPU lowers the loss, and the learned jumps on held-out
states beat the plain method in sweeps to the plain fidelity goal.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import mlp_oracle as M
from oracle import qmprs_oracle as O

pytestmark = pytest.mark.gpu


def _net(width, seed=0):
    from qmprs_amd import learned
    return learned.WarmStartMLP(width=width, seed=seed)


def _rand_params(net, seed):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(net.count) * 0.2).astype(np.float32)
    net.params.copy_(torch.from_numpy(p))
    net.sync_bf16()
    return p.astype(np.float64)


def _rand_x(rows, in_dim, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((rows, in_dim)).astype(np.float32)
    x[:, 50:] = 0.0
    bits = M.bf16_bits(x)
    return torch.from_numpy(bits).cuda(), M.bits_to_float(bits)


@pytest.mark.parametrize("rows,width", [(300, 128), (512, 512), (17, 64)])
def test_mlp_forward_matches_oracle(gpu, rows, width):
    net = _net(width)
    p = _rand_params(net, 1)
    xg, x = _rand_x(rows, net.in_dim, 2)
    y = net.forward(xg).cpu().numpy().astype(np.float64)
    ref, _ = M.forward(x, p, (net.in_dim, width, net.out_dim))
    err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    print(f"rows {rows} width {width} forward rel err {err:.2e}")
    assert err < 5e-3


@pytest.mark.parametrize("rows,width", [(256, 128), (512, 512)])
def test_mlp_backward_matches_oracle(gpu, rows, width):
    """The gradient check: GPU backward kernels vs the oracle's FP64 backprop."""
    net = _net(width)
    p = _rand_params(net, 3)
    xg, x = _rand_x(rows, net.in_dim, 4)
    rng = np.random.default_rng(5)
    dy = rng.standard_normal((rows, net.out_dim)).astype(np.float32)
    net.forward(xg)
    g = net.backward(xg, torch.from_numpy(dy).cuda()).cpu().numpy().astype(np.float64)
    dims = (net.in_dim, width, net.out_dim)
    _, cache = M.forward(x, p, dims)
    ref = M.backward(dy.astype(np.float64), cache, dims)
    names = ("W1", "b1", "W2", "b2", "W3", "b3")
    for name, a, b in zip(names, M.unpack(g, *dims), M.unpack(ref, *dims)):
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(f"{name}: rel err {err:.2e}")
        assert err < 2e-2, name


def test_adam_matches_oracle(gpu):
    net = _net(64)
    p = _rand_params(net, 6)
    rng = np.random.default_rng(7)
    m = np.zeros(net.count)
    v = np.zeros(net.count)
    for step in range(1, 4):
        g = rng.standard_normal(net.count).astype(np.float32)
        net.adam(torch.from_numpy(g).cuda(), 1e-2)
        p, m, v = M.adam(p, g.astype(np.float64), m, v, step, 1e-2)
    np.testing.assert_allclose(net.params.cpu().numpy(), p, rtol=1e-5, atol=1e-7)
    np.testing.assert_array_equal(net.params_bf16.cpu().numpy(), M.bf16_bits(net.params.cpu().numpy()))


def _random_gate_arrays(b, nl, n, seed):
    rng = np.random.default_rng(seed)

    def haar(d):
        q, r = np.linalg.qr(rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d)))
        return q * (np.diag(r) / np.abs(np.diag(r)))

    g = np.zeros((b, nl, n, 4, 4), complex)
    w = np.full((b, nl, n), 2, np.int32)
    w[:, :, n - 1] = 1
    w[:, 0, 2] = 1
    for idx in np.ndindex(b, nl, n):
        d = 4 if w[idx] == 2 else 2
        g[idx][:d, :d] = haar(d)
    used = np.array([nl, nl - 1][:b] + [nl] * max(0, b - 2), np.int32)
    return g, w, used


def test_gate_maps_match_oracle(gpu):
    from qmprs_amd import learned
    b, nl, n = 2, 3, 6
    ga, w, used = _random_gate_arrays(b, nl, n, 8)
    gb, _, _ = _random_gate_arrays(b, nl, n, 9)
    rng = np.random.default_rng(10)
    # G_b = G_a exp(i H) with moderate H, so the principal logarithm is H
    cs = rng.standard_normal((b, nl, n, 15)) * 0.4
    for idx in np.ndindex(b, nl, n):
        if w[idx] == 2:
            gb[idx] = ga[idx] @ M.exp_coeffs(cs[idx])
    t = lambda a: torch.from_numpy(a).cuda()
    GA, GB, W, U = t(ga), t(gb), t(w), t(used)
    logs = learned.gate_log(GA, GB, W, U).cpu().numpy()
    feats, wt = learned.gate_features(GA, GB, W, U)
    feats = M.bits_to_float(feats.cpu().numpy())
    wt = wt.cpu().numpy()
    row = 0
    for bb in range(b):
        for l in range(nl):
            for site in range(n - 1):
                live = l < used[bb] and w[bb, l, site] == 2
                assert wt[row] == (1.0 if live else 0.0)
                if live:
                    np.testing.assert_allclose(logs[row, :15], cs[bb, l, site], atol=1e-6)
                    ref = M.bf16_round(M.features(ga[bb, l, site], gb[bb, l, site], l, nl, site, n)
                                       .astype(np.float32))
                    np.testing.assert_allclose(feats[row], ref, rtol=1e-2, atol=2e-3)
                else:
                    assert not logs[row].any() and not feats[row].any()
                row += 1
    # exp map: G <- G exp(i sum c l), FP64
    coef = (rng.standard_normal((row, 16)) * 0.3).astype(np.float32)
    G = t(ga.copy())
    learned.gate_expmap(G, W, U, torch.from_numpy(coef).cuda())
    got = G.cpu().numpy()
    row = 0
    for bb in range(b):
        for l in range(nl):
            for site in range(n - 1):
                live = l < used[bb] and w[bb, l, site] == 2
                ref = ga[bb, l, site] @ M.exp_coeffs(coef[row, :15].astype(np.float64)) if live else ga[bb, l, site]
                np.testing.assert_allclose(got[bb, l, site], ref, atol=1e-12)
                row += 1


def test_training_lowers_loss(gpu):
    """A synthetic regression (labels a fixed smooth function of the
    features) is learned by the GPU training loop."""
    from qmprs_amd import learned
    rows = 4096
    rng = np.random.default_rng(11)
    x = rng.standard_normal((rows, learned.IN_DIM)).astype(np.float32)
    x[:, 50:] = 0
    a = rng.standard_normal((learned.IN_DIM, 15)).astype(np.float32) * 0.2
    y = np.zeros((rows, learned.OUT_DIM), np.float32)
    y[:, :15] = np.tanh(x @ a)
    data = learned.TrainingSet(torch.from_numpy(M.bf16_bits(x)).cuda(), torch.from_numpy(y).cuda(),
                               torch.ones(rows, device="cuda"))
    net = learned.WarmStartMLP(width=256, seed=1)
    before = learned.evaluate_loss(net, data)
    hist = learned.train(net, data, epochs=15, batch=256, lr=2e-3)
    after = learned.evaluate_loss(net, data)
    print(f"loss {before:.4e} -> {after:.4e}; epochs {['%.3e' % h for h in hist]}")
    assert after < 0.1 * before


def test_traj_maps_match_oracle(gpu):
    """Trajectory features, targets and jumps against the oracle: features to
    BF16 rounding, targets to FP32 rounding, jumps to 1e-10 (FP64 gates)."""
    from qmprs_amd import learned
    b, nl, n = 2, 3, 6
    ga, w, used = _random_gate_arrays(b, nl, n, 18)
    gb, gc, gd = ga.copy(), ga.copy(), ga.copy()
    rng = np.random.default_rng(19)
    for idx in np.ndindex(b, nl, n):
        if w[idx] == 2:
            gb[idx] = ga[idx] @ M.exp_coeffs(rng.standard_normal(15) * 0.05)
            gc[idx] = gb[idx] @ M.exp_coeffs(rng.standard_normal(15) * 0.03)
            gd[idx] = gc[idx] @ M.exp_coeffs(rng.standard_normal(15) * 0.2)
    t = lambda a: torch.from_numpy(a).cuda()
    W, U = t(w), t(used)
    feat, wt, sc = learned.traj_features(t(ga), t(gb), t(gc), W, U, 0.4)
    tg = learned.traj_targets(t(ga), t(gb), t(gc), t(gd), W, U).cpu().numpy()
    feat = M.bits_to_float(feat.cpu().numpy())
    wt, sc = wt.cpu().numpy(), sc.cpu().numpy()
    coef = (rng.standard_normal((feat.shape[0], 16)) * 3).astype(np.float32)
    G = t(gc.copy())
    learned.traj_jump(t(ga), t(gb), G, W, U, torch.from_numpy(coef).cuda())
    got = G.cpu().numpy()
    row = 0
    for bb in range(b):
        for l in range(nl):
            for site in range(n - 1):
                live = l < used[bb] and w[bb, l, site] == 2
                assert wt[row] == (1.0 if live else 0.0)
                if live:
                    f, s_ref = M.traj_features(ga[bb, l, site], gb[bb, l, site], gc[bb, l, site], l, nl, site, n, 0.4)
                    np.testing.assert_allclose(sc[row], s_ref, rtol=1e-6)
                    np.testing.assert_allclose(feat[row], M.bf16_round(f.astype(np.float32)), rtol=1e-2, atol=2e-3)
                    np.testing.assert_allclose(tg[row, :2], M.traj_target(ga[bb, l, site], gb[bb, l, site],
                                                                          gc[bb, l, site], gd[bb, l, site]),
                                               rtol=1e-5, atol=1e-5)
                    assert not tg[row, 2:].any()
                    ref = M.traj_jump(ga[bb, l, site], gb[bb, l, site], gc[bb, l, site], coef[row])
                    np.testing.assert_allclose(got[bb, l, site], ref, atol=1e-10)
                else:
                    assert not feat[row].any() and not tg[row].any()
                    np.testing.assert_array_equal(got[bb, l, site], gc[bb, l, site])
                row += 1


def test_warm_start_against_plain(gpu):
    """Train the trajectory network on plain-method pairs (N=12, chi=16, 6
    layers, 30 sweeps, 96 states), then encode 16 held-out states with learned
    jumps against the plain method's own 30-sweep fidelity - 1e-4 (complex64
    both).  Asserted: the network beats predicting no jump on held-out
    pairs, some jumps are kept, every state reaches the goal, and it takes
    fewer sweeps on average than the plain method."""
    from qmprs_amd import learned
    n, chi, L, S = 12, 16, 6, 30
    train_sv = np.stack([O.random_mps_state(n, 32, 1000 + s) for s in range(96)])
    test_sv = np.stack([O.random_mps_state(n, 32, 2000 + s) for s in range(16)])
    data = learned.make_trajectory_set(train_sv, chi, L, S)
    held = learned.make_trajectory_set(test_sv, chi, L, S)
    net = learned.WarmStartMLP(width=256, seed=2)
    v = learned.TRAJ_OUT
    hist = learned.train(net, data, epochs=12, batch=256, lr=1e-3, valid=v)
    tr, ho = learned.evaluate_loss(net, data, v), learned.evaluate_loss(net, held, v)
    zero = learned.zero_loss(held, v)
    _, _, fid = learned.plain_snapshots(test_sv, chi, L, S)
    goal = (fid[:, -1] - 1e-4).double().contiguous()
    plain_steps = torch.tensor([int(torch.nonzero(fid[b] >= goal[b])[0]) for b in range(len(test_sv))])
    res = learned.encode_warm(test_sv, chi, L, 45, net, goal=goal)
    warm_steps = res.sweeps.cpu()
    print(f"held-out loss: zero predictor {zero:.4e}, trained {ho:.4e} (train {tr:.4e}, epochs {hist[0]:.3e} -> "
          f"{hist[-1]:.3e}); jumps kept {int(res.accepted.sum())} of {res.tried * len(test_sv)}; sweeps to "
          f"F_ref - 1e-4: plain {plain_steps.float().mean():.2f}, learned {warm_steps.float().mean():.2f} "
          f"(max {int(warm_steps.max())})")
    assert ho < zero
    assert int(res.accepted.sum()) > 0
    assert torch.all(res.fidelity >= goal - 1e-9)
    assert warm_steps.float().mean() < plain_steps.float().mean()


@pytest.mark.parametrize("first", [0, 1, 2])
def test_sweep_snapshots_equal_split_calls(gpu, first):
    """qmx_sweep_snapshots: the gates after `first` and `first + 1` sweeps of
    one call equal the gates of the same sweeps run as separate calls, bit for
    bit (no goal: every call runs the same launches), and the call's final
    gates and fidelity equal the split run's."""
    from qmprs_amd import engine
    n, chi, L, k = 11, 16, 3, 4
    svs = np.stack([O.random_mps_state(n, 24, 60 + s) for s in range(3)]).astype(np.complex64)
    p = engine.encode_prepare(svs, chi, L, sweep_dtype=torch.complex64)
    g1 = p.gates.clone()
    buf = torch.empty((2,) + tuple(g1.shape), dtype=g1.dtype, device=g1.device)
    r1 = engine.sweep(p.target, g1, p.widths, p.layers_used, k, psi_ref=p.psi, snapshots=(first, buf))
    g2 = p.gates.clone()
    snaps = []
    for done in range(k + 1):
        if done in (first, first + 1):
            snaps.append(g2.clone())
        if done < k:
            r2 = engine.sweep(p.target, g2, p.widths, p.layers_used, 1, psi_ref=p.psi)
    torch.cuda.synchronize()
    assert torch.equal(buf[0], snaps[0]) and torch.equal(buf[1], snaps[1])
    assert torch.equal(g1, g2)
    assert torch.equal(r1.fidelity[:, -1], r2.fidelity[:, -1])
