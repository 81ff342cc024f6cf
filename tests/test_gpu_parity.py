"""GPU parity: libqmx (through the C-ABI) against the FP64 oracle.

Every comparison runs the HIP path and the oracle on the same seeded input.
Tolerances (stated per test):
  complex128 paths  - gates and states to 1e-8 .. 1e-10 (rounding only)
  complex64 sweeps  - fidelity within 1e-6 and C|0..0> within 1e-5 (L2) of the
                      oracle, the bar of the encoder's precision split.
"""

import numpy as np
import pytest
import torch

from oracle import qmprs_oracle as O

pytestmark = pytest.mark.gpu


def _engine():
    from qmprs_amd import engine
    return engine


def _states(kind, n, seed, bond=8):
    if kind == "uniform":
        return O.random_uniform_state(n, seed)
    if kind == "mps":
        return O.random_mps_state(n, bond, seed)
    raise ValueError(kind)


def _prefix_error(n, gl_a, gl_b):
    """max_j || G_j..G_1|0> - G'_j..G'_1|0> ||: the circuit's action on every
    input it ever sees.  A gate is unique only there -- where its environment
    is rank deficient (qubits still in |0>) the polar factor is free on the
    unpopulated inputs, for LAPACK as for any other SVD."""
    a = np.zeros(2 ** n, dtype=np.complex128)
    a[0] = 1
    b = a.copy()
    worst = 0.0
    for (sa, wa, ga), (sb, wb, gb) in zip(gl_a, gl_b):
        assert (sa, wa) == (sb, wb)
        a = O.apply_gate(a, n, ga, sa, wa)
        b = O.apply_gate(b, n, gb, sb, wb)
        worst = max(worst, np.linalg.norm(a - b))
    return worst


def _gate_list_from_device(gates, widths, used):
    """Device arrays of one state -> oracle gate list (circuit order)."""
    eng = _engine()
    layers = eng.gates_to_layers(gates, widths, used)
    out = []
    for layer in layers:
        for start, end, mats in layer:
            for j in range(start, end + 1):
                out.append((j, 1 if j == end else 2, mats[j - start].data))
    return out


@pytest.mark.parametrize("n,chi,kind,seed", [
    (4, 2, "uniform", 1), (6, 4, "uniform", 2), (8, 8, "mps", 3), (10, 16, "mps", 4),
    (10, 512, "uniform", 5), (12, 24, "mps", 6), (14, 32, "mps", 7),
    (11, 150, "uniform", 9),     # Gram matrices up to 256 x 256: one-block global Jacobi
    (18, 300, "uniform", 8)])    # 600 x 600: block Jacobi
def test_chi_truncate_matches_oracle(gpu, n, chi, kind, seed):
    eng = _engine()
    sv = _states(kind, n, seed, bond=2 * chi)
    psi, target, spec = eng.chi_truncate(eng.as_device_states(sv), chi, spectra=True)
    torch.cuda.synchronize()
    ref, ref_spec = O.chi_truncate(O.ket(sv), n, chi)
    got = target[0].cpu().numpy()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"n={n} chi={chi} rel.err={err:.3e}")
    assert err < 1e-10
    np.testing.assert_allclose(psi[0].cpu().numpy(), O.ket(sv), atol=1e-14)
    sp = spec[0].cpu().numpy()
    for i, lam in enumerate(ref_spec):
        if lam is None:
            continue
        k = min(len(lam), chi)
        # retained squared singular values, relative to the largest
        np.testing.assert_allclose(sp[i, :k], lam[:k], rtol=1e-9, atol=1e-13 * lam[0])


def test_chi_truncate_rank_deficient_grams(gpu):
    """N=20 states from bond-96 random MPSs truncated to chi=64 (the bench
    workload's input).  Left of bond 7 every Gram matrix has more columns than
    rows (16 x 16 of rank 4 at bond 2), so most of its eigenvalues are
    numerically zero; the Jacobi must still converge, and each truncated state
    matches the oracle."""
    eng = _engine()
    n, chi = 20, 64
    svs = np.stack([O.random_mps_state(n, 96, s) for s in (1, 2, 3)])
    _, target, _ = eng.chi_truncate(eng.as_device_states(svs), chi, spectra=True)
    torch.cuda.synchronize()
    for b in range(len(svs)):
        ref, _ = O.chi_truncate(O.ket(svs[b]), n, chi)
        got = target[b].cpu().numpy()
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        print(f"state {b} rel.err={err:.3e}")
        assert err < 1e-10


def test_chi_truncate_grid_batch_of_unequal_gram_sizes(gpu):
    """Grid Jacobi over a batch whose Gram matrices differ in size at every
    bond (600 x 600 beside at most 192 x 192) and converge after different
    numbers of sweeps: each state still matches the oracle."""
    eng = _engine()
    n, chi = 18, 300
    svs = np.stack([O.random_uniform_state(n, 12), O.random_mps_state(n, 96, 13)])
    _, target, _ = eng.chi_truncate(eng.as_device_states(svs), chi, spectra=True)
    torch.cuda.synchronize()
    for b in range(2):
        ref, _ = O.chi_truncate(O.ket(svs[b]), n, chi)
        got = target[b].cpu().numpy()
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        print(f"state {b} rel.err={err:.3e}")
        assert err < 1e-10


@pytest.mark.parametrize("n,kind,seed", [(2, "uniform", 0), (3, "uniform", 1), (5, "uniform", 2),
                                         (8, "mps", 3), (11, "uniform", 4)])
def test_generate_layer_matches_oracle(gpu, n, kind, seed):
    eng = _engine()
    sv = O.ket(_states(kind, n, seed))
    g, w, bonds = eng.generate_layer(eng.as_device_states(sv, torch.complex128))
    torch.cuda.synchronize()
    ref_layer = O.generate_layer(O.bond2_mps(sv, n))
    got = eng.gates_to_layers(g.cpu().numpy(), w.cpu().numpy(), 1)[0]
    assert [(s, e) for s, e, _ in got] == [(s, e) for s, e, _ in ref_layer]
    worst = 0.0
    for (_, _, gm), (_, _, rm) in zip(got, ref_layer):
        for a, b in zip(gm, rm):
            worst = max(worst, np.abs(a.data - b).max())
    print(f"n={n} max gate diff {worst:.3e}")
    assert worst < 1e-10


def test_generate_layer_partial_entanglement(gpu):
    """Blocks follow the bond-2 bond dims (quick _get_submps_indices)."""
    eng = _engine()
    from qmprs_amd.circuit import GateCircuit
    c = GateCircuit(8)
    c.H(0); c.CX(0, 3); c.H(4); c.H(5); c.CX(5, 7)
    sv = c.get_statevector()
    g, w, _ = eng.generate_layer(eng.as_device_states(sv, torch.complex128))
    ref = O.generate_layer(O.bond2_mps(sv, 8))
    got = eng.gates_to_layers(g.cpu().numpy(), w.cpu().numpy(), 1)[0]
    assert [(s, e) for s, e, _ in got] == [(s, e) for s, e, _ in ref]
    for (_, _, gm), (_, _, rm) in zip(got, ref):
        for a, b in zip(gm, rm):
            assert np.abs(a.data - b).max() < 1e-10


@pytest.mark.parametrize("n,seed", [(3, 0), (7, 1), (12, 2), (13, 3)])
def test_apply_layer_matches_oracle(gpu, n, seed):
    eng = _engine()
    sv = O.ket(O.random_uniform_state(n, seed))
    layer = O.generate_layer(O.bond2_mps(sv, n))
    g, w = eng.layers_to_arrays([layer], n)
    dev_g = torch.from_numpy(g).to(gpu)
    dev_w = torch.from_numpy(w).to(gpu)
    for inverse in (True, False):
        st = eng.as_device_states(sv, torch.complex128)
        eng.apply_layer(st, dev_g, dev_w, inverse)
        got = st[0].cpu().numpy()
        if inverse:
            ref = O.apply_inverse_layer_dense(sv, n, layer)
        else:
            ref = sv.copy()
            for s, e, mats in layer:
                for j in range(s, e + 1):
                    ref = O.apply_gate(ref, n, mats[j - s], j, 1 if j == e else 2)
        err = np.linalg.norm(got - ref)
        print(f"n={n} inverse={inverse} err={err:.3e}")
        assert err < 1e-12


@pytest.mark.parametrize("n,chi,layers,kind,seed", [
    (4, 4, 3, "uniform", 0), (6, 8, 5, "uniform", 1), (8, 16, 6, "mps", 2), (10, 32, 4, "uniform", 3)])
def test_extract_layers_matches_oracle(gpu, n, chi, layers, kind, seed):
    eng = _engine()
    sv = _states(kind, n, seed)
    _, target, _ = eng.chi_truncate(eng.as_device_states(sv), chi)
    work = (target / torch.linalg.vector_norm(target)).contiguous()
    gates, widths, used, amp0, _ = eng.extract_layers(work, layers)
    torch.cuda.synchronize()
    ref_t, _ = O.chi_truncate(O.ket(sv), n, chi)
    ref_layers, info = O.get_unitary_layers(ref_t, n, layers, return_info=True)
    assert int(used[0]) == len(ref_layers)
    got = _gate_list_from_device(gates[0].cpu().numpy(), widths[0].cpu().numpy(), int(used[0]))
    ref = O.layer_gate_list(ref_layers)
    worst = max(np.abs(a[2] - b[2]).max() for a, b in zip(got, ref))
    print(f"n={n} L={layers} gates max diff {worst:.3e}")
    assert [(a[0], a[1]) for a in got] == [(b[0], b[1]) for b in ref]
    assert worst < 1e-8
    np.testing.assert_allclose(amp0[0, :len(info["amp0"])].cpu().numpy(), info["amp0"], atol=1e-9)


def test_extract_layers_early_exit(gpu):
    """A product state is disentangled by one layer: the loop stops at once
    (sequential.py:390)."""
    eng = _engine()
    n = 6
    sv = np.zeros(2 ** n, dtype=np.complex128)
    sv[0b101100] = 1.0
    work = eng.as_device_states(sv, torch.complex128)
    gates, widths, used, amp0, _ = eng.extract_layers(work, 5)
    assert int(used[0]) == 1
    assert abs(amp0[0, 0].item() - 1) < 1e-12


@pytest.mark.parametrize("dtype,n,layers,sweeps,seed", [
    (torch.complex128, 4, 2, 3, 0), (torch.complex128, 6, 3, 4, 1), (torch.complex128, 9, 3, 2, 2),
    (torch.complex64, 8, 4, 3, 3), (torch.complex64, 12, 3, 2, 4)])
def test_sweep_matches_oracle(gpu, dtype, n, layers, sweeps, seed):
    eng = _engine()
    sv = O.ket(O.random_uniform_state(n, seed))
    ref_t, _ = O.chi_truncate(sv, n, 2 ** n)
    ref_layers = O.get_unitary_layers(ref_t, n, layers)
    gl0 = O.layer_gate_list(ref_layers)
    trace = []
    gl = O.optimize_gate_list(ref_t, n, gl0, sweeps, trace)
    gl_j = O.optimize_gate_list(ref_t, n, gl0, sweeps, None, polar="jacobi")
    ref_state = O.circuit_state(n, gl)
    ref_fid = abs(np.vdot(sv, ref_state))
    g, w = eng.layers_to_arrays(ref_layers, n)
    gates = torch.from_numpy(g).to(gpu).unsqueeze(0).contiguous()
    widths = torch.from_numpy(w).to(gpu).unsqueeze(0).contiguous()
    used = torch.tensor([g.shape[0]], dtype=torch.int32, device=gpu)
    target = torch.from_numpy(ref_t).to(gpu).to(dtype).unsqueeze(0).contiguous()
    psi_ref = torch.from_numpy(sv).to(gpu).unsqueeze(0).contiguous()
    res = eng.sweep(target, gates, widths, used, sweeps, psi_ref=psi_ref, circuit_out=True)
    torch.cuda.synchronize()
    fid = abs(res.fidelity[0, -1].item())
    circ = res.circuit[0].cpu().numpy().astype(np.complex128)
    got = _gate_list_from_device(gates[0].cpu().numpy(), widths[0].cpu().numpy(), g.shape[0])
    # gate by gate against the oracle with the kernel's polar completion rule
    gdiff = max(np.linalg.norm(a[2] - b[2]) for a, b in zip(got, gl_j))
    serr = np.linalg.norm(circ - ref_state)
    print(f"{dtype} n={n} fid {fid:.12f} ref {ref_fid:.12f} gate err {gdiff:.3e} state err {serr:.3e}")
    for _, _, gm in got:   # every updated gate is unitary
        d = gm.shape[0]
        assert np.abs(gm @ gm.conj().T - np.eye(d)).max() < 1e-12
    np.testing.assert_allclose(res.overlap[0].cpu().numpy(), np.abs(trace), rtol=1e-6 if dtype == torch.complex64 else 1e-10)
    if dtype == torch.complex128:
        assert abs(fid - ref_fid) < 1e-10
        assert gdiff < 1e-8
        assert serr < 1e-9
    else:
        assert abs(fid - ref_fid) < 1e-6
        assert serr < 1e-5


@pytest.mark.parametrize("n,layers,sweeps,seed", [(6, 3, 3, 1), (10, 3, 3, 2), (11, 3, 2, 3)])
def test_relaxed_sweep_matches_oracle(gpu, n, layers, sweeps, seed):
    """Over-relaxed gate update (omega = 1.6; an option with no reference
    counterpart) on the per-gate path (N = 6 < 10) and the windowed path
    (N = 10, 11: window_finalize_kernel's relaxed update) against the
    oracle's relaxed sweep, complex128."""
    from qmprs_amd import _lib
    eng = _engine()
    omega = 1.6
    sv = O.ket(O.random_uniform_state(n, seed))
    ref_t, _ = O.chi_truncate(sv, n, 2 ** n)
    ref_layers = O.get_unitary_layers(ref_t, n, layers)
    gl0 = O.layer_gate_list(ref_layers)
    trace = []
    gl = O.optimize_gate_list(ref_t, n, gl0, sweeps, trace, polar="jacobi", omega=omega)
    ref_state = O.circuit_state(n, gl)
    ref_fid = abs(np.vdot(sv, ref_state))
    g, w = eng.layers_to_arrays(ref_layers, n)
    gates = torch.from_numpy(g).to(gpu).unsqueeze(0).contiguous()
    widths = torch.from_numpy(w).to(gpu).unsqueeze(0).contiguous()
    used = torch.tensor([g.shape[0]], dtype=torch.int32, device=gpu)
    target = torch.from_numpy(ref_t).to(gpu).unsqueeze(0).contiguous()
    psi_ref = torch.from_numpy(sv).to(gpu).unsqueeze(0).contiguous()
    _lib.set_sweep_relaxation(omega)
    try:
        res = eng.sweep(target, gates, widths, used, sweeps, psi_ref=psi_ref, circuit_out=True)
        torch.cuda.synchronize()
    finally:
        _lib.set_sweep_relaxation(1.0)
    fid = abs(res.fidelity[0, -1].item())
    circ = res.circuit[0].cpu().numpy().astype(np.complex128)
    got = _gate_list_from_device(gates[0].cpu().numpy(), widths[0].cpu().numpy(), g.shape[0])
    gdiff = max(np.linalg.norm(a[2] - b[2]) for a, b in zip(got, gl))
    serr = np.linalg.norm(circ - ref_state)
    print(f"n={n} omega={omega} fid {fid:.12f} ref {ref_fid:.12f} gate err {gdiff:.3e} state err {serr:.3e}")
    for _, _, gm in got:
        d = gm.shape[0]
        assert np.abs(gm @ gm.conj().T - np.eye(d)).max() < 1e-12
    np.testing.assert_allclose(res.overlap[0].cpu().numpy(), np.abs(trace), rtol=1e-10)
    assert abs(fid - ref_fid) < 1e-10
    assert gdiff < 1e-8
    assert serr < 1e-9


@pytest.mark.parametrize("n,chi,layers,sweeps,seed,sdtype", [
    (6, 64, 4, 3, 0, torch.complex128), (8, 8, 5, 4, 1, torch.complex64),
    (10, 512, 5, 3, 2, torch.complex64), (12, 16, 4, 2, 3, torch.complex64)])
def test_encode_end_to_end(gpu, n, chi, layers, sweeps, seed, sdtype):
    """Whole prepare_state through qmx_encode against the oracle's."""
    eng = _engine()
    sv = O.random_uniform_state(n, seed) if n != 12 else O.random_mps_state(n, 2 * chi, seed)
    res = eng.encode(sv, chi, layers, sweeps, sweep_dtype=sdtype)
    torch.cuda.synchronize()
    ref = O.prepare_state(sv, chi, num_layers=layers, num_sweeps=sweeps)
    fid = abs(res.fidelity[0, -1].item())
    gl = _gate_list_from_device(res.gates[0].cpu().numpy(), res.widths[0].cpu().numpy(),
                                int(res.layers_used[0]))
    state = O.circuit_state(n, gl)
    serr = np.linalg.norm(state - ref["state"])
    print(f"n={n} chi={chi} fid {fid:.10f} ref {ref['fidelity']:.10f} state err {serr:.3e}")
    assert int(res.layers_used[0]) == len(ref["layers"])
    assert abs(fid - ref["fidelity"]) < 1e-6
    assert serr < 1e-5
    # the reported fidelity is the overlap of the emitted circuit with the input
    assert abs(abs(np.vdot(O.ket(sv), state)) - fid) < 1e-6


def test_encode_batch_equals_single(gpu):
    """A batch gives the same gates as each state alone (no cross-talk)."""
    eng = _engine()
    n, chi = 9, 8
    svs = np.stack([O.random_mps_state(n, 16, s) for s in range(5)])
    batch = eng.encode(svs, chi, 3, 2, sweep_dtype=torch.complex128)
    for b in (0, 3):
        one = eng.encode(svs[b], chi, 3, 2, sweep_dtype=torch.complex128)
        assert torch.equal(batch.layers_used[b:b + 1], one.layers_used)
        assert (batch.gates[b] - one.gates[0]).abs().max().item() < 1e-12
        assert abs(batch.fidelity[b, -1].item() - one.fidelity[0, -1].item()) < 1e-12


@pytest.mark.parametrize("dtype,n,layers,sweeps,seed", [
    (torch.complex128, 10, 3, 2, 5), (torch.complex128, 11, 4, 2, 6), (torch.complex64, 13, 3, 2, 7)])
def test_window_sweep_equals_step_sweep(gpu, dtype, n, layers, sweeps, seed):
    """The windowed sweep (one HBM pass per gate window, 16x16 MFMA
    environment; N >= 10) is the gate-by-gate sweep regrouped: same gates,
    same state."""
    eng = _engine()
    from qmprs_amd import _lib
    sv = O.ket(O.random_mps_state(n, 16, seed))
    ref_t, _ = O.chi_truncate(sv, n, 2 ** n)
    ref_layers = O.get_unitary_layers(ref_t, n, layers)
    g, w = eng.layers_to_arrays(ref_layers, n)
    out = {}
    try:
        for mode in ("step", "window"):
            _lib.set_sweep_mode(mode)
            gates = torch.from_numpy(g).to(gpu).unsqueeze(0).contiguous()
            widths = torch.from_numpy(w).to(gpu).unsqueeze(0).contiguous()
            used = torch.tensor([g.shape[0]], dtype=torch.int32, device=gpu)
            target = torch.from_numpy(ref_t).to(gpu).to(dtype).unsqueeze(0).contiguous()
            psi_ref = torch.from_numpy(sv).to(gpu).unsqueeze(0).contiguous()
            res = eng.sweep(target, gates, widths, used, sweeps, psi_ref=psi_ref, circuit_out=True)
            torch.cuda.synchronize()
            out[mode] = (gates.cpu().numpy(), res.circuit[0].cpu().numpy(),
                         res.fidelity[0].abs().cpu().numpy(), res.overlap[0].cpu().numpy())
    finally:
        _lib.set_sweep_mode("auto")
    gd = np.abs(out["step"][0] - out["window"][0]).max()
    sd = np.linalg.norm(out["step"][1] - out["window"][1])
    print(f"{dtype} n={n}: gate diff {gd:.3e} state diff {sd:.3e} fid {out['window'][2]}")
    if dtype == torch.complex128:
        assert gd < 1e-10
    # complex64: the environments' structurally zero singular values sit at
    # float rounding (~1e-8), so the polar factor is pinned only on the inputs
    # the circuit populates; states and fidelities are the comparison
    assert sd < (1e-10 if dtype == torch.complex128 else 1e-5)
    np.testing.assert_allclose(out["window"][2], out["step"][2], atol=1e-12 if dtype == torch.complex128 else 1e-6)
    np.testing.assert_allclose(out["window"][3], out["step"][3], rtol=1e-10 if dtype == torch.complex128 else 1e-6)


def test_prepare_then_sweeps_equals_encode(gpu):
    """qmx_encode == qmx_encode_prepare + qmx_sweep, and S sweeps in one call
    == 1 + (S - 1) sweeps in two calls, bit for bit (every sweep restarts from
    the target and the current gates)."""
    eng = _engine()
    n, chi, L, S = 12, 16, 4, 5
    svs = np.stack([O.random_mps_state(n, 24, 40 + s) for s in range(3)])
    ref = eng.encode(svs, chi, L, S, sweep_dtype=torch.complex64)
    p = eng.encode_prepare(svs, chi, L, sweep_dtype=torch.complex64)
    r1 = eng.sweep(p.target, p.gates, p.widths, p.layers_used, 1, psi_ref=p.psi)
    r2 = eng.sweep(p.target, p.gates, p.widths, p.layers_used, S - 1, psi_ref=p.psi)
    assert torch.equal(p.layers_used, ref.layers_used)
    assert torch.equal(p.gates, ref.gates)
    assert torch.equal(r2.fidelity[:, -1], ref.fidelity[:, -1])
    assert torch.equal(r1.fidelity[:, 0], ref.fidelity[:, 0])


@pytest.mark.parametrize("batch", [2, 5])
def test_split_sweeps_equal_single_stream(gpu, batch):
    """Two interleaved half-batches on two streams give the single-stream
    results (up to the order of the per-block partial sums)."""
    from qmprs_amd import _lib
    eng = _engine()
    n, chi, L, S = 11, 16, 3, 3
    svs = np.stack([O.random_mps_state(n, 24, 60 + s) for s in range(batch)])
    _lib.set_split_sweeps(False)
    try:
        one = eng.encode(svs, chi, L, S, sweep_dtype=torch.complex128)
    finally:
        _lib.set_split_sweeps(True)
    two = eng.encode(svs, chi, L, S, sweep_dtype=torch.complex128)
    torch.cuda.synchronize()
    assert torch.equal(one.layers_used, two.layers_used)
    assert (one.gates - two.gates).abs().max().item() < 1e-10
    assert (one.fidelity - two.fidelity).abs().max().item() < 1e-12


@pytest.mark.parametrize("dtype,tol", [(torch.complex64, 1e-5), (torch.complex128, 1e-10)])
def test_encode_mfma_rebuild_matches_vector_rebuild(gpu, dtype, tol):
    """An encode (N=14, 3 sweeps) with the circuit rebuilds on the matrix
    cores and with the vector-unit rebuilds: the same layers, and fidelities
    within the rounding that separates the two (complex64: the sweeps start
    from rebuilt states that differ by ~1e-7; complex128: by ~1e-16); both
    encodes' final gates re-evaluated in FP64 give the same fidelity.  Gate
    entries are not compared: where a gate's environment is rank-deficient
    its polar factor is not unique, and rounding picks a different completion
    of the same circuit fidelity (entries differ by O(1) there)."""
    from qmprs_amd import _lib
    eng = _engine()
    n, chi, L, S = 14, 16, 3, 3
    svs = np.stack([O.random_mps_state(n, 24, 500 + s) for s in range(3)])
    try:
        _lib.set_option(_lib.QMX_OPT_FWD_MFMA, 0)
        vec = eng.encode(svs, chi, L, S, sweep_dtype=dtype)
        _lib.set_option(_lib.QMX_OPT_FWD_MFMA, 1)
        mat = eng.encode(svs, chi, L, S, sweep_dtype=dtype)
    finally:
        _lib.set_option(_lib.QMX_OPT_FWD_MFMA, _lib.FWD_MFMA_DEFAULT)
    torch.cuda.synchronize()
    assert torch.equal(vec.layers_used, mat.layers_used)
    dg = (vec.gates - mat.gates).abs().max().item()
    df = (vec.fidelity.abs() - mat.fidelity.abs()).abs().max().item()
    psi = torch.from_numpy(svs).to(gpu)
    _, fv = eng.circuit_state(vec.gates, vec.widths, vec.layers_used, psi_ref=psi)
    _, fm = eng.circuit_state(mat.gates, mat.widths, mat.layers_used, psi_ref=psi)
    d64 = (fv.abs() - fm.abs()).abs().max().item()
    print(f"max |dgate| {dg:.2e}, max |dF| {df:.2e}, FP64 re-evaluation |dF| {d64:.2e}")
    assert df < tol and d64 < tol


@pytest.mark.parametrize("omega", [1.0, 1.7])
def test_encode_ragged_equals_bucket_encodes(gpu, omega):
    """qmx_encode_ragged: states of different (N, chi), shuffled, give what
    qmx_encode gives on each (N, chi) bucket (bit for bit: same batches);
    the caller handle's sweep relaxation reaches the bucket workers."""
    from qmprs_amd import _lib
    eng = _engine()
    spec = [(9, 8), (11, 16), (9, 8), (12, 8), (11, 16), (10, 4), (9, 8)]
    svs = [O.random_mps_state(n, 2 * chi, 70 + i).astype(np.complex64) for i, (n, chi) in enumerate(spec)]
    L, S = 3, 2
    buckets = {}
    for i, (n, chi) in enumerate(spec):
        buckets.setdefault((n, chi), []).append(i)
    _lib.set_sweep_relaxation(omega)
    try:
        res = eng.encode_ragged(svs, [c for _, c in spec], L, S, max_concurrency=3)
        refs = {key: eng.encode(np.stack([svs[i] for i in idx]), key[1], L, S)
                for key, idx in buckets.items()}
    finally:
        _lib.set_sweep_relaxation(1.0)
    for key, idx in buckets.items():
        ref = refs[key]
        for k, i in enumerate(idx):
            assert res[i].layers_used == int(ref.layers_used[k])
            assert torch.equal(res[i].gates, ref.gates[k])
            assert torch.equal(res[i].widths, ref.widths[k])
            assert torch.equal(res[i].fidelity, ref.fidelity[k])
    if omega != 1.0:   # and the relaxed gates differ from the plain ones
        plain = eng.encode(np.stack([svs[i] for i in buckets[(11, 16)]]), 16, L, S)
        assert not torch.equal(plain.gates, refs[(11, 16)].gates)


def test_encode_grouped_one_group_equals_encode(gpu):
    """qmx_encode_grouped with one bond dimension is qmx_encode, bit for bit."""
    eng = _engine()
    svs = np.stack([O.random_mps_state(10, 16, 90 + i).astype(np.complex64) for i in range(3)])
    L, S = 3, 2
    a = eng.encode(svs, 8, L, S)
    b = eng.encode_grouped(svs, [8, 8, 8], L, S)
    torch.cuda.synchronize()
    assert torch.equal(a.layers_used, b.layers_used)
    assert torch.equal(a.gates, b.gates)
    assert torch.equal(a.widths, b.widths)
    assert torch.equal(a.fidelity, b.fidelity)


def test_encode_ragged_mixed_bond_dimensions(gpu):
    """States of one qubit count but different bond dimensions share one
    encode (per-group chi-truncation, one layer extraction and sweep batch):
    each state matches qmx_encode of its own (N, chi) bucket up to the
    summation order of the batch-sized reductions (FP64 sweeps: gates within
    1e-9, fidelities within 1e-10)."""
    eng = _engine()
    spec = [(10, 4), (10, 16), (11, 8), (10, 4), (11, 32), (10, 8), (11, 8)]
    svs = [O.random_mps_state(n, 2 * chi, 110 + i).astype(np.complex64) for i, (n, chi) in enumerate(spec)]
    L, S = 3, 3
    res = eng.encode_ragged(svs, [c for _, c in spec], L, S, sweep_dtype=torch.complex128, max_concurrency=2)
    buckets = {}
    for i, (n, chi) in enumerate(spec):
        buckets.setdefault((n, chi), []).append(i)
    worst_g = worst_f = 0.0
    for (n, chi), idx in buckets.items():
        ref = eng.encode(np.stack([svs[i] for i in idx]), chi, L, S, sweep_dtype=torch.complex128)
        for k, i in enumerate(idx):
            assert res[i].layers_used == int(ref.layers_used[k])
            assert torch.equal(res[i].widths, ref.widths[k])
            worst_g = max(worst_g, (res[i].gates - ref.gates[k]).abs().max().item())
            worst_f = max(worst_f, (res[i].fidelity - ref.fidelity[k]).abs().max().item())
    print(f"max gate diff {worst_g:.2e}, max fidelity diff {worst_f:.2e}")
    assert worst_g < 1e-9
    assert worst_f < 1e-10


def test_encode_ragged_rejects_bad_input(gpu):
    eng = _engine()
    with pytest.raises(ValueError):
        eng.encode_ragged([np.ones(8, np.complex64)], [4, 4], 2, 1)
    with pytest.raises(ValueError):
        eng.encode_ragged([np.ones(8, np.complex64)], [0], 2, 1)


@pytest.mark.parametrize("dtype,mode,n,count", [(torch.complex64, 1, 13, 3), (torch.complex128, 1, 13, 3),
                                                (torch.complex64, 2, 13, 3), (torch.complex128, 2, 13, 3),
                                                (torch.complex64, 1, 20, 8), (torch.complex128, 1, 20, 8)])
def test_staged_windows_equal_direct(gpu, dtype, mode, n, count):
    """Windows whose gate bits are the lowest state bits move their tiles
    through LDS in tile order (mode 1: tiles with a contiguous low run of
    2^7 amplitudes, mode 2: from 2^3); the arithmetic is the same as the
    direct group loads, so results are identical bit for bit.  At N=20 with
    8 states each block takes two of a state's 512 tiles, so the next tile's
    amplitudes are loaded while the current one is processed."""
    from qmprs_amd import _lib
    eng = _engine()
    chi, L, S = 16, (3 if n < 20 else 2), (2 if n < 20 else 1)
    svs = np.stack([O.random_mps_state(n, 24, 80 + s) for s in range(count)])
    _lib.set_option(_lib.QMX_OPT_STAGED_WINDOWS, 0)
    try:
        direct = eng.encode(svs, chi, L, S, sweep_dtype=dtype)
        _lib.set_option(_lib.QMX_OPT_STAGED_WINDOWS, mode)
        staged = eng.encode(svs, chi, L, S, sweep_dtype=dtype)
    finally:
        _lib.set_option(_lib.QMX_OPT_STAGED_WINDOWS, 1)
    torch.cuda.synchronize()
    assert torch.equal(direct.gates, staged.gates)
    assert torch.equal(direct.fidelity, staged.fidelity)


def _random_unitary(rng, d):
    q, r = np.linalg.qr(rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)))
    return q * (np.diag(r) / np.abs(np.diag(r)))


@pytest.mark.parametrize("n,layers,seed", [(13, 3, 0), (16, 4, 1), (20, 3, 2)])
def test_circuit_state_large_n(gpu, n, layers, seed):
    """qmx_circuit_state at sizes with many tiles (the forward passes run over
    the light cone of 8-bit windows and skip tiles outside the state's
    support) against the oracle's dense forward application, for two states
    of different layer counts in one batch; random unitaries, blocks ending
    at random sites."""
    eng = _engine()
    rng = np.random.default_rng(seed)
    B = 2
    g = np.zeros((B, layers, n, 4, 4), dtype=np.complex128)
    w = np.full((B, layers, n), 2, dtype=np.int32)
    for b in range(B):
        for l in range(layers):
            w[b, l, n - 1] = 1
            for s in rng.choice(n - 1, size=2, replace=False):
                w[b, l, s] = 1
            for s in range(n):
                d = 4 if w[b, l, s] == 2 else 2
                g[b, l, s, :d, :d] = _random_unitary(rng, d)
    used = np.array([layers, layers - 1], dtype=np.int32)
    gates = torch.from_numpy(g).to(gpu).contiguous()
    widths = torch.from_numpy(w).to(gpu).contiguous()
    lu = torch.from_numpy(used).to(gpu)
    state = eng.circuit_state(gates, widths, lu, dtype=torch.complex128)
    torch.cuda.synchronize()
    for b in range(B):
        gl = _gate_list_from_device(g[b], w[b], int(used[b]))
        ref = O.circuit_state(n, gl)
        err = np.abs(state[b].cpu().numpy() - ref).max()
        print(f"n={n} state {b} layers {used[b]} max err {err:.3e}")
        assert err < 1e-12


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("n,layers,seed", [(12, 3, 0), (12, 10, 4), (13, 4, 1), (16, 3, 2), (20, 2, 3), (22, 2, 5)])
def test_circuit_state_mfma_matches_oracle(gpu, dtype, n, layers, seed):
    """The forward passes on the matrix cores (each stage's gates fused into
    a 16 x 16 matrix; FP32 MFMA products for complex64 states, FP64 for
    complex128) against the oracle's dense forward application and against
    the vector-unit passes that apply the gates one by one: rounding apart,
    the same state (max abs error below 2e-6 in complex64, 1e-12 in
    complex128, on states of norm 1) and the same overlap with a reference.
    The plans include one-stage passes (N=13, 4 layers: HBM to HBM, no LDS
    exchange) and twelve-stage passes (N=12, 10 layers: the most stage
    matrices a block holds)."""
    from qmprs_amd import _lib
    eng = _engine()
    rng = np.random.default_rng(100 + seed)
    B = 2
    g = np.zeros((B, layers, n, 4, 4), dtype=np.complex128)
    w = np.full((B, layers, n), 2, dtype=np.int32)
    for b in range(B):
        for l in range(layers):
            w[b, l, n - 1] = 1
            for s in rng.choice(n - 1, size=2, replace=False):
                w[b, l, s] = 1
            for s in range(n):
                d = 4 if w[b, l, s] == 2 else 2
                g[b, l, s, :d, :d] = _random_unitary(rng, d)
    used = np.array([layers, layers - 1], dtype=np.int32)
    gates = torch.from_numpy(g).to(gpu).contiguous()
    widths = torch.from_numpy(w).to(gpu).contiguous()
    lu = torch.from_numpy(used).to(gpu)
    psi = np.stack([O.random_mps_state(n, 8, 300 + seed + b) for b in range(B)])
    psi_ref = torch.from_numpy(psi).to(gpu).contiguous()
    try:
        _lib.set_option(_lib.QMX_OPT_FWD_MFMA, 1)
        s_mfma, f_mfma = eng.circuit_state(gates, widths, lu, dtype=dtype, psi_ref=psi_ref)
        _lib.set_option(_lib.QMX_OPT_FWD_MFMA, 0)
        s_valu, f_valu = eng.circuit_state(gates, widths, lu, dtype=dtype, psi_ref=psi_ref)
    finally:
        _lib.set_option(_lib.QMX_OPT_FWD_MFMA, _lib.FWD_MFMA_DEFAULT)
    torch.cuda.synchronize()
    tol, ftol = (2e-6, 1e-5) if dtype == torch.complex64 else (1e-12, 1e-12)
    for b in range(B):
        gl = _gate_list_from_device(g[b], w[b], int(used[b]))
        ref = O.circuit_state(n, gl)
        sm = s_mfma[b].cpu().numpy().astype(np.complex128)
        sv = s_valu[b].cpu().numpy().astype(np.complex128)
        em, ev = np.abs(sm - ref).max(), np.abs(sv - ref).max()
        fr = np.vdot(psi[b], ref)
        print(f"n={n} state {b} layers {used[b]}: mfma err {em:.3e}, vector err {ev:.3e}")
        assert em < tol and ev < tol
        assert np.abs(sm - sv).max() < tol
        assert abs(complex(f_mfma[b]) - fr) < ftol and abs(complex(f_valu[b]) - fr) < ftol


def test_circuit_state_mfma_fp64_largest_size(gpu):
    """The complex128 matrix-core forward passes at their largest size, N=27
    (32-bit byte offsets up to 2^31 - 16), against the vector-unit passes on
    one state of 2^27 amplitudes: the same state to 1e-12 (max abs), norm 1
    to 1e-12 (the circuit is unitary) and the same overlap with a random
    reference state.  The dense oracle is out of reach at this size; the
    vector passes are checked against it at N <= 22 above."""
    from qmprs_amd import _lib
    import torch
    eng = _engine()
    n, layers = 27, 2
    rng = np.random.default_rng(2027)
    g = np.zeros((1, layers, n, 4, 4), dtype=np.complex128)
    w = np.full((1, layers, n), 2, dtype=np.int32)
    for l in range(layers):
        w[0, l, n - 1] = 1
        for s in rng.choice(n - 1, size=3, replace=False):
            w[0, l, s] = 1
        for s in range(n):
            d = 4 if w[0, l, s] == 2 else 2
            g[0, l, s, :d, :d] = _random_unitary(rng, d)
    gates = torch.from_numpy(g).to(gpu).contiguous()
    widths = torch.from_numpy(w).to(gpu).contiguous()
    lu = torch.tensor([layers], dtype=torch.int32, device=gpu)
    gen = torch.Generator(device=gpu)
    gen.manual_seed(27)
    ref = torch.randn((1, 1 << n), dtype=torch.complex128, device=gpu, generator=gen)
    ref /= torch.linalg.vector_norm(ref)
    try:
        _lib.set_option(_lib.QMX_OPT_FWD_MFMA, 1)
        s_mfma, f_mfma = eng.circuit_state(gates, widths, lu, dtype=torch.complex128, psi_ref=ref)
        _lib.set_option(_lib.QMX_OPT_FWD_MFMA, 0)
        s_valu, f_valu = eng.circuit_state(gates, widths, lu, dtype=torch.complex128, psi_ref=ref)
    finally:
        _lib.set_option(_lib.QMX_OPT_FWD_MFMA, _lib.FWD_MFMA_DEFAULT)
    torch.cuda.synchronize()
    diff = (s_mfma - s_valu).abs().max().item()
    norm = torch.linalg.vector_norm(s_mfma).item()
    df = abs(complex(f_mfma[0]) - complex(f_valu[0]))
    print(f"N=27 complex128: max |mfma - vector| {diff:.2e}, norm - 1 {norm - 1:.2e}, overlap diff {df:.2e}")
    assert diff < 1e-12 and abs(norm - 1.0) < 1e-12 and df < 1e-12


@pytest.mark.parametrize("schedule", ["IterDiOi", "IterDiOall"])
@pytest.mark.parametrize("n,chi,layers,sweeps,seed", [(6, 8, 3, 3, 0), (9, 16, 3, 2, 1)])
def test_schedule_matches_oracle(gpu, schedule, n, chi, layers, sweeps, seed):
    """The iterative schedules (layer / sweep / inverse-layer kernels driven by
    engine.schedule_layers) against the oracle's restatement, complex128
    sweeps: same layer count, gates to 1e-7, fidelity to 1e-9, C|0..0> to
    1e-8 (L2)."""
    eng = _engine()
    sv = O.random_uniform_state(n, seed)
    gates, widths, used, fid = eng.encode_schedule(sv, chi, layers, sweeps, schedule=schedule,
                                                   sweep_dtype=torch.complex128)
    torch.cuda.synchronize()
    ref = O.prepare_state_schedule(sv, chi, layers, sweeps, schedule=schedule, polar="jacobi")
    assert int(used[0]) == len(ref["layers"])
    gl = _gate_list_from_device(gates[0].cpu().numpy(), widths[0].cpu().numpy(), int(used[0]))
    gdiff = max(np.abs(a[2] - b[2]).max() for a, b in zip(gl, ref["gates"]))
    state = O.circuit_state(n, gl)
    serr = np.linalg.norm(state - ref["state"])
    print(f"{schedule} n={n} fid {float(fid[0]):.12f} ref {ref['fidelity']:.12f} gate err {gdiff:.2e} "
          f"state err {serr:.2e}")
    assert gdiff < 1e-7
    assert abs(float(fid[0]) - ref["fidelity"]) < 1e-9
    assert serr < 1e-8


def test_schedule_batch_early_exit_and_sequential_api(gpu):
    """A batch where one state stops early (an exactly encodable product
    state) keeps that state's layers while the other continues; the
    Sequential API's schedule keyword returns a circuit whose state has the
    reported fidelity."""
    from qmprs_amd.circuit import GateCircuit
    from qmprs_amd.synthesis.mps_encoding import Sequential
    eng = _engine()
    n = 6
    prod = np.zeros(2 ** n, dtype=np.complex128)
    prod[0] = 1.0
    rnd = O.random_uniform_state(n, 7)
    svs = np.stack([prod, rnd])
    gates, widths, used, fid = eng.encode_schedule(svs, 8, 3, 2, schedule="IterDiOall",
                                                   sweep_dtype=torch.complex128)
    torch.cuda.synchronize()
    for b, sv in enumerate(svs):
        ref = O.prepare_state_schedule(sv, 8, 3, 2, schedule="IterDiOall", polar="jacobi")
        assert int(used[b]) == len(ref["layers"])
        assert abs(float(fid[b]) - ref["fidelity"]) < 1e-9
    assert int(used[0]) == 1 and int(used[1]) == 3
    enc = Sequential(circuit_framework=GateCircuit)
    circ = enc.prepare_state(rnd, 8, num_layers=3, num_sweeps=2, schedule="IterDiOi",
                             sweep_dtype=torch.complex128)
    got = abs(np.vdot(O.ket(rnd), circ.get_statevector()))
    assert abs(got - enc.last_info["fidelity"]) < 1e-9
    with pytest.raises(ValueError):
        enc.prepare_state(rnd, 8, num_layers=2, schedule="Brickwall")


def test_sequential_sweep_relaxation(gpu):
    """``sweep_relaxation`` through the Sequential API: the circuit is the
    oracle's over-relaxed encoding (complex128 sweeps), the option is reset
    afterwards (a plain encode follows the reference's update again), and
    values outside [1, 1.95] are rejected."""
    from qmprs_amd import _lib
    from qmprs_amd.circuit import GateCircuit
    from qmprs_amd.synthesis.mps_encoding import Sequential
    n, chi, layers, sweeps = 10, 16, 3, 3    # N >= 10: the windowed sweep
    sv = O.ket(O.random_mps_state(n, 24, 5))
    enc = Sequential(circuit_framework=GateCircuit)
    target, _ = O.chi_truncate(sv, n, chi)
    gl0 = O.layer_gate_list(O.get_unitary_layers(target, n, layers))
    for omega in (1.7, 1.0):
        circ = enc.prepare_state(sv, chi, num_layers=layers, num_sweeps=sweeps,
                                 sweep_relaxation=omega, sweep_dtype=torch.complex128)
        gl = O.optimize_gate_list(target, n, gl0, sweeps, polar="jacobi", omega=omega)
        ref = O.circuit_state(n, gl)
        err = np.linalg.norm(circ.get_statevector() - ref)
        print(f"omega {omega}: state err {err:.3e}")
        assert err < 1e-9
    for bad in (0.9, 2.0, True, "1.5"):
        with pytest.raises(ValueError):
            enc.prepare_state(sv, chi, num_layers=1, sweep_relaxation=bad)
    with pytest.raises(ValueError):
        _lib.set_sweep_relaxation(2.5)


def test_large_grid_window_sweep_equals_step_sweep(gpu):
    """N=23, one state: the window pass runs on 512 blocks and the finalize
    reads the partial environments pre-summed by groups of 32; the result is
    the gate-by-gate sweep's (complex128)."""
    eng = _engine()
    from qmprs_amd import _lib
    n = 23
    sv = O.random_mps_state(n, 16, 11)
    prep = eng.encode_prepare(sv, 16, 2, sweep_dtype=torch.complex128)
    out = {}
    try:
        for mode in ("step", "window"):
            _lib.set_sweep_mode(mode)
            gates = prep.gates.clone()
            res = eng.sweep(prep.target, gates, prep.widths, prep.layers_used, 1, psi_ref=prep.psi,
                            circuit_out=True)
            torch.cuda.synchronize()
            out[mode] = (gates, res.circuit[0], res.fidelity[0].abs())
    finally:
        _lib.set_sweep_mode("auto")
    gd = (out["step"][0] - out["window"][0]).abs().max().item()
    sd = torch.linalg.vector_norm(out["step"][1] - out["window"][1]).item()
    fd = (out["step"][2] - out["window"][2]).abs().max().item()
    print(f"n={n}: gate diff {gd:.3e} state diff {sd:.3e} fidelity diff {fd:.3e}")
    assert gd < 1e-10 and sd < 1e-10 and fd < 1e-12


@pytest.mark.parametrize("n,chi", [(2, 2), (3, 2), (4, 4), (5, 4), (9, 32)])
def test_encode_small_sizes(gpu, n, chi):
    """The smallest states (gate-by-gate sweeps and tile-sized forward passes
    below 5 qubits; light-cone rebuilds covering the whole state from 5 to
    12) end to end against the oracle, complex128 sweeps."""
    eng = _engine()
    sv = O.random_uniform_state(n, 40 + n)
    res = eng.encode(sv, chi, 3, 2, sweep_dtype=torch.complex128)
    torch.cuda.synchronize()
    ref = O.prepare_state(sv, chi, num_layers=3, num_sweeps=2, polar="jacobi")
    assert int(res.layers_used[0]) == len(ref["layers"])
    fid = abs(res.fidelity[0, -1].item())
    gl = _gate_list_from_device(res.gates[0].cpu().numpy(), res.widths[0].cpu().numpy(),
                                int(res.layers_used[0]))
    serr = np.linalg.norm(O.circuit_state(n, gl) - ref["state"])
    print(f"n={n} chi={chi} fid {fid:.12f} ref {ref['fidelity']:.12f} state err {serr:.2e}")
    assert abs(fid - ref["fidelity"]) < 1e-9
    assert serr < 1e-8
