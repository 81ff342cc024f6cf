"""The workload configurations at their own sizes, and the error paths.

* C2 (N=20, chi=64, 10 layers, bond-96 random-MPS input): two sweeps against
  the FP64 oracle; the full 30 sweeps in the complex64 throughput mode
  against the FP64 (complex128) path's fidelity goal.
* C3 (ragged batch): states of N <= 14 with chi in {16, 32, 64} against the
  oracle, and the 256-state batch (N 12..20) through size-independent
  properties.
* C5 (N=26, chi=512, 20 layers, 50 sweeps): one full-size property run.
* Host-forced failures: the unitarity check of mps.py:837-839 and the
  eigensolver's sweep limit, with every kernel running to its normal end.

Bars (written in each test): fidelity within 1e-6 and reconstructed states
within 1e-5 (L2) of the oracle; chi-states within 1e-10 relative; complex64
runs judged by the FP64 goal F_ref - 1e-4.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import qmprs_oracle as O

pytestmark = pytest.mark.gpu


def _engine():
    from qmprs_amd import engine
    return engine


def _say(msg):
    print(msg, flush=True)


def _gate_list(gates, widths, used):
    eng = _engine()
    out = []
    for layer in eng.gates_to_layers(gates, widths, used):
        for start, end, mats in layer:
            for j in range(start, end + 1):
                out.append((j, 1 if j == end else 2, mats[j - start].data))
    return out


def _assert_unitary(gates, widths, used, tol=1e-12):
    g = gates.cpu().numpy()
    w = widths.cpu().numpy()
    for b in range(g.shape[0]):
        for l in range(int(used[b])):
            for i in range(g.shape[2]):
                d = 4 if w[b, l, i] == 2 else 2
                m = g[b, l, i, :d, :d]
                assert np.abs(m @ m.conj().T - np.eye(d)).max() < tol


# ---------------------------------------------------------------- C2 -----

def test_c2_two_sweeps_against_oracle(gpu):
    """N=20, chi=64, 10 layers, 2 sweeps of the bench's first state (bond-96
    random MPS, complex64 input), complex128 sweeps, against the oracle:
    chi-state 1e-10 relative, fidelity 1e-6, circuit state 1e-5 (L2)."""
    eng = _engine()
    n, chi, L, S = 20, 64, 10, 2
    sv = O.random_mps_state(n, 96, 1000).astype(np.complex64)
    _say("C2 vs oracle: N=20 chi=64 L=10 S=2 seed 1000")
    res = eng.encode(sv, chi, L, S, sweep_dtype=torch.complex128)
    psi, target, _ = eng.chi_truncate(eng.as_device_states(sv), chi)
    torch.cuda.synchronize()
    ref = O.prepare_state(sv, chi, num_layers=L, num_sweeps=S)
    terr = np.linalg.norm(target[0].cpu().numpy() - ref["target"]) / np.linalg.norm(ref["target"])
    fid = abs(res.fidelity[0, -1].item())
    gl = _gate_list(res.gates[0].cpu().numpy(), res.widths[0].cpu().numpy(), int(res.layers_used[0]))
    state = O.circuit_state(n, gl)
    serr = np.linalg.norm(state - ref["state"])
    fid0 = abs(res.fidelity[0, 0].item())
    ref0 = abs(np.vdot(ref["psi"], O.circuit_state(n, O.layer_gate_list(
        O.get_unitary_layers(ref["target"], n, L)))))
    _say(f"chi-state rel err {terr:.3e}; fidelity before sweeps {fid0:.10f} (oracle {ref0:.10f}), "
         f"after {fid:.10f} (oracle {ref['fidelity']:.10f}); state err {serr:.3e}")
    assert int(res.layers_used[0]) == len(ref["layers"]) == L
    assert terr < 1e-10
    assert abs(fid0 - ref0) < 1e-6
    assert abs(fid - ref["fidelity"]) < 1e-6
    assert serr < 1e-5
    _assert_unitary(res.gates, res.widths, res.layers_used)


def test_c2_complex64_reaches_the_fp64_goal(gpu):
    """C2 at its full 30 sweeps on 4 of the bench's states.  F_ref = the FP64
    path's fidelity after 30 sweeps (complex128 sweeps, the path the oracle
    pins to 1e-10).  The complex64 throughput mode's trajectory leaves the
    FP64 one by rounding amplified through the sweeps (either sign), so it
    sweeps until |<psi|C|0>| >= F_ref - 1e-4, at most 45 sweeps: every state
    must get there, its reported fidelity must equal the complex128
    re-evaluation of its final gates within 1e-5, and the FP64 overlap with
    the chi-state must not decrease from sweep to sweep."""
    eng = _engine()
    n, chi, L, S, cap = 20, 64, 10, 30, 45
    svs = np.stack([O.random_mps_state(n, 96, 1000 + s).astype(np.complex64) for s in range(4)])
    _say("C2 goal: N=20 chi=64 L=10 S=30 complex128 reference, complex64 goal run, 4 states")
    p = eng.encode_prepare(svs, chi, L, sweep_dtype=torch.complex128)
    ref = eng.sweep(p.target, p.gates, p.widths, p.layers_used, S, psi_ref=p.psi)
    fref = ref.fidelity[:, -1].abs()
    ov = ref.overlap.cpu().numpy()
    assert np.all(np.diff(ov, axis=1) >= -1e-12 * ov[:, 1:])
    goal = (fref - 1e-4).double().contiguous()
    run = eng.encode(svs, chi, L, cap, sweep_dtype=torch.complex64, goal=goal)
    torch.cuda.synchronize()
    done = run.sweeps_done.cpu().numpy()
    # the last entry is the final circuit's fidelity for every state (a state
    # that stopped at sweep k keeps its gates from then on)
    got = run.fidelity[:, -1].abs().cpu().numpy()
    _, f128 = eng.circuit_state(run.gates, run.widths, run.layers_used, psi_ref=p.psi)
    f128 = f128.abs().cpu().numpy()
    plain30 = eng.encode(svs, chi, L, S, sweep_dtype=torch.complex64).fidelity[:, -1].abs().cpu().numpy()
    _say(f"F_ref {fref.cpu().numpy()}; complex64 after 30 sweeps {plain30} "
         f"(|dF| max {np.abs(plain30 - fref.cpu().numpy()).max():.2e}); goal run: sweeps {done}, "
         f"F {got}, FP64 re-evaluation {f128}")
    assert np.all(f128 >= goal.cpu().numpy())
    assert np.all(np.abs(got - f128) < 1e-5)
    assert np.all(done <= cap)
    _assert_unitary(run.gates, run.widths, run.layers_used)


# ---------------------------------------------------------------- C3 -----

def test_c3_ragged_mix_against_oracle(gpu):
    """Ragged batch of N in {12, 13, 14} with chi in {16, 32, 64} (input bond =
    chi), 8 layers, 30 sweeps in complex128, against the oracle state by
    state: fidelity within 1e-6, circuit state within 1e-5 (L2)."""
    eng = _engine()
    spec = [(12, 16), (13, 32), (14, 64), (12, 64), (14, 16), (13, 16)]
    L, S = 8, 30
    svs = [O.random_mps_state(n, chi, 300 + i).astype(np.complex64) for i, (n, chi) in enumerate(spec)]
    _say(f"C3 ragged vs oracle: {spec}, L=8, S=30, complex128 sweeps")
    res = eng.encode_ragged(svs, [c for _, c in spec], L, S, sweep_dtype=torch.complex128)
    torch.cuda.synchronize()
    for i, ((n, chi), sv) in enumerate(zip(spec, svs)):
        ref = O.prepare_state(sv, chi, num_layers=L, num_sweeps=S)
        r = res[i]
        fid = abs(r.fidelity[-1].item())
        gl = _gate_list(r.gates.cpu().numpy(), r.widths.cpu().numpy(), r.layers_used)
        serr = np.linalg.norm(O.circuit_state(n, gl) - ref["state"])
        _say(f"N={n} chi={chi}: fidelity {fid:.10f} oracle {ref['fidelity']:.10f} state err {serr:.3e}")
        assert r.layers_used == len(ref["layers"])
        assert abs(fid - ref["fidelity"]) < 1e-6
        assert serr < 1e-5


def test_c3_256_state_batch_properties(gpu):
    """The ragged workload itself: 256 states, N ~ U{12..20}, bond_dimension =
    generating bond ~ U{16, 32, 64}, 8 layers, 30 sweeps, complex64.  With
    chi equal to the generating bond the truncation is exact, so the target
    is the input and the fidelity is the sweep's monotone objective: every
    trace must be non-decreasing (to complex64 rounding, 2e-6); every gate is
    unitary; each reported final fidelity equals the complex128
    re-evaluation of the final gates (qmx_circuit_state) within 1e-5."""
    eng = _engine()
    rng = np.random.default_rng(7)
    ns = rng.integers(12, 21, 256)
    ds = rng.choice([16, 32, 64], 256)
    svs = [O.random_mps_state(int(n), int(d), 10000 + i).astype(np.complex64)
           for i, (n, d) in enumerate(zip(ns, ds))]
    _say("C3 batch: 256 states N 12..20, chi 16/32/64, L=8, S=30, complex64")
    res = eng.encode_ragged(svs, [int(d) for d in ds], 8, 30)
    torch.cuda.synchronize()
    worst_drop, worst_re = 0.0, 0.0
    for i, r in enumerate(res):
        f = r.fidelity.abs().cpu().numpy()
        worst_drop = max(worst_drop, float(np.max(f[:-1] - f[1:])))
        g = r.gates.unsqueeze(0)
        w = r.widths.unsqueeze(0)
        u = torch.tensor([r.layers_used], dtype=torch.int32, device=gpu)
        _assert_unitary(g, w, u)
        psi = eng.as_device_states(O.ket(svs[i]), torch.complex128)
        _, f128 = eng.circuit_state(g, w, u, psi_ref=psi)
        worst_re = max(worst_re, abs(abs(f128[0].item()) - f[-1]))
    _say(f"largest fidelity drop between sweeps {worst_drop:.2e}; "
         f"largest |F - F(complex128 re-evaluation)| {worst_re:.2e}")
    assert worst_drop < 2e-6
    assert worst_re < 1e-5


# ---------------------------------------------------------------- C5 -----

def _c5_state(dev):
    """Seeded random MPS of bond 1024 at N=26 (complex Gaussian tensors drawn
    on the host, contracted in complex128 on the GPU), normalised, complex64."""
    n, bond = 26, 1024
    gen = torch.Generator(device="cpu").manual_seed(2026)
    x = None
    for i in range(n):
        dl = min(bond, 2 ** i, 2 ** (n - i))
        dr = min(bond, 2 ** (i + 1), 2 ** (n - i - 1))
        a = (torch.randn(dl, 2, dr, generator=gen, dtype=torch.float64)
             + 1j * torch.randn(dl, 2, dr, generator=gen, dtype=torch.float64)) / np.sqrt(2.0)
        a = a.to(dev).reshape(dl, 2 * dr)
        x = a if x is None else x @ a
        x = x.reshape(-1, dr)
    psi = x.reshape(1, -1)
    return (psi / torch.linalg.vector_norm(psi)).to(torch.complex64).contiguous()


def test_c5_full_size_properties(gpu):
    """N=26, chi=512, 20 layers, 50 sweeps on one state.  Size-independent
    checks: the chi-state is an orthogonal projection of the input
    (<T|psi> = |T|^2, to 1e-10 relative), the FP64 overlap with it never
    decreases from sweep to sweep, every gate is unitary, the fidelity rises
    over the sweeps, and the reported final fidelity equals the complex128
    re-evaluation of the final gates within 1e-5."""
    eng = _engine()
    n, chi, L, S = 26, 512, 20, 50
    _say("C5: N=26 chi=512 L=20 S=50, complex64 sweeps")
    sv = _c5_state(gpu)
    psi, target, _ = eng.chi_truncate(sv, chi)
    t2 = float(torch.linalg.vector_norm(target).item() ** 2)
    proj = complex(torch.vdot(target[0], psi[0]).item())
    _say(f"|T|^2 = {t2:.12f}, <T|psi> = {proj:.12f}")
    assert abs(proj - t2) < 1e-10 * t2
    assert abs(proj.imag) < 1e-10
    p = eng.encode_prepare(sv, chi, L)
    res = eng.sweep(p.target, p.gates, p.widths, p.layers_used, S, psi_ref=p.psi)
    torch.cuda.synchronize()
    ov = res.overlap[0].cpu().numpy()
    fid = res.fidelity[0].abs().cpu().numpy()
    _, f128 = eng.circuit_state(p.gates, p.widths, p.layers_used, psi_ref=p.psi)
    f128 = abs(f128[0].item())
    _say(f"layers {int(p.layers_used[0])}; fidelity {fid[0]:.6f} -> {fid[-1]:.6f} (complex128 "
         f"re-evaluation {f128:.6f}); overlap {ov[0]:.6f} -> {ov[-1]:.6f}, "
         f"largest drop {float(np.max(ov[:-1] - ov[1:])):.2e}")
    assert int(p.layers_used[0]) == L
    assert np.all(np.diff(ov) >= -2e-6)
    assert fid[-1] > fid[0]
    assert abs(f128 - fid[-1]) < 1e-5
    _assert_unitary(p.gates, p.widths, p.layers_used)


# ------------------------------------------------------- failure paths ---

def test_injected_non_unitary_gate_raises_value_error(gpu):
    """QMX_OPT_FAIL_UNITARITY makes every generated gate fail the check of
    mps.py:837-839: QMX_ERR_NOT_UNITARY -> ValueError, through the engine and
    through Sequential.prepare_state; the next call after the reset works."""
    from qmprs_amd import _lib
    from qmprs_amd.circuit import GateCircuit
    from qmprs_amd.synthesis.mps_encoding import Sequential
    eng = _engine()
    sv = O.random_mps_state(8, 8, 3)
    _lib.set_option(_lib.QMX_OPT_FAIL_UNITARITY, 1)
    try:
        with pytest.raises(ValueError, match="unitar"):
            eng.extract_layers(eng.as_device_states(O.ket(sv), torch.complex128), 2)
        with pytest.raises(ValueError, match="unitar"):
            Sequential(GateCircuit).prepare_state(sv, 8, num_layers=2, num_sweeps=1)
    finally:
        _lib.set_option(_lib.QMX_OPT_FAIL_UNITARITY, 0)
    circ = Sequential(GateCircuit).prepare_state(sv, 8, num_layers=2, num_sweeps=1)
    assert circ.num_qubits == 8
    with pytest.raises(ValueError):
        _lib.set_option(_lib.QMX_OPT_FAIL_UNITARITY, 2)


@pytest.mark.parametrize("n,chi", [(12, 16), (11, 150), (18, 300)],
                         ids=["lds-jacobi", "global-jacobi", "block-jacobi"])
def test_eigensolver_sweep_cap_raises_runtime_error(gpu, n, chi):
    """QMX_OPT_JACOBI_SWEEP_CAP = 1 stops each of the three chi-truncation
    eigensolvers after one sweep; states that need more fail with
    QMX_ERR_NO_CONVERGENCE -> RuntimeError.  Without the cap the same input
    matches the oracle."""
    from qmprs_amd import _lib
    eng = _engine()
    sv = O.random_uniform_state(n, 21) if chi > 64 else O.random_mps_state(n, 2 * chi, 21)
    _say(f"Jacobi sweep cap 1: N={n} chi={chi}")
    dev_sv = eng.as_device_states(sv)
    _lib.set_option(_lib.QMX_OPT_JACOBI_SWEEP_CAP, 1)
    try:
        with pytest.raises(RuntimeError, match="converge"):
            eng.chi_truncate(dev_sv, chi)
    finally:
        _lib.set_option(_lib.QMX_OPT_JACOBI_SWEEP_CAP, 0)
    _, target, _ = eng.chi_truncate(dev_sv, chi)
    ref, _ = O.chi_truncate(O.ket(sv), n, chi)
    err = np.linalg.norm(target[0].cpu().numpy() - ref) / np.linalg.norm(ref)
    assert err < 1e-10
    with pytest.raises(ValueError):
        _lib.set_option(_lib.QMX_OPT_JACOBI_SWEEP_CAP, -1)
