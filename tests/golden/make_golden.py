"""Write the golden fixtures of the encoder (run from the repository root).

    python tests/golden/make_golden.py

The reference ships no golden vectors and cannot run here (quimb and
quick-core are absent), so the fixtures are produced by the FP64 oracle on
seeded inputs and committed as small .npz files.  They freeze the oracle's
outputs: the CPU suite checks that the oracle still reproduces them and the
GPU suite checks the HIP path against them.

Each case stores its inputs (statevector, sizes) and outputs: the chi-MPS
state and bond spectra, the extracted layers as gate arrays in extraction
order (gates[l, site] 4x4 out x in, width 1 gates in the top-left 2x2), the
per-layer <0|psi> amplitudes, the swept gates, the fidelity trace and the
final circuit state C|0..0>.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import qmprs_oracle as O  # noqa: E402

# name -> (num_qubits, bond_dimension, num_layers, num_sweeps, kind, seed, generator bond)
CASES = {
    # the reference's CPU-runnable headline case: N=10, U[0,1) complex128,
    # chi=512 (exact), 15 layers, 50 sweeps
    "c1_n10_chi512_l15_s50": (10, 512, 15, 50, "uniform", 0, None),
    "n6_chi8_l3_s4": (6, 8, 3, 4, "uniform", 1, None),
    "n9_chi8_l4_s3_mps": (9, 8, 4, 3, "mps", 2, 12),
    "n12_chi16_l3_s2_mps": (12, 16, 3, 2, "mps", 3, 24),
}


def _arrays(layers, n):
    """Circuit-order layers -> (gates (L, N, 4, 4), widths (L, N)) in extraction order."""
    nl = len(layers)
    gates = np.zeros((nl, n, 4, 4), dtype=np.complex128)
    widths = np.ones((nl, n), dtype=np.int32)
    for l, layer in enumerate(reversed(layers)):
        for start, end, mats in layer:
            for j in range(start, end + 1):
                m = mats[j - start]
                d = m.shape[0]
                gates[l, j, :d, :d] = m
                widths[l, j] = 2 if d == 4 else 1
    return gates, widths


def make_case(n, chi, num_layers, num_sweeps, kind, seed, gen_bond):
    if kind == "uniform":
        sv = O.random_uniform_state(n, seed)
    else:
        sv = O.random_mps_state(n, gen_bond, seed)
    psi = O.ket(sv)
    target, spectra = O.chi_truncate(psi, n, chi)
    layers, info = O.get_unitary_layers(target, n, num_layers, return_info=True)
    g0, w0 = _arrays(layers, n)
    gl = O.layer_gate_list(layers)
    fid = [abs(np.vdot(psi, O.circuit_state(n, gl)))]
    overlaps = []
    for _ in range(num_sweeps):
        gl, ov = O.sweep_once(target, n, gl)
        overlaps.append(abs(ov))
        fid.append(abs(np.vdot(psi, O.circuit_state(n, gl))))
    swept = O.gate_list_to_layers(gl, layers)
    g1, _ = _arrays(swept, n)
    state = O.circuit_state(n, gl)
    spec = np.zeros((n + 1, max(len(s) for s in spectra if s is not None) if any(
        s is not None for s in spectra) else 1))
    for i, s in enumerate(spectra):
        if s is not None and i < spec.shape[0]:
            spec[i, :len(s)] = s
    return dict(
        statevector=sv.astype(np.complex128), num_qubits=n, bond_dimension=chi,
        num_layers=num_layers, num_sweeps=num_sweeps,
        target=target, spectra=spec, layers_used=len(layers),
        gates=g0, widths=w0, amp0=np.asarray(info["amp0"], dtype=np.complex128),
        swept_gates=g1, fidelity=np.asarray(fid), overlap=np.asarray(overlaps),
        state=state)


def main():
    for name, args in CASES.items():
        out = make_case(*args)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print(f"{name}: layers {out['layers_used']} fidelity {out['fidelity'][0]:.9f} -> "
              f"{out['fidelity'][-1]:.9f}  ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
