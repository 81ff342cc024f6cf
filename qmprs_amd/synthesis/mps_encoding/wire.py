"""Gate-list wire format and circuit emission.

The reference hands the unitary layers straight to a quick-core circuit
(sequential.py:155-213) and has no serialised form.  This module gives the
``list[UnitaryLayer]`` (mps.py:46-47) a file format and emits it into any
circuit backend:

npz (no pickles), format version 1:
    gates    (K, 4, 4) complex128   each gate out x in; one-qubit gates in the
                                    top-left 2x2, zeros elsewhere
    widths   (K,) int32             1 or 2
    sites    (K,) int32             first MPS site of the gate
    layers   (K,) int32             index of the layer in circuit order
    blocks   (M, 3) int32           (layer, start, end) of every block
    qubits   (K, 2) int32           circuit qubits: [N-2-site, N-1-site] for a
                                    two-qubit gate, [N-1-site, -1] otherwise
    num_qubits ()                   N
    version  ()                     1
Gates are stored in circuit order (the order ``emit`` applies them).
"""

from __future__ import annotations

import numpy as np

from qmprs_amd.primitives.mps import Tensor, UnitaryLayer

VERSION = 1


def to_arrays(layers: list[UnitaryLayer], num_qubits: int) -> dict:
    gates, widths, sites, lidx, blocks, qubits = [], [], [], [], [], []
    for l, layer in enumerate(layers):
        for start, end, mats in layer:
            blocks.append((l, start, end))
            for site in range(start, end + 1):
                m = np.asarray(getattr(mats[site - start], "data", mats[site - start]), dtype=np.complex128)
                g = np.zeros((4, 4), dtype=np.complex128)
                g[:m.shape[0], :m.shape[1]] = m
                w = 2 if m.shape[0] == 4 else 1
                gates.append(g)
                widths.append(w)
                sites.append(site)
                lidx.append(l)
                qubits.append((num_qubits - 2 - site, num_qubits - 1 - site) if w == 2
                              else (num_qubits - 1 - site, -1))
    return {
        "gates": np.array(gates, dtype=np.complex128).reshape(-1, 4, 4),
        "widths": np.array(widths, dtype=np.int32),
        "sites": np.array(sites, dtype=np.int32),
        "layers": np.array(lidx, dtype=np.int32),
        "blocks": np.array(blocks, dtype=np.int32).reshape(-1, 3),
        "qubits": np.array(qubits, dtype=np.int32).reshape(-1, 2),
        "num_qubits": np.int32(num_qubits),
        "version": np.int32(VERSION),
    }


def from_arrays(arrays: dict) -> tuple[list[UnitaryLayer], int]:
    if int(arrays["version"]) != VERSION:
        raise ValueError(f"unsupported gate-list format version {int(arrays['version'])}")
    n = int(arrays["num_qubits"])
    gates, widths, sites = arrays["gates"], arrays["widths"], arrays["sites"]
    nl = int(arrays["layers"].max()) + 1 if len(arrays["layers"]) else 0
    layers: list[UnitaryLayer] = [[] for _ in range(nl)]
    k = 0
    for l, start, end in arrays["blocks"]:
        mats = []
        for site in range(int(start), int(end) + 1):
            if int(sites[k]) != site:
                raise ValueError("gate list and block table disagree")
            d = 4 if int(widths[k]) == 2 else 2
            mats.append(Tensor(np.array(gates[k, :d, :d])))
            k += 1
        layers[int(l)].append((int(start), int(end), mats))
    if k != len(gates):
        raise ValueError("gate list and block table disagree")
    return layers, n


def save(path, layers: list[UnitaryLayer], num_qubits: int) -> None:
    np.savez_compressed(path, **to_arrays(layers, num_qubits))


def load(path) -> tuple[list[UnitaryLayer], int]:
    with np.load(path, allow_pickle=False) as z:
        return from_arrays({k: z[k] for k in z.files})


def emit(layers: list[UnitaryLayer], num_qubits: int, circuit_framework):
    """Gates into ``circuit_framework(num_qubits)`` through ``.unitary``
    (sequential.py:173-213): a two-qubit gate at site i on qubits
    [N-2-i, N-1-i], a one-qubit gate on N-1-i."""
    circuit = circuit_framework(num_qubits)
    for layer in layers:
        for start, end, mats in layer:
            for site in range(start, end + 1):
                m = np.asarray(getattr(mats[site - start], "data", mats[site - start]))
                if m.shape[0] == 4:
                    circuit.unitary(m, [num_qubits - 2 - site, num_qubits - 1 - site])
                else:
                    circuit.unitary(m, num_qubits - 1 - site)
    return circuit


def to_qiskit(layers: list[UnitaryLayer], num_qubits: int):
    """A qiskit QuantumCircuit with the same gates (qiskit orders qubits LSB
    first like the circuit protocol).  Raises ImportError without qiskit."""
    from qiskit import QuantumCircuit   # optional dependency

    qc = QuantumCircuit(num_qubits)
    for layer in layers:
        for start, end, mats in layer:
            for site in range(start, end + 1):
                m = np.asarray(getattr(mats[site - start], "data", mats[site - start]))
                if m.shape[0] == 4:
                    qc.unitary(m, [num_qubits - 2 - site, num_qubits - 1 - site])
                else:
                    qc.unitary(m, [num_qubits - 1 - site])
    return qc
