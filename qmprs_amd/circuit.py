"""Minimal circuit backend implementing the protocol the encoder emits into.

The reference emits gates into a quick-core ``Circuit`` subclass
(``circuit_framework(num_qubits)``, ``.num_qubits``, ``.unitary(matrix,
qubits)``; sequential.py:155-213) and its tests read ``.get_statevector()`` and
``.get_depth()`` (tests/synthesis/mps_encoding/test_sequential_encoding.py).
quick-core is not part of this package; any class with that protocol can be
passed to ``Sequential``.  ``GateCircuit`` is a dependency-free one.

Conventions follow quick/qiskit: qubit q is bit q of the amplitude index (LSB
first), and ``unitary(M, [q0, q1])`` treats q0 as the least significant bit of
M's row/column index.
"""

from __future__ import annotations

import numpy as np


class GateCircuit:
    """Records unitaries; simulates them on demand for read-back."""

    def __init__(self, num_qubits: int) -> None:
        if not isinstance(num_qubits, int) or num_qubits < 1:
            raise ValueError("num_qubits must be a positive integer")
        self.num_qubits = num_qubits
        self.gates: list[tuple[np.ndarray, tuple[int, ...]]] = []

    # -- gate emission -----------------------------------------------------
    def unitary(self, matrix, qubit_indices) -> None:
        qubits = (qubit_indices,) if isinstance(qubit_indices, (int, np.integer)) else tuple(qubit_indices)
        m = np.asarray(matrix, dtype=np.complex128)
        if m.shape != (2 ** len(qubits), 2 ** len(qubits)):
            raise ValueError("matrix shape does not match the number of qubits")
        if any(q < 0 or q >= self.num_qubits for q in qubits) or len(set(qubits)) != len(qubits):
            raise ValueError("invalid qubit indices")
        self.gates.append((m, tuple(int(q) for q in qubits)))

    def H(self, q: int) -> None:
        self.unitary(np.array([[1, 1], [1, -1]]) / np.sqrt(2.0), q)

    def X(self, q: int) -> None:
        self.unitary(np.array([[0, 1], [1, 0]]), q)

    def CX(self, control: int, target: int) -> None:
        m = np.eye(4, dtype=np.complex128)
        # basis index = 2 * bit(target) + bit(control) with qubits [control, target]
        m[[1, 3]] = m[[3, 1]]
        self.unitary(m, [control, target])

    # -- read-back ---------------------------------------------------------
    def get_statevector(self) -> np.ndarray:
        n = self.num_qubits
        psi = np.zeros(2 ** n, dtype=np.complex128)
        psi[0] = 1.0
        for m, qubits in self.gates:
            k = len(qubits)
            # axes of psi reshaped to [2]*n in MSB-first order: qubit q <-> axis n-1-q
            t = psi.reshape([2] * n)
            axes = [n - 1 - q for q in reversed(qubits)]   # MSB of M first
            g = m.reshape([2] * (2 * k))
            t = np.tensordot(g, t, axes=(list(range(k, 2 * k)), axes))
            t = np.moveaxis(t, list(range(k)), axes)
            psi = t.reshape(-1)
        return psi

    def get_depth(self) -> int:
        """Depth counting each recorded unitary as one layer on its qubits."""
        level = [0] * self.num_qubits
        for _, qubits in self.gates:
            d = max(level[q] for q in qubits) + 1
            for q in qubits:
                level[q] = d
        return max(level) if level else 0

    def __len__(self) -> int:
        return len(self.gates)
