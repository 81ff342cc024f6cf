"""Encoder ABC (mirror of qmprs/synthesis/mps_encoding/base.py)."""

from __future__ import annotations

__all__ = ["MPSEncoder"]

from abc import ABC, abstractmethod
from typing import Literal

import numpy as np
import torch

from qmprs_amd import engine
from qmprs_amd.primitives import MPS


class MPSEncoder(ABC):
    """``MPSEncoder(circuit_framework)`` (base.py:32-56).

    ``circuit_framework`` is any class following the circuit protocol
    ``cls(num_qubits)``, ``.num_qubits``, ``.unitary(matrix, qubits)`` (for
    example a quick-core circuit class or ``qmprs_amd.circuit.GateCircuit``).
    """

    def __init__(self, circuit_framework) -> None:
        self.circuit_framework = circuit_framework

    def prepare_state(self, statevector, bond_dimension: int, compression_percentage: float = 0.0,
                      index_type: Literal["row", "snake"] = "row", **kwargs):
        """base.py:58-106: statevector -> re-indexing / compression (quick Ket
        semantics, on the GPU) -> MPS (chi-truncated on the GPU) -> circuit.

        ``index_type="snake"`` reverses the second half of the amplitudes
        (N >= 3) and ``compression_percentage`` > 0 zeroes the
        int(pct / 100 * 2^N) smallest-magnitude amplitudes, the rules assumed
        for quick's ``Ket.change_indexing`` / ``Ket.compress``."""
        if index_type not in ("row", "snake"):
            raise ValueError("Index type not supported.")
        if hasattr(statevector, "data") and not isinstance(statevector, (np.ndarray, torch.Tensor)):
            statevector = statevector.data
        if index_type == "snake" or compression_percentage > 0.0:
            statevector = engine.prepare_input(statevector, index_type, compression_percentage)
        mps = MPS(statevector, bond_dimension=bond_dimension)
        return self.prepare_mps(mps, **kwargs)

    @abstractmethod
    def prepare_mps(self, mps: MPS, **kwargs):
        """base.py:108-131."""
