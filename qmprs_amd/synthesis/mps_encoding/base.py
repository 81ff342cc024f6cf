"""Encoder ABC (mirror of qmprs/synthesis/mps_encoding/base.py)."""

from __future__ import annotations

__all__ = ["MPSEncoder"]

from abc import ABC, abstractmethod
from typing import Literal

import numpy as np

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
        """base.py:58-106: statevector -> MPS (chi-truncated on the GPU) -> circuit."""
        if index_type != "row":
            raise ValueError("Only index_type='row' is supported.")
        if compression_percentage > 0.0:
            raise ValueError("compression_percentage > 0 is not supported.")
        if hasattr(statevector, "data") and not isinstance(statevector, np.ndarray):
            statevector = statevector.data
        mps = MPS(statevector, bond_dimension=bond_dimension)
        return self.prepare_mps(mps, **kwargs)

    @abstractmethod
    def prepare_mps(self, mps: MPS, **kwargs):
        """base.py:108-131."""
