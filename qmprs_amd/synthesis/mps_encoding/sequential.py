"""Sequential encoder (mirror of qmprs/synthesis/mps_encoding/sequential.py).

Ran's analytic disentangling layers followed by Rudolph et al.'s environment
sweeps (DallOall), with every arithmetic step on the GPU through libqmx:

* ``_get_unitary_layers``      sequential.py:330-398 -> qmx_extract_layers
* ``_optimize_unitary_layers`` sequential.py:509-541 -> qmx_sweep
* ``_circuit_from_unitary_layers`` sequential.py:189-213 (host: gate emission)
"""

from __future__ import annotations

__all__ = ["Sequential"]

import torch

from qmprs_amd import _lib, engine
from qmprs_amd.primitives.mps import MPS, UnitaryLayer
from qmprs_amd.synthesis.mps_encoding.base import MPSEncoder


class Sequential(MPSEncoder):
    """``Sequential(circuit_framework)`` (sequential.py:31-153).

    Extra keyword arguments of ``prepare_mps`` / ``prepare_state`` beyond the
    reference's ``num_layers`` and ``num_sweeps``:
    ``sweep_dtype`` (torch.complex128 default, the FP64 parity mode that
    reproduces the reference's sweeps; torch.complex64 stores the sweep's
    dense states in complex64 with FP32 gate arithmetic and FP64 reductions,
    twice the throughput, whose gate trajectory leaves the FP64 one by up to
    ~1e-3 in fidelity after 30 sweeps at N=20 -- rounding amplified by the
    sweep dynamics, with either sign);
    ``schedule`` ("DallOall" default, the reference's method; "IterDiOi" or
    "IterDiOall", the iterative variants its notebook names) -- see
    ``qmprs_amd.engine.schedule_layers``;
    ``sweep_relaxation`` (1.0 default, the reference's gate update; 1 < omega
    < 1.95) -- over-relaxed sweep updates polar(G' + (omega - 1)(G' - G)),
    which reach a given fidelity in fewer sweeps but leave the reference's
    gate trajectory.
    """

    def __init__(self, circuit_framework) -> None:
        super().__init__(circuit_framework)
        self._fidelity_threshold = 1 - 1e-6
        self.last_info: dict = {}

    @property
    def fidelity_threshold(self) -> float:
        return self._fidelity_threshold

    @fidelity_threshold.setter
    def fidelity_threshold(self, threshold: float) -> None:
        if not isinstance(threshold, (int, float)) or threshold < 0 or threshold > 1:
            raise ValueError("The fidelity threshold must be a float between 0 and 1.")
        self._fidelity_threshold = threshold

    # -- emission (sequential.py:155-213) -----------------------------------
    @staticmethod
    def _apply_unitary_layer_to_circuit(circuit, unitary_layer: UnitaryLayer) -> None:
        """Gates go to qubit N-1-site (LSB convention of the circuit backend):
        a two-site gate on sites (i, i+1) to qubits [N-2-i, N-1-i]."""
        n = circuit.num_qubits
        for start, end, block in unitary_layer:
            for site in range(start, end + 1):
                matrix = block[site - start].data
                if site == end:
                    circuit.unitary(matrix, n - 1 - site)
                else:
                    circuit.unitary(matrix, [n - 2 - site, n - 1 - site])

    def _circuit_from_unitary_layers(self, mps: MPS, unitary_layers: list[UnitaryLayer]):
        circuit = self.circuit_framework(mps.num_sites)
        for layer in unitary_layers:
            Sequential._apply_unitary_layer_to_circuit(circuit, layer)
        return circuit

    # -- GPU phases ---------------------------------------------------------
    def _device_layers(self, mps: MPS, num_layers: int):
        work = (mps.state / mps.norm).contiguous()
        gates, widths, used, amp0, _ = engine.extract_layers(
            work, num_layers, self.fidelity_threshold)
        return gates, widths, used, amp0

    def _get_unitary_layers(self, mps: MPS, num_layers: int) -> list[UnitaryLayer]:
        gates, widths, used, _ = self._device_layers(mps, num_layers)
        return engine.gates_to_layers(gates[0].cpu().numpy(), widths[0].cpu().numpy(),
                                      int(used[0].item()))

    def _optimize_unitary_layers(self, mps: MPS, unitary_layers: list[UnitaryLayer],
                                 num_sweeps: int, sweep_dtype=torch.complex128) -> list[UnitaryLayer]:
        g, w = engine.layers_to_arrays(unitary_layers, mps.num_sites)
        dev = mps.state.device
        gates = torch.from_numpy(g).to(dev).unsqueeze(0).contiguous()
        widths = torch.from_numpy(w).to(dev).unsqueeze(0).contiguous()
        used = torch.tensor([g.shape[0]], dtype=torch.int32, device=dev)
        target = mps.state.to(sweep_dtype).contiguous()
        res = engine.sweep(target, gates, widths, used, num_sweeps, psi_ref=mps._psi)
        self.last_info = {"fidelity_trace": res.fidelity[0].abs().cpu().numpy(),
                          "overlap_trace": res.overlap[0].cpu().numpy()}
        return engine.gates_to_layers(gates[0].cpu().numpy(), widths[0].cpu().numpy(), g.shape[0])

    def _sequential_unitary_circuit(self, mps: MPS, num_layers: int, num_sweeps: int = 0,
                                    sweep_dtype=torch.complex128):
        """sequential.py:543-586 (DallOall) with the gates kept on the GPU
        between the layer and sweep phases."""
        gates, widths, used, amp0 = self._device_layers(mps, num_layers)
        target = mps.state.to(sweep_dtype).contiguous()
        res = engine.sweep(target, gates, widths, used, num_sweeps, psi_ref=mps._psi)
        fid = res.fidelity[0].abs().cpu().numpy()
        self.last_info = {"fidelity": float(fid[-1]), "fidelity_trace": fid,
                          "overlap_trace": res.overlap[0].cpu().numpy(),
                          "layers_used": int(used[0].item()),
                          "amp0": amp0[0].cpu().numpy()}
        layers = engine.gates_to_layers(gates[0].cpu().numpy(), widths[0].cpu().numpy(),
                                        int(used[0].item()))
        return self._circuit_from_unitary_layers(mps, layers)

    def _scheduled_circuit(self, mps: MPS, num_layers: int, num_sweeps: int, schedule: str,
                           sweep_dtype=torch.complex128):
        """IterDiOi / IterDiOall on the GPU (engine.schedule_layers)."""
        psi = mps._psi.reshape(1, -1).to(torch.complex128).contiguous()
        target = mps.state.reshape(1, -1).to(torch.complex128).contiguous()
        gates, widths, used, fid = engine.schedule_layers(
            psi, target, num_layers, num_sweeps, schedule, self.fidelity_threshold,
            sweep_dtype=sweep_dtype)
        self.last_info = {"fidelity": float(fid[0].item()), "layers_used": int(used[0].item()),
                          "schedule": schedule}
        layers = engine.gates_to_layers(gates[0].cpu().numpy(), widths[0].cpu().numpy(),
                                        int(used[0].item()))
        return self._circuit_from_unitary_layers(mps, layers)

    def prepare_mps(self, mps: MPS, **kwargs):
        """sequential.py:588-600."""
        num_layers = kwargs.get("num_layers", 1)
        num_sweeps = kwargs.get("num_sweeps", 0)
        if not isinstance(num_layers, int) or num_layers < 1:
            raise ValueError("The number of layers must be a positive integer.")
        if not isinstance(num_sweeps, int) or num_sweeps < 0:
            raise ValueError("The number of sweeps must be a non-negative integer.")
        schedule = kwargs.get("schedule", "DallOall")
        if schedule not in engine.SCHEDULES:
            raise ValueError(f"schedule must be one of {engine.SCHEDULES}")
        sweep_dtype = kwargs.get("sweep_dtype", torch.complex128)
        if sweep_dtype not in (torch.complex64, torch.complex128):
            raise ValueError("sweep_dtype must be torch.complex64 or torch.complex128.")
        omega = kwargs.get("sweep_relaxation", 1.0)
        if isinstance(omega, bool) or not isinstance(omega, (int, float)) or not 1.0 <= omega <= 1.95:
            raise ValueError("sweep_relaxation must be a number in [1, 1.95].")
        if omega == 1.0:
            return self._encode(mps, num_layers, num_sweeps, schedule, sweep_dtype)
        _lib.set_sweep_relaxation(float(omega))
        try:
            return self._encode(mps, num_layers, num_sweeps, schedule, sweep_dtype)
        finally:
            _lib.set_sweep_relaxation(1.0)

    def _encode(self, mps: MPS, num_layers: int, num_sweeps: int, schedule: str, sweep_dtype):
        if schedule != "DallOall":
            return self._scheduled_circuit(mps, num_layers, num_sweeps, schedule, sweep_dtype)
        return self._sequential_unitary_circuit(mps, num_layers, num_sweeps, sweep_dtype)
