"""Batched GPU engine: thin Python wrappers over the libqmx C-ABI.

Every function takes and returns torch tensors on the current ROCm device and
enqueues its kernels on the current stream.  A batch holds states of equal
qubit count, one per row.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from qmprs_amd import _lib

CUTOFF = 1e-10
DEFAULT_THRESHOLD = 1 - 1e-6

_DTYPES = {torch.complex64: _lib.QMX_C64, torch.complex128: _lib.QMX_C128}


def num_qubits_of(length: int) -> int:
    n = int(length).bit_length() - 1
    if length < 2 or (1 << n) != length:
        raise ValueError("The statevector length must be a power of two.")
    return n


def as_device_states(statevectors, dtype=None) -> torch.Tensor:
    """(B, 2^N) complex tensor on the GPU (accepts numpy / torch / 1-D)."""
    dev = _lib.require_gpu()
    if isinstance(statevectors, torch.Tensor):
        t = statevectors
    else:
        arr = np.asarray(statevectors)
        if not np.iscomplexobj(arr):
            arr = arr.astype(np.complex128)
        t = torch.from_numpy(np.ascontiguousarray(arr))
    if t.dim() == 1:
        t = t.unsqueeze(0)
    if not t.is_complex():
        t = t.to(torch.complex128)
    if dtype is not None:
        t = t.to(dtype)
    if t.dtype not in _DTYPES:
        t = t.to(torch.complex128)
    return t.to(dev).contiguous()


INDEX_TYPES = {"row": _lib.QMX_INDEX_ROW, "snake": _lib.QMX_INDEX_SNAKE}


def prepare_input(statevectors, index_type: str = "row", compression_percentage: float = 0.0) -> torch.Tensor:
    """quick Ket re-indexing and compression of prepare_state's input
    (base.py:99-102) on the GPU (qmx_prepare_input): a new (B, 2^N) tensor of
    the input's complex dtype."""
    if index_type not in INDEX_TYPES:
        raise ValueError("Index type not supported.")
    sv = as_device_states(statevectors)
    b, length = sv.shape
    n = num_qubits_of(length)
    lib, h = _lib.load(), _lib.handle()
    out = torch.empty_like(sv)
    ws = _lib.workspace(lib.qmx_prepare_input_workspace(n, b), sv.device)
    st = lib.qmx_prepare_input(h, n, b, _lib.ptr(sv), _DTYPES[sv.dtype], INDEX_TYPES[index_type],
                               float(compression_percentage), _lib.ptr(out), _lib.ptr(ws), ws.numel(),
                               _lib.stream_ptr())
    _lib.check(st, h)
    return out


def mps_to_state(tensors) -> torch.Tensor:
    """Dense state (1, 2^N) complex128 of 'lpr' site tensors (l, 2, r) with
    end bonds 1 (qmx_mps_to_state: the MPS(mps=...) entry, mps.py:183-197)."""
    ts = [np.asarray(a, dtype=np.complex128) for a in tensors]
    n = len(ts)
    if n < 2 or any(a.ndim != 3 for a in ts):
        raise ValueError("an MPS of at least 2 site tensors of shape (l, 2, r) is needed")
    if any(a.shape[1] != 2 for a in ts):
        raise ValueError("Only supports MPS with physical dimension of 2.")
    bonds = [ts[0].shape[0]] + [a.shape[2] for a in ts]
    if bonds[0] != 1 or bonds[-1] != 1 or any(ts[i].shape[0] != bonds[i] for i in range(n)):
        raise ValueError("site tensor bond dimensions do not match (end bonds must be 1)")
    dev = _lib.require_gpu()
    flat = torch.from_numpy(np.concatenate([a.reshape(-1) for a in ts])).to(dev)
    bd = (ctypes.c_int * (n + 1))(*bonds)
    lib, h = _lib.load(), _lib.handle()
    need = lib.qmx_mps_to_state_workspace(n, bd)
    if need == 0:
        raise ValueError("invalid bond dimensions")
    ws = _lib.workspace(need, dev)
    out = torch.empty((1, 1 << n), dtype=torch.complex128, device=dev)
    st = lib.qmx_mps_to_state(h, n, bd, _lib.ptr(flat), _lib.ptr(out), _lib.ptr(ws), ws.numel(),
                              _lib.stream_ptr())
    _lib.check(st, h)
    return out


def chi_truncate(states: torch.Tensor, bond_dimension: int, cutoff: float = CUTOFF,
                 spectra: bool = False):
    """(a)+(b): normalised inputs and the chi-MPS states (complex128)."""
    b, length = states.shape
    n = num_qubits_of(length)
    lib, h = _lib.load(), _lib.handle()
    psi = torch.empty((b, length), dtype=torch.complex128, device=states.device)
    target = torch.empty_like(psi)
    spec = (torch.zeros((b, n + 1, 2 * bond_dimension), dtype=torch.float64, device=states.device)
            if spectra else None)
    ws = _lib.workspace(lib.qmx_chi_truncate_workspace(n, b, bond_dimension), states.device)
    st = lib.qmx_chi_truncate(h, n, b, bond_dimension, cutoff, _lib.ptr(states), _DTYPES[states.dtype],
                              _lib.ptr(psi), _lib.ptr(target), _lib.ptr(spec), _lib.ptr(ws), ws.numel(),
                              _lib.stream_ptr())
    _lib.check(st, h)
    return psi, target, spec


def extract_layers(psi_work: torch.Tensor, num_layers: int,
                   fidelity_threshold: float = DEFAULT_THRESHOLD, cutoff: float = CUTOFF,
                   with_bonds: bool = False):
    """(b)-(d): layer loop of _get_unitary_layers on normalised states (in place)."""
    b, length = psi_work.shape
    n = num_qubits_of(length)
    lib, h = _lib.load(), _lib.handle()
    dev = psi_work.device
    gates = torch.zeros((b, num_layers, n, 4, 4), dtype=torch.complex128, device=dev)
    widths = torch.ones((b, num_layers, n), dtype=torch.int32, device=dev)
    bonds = torch.zeros((b, num_layers, n), dtype=torch.int32, device=dev) if with_bonds else None
    used = torch.zeros(b, dtype=torch.int32, device=dev)
    amp0 = torch.zeros((b, num_layers), dtype=torch.complex128, device=dev)
    ws = _lib.workspace(lib.qmx_extract_layers_workspace(n, b), dev)
    st = lib.qmx_extract_layers(h, n, b, num_layers, float(fidelity_threshold), cutoff,
                                _lib.ptr(psi_work), _lib.ptr(gates), _lib.ptr(widths), _lib.ptr(bonds),
                                _lib.ptr(used), _lib.ptr(amp0), _lib.ptr(ws), ws.numel(),
                                _lib.stream_ptr())
    _lib.check(st, h)
    return gates, widths, used, amp0, bonds


def generate_layer(psi_work: torch.Tensor, cutoff: float = CUTOFF):
    """(b)+(c): one Ran layer (bond-2 compression + unitary completion)."""
    b, length = psi_work.shape
    n = num_qubits_of(length)
    lib, h = _lib.load(), _lib.handle()
    dev = psi_work.device
    gates = torch.zeros((b, 1, n, 4, 4), dtype=torch.complex128, device=dev)
    widths = torch.ones((b, 1, n), dtype=torch.int32, device=dev)
    bonds = torch.zeros((b, 1, n), dtype=torch.int32, device=dev)
    ws = _lib.workspace(lib.qmx_layer_workspace(n, b), dev)
    st = lib.qmx_generate_layer(h, n, b, 1, 0, cutoff, _lib.ptr(psi_work), _lib.ptr(gates),
                                _lib.ptr(widths), _lib.ptr(bonds), _lib.ptr(ws), ws.numel(),
                                _lib.stream_ptr())
    _lib.check(st, h)
    return gates[:, 0], widths[:, 0], bonds[:, 0]


def apply_layer(psi: torch.Tensor, gates: torch.Tensor, widths: torch.Tensor, inverse: bool):
    """(d): apply one layer (gates (B, N, 4, 4), widths (B, N)) in place."""
    b, length = psi.shape
    n = num_qubits_of(length)
    if psi.dtype != torch.complex128:
        raise ValueError("apply_layer works on complex128 states")
    lib, h = _lib.load(), _lib.handle()
    g = gates.reshape(b, 1, n, 4, 4).to(torch.complex128).contiguous()
    w = widths.reshape(b, 1, n).to(torch.int32).contiguous()
    st = lib.qmx_apply_layer(h, n, b, 1, 0, 1 if inverse else 0, _lib.ptr(psi), _lib.ptr(g), _lib.ptr(w),
                             _lib.stream_ptr())
    _lib.check(st, h)
    return psi


@dataclass
class SweepResult:
    fidelity: torch.Tensor       # (B, S+1) complex128: <psi|C|0> before each sweep and at the end
    overlap: torch.Tensor        # (B, S) float64: <T|C'|0> after each sweep
    sweeps_done: torch.Tensor    # (B,) int32
    circuit: torch.Tensor | None


def sweep(target: torch.Tensor, gates: torch.Tensor, widths: torch.Tensor, layers_used: torch.Tensor,
          num_sweeps: int, psi_ref: torch.Tensor | None = None, goal: torch.Tensor | None = None,
          circuit_out: bool = False, snapshots: tuple[int, torch.Tensor] | None = None) -> SweepResult:
    """(e): environment sweeps, gates updated in place.  ``snapshots`` =
    (first, buffer of shape (2,) + gates.shape): the gates after ``first``
    and ``first + 1`` completed sweeps of this call are copied into buffer[0]
    and buffer[1] (qmx_sweep_snapshots)."""
    b, length = target.shape
    n = num_qubits_of(length)
    num_layers = gates.shape[1]
    lib, h = _lib.load(), _lib.handle()
    dev = target.device
    dt = _DTYPES[target.dtype]
    fid = torch.zeros((b, num_sweeps + 1), dtype=torch.complex128, device=dev) if psi_ref is not None else None
    ov = torch.zeros((b, max(num_sweeps, 1)), dtype=torch.float64, device=dev)
    done = torch.zeros(b, dtype=torch.int32, device=dev)
    circ = torch.empty_like(target) if circuit_out else None
    ws = _lib.workspace(lib.qmx_sweep_workspace(n, b, dt), dev)
    if snapshots is None:
        st = lib.qmx_sweep(h, n, b, num_layers, _lib.ptr(layers_used), num_sweeps, _lib.ptr(target), dt,
                           _lib.ptr(gates), _lib.ptr(widths), _lib.ptr(psi_ref), _lib.ptr(fid),
                           _lib.ptr(ov if num_sweeps > 0 else None), _lib.ptr(goal), _lib.ptr(done),
                           _lib.ptr(circ), _lib.ptr(ws), ws.numel(), _lib.stream_ptr())
    else:
        first, buf = snapshots
        if tuple(buf.shape) != (2,) + tuple(gates.shape) or buf.dtype != gates.dtype or not buf.is_contiguous():
            raise ValueError("snapshots: a contiguous buffer of shape (2,) + gates.shape")
        st = lib.qmx_sweep_snapshots(h, n, b, num_layers, _lib.ptr(layers_used), num_sweeps, _lib.ptr(target),
                                     dt, _lib.ptr(gates), _lib.ptr(widths), _lib.ptr(psi_ref), _lib.ptr(fid),
                                     _lib.ptr(ov if num_sweeps > 0 else None), _lib.ptr(goal), _lib.ptr(done),
                                     _lib.ptr(circ), int(first), _lib.ptr(buf), _lib.ptr(ws), ws.numel(),
                                     _lib.stream_ptr())
    _lib.check(st, h)
    return SweepResult(fid, ov[:, :num_sweeps], done, circ)


def sweep_noalloc(target: torch.Tensor, gates: torch.Tensor, widths: torch.Tensor,
                  layers_used: torch.Tensor, num_sweeps: int, fidelity: torch.Tensor | None = None,
                  overlap: torch.Tensor | None = None, psi_ref: torch.Tensor | None = None,
                  goal: torch.Tensor | None = None, sweeps_done: torch.Tensor | None = None,
                  circuit_out: torch.Tensor | None = None) -> None:
    """``sweep`` through qmx_sweep_noalloc: the library allocates nothing and
    takes its scratch from its own per-device array (problems up to
    qmx_sweep_workspace() <= 64 MiB); the caller passes every output buffer
    (fidelity (B, S+1) complex128 with psi_ref, overlap (B, S) float64,
    sweeps_done (B,) int32 with goal, circuit_out like target).  Overlapping
    calls on one GPU race on that scratch."""
    b, length = target.shape
    n = num_qubits_of(length)
    num_layers = gates.shape[1]
    lib, h = _lib.load(), _lib.handle()
    st = lib.qmx_sweep_noalloc(h, n, b, num_layers, _lib.ptr(layers_used), num_sweeps, _lib.ptr(target),
                               _DTYPES[target.dtype], _lib.ptr(gates), _lib.ptr(widths), _lib.ptr(psi_ref),
                               _lib.ptr(fidelity), _lib.ptr(overlap), _lib.ptr(goal), _lib.ptr(sweeps_done),
                               _lib.ptr(circuit_out), _lib.stream_ptr())
    _lib.check(st, h)


def circuit_state(gates: torch.Tensor, widths: torch.Tensor, layers_used: torch.Tensor,
                  dtype=torch.complex128, psi_ref: torch.Tensor | None = None):
    """C|0..0> of each state's gate list on the GPU (qmx_circuit_state);
    with psi_ref also returns <psi_ref|C|0> per state."""
    b, nl, n = widths.shape
    lib, h = _lib.load(), _lib.handle()
    dev = gates.device
    dt = _DTYPES[dtype]
    out = torch.empty((b, 1 << n), dtype=dtype, device=dev)
    fid = torch.zeros((b, 1), dtype=torch.complex128, device=dev) if psi_ref is not None else None
    ws = _lib.workspace(lib.qmx_sweep_workspace(n, b, dt), dev)
    st = lib.qmx_circuit_state(h, n, b, nl, _lib.ptr(layers_used), _lib.ptr(gates), _lib.ptr(widths), dt,
                               _lib.ptr(psi_ref), _lib.ptr(fid), _lib.ptr(out), _lib.ptr(ws), ws.numel(),
                               _lib.stream_ptr())
    _lib.check(st, h)
    return (out, fid[:, 0]) if fid is not None else out


@dataclass
class EncodeResult:
    gates: torch.Tensor          # (B, L, N, 4, 4) complex128, layer = extraction order
    widths: torch.Tensor         # (B, L, N) int32
    layers_used: torch.Tensor    # (B,) int32
    fidelity: torch.Tensor       # (B, S+1) complex128
    sweeps_done: torch.Tensor    # (B,) int32


def encode(statevectors, bond_dimension: int, num_layers: int = 1, num_sweeps: int = 0,
           fidelity_threshold: float = DEFAULT_THRESHOLD, cutoff: float = CUTOFF,
           sweep_dtype=torch.complex64, goal: torch.Tensor | None = None,
           workspace: torch.Tensor | None = None) -> EncodeResult:
    """Whole prepare_state for a batch of equal-size states in one C-ABI call."""
    sv = as_device_states(statevectors)
    b, length = sv.shape
    n = num_qubits_of(length)
    lib, h = _lib.load(), _lib.handle()
    cfg = _lib.EncodeConfig(n, b, int(bond_dimension), int(num_layers), int(num_sweeps),
                            _DTYPES[sv.dtype], _DTYPES[sweep_dtype], float(fidelity_threshold),
                            float(cutoff))
    need = lib.qmx_encode_workspace(ctypes.byref(cfg))
    if workspace is None or workspace.numel() < need:
        workspace = _lib.workspace(need, sv.device)
    gates = torch.zeros((b, num_layers, n, 4, 4), dtype=torch.complex128, device=sv.device)
    widths = torch.ones((b, num_layers, n), dtype=torch.int32, device=sv.device)
    used = torch.zeros(b, dtype=torch.int32, device=sv.device)
    fid = torch.zeros((b, num_sweeps + 1), dtype=torch.complex128, device=sv.device)
    done = torch.zeros(b, dtype=torch.int32, device=sv.device)
    st = lib.qmx_encode(h, ctypes.byref(cfg), _lib.ptr(sv), _lib.ptr(gates), _lib.ptr(widths),
                        _lib.ptr(used), _lib.ptr(fid), _lib.ptr(goal), _lib.ptr(done),
                        _lib.ptr(workspace), workspace.numel(), _lib.stream_ptr())
    _lib.check(st, h)
    return EncodeResult(gates, widths, used, fid, done)


def bond_groups(bond_dimensions) -> tuple[list[int], list[int], list[int]]:
    """Runs of equal bond dimension: (first, count, chi) per run, in order."""
    first, count, chi = [], [], []
    for i, d in enumerate(int(x) for x in bond_dimensions):
        if chi and chi[-1] == d:
            count[-1] += 1
        else:
            first.append(i)
            count.append(1)
            chi.append(d)
    return first, count, chi


def encode_grouped(statevectors, bond_dimensions, num_layers: int = 1, num_sweeps: int = 0,
                   fidelity_threshold: float = DEFAULT_THRESHOLD, cutoff: float = CUTOFF,
                   sweep_dtype=torch.complex64) -> EncodeResult:
    """``encode`` of equal-size states with one bond dimension per state
    (qmx_encode_grouped): the chi-truncation runs per run of equal bond
    dimension, everything after it once over the batch."""
    sv = as_device_states(statevectors)
    b, length = sv.shape
    n = num_qubits_of(length)
    if len(bond_dimensions) != b:
        raise ValueError("one bond dimension per state")
    first, count, chi = bond_groups(bond_dimensions)
    arr = lambda v: (ctypes.c_int * len(v))(*v)
    f_a, c_a, x_a = arr(first), arr(count), arr(chi)
    lib, h = _lib.load(), _lib.handle()
    cfg = _lib.EncodeConfig(n, b, max(chi), int(num_layers), int(num_sweeps), _DTYPES[sv.dtype],
                            _DTYPES[sweep_dtype], float(fidelity_threshold), float(cutoff))
    need = lib.qmx_encode_grouped_workspace(ctypes.byref(cfg), f_a, c_a, x_a, len(chi))
    if need == 0:
        raise ValueError("invalid bond dimensions")
    workspace = _lib.workspace(need, sv.device)
    gates = torch.zeros((b, num_layers, n, 4, 4), dtype=torch.complex128, device=sv.device)
    widths = torch.ones((b, num_layers, n), dtype=torch.int32, device=sv.device)
    used = torch.zeros(b, dtype=torch.int32, device=sv.device)
    fid = torch.zeros((b, num_sweeps + 1), dtype=torch.complex128, device=sv.device)
    st = lib.qmx_encode_grouped(h, ctypes.byref(cfg), f_a, c_a, x_a, len(chi), _lib.ptr(sv), _lib.ptr(gates),
                                _lib.ptr(widths), _lib.ptr(used), _lib.ptr(fid), _lib.ptr(workspace),
                                workspace.numel(), _lib.stream_ptr())
    _lib.check(st, h)
    return EncodeResult(gates, widths, used, fid, torch.full((b,), int(num_sweeps), dtype=torch.int32,
                                                             device=sv.device))


@dataclass
class Prepared:
    psi: torch.Tensor          # (B, 2^N) complex128 normalised input
    target: torch.Tensor       # (B, 2^N) sweep dtype: the chi-MPS state
    gates: torch.Tensor        # (B, L, N, 4, 4) complex128 Ran gates
    widths: torch.Tensor       # (B, L, N) int32
    layers_used: torch.Tensor  # (B,) int32


def encode_prepare(statevectors, bond_dimension: int, num_layers: int = 1,
                   fidelity_threshold: float = DEFAULT_THRESHOLD, cutoff: float = CUTOFF,
                   sweep_dtype=torch.complex64, workspace: torch.Tensor | None = None) -> Prepared:
    """prepare_state up to the sweeps (qmx_encode_prepare); ``sweep`` on the
    result with psi_ref=psi reproduces ``encode`` exactly."""
    sv = as_device_states(statevectors)
    b, length = sv.shape
    n = num_qubits_of(length)
    lib, h = _lib.load(), _lib.handle()
    cfg = _lib.EncodeConfig(n, b, int(bond_dimension), int(num_layers), 0, _DTYPES[sv.dtype],
                            _DTYPES[sweep_dtype], float(fidelity_threshold), float(cutoff))
    need = lib.qmx_encode_workspace(ctypes.byref(cfg))
    if workspace is None or workspace.numel() < need:
        workspace = _lib.workspace(need, sv.device)
    gates = torch.zeros((b, num_layers, n, 4, 4), dtype=torch.complex128, device=sv.device)
    widths = torch.ones((b, num_layers, n), dtype=torch.int32, device=sv.device)
    used = torch.zeros(b, dtype=torch.int32, device=sv.device)
    psi = torch.empty((b, length), dtype=torch.complex128, device=sv.device)
    target = torch.empty((b, length), dtype=sweep_dtype, device=sv.device)
    st = lib.qmx_encode_prepare(h, ctypes.byref(cfg), _lib.ptr(sv), _lib.ptr(gates), _lib.ptr(widths),
                                _lib.ptr(used), _lib.ptr(psi), _lib.ptr(target), _lib.ptr(workspace),
                                workspace.numel(), _lib.stream_ptr())
    _lib.check(st, h)
    return Prepared(psi, target, gates, widths, used)


SCHEDULES = ("DallOall", "IterDiOi", "IterDiOall")


def schedule_layers(psi: torch.Tensor, target: torch.Tensor, num_layers: int, num_sweeps: int,
                    schedule: str, fidelity_threshold: float = DEFAULT_THRESHOLD,
                    cutoff: float = CUTOFF, sweep_dtype=torch.complex64):
    """Iterative schedules named by the reference's notebook (markdown cell 22,
    after Rudolph et al., arXiv:2209.00595), from the layer and sweep kernels:

    IterDiOi   -- layer i from the current residual (qmx_generate_layer), swept
                  alone against that residual (qmx_sweep), then peeled off
                  (qmx_apply_layer, inverse);
    IterDiOall -- layer i from the residual, then every layer so far swept
                  against the chi-MPS target and the residual recomputed as
                  C^dag |target> / |target|.
    The early exit of sequential.py:384-391 is checked on each state's
    residual after every layer; finished states keep their layers.
    ``psi`` (normalised input) and ``target`` (chi-MPS state) are (B, 2^N)
    complex128.  Returns (gates, widths, layers_used, fidelity |<psi|C|0>|)."""
    if schedule not in ("IterDiOi", "IterDiOall"):
        raise ValueError(f"unknown schedule {schedule!r}; expected one of {SCHEDULES}")
    b, length = target.shape
    n = num_qubits_of(length)
    dev = target.device
    unit = (target / torch.linalg.vector_norm(target, dim=1, keepdim=True)).contiguous()
    resid = unit.clone()
    gates = torch.zeros((b, num_layers, n, 4, 4), dtype=torch.complex128, device=dev)
    widths = torch.ones((b, num_layers, n), dtype=torch.int32, device=dev)
    used = torch.zeros(b, dtype=torch.int32, device=dev)
    active = torch.ones(b, dtype=torch.bool, device=dev)
    sweep_target = target.to(sweep_dtype).contiguous()
    for i in range(num_layers):
        g, w, _ = generate_layer(resid.clone(), cutoff)
        if schedule == "IterDiOi":
            lg, lw = g.unsqueeze(1).contiguous(), w.unsqueeze(1).contiguous()
            if num_sweeps > 0:
                one = active.to(torch.int32)
                sweep(resid.to(sweep_dtype).contiguous(), lg, lw, one, num_sweeps)
            g, w = lg[:, 0], lw[:, 0]
        gates[:, i] = torch.where(active[:, None, None, None], g, gates[:, i])
        widths[:, i] = torch.where(active[:, None], w, widths[:, i])
        used = torch.where(active, torch.full_like(used, i + 1), used)
        if schedule == "IterDiOi":
            new = apply_layer(resid.clone(), g, w, inverse=True)
        else:
            if num_sweeps > 0:
                sweep(sweep_target, gates, widths, torch.where(active, used, torch.zeros_like(used)),
                      num_sweeps)
            new = unit.clone()
            for l in range(i + 1):
                apply_layer(new, gates[:, l], widths[:, l], inverse=True)
        resid = torch.where(active[:, None], new, resid).contiguous()
        amp0 = resid[:, 0].conj()
        active &= (amp0 - 1).abs() > (1 - fidelity_threshold) + 1e-5
        if not bool(active.any()):
            break
    _, fid = circuit_state(gates, widths, used, psi_ref=psi.contiguous())
    return gates, widths, used, fid.abs()


def encode_schedule(statevectors, bond_dimension: int, num_layers: int = 1, num_sweeps: int = 0,
                    schedule: str = "IterDiOall", fidelity_threshold: float = DEFAULT_THRESHOLD,
                    cutoff: float = CUTOFF, sweep_dtype=torch.complex64):
    """prepare_state of a batch under ``schedule`` (SCHEDULES); "DallOall" is
    ``encode``.  Returns (gates, widths, layers_used, fidelity)."""
    if schedule == "DallOall":
        r = encode(statevectors, bond_dimension, num_layers, num_sweeps, fidelity_threshold, cutoff,
                   sweep_dtype)
        return r.gates, r.widths, r.layers_used, r.fidelity[:, -1].abs()
    sv = as_device_states(statevectors)
    psi, target, _ = chi_truncate(sv, bond_dimension, cutoff)
    return schedule_layers(psi, target, num_layers, num_sweeps, schedule, fidelity_threshold, cutoff,
                           sweep_dtype)


@dataclass
class RaggedResult:
    gates: torch.Tensor        # (L, N, 4, 4) complex128
    widths: torch.Tensor       # (L, N) int32
    layers_used: int
    fidelity: torch.Tensor     # (S + 1,) complex128


def encode_ragged(statevectors, bond_dimensions, num_layers: int = 1, num_sweeps: int = 0,
                  fidelity_threshold: float = DEFAULT_THRESHOLD, cutoff: float = CUTOFF,
                  sweep_dtype=torch.complex64, max_concurrency: int = 6) -> list[RaggedResult]:
    """prepare_state for states of different sizes in one call (qmx_encode_ragged).

    ``statevectors``: sequence of 1-D arrays / tensors (any power-of-two
    lengths); ``bond_dimensions``: one int per state, or one int for all."""
    dev = _lib.require_gpu()
    svs = [as_device_states(v) for v in statevectors]
    dtype = torch.complex128 if any(v.dtype == torch.complex128 for v in svs) else torch.complex64
    svs = [v.to(dtype).contiguous() for v in svs]
    if isinstance(bond_dimensions, int):
        bond_dimensions = [bond_dimensions] * len(svs)
    if len(bond_dimensions) != len(svs):
        raise ValueError("one bond dimension per state")
    outs = []
    arr = (_lib.RaggedState * len(svs))()
    for i, (v, chi) in enumerate(zip(svs, bond_dimensions)):
        n = num_qubits_of(v.shape[1])
        g = torch.zeros((num_layers, n, 4, 4), dtype=torch.complex128, device=dev)
        w = torch.ones((num_layers, n), dtype=torch.int32, device=dev)
        u = torch.zeros(1, dtype=torch.int32, device=dev)
        f = torch.zeros(num_sweeps + 1, dtype=torch.complex128, device=dev)
        arr[i] = _lib.RaggedState(n, int(chi), v.data_ptr(), g.data_ptr(), w.data_ptr(), u.data_ptr(),
                                  f.data_ptr())
        outs.append((g, w, u, f))
    lib, h = _lib.load(), _lib.handle()
    sv_dt, sw_dt = _DTYPES[dtype], _DTYPES[sweep_dtype]
    need = lib.qmx_encode_ragged_workspace(len(svs), arr, num_layers, num_sweeps, sv_dt, sw_dt)
    if need == 0:
        raise ValueError("invalid ragged batch")
    ws = _lib.workspace(need, dev)
    st = lib.qmx_encode_ragged(h, len(svs), arr, num_layers, num_sweeps, sv_dt, sw_dt,
                               float(fidelity_threshold), float(cutoff), int(max_concurrency),
                               _lib.ptr(ws), ws.numel(), _lib.stream_ptr())
    _lib.check(st, h)
    return [RaggedResult(g, w, int(u.item()), f) for g, w, u, f in outs]


def gates_to_layers(gates: np.ndarray, widths: np.ndarray, layers_used: int):
    """Device gate arrays of one state -> list[UnitaryLayer] in circuit order.

    Blocks follow the widths: a block ends at every one-qubit gate
    (mps.py:816-834); the list is reversed as sequential.py:396 does.
    """
    from qmprs_amd.primitives.mps import Tensor
    n = gates.shape[1]
    layers = []
    for l in range(int(layers_used)):
        blocks = []
        start = 0
        for i in range(n):
            if int(widths[l, i]) == 1:
                mats = []
                for j in range(start, i + 1):
                    d = 4 if j < i else 2
                    mats.append(Tensor(np.array(gates[l, j, :d, :d], dtype=np.complex128)))
                blocks.append((start, i, mats))
                start = i + 1
        layers.append(blocks)
    layers.reverse()
    return layers


def layers_to_arrays(layers, num_sites: int):
    """list[UnitaryLayer] (circuit order) -> (gates (L,N,4,4), widths (L,N)) in extraction order."""
    nl = len(layers)
    gates = np.zeros((nl, num_sites, 4, 4), dtype=np.complex128)
    widths = np.ones((nl, num_sites), dtype=np.int32)
    for l, layer in enumerate(reversed(layers)):
        for start, end, mats in layer:
            for j in range(start, end + 1):
                m = np.asarray(getattr(mats[j - start], "data", mats[j - start]), dtype=np.complex128)
                d = m.shape[0]
                gates[l, j, :d, :d] = m
                widths[l, j] = 2 if d == 4 else 1
    return gates, widths
