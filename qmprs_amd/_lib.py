"""ctypes binding of libqmx.so (the C-ABI declared in include/qmx.h).

Device buffers are torch tensors on a ROCm GPU; their ``data_ptr()`` and the
current stream handle are passed straight through.  There is no CPU path: if
the library or the GPU is missing every entry point raises.
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().with_name("libqmx.so")

QMX_OK = 0
QMX_ERR_INVALID = 1
QMX_ERR_NOT_UNITARY = 2
QMX_ERR_NO_CONVERGENCE = 3
QMX_ERR_HIP = 4
QMX_ERR_WORKSPACE = 5
QMX_ERR_BASIS = 6

QMX_C64 = 0
QMX_C128 = 1

NOALLOC_SCRATCH_BYTES = 64 << 20   # QMX_NOALLOC_SCRATCH_BYTES

QMX_INDEX_ROW = 0
QMX_INDEX_SNAKE = 1

PHASES = ("chi", "layers", "rebuild", "sweep", "window_pass")

QMX_OPT_SWEEP_MODE = 1
SWEEP_MODES = {"auto": 0, "step": 1, "window": 2}

_VP = ctypes.c_void_p
_INT = ctypes.c_int
_DBL = ctypes.c_double
_FLT = ctypes.c_float
_SZ = ctypes.c_size_t


class EncodeConfig(ctypes.Structure):
    """Mirror of ``qmx_encode_config``."""
    _fields_ = [
        ("num_qubits", _INT),
        ("batch", _INT),
        ("bond_dimension", _INT),
        ("num_layers", _INT),
        ("num_sweeps", _INT),
        ("sv_dtype", _INT),
        ("sweep_dtype", _INT),
        ("fidelity_threshold", _DBL),
        ("cutoff", _DBL),
    ]


class RaggedState(ctypes.Structure):
    """Mirror of ``qmx_ragged_state``."""
    _fields_ = [
        ("num_qubits", _INT),
        ("bond_dimension", _INT),
        ("statevector", _VP),
        ("gates", _VP),
        ("widths", _VP),
        ("layers_used", _VP),
        ("fidelity", _VP),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "qmx_version": (_INT, []),
    "qmx_create": (_INT, [_INT, ctypes.POINTER(_VP)]),
    "qmx_destroy": (_INT, [_VP]),
    "qmx_last_error": (ctypes.c_char_p, [_VP]),
    "qmx_set_option": (_INT, [_VP, _INT, _INT]),
    "qmx_profile_enable": (_INT, [_VP, _INT]),
    "qmx_profile_read": (_INT, [_VP, ctypes.POINTER(_DBL), ctypes.POINTER(ctypes.c_longlong)]),
    "qmx_chi_truncate_workspace": (_SZ, [_INT, _INT, _INT]),
    "qmx_chi_truncate": (_INT, [_VP, _INT, _INT, _INT, _DBL, _VP, _INT, _VP, _VP, _VP, _VP, _SZ, _VP]),
    "qmx_layer_workspace": (_SZ, [_INT, _INT]),
    "qmx_generate_layer": (_INT, [_VP, _INT, _INT, _INT, _INT, _DBL, _VP, _VP, _VP, _VP, _VP, _SZ, _VP]),
    "qmx_apply_layer": (_INT, [_VP, _INT, _INT, _INT, _INT, _INT, _VP, _VP, _VP, _VP]),
    "qmx_extract_layers_workspace": (_SZ, [_INT, _INT]),
    "qmx_extract_layers": (_INT, [_VP, _INT, _INT, _INT, _DBL, _DBL, _VP, _VP, _VP, _VP, _VP, _VP,
                                  _VP, _SZ, _VP]),
    "qmx_sweep_workspace": (_SZ, [_INT, _INT, _INT]),
    "qmx_sweep": (_INT, [_VP, _INT, _INT, _INT, _VP, _INT, _VP, _INT, _VP, _VP, _VP, _VP, _VP, _VP,
                         _VP, _VP, _VP, _SZ, _VP]),
    "qmx_sweep_snapshots": (_INT, [_VP, _INT, _INT, _INT, _VP, _INT, _VP, _INT, _VP, _VP, _VP, _VP, _VP, _VP,
                                   _VP, _VP, _INT, _VP, _VP, _SZ, _VP]),
    "qmx_sweep_noalloc": (_INT, [_VP, _INT, _INT, _INT, _VP, _INT, _VP, _INT, _VP, _VP, _VP, _VP, _VP, _VP,
                                 _VP, _VP, _VP]),
    "qmx_prepare_input_workspace": (_SZ, [_INT, _INT]),
    "qmx_prepare_input": (_INT, [_VP, _INT, _INT, _VP, _INT, _INT, _DBL, _VP, _VP, _SZ, _VP]),
    "qmx_mps_to_state_workspace": (_SZ, [_INT, ctypes.POINTER(_INT)]),
    "qmx_mps_to_state": (_INT, [_VP, _INT, ctypes.POINTER(_INT), _VP, _VP, _VP, _SZ, _VP]),
    "qmx_circuit_state": (_INT, [_VP, _INT, _INT, _INT, _VP, _VP, _VP, _INT, _VP, _VP, _VP, _VP, _SZ, _VP]),
    "qmx_forward_plan": (_INT, [_INT, _INT, _VP, _SZ, _VP, _SZ, ctypes.POINTER(_INT), ctypes.POINTER(_INT)]),
    "qmx_encode_workspace": (_SZ, [ctypes.POINTER(EncodeConfig)]),
    "qmx_encode": (_INT, [_VP, ctypes.POINTER(EncodeConfig), _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP,
                          _SZ, _VP]),
    "qmx_encode_prepare": (_INT, [_VP, ctypes.POINTER(EncodeConfig), _VP, _VP, _VP, _VP, _VP, _VP, _VP,
                                  _SZ, _VP]),
    "qmx_encode_grouped_workspace": (_SZ, [ctypes.POINTER(EncodeConfig), ctypes.POINTER(_INT),
                                           ctypes.POINTER(_INT), ctypes.POINTER(_INT), _INT]),
    "qmx_encode_grouped": (_INT, [_VP, ctypes.POINTER(EncodeConfig), ctypes.POINTER(_INT), ctypes.POINTER(_INT),
                                  ctypes.POINTER(_INT), _INT, _VP, _VP, _VP, _VP, _VP, _VP, _SZ, _VP]),
    "qmx_encode_ragged_workspace": (_SZ, [_INT, ctypes.POINTER(RaggedState), _INT, _INT, _INT, _INT]),
    "qmx_encode_ragged": (_INT, [_VP, _INT, ctypes.POINTER(RaggedState), _INT, _INT, _INT, _INT, _DBL, _DBL,
                                 _INT, _VP, _SZ, _VP]),
    "qmx_mlp_param_count": (_SZ, [_INT, _INT, _INT]),
    "qmx_mlp_workspace": (_SZ, [_INT, _INT, _INT, _INT]),
    "qmx_mlp_forward": (_INT, [_VP, _INT, _INT, _INT, _INT, _VP, _VP, _VP, _VP, _VP, _SZ, _VP]),
    "qmx_mlp_backward": (_INT, [_VP, _INT, _INT, _INT, _INT, _VP, _VP, _VP, _VP, _VP, _VP, _SZ, _VP]),
    "qmx_mlp_adam": (_INT, [_VP, _SZ, _VP, _VP, _VP, _VP, _VP, _FLT, _FLT, _FLT, _FLT, _INT, _VP]),
    "qmx_mlp_to_bf16": (_INT, [_VP, _SZ, _VP, _VP, _VP]),
    "qmx_mse_workspace": (_SZ, []),
    "qmx_mse_loss": (_INT, [_VP, _INT, _INT, _INT, _VP, _VP, _VP, _FLT, _VP, _VP, _VP, _SZ, _VP]),
    "qmx_gate_features": (_INT, [_VP, _INT, _INT, _INT, _VP, _VP, _VP, _VP, _VP, _INT, _VP, _VP]),
    "qmx_gate_log": (_INT, [_VP, _INT, _INT, _INT, _VP, _VP, _VP, _VP, _VP, _INT, _VP]),
    "qmx_traj_features": (_INT, [_VP, _INT, _INT, _INT, _VP, _VP, _VP, _VP, _VP, _FLT, _VP, _INT, _VP, _VP, _VP]),
    "qmx_traj_targets": (_INT, [_VP, _INT, _INT, _INT, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _INT, _VP]),
    "qmx_traj_jump": (_INT, [_VP, _INT, _INT, _INT, _VP, _VP, _VP, _VP, _INT, _VP, _VP, _VP, _VP]),
    "qmx_gate_expmap": (_INT, [_VP, _INT, _INT, _INT, _VP, _VP, _INT, _VP, _VP, _VP, _VP]),
}

_lib = None
_lock = threading.Lock()
_local = threading.local()


def load() -> ctypes.CDLL:
    """Load libqmx.so and declare every exported signature."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(
                    f"{LIB_PATH} is not built; run `python -m qmprs_amd.build` (hipcc, gfx950)")
            lib = ctypes.CDLL(str(LIB_PATH))
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("qmprs_amd runs on an AMD GPU (ROCm); no GPU is visible")
    return torch.device("cuda", torch.cuda.current_device())


def handle() -> ctypes.c_void_p:
    """Per-thread, per-device qmx handle."""
    dev = require_gpu()
    lib = load()
    cache = getattr(_local, "handles", None)
    if cache is None:
        cache = _local.handles = {}
    h = cache.get(dev.index)
    if h is None:
        h = _VP()
        st = lib.qmx_create(dev.index, ctypes.byref(h))
        if st != QMX_OK:
            raise RuntimeError(f"qmx_create failed on device {dev.index} (status {st})")
        cache[dev.index] = h
    return h


def stream_ptr() -> ctypes.c_void_p:
    return _VP(torch.cuda.current_stream().cuda_stream)


def ptr(t) -> ctypes.c_void_p:
    return _VP(0 if t is None else t.data_ptr())


def check(status: int, h) -> None:
    """Map a QMX_* status to the reference's exception types."""
    if status == QMX_OK:
        return
    msg = load().qmx_last_error(h)
    msg = msg.decode() if msg else f"qmx status {status}"
    if status in (QMX_ERR_INVALID, QMX_ERR_NOT_UNITARY, QMX_ERR_WORKSPACE):
        raise ValueError(msg)
    raise RuntimeError(msg)


def workspace(nbytes: int, device) -> "torch.Tensor":
    if nbytes == 0:
        raise ValueError("invalid problem size for the qmx workspace")
    return torch.empty(int(nbytes), dtype=torch.uint8, device=device)


def profile_enable(enable: bool = True) -> None:
    h = handle()
    check(load().qmx_profile_enable(h, 1 if enable else 0), h)


def profile_read() -> dict:
    """{phase: (milliseconds, launches)} since the last read (synchronises)."""
    h = handle()
    ms = (ctypes.c_double * len(PHASES))()
    ln = (ctypes.c_longlong * len(PHASES))()
    check(load().qmx_profile_read(h, ms, ln), h)
    return {p: (ms[i], ln[i]) for i, p in enumerate(PHASES)}


QMX_OPT_SPLIT_SWEEPS = 2
QMX_OPT_WINDOW_BLOCKS_PER_CU = 3
QMX_OPT_FWD_BLOCKS_PER_CU = 4
QMX_OPT_STAGED_WINDOWS = 5
QMX_OPT_SWEEP_RELAXATION = 6
QMX_OPT_JACOBI_SWEEP_CAP = 7
QMX_OPT_FAIL_UNITARITY = 8
QMX_OPT_FWD_MFMA = 9
FWD_MFMA_DEFAULT = 1  # the library's default for QMX_OPT_FWD_MFMA


def set_option(option: int, value: int) -> None:
    h = handle()
    check(load().qmx_set_option(h, option, int(value)), h)


def forward_plan(num_qubits: int, num_layers: int):
    """Schedule of the forward circuit passes (host only): (records, masks)
    with records an (K, 3) int array of (pass, stage, forward position) in
    application order and masks the 4-bit state mask of each stage."""
    import numpy as np

    lib = load()
    nr, ns = _INT(0), _INT(0)
    lib.qmx_forward_plan(num_qubits, num_layers, None, 0, None, 0, ctypes.byref(nr), ctypes.byref(ns))
    rec = np.zeros((max(nr.value, 1), 3), dtype=np.int32)
    masks = np.zeros(max(ns.value, 1), dtype=np.uint64)
    st = lib.qmx_forward_plan(num_qubits, num_layers, rec.ctypes.data, rec.size, masks.ctypes.data,
                              masks.size, ctypes.byref(nr), ctypes.byref(ns))
    if st != 0:
        raise ValueError(f"qmx_forward_plan failed with status {st}")
    return rec[:nr.value], masks[:ns.value]


def set_split_sweeps(enable: bool) -> None:
    """Run batches of >= 2 states as two interleaved half-batches (default on)."""
    h = handle()
    check(load().qmx_set_option(h, QMX_OPT_SPLIT_SWEEPS, 1 if enable else 0), h)


def set_sweep_relaxation(omega: float) -> None:
    """Gate update of the sweeps: omega = 1 (default) is the reference's
    G' = conj(polar(E)); 1 < omega < 1.95 the over-relaxed
    polar(G' + (omega - 1)(G' - G)), which reaches a given fidelity in fewer
    sweeps but leaves the reference's gate trajectory."""
    h = handle()
    check(load().qmx_set_option(h, QMX_OPT_SWEEP_RELAXATION, int(round(1000 * omega))), h)


def set_sweep_mode(mode: str) -> None:
    """'auto' (windowed passes for N >= 10), 'step' (one pass per gate) or
    'window' (windowed passes; below N = 10 it also runs one pass per gate)."""
    h = handle()
    check(load().qmx_set_option(h, QMX_OPT_SWEEP_MODE, SWEEP_MODES[mode]), h)
