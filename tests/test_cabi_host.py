"""C-ABI and host logic on the CPU (no GPU calls).

* libqmx.so loads and exports every function include/qmx.h declares, and the
  ctypes binding declares exactly those;
* workspace queries (pure host arithmetic) accept valid sizes and reject
  invalid ones;
* the Python API validates its arguments like the reference
  (sequential.py:120-153, 594-595; mps.py:151-216; base.py:58-106) before any
  device work, and the gate-array <-> UnitaryLayer conversion round-trips.
"""

from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qmx.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w[\w\s\*]*?\b(qmx_\w+)\s*\(", text, flags=re.M)))


def _lib():
    from qmprs_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("libqmx.so not built")
    return _lib, _lib.load()


def test_header_declares_the_entry_points():
    names = _declared()
    for n in ("qmx_create", "qmx_destroy", "qmx_last_error", "qmx_chi_truncate",
              "qmx_generate_layer", "qmx_apply_layer", "qmx_extract_layers", "qmx_sweep",
              "qmx_encode", "qmx_encode_workspace"):
        assert n in names


def test_library_exports_every_declared_symbol():
    _, lib = _lib()
    for name in _declared():
        assert hasattr(lib, name), name


def test_binding_matches_header():
    mod, _ = _lib()
    assert sorted(mod.SIGNATURES) == _declared()


def test_version_and_workspace_queries():
    mod, lib = _lib()
    assert lib.qmx_version() >= 1
    assert lib.qmx_chi_truncate_workspace(20, 4, 64) > 0
    assert lib.qmx_chi_truncate_workspace(0, 4, 64) == 0
    assert lib.qmx_layer_workspace(12, 2) > 0
    assert lib.qmx_extract_layers_workspace(12, 2) > 0
    assert lib.qmx_sweep_workspace(20, 8, mod.QMX_C64) > 0
    assert lib.qmx_sweep_workspace(20, 8, 7) == 0
    cfg = mod.EncodeConfig(20, 8, 64, 10, 30, mod.QMX_C64, mod.QMX_C64, 1 - 1e-6, 1e-10)
    big = lib.qmx_encode_workspace(ctypes.byref(cfg))
    assert big > 0
    cfg.batch = 16
    assert lib.qmx_encode_workspace(ctypes.byref(cfg)) > big
    cfg.num_layers = 0
    assert lib.qmx_encode_workspace(ctypes.byref(cfg)) == 0


def test_grouped_workspace_queries():
    """qmx_encode_grouped_workspace: groups must tile the batch in order."""
    mod, lib = _lib()
    cfg = mod.EncodeConfig(12, 5, 32, 4, 3, mod.QMX_C64, mod.QMX_C64, 1 - 1e-6, 1e-10)
    arr = lambda v: (ctypes.c_int * len(v))(*v)
    ok = lib.qmx_encode_grouped_workspace(ctypes.byref(cfg), arr([0, 2]), arr([2, 3]), arr([8, 32]), 2)
    assert ok > 0
    one = lib.qmx_encode_grouped_workspace(ctypes.byref(cfg), arr([0]), arr([5]), arr([32]), 1)
    assert one > 0
    for first, count, chi in (([0, 3], [2, 3], [8, 32]),     # gap
                              ([0, 2], [2, 2], [8, 32]),     # counts short of the batch
                              ([0, 2], [2, 3], [8, 0]),      # bond dimension 0
                              ([1], [5], [32])):             # not starting at 0
        assert lib.qmx_encode_grouped_workspace(ctypes.byref(cfg), arr(first), arr(count), arr(chi),
                                                len(chi)) == 0
    cfg20 = mod.EncodeConfig(20, 2, 64, 4, 3, mod.QMX_C64, mod.QMX_C64, 1 - 1e-6, 1e-10)
    assert lib.qmx_encode_grouped_workspace(ctypes.byref(cfg20), arr([0, 1]), arr([1, 1]), arr([64, 600]),
                                            2) == 0      # truncating bond dimension above 512


def test_bond_groups_are_runs_in_order():
    from qmprs_amd.engine import bond_groups
    assert bond_groups([4, 4, 16, 4]) == ([0, 2, 3], [2, 1, 1], [4, 16, 4])
    assert bond_groups([8]) == ([0], [1], [8])
    assert bond_groups([]) == ([], [], [])


def test_encode_config_layout_matches_c_struct():
    mod, _ = _lib()
    # 7 ints then 2 doubles: doubles start at offset 32 (8-byte aligned)
    assert mod.EncodeConfig.fidelity_threshold.offset == 32
    assert ctypes.sizeof(mod.EncodeConfig) == 48


def test_null_handle_calls_are_rejected():
    mod, lib = _lib()
    assert lib.qmx_destroy(None) == mod.QMX_OK
    assert lib.qmx_set_option(None, 1, 0) == mod.QMX_ERR_INVALID
    assert lib.qmx_create(0, None) == mod.QMX_ERR_INVALID


# ---------------------------------------------------------- host logic ---

def test_fidelity_threshold_property():
    from qmprs_amd.circuit import GateCircuit
    from qmprs_amd.synthesis.mps_encoding import Sequential
    enc = Sequential(GateCircuit)
    assert enc.fidelity_threshold == 1 - 1e-6
    enc.fidelity_threshold = 0.5
    assert enc.fidelity_threshold == 0.5
    for bad in (-0.1, 1.5, "x"):
        with pytest.raises(ValueError):
            enc.fidelity_threshold = bad


def test_prepare_mps_argument_validation():
    from qmprs_amd.circuit import GateCircuit
    from qmprs_amd.synthesis.mps_encoding import Sequential
    enc = Sequential(GateCircuit)
    for kw in ({"num_layers": 0}, {"num_layers": 1.5}, {"num_layers": 2, "num_sweeps": -1}):
        with pytest.raises(ValueError):
            enc.prepare_mps(object(), **kw)


def test_prepare_state_argument_validation():
    from qmprs_amd.circuit import GateCircuit
    from qmprs_amd.synthesis.mps_encoding import Sequential
    enc = Sequential(GateCircuit)
    sv = np.ones(8) / np.sqrt(8)
    with pytest.raises(ValueError, match="Index type"):
        enc.prepare_state(sv, 4, index_type="column")


def test_mps_argument_validation():
    from qmprs_amd.primitives import MPS
    for bad in (0, -3, 2.5, True):
        with pytest.raises(ValueError):
            MPS(statevector=np.ones(4), bond_dimension=bad)
    with pytest.raises(ValueError):
        MPS()
    with pytest.raises(ValueError):
        MPS(statevector=np.ones(4), mps=[np.ones((1, 2, 1))] * 2)
    with pytest.raises(TypeError):
        MPS(mps=5)
    with pytest.raises(ValueError):
        MPS(mps=[np.ones((1, 2, 1))])


def test_statevector_length_checks():
    from qmprs_amd import engine
    assert engine.num_qubits_of(1024) == 10
    for bad in (0, 1, 3, 1000):
        with pytest.raises(ValueError):
            engine.num_qubits_of(bad)


def test_gate_arrays_round_trip():
    """(L, N, 4, 4) extraction-order arrays <-> circuit-order UnitaryLayers,
    with a block ending at every width-1 gate (mps.py:816-834)."""
    from qmprs_amd import engine
    rng = np.random.default_rng(0)
    n, L = 6, 3
    gates = rng.standard_normal((L, n, 4, 4)) + 1j * rng.standard_normal((L, n, 4, 4))
    widths = np.full((L, n), 2, dtype=np.int32)
    widths[:, n - 1] = 1
    widths[1, 2] = 1
    for l in range(L):
        for s in range(n):
            if widths[l, s] == 1:
                gates[l, s, 2:, :] = 0
                gates[l, s, :, 2:] = 0
    layers = engine.gates_to_layers(gates, widths, L)
    assert [(s, e) for s, e, _ in layers[-2]] == [(0, 2), (3, 5)]   # layer 1, reversed order
    g2, w2 = engine.layers_to_arrays(layers, n)
    np.testing.assert_array_equal(w2, widths)
    np.testing.assert_allclose(g2, gates)


def test_gate_circuit_rejects_bad_input():
    from qmprs_amd.circuit import GateCircuit
    c = GateCircuit(3)
    with pytest.raises(ValueError):
        c.unitary(np.eye(4), [0, 0])
    with pytest.raises(ValueError):
        c.unitary(np.eye(2), [0, 1])
    with pytest.raises(ValueError):
        c.unitary(np.eye(2), 3)
    with pytest.raises(ValueError):
        GateCircuit(0)


def test_wire_format_round_trip(tmp_path):
    """npz gate-list format: save -> load keeps every gate, block and mapping,
    and the emitted circuit reproduces the dense circuit state."""
    from oracle import qmprs_oracle as O
    from qmprs_amd.circuit import GateCircuit
    from qmprs_amd.primitives.mps import Tensor
    from qmprs_amd.synthesis.mps_encoding import wire
    n = 7
    psi = O.random_uniform_state(n, 3)
    target, _ = O.chi_truncate(psi, n, 8)
    raw = O.get_unitary_layers(target, n, 3)
    layers = [[(s, e, [Tensor(g) for g in gs]) for s, e, gs in layer] for layer in raw]
    path = tmp_path / "layers.npz"
    wire.save(path, layers, n)
    back, n2 = wire.load(path)
    assert n2 == n and len(back) == len(layers)
    for la, lb in zip(layers, back):
        assert [(s, e) for s, e, _ in la] == [(s, e) for s, e, _ in lb]
        for (_, _, ma), (_, _, mb) in zip(la, lb):
            for a, b in zip(ma, mb):
                np.testing.assert_array_equal(a.data, b.data)
    arrays = wire.to_arrays(layers, n)
    two = arrays["widths"] == 2
    np.testing.assert_array_equal(arrays["qubits"][two, 1], n - 1 - arrays["sites"][two])
    c = wire.emit(back, n, GateCircuit)
    np.testing.assert_allclose(c.get_statevector(), O.circuit_state(n, O.layer_gate_list(raw)), atol=1e-12)
    bad = dict(arrays)
    bad["version"] = np.int32(99)
    with pytest.raises(ValueError):
        wire.from_arrays(bad)


def test_qiskit_adapter_is_optional():
    from qmprs_amd.synthesis.mps_encoding import wire
    try:
        import qiskit  # noqa: F401
    except ImportError:
        with pytest.raises(ImportError):
            wire.to_qiskit([], 2)


def _worst_bits(n, site):
    m = 1 << (n - 1 - site)
    if site < n - 1:
        m |= 1 << (n - 2 - site)
    return m


@pytest.mark.parametrize("n,layers", [(5, 1), (5, 4), (8, 3), (12, 6), (13, 5), (16, 10), (20, 10),
                                      (24, 3)])
def test_forward_plan_is_a_valid_schedule(n, layers):
    """Every gate exactly once, inside its stage's 4 contiguous bits, stages of
    a pass inside one window of at most 8 bits (n > 12), and for every qubit
    the gates touching it in forward order (the reference applies layers in
    order, sequential.py:173-213; sites of a block in order)."""
    mod, _ = _lib()
    rec, masks = mod.forward_plan(n, layers)
    k = layers * n
    assert sorted(rec[:, 2].tolist()) == list(range(k))
    for m in masks.tolist():
        assert bin(m).count("1") == 4
        low = (m & -m).bit_length() - 1
        assert m == 15 << low
    union = {}
    for p, st, f in rec.tolist():
        assert _worst_bits(n, f % n) & ~int(masks[st]) == 0
        union[p] = union.get(p, 0) | int(masks[st])
    if n > 12:
        for u in union.values():
            assert bin(u).count("1") <= 8
    last = [-1] * n
    for _, _, f in rec.tolist():
        b = _worst_bits(n, f % n)
        for q in range(n):
            if (b >> q) & 1:
                assert f > last[q]
                last[q] = f
    # stage indices and pass indices only grow
    assert np.all(np.diff(rec[:, 0]) >= 0) and np.all(np.diff(rec[:, 1]) >= 0)
    if (n, layers) == (20, 10):
        # the light cone of an 8-bit window: 11 passes instead of one per
        # 7 gates of a layer (29)
        assert rec[-1, 0] + 1 <= 12


def test_forward_plan_applies_the_same_circuit():
    """Applying random gates in the plan's order gives the state of applying
    them in forward order (n = 13: more than one tile window)."""
    mod, _ = _lib()
    rng = np.random.default_rng(5)
    n, layers = 13, 4
    k = layers * n
    gates = []
    for f in range(k):
        site = f % n
        d = 4 if site < n - 1 else 2
        q, _ = np.linalg.qr(rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)))
        gates.append(q)

    def apply(psi, f):
        site = f % n
        t = psi.reshape([2] * n)   # axis a = qubit n - 1 - a (site a)
        g = gates[f]
        if g.shape[0] == 4:
            t = np.moveaxis(t, (site, site + 1), (0, 1)).reshape(4, -1)
            t = (g @ t).reshape([2, 2] + [2] * (n - 2))
            t = np.moveaxis(t, (0, 1), (site, site + 1))
        else:
            t = np.moveaxis(t, site, 0).reshape(2, -1)
            t = np.moveaxis((g @ t).reshape([2] + [2] * (n - 1)), 0, site)
        return t.reshape(-1)

    psi0 = np.zeros(1 << n, complex)
    psi0[0] = 1
    ref = psi0
    for f in range(k):
        ref = apply(ref, f)
    rec, _ = mod.forward_plan(n, layers)
    out = psi0
    for f in rec[:, 2].tolist():
        out = apply(out, f)
    assert np.allclose(out, ref, atol=1e-12)
    assert not np.allclose(ref, psi0)


def test_forward_plan_rejects_bad_sizes():
    mod, lib = _lib()
    nr, ns = ctypes.c_int(0), ctypes.c_int(0)
    assert lib.qmx_forward_plan(4, 2, None, 0, None, 0, ctypes.byref(nr), ctypes.byref(ns)) != 0
    assert lib.qmx_forward_plan(31, 2, None, 0, None, 0, ctypes.byref(nr), ctypes.byref(ns)) != 0
    st = lib.qmx_forward_plan(10, 2, None, 0, None, 0, ctypes.byref(nr), ctypes.byref(ns))
    assert st == mod.QMX_ERR_WORKSPACE and nr.value == 20 and ns.value >= 1


def test_every_package_module_imports():
    """Every module of the package and bench.py import without a GPU (the GPU
    is touched only by calls)."""
    import importlib
    import pkgutil

    import qmprs_amd
    for m in pkgutil.walk_packages(qmprs_amd.__path__, "qmprs_amd."):
        if m.name.endswith("libqmx"):     # the C library itself (loaded by ctypes)
            continue
        importlib.import_module(m.name)
    importlib.import_module("bench")
