"""The oracle reproduces its committed golden fixtures (tests/golden/*.npz).

The fixtures were written by tests/golden/make_golden.py; a change of the
oracle that moves any output beyond rounding fails here.  The reference
itself ships no golden vectors (its tests hold thresholds only).
"""

from __future__ import annotations

import glob
import os

import numpy as np
import pytest

from oracle import qmprs_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def _load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_fixtures_present():
    assert len(FIXTURES) >= 4


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_oracle_reproduces_fixture(path):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    g = _load(path)
    n, chi = int(g["num_qubits"]), int(g["bond_dimension"])
    L, S = int(g["num_layers"]), int(g["num_sweeps"])
    s_run = S
    psi = O.ket(g["statevector"])
    target, _ = O.chi_truncate(psi, n, chi)
    np.testing.assert_allclose(target, g["target"], atol=1e-12)
    layers, info = O.get_unitary_layers(target, n, L, return_info=True)
    assert len(layers) == int(g["layers_used"])
    gates, widths = mg._arrays(layers, n)
    np.testing.assert_array_equal(widths, g["widths"])
    np.testing.assert_allclose(gates, g["gates"], atol=1e-10)
    np.testing.assert_allclose(info["amp0"], g["amp0"], atol=1e-12)
    gl = O.layer_gate_list(layers)
    for s in range(s_run):
        gl, ov = O.sweep_once(target, n, gl)
        assert abs(abs(ov) - g["overlap"][s]) < 1e-11
        assert abs(abs(np.vdot(psi, O.circuit_state(n, gl))) - g["fidelity"][s + 1]) < 1e-11
    if s_run == S:
        np.testing.assert_allclose(O.circuit_state(n, gl), g["state"], atol=1e-10)


def test_fixture_invariants():
    """Size-independent facts every fixture must satisfy: unitary gates,
    monotone overlap trace, final state = C|0..0> of the swept gates."""
    for path in FIXTURES:
        g = _load(path)
        n = int(g["num_qubits"])
        for mats in (g["gates"], g["swept_gates"]):
            for l in range(mats.shape[0]):
                for site in range(n):
                    d = 4 if g["widths"][l, site] == 2 else 2
                    m = mats[l, site, :d, :d]
                    np.testing.assert_allclose(m @ m.conj().T, np.eye(d), atol=1e-12)
        assert np.all(np.diff(g["overlap"]) >= -1e-12)
        assert abs(np.linalg.norm(g["state"]) - 1) < 1e-12
        assert abs(abs(np.vdot(O.ket(g["statevector"]), g["state"])) - g["fidelity"][-1]) < 1e-12
