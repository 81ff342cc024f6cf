"""Input side of prepare_state on the GPU against the oracle.

* quick Ket re-indexing / compression (base.py:99-102; the rules assumed for
  quick, which is not installed: the oracle's change_indexing and
  compress_amplitudes; parity against quick itself is unpinned): bit-exact,
  including inputs with many equal magnitudes (quantised "images"), where the
  zeroed amplitudes among equals are taken in index order.
* MPS(mps=...) (mps.py:183-197): the dense state of given site tensors
  (qmx_mps_to_state) to 1e-12 relative, and prepare_mps on it.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import qmprs_oracle as O

pytestmark = pytest.mark.gpu


def _image(n, seed, levels=8):
    """Quantised non-negative amplitudes: many exact ties in magnitude."""
    rng = np.random.default_rng(seed)
    return rng.integers(0, levels, 2 ** n).astype(np.float64) + 0j


@pytest.mark.parametrize("n", [3, 8, 12, 20])
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128], ids=["c64", "c128"])
@pytest.mark.parametrize("kind", ["gauss", "image"])
def test_prepare_input_matches_oracle(gpu, n, dtype, kind):
    from qmprs_amd import engine
    if kind == "image":
        sv = _image(n, n).astype(dtype)
    else:
        rng = np.random.default_rng(100 + n)
        sv = (rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)).astype(dtype)
    for index_type in ("row", "snake"):
        for pct in (0.0, 10.0, 37.5, 99.9):
            print(f"prepare_input N={n} {kind} {np.dtype(dtype).name} {index_type} {pct}%", flush=True)
            got = engine.prepare_input(sv, index_type, pct)[0].cpu().numpy()
            want = O.compress_amplitudes(O.change_indexing(sv, index_type), pct) if pct > 0 \
                else O.change_indexing(sv, index_type)
            assert got.dtype == sv.dtype
            np.testing.assert_array_equal(got, want)


def test_prepare_input_batch_and_errors(gpu):
    from qmprs_amd import engine
    svs = np.stack([_image(10, s) for s in range(3)])
    got = engine.prepare_input(svs, "snake", 50.0).cpu().numpy()
    for b in range(3):
        np.testing.assert_array_equal(got[b], O.compress_amplitudes(O.change_indexing(svs[b], "snake"), 50.0))
    with pytest.raises(ValueError, match="every amplitude"):
        engine.prepare_input(svs, "row", 100.0)
    with pytest.raises(ValueError, match="Index type"):
        engine.prepare_input(svs, "column", 0.0)


def test_prepare_state_with_snake_and_compression(gpu):
    """prepare_state(sv, chi, compression_percentage, index_type="snake") through
    the Sequential API equals the oracle's encoding of the preprocessed input
    (complex128 sweeps): fidelity to 1e-9, circuit state to 1e-8."""
    from qmprs_amd.circuit import GateCircuit
    from qmprs_amd.synthesis.mps_encoding import Sequential
    n, chi, L, S, pct = 8, 16, 3, 2, 30.0
    img = _image(n, 7, levels=32)
    enc = Sequential(GateCircuit)
    circ = enc.prepare_state(img, chi, pct, "snake", num_layers=L, num_sweeps=S)
    prep = O.compress_amplitudes(O.change_indexing(img, "snake"), pct)
    ref = O.prepare_state(prep, chi, num_layers=L, num_sweeps=S, polar="jacobi")
    state = circ.get_statevector()
    fid = enc.last_info["fidelity"]
    print(f"fidelity {fid:.12f} oracle {ref['fidelity']:.12f} state err {np.linalg.norm(state - ref['state']):.2e}")
    assert abs(fid - ref["fidelity"]) < 1e-9
    assert np.linalg.norm(state - ref["state"]) < 1e-8


def _random_mps(bonds, seed):
    rng = np.random.default_rng(seed)
    return [(rng.standard_normal((bonds[i], 2, bonds[i + 1]))
             + 1j * rng.standard_normal((bonds[i], 2, bonds[i + 1]))) / np.sqrt(2)
            for i in range(len(bonds) - 1)]


@pytest.mark.parametrize("bonds", [[1, 2, 1], [1, 2, 4, 3, 5, 8, 5, 4, 2, 2, 1],
                                   [1] + [min(2 ** i, 2 ** (23 - i), 16) for i in range(1, 23)] + [1]],
                         ids=["n2", "n10-ragged", "n23"])
def test_mps_entry_state_matches_oracle(gpu, bonds):
    """MPS(mps=site tensors): the dense state (qmx_mps_to_state, a chain of FP64
    GEMMs, row-chunked above 2^21 rows) equals the oracle's contraction."""
    from qmprs_amd.primitives import MPS
    ts = _random_mps(bonds, len(bonds))
    print(f"MPS entry: N={len(ts)} bonds {bonds[:6]}...", flush=True)
    m = MPS(mps=ts)
    got = m.state[0].cpu().numpy()
    want = O.mps_to_dense(ts)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"rel err {err:.2e}")
    assert err < 1e-12
    assert m.num_sites == len(ts)
    np.testing.assert_allclose(np.linalg.norm(m._psi[0].cpu().numpy()), 1.0, rtol=1e-14)


def test_prepare_mps_from_site_tensors(gpu):
    """Sequential.prepare_mps on MPS(mps=...) (used as given, not truncated,
    as mps.py:190-197 keeps it) equals the oracle's encoding of that state."""
    from qmprs_amd.circuit import GateCircuit
    from qmprs_amd.primitives import MPS
    from qmprs_amd.synthesis.mps_encoding import Sequential
    bonds = [1, 2, 4, 8, 8, 8, 4, 2, 1]
    ts = _random_mps(bonds, 3)
    n, L, S = len(ts), 3, 2
    enc = Sequential(GateCircuit)
    circ = enc.prepare_mps(MPS(mps=ts), num_layers=L, num_sweeps=S)
    ref = O.prepare_state(O.mps_to_dense(ts), 2 ** n, num_layers=L, num_sweeps=S, polar="jacobi")
    err = np.linalg.norm(circ.get_statevector() - ref["state"])
    print(f"fidelity {enc.last_info['fidelity']:.12f} oracle {ref['fidelity']:.12f} state err {err:.2e}")
    assert abs(enc.last_info["fidelity"] - ref["fidelity"]) < 1e-9
    assert err < 1e-8
