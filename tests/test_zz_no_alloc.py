"""qmx_sweep_noalloc: the allocation-free variant of qmx_sweep.

Two calls in a row on the library's own scratch, the second on a smaller
problem (so it runs on whatever the first left there), each compared with
qmx_sweep on a caller workspace: bit for bit against the single-stream
qmx_sweep (QMX_OPT_SPLIT_SWEEPS = 0, the same kernels in the same order),
and against the default two-stream qmx_sweep to the complex128 bar (gates
1e-10, fidelities 1e-12).  Problems over the 64 MiB scratch are refused
before any GPU work.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import qmprs_oracle as O

pytestmark = pytest.mark.gpu


def _inputs(n, chi, layers, batch, dtype, seed):
    from qmprs_amd import engine
    svs = np.stack([O.random_mps_state(n, 2 * chi, seed + s) for s in range(batch)])
    return engine.encode_prepare(svs, chi, layers, sweep_dtype=dtype)


def _reference(p, sweeps, goal, split):
    from qmprs_amd import _lib, engine
    gates = p.gates.clone()
    _lib.set_split_sweeps(split)
    try:
        r = engine.sweep(p.target, gates, p.widths, p.layers_used, sweeps, psi_ref=p.psi, goal=goal,
                         circuit_out=True)
        torch.cuda.synchronize()
    finally:
        _lib.set_split_sweeps(True)
    return gates, r


def _noalloc(p, sweeps, goal):
    from qmprs_amd import engine
    b = p.target.shape[0]
    dev = p.target.device
    gates = p.gates.clone()
    fid = torch.zeros((b, sweeps + 1), dtype=torch.complex128, device=dev)
    ov = torch.zeros((b, sweeps), dtype=torch.float64, device=dev)
    done = torch.zeros(b, dtype=torch.int32, device=dev)
    circ = torch.empty_like(p.target)
    engine.sweep_noalloc(p.target, gates, p.widths, p.layers_used, sweeps, fidelity=fid, overlap=ov,
                         psi_ref=p.psi, goal=goal, sweeps_done=done, circuit_out=circ)
    torch.cuda.synchronize()
    return gates, fid, ov, done, circ


@pytest.mark.parametrize("order", ["large-then-small"])
def test_noalloc_sweep_equals_workspace_sweep(gpu, order):
    from qmprs_amd import _lib
    # call 1: N=18, 2 states, complex128 sweeps, 3 sweeps (32 MiB of scratch)
    # call 2: N=12, 1 state, complex64 sweeps, goal-stopped (smaller, reuses the scratch)
    cases = [(18, 16, 3, 2, torch.complex128, 3, 11), (12, 16, 4, 1, torch.complex64, 4, 12)]
    for k, (n, chi, layers, batch, dtype, sweeps, seed) in enumerate(cases):
        print(f"no-alloc call {k + 1}: N={n} batch={batch} {dtype} sweeps={sweeps}", flush=True)
        p = _inputs(n, chi, layers, batch, dtype, seed)
        need = _lib.load().qmx_sweep_workspace(n, batch, 1 if dtype == torch.complex128 else 0)
        assert need <= _lib.NOALLOC_SCRATCH_BYTES
        goal = None
        if k == 1:
            # a goal met after the first sweep: the goal path (skip flags, sweeps_done) runs too
            g1, r1 = _reference(p, 1, None, False)
            goal = r1.fidelity[:, -1].abs().double().contiguous()
        ga, fa, oa, da, ca = _noalloc(p, sweeps, goal)
        gr, rr = _reference(p, sweeps, goal, False)
        assert torch.equal(ga, gr)
        assert torch.equal(fa, rr.fidelity)
        assert torch.equal(oa, rr.overlap)
        assert torch.equal(ca, rr.circuit)
        if goal is not None:
            assert torch.equal(da, rr.sweeps_done)
            assert int(da.max()) == 1
        gs, rs = _reference(p, sweeps, goal, True)
        gdiff = (ga - gs).abs().max().item()
        fdiff = (fa - rs.fidelity).abs().max().item()
        print(f"  bitwise equal to the single-stream qmx_sweep; vs two-stream: gates {gdiff:.2e}, "
              f"fidelity {fdiff:.2e}", flush=True)
        tol_g, tol_f = (1e-10, 1e-12) if dtype == torch.complex128 else (1e-4, 1e-6)
        assert gdiff < tol_g
        assert fdiff < tol_f


def test_noalloc_sweep_refuses_oversized_problems(gpu):
    """qmx_sweep_workspace(20, 4, complex128) > 64 MiB: QMX_ERR_WORKSPACE ->
    ValueError, raised on the host before any kernel runs."""
    from qmprs_amd import _lib, engine
    n, b = 20, 4
    assert _lib.load().qmx_sweep_workspace(n, b, 1) > _lib.NOALLOC_SCRATCH_BYTES
    dev = gpu
    target = torch.zeros((b, 1 << n), dtype=torch.complex128, device=dev)
    gates = torch.zeros((b, 1, n, 4, 4), dtype=torch.complex128, device=dev)
    widths = torch.ones((b, 1, n), dtype=torch.int32, device=dev)
    used = torch.ones(b, dtype=torch.int32, device=dev)
    with pytest.raises(ValueError, match="scratch"):
        engine.sweep_noalloc(target, gates, widths, used, 1)
