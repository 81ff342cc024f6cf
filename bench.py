"""Benchmark: state-preps/sec of the Sequential encoder on one MI355X.

Workload (default): N=20 qubits, statevectors of complex64 amplitudes from
seeded random MPS of bond 96 (complex Gaussian tensors, contracted in
complex128, normalised), bond_dimension=64, num_layers=10, num_sweeps=30,
a batch of 256 states (seeds 1000..1255) resident in HBM, swept as two
half-batches of 128 on two streams (throughput saturates from 192 states:
the goal-stopped rate measured 51.1 / 53.1 / 53.6 state-preps/s at 128 /
192 / 256 states).

The metric is state-preps/sec at fidelity >= F_ref - 1e-4, with F_ref the
fidelity of the reference algorithm in FP64: the FP64 path (complex128
sweep storage, the path the FP64 oracle pins gate by gate) after 30 sweeps
on the same states.  It is measured first (the ``fp64_reference`` object,
also timed: the parity mode's own throughput).  The headline ``value`` is
the plain method in its throughput mode (complex64 sweep storage, FP32
gate arithmetic, FP64 eigensolves / layers / polar / reductions): a step is
one whole prepare_state of the batch (chi-truncation, layer extraction,
sweeps, final fidelity) whose sweeps stop per state once its fidelity
reaches F_ref - 1e-4, at most ``--max-sweeps`` (45).  The complex64
trajectory leaves the FP64 one by rounding that the sweeps amplify (either
sign), so it may need a sweep more or fewer than 30; ``fidelity_vs_fp64``
states that drift after exactly 30 sweeps and checks the goal with an FP64
re-evaluation of every state's final gates.

``roofline`` describes the dominant kernel, the windowed sweep's streaming
pass (window_pass_kernel): per launch it reads T' and phi and writes both
back, 32 B per complex64 amplitude per active state (24 B for the first
window of a sweep, which reads the target and writes no T', 16 B for the
last, whose results are not read again), so algorithmic bytes = state-sweeps
x 2^N x (32 (W - 2) + 24 + 16) over the W windows of a sweep.  Divided by the
summed launch durations, measured with HIP events recorded by the library
on the stream the kernel is launched on, launch by launch, over the timed
steps.  ``traffic`` comes from two rocprofv3 counter runs (FETCH_SIZE, then
WRITE_SIZE) of a one-sweep probe of this same code and batch, calibrated by
the same counters on a kernel that reads and writes 1 GiB (null when
rocprofv3 is absent).
``cpu_baseline`` times the FP64 NumPy oracle (a port of the reference
algorithm) on the host cores: chi-truncation + all layers + one sweep per
state, the other sweeps scaled from the timed one, median over 3 states.
"""

from __future__ import annotations

import argparse
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
WORKLOAD = {"num_qubits": 20, "bond_dimension": 64, "num_layers": 10, "num_sweeps": 30,
            "generator_bond": 96}
GOAL_TOL = 1e-4
DTYPE_C64 = ("complex64 sweep storage, FP32 gate arithmetic; FP64 eigensolves, layers, "
             "polar updates and reductions")
DTYPE_C128 = "complex128 throughout (FP64 arithmetic)"


def _states(n: int, bond: int, count: int, seed0: int) -> np.ndarray:
    rng_states = []
    for s in range(count):
        rng = np.random.default_rng(seed0 + s)
        dims = [1] + [min(bond, 2 ** i, 2 ** (n - i)) for i in range(1, n)] + [1]
        psi = np.ones((1, 1), dtype=np.complex128)
        for i in range(n):
            a = (rng.standard_normal((dims[i], 2, dims[i + 1]))
                 + 1j * rng.standard_normal((dims[i], 2, dims[i + 1]))) / np.sqrt(2.0)
            psi = (psi @ a.reshape(dims[i], 2 * dims[i + 1])).reshape(-1, dims[i + 1])
        psi = psi.reshape(-1)
        rng_states.append((psi / np.linalg.norm(psi)).astype(np.complex64))
    return np.stack(rng_states)


def _host_cores() -> tuple[int, int]:
    """(cores this process may run on, its CPU share): the affinity mask,
    capped by OMP_NUM_THREADS where the host sets it to the process's share."""
    aff = len(os.sched_getaffinity(0))
    share = aff
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        share = min(aff, int(env))
    return aff, share


def cpu_baseline(svs: np.ndarray, wl: dict, gpu_fid_trace=None, seeds: int = 3, gpu_chi=None) -> dict:
    """FP64 oracle timing on ``seeds`` states of the workload (bounded sample,
    each scaled to a full prep), median state-preps/s.  On the first state
    the oracle's results also check the GPU's: fidelity after 0 and 1 FP64
    sweeps (``gpu_fid_trace``) and the chi-truncation (``gpu_chi`` = (target,
    spectra) of qmx_chi_truncate): the truncated state's relative error and
    the retained squared singular values per bond."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from threadpoolctl import threadpool_limits
    from oracle import qmprs_oracle as O
    n, chi, L, S = wl["num_qubits"], wl["bond_dimension"], wl["num_layers"], wl["num_sweeps"]
    aff, cores = _host_cores()
    rates, parts = [], []
    parity = None
    with threadpool_limits(limits=cores, user_api="blas"):
        for k in range(seeds):
            t0 = time.perf_counter()
            psi = O.ket(svs[k].astype(np.complex128))
            target, spectra = O.chi_truncate(psi, n, chi)
            t1 = time.perf_counter()
            layers = O.get_unitary_layers(target, n, L)
            t2 = time.perf_counter()
            gl = O.layer_gate_list(layers)
            f0 = abs(np.vdot(psi, O.circuit_state(n, gl)))
            t3 = time.perf_counter()
            gl, _ = O.sweep_once(target, n, gl)
            t4 = time.perf_counter()
            per_state = (t1 - t0) + (t2 - t1) + S * (t4 - t3)
            rates.append(1.0 / per_state)
            parts.append((t1 - t0, t2 - t1, t4 - t3))
            if k == 0 and gpu_fid_trace is not None:
                f1 = abs(np.vdot(psi, O.circuit_state(n, gl)))
                parity = {"oracle_fid_sweep0": f0, "oracle_fid_sweep1": f1,
                          "gpu_fp64_fid_sweep0": float(gpu_fid_trace[0]),
                          "gpu_fp64_fid_sweep1": float(gpu_fid_trace[1])}
            if k == 0 and gpu_chi is not None:
                g_target, g_spec = gpu_chi
                parity = parity or {}
                parity["chi_target_rel_err"] = float(np.linalg.norm(g_target - target) / np.linalg.norm(target))
                # retained squared singular values (eigenvalues of the bond Gram
                # matrices), largest error relative to each bond's largest
                errs = []
                for i, lam in enumerate(spectra):
                    if lam is None:
                        continue
                    kk = min(len(lam), chi)
                    errs.append(float(np.max(np.abs(g_spec[i, :kk] - lam[:kk])) / lam[0]))
                parity["chi_spectra_max_rel_err"] = max(errs) if errs else 0.0
                parity["chi_bonds_checked"] = len(errs)
    med = float(np.median(rates))
    c, l, w = (float(np.median([p[i] for p in parts])) for i in range(3))
    out = {"value": med, "unit": "state-preps/s", "cores": cores, "kind": "port",
           "sample": (f"{seeds} states of the workload, FP64 NumPy oracle on {cores} BLAS threads "
                      f"(affinity {aff} cores): median chi-truncation {c:.1f}s + {L} layers {l:.1f}s + "
                      f"1 of {S} sweeps {w:.1f}s timed; per-state time = chi + layers + {S} x sweep; "
                      f"value = median over the states"),
           "per_state": [round(r, 6) for r in rates]}
    if parity is not None:
        out["parity"] = parity
    return out


def _algorithmic_bytes(n: int, windows: int, state_sweeps: float) -> float:
    amp_bytes = 16.0 if windows == 1 else (32.0 * (windows - 2) + 24.0 + 16.0)
    return state_sweeps * (2 ** n) * amp_bytes


def traffic_probe(args) -> int:
    """Counter-run body: a 1 GiB elementwise copy (calibration), then one encode of
    the bench batch with one sweep.  Run under rocprofv3 --pmc by
    ``measure_traffic``."""
    import torch
    from qmprs_amd import _lib, engine
    wl = WORKLOAD
    dev = _lib.require_gpu()
    # calibration: one elementwise kernel that reads 1 GiB and writes 1 GiB
    src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    torch.neg(src, out=dst)
    torch.cuda.synchronize()
    del src, dst
    svs = _states(wl["num_qubits"], wl["generator_bond"], args.batch, seed0=1000)
    engine.encode(engine.as_device_states(svs), wl["bond_dimension"], wl["num_layers"], 1,
                  sweep_dtype=torch.complex64)
    torch.cuda.synchronize()
    print("traffic probe done", flush=True)
    return 0


def rebuild_probe(args) -> int:
    """Profiling body for the circuit rebuild: the bench batch's gates (chi
    truncation and layer extraction of the seeded states, no sweeps), then
    qmx_circuit_state in ``--rebuild-dtype`` (complex64 or complex128) three
    times with the forward passes on the matrix cores (QMX_OPT_FWD_MFMA=1)
    and three times on the vector units (QMX_OPT_FWD_MFMA=0), each after one
    untimed call.  Prints the mean wall time per call of each."""
    import time
    import torch
    from qmprs_amd import _lib, engine
    wl = WORKLOAD
    _lib.require_gpu()
    svs = _states(wl["num_qubits"], wl["generator_bond"], args.batch, seed0=1000)
    r = engine.encode(engine.as_device_states(svs), wl["bond_dimension"], wl["num_layers"], 0,
                      sweep_dtype=torch.complex64)
    torch.cuda.synchronize()
    dtype = getattr(torch, args.rebuild_dtype)
    try:
        for mode in (1, 0):
            _lib.set_option(_lib.QMX_OPT_FWD_MFMA, mode)
            engine.circuit_state(r.gates, r.widths, r.layers_used, dtype=dtype)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(3):
                engine.circuit_state(r.gates, r.widths, r.layers_used, dtype=dtype)
            torch.cuda.synchronize()
            print(f"QMX_OPT_FWD_MFMA={mode}: {(time.perf_counter() - t0) / 3 * 1e3:.2f} ms per "
                  f"{args.rebuild_dtype} circuit_state of {args.batch} states", flush=True)
    finally:
        _lib.set_option(_lib.QMX_OPT_FWD_MFMA, _lib.FWD_MFMA_DEFAULT)
    return 0


def _counter_means(csv_dir: str) -> dict:
    """{kernel-name key: (mean counter value, dispatches)} from rocprofv3's csv."""
    import csv
    files = glob.glob(os.path.join(csv_dir, "**", "*counter_collection.csv"), recursive=True)
    sums: dict = {}
    for f in files:
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                name = row.get("Kernel_Name", "")
                key = ("window_pass" if "window_pass_kernel" in name else
                       "copy" if "neg_kernel" in name else None)
                if key is None:
                    continue
                v = float(row.get("Counter_Value", "nan"))
                s, c = sums.get(key, (0.0, 0))
                sums[key] = (s + v, c + 1)
    return {k: (s / c, c) for k, (s, c) in sums.items() if c}


def measure_traffic(args, n: int, states_per_launch: int) -> dict | None:
    """HBM bytes per window-pass launch from two rocprofv3 counter runs of
    ``traffic_probe`` (FETCH_SIZE, then WRITE_SIZE; one counter per run),
    each scaled by what the same counter reads on the probe's 1 GiB
    elementwise copy (torch.neg of 2^28 floats: 1 GiB read, 1 GiB written)."""
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None
    here = os.path.dirname(os.path.abspath(__file__))
    out = {}
    env = dict(os.environ, TMPDIR=tempfile.gettempdir())
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix=f"qmx_{counter.lower()}_")
        cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "probe", "--",
               sys.executable, os.path.join(here, "bench.py"), "--traffic-probe", "--batch", str(args.batch)]
        try:
            r = subprocess.run(cmd, cwd=here, env=env, capture_output=True, text=True, timeout=240)
        except subprocess.TimeoutExpired:
            return {"error": f"{counter} run timed out"}
        means = _counter_means(d)
        shutil.rmtree(d, ignore_errors=True)
        if r.returncode not in (0, -11, 139) or "window_pass" not in means or "copy" not in means:
            return {"error": f"{counter} run: rc {r.returncode}, kernels {sorted(means)}"}
        copy_kib = means["copy"][0]
        factor = 1048576.0 / copy_kib
        out[counter] = {"raw_kib": means["window_pass"][0], "dispatches": means["window_pass"][1],
                        "copy_kib": copy_kib, "factor": round(factor, 4),
                        "kib": means["window_pass"][0] * factor}
    # the probe's window passes: one sweep of full launches
    windows = out["FETCH_SIZE"]["dispatches"] // 2 if args.batch >= 2 else out["FETCH_SIZE"]["dispatches"]
    alg = _algorithmic_bytes(n, windows, states_per_launch) / max(windows, 1)
    traffic = (out["FETCH_SIZE"]["kib"] + out["WRITE_SIZE"]["kib"]) * 1024.0
    return {"bytes_per_launch": traffic, "algorithmic_bytes_per_launch": alg,
            "ratio": traffic / alg, "counters": out,
            "probe": f"one sweep of the batch ({states_per_launch} states per launch), "
                     f"{windows} window passes per half-batch"}


# learned warm start: trajectory pairs from the plain method on N=16, chi=32,
# 8 layers, 30 sweeps (416 states x 8 layers x 15 two-qubit slots x 11 sweeps
# = 549120 pairs; 32 held-out states for the held-out loss)
WARM_TRAIN = {"num_qubits": 16, "bond_dimension": 32, "num_layers": 8, "num_sweeps": 30,
              "generator_bond": 48, "states": 416, "held_out": 32, "epochs": 10, "batch": 512, "lr": 1e-3,
              "width": 512}


def _fp64_check(res, psi128, goal) -> dict:
    """FP64 re-evaluation of each state's final gates (qmx_circuit_state in
    complex128): |<psi|C|0>| against the goal."""
    from qmprs_amd import engine
    _, f = engine.circuit_state(res.gates, res.widths, res.layers_used, psi_ref=psi128)
    f = f.abs()
    return {"goal_reached": float((f >= goal).double().mean()),
            "fidelity_mean": float(f.mean()), "fidelity_min": float(f.min())}


def warm_start_leg(svs, psi128, goal, wl, args) -> dict:
    """Train the trajectory network on the GPU (untimed), then time the
    encode with learned jumps of the same batch against the FP64 goal."""
    import torch
    from qmprs_amd import learned
    tw = WARM_TRAIN
    t0 = time.perf_counter()
    train_sv = _states(tw["num_qubits"], tw["generator_bond"], tw["states"], seed0=5000)
    held_sv = _states(tw["num_qubits"], tw["generator_bond"], tw["held_out"], seed0=9000)
    data = learned.make_trajectory_set(train_sv, tw["bond_dimension"], tw["num_layers"], tw["num_sweeps"])
    held = learned.make_trajectory_set(held_sv, tw["bond_dimension"], tw["num_layers"], tw["num_sweeps"])
    t1 = time.perf_counter()
    net = learned.WarmStartMLP(width=tw["width"], seed=0)
    hist = learned.train(net, data, epochs=tw["epochs"], batch=tw["batch"], lr=tw["lr"], valid=learned.TRAJ_OUT)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    chi, L = wl["bond_dimension"], wl["num_layers"]
    sv_dev = torch.from_numpy(svs).cuda()
    steps = args.leg_steps
    for _ in range(args.warmup):
        res = learned.encode_warm(sv_dev, chi, L, args.max_sweeps, net, goal=goal)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    for _ in range(steps):
        res = learned.encode_warm(sv_dev, chi, L, args.max_sweeps, net, goal=goal)
    torch.cuda.synchronize()
    t4 = time.perf_counter()
    out = {
        "value": round(len(svs) * steps / (t4 - t3), 3), "unit": "state-preps/s", "steps": steps,
        "ms_per_step": round(1000.0 * (t4 - t3) / steps, 2),
        "method": f"plain complex64 sweeps with a learned jump after sweep {learned.JUMP_FIRST} and every "
                  f"{learned.JUMP_EVERY} sweeps after it (each predicting {learned.TRAJ_K} sweeps ahead), kept "
                  f"per state if the fidelity does not drop",
        "mean_sweeps": float(res.sweeps.float().mean()), "max_sweeps": int(res.sweeps.max()),
        "jumps_kept_per_state": float(res.accepted.float().mean()), "jump_points": res.tried,
        "train": {"pairs": int(data.weights.sum().item()), "config": {k: tw[k] for k in (
            "num_qubits", "bond_dimension", "num_layers", "num_sweeps", "states", "epochs", "width")},
            "data_s": round(t1 - t0, 1), "train_s": round(t2 - t1, 1),
            "loss_first_epoch": hist[0], "loss_last_epoch": hist[-1],
            "held_out_mse": learned.evaluate_loss(net, held, learned.TRAJ_OUT),
            "held_out_mse_zero_predictor": learned.zero_loss(held, learned.TRAJ_OUT)},
    }
    out.update(_fp64_check(res, psi128, goal))
    return out


def relaxed_leg(sv_dev, psi128, goal, wl, args) -> dict:
    """The same batch with the over-relaxed gate update (omega =
    ``--relax-omega``; an option with no reference counterpart) against the
    FP64 goal, at most ``--max-sweeps`` sweeps."""
    import torch
    from qmprs_amd import _lib, engine
    chi, L = wl["bond_dimension"], wl["num_layers"]
    steps = args.leg_steps
    _lib.set_sweep_relaxation(args.relax_omega)
    try:
        for _ in range(args.warmup):
            res = engine.encode(sv_dev, chi, L, args.max_sweeps, sweep_dtype=torch.complex64, goal=goal)
        torch.cuda.synchronize()
        # phase timers on, as in the plain leg's timed steps (same overhead)
        _lib.profile_enable(True)
        _lib.profile_read()
        t0 = time.perf_counter()
        for _ in range(steps):
            res = engine.encode(sv_dev, chi, L, args.max_sweeps, sweep_dtype=torch.complex64, goal=goal)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        prof = _lib.profile_read()
    finally:
        _lib.profile_enable(False)
        _lib.set_sweep_relaxation(1.0)
    out = {
        "value": round(sv_dev.shape[0] * steps / (t1 - t0), 3), "unit": "state-preps/s", "steps": steps,
        "ms_per_step": round(1000.0 * (t1 - t0) / steps, 2),
        "omega": args.relax_omega,
        "mean_sweeps": float(res.sweeps_done.double().mean()),
        "max_sweeps": int(res.sweeps_done.max()),
        "phases_ms_per_step": {p: {"ms": round(v[0] / steps, 2), "launches": int(v[1] / steps)}
                               for p, v in prof.items()},
    }
    out.update(_fp64_check(res, psi128, goal))
    return out


# ragged batch: 256 states, N ~ U{12..20}, generating bond D ~ U{16, 32, 64},
# bond_dimension = D, 8 layers, 30 sweeps, complex64
RAGGED = {"states": 256, "num_qubits": (12, 20), "bonds": (16, 32, 64), "num_layers": 8,
          "num_sweeps": 30, "seed": 7}


def ragged_main(args) -> int:
    import torch
    from qmprs_amd import _lib, engine
    rg = RAGGED
    rng = np.random.default_rng(rg["seed"])
    ns = rng.integers(rg["num_qubits"][0], rg["num_qubits"][1] + 1, rg["states"])
    ds = rng.choice(rg["bonds"], rg["states"])
    svs = [_states(int(n), int(d), 1, seed0=10000 + i)[0] for i, (n, d) in enumerate(zip(ns, ds))]
    _lib.require_gpu()
    dev_svs = [engine.as_device_states(v) for v in svs]
    L, S = rg["num_layers"], rg["num_sweeps"]

    def run(conc):
        return engine.encode_ragged(dev_svs, [int(d) for d in ds], L, S, max_concurrency=conc)

    out = {}
    for conc in (1, 6):
        for _ in range(args.warmup):
            run(conc)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            res = run(conc)
        torch.cuda.synchronize()
        out[conc] = (time.perf_counter() - t0) / args.steps
    fid = np.array([abs(r.fidelity[-1].item()) for r in res])
    buckets = len(set(ns.tolist()))   # one encode per qubit count (bond dimensions grouped inside)
    value = rg["states"] / out[6]
    line = {
        "metric": "state-preps/sec, ragged batch of 256 states (N 12..20, chi 16/32/64)",
        "value": round(value, 3), "unit": "state-preps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1000 * out[6], 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE_C64,
        "data": "synthetic: seeded random-MPS statevectors (complex Gaussian, generating bond = chi)",
        "config": {"workload": "ragged batch: 256 states, N ~ U{12..20}, bond_dimension = generating "
                               "bond ~ U{16,32,64}, num_layers=8, num_sweeps=30", "buckets": buckets,
                   "max_concurrency": 6},
        "sequential_buckets": {"value": round(rg["states"] / out[1], 3), "ms_per_step": round(1000 * out[1], 2)},
        "fidelity": {"mean": float(fid.mean()), "min": float(fid.min()), "max": float(fid.max())},
    }
    print(json.dumps(line))
    return 0


LARGE = {"num_qubits": 26, "bond_dimension": 512, "num_layers": 20, "num_sweeps": 50,
         "generator_bond": 1024, "seed": 2026}


def _large_state(dev):
    """Seeded random MPS of bond 1024 at N=26 (complex Gaussian tensors drawn
    on the host, contracted left to right in complex128 on the GPU),
    normalised, cast to complex64: 2^26 amplitudes, 512 MB."""
    import torch
    lg = LARGE
    n, bond = lg["num_qubits"], lg["generator_bond"]
    gen = torch.Generator(device="cpu").manual_seed(lg["seed"])
    x = None
    for i in range(n):
        dl = min(bond, 2 ** i, 2 ** (n - i))
        dr = min(bond, 2 ** (i + 1), 2 ** (n - i - 1))
        a = (torch.randn(dl, 2, dr, generator=gen, dtype=torch.float64)
             + 1j * torch.randn(dl, 2, dr, generator=gen, dtype=torch.float64)) / np.sqrt(2.0)
        a = a.to(dev).reshape(dl, 2 * dr)
        x = a if x is None else x @ a
        x = x.reshape(-1, dr)
    psi = x.reshape(1, -1)
    return (psi / torch.linalg.vector_norm(psi)).to(torch.complex64).contiguous()


def large_main(args) -> int:
    """One large state (N=26, chi=512, 20 layers, 50 sweeps) per step."""
    import torch
    from qmprs_amd import _lib, engine
    lg = LARGE
    n, chi, L, S = lg["num_qubits"], lg["bond_dimension"], lg["num_layers"], lg["num_sweeps"]
    dev = _lib.require_gpu()
    sv = _large_state(dev)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        engine.encode(sv, chi, L, S)
    torch.cuda.synchronize()
    _lib.profile_enable(True)
    _lib.profile_read()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = engine.encode(sv, chi, L, S)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    prof = _lib.profile_read()
    _lib.profile_enable(False)
    wp_ms, wp_launches = prof["window_pass"]
    windows = max(1, round(wp_launches / (args.steps * S)))
    amp_bytes = 16.0 if windows == 1 else (32.0 * (windows - 2) + 24.0 + 16.0) / windows
    total_bytes = args.steps * S * windows * (2 ** n) * amp_bytes
    achieved = total_bytes / (wp_ms / 1000.0) / 1e9
    fid = res.fidelity[0].abs().cpu().numpy()
    line = {
        "metric": "state-preps/sec, one large state (N=26, chi=512, 20 layers, 50 sweeps)",
        "value": round(args.steps / total, 5), "unit": "state-preps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1000 * total / args.steps, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE_C64,
        "data": "synthetic: seeded random-MPS statevector (bond 1024, complex Gaussian)",
        "config": {"workload": "large single state: N=26, bond_dimension=512, num_layers=20, num_sweeps=50, "
                               "complex64 input", "batch": 1},
        "fidelity": {"before_sweeps": float(fid[0]), "final": float(fid[-1]),
                     "layers_used": int(res.layers_used[0])},
        "phases_ms_per_step": {p: {"ms": round(v[0] / args.steps, 1), "launches": int(v[1] / args.steps)}
                               for p, v in prof.items()},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "window_pass_kernel<float2, 4>",
                     "avg_launch_us": round(1000.0 * wp_ms / max(wp_launches, 1), 2),
                     "launches": int(wp_launches), "windows_per_sweep": windows, "states_per_launch": 1,
                     "bytes_per_launch": int(total_bytes / max(wp_launches, 1))},
    }
    if args.relax_omega != 1.0:
        # goal: the plain run's final fidelity - 1e-4; the relaxed run's final
        # gates are re-evaluated in FP64 against the complex128 input
        goal = (res.fidelity[:, -1].abs() - 1e-4).double().contiguous()
        args.max_sweeps = max(args.max_sweeps, S)
        line["relaxed"] = relaxed_leg(sv, sv.to(torch.complex128), goal, lg, args)
    print(json.dumps(line))
    return 0


def _timed(fn, steps: int, warmup: int):
    import torch
    for _ in range(warmup):
        res = fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        res = fn()
    torch.cuda.synchronize()
    return res, time.perf_counter() - t0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=256,
                    help="states per step (held-out seeds 1000..; at least 64)")
    ap.add_argument("--max-sweeps", type=int, default=45,
                    help="sweep cap of the goal-stopped legs (the goal is F_ref - 1e-4)")
    ap.add_argument("--leg-steps", type=int, default=None,
                    help="timed steps of the secondary legs (default min(steps, 5))")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter runs that measure roofline.traffic")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--rebuild-probe", action="store_true",
                    help="profiling body: circuit rebuilds of the bench batch, matrix-core and "
                         "vector-unit forward passes (no JSON line)")
    ap.add_argument("--rebuild-dtype", choices=("complex64", "complex128"), default="complex64",
                    help="state type of --rebuild-probe")
    ap.add_argument("--no-warm-start", action="store_true",
                    help="skip the learned warm-start leg (training on N=16 plain-method "
                         "pairs, then the N=20 warm-started encode against the FP64 goal)")
    ap.add_argument("--relax-omega", type=float, default=1.7,
                    help="omega of the over-relaxed leg (the same batch with the gate update "
                         "polar(G' + (omega - 1)(G' - G)) and the FP64 goal); 1 skips the leg")
    ap.add_argument("--workload", choices=("single-size", "ragged", "large"), default="single-size",
                    help="single-size: the N=20, chi=64 batch (default); ragged: 256 states of "
                         "mixed N and chi through the ragged scheduler; large: one N=26, chi=512 "
                         "state, 20 layers, 50 sweeps (use --steps 1 --warmup 0)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this encoder shards across GPUs by state "
              "batches, and no multi-GPU figure is produced here", file=sys.stderr)
        return 2
    if args.leg_steps is None:
        args.leg_steps = max(1, min(args.steps, 5))

    import torch
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from qmprs_amd import _lib, engine
    if args.traffic_probe:
        return traffic_probe(args)
    if args.rebuild_probe:
        return rebuild_probe(args)
    if args.workload == "ragged":
        return ragged_main(args)
    if args.workload == "large":
        return large_main(args)

    wl = WORKLOAD
    n, chi, L, S = wl["num_qubits"], wl["bond_dimension"], wl["num_layers"], wl["num_sweeps"]
    B = args.batch
    svs = _states(n, wl["generator_bond"], B, seed0=1000)
    dev = _lib.require_gpu()
    sv_dev = engine.as_device_states(svs)

    # 1. FP64 reference: the reference algorithm's 30 sweeps with complex128
    #    storage; F_ref per state, and the parity mode's own throughput
    ref, ref_s = _timed(lambda: engine.encode(sv_dev, chi, L, S, sweep_dtype=torch.complex128),
                        args.leg_steps, 1)
    fref = ref.fidelity[:, -1].abs()
    goal = (fref - GOAL_TOL).double().contiguous()
    psi128 = engine.as_device_states(svs, torch.complex128)
    psi128 = psi128 / torch.linalg.vector_norm(psi128, dim=1, keepdim=True)
    fp64_reference = {
        "value": round(B * args.leg_steps / ref_s, 3), "unit": "state-preps/s", "steps": args.leg_steps,
        "ms_per_step": round(1000.0 * ref_s / args.leg_steps, 2), "dtype": DTYPE_C128,
        "num_sweeps": S, "fidelity_mean": float(fref.mean()), "fidelity_min": float(fref.min()),
        "fidelity_max": float(fref.max())}

    # 2. complex64 after exactly S sweeps: the drift against F_ref
    c64_30 = engine.encode(sv_dev, chi, L, S, sweep_dtype=torch.complex64)
    d30 = (c64_30.fidelity[:, -1].abs() - fref).cpu().numpy()

    # 3. headline: complex64, goal-stopped at F_ref - 1e-4 (at most max_sweeps)
    cfg = _lib.EncodeConfig(n, B, chi, L, args.max_sweeps, _lib.QMX_C64, _lib.QMX_C64, 1 - 1e-6, 1e-10)
    import ctypes
    ws = _lib.workspace(_lib.load().qmx_encode_workspace(ctypes.byref(cfg)), dev)

    def step():
        return engine.encode(sv_dev, chi, L, args.max_sweeps, sweep_dtype=torch.complex64, goal=goal,
                             workspace=ws)

    for _ in range(args.warmup):
        res = step()
    torch.cuda.synchronize()
    _lib.profile_enable(True)
    _lib.profile_read()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = _lib.profile_read()
    _lib.profile_enable(False)
    total = t1 - t0
    value = B * args.steps / total
    check = _fp64_check(res, psi128, goal)
    sweeps_done = res.sweeps_done.cpu().numpy()
    fid_c64 = res.fidelity[:, -1].abs().cpu().numpy()

    # dominant kernel: the windowed sweep's streaming pass.  The library runs
    # a batch as two half-batches whose passes alternate (never overlap); a
    # launch covers the half-batch's states that have not reached the goal.
    wp_ms, wp_launches = prof["window_pass"]
    jobs = 2 if B >= 2 else 1
    # a half-batch launches its W window passes in every sweep in which one of
    # its states is still short of the goal: launches = steps x W x sum over
    # the halves of their largest sweep count
    b1 = (B + 1) // 2
    half_sweeps = (int(sweeps_done[:b1].max()) + int(sweeps_done[b1:].max())) if jobs == 2 \
        else int(sweeps_done.max())
    windows = max(1, round(wp_launches / max(1, args.steps * half_sweeps)))
    state_sweeps = float(sweeps_done.sum()) * args.steps
    total_bytes = _algorithmic_bytes(n, windows, state_sweeps)
    bytes_per_launch = total_bytes / max(wp_launches, 1)
    achieved = total_bytes / (wp_ms / 1000.0) / 1e9
    traffic = None if args.no_traffic else measure_traffic(args, n, B // jobs)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": (int(traffic["bytes_per_launch"]) if traffic and "bytes_per_launch" in traffic
                            else None),
                "kernel": "window_pass_kernel<float2, 4>",
                "avg_launch_us": round(1e3 * wp_ms / max(wp_launches, 1), 2),
                "launches": int(wp_launches), "windows_per_sweep": windows,
                "states_per_launch_max": B // jobs,
                "bytes_per_launch": int(bytes_per_launch)}
    if traffic:
        roofline["traffic_detail"] = traffic
    step_bytes = total_bytes / args.steps
    phases = {p: {"ms": round(v[0] / args.steps, 2), "launches": int(v[1] / args.steps)}
              for p, v in prof.items()}
    out = {
        "metric": "state-preps/sec at fidelity >=ref-1e-4 (N=20, chi=64): plain vs NN-warm-started",
        "value": round(value, 3),
        "unit": "state-preps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * total / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE_C64,
        "data": "synthetic: seeded random-MPS statevectors (bond 96, complex Gaussian)",
        "config": {"workload": "plain Sequential.prepare_state, N=20, bond_dimension=64, "
                               "num_layers=10, num_sweeps=30 (F_ref), complex64 input",
                   "batch": B, "path": "plain, complex64 sweeps stopped per state at F_ref(FP64) - 1e-4",
                   "max_sweeps": args.max_sweeps},
        "fidelity_vs_fp64": {
            "goal": "F_ref - 1e-4 per state, F_ref = FP64 (complex128) path after 30 sweeps",
            "goal_reached": check["goal_reached"],
            "fp64_reevaluated_fidelity_mean": check["fidelity_mean"],
            "sweeps_mean": float(sweeps_done.mean()), "sweeps_max": int(sweeps_done.max()),
            "sweeps_min": int(sweeps_done.min()),
            "max_abs_dF_after_30_sweeps": float(np.abs(d30).max()),
            "mean_dF_after_30_sweeps": float(d30.mean()),
            "states_below_goal_after_30_sweeps": int((d30 < -GOAL_TOL).sum()),
            "reported_vs_fp64_reevaluation_max": float(np.abs(
                fid_c64 - engine.circuit_state(res.gates, res.widths, res.layers_used,
                                               psi_ref=psi128)[1].abs().cpu().numpy()).max())},
        "fp64_reference": fp64_reference,
        "phases_ms_per_step": phases,
        "roofline": roofline,
        "step_roofline": {"algorithmic_window_bytes_per_step": step_bytes,
                          "frac_of_step": round(step_bytes / (total / args.steps) / 1e9 / HBM_PEAK_GBS, 4),
                          "note": "window-pass bytes only, over the whole step time"},
    }
    if not args.no_warm_start:
        out["warm_start"] = warm_start_leg(svs, psi128, goal, wl, args)
    if args.relax_omega != 1.0:
        out["relaxed"] = relaxed_leg(sv_dev, psi128, goal, wl, args)
    if not args.no_cpu_baseline:
        # the GPU chi-truncation (FP64) of the first state, checked against the
        # oracle's inside the baseline leg
        _, tgt0, spec0 = engine.chi_truncate(sv_dev[:1].contiguous(), wl["bond_dimension"], spectra=True)
        out["cpu_baseline"] = cpu_baseline(svs, wl, ref.fidelity[0].abs().cpu().numpy()
                                           if ref.fidelity.shape[1] > 1 else None,
                                           gpu_chi=(tgt0[0].cpu().numpy(), spec0[0].cpu().numpy()))
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
