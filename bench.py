"""Benchmark: state-preps/sec of the Sequential encoder on one MI355X.

Workload (default): N=20 qubits, statevectors of complex64 amplitudes from
seeded random MPS of bond 96 (complex Gaussian tensors, contracted in
complex128, normalised), bond_dimension=64, num_layers=10, num_sweeps=30.
A step is one whole prepare_state (chi-truncation, layer extraction, 30
sweeps, final fidelity) of one batch of such states on the GPU.

Prints one JSON line (rank 0).  ``roofline`` describes the dominant kernel,
the windowed sweep's streaming pass (window_pass_kernel): per launch it reads
T' and phi and writes both back, 32 B per complex64 amplitude per state (24 B
for the first window of a sweep, which reads the target and writes no T',
16 B for the last, whose results are not read again), so algorithmic bytes
per launch = B * 2^N * (32 (W - 2) + 24 + 16) / W averaged over the W
windows of a sweep.  It is divided by the average launch duration,
measured with HIP events recorded by the library on the stream the kernel is
launched on, launch by launch, over the timed steps.  ``traffic`` is filled
from rocprofv3 FETCH_SIZE / WRITE_SIZE counters when --traffic-kib gives them
(KiB per dispatch, already calibrated against a plain copy).
``cpu_baseline`` times the FP64 NumPy oracle (a port of the reference
algorithm) on one state: chi-truncation + all layers + one sweep, with the
other sweeps scaled from the timed one.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
# Calibrated HBM traffic of one window_pass_kernel dispatch on this workload,
# from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs (FETCH_SIZE x 2 and
# WRITE_SIZE x 1, the factors a plain 1 GiB copy reads on the same counters):
# states per launch -> (read KiB, write KiB).  Summaries: profiles/window_pass_n20_*.md.
MEASURED_TRAFFIC_KIB = {32: (524977.0, 528714.0)}
WORKLOAD = {"num_qubits": 20, "bond_dimension": 64, "num_layers": 10, "num_sweeps": 30,
            "generator_bond": 96}


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


def _blas_threads() -> int:
    """Threads the oracle's BLAS calls use (NumPy's own loops are serial)."""
    try:
        from threadpoolctl import threadpool_info
        n = [i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"]
        return int(max(n)) if n else 1
    except Exception:
        return 1


def cpu_baseline(sv: np.ndarray, wl: dict, gpu_fid_trace=None) -> dict:
    """Oracle timing on one state (bounded sample, scaled to a full prep)."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import qmprs_oracle as O
    n, chi, L, S = wl["num_qubits"], wl["bond_dimension"], wl["num_layers"], wl["num_sweeps"]
    t0 = time.perf_counter()
    psi = O.ket(sv.astype(np.complex128))
    target, _ = O.chi_truncate(psi, n, chi)
    t1 = time.perf_counter()
    layers = O.get_unitary_layers(target, n, L)
    t2 = time.perf_counter()
    gl = O.layer_gate_list(layers)
    f0 = abs(np.vdot(psi, O.circuit_state(n, gl)))
    t3 = time.perf_counter()
    gl, _ = O.sweep_once(target, n, gl)
    t4 = time.perf_counter()
    f1 = abs(np.vdot(psi, O.circuit_state(n, gl)))
    per_state = (t1 - t0) + (t2 - t1) + S * (t4 - t3)
    out = {"value": 1.0 / per_state, "unit": "state-preps/s",
           "cores": _blas_threads(), "kind": "port",
           "sample": (f"1 state, FP64 NumPy oracle: chi-truncation {t1 - t0:.1f}s + {L} layers "
                      f"{t2 - t1:.1f}s + 1 of {S} sweeps {t4 - t3:.1f}s timed; "
                      f"per-state time = chi + layers + {S} x sweep")}
    if gpu_fid_trace is not None:
        out["parity"] = {"oracle_fid_sweep0": f0, "oracle_fid_sweep1": f1,
                         "gpu_fid_sweep0": float(gpu_fid_trace[0]),
                         "gpu_fid_sweep1": float(gpu_fid_trace[1])}
    return out


# learned warm start: training pairs from the plain method on N=16, chi=32,
# 8 layers (416 states x 8 layers x 15 two-qubit slots = 49920 pairs)
WARM_TRAIN = {"num_qubits": 16, "bond_dimension": 32, "num_layers": 8, "num_sweeps": 30,
              "generator_bond": 48, "states": 416, "epochs": 20, "batch": 512, "lr": 1e-3,
              "width": 512}


def warm_start_leg(svs, plain_res, wl, args) -> dict:
    """Train the warm-start MLP on the GPU (untimed), then time the
    warm-started encode of the same batch with goal F_ref - 1e-4 per state,
    F_ref = the plain method's fidelity after all sweeps."""
    import torch
    from qmprs_amd import learned
    tw = WARM_TRAIN
    t0 = time.perf_counter()
    train_sv = _states(tw["num_qubits"], tw["generator_bond"], tw["states"], seed0=5000)
    data = learned.make_training_set(train_sv, tw["bond_dimension"], tw["num_layers"], tw["num_sweeps"])
    t1 = time.perf_counter()
    net = learned.WarmStartMLP(width=tw["width"], seed=0)
    hist = learned.train(net, data, epochs=tw["epochs"], batch=tw["batch"], lr=tw["lr"])
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    n, chi, L, S = wl["num_qubits"], wl["bond_dimension"], wl["num_layers"], wl["num_sweeps"]
    fref = plain_res.fidelity[:, -1].abs()
    goal = (fref - 1e-4).double().contiguous()
    sv_dev = torch.from_numpy(svs).cuda()
    for _ in range(args.warmup):
        res = learned.encode_warm(sv_dev, chi, L, S, net, goal=goal)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    for _ in range(args.steps):
        res = learned.encode_warm(sv_dev, chi, L, S, net, goal=goal)
    torch.cuda.synchronize()
    t4 = time.perf_counter()
    reached = (res.fidelity >= goal.float()).float()
    return {
        "value": round(len(svs) * args.steps / (t4 - t3), 3), "unit": "state-preps/s",
        "ms_per_step": round(1000.0 * (t4 - t3) / args.steps, 2),
        "goal": "F_ref - 1e-4 per state (F_ref = plain fidelity after 30 sweeps)",
        "goal_reached": float(reached.mean()),
        "mean_sweeps": float(res.sweeps.float().mean()),
        "proposals_accepted": float(res.accepted.float().mean()),
        "fidelity_mean": float(res.fidelity.mean()),
        "train": {"pairs": int(data.weights.sum().item()), "config": {k: tw[k] for k in (
            "num_qubits", "bond_dimension", "num_layers", "num_sweeps", "states", "epochs", "width")},
            "data_s": round(t1 - t0, 1), "train_s": round(t2 - t1, 1),
            "loss_first_epoch": hist[0], "loss_last_epoch": hist[-1]},
    }


def relaxed_leg(sv_dev, plain_res, wl, args) -> dict:
    """The same batch with the over-relaxed gate update (omega =
    ``--relax-omega``; an option with no reference counterpart): sweeps stop
    per state once its fidelity reaches F_ref - 1e-4, F_ref = the plain
    method's fidelity after all sweeps; at most the plain sweep count
    (``wl`` names the workload's sizes)."""
    import torch
    from qmprs_amd import _lib, engine
    n, chi, L, S = wl["num_qubits"], wl["bond_dimension"], wl["num_layers"], wl["num_sweeps"]
    fref = plain_res.fidelity[:, -1].abs()
    goal = (fref - 1e-4).double().contiguous()
    _lib.set_sweep_relaxation(args.relax_omega)
    try:
        for _ in range(args.warmup):
            res = engine.encode(sv_dev, chi, L, S, sweep_dtype=torch.complex64, goal=goal)
        torch.cuda.synchronize()
        # phase timers on, as in the plain leg's timed steps (same overhead)
        _lib.profile_enable(True)
        _lib.profile_read()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            res = engine.encode(sv_dev, chi, L, S, sweep_dtype=torch.complex64, goal=goal)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        prof = _lib.profile_read()
    finally:
        _lib.profile_enable(False)
        _lib.set_sweep_relaxation(1.0)
    fid = res.fidelity[:, -1].abs()
    return {
        "value": round(sv_dev.shape[0] * args.steps / (t1 - t0), 3), "unit": "state-preps/s",
        "ms_per_step": round(1000.0 * (t1 - t0) / args.steps, 2),
        "omega": args.relax_omega,
        "goal": f"F_ref - 1e-4 per state (F_ref = plain fidelity after {S} sweeps)",
        "goal_reached": float((fid >= goal).double().mean()),
        "mean_sweeps": float(res.sweeps_done.double().mean()),
        "max_sweeps": int(res.sweeps_done.max()),
        "fidelity_mean": float(fid.mean()),
        "phases_ms_per_step": {p: {"ms": round(v[0] / args.steps, 2), "launches": int(v[1] / args.steps)}
                               for p, v in prof.items()},
    }


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
        "dtype": "complex64 sweep storage, fp64 arithmetic and reductions",
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
        "dtype": "complex64 sweep storage, fp64 arithmetic and reductions",
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
        line["relaxed"] = relaxed_leg(sv, res, lg, args)
    print(json.dumps(line))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic-kib", type=float, nargs=2, default=None,
                    metavar=("FETCH", "WRITE"),
                    help="calibrated rocprofv3 FETCH_SIZE and WRITE_SIZE of one "
                         "window pass dispatch, KiB")
    ap.add_argument("--no-warm-start", action="store_true",
                    help="skip the learned warm-start leg (training on N=16 plain-method "
                         "pairs, then the N=20 warm-started encode against the plain "
                         "fidelity goal)")
    ap.add_argument("--relax-omega", type=float, default=1.7,
                    help="omega of the over-relaxed leg (the same batch with the gate update "
                         "polar(G' + (omega - 1)(G' - G)) and the plain fidelity goal); "
                         "1 skips the leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="before the run, copy a 1 GiB buffer 3 times (a plain copy "
                         "kernel of known size for calibrating HBM counters)")
    ap.add_argument("--workload", choices=("single-size", "ragged", "large"), default="single-size",
                    help="single-size: the N=20, chi=64 batch (default); ragged: 256 states of "
                         "mixed N and chi through the ragged scheduler; large: one N=26, chi=512 "
                         "state, 20 layers, 50 sweeps (use --steps 1 --warmup 0)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this encoder shards across GPUs by state "
              "batches, and no multi-GPU figure is produced here", file=sys.stderr)
        return 2

    import torch
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from qmprs_amd import _lib, engine
    if args.workload == "ragged":
        return ragged_main(args)
    if args.workload == "large":
        return large_main(args)

    wl = WORKLOAD
    n, chi, L, S = wl["num_qubits"], wl["bond_dimension"], wl["num_layers"], wl["num_sweeps"]
    B = args.batch
    svs = _states(n, wl["generator_bond"], B, seed0=1000)
    dev = _lib.require_gpu()
    if args.copy_calibration:
        src = torch.empty(1 << 30, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    sv_dev = engine.as_device_states(svs)
    ws = None

    def step():
        return engine.encode(sv_dev, chi, L, S, sweep_dtype=torch.complex64, workspace=ws)

    res = step()   # sizes the workspace
    cfg = _lib.EncodeConfig(n, B, chi, L, S, _lib.QMX_C64, _lib.QMX_C64, 1 - 1e-6, 1e-10)
    import ctypes
    ws = _lib.workspace(_lib.load().qmx_encode_workspace(ctypes.byref(cfg)), dev)
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
    ms_per_step = 1000.0 * total / args.steps
    value = B * args.steps / total

    fid = res.fidelity.abs().cpu().numpy()
    layers_used = res.layers_used.cpu().numpy()
    # dominant kernel: the windowed sweep's streaming pass.  The library runs
    # a batch as two half-batches whose passes alternate (never overlap), so
    # a launch covers B / jobs states.
    wp_ms, wp_launches = prof["window_pass"]
    avg_launch_s = wp_ms / 1000.0 / max(wp_launches, 1)
    jobs = 2 if B >= 2 else 1
    windows = max(1, round(wp_launches / (args.steps * S * jobs)))
    amp_bytes = 16.0 if windows == 1 else (32.0 * (windows - 2) + 24.0 + 16.0) / windows
    total_bytes = args.steps * S * windows * B * (2 ** n) * amp_bytes
    bytes_per_launch = total_bytes / max(wp_launches, 1)
    achieved = total_bytes / (wp_ms / 1000.0) / 1e9
    traffic = None
    kib = args.traffic_kib or MEASURED_TRAFFIC_KIB.get(B // jobs)
    if kib:
        traffic = int((kib[0] + kib[1]) * 1024)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "window_pass_kernel<float2, 4>",
                "avg_launch_us": round(avg_launch_s * 1e6, 2),
                "launches": int(wp_launches), "windows_per_sweep": windows,
                "states_per_launch": B // jobs,
                "bytes_per_launch": int(bytes_per_launch)}
    phases = {p: {"ms": round(v[0] / args.steps, 2), "launches": int(v[1] / args.steps)}
              for p, v in prof.items()}
    out = {
        "metric": "state-preps/sec at fidelity >=ref-1e-4 (N=20, chi=64): plain vs NN-warm-started",
        "value": round(value, 3),
        "unit": "state-preps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex64 sweep storage, fp64 arithmetic and reductions",
        "data": "synthetic: seeded random-MPS statevectors (bond 96, complex Gaussian)",
        "config": {"workload": "plain Sequential.prepare_state, N=20, bond_dimension=64, "
                               "num_layers=10, num_sweeps=30, complex64 input",
                   "batch": B, "path": "plain"},
        "fidelity": {"mean": float(fid[:, -1].mean()), "min": float(fid[:, -1].min()),
                     "max": float(fid[:, -1].max()), "layers_used_max": int(layers_used.max())},
        "phases_ms_per_step": phases,
        "roofline": roofline,
    }
    if not args.no_warm_start:
        out["warm_start"] = warm_start_leg(svs, res, wl, args)
    if args.relax_omega != 1.0:
        out["relaxed"] = relaxed_leg(sv_dev, res, wl, args)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(svs[0], wl, fid[0])
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
