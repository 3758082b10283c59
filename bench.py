"""Benchmark: 24-qubit complex64 random circuit on one MI355X.

One step = one pass of the 200-gate circuit (120 Haar 1-qubit gates, 20 CNOT,
20 CZ, 20 Haar 4x4, 20 full-length diagonal phases) over one 2^24-amplitude
complex64 state resident in HBM.  value = amplitude updates per second
(gates x 2^24 / step time).  Prints one JSON line.

The roofline object is for the dominant kernel (dense 1-qubit gates), timed
with events on the stream the kernels run on, inside the timed steps.  The
cpu_baseline leg times the torch CPU restatement of the reference's op
sequence (oracle/unitair_torch.py) on a bounded prefix of the same circuit.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (ROOT, ROOT / "qcware-unitair_amd"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--qubits", type=int, default=24)
    p.add_argument("--seed", type=int, default=2)
    p.add_argument("--cpu-seconds", type=float, default=15.0,
                   help="target CPU time of the cpu_baseline sample (0 = skip)")
    p.add_argument("--no-fidelity", action="store_true")
    p.add_argument("--no-out-of-place", action="store_true")
    return p.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_steps(runner, state, steps, warmup, with_events):
    """Wall time per step (synchronised on both sides) and, with_events, the
    per-gate kernel durations from events recorded on the launch stream."""
    import torch
    from unitair_amd._hip import EventPool
    for _ in range(warmup):
        runner.run(state)
    pools = [EventPool(len(runner.kinds)) for _ in range(steps)] if with_events else None
    if pools is None:
        pools = [None] * steps
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        if pools[i] is None:
            runner.run(state)
        else:
            runner.run(state, events=pools[i])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    per_kind = {}
    if with_events:
        for pool in pools:
            for kind, ms in zip(runner.kinds, pool.elapsed_ms()):
                per_kind.setdefault(kind, []).append(ms)
    return (t1 - t0) / steps, per_kind


def time_graph(runner, state, steps, warmup):
    import torch
    runner.capture(state)
    for _ in range(warmup):
        runner.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        runner.replay()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def cpu_baseline(circ, n, seed, budget_s):
    """Torch CPU restatement of the reference op sequence on a prefix of the
    circuit, complex64, all host cores torch is given."""
    import numpy as np
    import torch
    from oracle import unitair_torch as T
    from unitair_amd.initializations import seeded_state

    cores = len(os.sched_getaffinity(0))
    env_threads = os.environ.get("OMP_NUM_THREADS")
    threads = min(cores, int(env_threads)) if env_threads else cores
    torch.set_num_threads(threads)
    psi = seeded_state(n, (), seed=seed + 10, dtype=torch.complex64)
    mats = {}
    done, t_used = 0, 0.0
    t0 = time.perf_counter()
    for g in circ.gates:
        if g.kind == "phase":
            a = torch.from_numpy(np.asarray(g.angles, dtype=np.float32))
            ts = time.perf_counter()
            psi = T.apply_phase(a, psi)
        else:
            m = mats.get(id(g))
            if m is None:
                m = torch.from_numpy(g.dense_matrix()).to(torch.complex64)
            ts = time.perf_counter()
            psi = T.apply_operator(m, g.qubits, psi)
        t_used += time.perf_counter() - ts
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    amp = done * 2 ** n
    return {"value": amp / t_used, "unit": "amplitude-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"first {done} of {len(circ.gates)} gates of the same {n}-qubit "
                      f"complex64 circuit, torch {torch.__version__} CPU restatement of the "
                      f"reference op sequence (permute/contiguous/einsum/exp), "
                      f"{t_used:.1f} s"}


def fidelity(circ, n, seed):
    import numpy as np
    import torch
    from oracle import circuits as oc
    from oracle import unitair_np as U
    from unitair_amd.circuits import CircuitRunner
    from unitair_amd.initializations import seeded_state_np
    psi = seeded_state_np(n, (), seed=seed + 20)
    ref = oc.run_closed_form(circ.gates, psi)
    r = CircuitRunner(circ, dtype=torch.complex64, device="cuda", in_place=True)
    out = r.run(torch.from_numpy(psi).to("cuda", torch.complex64)).cpu().numpy()
    return float(U.fidelity_error(ref, out)), U.max_abs_error(ref, out)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this "
              f"build (the state-vector path is single-GPU, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from unitair_amd.circuits import CircuitRunner
    from unitair_amd.initializations import seeded_state
    from unitair_amd.workloads import random_circuit_24q

    n = args.qubits
    log(f"building {n}-qubit circuit")
    circ = random_circuit_24q(seed=args.seed, n=n)
    ngates = len(circ.gates)
    amps = 2 ** n
    dev = torch.device("cuda:0")
    state = seeded_state(n, (), seed=args.seed + 10, dtype=torch.complex64, device=dev)
    runner = CircuitRunner(circ, dtype=torch.complex64, device=dev, in_place=True)
    log(f"timing {args.steps} steps (warmup {args.warmup}), in place")
    sec, per_kind = time_steps(runner, state, args.steps, args.warmup, with_events=True)
    value = ngates * amps / sec
    log(f"in place: {sec * 1e3:.3f} ms/step, {value:.4e} amp-updates/s; per kernel avg us: "
        + ", ".join(f"{k}={1e3 * sum(v) / len(v):.1f}" for k, v in per_kind.items()))

    extra = {}
    if not args.no_out_of_place:
        log("timing out-of-place form")
        r2 = CircuitRunner(circ, dtype=torch.complex64, device=dev, in_place=False)
        sec2, _ = time_steps(r2, state, max(3, args.steps // 2), 1, with_events=False)
        extra["out_of_place_amp_updates_per_s"] = ngates * amps / sec2
        extra["out_of_place_ms_per_step"] = sec2 * 1e3
        log(f"out of place: {sec2 * 1e3:.3f} ms/step")
        del r2
    log("timing graph replay")
    r3 = CircuitRunner(circ, dtype=torch.complex64, device=dev, in_place=True)
    sec3 = time_graph(r3, state, args.steps, args.warmup)
    extra["graph_replay_amp_updates_per_s"] = ngates * amps / sec3
    extra["graph_replay_ms_per_step"] = sec3 * 1e3
    log(f"graph replay: {sec3 * 1e3:.3f} ms/step")
    del r3

    log("timing fused passes")
    from unitair_amd.fusion import FusedProgram
    prog = FusedProgram(circ, dtype=torch.complex64, device=dev)
    sec4, _ = time_steps(prog, state, args.steps, args.warmup, with_events=False)
    sec5 = time_graph(prog, state, args.steps, args.warmup)
    extra["fused"] = {"passes": prog.num_passes,
                      "ms_per_step": sec4 * 1e3,
                      "amp_updates_per_s": ngates * amps / sec4,
                      "graph_ms_per_step": sec5 * 1e3,
                      "graph_amp_updates_per_s": ngates * amps / sec5}
    log(f"fused: {prog.num_passes} passes, {sec4 * 1e3:.3f} ms/step eager, "
        f"{sec5 * 1e3:.3f} ms/step graph")
    del prog

    kernels = {k: {"launches_per_step": len(v) // args.steps,
                   "avg_us": 1e3 * float(np.mean(v))} for k, v in per_kind.items()}
    dom = "dense1q"
    dom_bytes = 2 * amps * 8            # one read + one write of the state, complex64
    avg_s = kernels[dom]["avg_us"] * 1e-6
    achieved = dom_bytes / avg_s / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                "kernel": "dense_fast_kernel<complex64,k=1>",
                "bytes_per_launch": dom_bytes}
    # whole-step algorithmic bandwidth (the metric's "HBM GB/s")
    step_bytes = 0
    for g in circ.gates:
        step_bytes += 2 * amps * 8 + (amps * 4 if g.kind == "phase" else 0)
    extra["step_algorithmic_GBps"] = step_bytes / sec / 1e9
    extra["step_frac_of_hbm_peak"] = extra["step_algorithmic_GBps"] / HBM_PEAK_GBS
    extra["kernels"] = kernels

    if not args.no_fidelity:
        log("fidelity vs complex128 oracle")
        fe, mae = fidelity(circ, n, args.seed)
        extra["fidelity_error"] = fe
        extra["max_abs_error"] = mae

    cpu = None
    if args.cpu_seconds > 0:
        log("cpu baseline")
        cpu = cpu_baseline(circ, n, args.seed, args.cpu_seconds)

    line = {
        "metric": "amplitude-updates/s (24q complex64 random circuit)",
        "value": value,
        "unit": "amplitude-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic (seeded Gaussian state, Haar gates, uniform phases)",
        "config": {"workload": f"random circuit, {n} qubits, {ngates} gates "
                               f"({circ.counts()}), one complex64 state",
                   "qubits": n, "gates": ngates, "batch": 1},
        "roofline": roofline,
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
