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
    p.add_argument("--no-adjoint", action="store_true")
    p.add_argument("--no-serving", action="store_true")
    p.add_argument("--no-fused-dense", action="store_true",
                   help="skip the 28-qubit dense 4/5-qubit block workload")
    p.add_argument("--no-graph", action="store_true", help="skip the graph-replay leg")
    p.add_argument("--no-ops", action="store_true", help="skip the per-call drop-in leg")
    p.add_argument("--no-stream-ref", action="store_true",
                   help="skip timing torch's own streaming kernels on the same state size")
    p.add_argument("--no-fused", action="store_true", help="skip the fused-pass leg")
    p.add_argument("--no-28q", action="store_true",
                   help="skip the 28-qubit (beyond the Infinity Cache) per-gate and fused leg")
    p.add_argument("--latency-runs", type=int, default=100,
                   help="single synchronised circuit runs for the p50/p99 latency")
    p.add_argument("--copy-calibration-mib", type=int, default=0,
                   help="run one float32 multiply-by-one of this many MiB (read + write) "
                        "before the timed steps, as the known-size kernel that scales "
                        "FETCH_SIZE / WRITE_SIZE counter readings")
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
    pools = [EventPool(len(runner.kinds), shared=True) for _ in range(steps)] if with_events else None
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


def step_bytes_of(circ, amps, es=8, angle_es=4):
    """Algorithmic bytes of one pass of the circuit: one read and one write of
    the state per dense gate and phase (plus the phase's angles); in-place CZ
    and CNOT count only the amplitudes they touch (a quarter and a half)."""
    total = 0
    for g in circ.gates:
        full = 2 * amps * es
        if g.kind == "cz":
            total += full // 4
        elif g.kind == "cnot":
            total += full // 2
        elif g.kind == "phase":
            total += full + amps * angle_es
        else:
            total += full
    return total


def fused_step_bytes(circ, amps, passes, es=8):
    """Algorithmic bytes of one fused step: a read and a write of the state per
    pass, plus one read of each phase gate's precomputed factors (es bytes per
    amplitude)."""
    return passes * 2 * amps * es + circ.counts().get("phase", 0) * amps * es


def step_latencies(runner, state, steps):
    """Latency of single circuit runs (synchronised on both sides):
    p50 / p99 / max in ms over `steps` runs."""
    import numpy as np
    import torch
    runner.run(state)
    torch.cuda.synchronize()
    lat = []
    for _ in range(steps):
        t0 = time.perf_counter()
        runner.run(state)
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - t0) * 1e3)
    lat = np.asarray(lat)
    return {"runs": steps, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()),
            "mean_ms": float(lat.mean())}


def device_angles(circ, device, seed):
    """The circuit with every phase gate's angles drawn on the device
    (uniform [0, 2pi) float32, the same distribution as the seeded host
    angles): large states then need no host generation or upload."""
    import torch
    from unitair_amd.circuits import Circuit, Gate
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    gates = []
    for g in circ.gates:
        if g.kind == "phase":
            a = torch.rand(2 ** circ.num_qubits, generator=gen, device=device,
                           dtype=torch.float32) * (2 * 3.141592653589793)
            gates.append(Gate("phase", (), angles=a))
        else:
            gates.append(g)
    return Circuit(circ.num_qubits, gates)


def per_gate_28q(dev, seed, n=28, steps=3, warmup=1):
    """The same circuit recipe on a 28-qubit state (2 GiB, larger than the
    256 MB Infinity Cache, so every pass goes to HBM): per-gate launches in
    place, and the fused passes."""
    import numpy as np
    import torch
    from unitair_amd.circuits import CircuitRunner
    from unitair_amd.fusion import FusedProgram
    from unitair_amd.initializations import rand_state
    from unitair_amd.workloads import random_circuit_24q
    circ = device_angles(random_circuit_24q(seed=seed, n=n), dev, seed + 100)
    amps = 2 ** n
    state = rand_state(n, device=dev, dtype=torch.complex64, seed=seed + 10)
    runner = CircuitRunner(circ, dtype=torch.complex64, device=dev, in_place=True)
    sec, per_kind = time_steps(runner, state, steps, warmup, with_events=True)
    nbytes = step_bytes_of(circ, amps)
    dense1 = 1e3 * float(np.mean(per_kind["dense1q"]))
    out = {"workload": f"random circuit, {n} qubits, {len(circ.gates)} gates "
                       f"({circ.counts()}), one complex64 state (2 GiB)",
           "ms_per_step": sec * 1e3, "amp_updates_per_s": len(circ.gates) * amps / sec,
           "step_algorithmic_bytes": nbytes, "step_algorithmic_GBps": nbytes / sec / 1e9,
           "step_frac_of_hbm_peak": nbytes / sec / 1e9 / HBM_PEAK_GBS,
           "dense1q_avg_us": dense1,
           "dense1q_frac_of_hbm_peak": 2 * amps * 8 / dense1 / 1e3 / HBM_PEAK_GBS}
    del runner
    prog = FusedProgram(circ, dtype=torch.complex64, device=dev)
    sec_f, _ = time_steps(prog, state, steps, warmup, with_events=False)
    fb = fused_step_bytes(circ, amps, prog.num_passes)
    out["fused"] = {"passes": prog.num_passes, "ms_per_step": sec_f * 1e3,
                    "amp_updates_per_s": len(circ.gates) * amps / sec_f,
                    "ms_per_pass": sec_f * 1e3 / prog.num_passes,
                    "pass_algorithmic_GBps": 2 * amps * 8 * prog.num_passes / sec_f / 1e9,
                    "step_algorithmic_bytes_with_phase_data": fb,
                    "step_algorithmic_GBps": fb / sec_f / 1e9}
    del prog
    t0 = time.perf_counter()
    prog = FusedProgram(circ, dtype=torch.complex64, device=dev, compiled=True)
    compile_s = time.perf_counter() - t0
    sec_c, _ = time_steps(prog, state, steps, warmup, with_events=False)
    out["fused_compiled"] = {"passes": prog.num_passes, "compiled_passes": prog.num_compiled,
                             "compile_s": compile_s, "ms_per_step": sec_c * 1e3,
                             "amp_updates_per_s": len(circ.gates) * amps / sec_c,
                             "ms_per_pass": sec_c * 1e3 / prog.num_passes,
                             "pass_algorithmic_GBps":
                                 2 * amps * 8 * prog.num_passes / sec_c / 1e9,
                             "step_algorithmic_bytes_with_phase_data":
                                 fused_step_bytes(circ, amps, prog.num_passes),
                             "step_algorithmic_GBps":
                                 fused_step_bytes(circ, amps, prog.num_passes) / sec_c / 1e9}
    del prog, state, circ
    torch.cuda.empty_cache()
    return out


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


def adjoint_bench(dev, iters=3):
    """16 qubits, 1024 states, 10 layers of RZ.RY on every qubit + CZ ring,
    loss sum_i <Z_i>: forward + adjoint backward (complex64), against the
    plain per-gate path (unitair_amd.simulation.apply_operator per gate with
    torch autograd, one saved state per gate) at a batch that fits memory."""
    import numpy as np
    import torch
    from unitair_amd import adjoint
    from unitair_amd import simulation as sim
    n, L, B = 16, 10, 1024
    rng = np.random.default_rng(3)
    th = torch.from_numpy(rng.uniform(0, 2 * np.pi, (B, L, n))).to(dev, torch.float32)
    ph = torch.from_numpy(rng.uniform(0, 2 * np.pi, (B, L, n))).to(dev, torch.float32)

    def adj_step(t, p):
        t = t.detach().requires_grad_(True)
        p = p.detach().requires_grad_(True)
        adjoint.sum_z_loss(t, p, n, dtype=torch.complex64).sum().backward()
        return t.grad, p.grad

    for _ in range(2):          # the first calls also plan, allocate and ramp clocks
        adj_step(th, ph)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(iters):
        adj_step(th, ph)
    torch.cuda.synchronize()
    adj_s = (time.perf_counter() - t0) / iters
    moved = adjoint.last_traffic["forward_bytes"] + adjoint.last_traffic["backward_bytes"]

    def adj_step_compiled(t, p):
        t = t.detach().requires_grad_(True)
        p = p.detach().requires_grad_(True)
        adjoint.sum_z_loss(t, p, n, dtype=torch.complex64, compiled=True).sum().backward()
        return t.grad, p.grad

    t0 = time.perf_counter()
    g_ref = adj_step(th, ph)
    g_c = adj_step_compiled(th, ph)         # the first call compiles the passes
    torch.cuda.synchronize()
    first_compiled_s = time.perf_counter() - t0
    adj_step_compiled(th, ph)
    compiled_vs_interp = max(float((a - b).abs().max()) for a, b in zip(g_ref, g_c))
    t0 = time.perf_counter()
    for _ in range(iters):
        adj_step_compiled(th, ph)
    torch.cuda.synchronize()
    adj_c_s = (time.perf_counter() - t0) / iters

    Bp = 64
    pop = torch.tensor([bin(v).count("1") for v in range(2 ** n)], device=dev,
                       dtype=torch.float32)
    diag = n - 2 * pop
    cz = torch.diag(torch.tensor([1, 1, 1, -1], dtype=torch.complex64, device=dev))

    def plain_step(t, p):
        t = t.detach().requires_grad_(True)
        p = p.detach().requires_grad_(True)
        psi = torch.zeros(Bp, 2 ** n, dtype=torch.complex64, device=dev)
        psi[:, 0] = 1
        for l in range(L):
            for q in range(n):
                e = torch.polar(torch.ones_like(p[:, l, q]), -p[:, l, q] / 2)
                z = torch.zeros_like(e)
                rz = torch.stack([torch.stack([e, z], -1), torch.stack([z, e.conj()], -1)], -2)
                psi = sim.apply_operator(rz, [q], psi)
            for q in range(n):
                c, s_ = torch.cos(t[:, l, q] / 2), torch.sin(t[:, l, q] / 2)
                ry = torch.stack([torch.stack([c, -s_], -1), torch.stack([s_, c], -1)], -2)
                psi = sim.apply_operator(ry.to(torch.complex64), [q], psi)
            for q in range(n):
                psi = sim.apply_operator(cz, [q, (q + 1) % n], psi)
        ((psi.abs() ** 2) * diag).sum().backward()
        return t.grad, p.grad

    plain_step(th[:Bp], ph[:Bp])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain_step(th[:Bp], ph[:Bp])
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    return {"workload": f"{n} qubits, {L} layers (RZ.RY on every qubit + CZ ring), loss "
                        f"sum_i <Z_i>, complex64, forward + gradient of all {2 * L * n} "
                        f"angles per state",
            "adjoint_batch": B, "adjoint_ms": adj_s * 1e3,
            "adjoint_states_per_s": B / adj_s,
            "adjoint_algorithmic_bytes": moved,
            "adjoint_algorithmic_GBps": moved / adj_s / 1e9,
            "adjoint_frac_of_hbm_peak": moved / adj_s / 1e9 / HBM_PEAK_GBS,
            "compiled_passes": {"adjoint_ms": adj_c_s * 1e3,
                                "adjoint_states_per_s": B / adj_c_s,
                                "adjoint_algorithmic_GBps": moved / adj_c_s / 1e9,
                                "first_call_with_compile_s": first_compiled_s,
                                "max_grad_difference_vs_interpreted": compiled_vs_interp},
            "plain_per_gate_autograd_batch": Bp, "plain_per_gate_autograd_ms": plain_s * 1e3,
            "plain_per_gate_autograd_states_per_s": Bp / plain_s,
            "speedup_states_per_s": (B / adj_s) / (Bp / plain_s)}


def serving_bench(dev, num=200, seed=11, cap_sample=40):
    """Poisson arrivals of random layered circuits (n log-uniform in [10, 26],
    depth uniform in [10, 200], complex64), `num` requests.  The arrival rate
    is 0.7 x the one-at-a-time per-gate capacity (measured on the first
    `cap_sample` requests of the same stream).  The continuous-batching server
    and two one-at-a-time baselines -- per-gate launches, and the server's own
    fused programs with one request on the GPU -- see identical arrivals."""
    from unitair_amd.serving import (Server, poisson_workload, run_one_at_a_time,
                                     run_one_at_a_time_fused, summarize)
    warm = lambda: poisson_workload(4, rate=0.0, seed=seed + 1, n_range=(10, 14),
                                    depth_range=(10, 20))
    Server(device=dev).run(warm(), realtime=False)
    run_one_at_a_time(warm(), device=dev, realtime=False)
    run_one_at_a_time_fused(warm(), device=dev, realtime=False)
    cap_reqs = poisson_workload(num, rate=0.0, seed=seed)[:cap_sample]
    run_one_at_a_time(cap_reqs, device=dev, realtime=False)
    capacity = len(cap_reqs) / max(r.done for r in cap_reqs)
    sat = poisson_workload(num, rate=0.0, seed=seed)
    Server(device=dev).run(sat, realtime=False)
    capacity_cb = num / max(r.done for r in sat)
    sat1 = poisson_workload(num, rate=0.0, seed=seed)
    run_one_at_a_time_fused(sat1, device=dev, realtime=False)
    capacity_f1 = num / max(r.done for r in sat1)
    rate = 0.7 * capacity
    log(f"serving: per-gate capacity {capacity:.1f}/s, arrivals at {rate:.1f}/s")
    served = Server(device=dev).run(poisson_workload(num, rate=rate, seed=seed), realtime=True)
    fused1 = run_one_at_a_time_fused(poisson_workload(num, rate=rate, seed=seed), device=dev,
                                     realtime=True)
    plain = run_one_at_a_time(poisson_workload(num, rate=rate, seed=seed), device=dev,
                              realtime=True)
    # heavier load: 0.9 x the capacity of one request at a time on fused programs
    rate_hi = 0.9 * capacity_f1
    log(f"serving: fused one-at-a-time capacity {capacity_f1:.1f}/s, arrivals at {rate_hi:.1f}/s")
    served_hi = Server(device=dev).run(poisson_workload(num, rate=rate_hi, seed=seed),
                                       realtime=True)
    fused1_hi = run_one_at_a_time_fused(poisson_workload(num, rate=rate_hi, seed=seed),
                                        device=dev, realtime=True)
    same_fused = max(max(abs(a.z0 - b.z0), abs(a.norm2 - b.norm2))
                     for a, b in zip(served, fused1))
    same_plain = max(max(abs(a.z0 - b.z0), abs(a.norm2 - b.norm2))
                     for a, b in zip(served, plain))
    return {"workload": f"{num} requests, n log-uniform [10,26], depth U[10,200] layers of "
                        f"Haar 1q on every qubit + CZ on n//2 adjacent pairs, complex64",
            "one_at_a_time_capacity_circuits_per_s": capacity,
            "capacity_sample_requests": len(cap_reqs),
            "continuous_batching_capacity_circuits_per_s": capacity_cb,
            "one_at_a_time_fused_capacity_circuits_per_s": capacity_f1,
            "capacity_ratio": capacity_cb / capacity,
            "capacity_ratio_vs_fused_one_at_a_time": capacity_cb / capacity_f1,
            "arrival_rate_per_s": rate,
            "continuous_batching": summarize(served),
            "one_at_a_time_fused_programs": summarize(fused1),
            "one_at_a_time_per_gate": summarize(plain),
            "heavy_load": {"arrival_rate_per_s": rate_hi,
                           "continuous_batching": summarize(served_hi),
                           "one_at_a_time_fused_programs": summarize(fused1_hi)},
            "max_difference_vs_fused_alone": same_fused,
            "max_difference_vs_per_gate": same_plain}


MFMA_PEAK_TFLOPS = {"c64": 157.3, "c128": 78.6}   # dense FP32 / FP64 matrix-core peaks


def fused_dense_bench(dev, steps=3, warmup=1, seed=4, n=28):
    """One n-qubit state; 64 Haar 32x32 blocks on random 5-qubit subsets
    (including qubits 0-4 and n-5..n-1) and 16 Haar 16x16 blocks, in place,
    complex64 (FP32 matrix cores) and complex128 (FP64 matrix cores).  One
    step = the 80 gates."""
    import numpy as np
    import torch
    from unitair_amd.circuits import CircuitRunner
    from unitair_amd.initializations import rand_state
    from unitair_amd.workloads import fused_dense_circuit
    circ = fused_dense_circuit(seed=seed, n=n)
    amps = 2 ** n
    flops = sum(8 * 2 ** len(g.qubits) * amps for g in circ.gates)
    out = {"workload": f"{n} qubits, {circ.counts()}: 64 Haar 32x32 (5-qubit) + 16 Haar 16x16 "
                       f"(4-qubit) blocks on random subsets, one state, in place",
           "flops_per_step": flops}
    base = rand_state(n, device=dev, dtype=torch.complex128, seed=seed + 10)
    for dt, name in ((torch.complex64, "c64"), (torch.complex128, "c128")):
        state = base.to(dt) if dt != torch.complex128 else base
        es = state.element_size()
        peak = MFMA_PEAK_TFLOPS[name]

        def leg(runner, label):
            sec, per_kind = time_steps(runner, state, steps, warmup, with_events=True)
            launches = len(runner.kinds)
            step_bytes = launches * 2 * amps * es
            kernels = {}
            for kind, v in per_kind.items():
                ks = [int(c) for c in kind[len("blocks"):]] if kind.startswith("blocks") \
                    else [int(kind[len("dense"):-1])]
                us = 1e3 * float(np.mean(v))
                kflops = sum(8 * 2 ** k * amps for k in ks)
                kernels[kind] = {"launches_per_step": len(v) // steps, "avg_us": us,
                                 "algorithmic_GBps": 2 * amps * es / us / 1e3,
                                 "TFLOPs": kflops / us / 1e6,
                                 "mfma_frac": kflops / us / 1e6 / peak}
            res = {"ms_per_step": sec * 1e3, "passes_over_state": launches,
                   "amp_updates_per_s": len(circ.gates) * amps / sec,
                   "algorithmic_GBps": step_bytes / sec / 1e9,
                   "hbm_frac": step_bytes / sec / 1e9 / HBM_PEAK_GBS,
                   "TFLOPs": flops / sec / 1e12,
                   "mfma_frac": flops / sec / 1e12 / peak,
                   "roofline_ms": 1e3 * max(step_bytes / (HBM_PEAK_GBS * 1e9),
                                            flops / (peak * 1e12)),
                   "kernels": kernels}
            log(f"fused dense {name} {label}: {sec * 1e3:.2f} ms/step, {launches} passes, "
                f"{res['algorithmic_GBps']:.0f} GB/s, {res['TFLOPs']:.1f} TFLOP/s")
            return res

        blocks = CircuitRunner(circ, dtype=dt, device=dev, in_place=True, dense_blocks=True)
        res = leg(blocks, "block passes")
        del blocks
        per_gate = CircuitRunner(circ, dtype=dt, device=dev, in_place=True)
        res["per_gate"] = leg(per_gate, "one gate per pass")
        out[name] = res
        del per_gate, state
        torch.cuda.empty_cache()
    del base
    torch.cuda.empty_cache()
    return out


def fused_dense_slice_fidelity(dev, seed=4, slice_qubits=20):
    """The dense-block builder at slice_qubits qubits, run on the GPU in
    complex64 and complex128, against the complex128 oracle."""
    import torch
    from oracle import circuits as oc
    from oracle import unitair_np as U
    from unitair_amd.circuits import CircuitRunner
    from unitair_amd.initializations import seeded_state_np
    from unitair_amd.workloads import fused_dense_circuit
    sl = fused_dense_circuit(seed=seed, n=slice_qubits)
    psi = seeded_state_np(slice_qubits, (), seed=seed + 20)
    ref = oc.run_closed_form(sl.gates, psi)
    out = {"slice": f"same builder at {slice_qubits} qubits ({len(sl.gates)} gates)"}
    for dt, name in ((torch.complex64, "c64"), (torch.complex128, "c128")):
        for label, blocks in (("block_passes", True), ("per_gate", False)):
            r = CircuitRunner(sl, dtype=dt, device=dev, in_place=True, dense_blocks=blocks)
            got = r.run(torch.from_numpy(psi).to(dev, dt)).cpu().numpy()
            out.setdefault(name, {})[label] = {
                "fidelity_error": float(U.fidelity_error(ref, got)),
                "max_abs_error": float(U.max_abs_error(ref, got))}
    return out


def _ops_inputs(n, seed, device, dtype):
    """One n-qubit state, Haar gates and full-length float32 angles / diag."""
    import numpy as np
    import torch
    from unitair_amd.gates import haar_unitary
    from unitair_amd.initializations import seeded_state
    rng = np.random.default_rng(seed)
    psi = seeded_state(n, (), seed=seed, dtype=dtype, device=device)
    other = seeded_state(n, (), seed=seed + 1, dtype=dtype, device=device)
    u = {k: torch.from_numpy(haar_unitary(2 ** k, rng)).to(device, dtype)
         for k in (1, 2, 5, 6, 7, 8)}
    ang = torch.from_numpy(rng.uniform(0, 2 * np.pi, 2 ** n).astype(np.float32)).to(device)
    diag = torch.from_numpy(rng.standard_normal(2 ** n).astype(np.float32)).to(device)
    return psi, other, u, ang, diag


def _grad_call(apply_operator, u, psi):
    """Forward + backward of one 1-qubit apply_operator with respect to the
    state and the operator (the drop-in's autograd)."""
    import torch
    x = psi.detach().requires_grad_(True)
    op = u.detach().requires_grad_(True)
    y = apply_operator(op, [3], x)
    y.real.sum().backward()
    return x.grad


def _op_table(n, sim, states, psi, other, u, ang, diag, sample, rand_state):
    """(name, callable, algorithmic bytes) for each drop-in call, complex64."""
    import torch
    A, s = 2 ** n, 8
    rev = list(range(n))[::-1]
    ring = torch.tensor([[q, (q + 1) % n] for q in range(n)], device=psi.device)
    return [
        ("apply_operator 1q, qubit 0 (stride 2^(n-1))",
         lambda: sim.apply_operator(u[1], [0], psi), 2 * A * s),
        (f"apply_operator 1q, qubit {n - 1} (stride 1)",
         lambda: sim.apply_operator(u[1], [n - 1], psi), 2 * A * s),
        ("apply_operator 2q Haar, qubits (5, 17)",
         lambda: sim.apply_operator(u[2], [5, 17], psi), 2 * A * s),
        ("apply_operator 5q Haar, qubits (0..4)",
         lambda: sim.apply_operator(u[5], [0, 1, 2, 3, 4], psi), 2 * A * s),
        ("apply_operator 6q Haar, qubits (0, 7, 12, 3, 23, 18) (matrix cores, one-gate pass)",
         lambda: sim.apply_operator(u[6], [0, 7, 12, 3, 23, 18], psi), 2 * A * s),
        ("apply_operator 7q Haar, qubits (2, 9, 14, 0, 21, 5, 17) (matrix cores, 2 row slices)",
         lambda: sim.apply_operator(u[7], [2, 9, 14, 0, 21, 5, 17], psi), 2 * A * s),
        ("apply_operator 8q Haar, qubits (2, 9, 14, 0, 21, 5, 17, 11) (matrix cores, 8 row slices)",
         lambda: sim.apply_operator(u[8], [2, 9, 14, 0, 21, 5, 17, 11], psi), 2 * A * s),
        ("act_first_qubits 2q Haar",
         lambda: sim.act_first_qubits(u[2], psi), 2 * A * s),
        ("apply_phase, full-length float32 angles",
         lambda: sim.apply_phase(ang, psi), 2 * A * s + A * 4),
        ("apply_all_qubits Haar 2x2",
         lambda: sim.apply_all_qubits(u[1], psi), 2 * A * s),
        ("apply_to_qubits, 8 Haar 1q gates on 6 qubits (fused per qubit)",
         lambda: sim.apply_to_qubits([u[1]] * 8, [0, 5, 11, 12, 17, n - 1, 0, 5], psi),
         2 * A * s),
        ("swap qubits (0, n-1)", lambda: sim.swap(psi, (0, n - 1)), 2 * A * s),
        ("permute_qubits, reversed order", lambda: sim.permute_qubits(rev, psi), 2 * A * s),
        ("multi_cz, ring of n pairs", lambda: sim.multi_cz(ring, psi), 2 * A * s),
        ("measure: draw 2^20 samples (probabilities, prefix sums, draws)",
         lambda: sample(psi, 1 << 20), A * s + A * 8),
        ("rand_state (Gaussian, normalised per state)", lambda: rand_state(n), A * s),
        ("apply_operator 1q forward + backward (autograd, state and operator)",
         lambda: _grad_call(sim.apply_operator, u[1], psi), 6 * A * s),
        ("norm_squared", lambda: states.norm_squared(psi), A * s),
        ("abs_squared", lambda: states.abs_squared(psi), A * s + A * 4),
        ("inner_product", lambda: states.inner_product(psi, other), 2 * A * s),
        ("diag_expectation_value, float32 diag",
         lambda: states.diag_expectation_value(diag, psi), A * s + A * 4),
    ]


def ops_bench(dev, n=24, seed=5, reps=20):
    """Each drop-in call on one n-qubit complex64 state, out of place as the
    reference's contract requires: time per call (events around `reps`
    back-to-back calls on the current stream) and algorithmic GB/s."""
    import torch
    import unitair_amd as ua
    from unitair_amd import simulation as sim
    psi, other, u, ang, diag = _ops_inputs(n, seed, dev, torch.complex64)
    out = {}
    from unitair_amd.measurement import sample_indices
    from unitair_amd.initializations import rand_state
    for name, fn, nbytes in _op_table(n, sim, ua.states, psi, other, u, ang, diag,
                                      sample_indices,
                                      lambda k: rand_state(k, device=dev)):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) / reps * 1e3
        out[name] = {"us_per_call": us, "algorithmic_bytes": nbytes,
                     "algorithmic_GBps": nbytes / us / 1e3,
                     "frac_of_hbm_peak": nbytes / us / 1e3 / HBM_PEAK_GBS}
    return out


def cpu_ops_baseline(n=24, seed=5, budget_s=10.0):
    """The same calls through the torch CPU restatement of the reference
    (one call each, in table order, until budget_s is spent)."""
    import torch
    from oracle import unitair_torch as T
    psi, other, u, ang, diag = _ops_inputs(n, seed, "cpu", torch.complex64)

    class _Sim:
        apply_operator = staticmethod(lambda op, q, st: T.apply_operator(op, q, st))
        act_first_qubits = staticmethod(
            lambda op, st: T.apply_operator(op, list(range(op.shape[-1].bit_length() - 1)), st))
        apply_phase = staticmethod(T.apply_phase)
        apply_all_qubits = staticmethod(T.apply_all_qubits)
        apply_to_qubits = staticmethod(T.apply_to_qubits)
        swap = staticmethod(T.swap)
        permute_qubits = staticmethod(T.permute_qubits)
        multi_cz = staticmethod(T.multi_cz)

    class _States:
        norm_squared = staticmethod(T.norm_squared)
        abs_squared = staticmethod(T.abs_squared)
        inner_product = staticmethod(T.inner_product)
        diag_expectation_value = staticmethod(T.diag_expectation_value)

    out = {}
    t_start = time.perf_counter()
    def cpu_rand_state(k):
        x = torch.randn(2 ** k, dtype=torch.complex64)
        return x / x.abs().square().sum().sqrt()

    for name, fn, nbytes in _op_table(n, _Sim, _States, psi, other, u, ang, diag,
                                      T.sample_indices, cpu_rand_state):
        t0 = time.perf_counter()
        fn()
        us = (time.perf_counter() - t0) * 1e6
        out[name] = {"us_per_call": us, "algorithmic_GBps": nbytes / us / 1e3}
        if time.perf_counter() - t_start > budget_s:
            break
    return out


def stream_reference(dev, n, reps=50):
    """What plain streaming reaches on this GPU at the state's size: torch's
    in-place elementwise multiply and its device-to-device copy of one
    n-qubit complex64 state (read + write of 2^n x 8 B each)."""
    import torch
    x = torch.ones(2 ** n, dtype=torch.complex64, device=dev)
    y = torch.empty_like(x)
    nbytes = 2 * x.numel() * 8
    out = {}
    for name, fn in (("torch_inplace_mul", lambda: x.mul_(1.0)), ("torch_copy", lambda: y.copy_(x))):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) / reps * 1e3
        out[name] = {"us": us, "GBps": nbytes / us / 1e3}
    return out


def measured_traffic():
    """HBM-side bytes per launch of the 1-qubit dense kernels at 24 qubits,
    from the FETCH_SIZE / WRITE_SIZE counter passes (scaled by a copy kernel
    of known size) summarised in profiles/traffic_24q_circuit.md."""
    path = ROOT / "profiles" / "traffic_24q_circuit.md"
    try:
        for line in path.read_text().splitlines():
            if line.startswith("- traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


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
              f"build: it measures the state-vector kernels on one GPU", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from unitair_amd.circuits import CircuitRunner
    from unitair_amd.initializations import rand_state
    from unitair_amd.workloads import random_circuit_24q

    n = args.qubits
    log(f"building {n}-qubit circuit")
    circ = random_circuit_24q(seed=args.seed, n=n)
    ngates = len(circ.gates)
    amps = 2 ** n
    dev = torch.device("cuda:0")
    state = rand_state(n, device=dev, dtype=torch.complex64, seed=args.seed + 10)
    runner = CircuitRunner(circ, dtype=torch.complex64, device=dev, in_place=True)
    if args.copy_calibration_mib > 0:
        numel = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(numel, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
        log(f"copy calibration: {args.copy_calibration_mib} MiB read + written")
    log(f"timing {args.steps} steps (warmup {args.warmup}), in place")
    sec, per_kind = time_steps(runner, state, args.steps, args.warmup, with_events=True)
    value = ngates * amps / sec
    log(f"in place: {sec * 1e3:.3f} ms/step, {value:.4e} amp-updates/s; per kernel avg us: "
        + ", ".join(f"{k}={1e3 * sum(v) / len(v):.1f}" for k, v in per_kind.items()))

    extra = {}
    log(f"latency of single runs ({args.latency_runs} runs)")
    extra["latency_per_gate"] = step_latencies(runner, state, args.latency_runs)
    if not args.no_out_of_place:
        log("timing out-of-place form")
        r2 = CircuitRunner(circ, dtype=torch.complex64, device=dev, in_place=False)
        sec2, _ = time_steps(r2, state, max(3, args.steps // 2), 1, with_events=False)
        extra["out_of_place_amp_updates_per_s"] = ngates * amps / sec2
        extra["out_of_place_ms_per_step"] = sec2 * 1e3
        log(f"out of place: {sec2 * 1e3:.3f} ms/step")
        del r2
    if not args.no_graph:
        log("timing graph replay")
        r3 = CircuitRunner(circ, dtype=torch.complex64, device=dev, in_place=True)
        sec3 = time_graph(r3, state, args.steps, args.warmup)
        extra["graph_replay_amp_updates_per_s"] = ngates * amps / sec3
        extra["graph_replay_ms_per_step"] = sec3 * 1e3
        log(f"graph replay: {sec3 * 1e3:.3f} ms/step")
        del r3

    if not args.no_fused:
        log("timing fused passes")
        from unitair_amd.fusion import FusedProgram
        prog = FusedProgram(circ, dtype=torch.complex64, device=dev)
        sec4, _ = time_steps(prog, state, args.steps, args.warmup, with_events=False)
        sec5 = time_graph(prog, state, args.steps, args.warmup)
        extra["fused"] = {"passes": prog.num_passes,
                          "ms_per_step": sec4 * 1e3,
                          "step_algorithmic_GBps":
                              fused_step_bytes(circ, amps, prog.num_passes) / sec4 / 1e9,
                          "amp_updates_per_s": ngates * amps / sec4,
                          "graph_ms_per_step": sec5 * 1e3,
                          "graph_amp_updates_per_s": ngates * amps / sec5,
                          "ms_per_pass": sec4 * 1e3 / prog.num_passes,
                          "latency": step_latencies(prog, state, args.latency_runs)}
        log(f"fused: {prog.num_passes} passes, {sec4 * 1e3:.3f} ms/step eager, "
            f"{sec5 * 1e3:.3f} ms/step graph")
        del prog
        log("timing compiled fused passes")
        t0 = time.perf_counter()
        prog = FusedProgram(circ, dtype=torch.complex64, device=dev, compiled=True)
        compile_s = time.perf_counter() - t0
        sec6, _ = time_steps(prog, state, args.steps, args.warmup, with_events=False)
        sec7 = time_graph(prog, state, args.steps, args.warmup)
        extra["fused_compiled"] = {
            "passes": prog.num_passes, "compiled_passes": prog.num_compiled,
            "compile_s": compile_s, "ms_per_step": sec6 * 1e3,
            "amp_updates_per_s": ngates * amps / sec6,
            "step_algorithmic_GBps": fused_step_bytes(circ, amps, prog.num_passes) / sec6 / 1e9,
            "graph_ms_per_step": sec7 * 1e3,
            "graph_amp_updates_per_s": ngates * amps / sec7,
            "ms_per_pass": sec6 * 1e3 / prog.num_passes,
            "latency": step_latencies(prog, state, args.latency_runs)}
        log(f"compiled fused: {prog.num_passes} passes (compiled in {compile_s:.1f} s), "
            f"{sec6 * 1e3:.3f} ms/step eager, {sec7 * 1e3:.3f} ms/step graph")
        del prog

    kernels = {k: {"launches_per_step": len(v) // args.steps,
                   "avg_us": 1e3 * float(np.mean(v))} for k, v in per_kind.items()}
    dom = "dense1q"
    dom_bytes = 2 * amps * 8            # one read + one write of the state, complex64
    avg_s = kernels[dom]["avg_us"] * 1e-6
    achieved = dom_bytes / avg_s / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "traffic": measured_traffic() if n == 24 else None,
                "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of this "
                                  "command, scaled by a copy kernel of known size "
                                  "(profiles/traffic_24q_circuit.md); not measured in this run",
                "kernel": "dense_fast_kernel<complex64,k=1>",
                "bytes_per_launch": dom_bytes}
    # whole-step algorithmic bandwidth (the metric's "HBM GB/s")
    step_bytes = step_bytes_of(circ, amps)
    extra["step_algorithmic_bytes"] = step_bytes
    extra["step_algorithmic_GBps"] = step_bytes / sec / 1e9
    extra["step_frac_of_hbm_peak"] = extra["step_algorithmic_GBps"] / HBM_PEAK_GBS
    extra["kernels"] = kernels

    if not args.no_adjoint:
        log("adjoint gradient workload")
        extra["adjoint"] = adjoint_bench(dev)
        log(f"adjoint: {extra['adjoint']['adjoint_ms']:.1f} ms for 1024 states (compiled "
            f"passes {extra['adjoint']['compiled_passes']['adjoint_ms']:.1f} ms); plain "
            f"per-gate autograd {extra['adjoint']['plain_per_gate_autograd_ms']:.1f} ms for 64")

    if not args.no_serving:
        log("serving workload")
        extra["serving"] = serving_bench(dev)
        sv = extra["serving"]
        for tag, blk in (("", sv), ("heavy load ", sv["heavy_load"])):
            for key in ("continuous_batching", "one_at_a_time_fused_programs",
                        "one_at_a_time_per_gate"):
                if key in blk:
                    log(f"serving {tag}{key}: p50 {blk[key]['p50_ms']:.1f} ms p99 "
                        f"{blk[key]['p99_ms']:.1f} ms, {blk[key]['circuits_per_s']:.1f} circ/s")

    if not args.no_stream_ref:
        extra["stream_reference"] = stream_reference(dev, n)
        log("stream reference: " + ", ".join(f"{k} {v['GBps']:.0f} GB/s"
                                              for k, v in extra["stream_reference"].items()))

    if not args.no_28q and n < 28:
        log("28-qubit leg (per-gate and fused)")
        extra["circuit_28q"] = per_gate_28q(dev, args.seed)
        c28 = extra["circuit_28q"]
        log(f"28q: per-gate {c28['ms_per_step']:.1f} ms/step "
            f"({c28['step_frac_of_hbm_peak']:.3f} of HBM peak), dense1q "
            f"{c28['dense1q_avg_us']:.0f} us; fused {c28['fused']['passes']} passes "
            f"{c28['fused']['ms_per_step']:.2f} ms/step, compiled "
            f"{c28['fused_compiled']['ms_per_step']:.2f} ms/step")

    if not args.no_fused_dense:
        log("fused dense 4/5-qubit blocks, 28 qubits")
        extra["fused_dense_blocks"] = fused_dense_bench(dev)

    if not args.no_ops:
        log(f"drop-in calls, {n} qubits")
        extra["ops"] = ops_bench(dev, n=n)

    # ---- CPU leg: the oracle as the timed baseline and as the checker of
    # the GPU results (never as the thing measured on the GPU)
    if not args.no_fidelity:
        log("fidelity vs complex128 oracle")
        fe, mae = fidelity(circ, n, args.seed)
        extra["fidelity_error"] = fe
        extra["max_abs_error"] = mae
        if not args.no_fused_dense:
            extra["fused_dense_blocks"]["slice_check"] = fused_dense_slice_fidelity(dev)

    cpu = None
    if args.cpu_seconds > 0:
        log("cpu baseline")
        cpu = cpu_baseline(circ, n, args.seed, args.cpu_seconds)
        if not args.no_ops:
            cpu["ops"] = cpu_ops_baseline(n=n, budget_s=args.cpu_seconds)

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
        "data": "synthetic (seeded Gaussian state made on the device, Haar gates, uniform phases)",
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
