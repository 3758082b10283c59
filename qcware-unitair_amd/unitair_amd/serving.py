"""Continuous-batching request loop for circuits of uneven size and depth.

A request is a seeded random circuit on n qubits (complex64): `depth`
layers, each a Haar-random 1-qubit gate on every qubit followed by CZ on
floor(n/2) random adjacent pairs; the answer is ||psi||^2 and <Z_0> in
float64.  (unitair itself has no serving layer: a user loops over
apply_operator calls; README.rst:11-16.)

Design (one GPU):
  * Admission by footprint.  Same-n requests share a lane: one HBM buffer of
    `capacity` state slots (capacity x 2^n x 8 B).  A request is admitted
    when its lane has a free slot, or when a new lane fits the HBM budget;
    otherwise it waits in the queue.  Lanes are released when they empty.
  * Admission by work, in arrival order.  The passes are bound by the vector
    ALUs and HBM per amplitude, so running requests side by side saves
    launches (small n) but not arithmetic: past a few million amplitudes in
    flight, more requests only stretch every one of them (measured on 200
    mixed requests near saturation: p99 607 ms without a cap, 456 ms with
    2^22 or 2^24 amplitudes; the default 2^22 lets requests of 22 or more
    qubits run alone).  A request is
    admitted while the amplitudes of the active requests stay within
    `max_active_amplitudes` (the first one always is), and a request that
    does not fit holds back the ones behind it (no overtaking, so large
    requests are not starved).
  * Continuous batching.  Each lane step advances the lane by
    K = min(max_layers_per_step, layers left of its least advanced request)
    layers, with one batched fused-pass program over the lane's occupied
    slot range: per-slot gate data (each slot is at its own layer), so
    requests that arrive later join the lane's next step.  Empty slots inside
    the range run identity gates.
  * Asynchronous.  Each lane runs on its own HIP stream with up to two steps
    in flight (double-buffered gate staging).  The host never synchronises
    the device: a step ends with an event, and the scheduler polls the
    events, harvests finished requests (their reductions were queued right
    after their last step), admits waiting requests into freed slots and
    issues further steps while other lanes keep the GPU busy.  A request's
    completion time is its step's event time.
  * Graph replay.  Each (lane, K, occupied range, staging buffer) program is
    recorded into a HIP graph by stream capture on first use and replayed.
Each request's arithmetic is independent of its slot and of the other
requests, so its result is the same alone and in the mix.
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np
import torch

from . import _lib
from ._device import on_device
from ._hip import StreamGraph
from .fusion import TILE_BITS, FusedProgram, _Item, compile_in_background
from .gates import haar_unitary

BYTES_C64 = 8
# (qubit count, layers per step) -> future of the background compile of that
# lane program's passes (their structure depends on nothing else); shared by
# every lane and server of the process
_BACKGROUND_JOBS: Dict[tuple, object] = {}


@dataclass
class Request:
    rid: int
    n: int
    depth: int
    seed: int
    arrival: float = 0.0          # seconds after the start of the run
    # filled by the server
    norm2: Optional[float] = None
    z0: Optional[float] = None
    admitted: Optional[float] = None
    done: Optional[float] = None

    @property
    def footprint(self) -> int:
        return (2 ** self.n) * BYTES_C64

    @property
    def latency(self) -> float:
        return self.done - self.arrival

    def gate_data(self):
        """(depth, n, 2, 2) complex64 Haar gates, (depth,) int64 CZ masks.
        Mask bit j: CZ between index bits j, j+1, i.e. qubits (n-2-j, n-1-j)."""
        rng = np.random.default_rng(self.seed)
        mats = np.empty((self.depth, self.n, 2, 2), dtype=np.complex64)
        masks = np.zeros(self.depth, dtype=np.int64)
        for l in range(self.depth):
            for q in range(self.n):
                mats[l, q] = haar_unitary(2, rng)
            pairs = rng.choice(self.n - 1, size=self.n // 2, replace=False)
            for i in pairs:                       # pair (i, i+1) in qubit order
                masks[l] |= 1 << (self.n - 2 - int(i))
        return mats, masks

    def circuit(self):
        """The same request as a unitair_amd Circuit (one gate per call)."""
        from .circuits import Circuit, Gate
        mats, masks = self.gate_data()
        gates = []
        for l in range(self.depth):
            for q in range(self.n):
                gates.append(Gate("dense", (q,), matrix=mats[l, q].astype(np.complex128)))
            m = int(masks[l])
            for j in range(self.n - 1):
                if (m >> j) & 1:
                    gates.append(Gate("cz", (self.n - 2 - j, self.n - 1 - j)))
        return Circuit(self.n, gates)


def poisson_workload(num: int, rate: float, seed: int = 5, n_range=(10, 26),
                     depth_range=(10, 200)) -> List[Request]:
    """Requests with exponential inter-arrival times (rate per second), n
    log-uniform over n_range (inclusive), depth uniform over depth_range."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    lo, hi = n_range
    for i in range(num):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        n = int(math.floor(math.exp(rng.uniform(math.log(lo), math.log(hi + 1)))))
        n = min(max(n, lo), hi)
        depth = int(rng.integers(depth_range[0], depth_range[1] + 1))
        reqs.append(Request(i, n, depth, seed=int(rng.integers(2 ** 62)), arrival=t))
    return reqs


def _z0_diag(n: int, device) -> torch.Tensor:
    """1 - 2*(bit of qubit 0), qubit 0 = index bit n-1."""
    x = torch.arange(2 ** n, device=device)
    return (1 - 2 * ((x >> (n - 1)) & 1)).to(torch.float64)


class _Lane:
    NBUF = 2   # steps in flight per lane (gate staging buffers)

    def __init__(self, n: int, capacity: int, kmax: int, device, use_graph: bool,
                 compiled: bool = False):
        self.n = n
        self.compiled = compiled
        self.capacity = capacity
        self.kmax = kmax
        self.device = device
        self.use_graph = use_graph
        self.states = torch.empty(capacity, 2 ** n, dtype=torch.complex64, device=device)
        shape_m, shape_k = (kmax, n, capacity, 2, 2), (kmax, capacity)
        self.mats = [torch.empty(shape_m, dtype=torch.complex64, device=device)
                     for _ in range(self.NBUF)]
        self.masks = [torch.zeros(shape_k, dtype=torch.int64, device=device)
                      for _ in range(self.NBUF)]
        self.h_mats = [torch.empty(shape_m, dtype=torch.complex64, pin_memory=True)
                       for _ in range(self.NBUF)]
        self.h_masks = [torch.zeros(shape_k, dtype=torch.int64, pin_memory=True)
                        for _ in range(self.NBUF)]
        self.slots: List[Optional[Request]] = [None] * capacity
        self.pos = [0] * capacity           # next layer of each slot (as issued)
        self.data: Dict[int, tuple] = {}    # rid -> (mats, masks) host arrays
        self.stream = torch.cuda.Stream(device)   # default priority
        self.programs: Dict[tuple, object] = {}
        self.diag = _z0_diag(n, device)
        self.identity = np.eye(2, dtype=np.complex64)
        self.inflight = deque()             # (event, [(request, norm, z0)])
        self.next_buf = 0

    @property
    def bytes(self) -> int:
        return self.capacity * (2 ** self.n) * BYTES_C64

    def free_slot(self) -> Optional[int]:
        for i, r in enumerate(self.slots):
            if r is None:
                return i
        return None

    def active(self):
        return [i for i, r in enumerate(self.slots) if r is not None]

    def idle(self) -> bool:
        return not self.active() and not self.inflight

    def _program(self, K: int, used: int, buf: int):
        """FusedProgram (or its graph) for K layers over slots [0, used)
        reading staging buffer `buf`."""
        def layer_items():
            items = []
            for l in range(K):
                for q in range(self.n):
                    items.append(_Item("dense1", (self.n - 1 - q,), self.mats[buf][l, q], True))
                items.append(_Item("cz_layer", (), self.masks[buf][l], True))
            return items

        # compiled passes: always (compiled=True), or once a background
        # compile of this K's structure has finished ("background": until
        # then the lane runs the interpreted passes and never waits for it)
        compiled = self.compiled is True
        if self.compiled == "background" and self.n >= 12:
            job = _BACKGROUND_JOBS.get((self.n, K))
            if job is None:
                _BACKGROUND_JOBS[(self.n, K)] = compile_in_background(
                    layer_items(), self.n, torch.complex64, reorder=False)
            else:
                compiled = job.done() and job.exception() is None
        key = (K, used, buf, compiled)
        if key not in self.programs:
            items = layer_items()
            # built on the default stream: its op-record upload waits only
            # for that stream, not for the lane's queued steps
            # in order: a request's gates meet each amplitude in the same
            # order whatever K and the other slots, so its result is the same
            # alone and in the mix
            prog = FusedProgram.from_items(items, self.n, torch.complex64, self.device,
                                           tile_bits=TILE_BITS, reorder=False,
                                           compiled=compiled)
            # per-slot data: consecutive slots are 4 (2x2) elements apart in
            # mats[buf][l, q] and 1 mask apart in masks[buf][l], the defaults
            view = self.states[:used]
            entry = prog
            if self.use_graph:
                with torch.cuda.stream(self.stream):   # capture records, runs nothing
                    handle = _lib.stream_handle(self.device)
                    entry = (prog, StreamGraph(handle, lambda: prog.run(view)))
            self.programs[key] = entry
        return self.programs[key]

    def issue(self):
        """Queue the lane's next step (gate staging, the fused passes, the
        reductions of the requests it completes, an event) on the lane's
        stream.  Slots of completed requests are free for admission at once:
        later work on the stream is ordered after their reductions."""
        act = self.active()
        used = max(act) + 1
        K = min(self.kmax, min(self.slots[i].depth - self.pos[i] for i in act))
        b = self.next_buf
        self.next_buf = (b + 1) % self.NBUF
        hm, hk = self.h_mats[b], self.h_masks[b]
        for i in range(used):
            r = self.slots[i]
            if r is None:
                hm[:K, :, i] = torch.from_numpy(self.identity)
                hk[:K, i] = 0
                continue
            mats, masks = self.data[r.rid]
            p = self.pos[i]
            hm[:K, :, i] = torch.from_numpy(mats[p:p + K])
            hk[:K, i] = torch.from_numpy(masks[p:p + K])
        results = []
        entry = self._program(K, used, b)
        with torch.cuda.stream(self.stream):
            self.mats[b].copy_(hm, non_blocking=True)
            self.masks[b].copy_(hk, non_blocking=True)
            if isinstance(entry, tuple):
                entry[1].launch(_lib.stream_handle(self.device))
            else:
                entry.run(self.states[:used])
            for i in act:
                self.pos[i] += K
                r = self.slots[i]
                if self.pos[i] >= r.depth:
                    results.append((r,) + self.collect(i))
                    self.slots[i] = None
                    del self.data[r.rid]
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(self.stream)
        self.inflight.append((ev, results))

    def harvest(self, start_event):
        """Requests of completed steps, with their completion times (seconds
        after start_event) and results."""
        out = []
        while self.inflight and self.inflight[0][0].query():
            ev, results = self.inflight.popleft()
            t = start_event.elapsed_time(ev) / 1e3
            for r, norm, z0 in results:
                r.norm2 = float(norm)
                r.z0 = float(z0)
                r.done = t
                out.append(r)
        return out

    def collect(self, i: int):
        """float64 (||psi||^2, <Z_0>) of slot i, queued on the lane stream."""
        from .states import _reduce
        row = self.states[i:i + 1]
        return _reduce("norm", torch.float64, row), _reduce("diag", torch.float64, self.diag, row)


class Server:
    """Continuous-batching executor on one GPU (see the module docstring).

    max_active limits the requests on the GPU at once (None: slots and the
    HBM budget decide); max_active_amplitudes limits the sum of 2^n over the
    active requests (None: no limit); max_slots=1, max_active=1 runs one
    request at a time on the same fused programs.  compiled: lanes of 12 or
    more qubits run compiled passes (FusedProgram(compiled=True)); with True
    the first step of each (qubit count, layers per step) compiles its passes
    on the host, about a second each, which suits a steady mix of circuit
    shapes; with "background" that compile runs on a background host thread
    and the lane keeps the interpreted passes until it is done, so no step
    waits for it (compiled passes are kept for the life of the process).
    Results agree between the two forms to rounding."""

    def __init__(self, hbm_budget: int = 64 << 30, max_slots: int = 32,
                 max_layers_per_step: int = 8, device="cuda", use_graphs: bool = True,
                 max_active: Optional[int] = None,
                 max_active_amplitudes: Optional[int] = 1 << 22, compiled=False):
        self.budget = hbm_budget
        self.compiled = compiled
        self.max_slots = max_slots
        self.kmax = max_layers_per_step
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.use_graphs = use_graphs
        self.max_active = max_active
        self.max_amps = max_active_amplitudes
        self.lanes: Dict[int, _Lane] = {}
        self.used_bytes = 0
        self.active = 0
        self.active_amps = 0

    def _lane_capacity(self, n: int) -> int:
        fp = (2 ** n) * BYTES_C64
        return int(max(1, min(self.max_slots, (self.budget // 4) // fp)))

    def try_admit(self, r: Request, now: float) -> bool:
        if r.footprint > self.budget:
            raise ValueError(f"request {r.rid}: {r.footprint} B exceeds the HBM budget")
        if self.max_active is not None and self.active >= self.max_active:
            return False
        if self.max_amps is not None and self.active and \
                self.active_amps + (1 << r.n) > self.max_amps:
            return False
        lane = self.lanes.get(r.n)
        if lane is None:
            cap = self._lane_capacity(r.n)
            need = cap * r.footprint
            if self.used_bytes + need > self.budget:
                cap = (self.budget - self.used_bytes) // r.footprint
                if cap < 1:
                    return False
                need = cap * r.footprint
            with on_device(self.device):
                lane = _Lane(r.n, int(cap), self.kmax, self.device, self.use_graphs,
                             self.compiled)
            self.lanes[r.n] = lane
            self.used_bytes += need
        slot = lane.free_slot()
        if slot is None:
            return False
        lane.slots[slot] = r
        lane.pos[slot] = 0
        lane.data[r.rid] = r.gate_data()
        with torch.cuda.stream(lane.stream):
            lane.states[slot].zero_()
            lane.states[slot, 0] = 1
        r.admitted = now
        self.active += 1
        self.active_amps += 1 << r.n
        return True

    def run(self, requests: List[Request], realtime: bool = True) -> List[Request]:
        """Serve `requests` (arrival times in seconds from now).  Returns them
        with results and completion times filled in."""
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting: List[Request] = []
        finished = 0
        i = 0
        with on_device(self.device):
            start = torch.cuda.Event(enable_timing=True)
            start.record()
            t0 = time.perf_counter()
            while finished < len(requests):
                now = time.perf_counter() - t0 if realtime else float("inf")
                while i < len(pending) and pending[i].arrival <= now:
                    waiting.append(pending[i])
                    i += 1
                admitted = 0
                for r in waiting:                 # arrival order, no overtaking
                    if not self.try_admit(r, now):
                        break
                    admitted += 1
                waiting = waiting[admitted:]
                progress = False
                for lane in list(self.lanes.values()):
                    got = lane.harvest(start)
                    finished += len(got)
                    self.active -= len(got)
                    self.active_amps -= len(got) << lane.n
                    progress = progress or bool(got)
                    while lane.active() and len(lane.inflight) < _Lane.NBUF:
                        lane.issue()
                        progress = True
                for n in [n for n, l in self.lanes.items() if l.idle()]:
                    self.used_bytes -= self.lanes[n].bytes
                    del self.lanes[n]
                if not progress:
                    if waiting and not self.lanes and i >= len(pending):
                        raise RuntimeError("requests cannot be admitted under the budget")
                    time.sleep(2e-5)
        return requests


def run_one_at_a_time(requests: List[Request], device="cuda", realtime: bool = True):
    """Plain GPU execution: requests in arrival order, one gate per launch
    (ua_run_circuit), one request at a time.  Returns the requests with
    results and completion times."""
    from .circuits import CircuitRunner
    from .states import norm_squared_f64, _reduce
    dev = torch.device(device)
    t0 = time.perf_counter()
    for r in sorted(requests, key=lambda r: r.arrival):
        if realtime:
            while time.perf_counter() - t0 < r.arrival:
                time.sleep(0.0005)
        runner = CircuitRunner(r.circuit(), dtype=torch.complex64, device=dev)
        psi = torch.zeros(1, 2 ** r.n, dtype=torch.complex64, device=dev)
        psi[0, 0] = 1
        runner.run(psi)
        norm = norm_squared_f64(psi)
        z0 = _reduce("diag", torch.float64, _z0_diag(r.n, dev), psi)
        torch.cuda.synchronize(dev)
        r.done = time.perf_counter() - t0
        r.norm2, r.z0 = float(norm[0]), float(z0[0])
    return requests


def run_one_at_a_time_fused(requests: List[Request], device="cuda", realtime: bool = True,
                            compiled=False):
    """One request at a time on the server's own fused programs (one slot,
    one request on the GPU): the server's gain over this is batching and
    overlap alone, not fusion."""
    return Server(max_slots=1, max_active=1, device=device, compiled=compiled).run(
        requests, realtime=realtime)


def summarize(requests: List[Request]):
    lat = np.array([r.latency for r in requests])
    makespan = max(r.done for r in requests) - min(r.arrival for r in requests)
    amps = sum(r.depth * (r.n + r.n // 2) * (2 ** r.n) for r in requests)
    return {"requests": len(requests), "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "makespan_s": makespan,
            "circuits_per_s": len(requests) / makespan,
            "amp_updates_per_s": amps / makespan}
