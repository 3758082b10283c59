"""Compile a gate list into fused passes (ua_run_fused_pass) and run them.

A pass keeps m = min(n, 12) index bits local: bits 0..3 always (so tiles are
made of 128 B runs) plus the target bits of the gates it absorbs.  Gates are
taken in circuit order; a gate joins the current pass while the union of
local bits stays within m:
  dense 1-/2-qubit gate : needs its target bits local
  CNOT                  : needs its target bit local (control: any bit)
  CZ, phase, diagonal   : need nothing (applied element by element)
Wider dense gates end the pass and run as their own ua_apply_operator call.
Inside a pass, gates are grouped into stages: cube stages, whose dense
targets fit in a 4-bit cube, and diagonal stages that carry the full-length
phases (and diagonal gates next to them) in the coalesced tile order.  Consecutive 1-qubit gates on the same qubit are multiplied on
the host first (complex128), as unitair's fuse_single_qubit_operators
(unitair/gates/matrix_algebra.py:6-51) does for apply_to_qubits.
"""
from __future__ import annotations

import concurrent.futures
import ctypes
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .circuits import Circuit, Gate

TILE_BITS = 12
CUBE_BITS = 4
MAX_STAGES = _lib.UA_FUSED_MAX_STAGES

(FOP_DENSE1, FOP_DENSE2, FOP_CNOT, FOP_CZ, FOP_DIAG1, FOP_PHASE, FOP_CZ_LAYER,
 FOP_LAYER1Q, FOP_PHASE_FACTORS) = range(9)


FusedOp = _lib.FusedOp
FusedPass = _lib.FusedPass


@dataclass
class _Item:
    """A gate in state-bit terms.  kind: dense1/dense2/cnot/cz/phase/diag1/
    cz_layer (data: int64 mask(s) of adjacent index-bit pairs)."""
    kind: str
    bits: tuple                      # state bits (dense: gate MSB first)
    data: Optional[np.ndarray] = None
    per_state: bool = False          # data has a leading batch dim
    # how the planner derived this item from its input items (None: it is
    # one of them): ("as_dense1", item) or ("product", later, earlier)
    src: Optional[tuple] = field(default=None, repr=False, compare=False)


@dataclass
class _Pass:
    local: List[int]
    items: List[_Item] = field(default_factory=list)


def _merge_1q(items: List[_Item]) -> List[_Item]:
    """Multiply runs of 1-qubit gates on the same bit that are adjacent in
    the gate order (nothing in between touches that bit)."""
    out: List[_Item] = []
    last_on_bit = {}
    for it in items:
        touched = set(it.bits) if it.kind not in ("phase", "cz_layer") else None
        if it.kind == "dense1" and not it.per_state:
            j = last_on_bit.get(it.bits[0])
            if j is not None and out[j].kind == "dense1" and not out[j].per_state:
                out[j] = _Item("dense1", it.bits, it.data @ out[j].data)
                continue
            out.append(it)
            last_on_bit[it.bits[0]] = len(out) - 1
            continue
        out.append(it)
        if touched is None:          # full-state diagonal touches every bit
            last_on_bit.clear()
        else:
            for bt in touched:
                last_on_bit.pop(bt, None)
    return out


def _as_dense1(it: _Item) -> _Item:
    """A 1-qubit diagonal gate as a 2x2 matrix (host or per-state device data)."""
    if it.kind == "dense1":
        return it
    d = it.data
    m = torch.diag_embed(d).contiguous() if isinstance(d, torch.Tensor) else np.diag(np.asarray(d))
    return _Item("dense1", it.bits, m, it.per_state)


def _product(later: _Item, earlier: _Item) -> _Item:
    """later @ earlier for two 1-qubit gates on the same bit; per-state data
    (device tensors, one matrix per state) is multiplied state by state."""
    a, b = later.data, earlier.data
    if isinstance(a, torch.Tensor) or isinstance(b, torch.Tensor):
        dev = a.device if isinstance(a, torch.Tensor) else b.device
        dt = a.dtype if isinstance(a, torch.Tensor) else b.dtype
        a = torch.as_tensor(a).to(dev, dt)
        b = torch.as_tensor(b).to(dev, dt)
        return _Item("dense1", later.bits, torch.matmul(a, b).contiguous(),
                     later.per_state or earlier.per_state)
    return _Item("dense1", later.bits, np.asarray(a) @ np.asarray(b))


def _segment_stage(ops: List[_Item], cube_state_bits) -> list:
    """Rewrite one stage's gate list for the kernel: every maximal run of
    1-qubit gates on the stage's cube bits (diagonal ones included, as 2x2
    matrices) becomes ("layer", [[gates on cube bit 0], ..., [on bit 3]]),
    with consecutive gates on one bit multiplied together; 1-qubit gates on
    distinct bits commute, so the run can be applied bit by bit.  Other
    gates stay ("op", item) in place."""
    cube_pos = {sb: k for k, sb in enumerate(cube_state_bits)}
    out, run = [], None
    for it in ops:
        if it.kind in ("dense1", "diag1") and it.bits[0] in cube_pos:
            if run is None:
                run = [[] for _ in range(len(cube_state_bits))]
            lane = run[cube_pos[it.bits[0]]]
            g = _as_dense1(it)
            if g is not it:
                g.src = ("as_dense1", it)
            if lane:
                prod = _product(g, lane[-1])
                prod.src = ("product", g, lane[-1])
                lane[-1] = prod
            else:
                lane.append(g)
            continue
        if run is not None:
            out.append(("layer", run))
            run = None
        out.append(("op", it))
    if run is not None:
        out.append(("layer", run))
    return out


def _check_wide(k: int):
    """Dense gates wider than 2 qubits run in place through ua_apply_operator,
    which updates in place for 1..UA_MAX_GATE_QUBITS target qubits."""
    if not 1 <= k <= _lib.UA_MAX_GATE_QUBITS:
        raise ValueError(f"fused programs take dense gates on 1..{_lib.UA_MAX_GATE_QUBITS} "
                         f"qubits (got {k})")


def _items_from_circuit(circ: Circuit) -> List[_Item]:
    n = circ.num_qubits
    items = []
    for g in circ.gates:
        bits = tuple(n - 1 - q for q in g.qubits)
        if g.kind == "dense":
            k = len(g.qubits)
            _check_wide(k)
            items.append(_Item(f"dense{k}" if k <= 2 else "wide", bits,
                               np.asarray(g.matrix, dtype=np.complex128)))
        elif g.kind == "phase":
            items.append(_Item("phase", (), g.angles))
        else:
            items.append(_Item(g.kind, bits))
    return items


def _needs(it: _Item):
    if it.kind in ("dense1", "dense2"):
        return set(it.bits)
    if it.kind == "cnot":
        return {it.bits[1]}
    return set()


_DIAG = ("phase", "cz", "diag1", "cz_layer")


def _commute(a: _Item, b: _Item) -> bool:
    """True when the two gates commute exactly (as operators), so either order
    gives the same state: both diagonal; disjoint bits; a diagonal gate on a
    CNOT's control only; two CNOTs neither of which targets the other's
    control.  Full-length phases and CZ layers touch every bit."""
    da, db = a.kind in _DIAG, b.kind in _DIAG
    if da and db:
        return True
    if a.kind == "wide" or b.kind == "wide":
        return False
    ta = None if a.kind in ("phase", "cz_layer") else set(a.bits)
    tb = None if b.kind in ("phase", "cz_layer") else set(b.bits)
    if ta is None or tb is None:
        return False
    common = ta & tb
    if not common:
        return True
    for x, y, dy in ((a, b, db), (b, a, da)):
        if x.kind == "cnot" and dy and common <= {x.bits[0]}:
            return True
    if a.kind == "cnot" and b.kind == "cnot":
        return a.bits[1] != b.bits[0] and b.bits[1] != a.bits[0]
    return False


def _bit_actions(it: _Item, n: int):
    """How the gate acts on each bit it touches: "Z" (diagonal there), "X"
    (a CNOT target) or "G" (anything else).  Full-length phases and CZ layers
    are diagonal on every bit."""
    if it.kind in ("phase", "cz_layer"):
        return [(b, "Z") for b in range(n)]
    if it.kind in ("cz", "diag1"):
        return [(b, "Z") for b in it.bits]
    if it.kind == "cnot":
        return [(it.bits[0], "Z"), (it.bits[1], "X")]
    return [(b, "G") for b in it.bits]


def _predecessors(items: Sequence[_Item], n: int):
    """preds[j]: earlier items that item j must follow.  On one bit, gates of
    the same action type Z or X commute (the rule of _commute), so a run of
    them forms a group; a gate depends, on each of its bits, on the last
    group it does not join (and a wide gate on everything before it).  Every
    earlier gate that does not commute with j is an ancestor of preds[j]."""
    groups = {}                      # bit -> (type, members, previous group)
    preds = []
    for j, it in enumerate(items):
        if it.kind == "wide":
            preds.append(set(range(j)))
            for b in range(n):
                groups[b] = ("G", [j], [])
            continue
        deps = set()
        for b, t in _bit_actions(it, n):
            g = groups.get(b)
            if g is None:
                groups[b] = (t, [j], [])
                continue
            gt, members, prev = g
            if t == gt and t in ("Z", "X"):
                deps.update(prev)
                members.append(j)
            else:
                deps.update(members)
                groups[b] = (t, [j], members)
        preds.append(deps)
    return preds


class _Ready:
    """Items whose predecessors are all placed, kept up to date by counters."""

    def __init__(self, preds):
        self.succ = [[] for _ in preds]
        for j, ps in enumerate(preds):
            for i in ps:
                self.succ[i].append(j)
        self.waiting = [len(ps) for ps in preds]
        self.ready = [j for j in range(len(preds)) if self.waiting[j] == 0]

    def place(self, j):
        self.ready.remove(j)
        for k in self.succ[j]:
            self.waiting[k] -= 1
            if self.waiting[k] == 0:
                self.ready.append(k)

    def sorted(self):
        return sorted(self.ready)


def _plan_passes_in_order(items: Sequence[_Item], n: int, tile_bits: int = TILE_BITS):
    """Greedy split of the item list, in its order, into fused passes and
    wide gates."""
    m = min(n, tile_bits)
    base = set(range(min(n, CUBE_BITS)))
    out = []
    cur: Optional[_Pass] = None
    cur_local = set()
    for it in items:
        if it.kind == "wide":
            if cur is not None:
                out.append(cur)
                cur = None
            out.append(it)
            continue
        need = _needs(it)
        if cur is not None and len(cur_local | need) <= m:
            cur_local |= need
            cur.items.append(it)
            continue
        if cur is not None:
            out.append(cur)
        cur_local = base | need
        cur = _Pass(local=[], items=[it])
    if cur is not None:
        out.append(cur)
    # pad each pass's local set to m bits and fix its order
    for p in out:
        if isinstance(p, _Pass):
            loc = set(range(min(n, CUBE_BITS)))
            for it in p.items:
                loc |= _needs(it)
            for bt in range(n):
                if len(loc) >= m:
                    break
                loc.add(bt)
            p.local = sorted(loc)
    return out


def plan_passes(items: Sequence[_Item], n: int, tile_bits: int = TILE_BITS,
                reorder: bool = True):
    """Split the item list into fused passes and wide gates.

    List scheduling over the commutation order: a pass keeps taking any gate
    whose non-commuting predecessors are all placed and whose local bits
    fit (gates needing no new local bit first), so gates that commute with a
    pass-ending gate still join the pass.  Wide gates run alone, in order.
    reorder=False keeps the item order (every amplitude then sees its gates
    in the same order whatever the pass grouping, so results do not depend
    on how a gate list is cut into programs)."""
    if not reorder:
        return _plan_passes_in_order(items, n, tile_bits)
    m = min(n, tile_bits)
    base = set(range(min(n, CUBE_BITS)))
    items = list(items)
    rd = _Ready(_predecessors(items, n))
    out = []
    while rd.ready:
        order = rd.sorted()
        if all(items[j].kind == "wide" for j in order):
            j = order[0]
            rd.place(j)
            out.append(items[j])
            continue
        local = set(base)
        taken = []
        while True:
            best = None
            for j in rd.sorted():
                if items[j].kind == "wide":
                    continue
                grow = len(_needs(items[j]) - local)
                if len(local) + grow > m:
                    continue
                if best is None or grow < best[0]:
                    best = (grow, j)
                    if grow == 0:
                        break
            if best is None:
                break
            j = best[1]
            local |= _needs(items[j])
            rd.place(j)
            taken.append(items[j])
        out.append(_Pass(local=[], items=taken))
    # pad each pass's local set to m bits and fix its order
    for p in out:
        if isinstance(p, _Pass):
            loc = set(range(min(n, CUBE_BITS)))
            for it in p.items:
                loc |= _needs(it)
            for bt in range(n):
                if len(loc) >= m:
                    break
                loc.add(bt)
            p.local = sorted(loc)
    return out


def _stages_in_order(p: _Pass):
    """Split a pass into stages: [(cube, [items])].

    cube is the sorted list of 4 tile bits of a cube stage, or None for a
    diagonal stage.  Full-length phases always run in diagonal stages (the
    kernel reads their angles in the coalesced tile order); CZ, 1-qubit
    diagonals and CZ layers that follow a phase join its diagonal stage
    (diagonal gates commute), otherwise they ride in the current cube stage,
    which needs no cube bits for them.  Dense gates and CNOT need their
    target bits in the stage's cube."""
    tile_of = {s: i for i, s in enumerate(p.local)}
    m = len(p.local)
    stages = []
    cur = None                       # [cube set or None, items]
    for it in p.items:
        if it.kind == "phase" or (it.kind in _DIAG and cur is not None and
                                  cur[0] is None):
            if cur is None or cur[0] is not None:
                if cur is not None:
                    stages.append(cur)
                cur = [None, []]
            cur[1].append(it)
            continue
        need = {tile_of[b] for b in _needs(it)}
        if cur is not None and cur[0] is not None and len(cur[0] | need) <= CUBE_BITS:
            cur[0] |= need
            cur[1].append(it)
            continue
        if cur is not None:
            stages.append(cur)
        cur = [set(need), [it]]
    if cur is not None:
        stages.append(cur)
    result = []
    for cube, ops in stages:
        if cube is None:
            result.append((None, ops))
            continue
        cube = set(cube)
        for t in range(m):
            if len(cube) >= CUBE_BITS:
                break
            cube.add(t)
        result.append((sorted(cube), ops))
    return result


def _stages(p: _Pass, reorder: bool = True):
    """Split a pass into stages: [(cube, [items])], reordering gates only
    across gates they commute with.

    cube is the sorted list of 4 tile bits of a cube stage, or None for a
    diagonal stage.  Full-length phases always run in diagonal stages (the
    kernel reads their factors in the coalesced tile order), together with
    every other diagonal gate that is ready then.  A cube stage starts from
    the first ready dense gate or CNOT and keeps taking ready gates whose
    target bits fit its 4 cube bits (diagonal gates other than phases need
    none); a diagonal stage is made only when no such gate is ready.
    reorder=False keeps the item order."""
    if not reorder:
        return _stages_in_order(p)
    tile_of = {s: i for i, s in enumerate(p.local)}
    m = len(p.local)
    items = list(p.items)
    nbits = 1 + max([b for it in items if it.kind not in ("phase", "cz_layer")
                     for b in it.bits] + list(p.local) + [0])
    rd = _Ready(_predecessors(items, nbits))
    stages = []
    while rd.ready:
        if any(items[j].kind != "phase" for j in rd.ready):
            cube, ops = set(), []
            while True:
                best = None
                for j in rd.sorted():
                    if items[j].kind == "phase":
                        continue
                    grow = len({tile_of[b] for b in _needs(items[j])} - cube)
                    if len(cube) + grow > CUBE_BITS:
                        continue
                    if best is None or grow < best[0]:
                        best = (grow, j)
                        if grow == 0:
                            break
                if best is None:
                    break
                j = best[1]
                cube |= {tile_of[b] for b in _needs(items[j])}
                rd.place(j)
                ops.append(items[j])
            stages.append((cube, ops))
            continue
        ops = []
        while True:
            diag = [j for j in rd.sorted() if items[j].kind in _DIAG]
            if not diag:
                break
            for j in diag:
                rd.place(j)
                ops.append(items[j])
        stages.append((None, ops))
    result = []
    for cube, ops in stages:
        if cube is None:
            result.append((None, ops))
            continue
        cube = set(cube)
        for t in range(m):
            if len(cube) >= CUBE_BITS:
                break
            cube.add(t)
        result.append((sorted(cube), ops))
    return result


def _pass_structs(items, n: int, tile_bits: int = TILE_BITS, reorder: bool = True):
    """The plan of `items` as kernel structures, without touching a device:
    steps ("fused", FusedPass, first op, op count) | ("wide", qubits, item),
    the flat FusedOp list (data pointers unset) and the (op index, item) pairs
    whose data the ops read."""
    plan = plan_passes(items, n, tile_bits, reorder)
    steps, op_structs = [], []
    pending = []             # (op index, item) whose data pointer is set later
    for p in plan:
        if isinstance(p, _Item):
            qubits = [n - 1 - b for b in p.bits]
            steps.append(("wide", qubits, p))
            continue
        stages = _stages(p, reorder)
        for chunk in range(0, len(stages), MAX_STAGES):
            part = stages[chunk:chunk + MAX_STAGES]
            fp = FusedPass()
            fp.tile_bits = len(p.local)
            for i, s in enumerate(p.local):
                fp.local_bits[i] = s
            fp.num_stages = len(part)
            first_op = len(op_structs)
            needs_index = 0
            tile_of = {s: i for i, s in enumerate(p.local)}
            for si, (cube, ops) in enumerate(part):
                for k in range(CUBE_BITS):
                    fp.stage_bits[si][k] = cube[k] if cube is not None else -1
                fp.stage_ops[si] = len(op_structs) - first_op
                if cube is None:
                    pos = {}
                    entries = [("op", it) for it in ops]
                else:
                    pos = {t: i for i, t in enumerate(cube)}
                    entries = _segment_stage(ops, [p.local[t] for t in cube])
                flat_ops = []
                for kind, x in entries:
                    if kind == "op":
                        flat_ops.append(x)
                        continue
                    head = FusedOp()
                    head.kind = FOP_LAYER1Q
                    head.a = sum(len(lane) << (8 * k) for k, lane in enumerate(x))
                    flat_ops.append(head)
                    for lane in x:
                        flat_ops.extend(lane)
                for it in flat_ops:
                    if isinstance(it, FusedOp):       # a layer head
                        op_structs.append(it)
                        continue
                    o = FusedOp()
                    if it.kind == "dense1":
                        o.kind, o.a = FOP_DENSE1, pos[tile_of[it.bits[0]]]
                    elif it.kind == "dense2":
                        o.kind = FOP_DENSE2
                        o.a, o.b = pos[tile_of[it.bits[0]]], pos[tile_of[it.bits[1]]]
                    elif it.kind == "cnot":
                        o.kind, o.a, o.b = FOP_CNOT, it.bits[0], pos[tile_of[it.bits[1]]]
                        needs_index = 1
                    elif it.kind == "cz":
                        o.kind, o.a, o.b = FOP_CZ, it.bits[0], it.bits[1]
                        needs_index = 1
                    elif it.kind == "diag1":
                        o.kind, o.a = FOP_DIAG1, it.bits[0]
                        needs_index = 1
                    elif it.kind == "phase":
                        o.kind = FOP_PHASE_FACTORS
                        needs_index = 1
                    elif it.kind == "cz_layer":
                        o.kind = FOP_CZ_LAYER
                        needs_index = 1
                    else:
                        raise ValueError(it.kind)
                    if it.data is not None:
                        pending.append((len(op_structs), it))
                    op_structs.append(o)
            fp.stage_ops[len(part)] = len(op_structs) - first_op
            fp.needs_index = needs_index
            steps.append(("fused", fp, first_op, len(op_structs) - first_op))
    return steps, op_structs, pending


class PlanTemplate:
    """The plan of a gate-item list with its data left open: the kernel
    structures of _pass_structs, and for each op that reads data a recipe in
    terms of the input items' positions ("item", i) | ("as_dense1", r) |
    ("product", r_later, r_earlier).  FusedProgram.from_items(items,
    template=t) builds a program for any item list of the same kinds, bits
    and per-state flags without planning again."""

    def __init__(self, items, n: int, tile_bits: int = TILE_BITS, reorder: bool = True):
        items = list(items)
        self.n, self.tile_bits, self.reorder = n, tile_bits, reorder
        self.signature = [(it.kind, tuple(it.bits), bool(it.per_state)) for it in items]
        index = {id(it): i for i, it in enumerate(items)}

        def recipe(it):
            if it.src is None:
                return ("item", index[id(it)])
            if it.src[0] == "as_dense1":
                return ("as_dense1", recipe(it.src[1]))
            return ("product", recipe(it.src[1]), recipe(it.src[2]))

        steps, self.op_structs, pending = _pass_structs(items, n, tile_bits, reorder)
        self.steps = [st if st[0] == "fused" else ("wide", st[1], index[id(st[2])])
                      for st in steps]
        self.pending = [(idx, recipe(it)) for idx, it in pending]

    def check(self, items):
        sig = [(it.kind, tuple(it.bits), bool(it.per_state)) for it in items]
        if sig != self.signature:
            raise ValueError("the items differ from the template's in kind, bits or "
                             "per-state data")

    @staticmethod
    def evaluate(recipe, items):
        if recipe[0] == "item":
            return items[recipe[1]]
        if recipe[0] == "as_dense1":
            return _as_dense1(PlanTemplate.evaluate(recipe[1], items))
        return _product(PlanTemplate.evaluate(recipe[1], items),
                        PlanTemplate.evaluate(recipe[2], items))


def _structure_key(code: int, n: int, fp, host_ops: int) -> tuple:
    """What a compiled pass depends on: dtype, n, the pass record and the kind
    and bits of each op (not the data pointers or strides)."""
    count = fp.stage_ops[fp.num_stages]
    ops = (FusedOp * max(1, count)).from_address(host_ops)
    return (code, n, bytes(fp), tuple((ops[j].kind, ops[j].a, ops[j].b) for j in range(count)))


# default tiles of compiled passes: 2^13 amplitudes (64 KiB of LDS for
# complex64, 128 KiB for complex128).  ua_compile_fused_pass also takes
# 2^14-amplitude complex64 tiles (fewer passes, but measured slower on the
# 24- and 28-qubit circuits: 1024-thread workgroups, one per CU).
COMPILED_TILE_BITS = {torch.complex64: 13, torch.complex128: 13}

_DUMMY_DATA = {"dense1": np.eye(2), "dense2": np.eye(4), "diag1": np.ones(2),
               "cz_layer": np.zeros(1, dtype=np.int64)}
_BACKGROUND = []     # the executor of compile_in_background, made on first use


def compile_in_background(items, n: int, dtype: torch.dtype = torch.complex64,
                          reorder: bool = True) -> concurrent.futures.Future:
    """Compile the passes of `items` (those with 12 tile bits) on a background
    host thread, for FusedProgram(compiled=True) to pick up later without
    compiling: the plan depends only on the items' kinds and bits, so it is
    made from placeholder data and nothing touches the GPU.  The future's
    result is the number of passes compiled."""
    code = _lib.DTYPE_CODES[dtype]
    shapes = [_Item(it.kind, tuple(it.bits), _DUMMY_DATA.get(it.kind), False) for it in items]

    def job():
        lib = _lib.load()
        template = PlanTemplate(shapes, n, reorder=reorder)
        done = 0
        for st in template.steps:
            if st[0] != "fused" or st[1].tile_bits < TILE_BITS:
                continue
            _, fp, first, count = st
            ops = (FusedOp * max(1, count))(*[FusedOp.from_buffer_copy(o)
                                               for o in template.op_structs[first:first + count]])
            for o in ops:
                if o.kind not in (FOP_CZ, FOP_CNOT, FOP_LAYER1Q):
                    o.data = 0x10000      # any aligned non-NULL address: nothing runs
            key = _structure_key(code, n, fp, ctypes.addressof(ops))
            if key in _COMPILED:
                continue
            handle = ctypes.c_void_p()
            _lib.check(lib.ua_compile_fused_pass(code, n, ctypes.byref(fp), ctypes.addressof(ops),
                                                 ctypes.byref(handle)), "compiling a fused pass")
            _COMPILED[key] = handle
            done += 1
        return done

    if not _BACKGROUND:
        _BACKGROUND.append(concurrent.futures.ThreadPoolExecutor(
            max_workers=max(1, min(8, (os.cpu_count() or 2) // 2))))
    return _BACKGROUND[0].submit(job)


# structure key -> handle of its compiled pass (handles live as long as the
# process; the library keeps the same table keyed by the generated source).
# The key leaves out UA_TUNE_COMPILED_WAVES: set that knob before the first
# compile of a structure.
_COMPILED = {}


class FusedProgram:
    """A circuit compiled into fused passes for one (n, dtype, device).

    Built from a Circuit, or with from_items() from gate items whose data may
    be per state (a device tensor with a leading batch dimension)."""

    def __init__(self, circuit: Circuit, dtype: torch.dtype = torch.complex64,
                 device="cuda", tile_bits: Optional[int] = None, merge_1q: bool = True,
                 reorder: bool = True, compiled: bool = False, phase_angles: bool = False):
        items = _items_from_circuit(circuit)
        if merge_1q:
            items = _merge_1q(items)
        self._build(items, circuit.num_qubits, dtype, device, tile_bits, reorder, compiled,
                    phase_angles)
        self.num_gates = len(circuit.gates)

    @classmethod
    def from_items(cls, items: List[_Item], n: int, dtype: torch.dtype = torch.complex64,
                   device="cuda", tile_bits: Optional[int] = None,
                   reorder: bool = True, compiled: bool = False,
                   phase_angles: bool = False,
                   template: Optional[PlanTemplate] = None) -> "FusedProgram":
        """template: a PlanTemplate of an item list with the same kinds, bits
        and per-state flags (its n, tile_bits and reorder apply); the plan is
        then taken from it instead of planned again."""
        prog = cls.__new__(cls)
        prog._build(list(items), n, dtype, device, tile_bits, reorder, compiled, phase_angles,
                    template)
        prog.num_gates = len(items)
        return prog

    def _build(self, items, n, dtype, device, tile_bits, reorder=True, compiled=False,
               phase_angles=False, template=None):
        if tile_bits is None:
            # compiled passes take larger tiles (fewer passes); the interpreter 12
            tile_bits = COMPILED_TILE_BITS.get(dtype, TILE_BITS) if compiled else TILE_BITS
        if dtype not in (torch.complex64, torch.complex128):
            raise ValueError("fused programs run complex64/complex128 states")
        for it in items:
            if it.kind == "wide":
                _check_wide(len(it.bits))
        self.lib = _lib.load()
        self.n = n
        if self.n < CUBE_BITS:
            raise ValueError("fused programs need at least 4 qubits")
        self.dtype = dtype
        self.code = _lib.DTYPE_CODES[dtype]
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        real = torch.float32 if dtype == torch.complex64 else torch.float64
        if template is None:
            template = PlanTemplate(items, self.n, tile_bits, reorder)
        else:
            if template.n != self.n:
                raise ValueError("the template was planned for another qubit count")
            template.check(items)
        # fresh op records per program (data pointers differ between programs)
        op_structs = [FusedOp.from_buffer_copy(o) for o in template.op_structs]
        pending = [(idx, PlanTemplate.evaluate(r, items)) for idx, r in template.pending]
        self._keep = []
        self.steps = []
        for st in template.steps:
            if st[0] == "wide":
                m = torch.from_numpy(np.asarray(items[st[2]].data)).to(self.device, dtype)
                self._keep.append(m)
                self.steps.append(("wide", st[1], m))
            else:
                self.steps.append(st)
        for o in op_structs:
            o.cache_slot = -1           # unused by the kernel (gate data is read as scalars)
        # upload operand data and fill data pointers
        for idx, it in pending:
            o = op_structs[idx]
            want = {"phase": real, "cz_layer": torch.int64}.get(it.kind, dtype)
            if isinstance(it.data, torch.Tensor) and it.per_state:
                t = it.data
                if t.device != self.device or t.dtype != want or not t.is_contiguous():
                    raise ValueError("device gate data must be contiguous, of the program's "
                                     "dtype and device")
            elif isinstance(it.data, torch.Tensor):
                t = it.data.to(self.device, want).contiguous()
            else:
                t = torch.as_tensor(np.ascontiguousarray(it.data)).to(self.device, want)
            if it.kind == "phase" and phase_angles:
                # the pass reads the angles and forms exp(-1j * angle) itself
                o.kind = FOP_PHASE
            elif it.kind == "phase":
                # exp(-1j * angles), formed once by the drop-in phase kernel
                # (the same factors apply_phase multiplies by); the pass
                # then multiplies by them
                from .simulation import apply_phase
                t = apply_phase(t, torch.ones(2 ** self.n, dtype=dtype, device=self.device))
                t = t.contiguous()
            per = {"phase": 2 ** self.n, "diag1": 2, "dense1": 4, "dense2": 16,
                   "cz_layer": 1}[it.kind]
            o.data_bstride = per if it.per_state else 0
            self._keep.append(t)
            o.data = t.data_ptr()
        arr = (FusedOp * max(1, len(op_structs)))(*op_structs)
        self._host_ops = arr
        raw = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8)
        self._ops_dev = raw.to(self.device)
        base_ptr = self._ops_dev.data_ptr()
        self._calls = []
        for st in self.steps:
            if st[0] == "fused":
                _, fp, first, count = st
                self._calls.append(("fused", fp, base_ptr + first * ctypes.sizeof(FusedOp),
                                    ctypes.addressof(arr) + first * ctypes.sizeof(FusedOp)))
            else:
                _, qubits, m = st
                self._calls.append(("wide", _lib.int_array(qubits), len(qubits), m))
        self.num_passes = len(self._calls)
        self.num_compiled = 0
        if compiled:
            self._compile()

    def _structure_key(self, i):
        return _structure_key(self.code, self.n, self._calls[i][1], self._calls[i][3])

    def _compile(self):
        """Replace every pass with 12 or more tile bits by its compiled kernel
        (ua_compile_fused_pass; new structures compile on parallel threads,
        known ones come from a per-process table without a library call)."""
        todo = [i for i, c in enumerate(self._calls)
                if c[0] == "fused" and c[1].tile_bits >= TILE_BITS]
        keys = {i: self._structure_key(i) for i in todo}
        new = [i for i in todo if keys[i] not in _COMPILED]

        def one(i):
            fp, host = self._calls[i][1], self._calls[i][3]
            handle = ctypes.c_void_p()
            rc = self.lib.ua_compile_fused_pass(self.code, self.n, ctypes.byref(fp),
                                                ctypes.c_void_p(host), ctypes.byref(handle))
            return rc, handle

        if new:
            workers = max(1, min(8, len(new), os.cpu_count() or 1))
            with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as ex:
                results = list(ex.map(one, new))
            for i, (rc, handle) in zip(new, results):
                _lib.check(rc, "compiling a fused pass")
                _COMPILED[keys[i]] = handle
        with torch.cuda.device(self.device):
            for i in todo:
                # load the code on this GPU now, so that no launch (and no graph
                # capture) has to
                _lib.check(self.lib.ua_load_compiled_pass(_COMPILED[keys[i]]),
                           "loading a compiled fused pass")
                self._calls[i] = ("compiled", _COMPILED[keys[i]], self._calls[i][2])
        self.num_compiled = len(todo)

    def run(self, state: torch.Tensor) -> torch.Tensor:
        """Apply the program to a contiguous (batch, 2**n) state in place."""
        if state.device != self.device or state.dtype != self.dtype or \
                state.shape[-1] != 2 ** self.n or not state.is_contiguous():
            raise ValueError("state must be a contiguous (batch, 2**n) tensor of the "
                             "program's dtype and device")
        batch = state.numel() // (2 ** self.n)
        stream = _lib.stream_handle(self.device)
        for c in self._calls:
            if c[0] == "compiled":
                rc = self.lib.ua_run_compiled_pass(c[1], _lib.ptr(state), batch,
                                                   ctypes.c_void_p(c[2]), stream)
            elif c[0] == "fused":
                rc = self.lib.ua_run_fused_pass(_lib.ptr(state), self.code, self.n, batch,
                                                ctypes.byref(c[1]), ctypes.c_void_p(c[2]),
                                                ctypes.c_void_p(c[3]), stream)
            else:
                rc = self.lib.ua_apply_operator(_lib.ptr(state), _lib.ptr(state),
                                                _lib.ptr(c[3]), self.code, self.n, c[2],
                                                c[1], batch, 2 ** self.n, 0, stream)
            if rc:
                _lib.check(rc, "fused program")
        return state

    def capture(self, state: torch.Tensor):
        """Record run(state) into a CUDA/HIP graph; replay() re-runs it."""
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self.run(state)
        torch.cuda.current_stream(self.device).wait_stream(s)
        self._graph = (g, state)
        return g

    def replay(self):
        self._graph[0].replay()
        return self._graph[1]
