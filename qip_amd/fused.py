"""Fused forms: gates fused into <= 6-qubit dense blocks. The precision form (FusedBackend) keeps a
complex64 state and applies the blocks on split-BF16 MFMA; Fused128Backend keeps a complex128 state and
applies the same blocks in FP64 (dense sweeps, FP64 MFMA tiles at 4 to 6 qubits).

``FusedBackend`` is a StateType (qip/backend.py:14-65) like GpuBackend, but kronselect_dot only
records: each engine call of the plan (targets, controls, matrix) becomes a dense gate on its
controls + targets, and the descriptor's gates go to the library in one qipgpu_fused_record call.
The library plans them into blocks of at most ``fusion_width`` qubits and applies each block with
the split-BF16 MFMA kernel (qipgpu_fused_apply); a run of consecutive blocks made only of
single-qubit gates is one product of per-qubit 2x2s on the FP64 product sweep instead. A block is
launched as soon as no later gate can change it (streaming), and everything recorded before the
state is next read or changed any other way (func_apply, measurement, probabilities, get_state,
ranges). Every MFMA block renormalises the previous block's output with its FP64 norm. Gates wider
than the fusion width (a multi-controlled gate, say) are applied unfused by the plain path, in order.

Light-cone rule. A gate waits for the previous gate on each of its qubits and is ready once those
are placed. A block starts from the earliest ready gate with its qubits as the block's set S, then
alternates until no ready gate fits beside S: every ready gate within S joins; then S grows by the
qubits of one ready gate that keeps it within the width, chosen by a two-step look-ahead (gates the
block would absorb per qubit added, plus half of the same for the best further gate), preferring a
gate that shares a qubit with S unless another scores more than 1.5 times as high. The next block
starts from what is left. A block's matrix is the ordered product of its gates (program order)
embedded in the block's qubits. Growing along shared qubits follows the gates' light cone: on the
30-qubit depth-40 brickwork circuit this gives 66 blocks for 1780 gates, where adding each gate to
the first block it fits in program order gives 128. ``plan_blocks``, ``fuse`` and
``StreamingFusion`` below state the rule in Python (the library's recorder is checked against them);
``compose`` is the library's block composition.

Streaming. A gate recorded later can change the plan only if it enters it: as a ready gate, which
needs all its qubits to be without a recorded gate still outside the blocks, or inside a look-ahead
that used up some qubit's recorded gates. So a block is final (it and the blocks before it are the
same whatever comes later) when every qubit of the state keeps such a gate and none of the block's
look-aheads used one up. The library plans the record again every max(16, pending / 8) gates and
launches the final leading blocks while the host records the rest; the blocks, their order and the
results are those of planning the whole record at the end.
"""
import ctypes
from typing import List, Sequence, Tuple

import numpy

from . import _capi, planner
from .backend import GpuBackend, check_windows

FUSED_MIN_QUBITS = 12       # the MFMA kernel tiles 12 state bits


class Block:
    __slots__ = ("qubits", "gates")

    def __init__(self):
        self.qubits = set()
        self.gates = []          # [(qubits tuple MSB-first, dense matrix)]

    def matrix(self) -> Tuple[List[int], numpy.ndarray]:
        """Ordered product of the block's gates on the block qubits (ascending, first = MSB),
        composed in FP64 by the library's host routine qipgpu_compose."""
        qs = sorted(self.qubits)
        return qs, compose(qs, self.gates)


def compose(qubits: Sequence[int], gates) -> numpy.ndarray:
    """G_last ... G_1 as a 2^k x 2^k complex128 matrix on ``qubits`` (first = MSB); each gate is
    (its qubits MSB-first, all among ``qubits``, its dense matrix)."""
    k = len(qubits)
    gk, gq6, mats = [], [], []
    for g, (gq, G) in enumerate(gates):
        L = len(gq)
        if L > 6 or G.size != 1 << (2 * L):
            raise ValueError("gate {} on qubits {} has a matrix of shape {}".format(g, gq, numpy.shape(G)))
        gk.append(L)
        gq6.extend(gq)
        gq6.extend((0,) * (6 - L))
        mats.append(G.ravel())
    flat = numpy.concatenate(mats).astype(numpy.complex128, copy=False) if mats else numpy.zeros(1, numpy.complex128)
    qs = numpy.array(qubits, dtype=numpy.int32)
    gk = numpy.array(gk or [0], dtype=numpy.int32)
    gq6 = numpy.array(gq6 or [0], dtype=numpy.int32)
    out = numpy.empty((1 << k, 1 << k), dtype=numpy.complex128)
    # the arrays stay referenced here for the whole call, so their raw addresses are safe to pass
    _capi.check(_capi.load().qipgpu_compose(k, qs.ctypes.data, len(gates), gk.ctypes.data, gq6.ctypes.data,
                                            flat.ctypes.data, out.ctypes.data), "compose")
    return out


def controlled_dense(call: planner.MatrixCall) -> Tuple[Tuple[int, ...], numpy.ndarray]:
    """The call as a dense gate on controls + targets (controls first, each diag(I, .))."""
    M = call.mat
    for _ in call.controls:
        d = M.shape[0]
        C = numpy.eye(2 * d, dtype=numpy.complex128)
        C[d:, d:] = M
        M = C
    return tuple(call.controls) + tuple(call.targets), M


def descriptor_gates(mats) -> List[Tuple[Tuple[int, ...], numpy.ndarray]]:
    """The dense gates, in order, that FusedBackend.kronselect_dot(mats) records (no call may span
    more than six qubits)."""
    out = []
    for call in planner.plan(mats):
        if isinstance(call, planner.ProductCall):
            out.extend(((q,), m) for q, m in zip(call.targets, call.mats))
        elif isinstance(call, planner.GeneralCall) or len(call.controls) + len(call.targets) > 6:
            raise ValueError("a call on {} qubits is applied unfused".format(len(call.controls) + len(call.targets)))
        else:
            out.append(controlled_dense(call))
    return out


def plan_blocks(qsets: Sequence[Sequence[int]], width: int, n: int = 64,
                stop_at_open: bool = False) -> Tuple[List[Tuple[frozenset, List[int]]], int]:
    """The light-cone rule over gates given by their qubit sets, in program order. Returns the
    blocks in launch order, each (qubit set, indices of its gates ascending), and how many leading
    blocks are final: no gate recorded later can change them. With ``stop_at_open`` the plan stops
    after the first block that is not final."""
    count = len(qsets)
    sets = [frozenset(q) for q in qsets]
    last, npred = {}, [0] * count
    succ = [[] for _ in range(count)]
    open_on = [0] * max(n, 1 + max((q for S in sets for q in S), default=0))
    for g, S in enumerate(sets):
        preds = {last[q] for q in S if q in last}
        npred[g] = len(preds)
        for p in preds:
            succ[p].append(g)
        for q in S:
            last[q] = g
            open_on[q] += 1
    ready = set(g for g in range(count) if npred[g] == 0)

    def absorb(S, commit, taken=None):
        """Absorb every ready gate within S, transitively; with commit=False only simulate it.
        Returns (count, True if every qubit of S keeps a recorded gate outside, the ready gates
        left outside S)."""
        pend = {}
        used = {}
        work = [g for g in ready if sets[g] <= S]
        left = [g for g in ready if not sets[g] <= S]
        c = 0
        while work:
            g = work.pop()
            c += 1
            if commit:
                ready.discard(g)
                taken.append(g)
            for q in sets[g]:
                used[q] = used.get(q, 0) + 1
            for t in succ[g]:
                v = pend.get(t, npred[t]) - 1
                pend[t] = v
                if v == 0:
                    if commit:
                        ready.add(t)
                    if sets[t] <= S:
                        work.append(t)
                    else:
                        left.append(t)
        if commit:
            for t, v in pend.items():
                npred[t] = v
            for q, u in used.items():
                open_on[q] -= u
            return c, True, left
        return c, all(open_on[q] - used.get(q, 0) > 0 for q in S), left

    def score(S, g):
        """Look-ahead score of growing S by gate g, as a fraction (num, den): the gates absorbed per
        qubit added, plus half of the same for the best further step. Also whether every simulation
        kept a recorded gate on each of its qubits."""
        S2 = S | sets[g]
        a = len(S2) - len(S)
        c, kept, left = absorb(S2, False)
        d, a2 = 0, 1
        for h in left:
            S3 = S2 | sets[h]
            if len(S3) > width:
                continue
            c3, kept3, _ = absorb(S3, False)
            kept = kept and kept3
            if (c3 - c) * a2 > d * (len(S3) - len(S2)):
                d, a2 = c3 - c, len(S3) - len(S2)
        return (2 * c * a2 + a * d, 2 * a * a2), kept

    blocks, absorbed, final, prefix_final = [], 0, 0, True
    while absorbed < count:
        S = set(sets[min(ready)])
        taken = []
        exhausted = False        # a look-ahead used up some qubit's recorded gates
        while True:
            absorbed += absorb(S, True, taken)[0]
            best_all = best_shared = None       # (score numerator, denominator, gate)
            for g in sorted(ready):
                if len(S | sets[g]) > width:
                    continue
                (num, den), kept = score(S, g)
                exhausted = exhausted or not kept
                if best_all is None or num * best_all[1] > best_all[0] * den:
                    best_all = (num, den, g)
                if sets[g] & S and (best_shared is None or num * best_shared[1] > best_shared[0] * den):
                    best_shared = (num, den, g)
            if best_all is None:
                break
            # a candidate that shares a qubit with the block, unless another scores more than 1.5
            # times as high
            if best_shared is None or 2 * best_all[0] * best_shared[1] > 3 * best_shared[0] * best_all[1]:
                S |= sets[best_all[2]]
            else:
                S |= sets[best_shared[2]]
        blocks.append((frozenset(S), sorted(taken)))
        # A later gate changes nothing up to here unless it enters the plan: as a ready gate, which
        # needs all its qubits without an open recorded gate, or through a look-ahead that used up
        # some qubit's recorded gates. So every qubit of the state must keep one, and no look-ahead
        # may have used one up.
        is_final = not exhausted and all(open_on[q] > 0 for q in range(n))
        prefix_final = prefix_final and is_final
        final += prefix_final
        if stop_at_open and not prefix_final:
            break
    return blocks, final


def fuse(gates: Sequence[Tuple[Tuple[int, ...], numpy.ndarray]], width: int) -> List[Block]:
    """The blocks of the whole record under the light-cone rule, in launch order."""
    out = []
    for S, taken in plan_blocks([gq for gq, _ in gates], width)[0]:
        b = Block()
        b.qubits = set(S)
        b.gates = [gates[i] for i in taken]
        out.append(b)
    return out


class StreamingFusion:
    """The rule applied while gates arrive: ``take_final`` removes and returns the leading blocks
    that are final so far; whenever it is called, the blocks taken plus those of the rest at the
    end are the blocks of ``fuse`` on the whole record."""

    def __init__(self, n: int, width: int):
        self.n = n
        self.width = width
        self.gates = []          # recorded, not taken

    def add(self, gq: Tuple[int, ...], G: numpy.ndarray) -> None:
        self.gates.append((gq, G))

    def _take(self, plan) -> List[Block]:
        out, used = [], set()
        for S, taken in plan:
            b = Block()
            b.qubits = set(S)
            b.gates = [self.gates[i] for i in taken]
            used.update(taken)
            out.append(b)
        self.gates = [g for i, g in enumerate(self.gates) if i not in used]
        return out

    def take_final(self) -> List[Block]:
        blocks, final = plan_blocks([gq for gq, _ in self.gates], self.width, self.n, stop_at_open=True)
        return self._take(blocks[:final])

    def take(self) -> List[Block]:
        """Every recorded gate, planned as the rest of the record."""
        return self._take(plan_blocks([gq for gq, _ in self.gates], self.width, self.n)[0])


class _GateBuffer:
    """Gates of one call packed for qipgpu_fused_record: widths, qubits in six slots, matrices."""
    __slots__ = ("k", "q", "mats")

    def __init__(self):
        self.k, self.q, self.mats = [], [], []

    def add(self, gq, G) -> None:
        L = len(gq)
        self.k.append(L)
        self.q.extend(gq)
        self.q.extend((0,) * (6 - L))
        self.mats.append(numpy.ravel(G))

    def record(self, be) -> None:
        if not self.k:
            return
        k = numpy.array(self.k, dtype=numpy.int32)
        q = numpy.array(self.q, dtype=numpy.int32)
        m = numpy.concatenate(self.mats).astype(numpy.complex128, copy=False)
        self.k, self.q, self.mats = [], [], []
        # the arrays stay referenced here for the whole call, so their raw addresses are safe to pass
        _capi.check(be._lib.qipgpu_fused_record(be._h, len(k), k.ctypes.data, q.ctypes.data, m.ctypes.data),
                    "record")


class FusedBackend(GpuBackend):
    """complex64 state; kronselect_dot records into the library's online fusion (module docstring)."""

    fusion_width = 6
    supports_gate_graphs = False      # gates are recorded for fusion, not launched one by one

    def __init__(self, n, handle, dtype):
        super().__init__(n, handle, dtype)
        self._streaming = True
        _capi.check(self._lib.qipgpu_fused_set_recording(self._h, self.fusion_width, 1), "set_recording")
        # staging for the one-gate descriptors most circuits are made of (kept for the state's life)
        self._one_k = numpy.zeros(1, dtype=numpy.int32)
        self._one_q = numpy.zeros(6, dtype=numpy.int32)
        self._one_m = numpy.zeros(1 << (2 * self.fusion_width), dtype=numpy.complex128)
        self._one_ptrs = (self._one_k.ctypes.data, self._one_q.ctypes.data, self._one_m.ctypes.data)

    @classmethod
    def make_state(cls, n, index_groups, feed_list, inputstartindex=None, inputendindex=None,
                   outputstartindex=None, outputendindex=None, statetype=numpy.complex64) -> "FusedBackend":
        if numpy.dtype(statetype) != numpy.complex64:
            statetype = numpy.complex64          # the precision form stores complex64
        return super(FusedBackend, cls).make_state(n, index_groups, feed_list, inputstartindex, inputendindex,
                                                   outputstartindex, outputendindex, statetype=statetype)

    @property
    def streaming(self) -> bool:
        """Launch each block as soon as it is final (True), or every block at the next read (False)."""
        return self._streaming

    @streaming.setter
    def streaming(self, on: bool) -> None:
        _capi.check(self._lib.qipgpu_fused_set_recording(self._h, self.fusion_width, int(bool(on))), "set_recording")
        self._streaming = bool(on)

    # -- recording ----------------------------------------------------------------------------
    def kronselect_dot(self, mats, input_offset: int = 0, output_offset: int = 0) -> None:
        check_windows(self.n, 2 ** self.n, input_offset, output_offset)
        if len(mats) == 1:
            # one plain matrix on int qubits: what the planner would record, without planning
            ((key, m),) = mats.items()
            gq = (key,) if type(key) is int else key
            if (type(m) is numpy.ndarray and type(gq) is tuple and 0 < len(gq) <= self.fusion_width
                    and m.ndim == 2 and m.shape[0] == m.shape[1] == 1 << len(gq) and m.dtype.kind in "fc"
                    and all(type(q) is int for q in gq)):
                self._record_one(gq, m)
                return
        rec = _GateBuffer()
        for call in planner.plan(mats):
            if isinstance(call, planner.ProductCall):
                for q, m in zip(call.targets, call.mats):
                    rec.add((q,), m)
            elif isinstance(call, planner.GeneralCall):
                rec.record(self)
                GpuBackend.apply_general(self, call)       # launches the record first
            elif len(call.controls) + len(call.targets) > self.fusion_width:
                rec.record(self)
                GpuBackend.apply_matrix(self, call.targets, call.mat, call.controls)   # launches the record first
            else:
                rec.add(*controlled_dense(call))
        rec.record(self)

    def _record_one(self, gq, G) -> None:
        L = len(gq)
        self._one_k[0] = L
        self._one_q[:L] = gq
        self._one_m[:1 << (2 * L)] = G.ravel()
        k, q, m = self._one_ptrs
        _capi.check(self._lib.qipgpu_fused_record(self._h, 1, k, q, m), "record")

    def apply_matrix(self, targets, mat, controls=()) -> None:
        if len(targets) + len(controls) > self.fusion_width:
            GpuBackend.apply_matrix(self, targets, mat, controls)
            return
        call = planner.MatrixCall(list(targets), list(controls), numpy.ascontiguousarray(mat, dtype=numpy.complex128))
        rec = _GateBuffer()
        rec.add(*controlled_dense(call))
        rec.record(self)

    def apply_product(self, targets, mats) -> None:
        rec = _GateBuffer()
        for q, m in zip(targets, mats):
            rec.add((q,), numpy.asarray(m).reshape(2, 2))
        rec.record(self)

    def _recorded(self) -> Tuple[int, int]:
        pending, launched = ctypes.c_int64(0), ctypes.c_int64(0)
        _capi.check(self._lib.qipgpu_fused_recorded(self._h, ctypes.byref(pending), ctypes.byref(launched)), "recorded")
        return int(pending.value), int(launched.value)

    def pending_gates(self) -> int:
        """Recorded gates not launched yet."""
        return self._recorded()[0]

    @property
    def blocks_applied(self) -> int:
        """Fused blocks launched on this state so far."""
        return self._recorded()[1]

    def flush(self) -> int:
        """Launch every recorded gate not launched yet; returns the number of blocks launched."""
        before = self.blocks_applied
        _capi.check(self._lib.qipgpu_fused_flush(self._h), "flush")
        return self.blocks_applied - before


class Fused128Backend(FusedBackend):
    """complex128 state with the same recording and fusion as FusedBackend, each block applied in FP64
    by the dense sweeps (registers up to 3 qubits, FP64 MFMA tiles at 4 to 6) instead
    of the split-BF16 kernel: the plain path's arithmetic with one dense sweep per block instead of one
    per gate (the results differ from the gate-by-gate ones only by FP64 rounding). Runs of
    single-qubit-only blocks take the product sweep, as in the precision form. There is no
    renormalisation: the blocks are applied in FP64 as composed."""

    @classmethod
    def make_state(cls, n, index_groups, feed_list, inputstartindex=None, inputendindex=None,
                   outputstartindex=None, outputendindex=None, statetype=numpy.complex128) -> "Fused128Backend":
        # complex128 whatever is asked: FusedBackend is the complex64 form
        return super(FusedBackend, cls).make_state(n, index_groups, feed_list, inputstartindex, inputendindex,
                                                   outputstartindex, outputendindex, statetype=numpy.complex128)
