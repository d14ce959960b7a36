"""Batching form: many independent circuits of uneven qubit counts, run together on one GPU.

A batch stands for `count` independent reference runs (qip/pipeline.py:71-135). Each run starts
from CythonBackend.make_state of |0...0> (qip/backend.py:73-104) and applies its gates one
kronselect_dot (qip/backend.py:112-116) at a time. Here the states are complex64 slices of one
device arena and a run applies every state's gate list at once (include/qipgpu.h, batching form):
  * states of n <= lds_max_qubits run whole inside one workgroup, the state in LDS, FP64 gate
    arithmetic;
  * larger states have their gates fused into <= fusion_width-qubit blocks (the precision form's
    light-cone rule), applied wave by wave with the split-BF16 MFMA tile kernel, each block
    renormalised by the FP64 norm of the previous one;
  * the run is one captured hipGraph, replayed.
An autotuner times candidate tunings (tiles per workgroup, LDS threshold, fusion width) on the
loaded circuits and keeps the fastest.

Gates are packed (``PackedCircuits``): for gate g its state, its qubits (first = MSB of the matrix
index, qubit q = state bit n-1-q, as in the reference) and its dense 2^k x 2^k matrix, controls
included, k <= 6. ``pack_descriptors`` turns kronselect_dot descriptors into this form with the
same planner as the single-state backends.
"""
import ctypes
from typing import Dict, List, Optional, Sequence, Tuple

import numpy

from . import _capi, planner
from .fused import controlled_dense

MAX_GATE_QUBITS = 6
DEFAULT_TUNING = {"fusion_width": 6, "lds_max_qubits": 13, "chunk_tiles": 8, "use_graph": 1}
INFO_KEYS = ("states", "lds_states", "tiled_states", "gates", "blocks", "waves", "chunks",
             "largest_wave", "arena_bytes", "algorithmic_bytes")


class PackedCircuits:
    """Gate lists of several circuits in the C ABI's packed form."""

    def __init__(self, n_qubits: Sequence[int], gate_state, gate_k, gate_qubits, mats):
        self.n_qubits = [int(n) for n in n_qubits]
        self.gate_state = numpy.ascontiguousarray(gate_state, dtype=numpy.int32)
        self.gate_k = numpy.ascontiguousarray(gate_k, dtype=numpy.int32)
        self.gate_qubits = numpy.ascontiguousarray(gate_qubits, dtype=numpy.int32).reshape(-1, 6)
        self.mats = numpy.ascontiguousarray(mats, dtype=numpy.complex128).reshape(-1)
        g = len(self.gate_state)
        if self.gate_k.shape != (g,) or self.gate_qubits.shape != (g, 6):
            raise ValueError("packed gate arrays disagree in length")
        if len(self.mats) != int(numpy.sum(4 ** self.gate_k.astype(numpy.int64))):
            raise ValueError("matrix data does not match the gate sizes")

    @property
    def count(self) -> int:
        return len(self.n_qubits)

    @property
    def n_gates(self) -> int:
        return len(self.gate_state)

    @classmethod
    def from_gates(cls, n_qubits: Sequence[int],
                   circuits: Sequence[Sequence[Tuple[Sequence[int], numpy.ndarray]]]) -> "PackedCircuits":
        """circuits[i] = [(qubits, dense matrix), ...] in program order."""
        if len(circuits) != len(n_qubits):
            raise ValueError("{} circuits for {} states".format(len(circuits), len(n_qubits)))
        states, ks, qs, ms = [], [], [], []
        for i, gates in enumerate(circuits):
            for gq, G in gates:
                k = len(gq)
                if not 1 <= k <= MAX_GATE_QUBITS:
                    raise NotImplementedError("batched gates act on 1..{} qubits (got {})".format(MAX_GATE_QUBITS, k))
                G = numpy.asarray(G, dtype=numpy.complex128)
                if G.shape != (2 ** k, 2 ** k):
                    raise ValueError("matrix shape {} does not match {} qubits".format(G.shape, k))
                states.append(i)
                ks.append(k)
                qs.append(list(gq) + [0] * (6 - k))
                ms.append(G.reshape(-1))
        mats = numpy.concatenate(ms) if ms else numpy.zeros(0, dtype=numpy.complex128)
        return cls(n_qubits, states, ks, numpy.array(qs, dtype=numpy.int32).reshape(-1, 6), mats)

    def gates_of(self, i: int) -> List[Tuple[Tuple[int, ...], numpy.ndarray]]:
        """The (qubits, matrix) list of circuit i, in program order."""
        offs = numpy.concatenate([[0], numpy.cumsum(4 ** self.gate_k.astype(numpy.int64))])
        out = []
        for g in numpy.nonzero(self.gate_state == i)[0]:
            k = int(self.gate_k[g])
            m = self.mats[offs[g]:offs[g + 1]].reshape(2 ** k, 2 ** k)
            out.append((tuple(int(q) for q in self.gate_qubits[g, :k]), m))
        return out


def descriptor_gates(mats) -> List[Tuple[Tuple[int, ...], numpy.ndarray]]:
    """One kronselect_dot descriptor as dense gates (controls first), in application order: the
    planner's calls, with products split into their single-qubit factors."""
    out = []
    for call in planner.plan(mats):
        if isinstance(call, planner.ProductCall):
            for q, m in zip(call.targets, call.mats):
                out.append(((q,), numpy.ascontiguousarray(m, dtype=numpy.complex128)))
            continue
        if isinstance(call, planner.GeneralCall):
            raise NotImplementedError("batched gates act on at most {} qubits, controls included "
                                      "(got {})".format(MAX_GATE_QUBITS, len(call.targets)))
        gq, G = controlled_dense(call)
        if len(gq) > MAX_GATE_QUBITS:
            raise NotImplementedError("batched gates act on at most {} qubits, controls included "
                                      "(got {})".format(MAX_GATE_QUBITS, len(gq)))
        out.append((gq, G))
    return out


def pack_descriptors(n_qubits: Sequence[int], circuits: Sequence[Sequence[Dict]]) -> PackedCircuits:
    """circuits[i] = list of kronselect_dot descriptors of circuit i, in order."""
    return PackedCircuits.from_gates(n_qubits, [[g for d in ops for g in descriptor_gates(d)] for ops in circuits])


def plan_fusion(packed: PackedCircuits, fusion_width: int = 6,
                lds_max_qubits: int = 13) -> Tuple[numpy.ndarray, numpy.ndarray]:
    """The engine's host plan, without a GPU: (block of each gate within its state or -1 for
    LDS states, blocks per state)."""
    lib = _capi.load()
    n = _capi.i32(packed.n_qubits)
    gate_block = numpy.zeros(max(packed.n_gates, 1), dtype=numpy.int32)
    state_blocks = numpy.zeros(packed.count, dtype=numpy.int32)
    _capi.check(lib.qipgpu_batch_plan(packed.count, _capi.ptr(n), packed.n_gates, _capi.ptr(packed.gate_state),
                                      _capi.ptr(packed.gate_k), _capi.ptr(packed.gate_qubits), fusion_width,
                                      lds_max_qubits, _capi.ptr(gate_block), _capi.ptr(state_blocks)),
                "batch_plan")
    return gate_block[:packed.n_gates], state_blocks


def _tuning_struct(tuning: Optional[Dict]) -> _capi.BatchTuning:
    t = dict(DEFAULT_TUNING)
    if tuning:
        unknown = set(tuning) - set(t)
        if unknown:
            raise ValueError("unknown tuning keys: {}".format(sorted(unknown)))
        t.update(tuning)
    return _capi.BatchTuning(int(t["fusion_width"]), int(t["lds_max_qubits"]), int(t["chunk_tiles"]),
                             int(t["use_graph"]))


def _tuning_dict(s: _capi.BatchTuning) -> Dict[str, int]:
    return {"fusion_width": s.fusion_width, "lds_max_qubits": s.lds_max_qubits, "chunk_tiles": s.chunk_tiles,
            "use_graph": s.use_graph}


class GpuBatch:
    """`count` complex64 states, each |0...0> until set or run, and their gate lists."""

    def __init__(self, n_qubits: Sequence[int]):
        self._lib = _capi.load()
        self.n_qubits = [int(n) for n in n_qubits]
        n = _capi.i32(self.n_qubits)
        h = ctypes.c_void_p()
        _capi.check(self._lib.qipgpu_batch_create(len(self.n_qubits), _capi.ptr(n), ctypes.byref(h)), "batch_create")
        self._h = h

    @property
    def count(self) -> int:
        return len(self.n_qubits)

    def load(self, packed: PackedCircuits, tuning: Optional[Dict] = None) -> None:
        if packed.n_qubits != self.n_qubits:
            raise ValueError("packed circuits were built for other state sizes")
        t = _tuning_struct(tuning)
        _capi.check(self._lib.qipgpu_batch_load(self._h, packed.n_gates, _capi.ptr(packed.gate_state),
                                                _capi.ptr(packed.gate_k), _capi.ptr(packed.gate_qubits),
                                                _capi.ptr(packed.mats.view(numpy.float64)), ctypes.byref(t)),
                    "batch_load")

    def load_descriptors(self, circuits: Sequence[Sequence[Dict]], tuning: Optional[Dict] = None) -> None:
        self.load(pack_descriptors(self.n_qubits, circuits), tuning)

    @property
    def tuning(self) -> Dict[str, int]:
        s = _capi.BatchTuning()
        _capi.check(self._lib.qipgpu_batch_get_tuning(self._h, ctypes.byref(s)), "batch_get_tuning")
        return _tuning_dict(s)

    def set_tuning(self, tuning: Dict) -> None:
        t = _tuning_struct(dict(self.tuning, **tuning))
        _capi.check(self._lib.qipgpu_batch_set_tuning(self._h, ctypes.byref(t)), "batch_set_tuning")

    def autotune(self, reps: int = 2) -> Tuple[Dict[str, int], float]:
        """Time candidate tunings on the loaded circuits and keep the fastest; returns it and its
        mean run time in ms. Every state is left at |0...0>."""
        s = _capi.BatchTuning()
        ms = ctypes.c_double()
        _capi.check(self._lib.qipgpu_batch_autotune(self._h, int(reps), ctypes.byref(s), ctypes.byref(ms)),
                    "batch_autotune")
        return _tuning_dict(s), float(ms.value)

    def reset(self) -> None:
        _capi.check(self._lib.qipgpu_batch_reset(self._h), "batch_reset")

    def run(self) -> None:
        _capi.check(self._lib.qipgpu_batch_run(self._h), "batch_run")

    def synchronize(self) -> None:
        _capi.check(self._lib.qipgpu_batch_synchronize(self._h), "batch_synchronize")

    def _index(self, i: int) -> int:
        i = int(i)
        if not 0 <= i < self.count:
            raise ValueError("state {} outside 0..{}".format(i, self.count - 1))
        return i

    def get_state(self, i: int) -> numpy.ndarray:
        i = self._index(i)
        out = numpy.empty(2 ** self.n_qubits[i], dtype=numpy.complex64)
        _capi.check(self._lib.qipgpu_batch_get_state(self._h, int(i), _capi.ptr(out)), "batch_get_state")
        return out

    def set_state(self, i: int, vec) -> None:
        i = self._index(i)
        v = numpy.ascontiguousarray(vec, dtype=numpy.complex64).reshape(-1)
        if len(v) != 2 ** self.n_qubits[i]:
            raise ValueError("state {} holds {} amplitudes, got {}".format(i, 2 ** self.n_qubits[i], len(v)))
        _capi.check(self._lib.qipgpu_batch_set_state(self._h, int(i), _capi.ptr(v)), "batch_set_state")

    def info(self) -> Dict[str, int]:
        out = numpy.zeros(len(INFO_KEYS), dtype=numpy.int64)
        _capi.check(self._lib.qipgpu_batch_info(self._h, _capi.ptr(out), len(INFO_KEYS)), "batch_info")
        return {k: int(v) for k, v in zip(INFO_KEYS, out)}

    @property
    def stream(self) -> int:
        s = ctypes.c_void_p()
        _capi.check(self._lib.qipgpu_batch_get_stream(self._h, ctypes.byref(s)), "batch_get_stream")
        return s.value or 0

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.qipgpu_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
