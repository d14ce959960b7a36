"""Register and measurement nodes (counterpart of qip/qip.py).

Rules kept from the reference because results depend on them:
  * every node draws a qid from one process-wide increasing counter, `Qubit.QID`
    (qip/qip.py:13,35-38), except that the parts of one split share the qid of the first part and a
    controlled op shares the qid of the op it controls. Sources take their qubit index ranges in
    qid order, and the walk breaks ties by lowest qid;
  * a node may feed into one plain Qubit node only, unless those share a qid (no cloning,
    qip/qip.py:91-99); gates and measurements may read a node freely;
  * Measure collapses through state.measure and keeps every qubit, annotating (bits, probability);
    StochasticMeasure reads state.measure_probabilities and leaves the state alone
    (qip/qip.py:135-185).
"""
import itertools
from typing import Any, Callable, Iterable, List, Optional, Sequence, Tuple, Type, Union

import numpy

from .backend import StateType
from .pipeline import PipelineObject
from .util import InitialState, flatten


def _read_by(node: PipelineObject, inputs: Sequence[PipelineObject]) -> None:
    """Register `node` as a reader (sink) of each of its inputs."""
    for src in inputs:
        src.set_sink(node)


def _segments(n: int, cuts: Iterable[int]) -> List[Tuple[int, int]]:
    """[lo, hi) index ranges covering 0..n-1, cut at each point of `cuts` inside (0, n)."""
    inner = sorted(cuts)
    if any(c < 0 or c > n for c in inner):
        raise ValueError("split points must lie within 0..{}".format(n))
    marks = [0] + [c for c in inner if 0 < c < n] + [n]
    return list(zip(marks[:-1], marks[1:]))


class Qubit(PipelineObject):
    """A register of `n` qubits; built on inputs it is the identity on them (qip/qip.py:8-102)."""
    QID = 0

    @staticmethod
    def _draw_qid() -> int:
        qid = Qubit.QID
        Qubit.QID = qid + 1
        return qid

    def __init__(self, *inputs: "Qubit", n: int = None, qid: int = None,
                 nosink: bool = False, default: InitialState = None, quantum: bool = True, **kwargs):
        super().__init__(quantum, default=default)
        width = n if n is not None else (sum(q.n for q in inputs) if inputs else 1)
        if type(default) == list and len(default) != 2 ** width:
            raise ValueError("Default state length must be 2**n")
        if width <= 0:
            raise Exception("Number of qubits must be greater than 0")
        self.n = width
        self.qid = self._draw_qid() if qid is None else qid
        self.inputs = inputs
        if not nosink:
            _read_by(self, inputs)

    def split(self, indices: Optional[Iterable[int]] = None) -> Tuple["Qubit", ...]:
        """Sub-registers cut at the given qubit offsets; by default one per input, and the node
        itself when it has exactly one input. All parts share one new qid."""
        if indices is None:
            if len(self.inputs) == 1:
                return self,
            cuts = list(itertools.accumulate(q.n for q in self.inputs))
        else:
            cuts = list(indices)
        parts: List[SplitQubit] = []
        for lo, hi in _segments(self.n, cuts):
            shared = parts[0].qid if parts else None
            parts.append(SplitQubit(list(range(lo, hi)), self, qid=shared))
        return tuple(parts)

    def extract_index(self, indices: Union[int, Iterable[int]]):
        """(selected, remaining) sub-registers of this node, sharing one qid."""
        chosen = [indices] if type(indices) == int else list(indices)
        picked = SplitQubit([i for i in chosen if 0 <= i < self.n], self)
        rest = SplitQubit([i for i in range(self.n) if i not in chosen], self, qid=picked.qid)
        return picked, rest

    def set_sink(self, sink: "PipelineObject"):
        if type(sink) is Qubit and self.sink and self.sink[0].qid != sink.qid:
            raise Exception("Qubits may only sink to one output (no cloning)")
        super().set_sink(sink)

    def __repr__(self):
        return "Q({})".format(self.qid)


class OpConstructor:
    """A callable building nodes of one kind (qip/qip.py:105-117). Wrappers such as C(...) ask it,
    through wrap_op_hook, whether it handles being wrapped itself; the base answer is no (None)."""

    def __init__(self, op: Union[Type[Qubit], Callable[..., Qubit]]):
        self.op = op

    def __call__(self, *args, **kwargs) -> Qubit:
        return self.op(*args, **kwargs)

    def wrap_op_hook(self, opconstructor: Callable[["OpConstructor"], "OpConstructor"],
                     consumed_inputs: Optional[Sequence[int]] = None) -> Optional["OpConstructor"]:
        return None

    def __repr__(self):
        return "Wrap({})".format(repr(self.op))


Q = OpConstructor(Qubit)


class SplitQubit(Qubit):
    """The qubits at positions `indices` of its input's register (qip/qip.py:123-132)."""

    def __init__(self, indices: Sequence[int], *inputs: Qubit, **kwargs):
        super().__init__(*inputs, n=len(indices), **kwargs)
        self.indices = indices

    def select_index(self, indices: Sequence[int]) -> Sequence[int]:
        return [indices[i] for i in self.indices]

    def __repr__(self) -> str:
        return "SplitQubit({})".format(",".join(map(repr, self.inputs)))


class _Readout(Qubit):
    """A node that reads its inputs' qubits, flattened in input order, from the state and annotates
    the result without removing qubits."""

    def __init__(self, *inputs: Qubit, nosink: bool = False, quantum: bool = True):
        super().__init__(*inputs, nosink=nosink, quantum=quantum)

    def read(self, state: StateType, indices: numpy.ndarray) -> Any:
        raise NotImplementedError

    def feed_indices(self, state: StateType, index_groups: Sequence[Sequence[int]], n: int) -> Tuple[int, Any]:
        return 0, self.read(state, numpy.array(flatten(index_groups), dtype=numpy.int32))


class Measure(_Readout):
    """Projective measurement of its inputs; annotation (bits, probability) (qip/qip.py:135-158)."""

    def __init__(self, *inputs: Qubit, measure_by: int = None, nosink: bool = False):
        super().__init__(*inputs, nosink=nosink)
        self.measure = measure_by

    def read(self, state: StateType, indices: numpy.ndarray) -> Tuple[int, float]:
        bits, prob = state.measure(indices)
        return bits, prob

    def __repr__(self):
        return "M({})".format(",".join(repr(i) for i in self.inputs))


class StochasticMeasure(_Readout):
    """The outcome distribution of its inputs without collapse (LSB-first, qubit j of the inputs at
    bit j); with top_k, (outcomes, probabilities) of the top_k most likely (qip/qip.py:161-185)."""

    def __init__(self, *inputs: Qubit, nosink: bool = False, top_k: int = 0):
        super().__init__(*inputs, nosink=nosink, quantum=False)
        self.top_k = top_k

    def read(self, state: StateType, indices: numpy.ndarray):
        return state.measure_probabilities(indices, top_k=self.top_k)[:]

    def __repr__(self) -> str:
        return "SM({})".format(",".join(repr(i) for i in self.inputs))
