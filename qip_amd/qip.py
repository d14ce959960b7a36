"""Qubit and measurement nodes (counterpart of qip/qip.py).

Semantics kept from the reference because results depend on them:
  * every node takes a process-wide increasing `qid` unless given one; input nodes get their
    qubit index ranges in qid order and execution breaks ties by lowest qid (qip/qip.py:13,35-38);
  * a Qubit may sink into one Qubit only (no cloning, qip/qip.py:91-99);
  * Measure collapses through state.measure and keeps every qubit; StochasticMeasure reads
    state.measure_probabilities without changing the state (qip/qip.py:135-185).
"""
from typing import Any, Callable, Iterable, Optional, Sequence, Tuple, Type, Union

import numpy

from .backend import StateType
from .pipeline import PipelineObject, graph_changed
from .util import InitialState, flatten


class Qubit(PipelineObject):
    """A register of `n` qubits; with inputs it is the identity on them (qip/qip.py:8-102)."""
    QID = 0

    def __init__(self, *inputs: "Qubit", n: int = None, qid: int = None,
                 nosink: bool = False, default: InitialState = None, quantum: bool = True, **kwargs):
        super().__init__(quantum, default=default)
        if n is None:
            n = sum(q.n for q in inputs) if len(inputs) > 0 else 1
        if type(default) == list and len(default) != 2 ** n:
            raise ValueError("Default state length must be 2**n")
        if n <= 0:
            raise Exception("Number of qubits must be greater than 0")
        self.n = n
        if qid is None:
            qid = Qubit.QID
            Qubit.QID += 1
        self.qid = qid
        self.inputs = inputs
        if not nosink:
            for item in inputs:
                item.set_sink(self)

    def split(self, indices: Optional[Iterable[int]] = None) -> Tuple["Qubit", ...]:
        """Split into sub-registers at the given qubit offsets (default: one per input)."""
        if indices is None:
            cuts = []
            acc = 0
            for q in self.inputs:
                acc += q.n
                cuts.append(acc)
            if len(cuts) == 1:
                return self,
        else:
            cuts = sorted(indices)
        if 0 not in cuts:
            cuts = [0] + cuts
        if self.n not in cuts:
            cuts = cuts + [self.n]
        parts = []
        qid = None
        for lo, hi in zip(cuts[:-1], cuts[1:]):
            sq = SplitQubit(list(range(lo, hi)), self, qid=qid)
            parts.append(sq)
            qid = sq.qid
        return tuple(parts)

    def extract_index(self, indices: Union[int, Iterable[int]]):
        """(selected, remaining) sub-registers (qip/qip.py:76-87)."""
        if type(indices) == int:
            indices = [indices]
        indices = list(indices)
        sq = SplitQubit([i for i in indices if 0 <= i < self.n], self)
        rest = SplitQubit([i for i in range(self.n) if i not in indices], self, qid=sq.qid)
        return sq, rest

    def set_sink(self, sink: "PipelineObject"):
        graph_changed()
        if type(sink) == Qubit:
            if len(self.sink) == 0 or self.sink[0].qid == sink.qid:
                self.sink += [sink]
            else:
                raise Exception("Qubits may only sink to one output (no cloning)")
        else:
            self.sink += [sink]

    def __repr__(self):
        return "Q({})".format(self.qid)


class OpConstructor:
    """Callable wrapper of a node class (qip/qip.py:103-118)."""

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
    """A selection of the qubit indices of its input (qip/qip.py:123-132)."""

    def __init__(self, indices: Sequence[int], *inputs: Qubit, **kwargs):
        super().__init__(*inputs, n=len(indices), **kwargs)
        self.indices = indices

    def select_index(self, indices: Sequence[int]) -> Sequence[int]:
        return [indices[i] for i in self.indices]

    def __repr__(self) -> str:
        return "SplitQubit({})".format(",".join(map(repr, self.inputs)))


class Measure(Qubit):
    """Projective measurement of its inputs; annotation (bits, probability) (qip/qip.py:135-158)."""

    def __init__(self, *inputs: Qubit, measure_by: int = None, nosink: bool = False):
        super().__init__(*inputs, nosink=nosink)
        self.inputs = inputs
        self.n = sum(q.n for q in self.inputs)
        self.measure = measure_by
        if not nosink:
            for item in inputs:
                item.set_sink(self)

    def feed_indices(self, state: StateType, index_groups: Sequence[Sequence[int]], n: int) -> Tuple[int, Any]:
        indices = numpy.array(flatten(index_groups), dtype=numpy.int32)
        bits, prob = state.measure(indices)
        return 0, (bits, prob)

    def __repr__(self):
        return "M({})".format(",".join(i.__repr__() for i in self.inputs))


class StochasticMeasure(Qubit):
    """Outcome distribution of its inputs without collapse; with top_k, (outcomes, probabilities)
    of the top_k most likely (qip/qip.py:161-185)."""

    def __init__(self, *inputs: Qubit, nosink: bool = False, top_k: int = 0):
        super().__init__(*inputs, quantum=False, nosink=nosink)
        self.top_k = top_k
        if not nosink:
            for item in inputs:
                item.set_sink(self)

    def feed_indices(self, state: StateType, index_groups: Sequence[Sequence[int]], n: int) -> Tuple[int, object]:
        indices = numpy.array(flatten(index_groups), dtype=numpy.int32)
        probs = state.measure_probabilities(indices, top_k=self.top_k)[:]
        return 0, probs

    def __repr__(self) -> str:
        return "SM({})".format(",".join(i.__repr__() for i in self.inputs))
