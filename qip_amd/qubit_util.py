"""Op wrappers and control contexts (counterpart of qip/qubit_util.py).

A context entry is (constructor, control qubits, positions). While entries are active, building a
wrapped op goes through three steps:
  1. insert: the entries' qubits are placed into the op's inputs at their positions, innermost
     entry first, so the outermost entry's qubits end up in front;
  2. wrap: the op constructor is wrapped by every entry's constructor, innermost first
     (C(C(op)) for two nested C entries);
  3. release: the wrapped op's outputs at the entries' positions are taken out, outermost entry
     first, and become the entries' qubits for the next op built in the context.
Entries live in frames, one per `with QubitWrapperContext(...)` block.

QubitOpWrapper reproduces one quirk on purpose, because circuit results depend on it: a call
builds the node with the keyword arguments given at wrap time and drops the caller's
(qip/qubit_util.py:142-143). So `Rm(m, q, negate=True)` builds the +2 pi / 2^m phase, which is what
QFFT's makeR asks for (qip/qfft.py:42-43): QFFT(rev=True) equals numpy.fft.ifft(x, norm="ortho").
`nosplit` and `nocontext` are named parameters of the call and are kept.
"""
import collections.abc
from typing import Any, Callable, List, Mapping, Optional, Sequence, Tuple, Type, Union

from .qip import OpConstructor, Qubit
from .util import flatten


class _Entry:
    __slots__ = ("constructor", "qubits", "positions")

    def __init__(self, constructor, qubits: Sequence[Qubit], positions: Sequence[int]):
        self.constructor = constructor
        self.qubits = list(qubits)
        self.positions = list(positions)


_FRAMES: List[List[_Entry]] = []


def _active() -> List[_Entry]:
    return [e for frame in _FRAMES for e in frame]


def _insert(entries: Sequence[_Entry], qubits: Sequence[Qubit]) -> List[Qubit]:
    out = list(qubits)
    for e in reversed(entries):
        for q, pos in zip(e.qubits, e.positions):
            out.insert(pos, q)
    return out


def _wrap(entries: Sequence[_Entry], op: OpConstructor) -> OpConstructor:
    for e in reversed(entries):
        op = e.constructor(op)
    return op


def _release(entries: Sequence[_Entry], outputs: Sequence[Qubit]) -> List[Qubit]:
    rest = list(outputs)
    for e in entries:
        e.qubits = [rest.pop(pos) for pos in e.positions]
    return rest


def _entry_groups(args: Sequence[Any]) -> List[List[Any]]:
    """Split context arguments into (constructor, qubits[, positions]) groups: every argument that
    is not iterable starts a new group, iterables join the current one as lists."""
    groups: List[List[Any]] = []
    for a in args:
        if isinstance(a, collections.abc.Iterable):
            if not groups:
                groups.append([])
            groups[-1].append(list(a))
        else:
            groups.append([a])
    return groups


class QubitWrapperContext:
    """`with QubitWrapperContext(C, [q]):` builds every wrapped op inside as C(op) with q inserted as
    its first input; arguments are constructors, each followed by its qubits and optionally the
    positions they take (default 0..len-1) (qip/qubit_util.py:7-128)."""

    def __init__(self, *args: Any):
        self.context: List[_Entry] = []
        self.context_size = 0
        for group in _entry_groups(args):
            self.add_context(*group)

    def add_context(self, constructor: OpConstructor, qubits: Sequence[Qubit], indices: Optional[Sequence[int]] = None):
        positions = list(range(len(qubits))) if indices is None else list(indices)
        if len(positions) != len(qubits):
            raise ValueError("Size of qubits_list incorrect: {} is not required {}".format(len(positions), len(qubits)))
        self.context.append(_Entry(constructor, qubits, positions))
        self.context_size += len(qubits)

    @staticmethod
    def apply_to_qubits_in_context(op_func: OpConstructor, *qubits: Qubit,
                                   op_args: Optional[Sequence[Any]] = None,
                                   op_kwargs: Optional[Mapping[str, Any]] = None) -> Tuple[Qubit, ...]:
        """Build op_func on `qubits` inside the active contexts; returns the op's split outputs
        with the context qubits taken out."""
        entries = _active()
        if entries:
            qubits = _insert(entries, qubits)
            op_func = _wrap(entries, op_func)
        outputs = op_func(*qubits, nocontext=True, nosplit=True).split()
        return tuple(_release(entries, outputs)) if entries else outputs

    def put_qubits_in_local_context_order(self, *qubits: Qubit) -> Tuple[Qubit, ...]:
        """`qubits` with this context's qubits put back at their positions."""
        return tuple(_insert(self.context, qubits))

    def __enter__(self):
        _FRAMES.append(self.context)
        return self

    def __exit__(self, exc_type, exc_val, exc_tb):
        _FRAMES.pop()


class QubitOpWrapper(OpConstructor):
    """Calling it builds the wrapped node inside any active contexts and returns the node's outputs,
    one per input (qip/qubit_util.py:131-168). Call-time keyword arguments are dropped (module
    docstring)."""

    def __init__(self, op: Union[Type[Qubit], Callable[..., Qubit]], *args, **kwargs):
        super().__init__(op)
        self.args = args
        self.kwargs = kwargs

    def __call__(self, *inputs: Qubit, nosplit: bool = False, nocontext: bool = False,
                 **dropped) -> Union[Qubit, Tuple[Qubit, ...]]:
        if not nocontext:
            outs = QubitWrapperContext.apply_to_qubits_in_context(self, *inputs)
            if not outs:
                raise ValueError("Somehow ended up with zero output qubits.")
            return outs[0] if len(outs) == 1 else outs
        node = self.op(*self.args, *inputs, nocontext=True, **self.kwargs)
        return node if (nosplit or len(node.inputs) <= 1) else node.split()

    def __repr__(self):
        if self.args:
            return "OpWrap({},+{} args)".format(self.op, len(self.args))
        return "OpWrap({})".format(self.op)


class QubitFuncWrapper(OpConstructor):
    """Turns a circuit-building function into an op that wrappers such as C(...) can take: each
    wrapper is remembered with the input positions it consumes, and a call runs the function inside
    a context made of them (qip/qubit_util.py:171-205)."""

    def __init__(self, func: Callable[..., Union[Qubit, Tuple[Qubit, ...]]]):
        super().__init__(func)
        self.wrapper_funcs: List[Tuple[Any, Sequence[int]]] = []

    def __call__(self, *inputs: Qubit, **kwargs) -> Union[Qubit, Tuple[Qubit, ...]]:
        remaining = list(inputs)
        peeled = []
        # the last wrapper consumed its inputs from what the earlier ones left: undo in reverse
        for constructor, consumed in reversed(self.wrapper_funcs):
            taken = [remaining.pop(i) for i in sorted(consumed, reverse=True)]
            peeled.append((constructor, taken[::-1], list(consumed)))
        ctx = QubitWrapperContext()
        for constructor, taken, consumed in reversed(peeled):
            ctx.add_context(constructor, taken, consumed)
        with ctx:
            outputs = self.op(*remaining, **kwargs)
            if isinstance(outputs, Qubit):
                outputs = (outputs,)
            ordered = ctx.put_qubits_in_local_context_order(*outputs)
        return ordered[0] if len(ordered) == 1 else ordered

    def wrap_op_hook(self, opconstructor, consumed_inputs=None) -> OpConstructor:
        self.wrapper_funcs.append((opconstructor, list(consumed_inputs or [])))
        return self

    @staticmethod
    def wrap(op_func: Callable[..., Qubit]) -> "QubitFuncWrapper":
        return QubitFuncWrapper(op_func)
