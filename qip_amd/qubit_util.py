"""Operator wrappers and control contexts (counterpart of qip/qubit_util.py).

QubitOpWrapper reproduces one quirk on purpose because circuit results depend on it: when called,
it rebuilds its keyword arguments from the ones given at wrap time and drops the caller's
(qip/qubit_util.py:142-143). So ``Rm(m, q, negate=True)`` builds the +2*pi/2^m phase, which is
what QFFT's makeR asks for (qip/qfft.py:42-43): QFFT(rev=True) therefore equals
numpy.fft.ifft(x, norm="ortho"). ``nosplit`` and ``nocontext`` are named parameters and are kept.
"""
import collections.abc
from typing import Any, Callable, Mapping, Optional, Sequence, Tuple, Type, Union

from .qip import OpConstructor, Qubit
from .util import flatten


class QubitWrapperContext:
    """Within `with QubitWrapperContext(C, [q]):` every wrapped op is built as C(op) with q
    inserted as its control (qip/qubit_util.py:7-128)."""
    TOTAL_CONTEXT_QUBITS = 0
    CONTEXT_STACK = []

    def __init__(self, *args: Any):
        super().__init__()
        self.context = []
        self.context_size = 0
        buff = []
        for arg in args:
            if isinstance(arg, collections.abc.Iterable):
                buff.append(list(arg))
            else:
                if len(buff) > 0:
                    self.add_context(*buff)
                    buff = []
                buff.append(arg)
        if len(buff) > 0:
            self.add_context(*buff)

    def add_context(self, constructor: OpConstructor, qubits: Sequence[Qubit], indices: Optional[Sequence[int]] = None):
        if indices is None:
            indices = list(range(len(qubits)))
        self.context.append((constructor, qubits, indices))
        self.context_size += len(qubits)

    @staticmethod
    def apply_to_qubits_in_context(op_func: OpConstructor, *qubits: Qubit,
                                   op_args: Optional[Sequence[Any]] = None,
                                   op_kwargs: Optional[Mapping[str, Any]] = None) -> Tuple[Qubit, ...]:
        qubits_list = list(qubits)
        in_context = len(QubitWrapperContext.CONTEXT_STACK) > 0
        if in_context:
            qubits_list = QubitWrapperContext.put_qubits_in_context_order(*qubits_list)
            op_func = QubitWrapperContext.make_context_constructor(op_func)
        kwargs = dict(op_kwargs) if op_kwargs else {}
        kwargs["nocontext"] = True
        kwargs["nosplit"] = True
        qubits_list = op_func(*qubits_list, **kwargs).split()
        if in_context:
            applied, qubits_list = QubitWrapperContext.split_context_and_circuit_qubits(*qubits_list)
            QubitWrapperContext.set_qubits(applied)
        return qubits_list

    def put_qubits_in_local_context_order(self, *qubits: Qubit) -> Tuple[Qubit, ...]:
        out = list(qubits)
        for _, cq, ci in reversed(self.context):
            for q, i in zip(cq, ci):
                out.insert(i, q)
        return tuple(out)

    @staticmethod
    def put_qubits_in_context_order(*qubits: Qubit) -> Tuple[Qubit, ...]:
        out = list(qubits)
        for _, cq, ci in reversed(QubitWrapperContext.get_context()):
            for q, i in zip(cq, ci):
                out.insert(i, q)
        return tuple(out)

    @staticmethod
    def split_context_and_circuit_qubits(*qubits: Qubit) -> Tuple[Sequence[Qubit], Sequence[Qubit]]:
        rest = list(qubits)
        applied = []
        for ctx in QubitWrapperContext.get_context():
            for i in ctx[2]:
                applied.append(rest.pop(i))
        return applied, rest

    @staticmethod
    def make_context_constructor(op_func: OpConstructor, op_args=None, op_kwargs=None) -> OpConstructor:
        for constructor, _, _ in reversed(QubitWrapperContext.get_context()):
            op_func = constructor(op_func)
        return op_func

    @staticmethod
    def get_context():
        return flatten(QubitWrapperContext.CONTEXT_STACK)

    @staticmethod
    def set_qubits(qubits_list: Sequence[Qubit]):
        if len(qubits_list) != QubitWrapperContext.TOTAL_CONTEXT_QUBITS:
            raise ValueError("Size of qubits_list incorrect: {} is not required {}"
                             .format(len(qubits_list), QubitWrapperContext.TOTAL_CONTEXT_QUBITS))
        pos = 0
        for context_set in QubitWrapperContext.CONTEXT_STACK:
            for i, ctx in enumerate(context_set):
                new_qubits = qubits_list[pos:pos + len(ctx[1])]
                context_set[i] = (ctx[0], new_qubits, ctx[2])
                pos += len(new_qubits)

    def __enter__(self):
        QubitWrapperContext.CONTEXT_STACK.append(self.context)
        QubitWrapperContext.TOTAL_CONTEXT_QUBITS += self.context_size
        return self

    def __exit__(self, exc_type, exc_val, exc_tb):
        QubitWrapperContext.CONTEXT_STACK.pop()
        QubitWrapperContext.TOTAL_CONTEXT_QUBITS -= self.context_size


class QubitOpWrapper(OpConstructor):
    """Wraps an op class: calling it builds the node (inside any active contexts) and returns its
    outputs split per input (qip/qubit_util.py:131-168)."""

    def __init__(self, op: Union[Type[Qubit], Callable[..., Qubit]], *args, **kwargs):
        super().__init__(op)
        self.args = args
        self.kwargs = kwargs

    def __call__(self, *inputs: Qubit, nosplit: bool = False, nocontext: bool = False,
                 **kwargs) -> Union[Qubit, Tuple[Qubit, ...]]:
        # the caller's keyword arguments are replaced by the wrap-time ones (module docstring)
        kwargs = dict(self.kwargs)
        if nocontext:
            q = self.op(*self.args, *inputs, nocontext=nocontext, **kwargs)
            if len(q.inputs) == 1 or nosplit:
                return q
            return q.split()
        outs = QubitWrapperContext.apply_to_qubits_in_context(self, *inputs)
        if len(outs) > 1:
            return outs
        if len(outs) == 1:
            return outs[0]
        raise ValueError("Somehow ended up with zero output qubits.")

    def wrap_op_hook(self, opconstructor, consumed_inputs=None) -> Optional[OpConstructor]:
        return None

    def __repr__(self):
        if len(self.args) == 0:
            return "OpWrap({})".format(self.op)
        return "OpWrap({},+{} args)".format(self.op, len(self.args))


class QubitFuncWrapper(OpConstructor):
    """Turns a circuit-building function into an op that C(...) can wrap: the wrappers applied to
    it become a context around the function body (qip/qubit_util.py:171-205)."""

    def __init__(self, func: Callable[..., Union[Qubit, Tuple[Qubit, ...]]]):
        super().__init__(func)
        self.wrapper_funcs = []

    def __call__(self, *inputs: Qubit, **kwargs) -> Union[Qubit, Tuple[Qubit, ...]]:
        remaining = list(inputs)
        taken = []
        for opconstructor, consumed in reversed(self.wrapper_funcs):
            consumed_qubits = reversed([remaining.pop(i) for i in reversed(consumed)])
            taken.append((opconstructor, consumed_qubits, consumed))
        taken = list(reversed(taken))
        with QubitWrapperContext(*flatten(taken)) as context:
            outputs = self.op(*remaining, **kwargs)
            if isinstance(outputs, Qubit):
                outputs = (outputs,)
            ordered = context.put_qubits_in_local_context_order(*outputs)
            if len(ordered) == 1:
                return ordered[0]
            return ordered

    def wrap_op_hook(self, opconstructor, consumed_inputs=None) -> OpConstructor:
        self.wrapper_funcs.append((opconstructor, consumed_inputs))
        return self

    @staticmethod
    def wrap(op_func: Callable[..., Qubit]) -> "QubitFuncWrapper":
        return QubitFuncWrapper(op_func)
