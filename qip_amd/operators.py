"""Gate nodes and operator descriptors (counterpart of qip/operators.py).

A gate node hands the backend a descriptor {index group: matrix} through
``state.kronselect_dot``; ``F`` hands it a classical function through ``state.func_apply``. The
descriptor objects CMat (controlled, `_kron_struct` 2) and SwapMat (`_kron_struct` 3) carry the
integer tags of the reference's MatrixType enum (qip/ext/kronprod.pyx:18-19), so descriptors
built here and by the reference are interchangeable.
"""
from typing import Any, Callable, Mapping, Sequence, Tuple, Union

import numpy

from .backend import StateType
from .qip import OpConstructor, Qubit
from .qubit_util import QubitOpWrapper
from .util import IndexType, MatrixType, flatten


class MatrixOp(Qubit):
    """A node applying a fixed operator; matrices are built once, on first use, then reused
    (qip/operators.py:9-34). dagger=True applies the conjugate transpose."""

    def __init__(self, *inputs: Qubit, **kwargs):
        super().__init__(*inputs, **kwargs)
        self.ms = None
        self.dagger = kwargs.get("dagger", False)

    def feed_indices(self, state: StateType, index_groups: Sequence[Sequence[int]], n: int) -> Tuple[int, int]:
        if self.ms is None:
            self.ms = self.makemats(index_groups)
            if self.dagger:
                self.ms = {k: v.conj().T for k, v in self.ms.items()}
        state.kronselect_dot(self.ms)
        return 0, 0

    def makemats(self, index_groups: Sequence[Sequence[int]]) -> Mapping[IndexType, MatrixType]:
        raise NotImplementedError("This method should never be called.")


def _per_qubit(mat, index_groups):
    return {i: mat for i in flatten(index_groups)}


class NotOp(MatrixOp):
    def makemats(self, index_groups):
        return _per_qubit(numpy.flip(numpy.eye(2), 0), index_groups)

    def __repr__(self) -> str:
        return "Not({})".format(self.qid)


Not = QubitOpWrapper(NotOp)


class XOp(NotOp):
    def __repr__(self):
        return "X({})".format(self.qid)


X = QubitOpWrapper(XOp)


class YOp(MatrixOp):
    def makemats(self, index_groups):
        return _per_qubit(numpy.array([[0.0, -1.0j], [1.0j, 0.0]]), index_groups)

    def __repr__(self) -> str:
        return "Y({})".format(self.qid)


Y = QubitOpWrapper(YOp)


class ZOp(MatrixOp):
    def makemats(self, index_groups):
        return _per_qubit(numpy.array([[1.0, 0.0], [0.0j, -1.0]]), index_groups)

    def __repr__(self) -> str:
        return "Z({})".format(self.qid)


Z = QubitOpWrapper(ZOp)


class HOp(MatrixOp):
    def makemats(self, index_groups):
        return _per_qubit((1 / numpy.sqrt(2)) * numpy.array([[1, 1], [1, -1]]), index_groups)

    def __repr__(self) -> str:
        return "H({})".format(self.qid)


H = QubitOpWrapper(HOp)


class ROp(MatrixOp):
    """diag(1, e^{i phi}) on every qubit of the inputs."""

    def __init__(self, phi: float, *inputs: Qubit, **kwargs):
        super().__init__(*inputs, **kwargs)
        self.exponented = numpy.exp(1.0j * phi)

    def makemats(self, index_groups):
        return _per_qubit(numpy.array([[1, 0], [0, self.exponented]]), index_groups)

    def __repr__(self) -> str:
        return "R({})".format(",".join(map(repr, self.inputs)))


R = QubitOpWrapper(ROp)


class RmOp(ROp):
    """R with phi = +-2 pi / 2^m."""

    def __init__(self, m: int, *inputs: Qubit, negate: bool = False, **kwargs):
        super().__init__((-2 if negate else 2) * numpy.pi / pow(2.0, m), *inputs, **kwargs)
        self.m = m

    def __repr__(self) -> str:
        return "R[{}]({})".format(self.m, ",".join(map(repr, self.inputs)))


Rm = QubitOpWrapper(RmOp)


class SwapOp(MatrixOp):
    """Exchange two equal-size registers."""

    def __init__(self, *inputs: Qubit, **kwargs):
        super().__init__(*inputs, **kwargs)
        if len(self.inputs) != 2:
            raise Exception("Swap can only take two inputs")
        if self.inputs[0].n != self.inputs[1].n:
            raise Exception("Inputs must be of equal size {}/{}".format(self.inputs[0], self.inputs[1]))

    def makemats(self, index_groups):
        return {tuple(flatten([index_groups[0], index_groups[1]])): SwapMat(self.inputs[0].n)}

    def __repr__(self) -> str:
        return "Swap({})".format(self.qid)


Swap = QubitOpWrapper(SwapOp)


class SwapMat(object):
    """The 2^(2n) x 2^(2n) operator exchanging two n-qubit blocks: entry (i, j) is 1 iff the low
    half of i equals the high half of j and vice versa (qip/operators.py:157-179)."""
    _kron_struct = 3

    def __init__(self, n: int):
        self.n = n
        self.shape = (2 ** (2 * n), 2 ** (2 * n))

    def __getitem__(self, item: Union[int, Tuple[int, int]]) -> complex:
        if type(item) == tuple and len(item) == 2:
            mask = (1 << self.n) - 1
            low_a, high_a = item[0] & mask, item[0] >> self.n
            low_b, high_b = item[1] & mask, item[1] >> self.n
            return 1.0 if low_a == high_b and low_b == high_a else 0.0
        raise ValueError("SwapMat can only be indexed with M[i,j] not M[{}]".format(item))

    def __repr__(self):
        return "SwapMat({})".format(self.n)


def C(op: OpConstructor) -> OpConstructor:
    """Controlled version of an op: the new first input is the control qubit."""
    if hasattr(op, "wrap_op_hook"):
        return op.wrap_op_hook(C, consumed_inputs=[0]) or QubitOpWrapper(COp, op)
    return QubitOpWrapper(COp, op)


class COp(MatrixOp):
    """Controlled op node: every matrix M of the inner op becomes CMat(M) with the control index
    prepended (qip/operators.py:195-230)."""

    def __init__(self, op: OpConstructor, *inputs: Qubit, **kwargs):
        if len(inputs) < 2:
            raise ValueError("Not enough input values given.")
        if inputs[0].n != 1:
            raise ValueError("Control bit can only be of size n=1")
        if len(set(inputs)) < len(inputs):
            raise ValueError("Cannot have repeated qubits in COp")
        nkwargs = dict(kwargs)
        nkwargs["nosink"] = True
        nkwargs["nosplit"] = True
        self.op = self.make_op(op, inputs, nkwargs)
        if not issubclass(type(self.op), MatrixOp):
            raise ValueError("C(Op) may only be applied to matrix ops and swaps, not {}".format(type(self.op)))
        super().__init__(*inputs, qid=self.op.qid, **kwargs)

    def make_op(self, op: OpConstructor, inputs: Sequence[Qubit], nkwargs: Mapping[str, Any]):
        return op(*inputs[1:], **nkwargs)

    def makemats(self, index_groups):
        inner = self.op.makemats(index_groups[1:])
        return {tuple(flatten([index_groups[0], idx])): CMat(inner[idx]) for idx in inner}

    def __repr__(self):
        return "C{}".format(self.op)


class CMat(object):
    """diag(I, m): the controlled form of m, control as the most significant bit."""
    _kron_struct = 2

    def __init__(self, mat: Sequence[Sequence[complex]]):
        self.m = numpy.array(mat) if type(mat) == list else mat
        self.shape = (self.m.shape[0] * 2, self.m.shape[1] * 2)

    def __getitem__(self, item: Union[int, Tuple[int, int]]):
        if type(item) == tuple and len(item) == 2:
            row, col = item
            h0, h1 = self.shape[0] // 2, self.shape[1] // 2
            if row < h0 and col < h1:
                return 1.0 if row == col else 0.0
            if row >= h0 and col >= h1:
                return self.m[row - h0, col - h1]
            return 0.0
        raise ValueError("CMat can only be indexed with M[i,j] not M[{}]".format(item))

    def conj(self):
        return CMat(self.m.conj())

    @property
    def T(self):
        return CMat(self.m.T)

    def numpy(self):
        out = numpy.zeros(shape=self.shape, dtype=numpy.complex128)
        for i in range(self.shape[0]):
            for j in range(self.shape[1]):
                out[i, j] = self[i, j]
        return out


class FOp(Qubit):
    """|x>|q> -> |x>|func(x) xor q> for registers reg1, reg2 (qip/operators.py:271-298)."""

    def __init__(self, func: Callable[[int], int], reg1: Qubit, reg2: Qubit, **kwargs):
        super().__init__(reg1, reg2, **kwargs)
        self.func = func
        self.reg1 = reg1
        self.reg2 = reg2

    def feed_indices(self, state: StateType, index_groups: Sequence[Sequence[int]], n: int) -> Tuple[int, int]:
        reg1 = numpy.array(index_groups[0], dtype=numpy.int32)
        reg2 = numpy.array(index_groups[1], dtype=numpy.int32)
        state.func_apply(reg1, reg2, self.func, n)
        return 0, 0

    def __repr__(self) -> str:
        return "F({}, {})".format(self.reg1, self.reg2)


F = QubitOpWrapper(FOp)
