"""Gate nodes and operator descriptors (counterpart of qip/operators.py).

A gate node hands the backend a descriptor {index group: matrix} through `state.kronselect_dot`;
F hands it a classical function through `state.func_apply`. The descriptor objects CMat
(controlled, `_kron_struct` 2) and SwapMat (`_kron_struct` 3) carry the integer tags of the
reference's MatrixType enum (qip/ext/kronprod.pyx:18-19), so descriptors built here and by the
reference are interchangeable.

The fixed gate matrices are the reference's values bit for bit (H's entries are 1/sqrt(2) computed
as 1.0 / numpy.sqrt(2), not sqrt(0.5), which differs in the last bit), since the amplitudes depend
on them.
"""
from typing import Any, Callable, Dict, Mapping, Sequence, Tuple, Union

import numpy

from . import planner
from .backend import StateType
from .qip import OpConstructor, Qubit
from .qubit_util import QubitOpWrapper
from .util import IndexType, MatrixType, flatten

_INV_SQRT2 = 1.0 / numpy.sqrt(2)


class MatrixOp(Qubit):
    """A node applying an operator descriptor made by `makemats(index_groups)`. The descriptor is
    made on first use and kept in `ms`, and made again only if the node's index groups change
    between runs (the reference keeps the first one, qip/operators.py:15-26). dagger=True applies
    the conjugate transpose of every matrix. The descriptor is a planner.Descriptor with read-only
    numpy matrices, so the engine calls planned on its first use serve the later runs."""

    def __init__(self, *inputs: Qubit, **kwargs):
        super().__init__(*inputs, **kwargs)
        self.dagger = kwargs.get("dagger", False)
        self.ms = None
        self._ms_groups = None

    def descriptor(self, index_groups: Sequence[Sequence[int]]) -> Mapping[IndexType, MatrixType]:
        if self.ms is None or self._ms_groups != index_groups:
            groups = [list(g) for g in index_groups]
            mats = self.makemats(groups)
            if self.dagger:
                mats = {k: v.conj().T for k, v in mats.items()}
            mats = planner.Descriptor(mats)
            for m in mats.values():
                planner.freeze(m)
            self.ms, self._ms_groups = mats, groups
        return self.ms

    def feed_indices(self, state: StateType, index_groups: Sequence[Sequence[int]], n: int) -> Tuple[int, int]:
        state.kronselect_dot(self.descriptor(index_groups))
        return 0, 0

    def makemats(self, index_groups: Sequence[Sequence[int]]) -> Mapping[IndexType, MatrixType]:
        raise NotImplementedError("This method should never be called.")


class _EachQubit(MatrixOp):
    """One 2x2 matrix, `gate()`, on every qubit of the inputs: one index group per qubit."""
    label = "?"

    def gate(self) -> numpy.ndarray:
        raise NotImplementedError

    def makemats(self, index_groups: Sequence[Sequence[int]]) -> Dict[int, numpy.ndarray]:
        m = self.gate()
        return {q: m for q in flatten(index_groups)}

    def __repr__(self) -> str:
        return "{}({})".format(self.label, self.qid)


class NotOp(_EachQubit):
    label = "Not"

    def gate(self):
        return numpy.array([[0.0, 1.0], [1.0, 0.0]])


class XOp(NotOp):
    label = "X"


class YOp(_EachQubit):
    label = "Y"

    def gate(self):
        return numpy.array([[0.0, -1.0j], [1.0j, 0.0]])


class ZOp(_EachQubit):
    label = "Z"

    def gate(self):
        return numpy.array([[1.0, 0.0], [0.0, -1.0]], dtype=numpy.complex128)


class HOp(_EachQubit):
    label = "H"

    def gate(self):
        return _INV_SQRT2 * numpy.array([[1, 1], [1, -1]])


class ROp(_EachQubit):
    """diag(1, e^{i phi}) on every qubit of the inputs."""

    def __init__(self, phi: float, *inputs: Qubit, **kwargs):
        super().__init__(*inputs, **kwargs)
        self.exponented = numpy.exp(1.0j * phi)

    def gate(self):
        return numpy.array([[1, 0], [0, self.exponented]])

    def __repr__(self) -> str:
        return "R({})".format(",".join(map(repr, self.inputs)))


class RmOp(ROp):
    """R with phi = 2 pi / 2^m, negated when negate=True (a keyword the op wrapper drops)."""

    def __init__(self, m: int, *inputs: Qubit, negate: bool = False, **kwargs):
        sign = -1.0 if negate else 1.0
        super().__init__(sign * 2.0 * numpy.pi / 2.0 ** m, *inputs, **kwargs)
        self.m = m

    def __repr__(self) -> str:
        return "R[{}]({})".format(self.m, ",".join(map(repr, self.inputs)))


class SwapOp(MatrixOp):
    """Exchange two equal-size registers: one SwapMat on the concatenated index groups."""

    def __init__(self, *inputs: Qubit, **kwargs):
        super().__init__(*inputs, **kwargs)
        if len(self.inputs) != 2:
            raise Exception("Swap can only take two inputs")
        a, b = self.inputs
        if a.n != b.n:
            raise Exception("Inputs must be of equal size {}/{}".format(a, b))

    def makemats(self, index_groups):
        first, second = index_groups[0], index_groups[1]
        return {tuple(list(first) + list(second)): SwapMat(self.inputs[0].n)}

    def __repr__(self) -> str:
        return "Swap({})".format(self.qid)


class SwapMat(object):
    """The 2^(2n) x 2^(2n) permutation exchanging the high and the low n-bit halves of an index:
    entry (i, j) is 1 exactly when j is i with its halves exchanged (qip/operators.py:157-179)."""
    _kron_struct = 3

    def __init__(self, n: int):
        self.n = n
        self.shape = (2 ** (2 * n), 2 ** (2 * n))

    def partner(self, i: int) -> int:
        half = 1 << self.n
        hi, lo = divmod(i, half)
        return lo * half + hi

    def __getitem__(self, item: Union[int, Tuple[int, int]]) -> complex:
        if not (type(item) == tuple and len(item) == 2):
            raise ValueError("SwapMat can only be indexed with M[i,j] not M[{}]".format(item))
        return 1.0 if self.partner(item[0]) == item[1] else 0.0

    def __repr__(self):
        return "SwapMat({})".format(self.n)


def _dense(m) -> numpy.ndarray:
    if isinstance(m, (CMat, SwapMat)):
        if isinstance(m, SwapMat):
            d = m.shape[0]
            out = numpy.zeros((d, d), dtype=numpy.complex128)
            out[numpy.arange(d), [m.partner(i) for i in range(d)]] = 1.0
            return out
        return m.numpy()
    return numpy.asarray(m, dtype=numpy.complex128)


class CMat(object):
    """diag(I, m): the controlled form of m, the control as the most significant bit. m may be
    another CMat (several controls) or a SwapMat (qip/operators.py:233-268)."""
    _kron_struct = 2

    def __init__(self, mat: Sequence[Sequence[complex]]):
        self.m = numpy.array(mat) if type(mat) == list else mat
        rows, cols = self.m.shape
        self.shape = (2 * rows, 2 * cols)

    def __getitem__(self, item: Union[int, Tuple[int, int]]):
        if not (type(item) == tuple and len(item) == 2):
            raise ValueError("CMat can only be indexed with M[i,j] not M[{}]".format(item))
        half_r, half_c = self.shape[0] // 2, self.shape[1] // 2
        (hr, r), (hc, c) = divmod(item[0], half_r), divmod(item[1], half_c)
        if hr != hc:
            return 0.0
        if hr == 0:
            return 1.0 if r == c else 0.0
        return self.m[r, c]

    def conj(self):
        return CMat(self.m.conj())

    @property
    def T(self):
        return CMat(self.m.T)

    def numpy(self) -> numpy.ndarray:
        inner = _dense(self.m)
        h = inner.shape[0]
        out = numpy.eye(2 * h, dtype=numpy.complex128)
        out[h:, h:] = inner
        return out


def C(op: OpConstructor) -> OpConstructor:
    """The controlled version of an op: its first input is the control qubit. An op that handles
    wrappers itself (a QubitFuncWrapper) takes the control through wrap_op_hook."""
    hook = getattr(op, "wrap_op_hook", None)
    own = hook(C, consumed_inputs=[0]) if hook is not None else None
    return own if own is not None else QubitOpWrapper(COp, op)


class COp(MatrixOp):
    """Controlled op: every matrix M of the inner op becomes CMat(M) on (control, its group). The
    node takes the inner op's qid (qip/operators.py:195-230)."""

    def __init__(self, op: OpConstructor, *inputs: Qubit, **kwargs):
        if len(inputs) < 2:
            raise ValueError("Not enough input values given.")
        if inputs[0].n != 1:
            raise ValueError("Control bit can only be of size n=1")
        if len(set(inputs)) < len(inputs):
            raise ValueError("Cannot have repeated qubits in COp")
        # the inner op is built detached: it registers no sinks of its own and is not split
        inner_kwargs = dict(kwargs, nosink=True, nosplit=True)
        self.op = self.make_op(op, inputs, inner_kwargs)
        if not isinstance(self.op, MatrixOp):
            raise ValueError("C(Op) may only be applied to matrix ops and swaps, not {}".format(type(self.op)))
        super().__init__(*inputs, qid=self.op.qid, **kwargs)

    def make_op(self, op: OpConstructor, inputs: Sequence[Qubit], nkwargs: Mapping[str, Any]):
        return op(*inputs[1:], **nkwargs)

    def makemats(self, index_groups):
        control = list(index_groups[0])
        inner = self.op.makemats(index_groups[1:])
        out = {}
        for group, m in inner.items():
            tail = [group] if isinstance(group, int) else list(group)
            out[tuple(control + tail)] = CMat(m)
        return out

    def __repr__(self):
        return "C{}".format(self.op)


class FOp(Qubit):
    """|x>|q> -> |x>|func(x) xor q> for registers reg1, reg2 (qip/operators.py:271-298). The qubit
    count is passed as func_apply's fourth positional argument, as the reference's FOp does."""

    def __init__(self, func: Callable[[int], int], reg1: Qubit, reg2: Qubit, **kwargs):
        super().__init__(reg1, reg2, **kwargs)
        self.func = func
        self.reg1 = reg1
        self.reg2 = reg2

    def feed_indices(self, state: StateType, index_groups: Sequence[Sequence[int]], n: int) -> Tuple[int, int]:
        x_reg = numpy.array(index_groups[0], dtype=numpy.int32)
        y_reg = numpy.array(index_groups[1], dtype=numpy.int32)
        state.func_apply(x_reg, y_reg, self.func, n)
        return 0, 0

    def __repr__(self) -> str:
        return "F({}, {})".format(self.reg1, self.reg2)


Not = QubitOpWrapper(NotOp)
X = QubitOpWrapper(XOp)
Y = QubitOpWrapper(YOp)
Z = QubitOpWrapper(ZOp)
H = QubitOpWrapper(HOp)
R = QubitOpWrapper(ROp)
Rm = QubitOpWrapper(RmOp)
Swap = QubitOpWrapper(SwapOp)
F = QubitOpWrapper(FOp)
