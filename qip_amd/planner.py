"""Operator-descriptor planning: a kronselect_dot `mats` dict -> engine calls.

The reference applies a whole descriptor {indices: matrix} in one out-of-place sweep whose cost
is 2^k columns per output row for the k distinct qubits involved (qip/ext/kronprod.pyx:157-197).
Its operator (kronprod.pyx:164-196) is

    out[r] = sum over c agreeing with r outside U of  prod_g M_g[r|_g, c|_g] * in[c]

with U the union of all index groups. Groups that share no qubit give a Kronecker product, i.e.
independent commuting factors; the planner splits the descriptor into such connected components
and emits in-place engine calls per component:

  * a C_MAT chain (operators.py:233-268, kronprod.pyx:215-225) becomes control qubits around its
    innermost matrix;
  * a SWAP_MAT(m) (operators.py:157-179, kronprod.pyx:227-231) over registers a, b becomes m
    two-qubit swaps a[i] <-> b[i] (same significance), each under the group's controls;
  * a numpy matrix is applied densely on its qubits (up to MAX_DENSE_TARGETS);
  * a component whose groups all begin with the same C_MAT control q (C(Op) on a register gives
    {(c,t1): CMat(X), (c,t2): CMat(X)}, operators.py:41-43,221-227) is |0><0|_q (x) I +
    |1><1|_q (x) O', O' the descriptor of the groups with that level peeled: q becomes a control of
    every call planned for O', whose groups may now fall apart into components;
  * any other component of several groups sharing qubits is expanded into its dense operator over
    the union by the product formula above (up to MAX_EXPANDED_QUBITS); the engine then recovers
    its controls;
  * what is left, and numpy matrices wider than MAX_DENSE_TARGETS, go to the engine's restatement
    of the formula itself (GeneralCall).
Every call of one descriptor acts on qubits no other call of it changes under different controls,
so the calls commute and their order is free.
"""
from typing import Dict, List, Mapping, Sequence, Tuple

import numpy

KRON_NUMPY, KRON_CMAT, KRON_SWAP = 1, 2, 3          # qip/ext/kronprod.pyx:18-19
MAX_EXPANDED_QUBITS = 10                             # dense expansion of shared-qubit components
MAX_DENSE_TARGETS = 12                               # QIPGPU_MAX_DENSE
KRON_MERGE_WIDTH = 3                                 # single-qubit factors merged per sweep

_SWAP4 = numpy.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=numpy.complex128)


class MatrixCall:
    """One engine call: dense `mat` (2^k x 2^k, first target = MSB) on `targets`, applied where
    every `controls` qubit is 1. `packed` holds a backend's ready-made call arguments."""
    __slots__ = ("targets", "controls", "mat", "packed")

    def __init__(self, targets: Sequence[int], controls: Sequence[int], mat: numpy.ndarray):
        self.targets = list(targets)
        self.controls = list(controls)
        self.mat = mat
        self.packed = None

    def __repr__(self):
        return "MatrixCall(t={}, c={})".format(self.targets, self.controls)


def normalise(mats: Mapping) -> Dict[Tuple[int, ...], object]:
    """qip/util.py:35-58: int keys become 1-tuples, list matrices become arrays, and every matrix
    must be square with side 2^(number of indices). Same exceptions and messages."""
    newmats = {}
    for indices in mats:
        if type(indices) != tuple and type(indices) != int:
            raise Exception("Type of indices must be tuple: {}".format(indices))
        m = mats[indices]
        if type(m) == list:
            m = numpy.array(m)
        key = indices if type(indices) == tuple else (indices,)
        newmats[key] = m
        if 2 ** len(key) != m.shape[0] or 2 ** len(key) != m.shape[1]:
            raise Exception("Shape of square submatrix must equal 2**(number of indices): "
                            "{}: {}".format(indices, m))
    return newmats


def _kind(m) -> int:
    """The MatStruct type of a matrix object (kronprod.pyx:89-116)."""
    if hasattr(m, "_kron_struct"):
        kind = m._kron_struct
    elif type(m) == numpy.ndarray:
        kind = KRON_NUMPY
    else:
        kind = 0
    if kind not in (KRON_NUMPY, KRON_CMAT, KRON_SWAP):
        raise ValueError("Cannot pass matrices which are not numpy, SwapMat, or CMat")
    return kind


def _peel(indices: Tuple[int, ...], m):
    """Strip a C_MAT chain: (controls, inner indices, inner kind, inner matrix)."""
    controls = []
    idx = list(indices)
    kind = _kind(m)
    while kind == KRON_CMAT:
        controls.append(idx[0])
        idx = idx[1:]
        m = m.m
        kind = _kind(m)
    return controls, idx, kind, m


def dense_of(m) -> numpy.ndarray:
    """Dense complex128 matrix of a NUMPY / C_MAT / SWAP_MAT object, entry by entry the
    calc_mat_entry rules (kronprod.pyx:205-235)."""
    kind = _kind(m)
    if kind == KRON_NUMPY:
        return numpy.ascontiguousarray(m, dtype=numpy.complex128)
    if kind == KRON_CMAT:
        inner = dense_of(m.m)
        d = inner.shape[0]
        out = numpy.zeros((2 * d, 2 * d), dtype=numpy.complex128)
        out[:d, :d] = numpy.eye(d)
        out[d:, d:] = inner
        return out
    w = m.n
    dim = 1 << (2 * w)
    i = numpy.arange(dim)
    low, high = i & ((1 << w) - 1), i >> w
    partner = (low << w) | high
    out = numpy.zeros((dim, dim), dtype=numpy.complex128)
    out[i, partner] = 1.0
    return out


def _expand_component(groups: List[Tuple[Tuple[int, ...], object]]) -> Tuple[List[int], numpy.ndarray]:
    """Dense operator over the union of the groups' qubits (ascending, first = MSB):
    O[r, c] = prod_g D_g[r|_g, c|_g] (kronprod.pyx:174-192)."""
    union = sorted(set(q for idx, _ in groups for q in idx))
    u = len(union)
    if u > MAX_EXPANDED_QUBITS:
        raise NotImplementedError("operator with shared indices spans {} qubits (at most {})"
                                  .format(u, MAX_EXPANDED_QUBITS))
    pos = {q: u - 1 - j for j, q in enumerate(union)}       # bit of the union index
    dim = 1 << u
    r = numpy.arange(dim)[:, None]
    c = numpy.arange(dim)[None, :]
    out = numpy.ones((dim, dim), dtype=numpy.complex128)
    for idx, m in groups:
        d = dense_of(m)
        L = len(idx)
        sr = numpy.zeros_like(r)
        sc = numpy.zeros_like(c)
        for j, q in enumerate(idx):
            sr |= ((r >> pos[q]) & 1) << (L - 1 - j)
            sc |= ((c >> pos[q]) & 1) << (L - 1 - j)
        out = out * d[sr, sc]
    return union, out


class GeneralCall:
    """A descriptor component applied by the engine's restatement of the reference's formula:
    `groups` as (indices, matrix) with C_MAT chains and SWAP_MAT kept."""
    __slots__ = ("groups",)
    controls = ()

    def __init__(self, groups: Sequence[Tuple[Tuple[int, ...], object]]):
        self.groups = list(groups)

    def packed(self):
        """(group lengths, indices, C_MAT levels, inner kinds, SWAP_MAT widths, inner NUMPY matrices
        concatenated as complex128) for qipgpu_apply_kron."""
        lens, idx, levels, kinds, params, mats = [], [], [], [], [], []
        for key, m in self.groups:
            peeled, inner_idx, kind, inner = _peel(tuple(key), m)
            lens.append(len(key))
            idx.extend(key)
            levels.append(len(peeled))
            kinds.append(kind)
            if kind == KRON_SWAP:
                if len(inner_idx) != 2 * inner.n:
                    raise Exception("Shape of square submatrix must equal 2**(number of indices): "
                                    "{}: {}".format(key, m))
                params.append(inner.n)
            else:
                params.append(0)
                mats.append(numpy.ascontiguousarray(inner, dtype=numpy.complex128).ravel())
        flat = numpy.concatenate(mats) if mats else numpy.zeros(1, dtype=numpy.complex128)
        return lens, idx, levels, kinds, params, flat

    @property
    def targets(self) -> List[int]:
        return sorted(set(q for idx, _ in self.groups for q in idx))

    def __repr__(self):
        return "GeneralCall(qubits={})".format(self.targets)


def _components(items: List[Tuple[Tuple[int, ...], object]]) -> List[List[Tuple[Tuple[int, ...], object]]]:
    """Groups joined by shared qubits (union-find), components in first-appearance order."""
    parent: Dict[int, int] = {}

    def find(q):
        while parent[q] != q:
            parent[q] = parent[parent[q]]
            q = parent[q]
        return q

    for key, _ in items:
        for q in key:
            parent.setdefault(q, q)
        for q in key[1:]:
            a, b = find(key[0]), find(q)
            if a != b:
                parent[b] = a
    comps: Dict[object, List[Tuple[Tuple[int, ...], object]]] = {}
    for n, (key, m) in enumerate(items):
        # a group without indices (a 1x1 matrix under C_MAT levels) is a scalar: its own component
        comps.setdefault(find(key[0]) if key else ("scalar", n), []).append((key, m))
    return list(comps.values())


def _wrap_controls(controls: Sequence[int], groups):
    """The groups with the peeled controls put back as C_MAT levels (for a GeneralCall)."""
    out = []
    for key, m in groups:
        for c in reversed(controls):
            key = (c,) + tuple(key)
            m = _CMatLevel(m)
        out.append((tuple(key), m))
    return out


class _CMatLevel:
    """A C_MAT level around m (same _kron_struct, m and shape as operators.CMat)."""
    _kron_struct = KRON_CMAT

    def __init__(self, m):
        self.m = m
        rows, cols = m.shape
        self.shape = (2 * rows, 2 * cols)


def _plan_component(groups, controls: List[int]) -> list:
    if len(groups) == 1:
        key, m = groups[0]
        peeled, idx, kind, inner = _peel(key, m)
        ctl = controls + peeled
        if kind == KRON_SWAP:
            w = inner.n
            if len(idx) != 2 * w:
                raise Exception("Shape of square submatrix must equal 2**(number of indices): "
                                "{}: {}".format(key, m))
            return [MatrixCall([a, b], ctl, _SWAP4) for a, b in zip(idx[:w], idx[w:])]
        if len(idx) <= MAX_DENSE_TARGETS:
            return [MatrixCall(idx, ctl, numpy.ascontiguousarray(inner, dtype=numpy.complex128))]
        return [GeneralCall(_wrap_controls(controls, groups))]
    heads = set(key[0] if _kind(m) == KRON_CMAT else None for key, m in groups)
    if len(heads) == 1 and None not in heads:
        q = heads.pop()
        if all(q not in key[1:] for key, _ in groups):
            inner = [(tuple(key[1:]), m.m) for key, m in groups]
            calls = []
            for comp in _components(inner):
                calls.extend(_plan_component(comp, controls + [q]))
            return calls
    union = set(q for key, _ in groups for q in key)
    if len(union) <= MAX_EXPANDED_QUBITS:
        qubits, dense = _expand_component(groups)
        return [MatrixCall(qubits, controls, dense)]
    return [GeneralCall(_wrap_controls(controls, groups))]


class Descriptor(dict):
    """A descriptor that keeps its plan: the graph frontend's gate nodes hand these to the backend
    and do not change them once made, and their numpy matrices are read-only, so the calls planned
    for the first run serve every later one. A change through the dict's own methods drops the
    kept plan; a descriptor holding a matrix that is not read-only (a list, a writeable array) is
    planned on every use."""
    __slots__ = ("_plan",)

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self._plan = None

    def __setitem__(self, key, value):
        self._plan = None
        super().__setitem__(key, value)

    def __delitem__(self, key):
        self._plan = None
        super().__delitem__(key)

    def __ior__(self, other):
        self._plan = None
        return super().__ior__(other)

    def clear(self):
        self._plan = None
        super().clear()

    def pop(self, *args):
        self._plan = None
        return super().pop(*args)

    def popitem(self):
        self._plan = None
        return super().popitem()

    def setdefault(self, *args):
        self._plan = None
        return super().setdefault(*args)

    def update(self, *args, **kwargs):
        self._plan = None
        super().update(*args, **kwargs)


def freeze(m) -> None:
    """Make the numpy matrix inside a descriptor entry (through C_MAT levels) read-only."""
    while getattr(m, "_kron_struct", None) == KRON_CMAT:
        m = m.m
    if type(m) == numpy.ndarray:
        m.flags.writeable = False


def _frozen(m) -> bool:
    while getattr(m, "_kron_struct", None) == KRON_CMAT:
        m = m.m
    if getattr(m, "_kron_struct", None) == KRON_SWAP:
        return True
    return type(m) == numpy.ndarray and not m.flags.writeable


def plan(mats: Mapping) -> list:
    """Engine calls equivalent to one kronselect_dot(mats): MatrixCall, ProductCall and GeneralCall.
    A Descriptor whose matrices are all read-only keeps its calls for the next use."""
    if type(mats) is Descriptor:
        if mats._plan is None:
            calls = _plan(mats)
            if all(_frozen(m) for m in mats.values()):
                mats._plan = tuple(calls)
            return calls
        return list(mats._plan)
    return _plan(mats)


def _plan(mats: Mapping) -> list:
    newmats = normalise(mats)
    items = []
    for key, m in newmats.items():
        if type(m) == numpy.ndarray:
            m = numpy.asarray(m, dtype=numpy.complex128)      # qip/util.py:63 (read-only here)
        _kind(m)
        items.append((key, m))
    if len(items) == 1 and items[0][0]:
        # one group is its own component, and its calls have nothing to merge with
        return _plan_component(items, [])
    calls: list = []
    for comp in _components(items):
        calls.extend(_plan_component(comp, []))
    return merge_single_qubit(calls)


class ProductCall:
    """The uncontrolled single-qubit factors of one descriptor (distinct qubits, so they commute),
    applied together: mats[j] (2x2) on targets[j]. `packed` holds a backend's ready-made call
    arguments."""
    __slots__ = ("targets", "mats", "packed")
    controls = ()

    def __init__(self, targets: Sequence[int], mats: Sequence[numpy.ndarray]):
        self.targets = list(targets)
        self.mats = list(mats)
        self.packed = None

    def matrix_calls(self, width: int = KRON_MERGE_WIDTH) -> List[MatrixCall]:
        """The same product as Kronecker blocks on up to `width` qubits each."""
        out = []
        for i in range(0, len(self.targets), width):
            mat = self.mats[i]
            for m in self.mats[i + 1:i + width]:
                mat = numpy.kron(mat, m)
            out.append(MatrixCall(self.targets[i:i + width], [], mat))
        return out

    def __repr__(self):
        return "ProductCall(t={})".format(self.targets)


def merge_single_qubit(calls: list) -> list:
    """Collect the uncontrolled single-qubit factors of one descriptor into one ProductCall (one
    sweep for all of them; the reference applies the whole descriptor in one sweep too,
    kronprod.pyx:157-197), and the single-qubit factors under one control set into Kronecker
    blocks of up to KRON_MERGE_WIDTH qubits under those controls."""
    singles, by_ctl, rest = [], {}, []
    for c in calls:
        if isinstance(c, MatrixCall) and len(c.targets) == 1:
            if not c.controls:
                singles.append(c)
            else:
                by_ctl.setdefault(tuple(sorted(c.controls)), []).append(c)
        else:
            rest.append(c)
    out = rest
    for ctl, group in by_ctl.items():
        if len(group) == 1:
            out.append(group[0])
            continue
        for i in range(0, len(group), KRON_MERGE_WIDTH):
            part = group[i:i + KRON_MERGE_WIDTH]
            mat = part[0].mat
            for c in part[1:]:
                mat = numpy.kron(mat, c.mat)
            out.append(MatrixCall([c.targets[0] for c in part], list(ctl), mat))
    if len(singles) >= 2:
        out.append(ProductCall([c.targets[0] for c in singles], [c.mat for c in singles]))
    else:
        out.extend(singles)
    return out
