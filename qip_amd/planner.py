"""Operator-descriptor planning: a kronselect_dot `mats` dict -> engine calls.

The reference applies a whole descriptor {indices: matrix} in one out-of-place sweep whose cost
is 2^k columns per output row for the k distinct qubits involved (qip/ext/kronprod.pyx:157-197).
Its operator (kronprod.pyx:164-196) is

    out[r] = sum over c agreeing with r outside U of  prod_g M_g[r|_g, c|_g] * in[c]

with U the union of all index groups. Groups that share no qubit give a Kronecker product, i.e.
independent commuting factors; the planner splits the descriptor into such connected components
and emits one in-place engine call per component:

  * a C_MAT chain (operators.py:233-268, kronprod.pyx:215-225) becomes control qubits around its
    innermost matrix;
  * a SWAP_MAT(m) (operators.py:157-179, kronprod.pyx:227-231) over registers a, b becomes m
    two-qubit swaps a[i] <-> b[i] (same significance), each under the group's controls;
  * a numpy matrix is applied densely on its qubits;
  * a component of several groups sharing qubits (e.g. C(Not) on a 2-qubit register gives
    {(c,t1): CMat(X), (c,t2): CMat(X)}, operators.py:41-43,221-227) is expanded into its dense
    operator over the union by the product formula above; the engine then recovers its controls.
"""
from typing import Dict, List, Mapping, Sequence, Tuple

import numpy

KRON_NUMPY, KRON_CMAT, KRON_SWAP = 1, 2, 3          # qip/ext/kronprod.pyx:18-19
MAX_EXPANDED_QUBITS = 10                             # dense expansion of shared-qubit components
KRON_MERGE_WIDTH = 3                                 # single-qubit factors merged per sweep

_SWAP4 = numpy.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=numpy.complex128)


class MatrixCall:
    """One engine call: dense `mat` (2^k x 2^k, first target = MSB) on `targets`, applied where
    every `controls` qubit is 1."""
    __slots__ = ("targets", "controls", "mat")

    def __init__(self, targets: Sequence[int], controls: Sequence[int], mat: numpy.ndarray):
        self.targets = list(targets)
        self.controls = list(controls)
        self.mat = mat

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


def plan(mats: Mapping) -> List[MatrixCall]:
    """Engine calls equivalent to one kronselect_dot(mats)."""
    newmats = normalise(mats)
    items = []
    for key, m in newmats.items():
        if type(m) == numpy.ndarray:
            m = m.astype(numpy.complex128)      # qip/util.py:63
        _kind(m)
        items.append((key, m))
    # connected components over shared qubits (union-find), in first-appearance order
    parent = {}

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
    comps: Dict[int, List[Tuple[Tuple[int, ...], object]]] = {}
    order = []
    for key, m in items:
        root = find(key[0])
        if root not in comps:
            comps[root] = []
            order.append(root)
        comps[root].append((key, m))

    calls: List[MatrixCall] = []
    for root in order:
        groups = comps[root]
        if len(groups) == 1:
            key, m = groups[0]
            controls, idx, kind, inner = _peel(key, m)
            if kind == KRON_SWAP:
                w = inner.n
                if len(idx) != 2 * w:
                    raise Exception("Shape of square submatrix must equal 2**(number of indices): "
                                    "{}: {}".format(key, m))
                for a, b in zip(idx[:w], idx[w:]):
                    calls.append(MatrixCall([a, b], controls, _SWAP4))
            else:
                calls.append(MatrixCall(idx, controls, numpy.ascontiguousarray(inner, dtype=numpy.complex128)))
        else:
            union, dense = _expand_component(groups)
            calls.append(MatrixCall(union, [], dense))
    return merge_single_qubit(calls)


class ProductCall:
    """The uncontrolled single-qubit factors of one descriptor (distinct qubits, so they commute),
    applied together: mats[j] (2x2) on targets[j]."""
    __slots__ = ("targets", "mats")
    controls = ()

    def __init__(self, targets: Sequence[int], mats: Sequence[numpy.ndarray]):
        self.targets = list(targets)
        self.mats = list(mats)

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


def merge_single_qubit(calls: List[MatrixCall]) -> list:
    """Collect the uncontrolled single-qubit factors of one descriptor into one ProductCall (one
    sweep for all of them; the reference applies the whole descriptor in one sweep too,
    kronprod.pyx:157-197)."""
    singles = [c for c in calls if len(c.targets) == 1 and not c.controls]
    if len(singles) < 2:
        return calls
    rest = [c for c in calls if not (len(c.targets) == 1 and not c.controls)]
    return rest + [ProductCall([c.targets[0] for c in singles], [c.mat for c in singles])]
