"""The descriptor planner on the CPU: the engine calls it plans for one kronselect_dot descriptor,
applied densely with numpy, equal the oracle's restatement of cdot_loop (qip/ext/kronprod.pyx:41-235)
on the same state. Covers a control shared by every group (peeled, possibly several levels deep),
controlled single-qubit factors merged into Kronecker blocks, SWAP_MAT under controls, overlapping
groups expanded densely, components left to the engine's column sum, and scalar groups."""
import numpy
import pytest

from oracle import oracle
from qip_amd import CMat, SwapMat, planner

RNG = numpy.random.default_rng(31337)


def rand_u(d):
    z = (RNG.standard_normal((d, d)) + 1j * RNG.standard_normal((d, d))) / numpy.sqrt(2)
    q, r = numpy.linalg.qr(z)
    return q * (numpy.diagonal(r) / numpy.abs(numpy.diagonal(r)))


def apply_dense(psi, n, targets, controls, mat):
    """mat on `targets` (first = MSB), where every control qubit is 1; qubit q is axis q."""
    t = psi.reshape((2,) * n).copy()
    k = len(targets)
    sel = [slice(None)] * n
    for c in controls:
        sel[c] = 1
    sub = t[tuple(sel)]
    # axes of the targets inside the controlled sub-array
    rest = [q for q in range(n) if q not in controls]
    axes = [rest.index(q) for q in targets]
    moved = numpy.moveaxis(sub, axes, list(range(k)))
    shp = moved.shape
    out = (numpy.asarray(mat).reshape(2 ** k, 2 ** k) @ moved.reshape(2 ** k, -1)).reshape(shp)
    t[tuple(sel)] = numpy.moveaxis(out, list(range(k)), axes)
    return t.reshape(-1)


def apply_plan(mats, psi, n):
    v = psi.copy()
    for call in planner.plan(mats):
        if isinstance(call, planner.ProductCall):
            for q, m in zip(call.targets, call.mats):
                v = apply_dense(v, n, [q], [], m)
        elif isinstance(call, planner.GeneralCall):
            out = numpy.zeros_like(v)
            oracle.kronselect_dot(dict(call.groups), v, n, out)
            v = out
        else:
            v = apply_dense(v, n, call.targets, call.controls, call.mat)
    return v


def check(mats, n, kinds=None):
    psi = RNG.standard_normal(2 ** n) + 1j * RNG.standard_normal(2 ** n)
    want = numpy.zeros_like(psi)
    oracle.kronselect_dot(mats, psi, n, want)
    got = apply_plan(mats, psi, n)
    assert numpy.max(numpy.abs(got - want)) < 1e-12
    if kinds is not None:
        assert sorted(type(c).__name__ for c in planner.plan(mats)) == sorted(kinds)


X = numpy.array([[0.0, 1.0], [1.0, 0.0]])
H = numpy.array([[1.0, 1.0], [1.0, -1.0]]) / numpy.sqrt(2.0)


def test_shared_control_is_peeled_into_merged_blocks():
    mats = {(0, i): CMat(X) for i in range(1, 8)}
    calls = planner.plan(mats)
    assert [(c.targets, c.controls) for c in calls] == [([1, 2, 3], [0]), ([4, 5, 6], [0]), ([7], [0])]
    check(mats, 9)


def test_two_levels_of_shared_controls():
    mats = {(4, 0, q): CMat(CMat(H)) for q in (1, 2, 3, 5, 6)}
    calls = planner.plan(mats)
    assert all(sorted(c.controls) == [0, 4] for c in calls)
    assert sorted(q for c in calls for q in c.targets) == [1, 2, 3, 5, 6]
    check(mats, 8)


def test_shared_control_over_swap_and_dense_groups():
    """C(Op) of an op whose descriptor has several groups of different kinds."""
    mats = {(0, 1, 2, 3, 4): CMat(SwapMat(2)), (0, 5, 6): CMat(rand_u(4)), (0, 7): CMat(H)}
    check(mats, 9, kinds=["MatrixCall"] * 4)


def test_control_that_is_not_shared_by_every_group():
    """(c, t1) controlled and (c, t2) plain share c without a common control: dense expansion."""
    mats = {(0, 1): CMat(X), (0, 2): rand_u(4)}
    calls = planner.plan(mats)
    assert len(calls) == 1 and sorted(calls[0].targets) == [0, 1, 2]
    check(mats, 5)


def test_overlapping_groups_beyond_the_expansion_go_to_the_column_sum():
    mats = {(i, i + 1): rand_u(4) for i in range(11)}
    check(mats, 12, kinds=["GeneralCall"])


def test_controlled_overlapping_groups_keep_the_peeled_controls():
    """A shared control peeled, then an overlapping chain over 11 qubits under it: the column sum
    gets the control back as a C_MAT level on every group."""
    inner = {(i, i + 1): rand_u(4) for i in range(1, 11)}
    mats = {(0,) + k: CMat(m) for k, m in inner.items()}
    calls = planner.plan(mats)
    assert len(calls) == 1 and isinstance(calls[0], planner.GeneralCall)
    assert all(key[0] == 0 for key, _ in calls[0].groups)
    check(mats, 12)


def test_wide_dense_matrix_goes_to_the_column_sum():
    mats = {tuple(range(13)): numpy.eye(2 ** 13) * 0.5 + 0.5 * numpy.diag(numpy.ones(2 ** 13 - 1), 1)}
    assert [type(c).__name__ for c in planner.plan(mats)] == ["GeneralCall"]


def test_scalar_group_under_a_control():
    """CMat of a 1x1 matrix: a phase on the control's 1 subspace, also inside a shared control."""
    mats = {(2,): CMat(numpy.array([[1j]]))}
    check(mats, 4)
    mats = {(0, 1): CMat(CMat(numpy.array([[-1.0]]))), (0, 3): CMat(H)}
    check(mats, 5)


def test_general_call_packing():
    g = planner.GeneralCall([((0, 1, 2), CMat(CMat(H))), ((2, 3, 4, 5), SwapMat(2)), ((5, 6), rand_u(4))])
    lens, idx, levels, kinds, params, mats = g.packed()
    assert lens == [3, 4, 2] and idx == [0, 1, 2, 2, 3, 4, 5, 5, 6]
    assert levels == [2, 0, 0] and kinds == [1, 3, 1] and params == [0, 2, 0]
    assert mats.shape == (4 + 16,)


@pytest.mark.parametrize("trial", range(25))
def test_random_descriptors(trial):
    """Random descriptors of up to 4 groups on up to 9 qubits, C_MAT chains of depth up to 2,
    SWAP_MAT and dense matrices, groups disjoint or sharing qubits."""
    n = int(RNG.integers(4, 10))
    mats = {}
    used = []
    for _ in range(int(RNG.integers(1, 5))):
        levels = int(RNG.integers(0, 3))
        kind = "swap" if RNG.random() < 0.25 else "dense"
        inner = 2 if kind == "swap" else int(RNG.integers(1, 3))
        size = levels + inner
        if size > n:
            continue
        if used and RNG.random() < 0.5:
            share = [int(RNG.choice(used))]
            others = [q for q in RNG.permutation(n) if q not in share][:size - 1]
            key = tuple(share + [int(q) for q in others]) if RNG.random() < 0.5 else tuple([int(q) for q in others] + share)
        else:
            key = tuple(int(q) for q in RNG.permutation(n)[:size])
        if key in mats:
            continue
        m = SwapMat(1) if kind == "swap" else rand_u(2 ** inner)
        for _ in range(levels):
            m = CMat(m)
        mats[key] = m
        used.extend(key)
    if not mats:
        mats = {0: H}
    check(mats, n)


def test_descriptor_keeps_its_plan_while_unchanged():
    """A Descriptor with read-only matrices is planned once; the kept calls are the fresh plan's;
    a change through the dict drops them; a writeable or list matrix is planned on every use."""
    m1, m2 = rand_u(4), rand_u(2)
    d = planner.Descriptor({(0, 1, 2): CMat(m1), 3: m2, (4, 5): SwapMat(1)})
    assert planner.plan(d) is not planner.plan(d)           # writeable matrices: nothing kept
    assert d._plan is None
    for m in d.values():
        planner.freeze(m)
    assert not m1.flags.writeable and not m2.flags.writeable
    first = planner.plan(d)
    kept = d._plan
    assert kept is not None
    again = planner.plan(d)
    assert again is not first and list(kept) == again        # a new list of the same calls
    assert [repr(c) for c in again] == [repr(c) for c in planner._plan(dict(d))]
    with pytest.raises(ValueError):
        m2[0, 0] = 0.0                                        # a read-only matrix cannot change under it
    for change in (lambda: d.__setitem__(6, H.copy()), lambda: d.pop(6), lambda: d.update({7: X}),
                   lambda: d.setdefault(8, X), lambda: d.popitem(), lambda: d.__delitem__(3)):
        planner.plan(d)
        change()
        assert d._plan is None
    d2 = planner.Descriptor({0: [[0.0, 1.0], [1.0, 0.0]]})  # a list matrix
    planner.plan(d2)
    assert d2._plan is None


def test_graph_gate_descriptors_are_frozen():
    """The frontend's gate nodes hand out Descriptors with read-only matrices, remade when the
    node's index groups change."""
    from qip_amd import C, H as HOp, Qubit
    from qip_amd.operators import MatrixOp
    op = C(HOp)(Qubit(n=1), Qubit(n=2))[1]
    node = op if isinstance(op, MatrixOp) else op.inputs[0]
    assert isinstance(node, MatrixOp)
    d = node.descriptor([[0], [1, 2]])
    assert type(d) is planner.Descriptor
    assert planner.plan(d) and d._plan is not None
    assert node.descriptor([[0], [1, 2]]) is d
    d3 = node.descriptor([[3], [4, 5]])
    assert d3 is not d and d3._plan is None
