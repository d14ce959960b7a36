"""CPU tests of the precision form's host logic: greedy fusion keeps program semantics (the
ordered product of block matrices equals the ordered product of the gates, both expanded to the
full 2^n operator independently), respects the width, and gives the expected block counts."""
import numpy
import pytest

from qip_amd import planner
from qip_amd.fused import controlled_dense, fuse
from qip_amd.workloads import algorithmic_bytes_plain, gate_counts, layered_random_circuit, touched_bytes_plain


def full_operator(gq, G, n):
    """2^n operator of gate G on qubits gq (first = MSB of G's index) by tensor contraction."""
    dim = 2 ** n
    L = len(gq)
    eye = numpy.eye(dim, dtype=complex).reshape([2] * n + [dim])
    out = numpy.tensordot(G.reshape([2] * (2 * L)), eye, axes=(list(range(L, 2 * L)), list(gq)))
    return numpy.moveaxis(out, list(range(L)), list(gq)).reshape(dim, dim)


def gates_of(ops):
    gates = []
    for d in ops:
        for call in planner.plan(d):
            if isinstance(call, planner.ProductCall):
                gates.extend(((q,), m) for q, m in zip(call.targets, call.mats))
            else:
                gates.append(controlled_dense(call))
    return gates


@pytest.mark.parametrize("width", [2, 3, 4, 6])
def test_fusion_preserves_operator(width):
    n, depth = 7, 4
    gates = gates_of(layered_random_circuit(n, depth, seed=5))
    want = numpy.eye(2 ** n, dtype=complex)
    for gq, G in gates:
        want = full_operator(gq, G, n) @ want
    blocks = fuse(gates, width)
    got = numpy.eye(2 ** n, dtype=complex)
    for b in blocks:
        assert len(b.qubits) <= width
        qs, M = b.matrix()
        got = full_operator(qs, M, n) @ got
    assert numpy.max(numpy.abs(got - want)) < 1e-12
    assert sum(len(b.gates) for b in blocks) == len(gates)


def test_block_counts_brickwork():
    gates = gates_of(layered_random_circuit(30, 40, seed=1))
    assert len(gates) == 1780
    assert len(fuse(gates, 6)) == 128


def test_circuit_shape():
    ops = layered_random_circuit(30, 40, seed=2024)
    assert gate_counts(ops) == (1200, 580)
    # the per-gate definition 2 S 2^(n - c), c = controls as written: 47,680 GiB and 5,560 GiB
    assert algorithmic_bytes_plain(ops, 30, 16) == (1200 * 32 + 580 * 16) * 2 ** 30 == 47680 * 2 ** 30
    assert touched_bytes_plain(ops, 30, 16) == (1200 * 32 + 300 * 16 + 280 * 8) * 2 ** 30
    ops2 = layered_random_circuit(28, 20, seed=1)
    assert gate_counts(ops2) == (560, 270)
    assert algorithmic_bytes_plain(ops2, 28, 16) == 5560 * 2 ** 30
    U = ops[0][0]
    assert numpy.allclose(U.conj().T @ U, numpy.eye(2), atol=1e-15)


@pytest.mark.parametrize("n,depth,seed", [(30, 40, 2024), (16, 12, 3), (13, 5, 8)])
def test_online_fusion_equals_batch_fusion(n, depth, seed):
    """Gate-by-gate fusion with final blocks taken out as soon as they are final gives the blocks of
    fuse() on the whole record, in the same order; taken blocks really are final (no later gate
    would have joined them)."""
    from qip_amd.fused import OnlineFusion
    gates = gates_of(layered_random_circuit(n, depth, seed=seed))
    ref = fuse(gates, 6)
    of = OnlineFusion(n, 6)
    got = []
    for gq, G in gates:
        of.add(gq, G)
        k = of.final_count()
        if k:
            got.extend(of.take(k))
    assert of.pending == sum(len(b.gates) for b in of.blocks)
    got.extend(of.take())
    assert of.pending == 0
    assert len(got) == len(ref)
    for a, b in zip(got, ref):
        assert a.qubits == b.qubits
        assert [g[0] for g in a.gates] == [g[0] for g in b.gates]
    if n == 30:
        assert len(got) == 128


def test_online_fusion_waits_for_untouched_qubits():
    """While some qubit has not been touched, a later gate on it could join the first block, so
    nothing is final."""
    from qip_amd.fused import OnlineFusion
    of = OnlineFusion(14, 6)
    for q in range(13):
        of.add((q,), numpy.eye(2))
        of.add((q, q + 1) if q < 12 else (q,), numpy.eye(4) if q < 12 else numpy.eye(2))
    assert of.final_count() == 0
