"""CPU tests of the precision form's host logic: the light-cone fusion keeps program semantics (the
ordered product of block matrices equals the ordered product of the gates, both expanded to the
full 2^n operator independently), respects the width and every dependency, takes out only final
blocks while streaming, and gives the expected block counts."""
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


def _random_gates(n, count, seed, kmax=3):
    """Gates of 1..kmax qubits on random qubit sets, random unitaries."""
    rng = numpy.random.default_rng(seed)
    gates = []
    for _ in range(count):
        L = int(rng.integers(1, kmax + 1))
        gq = tuple(int(q) for q in rng.choice(n, size=L, replace=False))
        Q, _ = numpy.linalg.qr(rng.standard_normal((2 ** L, 2 ** L)) + 1j * rng.standard_normal((2 ** L, 2 ** L)))
        gates.append((gq, Q))
    return gates


@pytest.mark.parametrize("width,shape", [(2, "layered"), (3, "layered"), (4, "layered"), (6, "layered"),
                                         (3, "random"), (4, "random"), (6, "random")])
def test_fusion_preserves_operator(width, shape):
    n, depth = 7, 4
    if shape == "layered":
        gates = gates_of(layered_random_circuit(n, depth, seed=5))
    else:
        gates = _random_gates(n, 40, seed=width, kmax=min(width, 3))
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
    """On the brickwork circuits the light-cone rule covers diamonds of the circuit: 66 six-qubit
    blocks for the 1780 gates at 30 qubits (128 when each gate joins the first block, in program
    order, that it fits), 32 for the 830 gates at 28 qubits depth 20."""
    gates = gates_of(layered_random_circuit(30, 40, seed=1))
    assert len(gates) == 1780
    assert len(fuse(gates, 6)) == 66
    assert len(fuse(gates_of(layered_random_circuit(28, 20, seed=2024)), 6)) == 32


def test_block_counts_qft():
    """QFT circuits (H, then controlled phases from every later qubit): the look-ahead lets a block take
    the qubits its next gates need, 28 blocks for 24 qubits (48 without it)."""
    gates = []
    for i in range(24):
        gates.append(((i,), numpy.eye(2)))
        gates.extend(((j, i), numpy.eye(4)) for j in range(i + 1, 24))
    assert len(fuse(gates, 6)) == 28


def _check_plan(gates, blocks, width):
    """Every gate in exactly one block, gates in program order inside a block, blocks within the
    width, and every dependency (two gates sharing a qubit) kept in order across blocks."""
    where = {}
    for b, blk in enumerate(blocks):
        assert len(blk.qubits) <= width
        prev = -1
        for gq, G in blk.gates:
            i = next(j for j, (q2, G2) in enumerate(gates) if G2 is G)
            assert i > prev
            prev = i
            assert set(gq) <= blk.qubits
            assert i not in where
            where[i] = b
    assert sorted(where) == list(range(len(gates)))
    last = {}
    for i, (gq, _) in enumerate(gates):
        for q in gq:
            if q in last:
                assert where[last[q]] <= where[i]
            last[q] = i


@pytest.mark.parametrize("seed", range(4))
def test_fusion_plan_is_valid(seed):
    n = 12 + seed
    gates = _random_gates(n, 150, seed, kmax=4)
    for width in (4, 6):
        _check_plan(gates, fuse(gates, width), width)
    gates = gates_of(layered_random_circuit(n, 10, seed=seed))
    _check_plan(gates, fuse(gates, 6), 6)


def test_fusion_tiny_cases():
    """Hand-checked plans: a chain of CNOTs along a line at width 3 grows along the chain; gates on
    disjoint qubits share a block up to the width."""
    from qip_amd.fused import plan_blocks
    blocks, _ = plan_blocks([(0, 1), (1, 2), (2, 3), (3, 4)], 3)
    assert [(sorted(S), t) for S, t in blocks] == [([0, 1, 2], [0, 1]), ([2, 3, 4], [2, 3])]
    blocks, _ = plan_blocks([(0,), (1,), (2,), (3,), (0,)], 2)
    assert [(sorted(S), t) for S, t in blocks] == [([0, 1], [0, 1, 4]), ([2, 3], [2, 3])]
    # a gate waits for its predecessors: (0, 1) is not ready before (1, 2) is placed, and (1, 2)
    # does not fit beside qubit 0 at width 2, so the first block keeps qubit 0 alone
    blocks, _ = plan_blocks([(0,), (1, 2), (0, 1)], 2)
    assert [(sorted(S), t) for S, t in blocks] == [([0], [0]), ([1, 2], [1]), ([0, 1], [2])]


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
def test_streaming_fusion_equals_batch_fusion(n, depth, seed):
    """Final blocks taken out while the gates arrive, at any moments, plus the plan of the rest at the
    end, are the blocks of fuse() on the whole record, in the same order."""
    import random
    from qip_amd.fused import StreamingFusion
    gates = gates_of(layered_random_circuit(n, depth, seed=seed))
    ref = fuse(gates, 6)
    for cadence in (0.05, 0.3):
        rnd = random.Random(seed)
        sf = StreamingFusion(n, 6)
        got = []
        for gq, G in gates:
            sf.add(gq, G)
            if rnd.random() < cadence:
                got.extend(sf.take_final())
        early = len(got)
        got.extend(sf.take())
        assert not sf.gates
        assert len(got) == len(ref)
        for a, b in zip(got, ref):
            assert a.qubits == b.qubits
            assert [id(g[1]) for g in a.gates] == [id(g[1]) for g in b.gates]
        if n == 30:
            assert len(got) == 66 and early >= 50


def test_streaming_fusion_waits_for_untouched_qubits():
    """(0, 1), (1, 2, 3), (0, 2) at width 3: the first block is {0, 1} (the three-qubit gate does not
    fit beside it) and every qubit keeps a recorded gate outside it. On 4 qubits it is final; on 5 it
    is not, since qubit 4 has no recorded gate and a later gate there could join it."""
    from qip_amd.fused import plan_blocks
    qs = [(0, 1), (1, 2, 3), (0, 2)]
    blocks, final = plan_blocks(qs, 3, 4)
    assert [(sorted(S), t) for S, t in blocks][0] == ([0, 1], [0])
    assert final == 1
    assert plan_blocks(qs, 3, 5)[1] == 0
    # qubits 0 and 1 have nothing left outside the first block: not final
    blocks, final = plan_blocks([(0, 1, 2), (2, 3), (0, 1)], 3, 4)
    assert [(sorted(S), t) for S, t in blocks][0] == ([0, 1, 2], [0, 2])
    assert final == 0
    # each of them keeps one here
    blocks, final = plan_blocks([(0, 1, 2), (2, 3), (0, 3), (1, 3)], 3, 4)
    assert [(sorted(S), t) for S, t in blocks][0] == ([0, 1, 2], [0])
    assert final == 1


def test_streaming_fusion_waits_after_exhausting_look_ahead():
    """(1, 2), (0, 1), (2, 3), (1, 3), (0, 1) at width 3: the first block is {0, 1, 2} and every qubit
    keeps a recorded gate outside it, but the look-ahead of (2, 3) absorbed qubit 2's last recorded
    gate. A later gate on qubit 2 changes that look-ahead, and the block: it is not final."""
    from qip_amd.fused import plan_blocks
    qs = [(1, 2), (0, 1), (2, 3), (1, 3), (0, 1)]
    blocks, final = plan_blocks(qs, 3, 4)
    assert [(sorted(S), t) for S, t in blocks][0] == ([0, 1, 2], [0, 1])
    assert final == 0
    assert sorted(plan_blocks(qs + [(2,)], 3, 4)[0][0][0]) == [1, 2, 3]


def test_streaming_fusion_random_records():
    """Random gate lists of 1..4 qubits, random widths, final blocks taken at random moments: always
    the blocks of fuse() on the whole record."""
    import random
    from qip_amd.fused import StreamingFusion
    for trial in range(60):
        rng = numpy.random.default_rng(trial)
        n, width = int(rng.integers(3, 14)), int(rng.integers(2, 7))
        gates = [(tuple(int(q) for q in rng.choice(n, int(rng.integers(1, min(width, 4, n) + 1)), replace=False)),
                  numpy.eye(2)) for _ in range(int(rng.integers(1, 120)))]
        gates = [(gq, numpy.eye(2 ** len(gq))) for gq, _ in gates]
        ref = fuse(gates, width)
        rnd = random.Random(trial)
        sf = StreamingFusion(n, width)
        got = []
        cadence = rnd.choice([0.05, 0.3, 1.0])
        for gq, G in gates:
            sf.add(gq, G)
            if rnd.random() < cadence:
                got.extend(sf.take_final())
        got.extend(sf.take())
        assert [(b.qubits, [id(g[1]) for g in b.gates]) for b in got] == \
               [(b.qubits, [id(g[1]) for g in b.gates]) for b in ref], trial


def _compose_tensordot(qs, gates):
    """The block product by tensor contraction (independent of the library's loop)."""
    k = len(qs)
    pos = {q: i for i, q in enumerate(qs)}
    M = numpy.eye(2 ** k, dtype=numpy.complex128).reshape([2] * k + [2 ** k])
    for gq, G in gates:
        L = len(gq)
        axes = [pos[q] for q in gq]
        M = numpy.tensordot(G.reshape([2] * (2 * L)), M, axes=(list(range(L, 2 * L)), axes))
        M = numpy.moveaxis(M, list(range(L)), axes)
    return M.reshape(2 ** k, 2 ** k)


@pytest.mark.parametrize("seed", range(6))
def test_compose_matches_full_operator(seed):
    """qipgpu_compose (host C++) on random gates of 1..k qubits in any order, against the product of
    the gates embedded element by element on the block qubits, and against tensor contraction."""
    from qip_amd.fused import compose
    rng = numpy.random.default_rng(seed)
    k = int(rng.integers(1, 7))
    qs = sorted(rng.choice(40, size=k, replace=False).tolist())
    gates = []
    for _ in range(int(rng.integers(0, 12))):
        L = int(rng.integers(1, k + 1))
        gq = tuple(int(q) for q in rng.permutation(qs)[:L])
        G = rng.standard_normal((2 ** L, 2 ** L)) + 1j * rng.standard_normal((2 ** L, 2 ** L))
        gates.append((gq, G))
    got = compose(qs, gates)
    # element by element: <r|G|c> = G[sub(r), sub(c)] when r and c agree off the gate's bits
    # (block qubit qs[i] is bit k-1-i of the block index, gate qubit j bit L-1-j of G's index)
    bit = {q: k - 1 - i for i, q in enumerate(qs)}
    D = 2 ** k
    want = numpy.eye(D, dtype=complex)
    for gq, G in gates:
        mask = sum(1 << bit[q] for q in gq)
        sub = [sum(((x >> bit[q]) & 1) << (len(gq) - 1 - j) for j, q in enumerate(gq)) for x in range(D)]
        E = numpy.zeros((D, D), dtype=complex)
        for r in range(D):
            for c in range(D):
                if (r & ~mask) == (c & ~mask):
                    E[r, c] = G[sub[r], sub[c]]
        want = E @ want
    scale = max(1.0, float(numpy.max(numpy.abs(want))))
    assert numpy.max(numpy.abs(got - want)) < 1e-12 * scale
    assert numpy.max(numpy.abs(got - _compose_tensordot(qs, gates))) < 1e-12 * scale


def test_compose_rejects_bad_blocks():
    from qip_amd.fused import compose
    H = numpy.array([[1, 1], [1, -1]]) / numpy.sqrt(2)
    with pytest.raises(ValueError):
        compose([0, 1], [((2,), H)])              # gate qubit outside the block
    with pytest.raises(ValueError):
        compose([0, 0], [((0,), H)])              # repeated block qubit
    with pytest.raises(ValueError):
        compose(list(range(7)), [((0,), H)])      # wider than six
    with pytest.raises(ValueError):
        compose([0, 1], [((0, 0), numpy.eye(4))])  # repeated gate qubit
    with pytest.raises(ValueError):
        compose([0, 1], [((0, 1), H)])            # matrix of the wrong size
    assert numpy.array_equal(compose([3], []), numpy.eye(2))


@pytest.mark.parametrize("k", [1, 3, 6])
def test_compose_single_qubit_blocks(k):
    """Blocks of single-qubit gates only (several per qubit, some qubits without one) compose to the
    Kronecker product of each qubit's ordered 2x2 product."""
    from qip_amd.fused import compose
    rng = numpy.random.default_rng(k)
    qs = list(range(10, 10 + k))
    gates, per = [], {q: numpy.eye(2, dtype=complex) for q in qs}
    for _ in range(3 * k):
        q = int(rng.choice(qs[:-1] if k > 1 else qs))      # the last qubit keeps no gate when k > 1
        G = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
        gates.append(((q,), G))
        per[q] = G @ per[q]
    want = numpy.ones((1, 1), dtype=complex)
    for q in qs:
        want = numpy.kron(want, per[q])
    assert numpy.max(numpy.abs(compose(qs, gates) - want)) < 1e-12 * max(1.0, numpy.max(numpy.abs(want)))


def _native_plan(n, width, gates):
    from qip_amd import _capi
    k = numpy.array([len(gq) for gq, _ in gates], dtype=numpy.int32)
    q = numpy.zeros((max(len(gates), 1), 6), dtype=numpy.int32)
    for i, (gq, _) in enumerate(gates):
        q[i, :len(gq)] = gq
    blk = numpy.zeros(max(len(gates), 1), dtype=numpy.int64)
    fin = numpy.zeros(max(len(gates), 1), dtype=numpy.int64)
    _capi.check(_capi.load().qipgpu_fuse_online(n, width, len(gates), _capi.ptr(k), _capi.ptr(q), _capi.ptr(blk),
                                                _capi.ptr(fin)), "fuse_online")
    return blk[:len(gates)], fin[:len(gates)]


def _python_gate_blocks(gates, width):
    where = {}
    for b, blk in enumerate(fuse(gates, width)):
        for _, G in blk.gates:
            where[id(G)] = b
    return [where[id(G)] for _, G in gates]


@pytest.mark.parametrize("n,depth,seed,width", [(30, 40, 2024, 6), (16, 12, 3, 6), (13, 5, 8, 4), (9, 6, 1, 2)])
def test_native_recorder_plan_equals_python_rule(n, depth, seed, width):
    """The library's recorder (qipgpu_fuse_online, the schedule qipgpu_fused_record runs) places every
    gate in the block the Python statement of the rule does; the blocks it has taken out after each
    gate are final for the gates recorded so far (the Python rule finds at least as many)."""
    from qip_amd.fused import plan_blocks
    gates = [(gq, G.copy()) for gq, G in gates_of(layered_random_circuit(n, depth, seed=seed))]
    blk, fin = _native_plan(n, width, gates)
    assert list(blk) == _python_gate_blocks(gates, width)
    assert all(a <= b for a, b in zip(fin[:-1], fin[1:]))
    qs = [gq for gq, _ in gates]
    for g in range(0, len(gates), max(1, len(gates) // 60)):
        assert plan_blocks(qs[:g + 1], width, n)[1] >= fin[g]
        assert all(blk[i] >= fin[g] for i in range(g + 1, len(gates)))   # taken blocks hold no later gate
    if n == 30:
        assert fin[-1] >= 50     # most blocks launch while the circuit is still being recorded


def test_native_plan_random_gates():
    """The same on gates of 1..4 qubits on random qubit sets, for widths 4 and 6."""
    for seed in range(3):
        gates = _random_gates(14 + seed, 300, seed, kmax=4)
        for width in (4, 6):
            blk, _ = _native_plan(14 + seed, width, gates)
            assert list(blk) == _python_gate_blocks(gates, width)


def test_native_recorder_plan_rejects():
    from qip_amd import _capi
    H = numpy.eye(2)
    with pytest.raises(ValueError):
        _native_plan(4, 6, [((4,), H)])             # qubit outside the state
    with pytest.raises(ValueError):
        _native_plan(4, 2, [((0, 1, 2), numpy.eye(8))])   # wider than the fusion width
    with pytest.raises(ValueError):
        _native_plan(4, 6, [((1, 1), numpy.eye(4))])    # repeated qubit
    blk, fin = _native_plan(3, 6, [])
    assert len(blk) == 0


def _native_launches(n, width, streaming, gates):
    from qip_amd import _capi
    import ctypes
    k = numpy.array([len(gq) for gq, _ in gates] or [0], dtype=numpy.int32)
    q = numpy.zeros((max(len(gates), 1), 6), dtype=numpy.int32)
    for i, (gq, _) in enumerate(gates):
        q[i, :len(gq)] = gq
    lob = numpy.full(max(len(gates), 1), -1, dtype=numpy.int64)
    nb = ctypes.c_int64(0)
    _capi.check(_capi.load().qipgpu_fuse_launches(n, width, int(streaming), len(gates), _capi.ptr(k), _capi.ptr(q),
                                                  _capi.ptr(lob), ctypes.byref(nb)), "fuse_launches")
    return lob[:nb.value]


def _mixed_gates(n, layers, seed):
    """Single-qubit layers on random subsets, with a few two-qubit gates between some of them."""
    rng = numpy.random.default_rng(seed)
    gates = []
    for layer in range(layers):
        for q in rng.permutation(n)[:int(rng.integers(1, n + 1))]:
            gates.append(((int(q),), numpy.eye(2)))
        if rng.random() < 0.5:
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            gates.append(((a, b), numpy.eye(4)))
    return gates


@pytest.mark.parametrize("n,seed", [(14, 1), (16, 2), (20, 3), (13, 4), (10, 5)])
def test_native_launches_runs_and_streaming(n, seed):
    """The recorder's launches (qipgpu_fuse_launches, the decisions qipgpu_fused_record makes): the
    blocks of the Python rule, one launch per block except maximal runs of consecutive single-qubit-
    only blocks (n >= 12), and the same launches streamed or deferred."""
    from qip_amd.fused import fuse
    gates = _mixed_gates(n, 12, seed)
    blocks = fuse(gates, 6)
    deferred = _native_launches(n, 6, False, gates)
    streamed = _native_launches(n, 6, True, gates)
    assert len(deferred) == len(blocks)
    assert numpy.array_equal(deferred, streamed)
    product = [all(len(gq) == 1 for gq, _ in b.gates) for b in blocks]
    expect, cur = [], -1
    for i, p in enumerate(product):
        if not (n >= 12 and p and i > 0 and product[i - 1]):
            cur += 1
        expect.append(cur)
    assert deferred.tolist() == expect


def test_native_launches_known_runs():
    """A 2-qubit gate, H on 17 of 18 qubits, then H on the even ones: the plan is [mixed, product,
    product] and the two product blocks share one launch; below 12 qubits every block is its own
    launch."""
    H = numpy.array([[1, 1], [1, -1]]) / numpy.sqrt(2)
    gates = [((3, 4), numpy.eye(4))] + [((q,), H) for q in range(17)] + [((q,), H) for q in range(0, 17, 2)]
    kinds = [all(len(gq) == 1 for gq, _ in b.gates) for b in fuse(gates, 6)]
    assert kinds == [False, True, True]
    assert _native_launches(18, 6, True, gates).tolist() == [0, 1, 1]
    assert _native_launches(18, 6, False, gates).tolist() == [0, 1, 1]
    small = [((3, 4), numpy.eye(4))] + [((q,), H) for q in range(10)]
    assert _native_launches(10, 6, True, small).tolist() == [0, 1]
