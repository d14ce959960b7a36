"""CPU tests of the batching form's host side: the packed circuit format, the generator of the
512-circuit workload, and the engine's host plan (qipgpu_batch_plan, no GPU call) against the
precision form's Python fusion rule."""
import numpy
import pytest

from qip_amd import CMat, SwapMat, batch, workloads
from qip_amd.fused import fuse

RNG = numpy.random.default_rng(7)


def rand_u(d):
    z = RNG.standard_normal((d, d)) + 1j * RNG.standard_normal((d, d))
    return numpy.linalg.qr(z)[0]


@pytest.mark.parametrize("n,depth,seed", [(5, 4, 3), (12, 7, 11), (1, 3, 2), (22, 2, 9)])
def test_generator_equals_descriptor_circuit(n, depth, seed):
    """The vectorised generator reproduces layered_random_circuit gate for gate, bit for bit."""
    fast = workloads.layered_circuit_gates(n, depth, seed)
    slow = [g for op in workloads.layered_random_circuit(n, depth, seed) for g in batch.descriptor_gates(op)]
    assert len(fast) == len(slow)
    for (qa, ma), (qb, mb) in zip(fast, slow):
        assert tuple(qa) == tuple(qb)
        assert numpy.array_equal(numpy.asarray(ma), numpy.asarray(mb))


def test_batch_workload_shape():
    ns, depths, seeds, packed = workloads.batch_layered_circuits(512, 2025)
    assert len(ns) == len(depths) == len(seeds) == packed.count == 512
    assert min(ns) >= 10 and max(ns) <= 22 and min(depths) >= 8 and max(depths) <= 32
    assert set(ns) == set(range(10, 23))
    expect = sum(n * d + ((d + 1) // 2) * (n // 2) + (d // 2) * ((n - 1) // 2) for n, d in zip(ns, depths))
    assert packed.n_gates == expect
    # the same seed gives the same batch
    ns2, _, seeds2, packed2 = workloads.batch_layered_circuits(512, 2025)
    assert ns2 == ns and seeds2 == seeds and numpy.array_equal(packed2.mats, packed.mats)


def test_packed_round_trip():
    gates = [[((0,), rand_u(2)), ((1, 0), rand_u(4))], [], [((2, 0, 1), rand_u(8))]]
    p = batch.PackedCircuits.from_gates([2, 4, 3], gates)
    assert p.count == 3 and p.n_gates == 3
    for i in range(3):
        got = p.gates_of(i)
        assert len(got) == len(gates[i])
        for (qa, ma), (qb, mb) in zip(got, gates[i]):
            assert qa == tuple(qb)
            assert numpy.array_equal(ma, mb)


def test_packed_rejects_bad_shapes():
    with pytest.raises(ValueError):
        batch.PackedCircuits.from_gates([2], [[((0,), numpy.eye(4))]])
    with pytest.raises(NotImplementedError):
        batch.PackedCircuits.from_gates([8], [[(tuple(range(7)), numpy.eye(128))]])
    with pytest.raises(ValueError):
        batch.PackedCircuits([2], [0], [1], numpy.zeros((1, 6)), numpy.zeros(3))


def test_descriptor_gates_cover_reference_descriptors():
    """CMat chains become dense controlled gates (controls first), SwapMat pairwise swaps, plain
    products single-qubit factors."""
    X = numpy.array([[0, 1], [1, 0]], dtype=complex)
    g = batch.descriptor_gates({(0, 1, 2): CMat(CMat(X))})
    assert len(g) == 1 and g[0][0] == (0, 1, 2)
    toffoli = numpy.eye(8, dtype=complex)
    toffoli[6:, 6:] = X
    assert numpy.allclose(g[0][1], toffoli)
    g = batch.descriptor_gates({(0, 1): SwapMat(1)})
    assert len(g) == 1 and sorted(g[0][0]) == [0, 1]
    swap = numpy.eye(4)[[0, 2, 1, 3]]
    assert numpy.allclose(g[0][1], swap)
    h = numpy.array([[1, 1], [1, -1]]) / numpy.sqrt(2)
    g = batch.descriptor_gates({0: h, 3: h})
    assert sorted(q for q, _ in g) == [(0,), (3,)]
    with pytest.raises(NotImplementedError):
        m = X
        for _ in range(6):
            m = CMat(m)
        batch.descriptor_gates({tuple(range(7)): m})


def python_blocks(gates, width):
    """Block index of every gate under the precision form's fuse()."""
    tagged = [(tuple(q), (i, m)) for i, (q, m) in enumerate(gates)]
    blocks = fuse(tagged, width)
    where = {}
    for b, blk in enumerate(blocks):
        for _, (i, _m) in blk.gates:
            where[i] = b
    return [where[i] for i in range(len(gates))], len(blocks)


@pytest.mark.parametrize("width", [2, 4, 5, 6])
def test_host_plan_equals_python_fusion(width):
    """The engine's C++ fusion assigns every gate to the block the Python rule picks."""
    ns = [14, 12, 20, 13, 16]
    circuits = [workloads.layered_circuit_gates(n, 6 + i, 100 + i) for i, n in enumerate(ns)]
    # sprinkle wider dense gates into two circuits
    circuits[2] = circuits[2][:40] + [((3, 9, 1), rand_u(8)), ((17, 2, 5, 11), rand_u(16))] + circuits[2][40:]
    circuits[4] = circuits[4][:10] + [((0, 15, 7, 3, 9, 12), rand_u(64))] + circuits[4][10:]
    packed = batch.PackedCircuits.from_gates(ns, circuits)
    gate_block, state_blocks = batch.plan_fusion(packed, fusion_width=width, lds_max_qubits=12)
    for i, n in enumerate(ns):
        mine = gate_block[packed.gate_state == i]
        if n <= 12:
            assert (mine == -1).all() and state_blocks[i] == 0
            continue
        ref, nblocks = python_blocks(circuits[i], width)
        assert state_blocks[i] == nblocks
        assert list(mine) == ref


def test_host_plan_workload_blocks():
    """30-qubit depth-40 circuit: the C++ plan gives the 66 blocks the single-state path uses."""
    gates = workloads.layered_circuit_gates(30, 40, 2024)
    packed = batch.PackedCircuits.from_gates([30], [gates])
    _, state_blocks = batch.plan_fusion(packed)
    assert state_blocks[0] == 66


@pytest.mark.parametrize("bad", [
    dict(gate_state=[1]),                 # state out of range
    dict(gate_k=[0]),                     # k = 0
    dict(gate_k=[7]),                     # k > 6
    dict(gate_qubits=[[4, 0, 0, 0, 0, 0]]),   # qubit >= n
    dict(gate_qubits=[[-1, 0, 0, 0, 0, 0]]),  # negative qubit
])
def test_host_plan_rejects_bad_gates(bad):
    args = dict(gate_state=[0], gate_k=[1], gate_qubits=[[0, 0, 0, 0, 0, 0]])
    args.update(bad)
    k = int(args["gate_k"][0])
    p = batch.PackedCircuits.__new__(batch.PackedCircuits)
    p.n_qubits = [4]
    p.gate_state = numpy.array(args["gate_state"], dtype=numpy.int32)
    p.gate_k = numpy.array(args["gate_k"], dtype=numpy.int32)
    p.gate_qubits = numpy.array(args["gate_qubits"], dtype=numpy.int32)
    p.mats = numpy.zeros(4 ** max(k, 0), dtype=numpy.complex128)
    with pytest.raises(ValueError):
        batch.plan_fusion(p)


def test_host_plan_rejects_repeated_qubit_and_bad_tuning():
    p = batch.PackedCircuits.from_gates([4], [[((0, 1), numpy.eye(4))]])
    p.gate_qubits[0, 1] = 0
    with pytest.raises(ValueError, match="repeated"):
        batch.plan_fusion(p)
    good = batch.PackedCircuits.from_gates([4], [[((0, 1), numpy.eye(4))]])
    for kw in (dict(fusion_width=0), dict(fusion_width=7), dict(lds_max_qubits=10), dict(lds_max_qubits=14)):
        with pytest.raises(ValueError):
            batch.plan_fusion(good, **kw)
    bad_n = batch.PackedCircuits.from_gates([31], [[]])
    with pytest.raises(ValueError):
        batch.plan_fusion(bad_n)
