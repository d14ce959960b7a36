"""GPU parity of the batching form (marked gpu): every state of a batch against the oracle's
complex128 restatement of cdot_loop applied gate by gate (small states) or the plain complex128
engine (large ones), under several tunings; graph replays against eager launches (bitwise);
runs on caller-set states; agreement with the single-state precision form; the config of 512
uneven circuits on a subset. Bars: 1 - F <= 1e-9 for LDS states (FP64 gate arithmetic, complex64
storage), <= 1e-8 for tiled states (split-BF16 MFMA blocks); max |amplitude error| <= 1e-5."""
import numpy
import pytest

from oracle import oracle
from qip_amd import GpuBackend, batch, workloads
from qip_amd.fused import FusedBackend

pytestmark = pytest.mark.gpu
RNG = numpy.random.default_rng(4242)


def rand_u(d):
    z = (RNG.standard_normal((d, d)) + 1j * RNG.standard_normal((d, d))) / numpy.sqrt(2)
    q, r = numpy.linalg.qr(z)
    return q * (numpy.diagonal(r) / numpy.abs(numpy.diagonal(r)))


def rand_state(n):
    v = RNG.standard_normal(2 ** n) + 1j * RNG.standard_normal(2 ** n)
    return v / numpy.linalg.norm(v)


def oracle_run(n, gates, psi=None):
    v = numpy.zeros(2 ** n, dtype=numpy.complex128)
    if psi is None:
        v[0] = 1.0
    else:
        v[:] = psi
    out = numpy.zeros_like(v)
    for gq, G in gates:
        oracle.kronselect_dot({tuple(int(q) for q in gq): numpy.asarray(G, dtype=numpy.complex128)}, v, n, out)
        v, out = out, v
    return v


def infid(ref, got):
    ref = numpy.asarray(ref, dtype=numpy.complex128)
    got = numpy.asarray(got, dtype=numpy.complex128)
    ov = numpy.vdot(ref, got)
    return 1.0 - abs(ov) ** 2 / (numpy.vdot(ref, ref).real * numpy.vdot(got, got).real)


def mixed_gates(n, depth, seed):
    """A layered circuit with random dense 3..6-qubit gates spliced in."""
    gates = workloads.layered_circuit_gates(n, depth, seed)
    rng = numpy.random.default_rng(seed + 1)
    for k in range(3, min(6, n) + 1):
        qs = tuple(int(q) for q in rng.permutation(n)[:k])
        at = int(rng.integers(0, len(gates) + 1))
        gates.insert(at, (qs, rand_u(2 ** k)))
    return gates


MIXED_NS = [1, 3, 6, 10, 11, 12, 13, 14, 15, 16]


def mixed_batch():
    circuits = [mixed_gates(n, 3 + (i % 4), 50 + i) for i, n in enumerate(MIXED_NS)]
    return circuits, batch.PackedCircuits.from_gates(MIXED_NS, circuits)


def check_state(n, ref, got, tiled):
    bar = 1e-8 if tiled else 1e-9
    got = numpy.asarray(got, dtype=numpy.complex128)      # FP64 sums below
    f = infid(ref, got)
    err = numpy.max(numpy.abs(ref - got))
    assert f < bar, (n, f)
    assert err < 1e-5, (n, err)
    assert abs(numpy.vdot(got, got).real - 1.0) < 1e-6


@pytest.mark.parametrize("tuning", [None, {"use_graph": 0}, {"lds_max_qubits": 11},
                                    {"lds_max_qubits": 12, "chunk_tiles": 1},
                                    {"chunk_tiles": 3, "fusion_width": 4},
                                    {"chunk_tiles": 64, "fusion_width": 5}],
                         ids=["default", "eager", "lds11", "lds12-chunk1", "chunk3-w4", "chunk64-w5"])
def test_batch_matches_oracle(tuning):
    print("batch vs oracle, tuning", tuning, flush=True)
    circuits, packed = mixed_batch()
    b = batch.GpuBatch(MIXED_NS)
    b.load(packed, tuning)
    lds_max = b.tuning["lds_max_qubits"]
    b.run()
    for i, n in enumerate(MIXED_NS):
        check_state(n, oracle_run(n, circuits[i]), b.get_state(i), n > lds_max)
    b.close()


def test_batch_graph_replays_match_eager_bitwise():
    """Replays of the captured graph, with eager work and synchronisation between them, give
    exactly what eager launches of the same plan give."""
    print("graph replays vs eager", flush=True)
    _, packed = mixed_batch()
    eager = batch.GpuBatch(MIXED_NS)
    eager.load(packed, {"use_graph": 0})
    eager.run()
    want = [eager.get_state(i) for i in range(len(MIXED_NS))]
    eager.close()
    g = batch.GpuBatch(MIXED_NS)
    g.load(packed, {"use_graph": 1})
    for rep in range(3):
        g.reset()
        g.run()
        g.synchronize()
        for i in range(len(MIXED_NS)):
            assert numpy.array_equal(g.get_state(i), want[i]), (rep, i)
    g.close()


def test_batch_runs_on_set_states_and_repeats():
    """run() applies the gates to the current contents: set random states, run twice, compare
    with the circuit applied twice by the oracle."""
    print("set_state + two runs", flush=True)
    circuits, packed = mixed_batch()
    b = batch.GpuBatch(MIXED_NS)
    b.load(packed)
    lds_max = b.tuning["lds_max_qubits"]
    psis = [rand_state(n) for n in MIXED_NS]
    for i, psi in enumerate(psis):
        b.set_state(i, psi)
    b.run()
    b.run()
    for i, n in enumerate(MIXED_NS):
        ref = oracle_run(n, circuits[i] + circuits[i], psis[i].astype(numpy.complex64).astype(numpy.complex128))
        check_state(n, ref, b.get_state(i), n > lds_max)
    b.close()


def test_states_without_gates_are_untouched():
    ns = [5, 14, 9, 13]
    gates = [[], [], [((0,), rand_u(2))], []]
    b = batch.GpuBatch(ns)
    b.load(batch.PackedCircuits.from_gates(ns, gates))
    psis = [(rand_state(n) * 0.5).astype(numpy.complex64) for n in ns]    # not normalised on purpose
    for i, psi in enumerate(psis):
        b.set_state(i, psi)
    b.run()
    for i in (0, 1, 3):
        assert numpy.array_equal(b.get_state(i), psis[i])
    ref = oracle_run(9, gates[2], psis[2].astype(numpy.complex128))
    got = b.get_state(2).astype(numpy.complex128)
    assert infid(ref, got) < 1e-12
    assert abs(numpy.vdot(got, got).real - 1.0) < 1e-6     # renormalised
    b.close()


@pytest.mark.parametrize("n", [14, 17])
def test_batch_matches_single_state_precision_form(n):
    """Same gates, same fusion: the batch's tiled state equals FusedBackend's to FP32 level (the
    block operators are composed in a different FP64 order)."""
    print("batch vs FusedBackend n", n, flush=True)
    gates = mixed_gates(n, 6, 900 + n)
    b = batch.GpuBatch([n])
    b.load(batch.PackedCircuits.from_gates([n], [gates]))
    b.run()
    got = b.get_state(0)
    fz = FusedBackend.make_state(n, [], [])
    for gq, G in gates:
        fz.apply_matrix(list(gq), G)
    want = fz.get_state()
    assert numpy.max(numpy.abs(want - got)) < 1e-6
    fz.close()
    b.close()


def test_autotune_keeps_results_correct():
    print("autotune", flush=True)
    ns = [10, 12, 13, 14, 16]
    circuits = [workloads.layered_circuit_gates(n, 8, 70 + n) for n in ns]
    b = batch.GpuBatch(ns)
    b.load(batch.PackedCircuits.from_gates(ns, circuits))
    tuning, ms = b.autotune(reps=1)
    assert ms > 0
    assert tuning["fusion_width"] in (5, 6) and tuning["lds_max_qubits"] in (11, 12, 13)
    assert tuning["chunk_tiles"] in (2, 4, 8, 16, 32)
    e0 = numpy.zeros(2 ** ns[0], dtype=numpy.complex64)
    e0[0] = 1
    assert numpy.array_equal(b.get_state(0), e0)         # autotune leaves |0...0>
    b.run()
    for i, n in enumerate(ns):
        check_state(n, oracle_run(n, circuits[i]), b.get_state(i), n > tuning["lds_max_qubits"])
    b.close()


def test_uneven_batch_subset_against_plain_complex128():
    """24 circuits of the 512-circuit workload (seed 2025): its 8 largest, 8 smallest and 8
    middle states, against the plain complex128 engine applying the same gates one launch each."""
    ns, depths, seeds, packed = workloads.batch_layered_circuits(512, 2025)
    order = sorted(range(512), key=lambda i: (ns[i], i))
    pick = order[:8] + order[252:260] + order[-8:]
    sub_ns = [ns[i] for i in pick]
    sub = [packed.gates_of(i) for i in pick]
    print("uneven subset n =", sub_ns, flush=True)
    b = batch.GpuBatch(sub_ns)
    b.load(batch.PackedCircuits.from_gates(sub_ns, sub))
    b.run()
    lds_max = b.tuning["lds_max_qubits"]
    for j, n in enumerate(sub_ns):
        ref = GpuBackend.make_state(n, [], [])
        for gq, G in sub[j]:
            ref.apply_matrix(list(gq), G)
        want = ref.get_state()
        ref.close()
        check_state(n, want, b.get_state(j), n > lds_max)
    b.close()


def test_batch_errors():
    with pytest.raises(ValueError):
        batch.GpuBatch([])
    with pytest.raises(ValueError):
        batch.GpuBatch([31])
    b = batch.GpuBatch([4, 13])
    bad = batch.PackedCircuits.from_gates([4, 13], [[((0, 1), numpy.eye(4))], []])
    bad.gate_qubits[0, 1] = 4
    with pytest.raises(ValueError):
        b.load(bad)
    good = batch.PackedCircuits.from_gates([4, 13], [[((0, 1), numpy.eye(4))], []])
    b.load(good)
    with pytest.raises(ValueError):
        b.get_state(2)
    with pytest.raises(ValueError):
        b.get_state(-1)
    with pytest.raises(ValueError):
        b.set_state(0, numpy.zeros(8))
    with pytest.raises(ValueError):
        b.set_tuning({"chunk_tiles": 0})
    with pytest.raises(ValueError):
        b.load(batch.PackedCircuits.from_gates([4, 12], [[], []]))
    b.close()


def test_batch_without_gates_and_rejected_retune():
    """A batch loaded with no gates runs as a no-op; a rejected tuning leaves the loaded plan in use."""
    ns = [5, 14]
    b = batch.GpuBatch(ns)
    b.load(batch.PackedCircuits.from_gates(ns, [[], []]))
    b.run()
    b.synchronize()
    e = numpy.zeros(2 ** 14, dtype=numpy.complex64)
    e[0] = 1
    assert numpy.array_equal(b.get_state(1), e)
    gates = [[((0,), rand_u(2))], [((3, 7), rand_u(4)), ((13,), rand_u(2))]]
    b.load(batch.PackedCircuits.from_gates(ns, gates))
    with pytest.raises(ValueError):
        b.set_tuning({"lds_max_qubits": 20})
    b.reset()
    b.run()
    for i, n in enumerate(ns):
        check_state(n, oracle_run(n, gates[i]), b.get_state(i), n > b.tuning["lds_max_qubits"])
    b.close()
