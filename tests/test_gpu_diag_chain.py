"""Held-back diagonal gates (marked gpu): diagonal operators on up to two targets wait on the state
and a run of them is applied in one sweep (k_diag_chain) before any other operation. The state must
be bit for bit what the same gates give when each is launched alone (a synchronisation after every
gate applies each held gate by itself), and within 1e-12 of the oracle; held gates must land before
reads, other gates, graph launches, fused recording and never on a re-initialised state."""
import numpy
import pytest

from oracle.oracle_backend import OracleBackend
from qip_amd import CMat, GpuBackend
from qip_amd.fused import Fused128Backend, FusedBackend

pytestmark = pytest.mark.gpu
RNG = numpy.random.default_rng(9090)
HM = numpy.array([[1.0, 1.0], [1.0, -1.0]]) / numpy.sqrt(2.0)
XM = numpy.array([[0.0, 1.0], [1.0, 0.0]])


def rand_state(n):
    psi = RNG.standard_normal(2 ** n) + 1j * RNG.standard_normal(2 ** n)
    return psi / numpy.linalg.norm(psi)


def phase(theta):
    return numpy.diag([1.0, numpy.exp(1j * theta)])


def rand_diag_desc(n, ctrl_max=3):
    """A descriptor that the engine applies as a diagonal on at most two targets: a phase, a
    controlled phase (as CMat or as one dense 4x4), a two-target diagonal, under up to ctrl_max
    controls."""
    kind = int(RNG.integers(0, 4))
    qs = [int(q) for q in RNG.permutation(n)[:2 + ctrl_max]]
    th = RNG.uniform(0, 2 * numpy.pi, 4)
    if kind == 0:
        m = phase(th[0])
    elif kind == 1:
        m = numpy.diag(numpy.exp(1j * th[:2]))
    elif kind == 2:
        m = numpy.diag(numpy.exp(1j * th[:4]))
    else:
        return {(qs[0], qs[1]): numpy.diag([1, 1, 1, numpy.exp(1j * th[0])])}
    k = 1 if m.shape[0] == 2 else 2
    nc = int(RNG.integers(0, ctrl_max + 1))
    inner = m
    for _ in range(nc):
        inner = CMat(inner)
    return {tuple(qs[:nc + k]): inner}


def gate_sequence(n, length, diag_share=0.8, ctrl_max=3):
    out = []
    for _ in range(length):
        if RNG.random() < diag_share:
            out.append(rand_diag_desc(n, ctrl_max))
        elif RNG.random() < 0.5:
            out.append({int(RNG.integers(0, n)): HM})
        else:
            a, b = (int(q) for q in RNG.permutation(n)[:2])
            out.append({(a, b): CMat(XM)})
    return out


def run_both(n, dtype, seq):
    psi = rand_state(n)
    held = GpuBackend.make_state(n, [list(range(n))], [psi], statetype=dtype)
    alone = GpuBackend.make_state(n, [list(range(n))], [psi], statetype=dtype)
    for d in seq:
        held.kronselect_dot(d)
        alone.kronselect_dot(d)
        alone.synchronize()                 # each held gate is applied by itself
    return psi, held, alone


@pytest.mark.parametrize("dtype", [numpy.complex128, numpy.complex64])
@pytest.mark.parametrize("n", [10, 16, 22])
def test_held_diagonals_equal_separate_launches(n, dtype):
    seq = gate_sequence(n, 90)
    seq += [rand_diag_desc(n) for _ in range(40)]        # runs longer than 16: full queues
    print("held diagonals n", n, numpy.dtype(dtype).name, len(seq), "gates", flush=True)
    psi, held, alone = run_both(n, dtype, seq)
    got = held.get_state()
    assert numpy.array_equal(got, alone.get_state())
    if dtype == numpy.complex128 and n <= 16:
        ora = OracleBackend.make_state(n, [list(range(n))], [psi])
        for d in seq:
            ora.kronselect_dot(d)
        assert numpy.max(numpy.abs(got - ora.get_state())) < 1e-12
    held.close()
    alone.close()


def test_gates_on_small_subspaces_of_a_large_state():
    """Gates under three controls each, on disjoint qubits of a 24-qubit state: separate sweeps move
    fewer bytes than one over the whole state, and give the same bits."""
    n = 24
    seq = []
    for i in range(3):
        qs = list(range(4 * i, 4 * i + 4))
        seq.append({tuple(qs): CMat(CMat(CMat(phase(0.3 + i))))})
    psi, held, alone = run_both(n, numpy.complex128, seq)
    assert numpy.array_equal(held.get_state(), alone.get_state())
    ref = psi.copy()
    for i in range(3):                                     # qubits 4i..4i+3 all 1: phase
        mask = sum(1 << (n - 1 - q) for q in range(4 * i, 4 * i + 4))
        idx = numpy.nonzero((numpy.arange(2 ** n) & mask) == mask)[0]
        ref[idx] *= numpy.exp(1j * (0.3 + i))
    assert numpy.max(numpy.abs(held.get_state() - ref)) < 1e-15
    held.close()
    alone.close()


def test_held_diagonals_before_reads_and_other_calls():
    n = 14
    psi = rand_state(n)
    be = GpuBackend.make_state(n, [list(range(n))], [psi])
    ora = OracleBackend.make_state(n, [list(range(n))], [psi])
    seq = [rand_diag_desc(n) for _ in range(5)]
    for d in seq:
        be.kronselect_dot(d)
        ora.kronselect_dot(d)
    p = be.measure_probabilities([0, 3, 7])                # a read applies what is held
    assert numpy.max(numpy.abs(numpy.asarray(p) - numpy.asarray(ora.measure_probabilities([0, 3, 7])))) < 1e-14
    for d in seq:
        be.kronselect_dot(d)
        ora.kronselect_dot(d)
    be.func_apply([0, 1], [5], lambda x: x % 2)            # another operation: held gates go first
    ora.func_apply([0, 1], [5], lambda x: x % 2)
    be.kronselect_dot({2: HM})
    ora.kronselect_dot({2: HM})
    assert numpy.max(numpy.abs(be.get_state() - ora.get_state())) < 1e-13
    be.kronselect_dot(rand_diag_desc(n))                   # held, then the state is re-initialised
    be.reinit([list(range(n))], [psi])
    assert numpy.array_equal(be.get_state(), GpuBackend.make_state(n, [list(range(n))], [psi]).get_state())
    be.close()


def test_held_diagonals_in_gate_graphs():
    """A gate held before a recording is applied when the recording begins; a recording that holds
    runs of diagonals between other gates and ends with some: launches of the graph equal the gates
    applied one by one, with eager held gates between launches."""
    n = 13
    psi = rand_state(n)
    seq = gate_sequence(n, 40, diag_share=0.7, ctrl_max=2) + [rand_diag_desc(n, 2) for _ in range(6)]
    d0 = rand_diag_desc(n, 2)
    be = GpuBackend.make_state(n, [list(range(n))], [psi])
    alone = GpuBackend.make_state(n, [list(range(n))], [psi])
    be.kronselect_dot(d0)                                  # held, then applied eagerly by capture_begin
    be.capture_begin()
    for d in seq:
        be.kronselect_dot(d)
    graph = be.capture_end()
    assert graph is not None
    alone.kronselect_dot(d0)
    alone.synchronize()
    assert numpy.array_equal(be.get_state(), alone.get_state())   # the recording has not run
    for launch in range(3):
        be.launch_graph(graph)
        for d in seq:
            alone.kronselect_dot(d)
            alone.synchronize()
        assert numpy.array_equal(be.get_state(), alone.get_state()), launch
        d = rand_diag_desc(n, 2)
        be.kronselect_dot(d)                               # held eagerly between launches
        alone.kronselect_dot(d)
        alone.synchronize()
    be.close()
    alone.close()


def test_held_diagonals_with_fused_recording():
    """Diagonals applied directly on a recording backend are held; recording then applies them
    before the recorded gates (complex64 precision form and complex128 fused form)."""
    n = 14
    psi = rand_state(n)
    for cls, tol in ((FusedBackend, 1e-9), (Fused128Backend, 1e-14)):
        be = cls.make_state(n, [list(range(n))], [psi])
        ora = OracleBackend.make_state(n, [list(range(n))], [psi])
        for r in range(4):
            d = rand_diag_desc(n, 2)
            GpuBackend.kronselect_dot(be, d)               # direct: held back
            ora.kronselect_dot(d)
            for q in range(0, n, 3):                       # recorded for fusion
                be.kronselect_dot({q: HM})
                ora.kronselect_dot({q: HM})
            be.kronselect_dot({(1, 2): CMat(phase(0.4 * r))})
            ora.kronselect_dot({(1, 2): CMat(phase(0.4 * r))})
        got = be.get_state().astype(numpy.complex128)
        ref = ora.get_state()
        inf = 1.0 - abs(numpy.vdot(ref, got)) ** 2 / (numpy.vdot(got, got).real * numpy.vdot(ref, ref).real)
        assert inf < tol * 20, (cls.__name__, inf)
        be.close()


def rand_u2():
    z = (RNG.standard_normal((2, 2)) + 1j * RNG.standard_normal((2, 2))) / numpy.sqrt(2)
    q, r = numpy.linalg.qr(z)
    return q * (numpy.diagonal(r) / numpy.abs(numpy.diagonal(r)))


def single_run_sequence(n, length):
    """Single-qubit gates (Haar, H, X), some under one or two controls, with a diagonal now and
    then: long runs of held single-qubit gates, tile runs split where their bits do not fit."""
    out = []
    for _ in range(length):
        r = RNG.random()
        qs = [int(q) for q in RNG.permutation(n)[:3]]
        if r < 0.55:
            out.append({qs[0]: rand_u2()})
        elif r < 0.7:
            out.append({qs[0]: HM if RNG.random() < 0.5 else XM})
        elif r < 0.85:
            out.append({(qs[1], qs[0]): CMat(rand_u2())})
        elif r < 0.93:
            out.append({(qs[2], qs[1], qs[0]): CMat(CMat(XM))})
        else:
            out.append(rand_diag_desc(n, 2))
    return out


@pytest.mark.parametrize("dtype", [numpy.complex128, numpy.complex64])
@pytest.mark.parametrize("n", [12, 15, 21])
def test_held_single_qubit_runs_equal_separate_launches(n, dtype):
    seq = single_run_sequence(n, 150)
    seq += [{q: rand_u2()} for q in range(n)]            # a whole layer, in qubit order
    seq += [{q: rand_u2()} for q in range(n - 4, n)] * 9  # 36 gates on the low bits: full runs
    print("held single-qubit runs n", n, numpy.dtype(dtype).name, len(seq), "gates", flush=True)
    psi, held, alone = run_both(n, dtype, seq)
    got = held.get_state()
    assert numpy.array_equal(got, alone.get_state())
    if dtype == numpy.complex128 and n <= 15:
        ora = OracleBackend.make_state(n, [list(range(n))], [psi])
        for d in seq:
            ora.kronselect_dot(d)
        assert numpy.max(numpy.abs(got - ora.get_state())) < 1e-12
    held.close()
    alone.close()


def test_layered_circuit_bits_and_graph():
    """A 20-qubit layered random circuit of the benchmarks (Haar layers, CNOT and CZ layers):
    free-running equals a synchronisation after every gate, bit for bit; and a recorded gate graph
    of single-qubit runs equals the same gates launched one by one."""
    from qip_amd.workloads import layered_random_circuit
    n = 20
    ops = layered_random_circuit(n, 6, 11)
    psi, held, alone = run_both(n, numpy.complex128, ops)
    assert numpy.array_equal(held.get_state(), alone.get_state())
    held.close()
    alone.close()
    n = 14
    psi = rand_state(n)
    seq = single_run_sequence(n, 60)
    be = GpuBackend.make_state(n, [list(range(n))], [psi])
    ref = GpuBackend.make_state(n, [list(range(n))], [psi])
    be.capture_begin()
    for d in seq:
        be.kronselect_dot(d)
    graph = be.capture_end()
    assert graph is not None
    for launch in range(3):
        be.launch_graph(graph)
        for d in seq:
            ref.kronselect_dot(d)
            ref.synchronize()
        assert numpy.array_equal(be.get_state(), ref.get_state()), launch
    be.close()
    ref.close()
