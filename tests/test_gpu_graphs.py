"""Gate graphs (marked gpu): gates recorded between GpuBackend.capture_begin and capture_end run
only when the graph is launched, give the bits the same gates give launched one by one, and launch
again on later states that hold the same buffer; the graph frontend plays replayed runs of plain
gates as graphs. Every launch is checked against the eager result, with eager work and
synchronisations between launches."""
import threading

import numpy
import pytest

from oracle.oracle_backend import OracleBackend
from qip_amd import CMat, FuncTable, GpuBackend, H, MatrixOp, QFFT, Qubit, StochasticMeasure, run

pytestmark = pytest.mark.gpu
RNG = numpy.random.default_rng(4242)
HM = numpy.array([[1.0, 1.0], [1.0, -1.0]]) / numpy.sqrt(2.0)
XM = numpy.array([[0.0, 1.0], [1.0, 0.0]])


def rand_u(d):
    z = (RNG.standard_normal((d, d)) + 1j * RNG.standard_normal((d, d))) / numpy.sqrt(2)
    q, r = numpy.linalg.qr(z)
    return q * (numpy.diagonal(r) / numpy.abs(numpy.diagonal(r)))


def rand_state(n):
    psi = RNG.standard_normal(2 ** n) + 1j * RNG.standard_normal(2 ** n)
    return psi / numpy.linalg.norm(psi)


def gate_list(n):
    """Descriptors whose calls all travel as kernel arguments: single-qubit layers (product sweeps
    from 12 qubits on), CNOT, controlled phases, a 3-qubit dense operator, a swap."""
    descs = [{q: HM for q in range(n)}]
    for q in range(0, n - 1, 2):
        descs.append({(q, q + 1): CMat(XM)})
    for q in range(1, n - 1, 3):
        descs.append({(q, q + 1): CMat(numpy.diag([1.0, numpy.exp(0.7j * q)]))})
    descs.append({(2, 5, 7): rand_u(8)})
    descs.append({q: rand_u(2) for q in range(0, n, 3)})
    descs.append({(0, n - 1): numpy.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]])})
    return descs


@pytest.mark.parametrize("n,dtype", [(10, numpy.complex128), (14, numpy.complex128), (14, numpy.complex64)])
def test_capture_then_launch_equals_eager(n, dtype):
    psi = rand_state(n)
    descs = gate_list(n)
    eager = GpuBackend.make_state(n, [list(range(n))], [psi], statetype=dtype)
    be = GpuBackend.make_state(n, [list(range(n))], [psi], statetype=dtype)
    start = be.get_state()
    print("gate graph n", n, numpy.dtype(dtype).name, len(descs), "descriptors", flush=True)
    be.capture_begin()
    for d in descs:
        be.kronselect_dot(d)
    graph = be.capture_end()
    assert graph is not None and be.graph_matches(graph)
    assert numpy.array_equal(be.get_state(), start)         # recorded, not run
    for launch in range(3):
        be.launch_graph(graph)
        for d in descs:
            eager.kronselect_dot(d)
        assert numpy.array_equal(be.get_state(), eager.get_state()), launch
        be.kronselect_dot({0: HM})                            # eager work between launches
        eager.kronselect_dot({0: HM})
        be.synchronize()
    ora = OracleBackend.make_state(n, [list(range(n))], [start.astype(numpy.complex128)])
    for _ in range(3):
        for d in descs:
            ora.kronselect_dot(d)
        ora.kronselect_dot({0: HM})
    tol = 1e-12 if dtype == numpy.complex128 else 1e-5
    assert numpy.max(numpy.abs(be.get_state() - ora.get_state())) < tol
    be.close()
    eager.close()


def test_refused_operations_drop_the_capture():
    """A 4-target dense operator (staged through pinned memory), func_apply and a read are refused
    while capturing; the capture is dropped and the state is as before it."""
    n = 13
    psi = rand_state(n)
    table = FuncTable(numpy.arange(4) % 2)
    warm = GpuBackend.make_state(n, [], [])
    warm.func_apply([0, 1], [2], table)                     # the table's device copy, made outside
    warm.close()
    be = GpuBackend.make_state(n, [list(range(n))], [psi])
    for refused in (lambda: be.kronselect_dot({(0, 1, 2, 3): rand_u(16)}),
                    lambda: be.func_apply([0, 1], [2], table),
                    lambda: be.measure_probabilities([0])):
        be.capture_begin()
        be.kronselect_dot({0: HM})
        with pytest.raises(NotImplementedError):
            refused()
        assert be.capture_end() is None
        assert numpy.array_equal(be.get_state(), psi)
    be.kronselect_dot({0: HM})                              # the state works as usual afterwards
    ora = OracleBackend.make_state(n, [list(range(n))], [psi])
    ora.kronselect_dot({0: HM})
    assert numpy.max(numpy.abs(be.get_state() - ora.get_state())) < 1e-13
    be.close()


def test_graph_serves_the_next_state_of_the_same_size():
    n = 12
    psi = rand_state(n)
    descs = gate_list(n)
    a = GpuBackend.make_state(n, [list(range(n))], [psi])
    a.capture_begin()
    for d in descs:
        a.kronselect_dot(d)
    graph = a.capture_end()
    a.close()
    b = GpuBackend.make_state(n, [list(range(n))], [psi])   # the destroyed state's buffer
    other = GpuBackend.make_state(n + 1, [], [])
    assert b.graph_matches(graph) and not other.graph_matches(graph)
    with pytest.raises(ValueError):
        other.launch_graph(graph)
    b.launch_graph(graph)
    eager = GpuBackend.make_state(n, [list(range(n))], [psi])
    for d in descs:
        eager.kronselect_dot(d)
    assert numpy.array_equal(b.get_state(), eager.get_state())
    for s in (b, other, eager):
        s.close()


def _qft_graph(n):
    q = Qubit(n=n)
    return q, StochasticMeasure(QFFT(q))


def test_qft_graph_runs_as_graphs():
    """The 16-qubit QFT graph: the first run walks it, the second captures its gates, later runs
    launch the graph; every run's state equals ifft(norm='ortho') and the first run's bits."""
    n = 16
    psi = rand_state(n)
    q, sm = _qft_graph(n)
    expect = numpy.fft.ifft(psi, norm="ortho")
    states = []
    for r in range(6):
        s, c = run(sm, feed={q: psi}, backend_constructor=GpuBackend.make_state)
        states.append(s)
        assert numpy.max(numpy.abs(s - expect)) < 1e-12, r
    assert all(numpy.array_equal(states[0], s) for s in states[1:])
    from qip_amd.pipeline import _schedule_for
    sched = _schedule_for((sm,), ((q,),))
    kept = [v for v in sched.graphs.values() if v]
    assert kept, "no gate graph was kept"


def test_changed_descriptor_is_captured_again():
    """A gate node whose descriptor changes through its dict drops the kept plan; the next run
    captures again and applies the new operator."""
    n = 12
    psi = rand_state(n)
    q = Qubit(n=n)
    out = H(q)
    for _ in range(10):
        out = H(out)                                            # 11 H layers: one run of gates
    for _ in range(3):
        s, _ = run(out, feed={q: psi})
    node = out if isinstance(out, MatrixOp) else out.inputs[0]
    d = node.ms
    d[next(iter(d))] = XM.copy()                                 # X instead of H on one qubit
    s2, _ = run(out, feed={q: psi})
    s3, _ = run(out, feed={q: psi})
    ora, _ = run(out, feed={q: psi}, backend_constructor=OracleBackend.make_state)
    assert numpy.max(numpy.abs(s2 - ora)) < 1e-12
    assert numpy.array_equal(s2, s3)


def test_wide_operator_runs_gate_by_gate():
    """A run of gates holding a 4-qubit operator cannot be captured: it is applied gate by gate on
    every run, with the oracle's result."""
    n = 12
    psi = rand_state(n)
    q = Qubit(n=n)
    u = rand_u(16)

    class Wide(MatrixOp):
        def makemats(self, index_groups):
            return {tuple(index_groups[0][:4]): u}

    out = H(q)
    for _ in range(9):
        out = H(out)
    out = Wide(out)
    ora, _ = run(out, feed={q: psi}, backend_constructor=OracleBackend.make_state)
    for _ in range(3):
        s, _ = run(out, feed={q: psi})
        assert numpy.max(numpy.abs(s - ora)) < 1e-12


def test_concurrent_runs_with_graphs():
    n = 14
    psi = rand_state(n)
    q, sm = _qft_graph(n)
    expect = numpy.fft.ifft(psi, norm="ortho")
    errors = []

    def worker():
        try:
            for _ in range(5):
                s, _ = run(sm, feed={q: psi}, backend_constructor=GpuBackend.make_state)
                if numpy.max(numpy.abs(s - expect)) >= 1e-12:
                    errors.append("mismatch")
        except Exception as e:                                  # reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=worker) for _ in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
