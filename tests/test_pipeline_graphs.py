"""The graph frontend's gate-graph replay on the CPU: a state backend that records gate calls between
capture_begin and capture_end and applies them when the recording is launched (the oracle does the
arithmetic) shows which runs of steps a schedule records, that a kept recording is launched on
later runs, recorded again when a gate node's descriptor changes, and that a run whose recording is
refused is applied gate by gate from then on, with the oracle's results throughout."""
import numpy

from oracle.oracle_backend import OracleBackend
from qip_amd import H, MatrixOp, QFFT, Qubit, StochasticMeasure, run
from qip_amd import planner
from qip_amd.pipeline import GRAPH_MIN_GATES, _schedule_for

RNG = numpy.random.default_rng(777)
XM = numpy.array([[0.0, 1.0], [1.0, 0.0]])


class Recording:
    def __init__(self, descriptors):
        self.descriptors = descriptors


class RecordingState(OracleBackend):
    """OracleBackend whose gate calls can be recorded and launched later; wider than 3 targets is
    refused while recording, as the engine refuses what it cannot put in a graph. Like GpuBackend,
    it plans every descriptor (a gate node's Descriptor then keeps its plan, which a kept recording
    is checked against)."""
    graphs_reusable = True
    log = []

    @staticmethod
    def make_state(n, index_groups, feed_list, inputstartindex=None, inputendindex=None, outputstartindex=None,
                   outputendindex=None, statetype=numpy.complex128):
        base = OracleBackend.make_state(n, index_groups, feed_list)
        s = RecordingState(n, base.state, base.arena)
        s.recording = None
        return s

    def capture_begin(self):
        self.recording = []
        RecordingState.log.append("capture")

    def capture_end(self):
        rec, self.recording = self.recording, None
        return None if rec is None or rec == "refused" else Recording(rec)

    def graph_matches(self, graph):
        return True

    def launch_graph(self, graph):
        RecordingState.log.append("launch")
        for d in graph.descriptors:
            OracleBackend.kronselect_dot(self, d)

    def kronselect_dot(self, mats, input_offset=0, output_offset=0):
        planner.plan(mats)                       # as GpuBackend does: a Descriptor keeps its plan
        if self.recording is None:
            return OracleBackend.kronselect_dot(self, mats, input_offset, output_offset)
        if self.recording == "refused" or any(len(k if type(k) == tuple else (k,)) > 3 for k in mats):
            self.recording = "refused"
            raise NotImplementedError("refused while recording")
        self.recording.append(mats)


def rand_state(n):
    psi = RNG.standard_normal(2 ** n) + 1j * RNG.standard_normal(2 ** n)
    return psi / numpy.linalg.norm(psi)


def test_qft_runs_record_once_then_launch():
    n = 6
    psi = rand_state(n)
    q = Qubit(n=n)
    sm = StochasticMeasure(QFFT(q))
    RecordingState.log.clear()
    expect = numpy.fft.ifft(psi, norm="ortho")
    for _ in range(4):
        s, _ = run(sm, feed={q: psi}, backend_constructor=RecordingState.make_state)
        assert numpy.max(numpy.abs(s - expect)) < 1e-12
    assert RecordingState.log.count("capture") == 1
    assert RecordingState.log.count("launch") == 4
    sched = _schedule_for((sm,), ((q,),))
    assert list(sched.segments.values()) and all(sched.graphs[k] for k in sched.segments)


def test_short_runs_are_not_recorded():
    """Fewer than GRAPH_MIN_GATES gate steps, or runs split by a function node: gate by gate."""
    n = 4
    psi = rand_state(n)
    q = Qubit(n=n)
    out = q
    for _ in range(GRAPH_MIN_GATES - 1):
        out = H(out)
    RecordingState.log.clear()
    for _ in range(2):
        run(out, feed={q: psi}, backend_constructor=RecordingState.make_state)
    assert RecordingState.log == []


def test_changed_descriptor_records_again():
    n = 5
    psi = rand_state(n)
    q = Qubit(n=n)
    out = q
    for _ in range(GRAPH_MIN_GATES + 2):
        out = H(out)
    RecordingState.log.clear()
    for _ in range(3):
        run(out, feed={q: psi}, backend_constructor=RecordingState.make_state)
    assert RecordingState.log.count("capture") == 1
    d = out.ms
    assert type(d) is planner.Descriptor and d._plan is not None
    d[next(iter(d))] = XM.copy()               # a change through the dict: the kept plan is dropped
    s, _ = run(out, feed={q: psi}, backend_constructor=RecordingState.make_state)
    ora, _ = run(out, feed={q: psi}, backend_constructor=OracleBackend.make_state)
    assert numpy.max(numpy.abs(s - ora)) < 1e-12
    assert RecordingState.log.count("capture") == 2
    # the changed descriptor holds a writeable matrix: no kept plan, so the run is not recorded again
    run(out, feed={q: psi}, backend_constructor=RecordingState.make_state)
    assert RecordingState.log.count("capture") == 2


def test_refused_recording_runs_gate_by_gate():
    n = 6
    psi = rand_state(n)
    q = Qubit(n=n)
    u = numpy.linalg.qr(RNG.standard_normal((16, 16)) + 1j * RNG.standard_normal((16, 16)))[0]

    class Wide(MatrixOp):
        def makemats(self, index_groups):
            return {tuple(index_groups[0][:4]): u}

    out = q
    for _ in range(GRAPH_MIN_GATES):
        out = H(out)
    out = Wide(out)
    ora, _ = run(out, feed={q: psi}, backend_constructor=OracleBackend.make_state)
    RecordingState.log.clear()
    for _ in range(3):
        s, _ = run(out, feed={q: psi}, backend_constructor=RecordingState.make_state)
        assert numpy.max(numpy.abs(s - ora)) < 1e-12
    assert RecordingState.log == ["capture"]     # tried once, then gate by gate
