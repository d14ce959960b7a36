"""A graph run again plays its compiled schedule (visiting order and index groups); the results must
equal those of a fresh walk of the graph, also after the graph grows, for nodes that override
`feed` (always walked), for nodes that modify the index groups they are given, and for runs from
several threads at once."""
import threading

import numpy

import qip_amd.pipeline as pipeline
from qip_amd import C, H, Measure, MatrixOp, Not, QFFT, Qubit, StochasticMeasure, Swap, run
from oracle.oracle_backend import OracleBackend

BC = OracleBackend.make_state


def _walked(*args, **kw):
    """The same run with every compiled schedule retired first (a fresh walk)."""
    pipeline.graph_changed()
    return run(*args, backend_constructor=BC, **kw)


def _compiled(node):
    return [s for s in getattr(node, "_compiled_schedules", {}).values()
            if s.epoch == pipeline._GRAPH_EPOCH.value]


def _random_state(n, seed):
    rng = numpy.random.default_rng(seed)
    psi = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    return psi / numpy.linalg.norm(psi)


def test_replay_matches_walk():
    a, b = Qubit(n=2), Qubit(n=2)
    x, y = Swap(a, b)
    qa, qb = H(x), Not(y)
    c1, c2 = qa.split([1])
    ca, cb = C(Not)(c1, qb)
    sm = StochasticMeasure(QFFT(cb), c2, ca)
    feed = {a: _random_state(2, 1), b: _random_state(2, 2)}
    first = run(sm, feed=feed, backend_constructor=BC)
    assert len(_compiled(sm)) == 1
    again = run(sm, feed=feed, backend_constructor=BC)
    fresh = _walked(sm, feed=feed)
    for got in (again, fresh):
        assert numpy.array_equal(got[0], first[0])
        assert numpy.array_equal(got[1][sm], first[1][sm])


def test_qft_replay_against_ifft():
    n = 8
    psi = _random_state(n, 3)
    q = Qubit(n=n)
    sm = StochasticMeasure(QFFT(q))
    for _ in range(3):
        state, classic = run(sm, feed={q: psi}, backend_constructor=BC)
        assert numpy.max(numpy.abs(state - numpy.fft.ifft(psi, norm="ortho"))) < 1e-12


def test_graph_growth_retires_schedule():
    q = Qubit(n=2)
    h = H(q)
    m = Measure(h)
    feed = {q: _random_state(2, 4)}
    run(m, feed=feed, backend_constructor=BC)
    assert len(_compiled(m)) == 1
    epoch = pipeline._GRAPH_EPOCH.value
    grown = Not(h)                      # a new sink on a node of the compiled graph
    assert pipeline._GRAPH_EPOCH.value > epoch
    assert _compiled(m) == []
    state, _ = run(grown, feed=feed, backend_constructor=BC)
    expect, _ = _walked(grown, feed=feed)
    assert numpy.array_equal(state, expect)


def test_node_overriding_feed_is_walked():
    calls = []

    class Tap(Qubit):
        def feed(self, state, qbitindex, n):
            calls.append([qbitindex[i] for i in self.inputs])
            return super().feed(state, qbitindex, n)

    q = Qubit(n=2)
    t = Tap(H(q))
    feed = {q: _random_state(2, 5)}
    s1, _ = run(t, feed=feed, backend_constructor=BC)
    s2, _ = run(t, feed=feed, backend_constructor=BC)
    assert len(calls) == 2
    assert _compiled(t) == []
    assert numpy.array_equal(s1, s2)


def test_replay_hands_out_fresh_index_groups():
    """A node that reverses the index groups it is given in place does not corrupt later runs."""
    seen = []

    class Scramble(MatrixOp):
        def makemats(self, index_groups):
            return {q: numpy.array([[0.0, 1.0], [1.0, 0.0]]) for g in index_groups for q in g}

        def feed_indices(self, state, index_groups, n):
            seen.append([list(g) for g in index_groups])
            for g in index_groups:
                g.reverse()
            index_groups.reverse()
            return super().feed_indices(state, index_groups, n)

    a, b = Qubit(n=2), Qubit(n=1)
    s = Scramble(H(a), b)
    feed = {a: _random_state(2, 6), b: [0.6, 0.8]}
    outs = [run(s, feed=feed, backend_constructor=BC)[0] for _ in range(3)]
    assert seen[0] == seen[1] == seen[2] == [[0, 1], [2]]
    assert numpy.array_equal(outs[0], outs[1]) and numpy.array_equal(outs[0], outs[2])


def test_concurrent_runs_of_one_graph():
    n = 6
    psi = _random_state(n, 7)
    q = Qubit(n=n)
    sm = StochasticMeasure(QFFT(q))
    want = numpy.fft.ifft(psi, norm="ortho")
    errors = []

    def worker():
        try:
            for _ in range(20):
                state, _ = run(sm, feed={q: psi}, backend_constructor=BC)
                assert numpy.max(numpy.abs(state - want)) < 1e-12
        except Exception as e:          # reported below
            errors.append(e)

    threads = [threading.Thread(target=worker) for _ in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert errors == []
