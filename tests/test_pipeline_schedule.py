"""A graph run again replays its stored schedule (visiting order and index groups); the results must
equal those of a fresh walk of the graph, also after the graph grows and for nodes that override
`feed`, which are always walked."""
import numpy

import qip_amd.pipeline as pipeline
from qip_amd import C, H, Measure, Not, QFFT, Qubit, StochasticMeasure, Swap, run
from oracle.oracle_backend import OracleBackend

BC = OracleBackend.make_state


def _walked(*args, **kw):
    """The same run with every stored schedule retired first (a fresh walk)."""
    pipeline.graph_changed()
    return run(*args, backend_constructor=BC, **kw)


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
    assert any(key[0] == (sm,) for key in pipeline._SCHEDULES)
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
    version = pipeline._GRAPH_VERSION[0]
    grown = Not(h)                      # a new sink on a node of the stored graph
    assert pipeline._GRAPH_VERSION[0] > version
    assert all(stored[0] != pipeline._GRAPH_VERSION[0] for stored in pipeline._SCHEDULES.values())
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
    assert not any(key[0] == (t,) for key in pipeline._SCHEDULES)
    assert numpy.array_equal(s1, s2)
