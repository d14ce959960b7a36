"""The reference's own test suite, restated against this package's graph frontend.

Each test cites the reference test it restates (tests/qiptest.py, tests/qfttest.py,
tests/qubit_util_test.py, tests/utiltest.py) and keeps its inputs and expectations. Every graph
test runs on the CPU oracle backend (the reference's algorithm) and, marked gpu, on the HIP
engine.
"""
import numpy
import pytest

from conftest import BACKENDS
from qip_amd import (C, F, H, Measure, Not, NotOp, QFFT, Qubit, QubitFuncWrapper, QubitWrapperContext, Rm, StochasticMeasure,
                     Swap, SwapOp, run)

pytestmark = pytest.mark.parametrize("backend_constructor", BACKENDS, indirect=True)


def R(*args, bc, **kw):
    return run(*args, backend_constructor=bc, **kw)


# ---- tests/qiptest.py ---------------------------------------------------------------------------

def test_simple_pipeline(backend_constructor):            # qiptest.py:9-13
    q = Qubit(n=1)
    o, _ = Not(q).run(feed={q: [1.0, 0.0]}, backend_constructor=backend_constructor)
    assert numpy.allclose(o, [0., 1.])


def test_simple_pipeline_inv(backend_constructor):        # qiptest.py:15-19
    q = Qubit(n=1)
    o, _ = Not(q).run(feed={q: [0.0, 1.0]}, backend_constructor=backend_constructor)
    assert numpy.allclose(o, [1., 0.])


def test_bell(backend_constructor):                       # qiptest.py:21-25
    q = Qubit(n=1)
    o, _ = H(q).run(feed={q: [1.0, 0.0]}, backend_constructor=backend_constructor)
    assert o[0] == o[1]


def test_h_first(backend_constructor):                    # qiptest.py:27-43
    bc = backend_constructor
    q1, q2 = Qubit(n=1), Qubit(n=1)
    q3, q4 = Qubit(H(q1), H(q2)).split()
    o1, _ = H(q3).run(feed={q1: [1.0, 0.0], q2: [1.0, 0.0]}, backend_constructor=bc)
    q1, q2 = Qubit(n=1), Qubit(n=1)
    q3, q4 = Qubit(q1, H(q2)).split()
    o2, _ = q3.run(feed={q1: [1.0, 0.0], q2: [1.0, 0.0]}, backend_constructor=bc)
    assert numpy.allclose(o1, o2)


def test_bell_inv(backend_constructor):                   # qiptest.py:45-49
    q = Qubit(n=1)
    o, _ = H(q).run(feed={q: [0.0, 1.0]}, backend_constructor=backend_constructor)
    assert o[0] == -o[1]


def test_multi_not(backend_constructor):                  # qiptest.py:51-56
    q1, q2 = Qubit(n=1), Qubit(n=1)
    s = NotOp(q1, q2)
    o, _ = s.run(feed={q1: [1.0, 0.0], q2: [1.0, 0.0]}, backend_constructor=backend_constructor)
    assert numpy.allclose(o, [0.0, 0.0, 0.0, 1.0])


def test_swap(backend_constructor):                       # qiptest.py:58-72
    q1, q2 = Qubit(n=1), Qubit(n=1)
    s = SwapOp(q1, q2)
    for f1, f2, want in (([1.0, 0.0], [1.0, 0.0], [1, 0, 0, 0]), ([1.0, 0.0], [0.0, 1.0], [0, 0, 1, 0]),
                         ([0.0, 1.0], [1.0, 0.0], [0, 1, 0, 0]), ([0.0, 1.0], [0.0, 1.0], [0, 0, 0, 1])):
        o, _ = s.run(feed={q1: f1, q2: f2}, backend_constructor=backend_constructor)
        assert numpy.allclose(o, want)


def test_swap_twobit(backend_constructor):                # qiptest.py:74-80
    q1, q2 = Qubit(n=2), Qubit(n=2)
    s = SwapOp(q1, q2)
    o, _ = s.run(feed={q1: [0.0, 1.0, 0.0, 0.0], q2: [1.0, 0.0, 0.0, 0.0]}, backend_constructor=backend_constructor)
    assert o[1] == 1.0
    assert sum(abs(o)) == 1.0


def test_split(backend_constructor):                      # qiptest.py:82-98
    bc = backend_constructor
    q1, q2 = Qubit(n=1), Qubit(n=1)
    n1, n2 = Not(q1, q2)
    f1, f2 = Not(n1), Not(n2)
    o, _ = f1.run(feed={q1: [1.0, 0.0], q2: [1.0, 0.0]}, backend_constructor=bc)
    assert numpy.allclose(o, [0.0, 1.0, 0.0, 0.0])
    o, _ = f2.run(feed={q1: [1.0, 0.0], q2: [1.0, 0.0]}, backend_constructor=bc)
    assert numpy.allclose(o, [0.0, 0.0, 1.0, 0.0])
    o, _ = R(f1, f2, feed={q1: [1.0, 0.0], q2: [1.0, 0.0]}, bc=bc)
    assert numpy.allclose(o, [1.0, 0.0, 0.0, 0.0])


def test_allsplit(backend_constructor):                   # qiptest.py:100-109
    n = 5
    q = Qubit(n=n)
    qs = q.split(range(n + 1))
    assert len(qs) == n
    newq = Qubit(*qs)
    state = numpy.cos(numpy.linspace(0, numpy.pi, 2 ** n))
    o, _ = R(newq, feed={q: state}, bc=backend_constructor)
    assert numpy.allclose(state, o)


def test_cnot(backend_constructor):                       # qiptest.py:111-118
    q1, q2 = Qubit(n=1), Qubit(n=1)
    c1, c2 = C(Not)(q1, q2)
    o, _ = R(c1, c2, feed={q1: [1.0, 0.0], q2: [1.0, 0.0]}, bc=backend_constructor)
    assert numpy.allclose(o, [1.0, 0.0, 0.0, 0.0])
    o, _ = R(c1, c2, feed={q1: [0.0, 1.0], q2: [1.0, 0.0]}, bc=backend_constructor)
    assert numpy.allclose(o, [0.0, 0.0, 0.0, 1.0])


def test_default(backend_constructor):                    # qiptest.py:120-129
    q1, q2 = Qubit(n=1, default=[0.0, 1.0]), Qubit(n=1)
    o1, _ = R(q1, q2, feed={q2: [1.0, 0.0]}, bc=backend_constructor)
    q3, q4 = Qubit(n=1), Qubit(n=1)
    o2, _ = R(q3, q4, feed={q3: [0.0, 1.0], q4: [1.0, 0.0]}, bc=backend_constructor)
    assert numpy.allclose(o1, o2)


def test_default2(backend_constructor):                   # qiptest.py:131-146
    n = 2
    bc = backend_constructor
    reg11, reg21 = Qubit(n=n, default=[1, 2, 3, 4]), Qubit(n=n, default=[1, 0, 0, 0])
    o1, _ = R(reg11, reg21, bc=bc)
    reg12, reg22 = Qubit(n=n, default=[1, 2, 3, 4]), Qubit(n=n)
    o2, _ = R(reg12, reg22, feed={reg22: [1, 0, 0, 0]}, bc=bc)
    reg13, reg23 = Qubit(n=n), Qubit(n=n)
    o3, _ = R(reg13, reg23, feed={reg13: [1, 2, 3, 4], reg23: [1, 0, 0, 0]}, bc=bc)
    assert numpy.allclose(o1, o2)
    assert numpy.allclose(o1, o3)


def _cswap_p0(o):
    return numpy.dot(o[:len(o) >> 1], o[:len(o) >> 1])


def test_cswap_compare(backend_constructor):              # qiptest.py:148-169
    q1, q2, q3 = Qubit(n=1), Qubit(n=1), Qubit(n=1)
    c1, c2, c3 = C(Swap)(H(q1), q2, q3)
    m1 = H(c1)
    o, _ = m1.run(feed={q1: [1.0, 0.0], q2: [0.0, 1.0], q3: [1.0, 0.0]}, backend_constructor=backend_constructor)
    assert abs(_cswap_p0(o) - 0.5) <= 1e-15
    o, _ = m1.run(feed={q1: [1.0, 0.0], q2: [1.0, 0.0], q3: [1.0, 0.0]}, backend_constructor=backend_constructor)
    assert abs(_cswap_p0(o) - 1.0) <= 1e-15


def test_cswap_twobitcompare(backend_constructor):        # qiptest.py:171-192
    q1, q2, q3 = Qubit(n=1), Qubit(n=2), Qubit(n=2)
    c1, c2, c3 = C(Swap)(H(q1), q2, q3)
    m1 = H(c1)
    o, _ = m1.run(feed={q1: [1.0, 0.0], q2: [0.0, 0.0, 0.0, 1.0], q3: [1.0, 0.0, 0.0, 0.0]},
                  backend_constructor=backend_constructor)
    assert abs(_cswap_p0(o) - 0.5) <= 1e-15
    o, _ = m1.run(feed={q1: [1.0, 0.0], q2: [1.0, 0.0, 0.0, 0.0], q3: [1.0, 0.0, 0.0, 0.0]},
                  backend_constructor=backend_constructor)
    assert abs(_cswap_p0(o) - 1.0) <= 1e-15


def test_cswap_5bitcompare(backend_constructor):          # qiptest.py:194-228
    q1, q2, q3 = Qubit(n=1), Qubit(n=5), Qubit(n=5)
    c1, c2, c3 = C(Swap)(H(q1), q2, q3)
    m1 = H(c1)
    s2, s3 = numpy.zeros((32,)), numpy.zeros((32,))
    s2[0], s3[1] = 1.0, 1.0
    o, _ = m1.run(feed={q1: [1.0, 0.0], q2: s2, q3: s3}, backend_constructor=backend_constructor)
    assert abs(_cswap_p0(o) - 0.5) <= 1e-15
    s2, s3 = numpy.zeros((32,)), numpy.zeros((32,))
    s2[0], s3[0] = 1.0, 1.0
    o, _ = m1.run(feed={q1: [1.0, 0.0], q2: s2, q3: s3}, backend_constructor=backend_constructor)
    assert abs(_cswap_p0(o) - 1.0) <= 1e-15


def test_many_bits(backend_constructor):                  # qiptest.py:230-235
    q1, q2 = Qubit(n=19), Qubit(n=1)
    o, _ = R(q1, Not(q2), feed={q2: [1.0, 0.0]}, bc=backend_constructor)
    assert o[1] == numpy.sum(numpy.abs(o))


def test_not_ordered(backend_constructor):                # qiptest.py:237-259
    feed3 = [1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0]
    q1, q2, q3 = Qubit(n=1), Qubit(n=1), Qubit(n=3)
    n1, n2, n3 = Not(q1), H(q2), Not(q3)
    o1, _ = R(n1, n2, n3, feed={q1: [1.0, 0.0], q2: [1.0, 0.0], q3: feed3}, bc=backend_constructor)
    q1, q2, q3 = Qubit(n=1), Qubit(n=1), Qubit(n=3)
    n2, n3, n1 = H(q2), Not(q3), Not(q1)
    o2, _ = R(n1, n2, n3, feed={q1: [1.0, 0.0], q2: [1.0, 0.0], q3: feed3}, bc=backend_constructor)
    assert numpy.allclose(o1, o2)


def test_cswap_measure(backend_constructor):              # qiptest.py:261-276
    q1, q2, q3 = Qubit(n=1), Qubit(n=2), Qubit(n=2)
    c1, c2, c3 = C(Swap)(H(q1), q2, q3)
    m1 = Measure(H(c1))
    o, classic = m1.run(feed={q1: [1.0, 0.0], q2: [0.0, 0.0, 0.0, 1.0], q3: [1.0, 0.0, 0.0, 0.0]},
                        backend_constructor=backend_constructor)
    measured, measured_prob = classic[m1]
    assert measured_prob == pytest.approx(0.5, abs=1e-7)


def test_cswap_reorder_measure(backend_constructor):      # qiptest.py:278-294
    q2, q1, q3 = Qubit(n=2), Qubit(n=1), Qubit(n=2)
    c1, c2, c3 = C(Swap)(H(q1), q2, q3)
    m1 = Measure(H(c1))
    o, classic = m1.run(feed={q1: [1.0, 0.0], q2: [0.0, 0.0, 0.0, 1.0], q3: [1.0, 0.0, 0.0, 0.0]},
                        backend_constructor=backend_constructor)
    measured, measured_prob = classic[m1]
    assert measured_prob == pytest.approx(0.5, abs=1e-7)


def test_measure_stochastic(backend_constructor):         # qiptest.py:296-302
    q1, q2 = Qubit(n=2), Qubit(n=2)
    m1 = StochasticMeasure(q2)
    _, c = R(m1, q1, feed={q1: [0.0, 0.0, 1.0, 0.0], q2: [0.5, 0.5, 0.5, 0.5]}, bc=backend_constructor)
    assert numpy.allclose(c[m1], [0.25, 0.25, 0.25, 0.25])


def test_measure_stochastic_top(backend_constructor):     # qiptest.py:304-315
    q1, q2 = Qubit(n=2), Qubit(n=2)
    m1 = StochasticMeasure(q2, top_k=3)
    _, c = R(m1, q1, feed={q1: [0.0, 0.0, 1.0, 0.0], q2: [0.6, 0.8, 0.0, 0.0]}, bc=backend_constructor)
    assert c[m1][0][0] == 1
    assert c[m1][0][1] == 0
    assert numpy.allclose(c[m1][1], [0.64, 0.36, 0.0])


def test_fop(backend_constructor):                        # qiptest.py:317-328
    n = 2
    reg1, reg2 = Qubit(n=n, default=[1, 2, 3, 4]), Qubit(n=n, default=[1, 0, 0, 0])
    freg1, freg2 = F(lambda x: 1, reg1, reg2)
    o1, _ = R(freg1, freg2, bc=backend_constructor)
    mock1, mock2 = Qubit(n=n, default=[1, 2, 3, 4]), Qubit(n=n, default=[0.0, 1.0, 0, 0])
    o2, _ = R(mock1, mock2, bc=backend_constructor)
    assert numpy.allclose(o1, o2)


def test_fop_withq(backend_constructor):                  # qiptest.py:330-342
    n = 2
    reg1, reg2 = Qubit(n=n, default=[1, 2, 3, 4]), Qubit(n=n, default=[0, 1, 0, 0])
    freg1, freg2 = F(lambda x: 1, reg1, reg2)
    o1, _ = R(freg1, freg2, bc=backend_constructor)
    mock1, mock2 = Qubit(n=n, default=[1, 2, 3, 4]), Qubit(n=n, default=[1, 0, 0, 0])
    o2, _ = R(mock1, mock2, bc=backend_constructor)
    assert numpy.allclose(o1, o2)


def test_fop_regsize(backend_constructor):                # qiptest.py:344-349 (asserts nothing there)
    n = 2
    reg1 = Qubit(n=2 * n, default=numpy.ones(2 ** (2 * n)) * pow(2 ** (2 * n), -0.5))
    reg2 = Qubit(n=n)
    freg1, freg2 = F(lambda x: pow(3, x, 2 ** n), reg1, reg2)
    o1, _ = R(freg1, freg2, bc=backend_constructor)
    # |x>|0> -> |x>|3^x mod 4>: each x keeps amplitude 1/4 on its one image
    expect = numpy.zeros(2 ** (3 * n))
    for x in range(2 ** (2 * n)):
        expect[(x << n) | pow(3, x, 2 ** n)] = 0.25
    assert numpy.allclose(o1, expect)


def test_basic_rop(backend_constructor):                  # qiptest.py:351-356
    q1, q2 = Qubit(n=1), Qubit(n=1)
    o, _ = R(q1, Rm(2, q2), bc=backend_constructor)
    assert numpy.allclose(numpy.abs(o), [1, 0, 0, 0])


def test_rop(backend_constructor):                        # qiptest.py:358-363
    q1, q2 = Qubit(n=1), Qubit(n=1, default=1)
    o, _ = R(q1, Rm(2, q2), bc=backend_constructor)
    assert numpy.allclose(numpy.abs(o), [0, 1.0, 0, 0])


def test_toffoli(backend_constructor):                    # qiptest.py:365-375
    q1, q2, q3 = Qubit(n=1), Qubit(n=1), Qubit(n=1)
    c1, c2, c3 = C(C(Not))(q1, q2, q3)
    o, c = R(c1, c2, c3, feed={q1: [pow(2, -0.5), pow(2, -0.5)], q2: [pow(2, -0.5), pow(2, -0.5)]},
             bc=backend_constructor)
    assert numpy.allclose(numpy.abs(o), [0.5, 0, 0.5, 0, 0.5, 0, 0, 0.5])


def test_rop_tuplefeed(backend_constructor):              # qiptest.py:377-387
    q1, q2 = Qubit(n=1), Qubit(n=1)
    r2 = Rm(2, q2)
    o1, _ = R(q1, r2, feed={q1: [1.0, 0.0], q2: [0.0, 1.0]}, bc=backend_constructor)
    o2, _ = R(q1, r2, feed={(q1, q2): [0.0, 1.0, 0.0, 0.0]}, bc=backend_constructor)
    o3, _ = R(q1, r2, feed={(q2, q1): [0.0, 0.0, 1.0, 0.0]}, bc=backend_constructor)
    for o in (o1, o2, o3):
        assert numpy.allclose(numpy.abs(o), [0, 1.0, 0, 0])


# ---- tests/qfttest.py ---------------------------------------------------------------------------

def test_qft_simple_pipeline(backend_constructor):        # qfttest.py:12-16
    q = Qubit(n=2)
    o, _ = QFFT(q).run(feed={q: [0.5, 0.5, 0.5, 0.5]}, backend_constructor=backend_constructor)
    assert numpy.allclose(o, [1., 0., 0., 0.])


def test_qft_cosine_pipeline(backend_constructor):        # qfttest.py:18-29
    freq, n = 5, 6
    q = Qubit(n=n)
    state = numpy.cos(numpy.linspace(0, 2 * freq * numpy.pi, 2 ** n))
    state = state / numpy.linalg.norm(state)
    o, _ = QFFT(q).run(feed={q: state}, backend_constructor=backend_constructor)
    out = numpy.power(numpy.abs(o), 2)
    assert numpy.argmax(out[:int(len(o) / 2)]) == freq


# ---- tests/qubit_util_test.py -------------------------------------------------------------------

S = [pow(2, -0.5), pow(2, -0.5)]


def test_basic_context(backend_constructor):              # qubit_util_test.py:10-27
    q1, q2 = Qubit(n=1, default=S), Qubit(n=1)
    c1, c2 = C(H)(q1, q2)
    o1, _ = R(c1, c2, bc=backend_constructor)
    q3, q4 = Qubit(n=1, default=S), Qubit(n=1)
    with QubitWrapperContext(C, [q3]):
        h = H(q4)
    o2, _ = R(h, bc=backend_constructor)
    assert numpy.allclose(o1, o2)


def test_explicit_context(backend_constructor):           # qubit_util_test.py:29-47
    q1, q2, q3 = Qubit(n=1, default=S), Qubit(n=1, default=S), Qubit(n=1)
    c1, c2, c3 = C(C(H))(q1, q2, q3)
    o1, _ = R(c1, c2, c3, bc=backend_constructor)
    q4, q5, q6 = Qubit(n=1, default=S), Qubit(n=1, default=S), Qubit(n=1)
    with QubitWrapperContext(C, [q4]):
        c4, c5 = C(H)(q5, q6)
    o2, _ = R(c4, c5, bc=backend_constructor)
    assert numpy.allclose(o1, o2)


def test_function_wrapper(backend_constructor):           # qubit_util_test.py:49-67
    @QubitFuncWrapper.wrap
    def circuit(*args):
        return H(*args)

    q1, q2 = Qubit(n=1, default=S), Qubit(n=2)
    c1, c2 = C(H)(q1, q2)
    o1, _ = R(c1, c2, bc=backend_constructor)
    q3, q4 = Qubit(n=1, default=S), Qubit(n=2)
    c3, c4 = C(circuit)(q3, q4)
    o2, _ = R(c3, c4, bc=backend_constructor)
    assert numpy.allclose(o1, o2)


def test_function_nested_c(backend_constructor):          # qubit_util_test.py:69-89
    @QubitFuncWrapper.wrap
    def circuit(*args):
        return C(H)(*args)

    q1, q2, q3 = Qubit(n=1, default=S), Qubit(n=1, default=S), Qubit(n=1)
    c1, c2, c3 = C(C(H))(q1, q2, q3)
    o1, _ = R(c1, c2, c3, bc=backend_constructor)
    q4, q5, q6 = Qubit(n=1, default=S), Qubit(n=1, default=S), Qubit(n=1)
    c4, c5, c6 = C(circuit)(q4, q5, q6)
    o2, _ = R(c4, c5, c6, bc=backend_constructor)
    assert numpy.allclose(o1, o2)


def test_function_nested_wrapper(backend_constructor):    # qubit_util_test.py:91-115
    @QubitFuncWrapper.wrap
    def circuit_b(*args):
        return H(*args)

    @QubitFuncWrapper.wrap
    def circuit_a(*args):
        return C(circuit_b)(*args)

    q1, q2, q3 = Qubit(n=1, default=S), Qubit(n=1, default=S), Qubit(n=1)
    c1, c2, c3 = C(C(H))(q1, q2, q3)
    o1, _ = R(c1, c2, c3, bc=backend_constructor)
    q4, q5, q6 = Qubit(n=1, default=S), Qubit(n=1, default=S), Qubit(n=1)
    c4, c5, c6 = C(circuit_a)(q4, q5, q6)
    o2, _ = R(c4, c5, c6, bc=backend_constructor)
    assert numpy.allclose(o1, o2)


def test_function_double_nested_c(backend_constructor):   # qubit_util_test.py:117-139
    @QubitFuncWrapper.wrap
    def circuit(*args):
        return C(C(H))(*args)

    q1, q2, q3, q4 = Qubit(n=1, default=S), Qubit(n=1, default=S), Qubit(n=1, default=S), Qubit(n=1)
    c1, c2, c3, c4 = C(C(C(H)))(q1, q2, q3, q4)
    o1, _ = R(c1, c2, c3, c4, bc=backend_constructor)
    q5, q6, q7, q8 = Qubit(n=1, default=S), Qubit(n=1, default=S), Qubit(n=1, default=S), Qubit(n=1)
    c5, c6, c7, c8 = C(circuit)(q5, q6, q7, q8)
    o2, _ = R(c5, c6, c7, c8, bc=backend_constructor)
    assert numpy.allclose(o1, o2)
