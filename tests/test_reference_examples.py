"""The reference's example programs (examples/*.py), restated on this package's graph frontend
without their diagram output: every graph runs on the CPU oracle backend and, marked gpu, on the
HIP engine, and the engine must agree with the oracle (states within 1e-12, probabilities within
1e-12, measured outcomes identical for the same seeded uniforms)."""
import math
import random

import numpy
import pytest

from conftest import BACKENDS
from oracle.oracle_backend import OracleBackend
from qip_amd import C, F, H, Measure, Not, QFFT, Qubit, QubitFuncWrapper, StochasticMeasure, run

pytestmark = pytest.mark.parametrize("backend_constructor", BACKENDS, indirect=True)


def shor_circuit(m, n, x, N):
    """examples/shors.py:101-125: a uniform m-qubit register, F(i -> x^i mod N) into an n-qubit
    register, QFFT of the first, then a stochastic measurement of it and a measurement."""
    reg1 = Qubit(n=m, default=numpy.ones(2 ** m) * (2 ** m) ** -0.5)
    reg2 = Qubit(n=n)
    ufreg1, ufreg2 = F(lambda i: pow(x, i, N), reg1, reg2)
    qft = QFFT(ufreg1)
    out1, out2 = Qubit(qft, ufreg2).split()
    return StochasticMeasure(out1), Measure(qft)


def continued_fractions(c, q):
    """examples/shors.py:73-98: the convergents d/r of c/q."""
    coefs = []
    while c > 0 and q > 0:
        a = c // q
        c, q = q, c - q * a
        coefs.append(a)
    d2, d1, r2, r1 = 0, 1, 1, 0
    out = []
    for a in coefs:
        d, r = a * d1 + d2, a * r1 + r2
        out.append((d, r))
        d2, d1, r2, r1 = d1, d, r1, r
    return out


def test_shor_21(backend_constructor):
    """Shor's order finding for x = 11, N = 21 (order 6) on 9 + 5 qubits: the QFT register's
    distribution peaks at multiples of 2^9 / 6, a seeded measurement reads one of them, and the
    convergents of that peak give the order."""
    m, n, x, N = 9, 5, 11, 21
    mqft, meas = shor_circuit(m, n, x, N)
    random.seed(3)
    state, classic = run(mqft, meas, backend_constructor=backend_constructor)
    probs = numpy.asarray(classic[mqft])
    peak = classic[meas][0]
    ref_state, ref_classic = None, None
    if backend_constructor is not OracleBackend.make_state:
        random.seed(3)
        ref_state, ref_classic = run(mqft, meas, backend_constructor=OracleBackend.make_state)
        assert numpy.max(numpy.abs(numpy.asarray(state) - ref_state)) < 1e-12
        assert numpy.max(numpy.abs(probs - numpy.asarray(ref_classic[mqft]))) < 1e-12
        assert ref_classic[meas][0] == peak
    # the order of 11 mod 21 is 6: probability mass sits near multiples of 512/6. The recorded
    # distribution is indexed LSB-first, the measured value MSB-first (qubit 0 = highest bit).
    q = 2 ** m
    perm = [int(format(i, "0{}b".format(m))[::-1], 2) for i in range(q)]
    natural = probs[perm]
    near = [k for k in range(q) if min(abs(k - j * q / 6) for j in range(7)) < 1.0]
    assert abs(probs.sum() - 1.0) < 1e-9
    assert natural[near].sum() > 0.9
    assert abs(natural[peak] - classic[meas][1]) < 1e-12
    assert peak in near
    assert pow(x, 6, N) == 1
    # the convergents of peak / 2^m hold a divisor of the order 6 (peak 0 carries no information)
    orders = [r for d, r in continued_fractions(peak, q) if r > 1]
    assert peak == 0 or any(6 % r == 0 for r in orders)


@QubitFuncWrapper.wrap
def basic_circuit(qa, qb):
    """examples/context_example.py:7-11."""
    h = H(qa)
    cnot = C(Not)
    return cnot(h, qb)


def test_context_example(backend_constructor):
    """examples/context_example.py:13-22: C(basic_circuit) on (c, a, b), i.e. H(a) and CNOT(a, b),
    each controlled on c. With c = |1>: a Bell pair on (a, b); with c = |0>: identity."""
    # C(...) on a QubitFuncWrapper registers the control on the wrapper itself
    # (qip/qubit_util.py:198-201), so the controlled op is built once and called for each input.
    controlled = C(basic_circuit)
    for cval in ([0.0, 1.0], [1.0, 0.0]):
        c, a, b = Qubit(n=1), Qubit(n=1), Qubit(n=1)
        cout, aout, bout = controlled(c, a, b)
        o, _ = run(cout, aout, bout, feed={c: cval, a: [1.0, 0.0], b: [1.0, 0.0]},
                   backend_constructor=backend_constructor)
        o = numpy.asarray(o)
        want = numpy.zeros(8, dtype=complex)
        if cval[1] == 1.0:            # control set: (|00> + |11>)/sqrt(2) on (a, b), c = 1
            want[0b100] = want[0b111] = 2 ** -0.5
        else:
            want[0b000] = 1.0
        assert numpy.max(numpy.abs(o - want)) < 1e-12


def test_qfft_example(backend_constructor):
    """examples/qfft_example.py: two cosines (frequencies 3 and 5.5) on 6 qubits, stochastically
    measured before and after QFFT(rev=True): the spectrum's peaks sit at the frequencies' bins."""
    n = 6
    q = Qubit(n=n)
    mq = StochasticMeasure(q)
    qft = QFFT(mq, rev=True)
    mqft = StochasticMeasure(qft)
    state = numpy.zeros(2 ** n, dtype=numpy.complex128)
    for freq in (3, 5.5):
        state = state + numpy.cos(numpy.linspace(0, 2 * freq * numpy.pi, 2 ** n))
    state = state / numpy.linalg.norm(state)
    o, c = run(mqft, feed={q: state}, backend_constructor=backend_constructor)
    expect = numpy.fft.ifft(state, norm="ortho")
    assert numpy.max(numpy.abs(numpy.asarray(o) - expect)) < 1e-12
    spec = numpy.asarray(c[mqft])
    perm = [int(format(i, "06b")[::-1], 2) for i in range(2 ** n)]
    assert numpy.max(numpy.abs(spec - (numpy.abs(expect) ** 2)[perm])) < 1e-12
    before = numpy.asarray(c[mq])
    assert numpy.max(numpy.abs(before - (numpy.abs(state) ** 2)[perm])) < 1e-12
    top = sorted(range(2 ** n), key=lambda i: -(numpy.abs(expect) ** 2)[i])[:4]
    assert {3, 2 ** n - 3} & set(top)
