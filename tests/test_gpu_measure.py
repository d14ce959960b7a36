"""Device top-k, the state-shrinking reduce_measure, and one-pass measurement (marked gpu).

* measure_top_probabilities (qip/ext/kronprod.pyx:264-315): the true top-k, descending, ties by
  ascending outcome, against the oracle's selection for small registers and, for 23 measured qubits
  of 24, against an independent numpy construction (axis sums of |psi|^2 in measurement order, then
  a lexsort). Selections up to 1024 run on the device, wider ones on the host; both are checked.
* reduce_measure with shrink=True (kronprod.pyx:449-490 with the arena resized): the outcome and its
  probability equal the oracle's for the same random.random() draw, the 2^(n-k) amplitudes equal the
  first 2^(n-k) of the oracle's tiled output, and gates applied afterwards act on the smaller state.
"""
import random

import numpy
import pytest

from oracle import oracle
from qip_amd import GpuBackend

pytestmark = pytest.mark.gpu
RNG = numpy.random.default_rng(2323)


def rand_state(n):
    v = RNG.standard_normal(2 ** n) + 1j * RNG.standard_normal(2 ** n)
    return v / numpy.linalg.norm(v)


def entwine_marginals(psi, n, indices):
    """Marginals with the smallest measured qubit as the outcome's MSB (qubit q = axis q)."""
    p = (numpy.abs(psi) ** 2).reshape((2,) * n)
    keep = sorted(indices)
    rest = tuple(q for q in range(n) if q not in keep)
    return p.sum(axis=rest).reshape(-1)


def numpy_top(marg, top_k):
    order = numpy.lexsort((numpy.arange(marg.size), -marg))[:top_k]
    return [int(m) for m in order], marg[order]


@pytest.mark.parametrize("k", [1, 3, 8, 12])
@pytest.mark.parametrize("dtype", [numpy.complex128, numpy.complex64])
def test_top_k_small_registers_against_oracle(k, dtype):
    n = 16
    psi = rand_state(n)
    idx = [int(q) for q in RNG.permutation(n)[:k]]
    be = GpuBackend.make_state(n, [list(range(n))], [psi], statetype=dtype)
    vec = be.get_state().astype(numpy.complex128)
    for top_k in sorted({1, 3, 50, 1024, 2 ** k, 2 ** k + 5}):
        ms, ps = be.measure_probabilities(idx, top_k=top_k)
        rms, rps = oracle.measure_top_probabilities(idx, n, min(top_k, 300), vec)
        assert len(ms) == min(top_k, 2 ** k)
        assert ms[:len(rms)] == rms
        assert numpy.max(numpy.abs(numpy.array(ps[:len(rps)]) - rps)) < 1e-14
        assert all(a >= b for a, b in zip(ps, ps[1:]))
    be.close()


def test_top_k_twenty_three_of_twenty_four_qubits():
    n = 24
    psi = rand_state(n)
    idx = [int(q) for q in RNG.permutation(n)[:23]]
    print("top-k on", idx, flush=True)
    be = GpuBackend.make_state(n, [list(range(n))], [psi])
    marg = entwine_marginals(psi, n, idx)
    for top_k in (1, 7, 1000, 1024, 1025, 3000):
        ms, ps = be.measure_probabilities(idx, top_k=top_k)
        rms, rps = numpy_top(marg, top_k)
        assert ms == rms, top_k
        assert numpy.max(numpy.abs(numpy.array(ps) - rps) / rps) < 1e-12
    be.close()


def test_top_k_ties_by_ascending_outcome():
    """Equal probabilities: the smaller outcome first, across chunk and round boundaries."""
    n = 14
    psi = numpy.ones(2 ** n) / 2 ** (n / 2)
    be = GpuBackend.make_state(n, [list(range(n))], [psi])
    idx = [2, 5, 0, 9, 11, 13, 7, 1, 3, 4, 6, 8]        # 12 qubits: 4096 equal outcomes, two chunks
    for top_k in (1, 5, 1000, 1024):
        ms, ps = be.measure_probabilities(idx, top_k=top_k)
        assert ms == list(range(top_k))
        assert all(p == ps[0] for p in ps)
    # pairs of equal values: outcome m and m ^ 1 on the last measured bit
    amp = numpy.repeat(numpy.linspace(1.0, 2.0, 2 ** (n - 1)), 2)
    amp = amp / numpy.linalg.norm(amp)
    be.overwrite_relative_range(0, 2 ** n, amp.astype(numpy.complex128))
    idx = list(range(n))
    ms, ps = be.measure_probabilities(idx, top_k=6)
    rms, _ = numpy_top(entwine_marginals(amp, n, idx), 6)
    assert ms == rms == [2 ** n - 2, 2 ** n - 1, 2 ** n - 4, 2 ** n - 3, 2 ** n - 6, 2 ** n - 5]
    be.close()


@pytest.mark.parametrize("n,k", [(12, 1), (12, 4), (12, 11), (18, 5), (20, 9)])
@pytest.mark.parametrize("dtype", [numpy.complex128, numpy.complex64])
def test_reduce_measure_shrink_against_oracle(n, k, dtype):
    psi = rand_state(n)
    idx = [int(q) for q in RNG.permutation(n)[:k]]
    print("shrink n", n, "indices", idx, numpy.dtype(dtype).name, flush=True)
    be = GpuBackend.make_state(n, [list(range(n))], [psi], statetype=dtype)
    vec = be.get_state().astype(numpy.complex128)
    random.seed(n * 100 + k)
    m, p = be.reduce_measure(idx, shrink=True)
    random.seed(n * 100 + k)
    out = numpy.zeros_like(vec)
    rm, rp = oracle.reduce_measure(idx, n, vec, out)
    assert m == rm
    assert abs(p - rp) < 1e-12
    assert be.n == n - k and be.get_state_size() == 2 ** (n - k)
    got = be.get_state().astype(numpy.complex128)
    want = out[:2 ** (n - k)]
    bar = 1e-12 if dtype == numpy.complex128 else 2e-6 * numpy.max(numpy.abs(want))
    assert numpy.max(numpy.abs(got - want)) < bar
    assert abs(be.total_prob() - 1.0) < (1e-12 if dtype == numpy.complex128 else 1e-5)
    # the smaller state goes on like any other
    U = numpy.array([[0.6, 0.8j], [0.8j, 0.6]])
    q = (n - k) // 2
    be.kronselect_dot({q: U})
    after = numpy.zeros_like(want)
    oracle.kronselect_dot({q: U}, want, n - k, after)
    assert numpy.max(numpy.abs(be.get_state().astype(numpy.complex128) - after)) < bar * 2
    be.close()


def test_reduce_measure_shrink_errors_and_tiled_default():
    n = 10
    psi = rand_state(n)
    be = GpuBackend.make_state(n, [list(range(n))], [psi])
    with pytest.raises(ValueError):
        be.reduce_measure(list(range(n)), shrink=True)          # nothing would be left
    assert be.n == n
    random.seed(5)
    m, p = be.reduce_measure([3, 7])                             # the reference's tiled result
    random.seed(5)
    out = numpy.zeros_like(psi)
    oracle.reduce_measure([3, 7], n, psi, out)
    assert numpy.max(numpy.abs(be.get_state() - out)) < 1e-12
    be.close()


def test_fused_backend_records_after_shrink():
    """The precision form keeps recording on the shrunk state: gates after a shrink act on n - k qubits."""
    from qip_amd import infidelity
    from qip_amd.fused import FusedBackend
    from qip_amd.workloads import layered_random_circuit
    n, k = 16, 3
    ops = layered_random_circuit(n, 3, seed=8)
    after = layered_random_circuit(n - k, 3, seed=9)
    fz = FusedBackend.make_state(n, [], [])
    pl = GpuBackend.make_state(n, [], [])
    for d in ops:
        fz.kronselect_dot(d)
        pl.kronselect_dot(d)
    for be in (fz, pl):
        be.reduce_measure([1, 6, 11], measured=5, measured_prob=0.125, shrink=True)
        for d in after:
            be.kronselect_dot(d)
    assert fz.n == pl.n == n - k
    assert infidelity(pl, fz) < 1e-8
    fz.close()
    pl.close()
