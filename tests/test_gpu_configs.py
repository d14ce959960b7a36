"""GPU parity of the five synthetic workloads at their stated sizes (marked gpu).

* 16-qubit QFT graph (qip/qfft.py:8-43) on a seeded Gaussian state: the state against
  ifft(norm="ortho") and the oracle, the StochasticMeasure probabilities against the oracle in the
  reference's LSB-first order (qip/ext/kronprod.pyx:238-261); the graph run again, and an identical
  graph built anew, give the same bits.
* 28-qubit depth-20 layered random circuit: plain complex128 against plain complex64 (unfused);
  the same construction at 24 qubits, both precisions, against the oracle's FP64 cdot_loop.
* 30-qubit depth-40 circuit: the precision form (fused split-BF16 MFMA, complex64) against plain
  complex128; at 24 qubits, depth 40, both against the oracle.
* All 512 circuits of the batched workload, each against the plain complex128 engine.

The 24-qubit oracle run is shared: the depth-20 circuit is the first 20 layers of the depth-40 one
(same seed, the same draws in the same order), so one FP64 pass through 40 layers gives both.
"""
import os

import numpy
import pytest

from qip_amd import GpuBackend, infidelity, max_abs_diff

pytestmark = pytest.mark.gpu


def _threads() -> int:
    return max(1, min(16, len(os.sched_getaffinity(0))))


def _host_infidelity(a: numpy.ndarray, b: numpy.ndarray) -> float:
    a = numpy.asarray(a, dtype=numpy.complex128)
    b = numpy.asarray(b, dtype=numpy.complex128)
    return float(1.0 - abs(numpy.vdot(a, b)) ** 2 / (numpy.vdot(a, a).real * numpy.vdot(b, b).real))


# ---- 16-qubit QFT graph ------------------------------------------------------------------------

def _qft16_graph():
    from qip_amd import QFFT, Qubit, StochasticMeasure
    q = Qubit(n=16)
    sm = StochasticMeasure(QFFT(q))
    return q, sm


def test_qft16_graph_state_and_probabilities():
    from oracle.oracle_backend import OracleBackend
    from qip_amd import run
    n = 16
    rng = numpy.random.default_rng(1234)
    psi = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    psi /= numpy.linalg.norm(psi)
    print("qft16 graph", flush=True)
    q, sm = _qft16_graph()
    s1, c1 = run(sm, feed={q: psi}, backend_constructor=GpuBackend.make_state)
    s2, c2 = run(sm, feed={q: psi}, backend_constructor=GpuBackend.make_state)
    ro, co = run(sm, feed={q: psi}, backend_constructor=OracleBackend.make_state)
    expect = numpy.fft.ifft(psi, norm="ortho")
    assert numpy.max(numpy.abs(s1 - expect)) < 1e-12
    assert numpy.max(numpy.abs(s1 - ro)) < 1e-12
    p1 = numpy.asarray(c1[sm])
    assert p1.shape == (2 ** n,)
    assert numpy.max(numpy.abs(p1 - numpy.asarray(co[sm]))) < 1e-14
    # LSB-first order: qubit j of the register at bit j of the outcome, i.e. the state-order
    # probabilities with their 16 index bits reversed
    rev = numpy.array([int(format(i, "016b")[::-1], 2) for i in range(2 ** n)])
    assert numpy.max(numpy.abs(p1[rev] - numpy.abs(expect) ** 2)) < 1e-14
    # the same graph run again, and an identical graph built anew, give the same bits
    assert numpy.array_equal(s1, s2)
    assert numpy.array_equal(p1, numpy.asarray(c2[sm]))
    q2, sm2 = _qft16_graph()
    s3, c3 = run(sm2, feed={q2: psi}, backend_constructor=GpuBackend.make_state)
    assert numpy.array_equal(s1, s3)
    assert numpy.array_equal(p1, numpy.asarray(c3[sm2]))


def test_qft16_graph_complex64_and_fused():
    """The same graph on the complex64 plain engine and on the precision form."""
    from qip_amd import run
    from qip_amd.fused import FusedBackend
    n = 16
    rng = numpy.random.default_rng(99)
    psi = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    psi /= numpy.linalg.norm(psi)
    expect = numpy.fft.ifft(psi, norm="ortho")
    q, sm = _qft16_graph()

    def c64(n_, groups, feeds, statetype=None, **kw):
        return GpuBackend.make_state(n_, groups, feeds, statetype=numpy.complex64, **kw)
    for ctor in (c64, FusedBackend.make_state):
        s, c = run(sm, feed={q: psi}, backend_constructor=ctor)
        assert s.dtype == numpy.complex64
        assert _host_infidelity(expect, s) < 1e-8
        assert numpy.max(numpy.abs(s - expect)) < 2e-5 * numpy.max(numpy.abs(expect))


# ---- 24-qubit oracle runs of the layered circuit -----------------------------------------------

@pytest.fixture(scope="module")
def oracle24():
    """FP64 states of the 24-qubit layered circuit (seed 24) after 20 and after 40 layers, from the
    oracle's cdot_loop restatement, one out-of-place pass per descriptor."""
    from oracle import oracle
    from qip_amd.workloads import layered_random_circuit
    n = 24
    ops40 = layered_random_circuit(n, 40, seed=24)
    ops20 = layered_random_circuit(n, 20, seed=24)
    assert all(a.keys() == b.keys() for a, b in zip(ops20, ops40))
    v = numpy.zeros(2 ** n, dtype=numpy.complex128)
    v[0] = 1.0
    out = numpy.zeros_like(v)
    snap = None
    print("oracle24: {} gates on {} threads".format(len(ops40), _threads()), flush=True)
    for i, d in enumerate(ops40):
        oracle.kronselect_dot(d, v, n, out, fast=True, nthreads=_threads())
        v, out = out, v
        if i + 1 == len(ops20):
            snap = v.copy()
        if (i + 1) % 200 == 0:
            print("oracle24: {} gates".format(i + 1), flush=True)
    return {"n": n, "ops20": ops20, "ops40": ops40, "s20": snap, "s40": v}


def _run_plain(n, ops, dtype):
    be = GpuBackend.make_state(n, [], [], statetype=dtype)
    for d in ops:
        be.kronselect_dot(d)
    return be


def test_circuit24_depth20_both_precisions_against_oracle(oracle24):
    n = oracle24["n"]
    want = oracle24["s20"]
    for dtype, bar_f, bar_a in ((numpy.complex128, 1e-13, 1e-12), (numpy.complex64, 1e-9, 2e-6)):
        be = _run_plain(n, oracle24["ops20"], dtype)
        got = be.get_state()
        be.close()
        f = _host_infidelity(want, got)
        e = float(numpy.max(numpy.abs(got - want)))
        print("24q depth 20 {}: 1-F {:.3e} max err {:.3e}".format(numpy.dtype(dtype).name, f, e), flush=True)
        assert f < bar_f and e < bar_a


def test_circuit24_depth40_fused_and_plain_against_oracle(oracle24):
    from qip_amd.fused import FusedBackend
    n = oracle24["n"]
    want = oracle24["s40"]
    be = _run_plain(n, oracle24["ops40"], numpy.complex128)
    got = be.get_state()
    be.close()
    assert _host_infidelity(want, got) < 1e-13
    assert numpy.max(numpy.abs(got - want)) < 1e-12
    fz = FusedBackend.make_state(n, [], [])
    for d in oracle24["ops40"]:
        fz.kronselect_dot(d)
    got = fz.get_state()
    blocks = fz.blocks_applied
    fz.close()
    f = _host_infidelity(want, got)
    e = float(numpy.max(numpy.abs(got.astype(numpy.complex128) - want)))
    print("24q depth 40 fused ({} blocks): 1-F {:.3e} max err {:.3e}".format(blocks, f, e), flush=True)
    assert f < 1e-8
    assert e < 1e-6


# ---- 28 and 30 qubits at full size ---------------------------------------------------------------

def test_circuit28_depth20_complex64_against_complex128():
    from qip_amd.workloads import layered_random_circuit
    n = 28
    ops = layered_random_circuit(n, 20, seed=2028)
    assert len(ops) == 830
    print("28q depth 20: plain c128", flush=True)
    a = _run_plain(n, ops, numpy.complex128)
    print("28q depth 20: plain c64", flush=True)
    b = _run_plain(n, ops, numpy.complex64)
    f = infidelity(a, b)
    e = max_abs_diff(a, b)
    na = a.total_prob()
    a.close()
    b.close()
    print("28q depth 20: 1-F(c64, c128) {:.3e}, max err {:.3e}".format(f, e), flush=True)
    assert abs(na - 1.0) < 1e-12
    assert f < 1e-10
    assert e < 2e-5 * 2.0 ** (-n / 2) * 64


def test_circuit30_depth40_fused_against_plain_complex128():
    from qip_amd.fused import FusedBackend
    from qip_amd.workloads import layered_random_circuit
    n = 30
    ops = layered_random_circuit(n, 40, seed=2024)
    assert len(ops) == 1780
    print("30q depth 40: fused c64", flush=True)
    fz = FusedBackend.make_state(n, [], [])
    for d in ops:
        fz.kronselect_dot(d)
    fz.flush()
    blocks = fz.blocks_applied
    print("30q depth 40: plain c128", flush=True)
    ref = _run_plain(n, ops, numpy.complex128)
    f = infidelity(ref, fz)
    e = max_abs_diff(ref, fz)
    nf = fz.total_prob()
    ref.close()
    fz.close()
    print("30q depth 40: {} blocks, 1-F {:.3e}, max err {:.3e}, norm {:.15f}".format(blocks, f, e, nf), flush=True)
    assert f < 1e-8
    assert e < 1e-7
    assert abs(nf - 1.0) < 1e-6


# ---- the 512-circuit batch -----------------------------------------------------------------------

def test_batch512_every_circuit_against_plain_complex128():
    from qip_amd import batch, workloads
    ns, depths, seeds, packed = workloads.batch_layered_circuits(512, 2025)
    print("batch512: load", flush=True)
    b = batch.GpuBatch(ns)
    b.load(packed)
    b.run()
    b.synchronize()
    worst_f, worst_e = 0.0, 0.0
    ref = {}
    for i, n in enumerate(ns):
        be = ref.get(n)
        if be is None:
            be = ref[n] = GpuBackend.make_state(n, [], [])
        else:
            be.reinit()
        for gq, G in packed.gates_of(i):
            be.apply_matrix(list(gq), G)
        want = be.get_state()
        got = b.get_state(i).astype(numpy.complex128)     # FP64 sums: a complex64 vdot is good to ~1e-4 at 2^22
        f = _host_infidelity(want, got)
        e = float(numpy.max(numpy.abs(got - want)))
        worst_f, worst_e = max(worst_f, f), max(worst_e, e)
        assert f < 1e-8, (i, n, depths[i], f)
        assert e < 1e-5, (i, n, depths[i], e)
        assert abs(numpy.vdot(got, got).real - 1.0) < 1e-5
        if (i + 1) % 128 == 0:
            print("batch512: {} circuits checked, worst 1-F {:.3e}, worst err {:.3e}".format(i + 1, worst_f, worst_e),
                  flush=True)
    for be in ref.values():
        be.close()
    b.close()
