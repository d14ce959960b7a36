"""Iterative Grover search on the GPU (marked gpu): func_apply oracles, both diffusion forms,
seeded measurement. Bars: measured outcome identical to the oracle's for the same uniform draw;
amplitudes within 1e-10 (complex128) / 1e-5 (complex64) of the closed form."""
import random

import numpy
import pytest

from oracle.oracle_backend import OracleBackend
from qip_amd import GpuBackend
from qip_amd.fused import FusedBackend
from qip_amd.workloads import apply_ops, grover_closed_form, grover_initial, grover_iteration_ops, grover_iterations

pytestmark = pytest.mark.gpu


def run_grover(make, n_search, x_star, diffusion, iterations=None, statetype=numpy.complex128):
    groups, feeds = grover_initial(n_search)
    be = make(n_search + 1, groups, feeds, statetype=statetype)
    t = grover_iterations(n_search) if iterations is None else iterations
    ops = grover_iteration_ops(n_search, x_star, diffusion)
    for _ in range(t):
        apply_ops(be, ops)
    return be, t


@pytest.mark.parametrize("diffusion", ["f", "mcz"])
def test_grover_small_matches_oracle(diffusion):
    n_search, x_star, seed = 10, 613, 20240611
    gpu, t = run_grover(GpuBackend.make_state, n_search, x_star, diffusion)
    ora, _ = run_grover(OracleBackend.make_state, n_search, x_star, diffusion)
    expect = grover_closed_form(n_search, x_star, t)
    assert numpy.max(numpy.abs(gpu.get_state() - expect)) < 1e-10
    assert numpy.max(numpy.abs(gpu.get_state() - ora.get_state())) < 1e-12
    search = list(range(n_search))
    random.seed(seed)
    m_g = gpu.measure(search)
    random.seed(seed)
    m_o = ora.measure(search)
    assert m_g[0] == m_o[0] == x_star
    assert m_g[1] == pytest.approx(m_o[1], rel=1e-12)


def expected_outcome(n_search, x_star, t, u):
    """Outcome of the sequential walk over the closed-form marginals (x natural order), plus the
    distance of u * total to the nearest boundary it crosses."""
    N = 2 ** n_search
    p = numpy.abs(grover_closed_form(n_search, x_star, t).reshape(N, 2)) ** 2
    marg = p.sum(axis=1)
    cdf = numpy.cumsum(marg)
    r = u * cdf[-1]
    m = int(numpy.searchsorted(cdf, r))
    margin = min(abs(cdf[m] - r), abs(r - (cdf[m - 1] if m else 0.0)))
    return m, marg[m], margin


def test_grover_24_qubits_c128():
    """The full configuration: 23 search qubits + ancilla, floor(pi/4 sqrt(2^23)) = 2274
    iterations, H/X/multi-controlled-Z diffusion, then a seeded measurement of the register."""
    n_search = 23
    x_star = int(numpy.random.default_rng(4).integers(0, 2 ** n_search))
    be, t = run_grover(GpuBackend.make_state, n_search, x_star, "mcz")
    assert t == 2274
    expect = grover_closed_form(n_search, x_star, t)
    got = be.get_state()
    assert numpy.max(numpy.abs(got - expect)) < 1e-10
    random.seed(4)
    u = random.random()
    want_m, want_p, margin = expected_outcome(n_search, x_star, t, u)
    assert margin > 1e-9                       # the draw is far from any CDF boundary
    random.seed(4)
    m, p = be.measure(list(range(n_search)))
    assert m == want_m == x_star
    assert p == pytest.approx(want_p, rel=1e-9)


def test_grover_24_qubits_c64_plain():
    """complex64 storage with FP64 sweep arithmetic: within 1e-5 of the closed form after 2274
    iterations (about 36k rounding sweeps), and the measurement finds x*."""
    n_search = 23
    x_star = int(numpy.random.default_rng(5).integers(0, 2 ** n_search))
    be, t = run_grover(GpuBackend.make_state, n_search, x_star, "f", statetype=numpy.complex64)
    got = be.get_state()
    assert got.dtype == numpy.complex64
    assert numpy.max(numpy.abs(got - grover_closed_form(n_search, x_star, t))) < 1e-5
    random.seed(5)
    assert be.measure(list(range(n_search)))[0] == x_star


def test_grover_24_qubits_c64_fused():
    """The precision form on a 2274-iteration loop. Its recorded H layers are blocks of single-qubit
    gates only, so each run of them is one product of per-qubit 2x2s on the FP64 product sweep (three
    passes per layer), not split-BF16 MFMA: the complex64 amplitudes stay within the 1e-5 bar, and the
    measurement finds x*."""
    n_search = 23
    x_star = int(numpy.random.default_rng(6).integers(0, 2 ** n_search))
    be, t = run_grover(FusedBackend.make_state, n_search, x_star, "f")
    got = be.get_state()
    assert got.dtype == numpy.complex64
    assert numpy.max(numpy.abs(got - grover_closed_form(n_search, x_star, t))) < 1e-5
    random.seed(6)
    assert be.measure(list(range(n_search)))[0] == x_star
