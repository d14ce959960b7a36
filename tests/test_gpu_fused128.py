"""The complex128 fused form (Fused128Backend, marked gpu): the precision form's recording and
light-cone fusion, each block applied in FP64 by the dense sweeps. It computes the same operator as
the plain gate-by-gate complex128 path, so the bars are FP64 rounding: amplitudes within 1e-12 of
the plain path or of the oracle, 1 - F within 1e-13."""
import numpy
import pytest

from oracle.oracle_backend import OracleBackend
from qip_amd import GpuBackend, QFFT, Qubit, infidelity, run
from qip_amd.fused import Fused128Backend
from qip_amd.workloads import (apply_ops, grover_closed_form, grover_initial, grover_iteration_ops,
                               grover_iterations, layered_random_circuit)

pytestmark = pytest.mark.gpu
RNG = numpy.random.default_rng(1281)


@pytest.mark.parametrize("n,depth", [(14, 10), (20, 12), (24, 6)])
def test_layered_circuit_matches_plain(n, depth):
    """Blocks of up to 6 qubits (FP64 MFMA tiles at 5 and 6) against one sweep per gate."""
    ops = layered_random_circuit(n, depth, seed=n)
    pl = GpuBackend.make_state(n, [], [])
    fz = Fused128Backend.make_state(n, [], [])
    for d in ops:
        pl.kronselect_dot(d)
        fz.kronselect_dot(d)
    got = fz.get_state()
    assert got.dtype == numpy.complex128
    assert 0 < fz.blocks_applied < len(ops) // 4
    assert numpy.max(numpy.abs(got - pl.get_state())) < 1e-12
    assert infidelity(pl, fz) < 1e-13


def test_streaming_equals_deferred():
    """Blocks launched while recording and blocks launched at the read are the same, bit for bit."""
    n = 16
    ops = layered_random_circuit(n, 12, seed=5)
    fz = Fused128Backend.make_state(n, [], [])
    deferred = Fused128Backend.make_state(n, [], [])
    deferred.streaming = False
    for d in ops:
        fz.kronselect_dot(d)
        deferred.kronselect_dot(d)
    assert deferred.blocks_applied == 0 and fz.blocks_applied > 0
    assert numpy.array_equal(fz.get_state(), deferred.get_state())


def test_small_state_against_oracle():
    """Below 12 qubits every block takes the plain dense sweep; the result against the oracle's
    gate-by-gate cdot_loop restatement."""
    n = 9
    ops = layered_random_circuit(n, 14, seed=9)
    fz = Fused128Backend.make_state(n, [], [])
    ora = OracleBackend.make_state(n, [], [])
    for d in ops:
        fz.kronselect_dot(d)
        ora.kronselect_dot(d)
    assert numpy.max(numpy.abs(fz.get_state() - ora.get_state())) < 1e-12


def test_qft_graph():
    """The 16-qubit QFT graph through run(): the state against ifft(norm='ortho'), as the plain
    path gives it; complex128 whatever statetype is asked for."""
    n = 16
    psi = RNG.standard_normal(2 ** n) + 1j * RNG.standard_normal(2 ** n)
    psi /= numpy.linalg.norm(psi)
    q = Qubit(n=n)
    out, _ = run(QFFT(q), feed={q: psi}, backend_constructor=Fused128Backend.make_state,
                 statetype=numpy.complex64)
    assert out.dtype == numpy.complex128
    assert numpy.max(numpy.abs(out - numpy.fft.ifft(psi, norm="ortho"))) < 1e-12


def test_grover_against_closed_form():
    """16 search qubits + ancilla, all 201 iterations with F-based diffusion: H layers as product
    runs, func_apply oracles between them, amplitudes within 1e-10 of the closed form."""
    ns = 16
    x_star = int(RNG.integers(0, 2 ** ns))
    groups, feeds = grover_initial(ns)
    be = Fused128Backend.make_state(ns + 1, groups, feeds)
    t = grover_iterations(ns)
    ops = grover_iteration_ops(ns, x_star, "f")
    for _ in range(t):
        apply_ops(be, ops)
    assert numpy.max(numpy.abs(be.get_state() - grover_closed_form(ns, x_star, t))) < 1e-10


def test_measurement_after_recording():
    """Reads launch what is recorded first: marginals after a recorded circuit equal the plain path's."""
    n = 14
    ops = layered_random_circuit(n, 8, seed=2)
    pl = GpuBackend.make_state(n, [], [])
    fz = Fused128Backend.make_state(n, [], [])
    for d in ops:
        pl.kronselect_dot(d)
        fz.kronselect_dot(d)
    idx = [1, 4, 9, 13]
    assert numpy.max(numpy.abs(fz.measure_probabilities(idx) - pl.measure_probabilities(idx))) < 1e-13
