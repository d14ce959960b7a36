"""GPU tests at the largest sizes (marked gpu): states of 2^32 and 2^33 complex64 amplitudes (32 and
64 GiB), beyond any 32-bit index. Size-independent properties only: U then U^dagger restores the
state (checked on windows at the lowest and the highest addresses), the FP64 norm is 1, marginals of
a uniform superposition are uniform, func_apply moves exactly the amplitudes it should, and the
fused and batched forms agree with the plain engine on a 30-qubit state placed after a small one in
the batch arena."""
import numpy
import pytest

from qip_amd import GpuBackend
from qip_amd.batch import GpuBatch, PackedCircuits
from qip_amd.workloads import layered_circuit_gates

pytestmark = pytest.mark.gpu
RNG = numpy.random.default_rng(3232)
H = numpy.array([[1, 1], [1, -1]], dtype=numpy.complex128) / numpy.sqrt(2)


def rand_u(d):
    z = (RNG.standard_normal((d, d)) + 1j * RNG.standard_normal((d, d))) / numpy.sqrt(2)
    q, r = numpy.linalg.qr(z)
    return q * (numpy.diagonal(r) / numpy.abs(numpy.diagonal(r)))


def windows(be, n):
    top = 2 ** n
    return numpy.concatenate([be.get_relative_range(0, 1024), be.get_relative_range(top - 1024, top),
                              be.get_relative_range(top // 2 - 512, top // 2 + 512)])


@pytest.mark.parametrize("n", [32, 33])
def test_plain_round_trip_beyond_32_bit_indices(n):
    print("plain complex64 n =", n, flush=True)
    be = GpuBackend.make_state(n, [], [], statetype=numpy.complex64)
    be.kronselect_dot({q: H for q in range(n)})                    # product sweep
    assert be.total_prob() == pytest.approx(1.0, abs=1e-6)
    # a window at the top of the state: the first and last qubits rotated, to make it non-uniform
    pre = [((0,), rand_u(2)), ((n - 1,), rand_u(2)), ((0, n - 1), rand_u(4)), ((n // 2, 1), rand_u(4))]
    for gq, U in pre:
        be.kronselect_dot({gq if len(gq) > 1 else gq[0]: U})
    start = windows(be, n)
    gates = [((0, n - 1), rand_u(4)), ((n - 1,), rand_u(2)), ((3, n - 2), rand_u(4)), ((n // 2,), rand_u(2)),
             ((n - 1, 0, 5), rand_u(8))]
    for gq, U in gates:
        be.kronselect_dot({gq if len(gq) > 1 else gq[0]: U})
    assert be.total_prob() == pytest.approx(1.0, abs=1e-5)
    for gq, U in reversed(gates):
        be.kronselect_dot({gq if len(gq) > 1 else gq[0]: U.conj().T})
    end = windows(be, n)
    scale = numpy.max(numpy.abs(start))
    assert numpy.max(numpy.abs(end - start)) < 1e-5 * scale
    probs = be.measure_probabilities([2, n - 3])                      # untouched qubits: uniform
    assert numpy.allclose(probs, 0.25, atol=1e-5)
    be.close()


def test_func_apply_and_measure_beyond_32_bit_indices():
    """|x>|y> with f(x) = x on reg1 = [0], reg2 = [n-1] (a CNOT from the top qubit to the bottom
    one, through the classical-function path) moves exactly the amplitudes with qubit 0 set."""
    n = 32
    print("func_apply n =", n, flush=True)
    be = GpuBackend.make_state(n, [], [], statetype=numpy.complex64)
    be.kronselect_dot({0: H})               # |+> on qubit 0 (state bit 31): amplitudes 0 and 2^31
    be.func_apply([0], [n - 1], lambda x: x)
    top = 2 ** (n - 1)
    lo = be.get_relative_range(0, 4)
    hi = be.get_relative_range(top, top + 4)
    r = numpy.float32(1 / numpy.sqrt(2))
    assert lo[0] == pytest.approx(r) and numpy.all(lo[1:] == 0)
    assert hi[0] == 0 and hi[1] == pytest.approx(r) and numpy.all(hi[2:] == 0)
    m, p = be.measure([0, n - 1], measured=3, measured_prob=0.5)
    assert m == 3
    assert be.total_prob() == pytest.approx(1.0, abs=1e-6)
    assert be.get_relative_range(top + 1, top + 2)[0] == pytest.approx(1.0, abs=1e-6)
    be.close()


def test_fused_round_trip_beyond_32_bit_indices():
    n = 32
    print("fused n =", n, flush=True)
    be = GpuBackend.make_state(n, [], [], statetype=numpy.complex64)
    be.kronselect_dot({q: H for q in range(n)})
    be.kronselect_dot({(0, n - 1): rand_u(4)})
    be.kronselect_dot({(1, n - 2): rand_u(4)})
    start = windows(be, n)
    blocks = [([0, 5, 9, n - 1, n - 3, 16], rand_u(64)), ([n - 1, n - 2, n - 4, 2, 1, 0], rand_u(64)),
              ([7, 11, 13], rand_u(8))]
    for qs, U in blocks:
        be.apply_fused(qs, U)
    for qs, U in reversed(blocks):
        be.apply_fused(qs, U.conj().T)
    be.fused_finish()
    end = windows(be, n)
    scale = numpy.max(numpy.abs(start))
    assert numpy.max(numpy.abs(end - start)) < 2e-4 * scale      # six split-BF16 blocks
    assert be.total_prob() == pytest.approx(1.0, abs=1e-6)
    be.close()


def test_batch_large_state_after_small_one():
    """A 30-qubit state placed after a 10-qubit one in the batch arena, against the plain engine."""
    ns = [10, 30]
    circuits = [layered_circuit_gates(10, 4, 1), layered_circuit_gates(30, 3, 2)]
    print("batch n =", ns, flush=True)
    b = GpuBatch(ns)
    b.load(PackedCircuits.from_gates(ns, circuits))
    b.run()
    got = b.get_state(1)
    b.close()
    ref = GpuBackend.make_state(30, [], [], statetype=numpy.complex128)
    for gq, G in circuits[1]:
        ref.apply_matrix(list(gq), G)
    want = ref.get_state()
    ref.close()
    ov = numpy.vdot(want, got.astype(numpy.complex128))
    inf = 1.0 - abs(ov) ** 2 / (numpy.vdot(want, want).real * numpy.vdot(got.astype(numpy.complex128),
                                                                          got.astype(numpy.complex128)).real)
    assert inf < 1e-8
    assert numpy.max(numpy.abs(got - want)) < 1e-6
