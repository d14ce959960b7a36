"""GPU parity of the product sweep (qipgpu_apply_product, marked gpu): a descriptor's single-qubit
factors applied in one pass, against the oracle's cdot_loop restatement."""
import numpy
import pytest

from oracle import oracle
from qip_amd import GpuBackend

pytestmark = pytest.mark.gpu
RNG = numpy.random.default_rng(31337)


def rand_u2():
    z = (RNG.standard_normal((2, 2)) + 1j * RNG.standard_normal((2, 2))) / numpy.sqrt(2)
    q, r = numpy.linalg.qr(z)
    return q * (numpy.diagonal(r) / numpy.abs(numpy.diagonal(r)))


@pytest.mark.parametrize("dtype,tol", [(numpy.complex128, 1e-13), (numpy.complex64, 1e-6)])
@pytest.mark.parametrize("n,k", [(12, 12), (13, 2), (14, 5), (16, 9), (16, 16), (20, 20)])
def test_product_matches_oracle(n, k, dtype, tol):
    psi = RNG.standard_normal(2 ** n) + 1j * RNG.standard_normal(2 ** n)
    psi /= numpy.linalg.norm(psi)
    qs = [int(q) for q in RNG.permutation(n)[:k]]
    mats = {q: rand_u2() for q in qs}
    be = GpuBackend.make_state(n, [list(range(n))], [psi], statetype=dtype)
    be.kronselect_dot(mats)              # one descriptor: a product sweep (two for k > 12)
    ref = numpy.zeros(2 ** n, dtype=numpy.complex128)
    if n <= 16:
        oracle.kronselect_dot(mats, psi, n, ref)
    else:                                # the oracle's 2^(n+k) cost is too high: one factor at a time
        ref = psi.copy()
        for q, m in mats.items():
            out = numpy.zeros_like(ref)
            oracle.kronselect_dot({q: m}, ref, n, out)
            ref = out
    assert numpy.max(numpy.abs(be.get_state() - ref)) < tol * 2 ** (-n / 2) * 50


def test_product_small_state_falls_back():
    """n < 12: Kronecker blocks of three through the dense kernel, same result."""
    n = 6
    psi = RNG.standard_normal(2 ** n) + 0j
    mats = {q: rand_u2() for q in range(n)}
    be = GpuBackend.make_state(n, [list(range(n))], [psi])
    be.kronselect_dot(mats)
    ref = numpy.zeros(2 ** n, dtype=numpy.complex128)
    oracle.kronselect_dot(mats, psi, n, ref)
    assert numpy.max(numpy.abs(be.get_state() - ref)) < 1e-13


def rand_o2():
    """A random real orthogonal 2x2 (the product sweep's real-matrix arithmetic)."""
    th = RNG.uniform(0, 2 * numpy.pi)
    return numpy.array([[numpy.cos(th), -numpy.sin(th)], [numpy.sin(th), numpy.cos(th)]])


@pytest.mark.parametrize("dtype,tol", [(numpy.complex128, 1e-13), (numpy.complex64, 1e-6)])
@pytest.mark.parametrize("n,k", [(13, 12), (16, 9)])
def test_product_real_and_complex_factors(n, k, dtype, tol):
    """Real factors (H, X, rotations) take the real-matrix arithmetic, mixed with complex ones in
    one sweep; the factors land on tile bits of every register layout."""
    psi = RNG.standard_normal(2 ** n) + 1j * RNG.standard_normal(2 ** n)
    psi /= numpy.linalg.norm(psi)
    qs = [int(q) for q in RNG.permutation(n)[:k]]
    H = numpy.array([[1, 1], [1, -1]]) / numpy.sqrt(2)
    X = numpy.array([[0.0, 1.0], [1.0, 0.0]])
    choice = [H, X, None, None]
    mats = {}
    for i, q in enumerate(qs):
        m = choice[i % 4]
        mats[q] = (rand_o2() if i % 4 == 2 else rand_u2()) if m is None else m
    be = GpuBackend.make_state(n, [list(range(n))], [psi], statetype=dtype)
    be.kronselect_dot(mats)
    ref = numpy.zeros(2 ** n, dtype=numpy.complex128)
    oracle.kronselect_dot(mats, psi, n, ref)
    assert numpy.max(numpy.abs(be.get_state() - ref)) < tol * 2 ** (-n / 2) * 50


@pytest.mark.parametrize("n,low", [(16, [0, 1, 2, 3]), (20, [1, 2, 3]), (18, [])])
def test_complex64_product_rounds_once(n, low):
    """On a complex64 state the factors are applied in FP64 and the tile crosses LDS between register
    layouts in FP64: every component must come out within one FP32 ulp of the FP64 product of the
    same complex64 input (H, X and complex factors on the tile bits of all three layouts, i.e. on the
    four lowest state bits and eight high ones)."""
    psi = (RNG.standard_normal(2 ** n) + 1j * RNG.standard_normal(2 ** n)).astype(numpy.complex64)
    psi /= numpy.float32(numpy.linalg.norm(psi))
    bits = low + list(range(n - 8, n))
    qs = [n - 1 - b for b in bits]
    H = numpy.array([[1, 1], [1, -1]]) / numpy.sqrt(2)
    X = numpy.array([[0.0, 1.0], [1.0, 0.0]])
    mats = {q: (H, X, rand_u2())[i % 3] for i, q in enumerate(qs)}
    be = GpuBackend.make_state(n, [list(range(n))], [psi], statetype=numpy.complex64)
    be.kronselect_dot(mats)
    got = be.get_state()
    ref = psi.astype(numpy.complex128).reshape((2,) * n)
    for q, m in mats.items():            # qubit q is axis q
        ref = numpy.moveaxis(numpy.tensordot(m, ref, axes=([1], [q])), 0, q)
    ref = ref.reshape(-1)
    for part in (numpy.real, numpy.imag):
        r = part(ref)
        ulp = numpy.spacing(numpy.abs(r).astype(numpy.float32)).astype(numpy.float64)
        assert numpy.all(numpy.abs(part(got).astype(numpy.float64) - r) <= ulp)


@pytest.mark.parametrize("dtype", [numpy.complex128, numpy.complex64])
@pytest.mark.parametrize("n", [16, 22])
def test_x_on_every_qubit_is_the_index_reversal(n, dtype):
    """X on all n qubits maps amplitude i to i xor (2^n - 1): the state reversed, bit for bit. X
    factors cost no arithmetic on most tile bits (the store addresses are flipped) and are real
    factors on the others; both must give exact values, across many workgroups at once."""
    psi = (RNG.standard_normal(2 ** n) + 1j * RNG.standard_normal(2 ** n)).astype(dtype)
    X = numpy.array([[0.0, 1.0], [1.0, 0.0]])
    be = GpuBackend.make_state(n, [list(range(n))], [psi], statetype=dtype)
    be.kronselect_dot({q: X for q in range(n)})
    assert numpy.array_equal(be.get_state(), psi[::-1])
    be.kronselect_dot({q: X for q in range(n)})
    assert numpy.array_equal(be.get_state(), psi)


@pytest.mark.parametrize("dtype,tol", [(numpy.complex128, 1e-13), (numpy.complex64, 1e-6)])
@pytest.mark.parametrize("low,high", [
    ("xxxx", "xxxxxxxx"),     # X on every tile bit: two of them on the wave bits, as arithmetic
    ("ccc", "hhhhhhxx"),      # more factors that need arithmetic than register bits
    ("", "hhxxxxxx"),         # more X than lane bits
    ("", "cccc"),             # the register bits only: no LDS
    ("chh", ""),              # the low bits only: straight to layout C and back
    ("hc", "cccc"),           # low bits and register bits: A -> C -> A
    ("h", "hhhhx"),           # low bits and register bits, X on a lane bit: A -> C -> A
    ("", "hhhhhhhh"),         # register bits and u bits 4..7: A -> B, stored from B
    ("x", "hhhhhhxx"),        # A -> B with X flipped on a low bit and on u bits 4, 5 at B's store
    ("xc", "hhhhhx"),         # every layout: A -> C -> B, X on a low bit and on u bit 4
])
def test_product_tile_bit_placement(low, high, dtype, tol):
    """Factors on the four lowest state bits and on higher bits, of each kind (H real, X, complex),
    in numbers that exercise every placement of targets on tile bits and every LDS path."""
    n = 16
    psi = RNG.standard_normal(2 ** n) + 1j * RNG.standard_normal(2 ** n)
    psi /= numpy.linalg.norm(psi)
    H = numpy.array([[1, 1], [1, -1]]) / numpy.sqrt(2)
    X = numpy.array([[0.0, 1.0], [1.0, 0.0]])
    kinds = {"h": lambda: H, "x": lambda: X, "c": rand_u2}
    low_bits = [int(b) for b in RNG.permutation(4)[:len(low)]]
    high_bits = [int(b) for b in 4 + RNG.permutation(n - 4)[:len(high)]]
    mats = {n - 1 - b: kinds[c]() for b, c in zip(low_bits + high_bits, low + high)}
    be = GpuBackend.make_state(n, [list(range(n))], [psi], statetype=dtype)
    be.kronselect_dot(mats)
    ref = numpy.zeros(2 ** n, dtype=numpy.complex128)
    oracle.kronselect_dot(mats, psi, n, ref)
    assert numpy.max(numpy.abs(be.get_state() - ref)) < tol * 2 ** (-n / 2) * 50
