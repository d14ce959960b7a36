"""GPU parity of the plain engine against the CPU oracle and closed forms (marked gpu).

Bars: complex128 amplitudes within 1e-12 of the oracle (the oracle and the engine order their
floating-point sums differently, so agreement is to rounding, not to the bit); complex64 within
1e-5; func_apply permutations, measured outcomes and product-state initialisation bit-exact.
"""
import random

import numpy
import pytest

from oracle import oracle
from oracle.oracle_backend import OracleBackend
from qip_amd import CMat, GpuBackend, SwapMat, infidelity, max_abs_diff
from qip_amd.planner import plan

pytestmark = pytest.mark.gpu
RNG = numpy.random.default_rng(4242)
TOL = {numpy.complex128: 1e-12, numpy.complex64: 2e-6}


def rand_state(n):
    v = RNG.standard_normal(2 ** n) + 1j * RNG.standard_normal(2 ** n)
    return v / numpy.linalg.norm(v)


def rand_u(d):
    z = (RNG.standard_normal((d, d)) + 1j * RNG.standard_normal((d, d))) / numpy.sqrt(2)
    q, r = numpy.linalg.qr(z)
    return q * (numpy.diagonal(r) / numpy.abs(numpy.diagonal(r)))


def gpu_from(psi, n, dtype=numpy.complex128):
    return GpuBackend.make_state(n, [list(range(n))], [psi], statetype=dtype)


def oracle_apply(mats, psi, n):
    out = numpy.zeros(2 ** n, dtype=numpy.complex128)
    oracle.kronselect_dot(mats, psi, n, out)
    return out


X = numpy.array([[0, 1], [1, 0]], dtype=complex)


def random_descriptor(n):
    qs = [int(q) for q in RNG.permutation(n)]
    kind = int(RNG.integers(0, 7))
    if kind == 0:
        return {qs[0]: rand_u(2)}
    if kind == 1:
        return {(qs[0], qs[1]): rand_u(4)}
    if kind == 2:
        return {(qs[0], qs[1]): CMat(rand_u(2)), qs[2]: rand_u(2)}
    if kind == 3:
        return {(qs[0], qs[1], qs[2]): CMat(CMat(X))}
    if kind == 4:
        w = int(RNG.integers(1, min(3, n // 2) + 1))
        return {tuple(qs[:2 * w]): SwapMat(w)} if RNG.integers(0, 2) else {tuple(qs[:2 * w + 1]): CMat(SwapMat(w))}
    if kind == 5:
        return {(qs[0], qs[1]): CMat(X), (qs[0], qs[2]): CMat(X)}
    return {(qs[0], qs[1], qs[2]): rand_u(8)}


@pytest.mark.parametrize("dtype", [numpy.complex128, numpy.complex64])
@pytest.mark.parametrize("trial", range(10))
def test_kronselect_random(trial, dtype):
    n = int(RNG.integers(5, 15))
    psi = rand_state(n)
    be = gpu_from(psi, n, dtype)
    ref = psi.copy()
    for _ in range(6):
        mats = random_descriptor(n)
        be.kronselect_dot(mats)
        ref = oracle_apply(mats, ref, n)
    got = be.get_state()
    assert got.dtype == dtype
    assert numpy.max(numpy.abs(got - ref)) < TOL[dtype] * 10
    be.close()


@pytest.mark.parametrize("k", [3, 4, 5, 6])
def test_dense_wide(k):
    n = 12
    psi = rand_state(n)
    qs = [int(q) for q in RNG.permutation(n)[:k]]
    U = rand_u(2 ** k)
    be = gpu_from(psi, n)
    be.kronselect_dot({tuple(qs): U})
    ref = oracle_apply({tuple(qs): U}, psi, n)
    assert numpy.max(numpy.abs(be.get_state() - ref)) < 1e-12


def test_structure_detection():
    """Diagonal phases, controlled phases and identity blocks take the diagonal / control paths;
    results must not change."""
    n = 9
    psi = rand_state(n)
    be = gpu_from(psi, n)
    ref = psi.copy()
    cases = [
        {(2, 5): numpy.diag([1, 1, 1, numpy.exp(0.3j)])},                    # controlled phase
        {(0, 1, 2): numpy.diag(numpy.exp(1j * RNG.uniform(0, 6, 8)))},       # general diagonal
        {4: numpy.eye(2)},                                                   # identity
        {(3, 7): numpy.kron(numpy.eye(2), rand_u(2))},                       # identity factor
        {(8, 6): CMat(numpy.diag([1.0, -1.0]))},                             # CZ
    ]
    for mats in cases:
        be.kronselect_dot(mats)
        ref = oracle_apply(mats, ref, n)
    assert numpy.max(numpy.abs(be.get_state() - ref)) < 1e-13


def test_init_product():
    """Which amplitudes are non-zero is exact; values agree to 1 ulp (the product is taken in
    the same order, but numpy may fuse the complex multiply where the kernel does not)."""
    n = 10
    groups = [[0, 3], [5], [9, 1, 2]]
    feeds = [rand_state(2), rand_state(1) * 3, rand_state(3)]
    be = GpuBackend.make_state(n, groups, feeds)
    got, ref = be.get_state(), oracle.make_state_vector(n, groups, feeds)
    assert numpy.array_equal(got != 0, ref != 0)
    assert numpy.max(numpy.abs(got - ref)) <= 2.3e-16 * numpy.max(numpy.abs(ref))
    be0 = GpuBackend.make_state(4, [], [])
    assert numpy.array_equal(be0.get_state(), numpy.eye(16)[0])


@pytest.mark.parametrize("dtype", [numpy.complex128, numpy.complex64])
@pytest.mark.parametrize("groups", [
    [[q] for q in range(16)],                  # one feed per qubit: qubits 0..3 (bits 15..12) lead
    [[15, 14], [0, 1, 2], [7], [3, 4]],        # a group on low bits first: no common prefix
    [[0, 2], [5, 9, 11], [13]],                # qubits 1, 3, 4, ... unfed
    [[0, 1, 2, 3], [8, 4], [15], [6, 5, 7]],
])
def test_init_product_chunks(groups, dtype):
    """From 12 qubits on, the product of the leading groups on bits above 11 is formed once per
    4096-amplitude chunk and continued per amplitude in the same group order: the same zeros, and
    each value within (groups) ulps of the oracle's product, which takes the same order but may fuse
    a multiply (complex64: within one rounding of the oracle's value rounded)."""
    n = 16
    feeds = [rand_state(len(g)) for g in groups]
    be = GpuBackend.make_state(n, groups, feeds, statetype=dtype)
    got = be.get_state().astype(numpy.complex128)
    ref = oracle.make_state_vector(n, groups, feeds)
    nz = ref != 0
    assert numpy.array_equal(got != 0, nz)
    rel = numpy.abs(got[nz] - ref[nz]) / numpy.abs(ref[nz])
    assert numpy.max(rel) <= (len(groups) * 1.2e-16 if dtype == numpy.complex128 else 1.2e-7)


def test_ranges():
    n = 8
    psi = rand_state(n)
    be = gpu_from(psi, n)
    assert numpy.array_equal(be.get_relative_range(5, 40), psi[5:40])
    data = rand_state(4)
    be.overwrite_relative_range(16, 32, data)
    psi[16:32] = data
    be.addto_relative_range(100, 116, data)
    psi[100:116] += data
    assert numpy.array_equal(be.get_state(), psi)
    assert be.get_state_size() == 256


@pytest.mark.parametrize("trial", range(6))
def test_func_apply_bit_exact(trial):
    n = int(RNG.integers(4, 14))
    perm = [int(q) for q in RNG.permutation(n)]
    n1 = int(RNG.integers(1, n))
    n2 = int(RNG.integers(1, n - n1 + 1))
    reg1, reg2 = perm[:n1], perm[n1:n1 + n2]
    table = RNG.integers(-1000, 1000, size=2 ** n1).astype(numpy.int32)
    psi = rand_state(n)
    be = gpu_from(psi, n)
    be.func_apply(numpy.array(reg1, dtype=numpy.int32), numpy.array(reg2, dtype=numpy.int32), table)
    ref = numpy.zeros_like(psi)
    oracle.func_apply(reg1, reg2, None, psi, n, ref, faithful=False, table=table)
    assert numpy.array_equal(be.get_state(), ref)


@pytest.mark.parametrize("dtype", [numpy.complex128, numpy.complex64])
@pytest.mark.parametrize("n,reg1,reg2,kind", [
    (12, list(range(4)), list(range(4, 12)), "dense"),           # reg2 on the 8 lowest state bits
    (14, list(range(6, 14)), list(range(6)), "dense"),           # reg2 high, reg1 low
    (14, [3, 9, 0, 12], [7, 1, 13], "dense"),                    # scattered, 7 rest qubits
    (16, list(range(15)), [15], "sparse"),                       # Grover-style: one x moves
    (16, list(range(15)), [15], "block"),                        # f != 0 on the first 1/16 of x only
    (14, list(range(6, 14)), list(range(6)), "block"),
    (16, [2, 5, 11, 14, 8], [0, 9, 4, 15, 1, 12, 6, 3], "dense"),
    (13, [0, 1, 2], [5, 6, 7, 8, 9, 10, 11, 12, 3], "dense"),    # reg2 of 9 qubits: per-pair kernel
])
def test_func_apply_tiles(n, reg1, reg2, kind, dtype):
    """func_apply from 12 qubits on with reg2 of up to 8 qubits (tiles of 4096 amplitudes permuted in
    LDS, untouched where f is 0 on the whole tile; the per-pair kernel where fewer than one x in 64
    moves): bit-exact against the oracle for registers on low, high and scattered state bits, with
    other qubits beside them, for a table that moves one x only and one that is 0 on most tiles."""
    psi = rand_state(n)
    be = gpu_from(psi, n, dtype)
    v = be.get_state().astype(numpy.complex128)
    if kind == "dense":
        table = RNG.integers(0, 2 ** 31 - 1, size=2 ** len(reg1)).astype(numpy.int32)
    elif kind == "block":
        x = numpy.arange(2 ** len(reg1))
        table = numpy.where(x < 2 ** len(reg1) // 16, x % 7 + 1, 0).astype(numpy.int32)
    else:
        table = numpy.zeros(2 ** len(reg1), dtype=numpy.int32)
        table[int(RNG.integers(0, 2 ** len(reg1)))] = 1
    be.func_apply(numpy.array(reg1, dtype=numpy.int32), numpy.array(reg2, dtype=numpy.int32), table)
    ref = numpy.zeros_like(v)
    oracle.func_apply(reg1, reg2, None, v, n, ref, faithful=False, table=table)
    assert numpy.array_equal(be.get_state().astype(numpy.complex128), ref)


@pytest.mark.parametrize("dtype", [numpy.complex128, numpy.complex64])
@pytest.mark.parametrize("n,reg1,reg2", [
    (9, [1, 4], [6]),                              # per-pair kernel (n < 12); rest 0, 2, 3, 5, 7, 8
    (14, [3, 9, 0, 12], [7, 1, 13]),               # tile kernel; 7 rest qubits, 4 of them dropped
    (16, list(range(12)), [15]),                   # one reg2 qubit, rest 12, 13, 14: two dropped
    (14, list(range(6, 14)), list(range(6))),      # no rest qubits: the flag changes nothing
    (13, [0, 1, 2], [5, 6, 7, 8, 9, 10, 11, 12, 3]),  # reg2 of 9 qubits, rest 4 only
])
def test_func_apply_reference_bound(n, reg1, reg2, dtype):
    """reference_func_bound = True: the reference's rest-qubit loop bound (func_apply.pyx:107-111),
    amplitudes with a dropped rest qubit at 1 end 0; bit-exact against the oracle's faithful mode
    on both kernels and both precisions."""
    psi = rand_state(n)
    be = gpu_from(psi, n, dtype)
    be.reference_func_bound = True
    v = be.get_state().astype(numpy.complex128)
    table = RNG.integers(0, 2 ** 31 - 1, size=2 ** len(reg1)).astype(numpy.int32)
    be.func_apply(numpy.array(reg1, dtype=numpy.int32), numpy.array(reg2, dtype=numpy.int32), table)
    ref = numpy.zeros_like(v)
    oracle.func_apply(reg1, reg2, None, v, n, ref, faithful=True, table=table)
    assert numpy.array_equal(be.get_state().astype(numpy.complex128), ref)
    rest = n - len(reg1) - len(reg2)
    if rest > len(reg2):
        assert numpy.count_nonzero(ref) < 2 ** n // 2


def test_func_apply_callable_and_cache():
    n = 6
    psi = rand_state(n)
    be = gpu_from(psi, n)
    f = lambda x: (3 * x + 1) % 8  # noqa: E731
    be.func_apply([0, 1, 2], [3, 4, 5], f)
    be.func_apply([0, 1, 2], [3, 4, 5], f)
    ref = numpy.zeros_like(psi)
    oracle.func_apply([0, 1, 2], [3, 4, 5], f, psi, n, ref)
    ref2 = numpy.zeros_like(psi)
    oracle.func_apply([0, 1, 2], [3, 4, 5], f, ref, n, ref2)
    assert numpy.array_equal(be.get_state(), ref2)


def test_func_apply_follows_a_changing_closure():
    """A closure whose values change between two func_apply calls: each call uses the values the
    function has at that call, as the reference's per-call tabulation does (func_apply.pyx:66-69).
    With cache_func_tables the first table is kept, which is the documented opt-in."""
    n = 7
    psi = rand_state(n)
    shift = [1]
    f = lambda x: (x + shift[0]) % 8  # noqa: E731
    want = psi
    be = gpu_from(psi, n)
    for s in (1, 5):
        shift[0] = s
        be.func_apply([0, 1, 2, 3], [4, 5, 6], f)
        out = numpy.zeros_like(want)
        oracle.func_apply([0, 1, 2, 3], [4, 5, 6], f, want, n, out)
        want = out
    assert numpy.array_equal(be.get_state(), want)
    be.close()

    class Cached(GpuBackend):
        cache_func_tables = True
    c = Cached.make_state(n, [list(range(n))], [psi])
    shift[0] = 1
    c.func_apply([0, 1, 2, 3], [4, 5, 6], f)
    shift[0] = 5
    c.func_apply([0, 1, 2, 3], [4, 5, 6], f)           # the table of shift 1 again
    shift[0] = 1
    once = numpy.zeros_like(psi)
    oracle.func_apply([0, 1, 2, 3], [4, 5, 6], f, psi, n, once)
    twice = numpy.zeros_like(psi)
    oracle.func_apply([0, 1, 2, 3], [4, 5, 6], f, once, n, twice)
    assert numpy.array_equal(c.get_state(), twice)
    c.close()


@pytest.mark.parametrize("trial", range(6))
def test_probabilities_and_marginals(trial):
    n = int(RNG.integers(3, 17))
    k = int(RNG.integers(1, n + 1))
    idx = [int(q) for q in RNG.permutation(n)[:k]]
    psi = rand_state(n)
    be = gpu_from(psi, n)
    got = be.measure_probabilities(idx)
    ref = oracle.measure_probabilities(idx, n, psi)
    assert numpy.max(numpy.abs(got - ref)) < 1e-14
    assert be.total_prob() == pytest.approx(1.0, abs=1e-13)
    ms, ps = be.measure_probabilities(idx, top_k=3)
    rms, rps = oracle.measure_top_probabilities(idx, n, 3, psi)
    assert ms == rms and numpy.allclose(ps, rps, rtol=1e-12, atol=0)


@pytest.mark.parametrize("dtype", [numpy.complex128, numpy.complex64])
@pytest.mark.parametrize("n,idx", [
    (12, [11, 10, 9]),                            # the lowest state bits
    (13, [0, 1]),                                 # the highest
    (18, [17, 3, 9, 0, 12]),                      # both sides of the 4096-amplitude chunk
    (18, list(range(18))),                        # every qubit: 2^18 outcomes
    (20, [19]), (20, [0]),
    (20, [5, 19, 2, 14, 8, 11, 17, 0, 13, 6]),
])
def test_marginals_one_pass(n, idx, dtype):
    """All-outcome marginals (one pass over the state from 12 qubits on) against the oracle in the
    LSB-first order, and against the definition of the measurement order (output bit j = the j-th
    lowest measured state bit) by bincount; the same bits on a second call."""
    psi = rand_state(n)
    be = gpu_from(psi, n, dtype)
    v = be.get_state().astype(numpy.complex128)
    got = be.measure_probabilities(idx)
    assert numpy.max(numpy.abs(got - oracle.measure_probabilities(idx, n, v))) < 1e-13
    pos = sorted(n - 1 - q for q in idx)
    i = numpy.arange(2 ** n, dtype=numpy.int64)
    m = numpy.zeros(2 ** n, dtype=numpy.int64)
    for j, b in enumerate(pos):
        m |= ((i >> b) & 1) << j
    want = numpy.bincount(m, weights=numpy.abs(v) ** 2, minlength=2 ** len(idx))
    marg = be.marginals(idx)
    assert numpy.max(numpy.abs(marg - want)) < 1e-13
    assert numpy.array_equal(marg, be.marginals(idx))
    assert numpy.array_equal(got, be.measure_probabilities(idx))


@pytest.mark.parametrize("trial", range(8))
def test_measure_bit_exact_outcomes(trial):
    n = int(RNG.integers(3, 15))
    k = int(RNG.integers(1, n + 1))
    idx = [int(q) for q in RNG.permutation(n)[:k]]
    psi = rand_state(n)
    seed = int(RNG.integers(0, 2 ** 31))
    be = gpu_from(psi, n)
    ob = OracleBackend.make_state(n, [list(range(n))], [psi])
    random.seed(seed)
    m_g, p_g = be.measure(idx)
    random.seed(seed)
    m_o, p_o = ob.measure(idx)
    assert m_g == m_o
    assert p_g == pytest.approx(p_o, rel=1e-12)
    assert numpy.max(numpy.abs(be.get_state() - ob.get_state())) < 1e-12
    random.seed(seed + 1)
    s_g = be.soft_measure(idx)
    random.seed(seed + 1)
    s_o = ob.soft_measure(idx)
    assert s_g[0] == s_o[0] and s_g[1] == pytest.approx(s_o[1], rel=1e-12)


def test_measure_forced_and_errors():
    n = 6
    psi = rand_state(n)
    be = gpu_from(psi, n)
    with pytest.raises(ValueError, match="Measured value must be less"):
        be.measure([0, 1], measured=4)
    with pytest.raises(ValueError, match="measured_prob must be"):
        be.measure([0], measured=0, measured_prob=0.0)
    m, p = be.measure([2], measured=1, measured_prob=0.5)
    assert (m, p) == (1, 0.5)
    z = GpuBackend.make_state(2, [], [])
    with pytest.raises(ZeroDivisionError):
        z.measure([0], measured=1)


def test_reduce_measure_tiled():
    n = 7
    psi = rand_state(n)
    idx = [5, 2]
    be = gpu_from(psi, n)
    ob = OracleBackend.make_state(n, [list(range(n))], [psi])
    random.seed(11)
    r_g = be.reduce_measure(idx)
    random.seed(11)
    r_o = ob.reduce_measure(idx)
    assert r_g[0] == r_o[0]
    assert numpy.max(numpy.abs(be.get_state() - ob.get_state())) < 1e-12


def test_bad_arguments_raise():
    be = gpu_from(rand_state(4), 4)
    with pytest.raises(ValueError):
        be.kronselect_dot({7: numpy.eye(2)})
    with pytest.raises(ValueError, match="Input vector size plus offset"):
        be.kronselect_dot({0: numpy.eye(2)}, input_offset=3)
    with pytest.raises(Exception, match="Shape of square submatrix"):
        be.kronselect_dot({(0, 1): numpy.eye(2)})
    with pytest.raises(ValueError):
        be.func_apply([0, 0], [1], lambda x: x)


@pytest.mark.parametrize("n", [24, 26])
def test_large_state_properties(n):
    """Size-independent properties at full-scale widths: U then U^dagger restores the state,
    norms are preserved, and complex64 tracks complex128 to single precision."""
    rng = numpy.random.default_rng(n)
    be = GpuBackend.make_state(n, [], [])
    be.kronselect_dot({q: H_ for q, H_ in zip(range(n), [numpy.array([[1, 1], [1, -1]]) / numpy.sqrt(2)] * n)})
    be64 = GpuBackend.make_state(n, [], [], statetype=numpy.complex64)
    be64.kronselect_dot({q: numpy.array([[1, 1], [1, -1]]) / numpy.sqrt(2) for q in range(n)})
    start = be.get_relative_range(0, 1024).copy()
    gates = []
    for _ in range(12):
        a, b = [int(q) for q in rng.permutation(n)[:2]]
        U = rand_u(4) if rng.integers(0, 2) else numpy.kron(rand_u(2), numpy.eye(2))
        gates.append(((a, b), U))
    for g, U in gates:
        be.kronselect_dot({g: U})
        be64.kronselect_dot({g: U})
    assert be.total_prob() == pytest.approx(1.0, abs=1e-12)
    assert infidelity(be, be64) < 1e-10
    assert max_abs_diff(be, be64) < 1e-6
    for g, U in reversed(gates):
        be.kronselect_dot({g: U.conj().T})
    assert numpy.max(numpy.abs(be.get_relative_range(0, 1024) - start)) < 1e-13
    probs = be.measure_probabilities([0, n - 1, 3])
    assert numpy.allclose(probs, 1 / 8, atol=1e-12)


def test_plan_controls_reach_engine():
    """A Toffoli descriptor becomes one controlled call; the engine result matches the oracle."""
    calls = plan({(0, 1, 2): CMat(CMat(X))})
    assert len(calls) == 1 and calls[0].controls == [0, 1]
    n = 3
    psi = rand_state(n)
    be = gpu_from(psi, n)
    be.kronselect_dot({(0, 1, 2): CMat(CMat(X))})
    assert numpy.max(numpy.abs(be.get_state() - oracle_apply({(0, 1, 2): CMat(CMat(X))}, psi, n))) < 1e-15
