"""Pins the CPU oracle (oracle/qip_oracle.c + oracle/oracle.py) before anything is compared with it.

Sources of truth, none of them the oracle's own formulas:
  * the reference's known-answer tests (tests/utiltest.py: numpy.kron columns, offsets windows,
    func_apply amplitudes);
  * dense operators built independently with numpy tensor contractions over the [2]*n view of the
    state (qubit q = axis q), numpy.block for controlled matrices and axis transposition for
    register swaps;
  * closed forms: QFT == ifft(norm="ortho") (qip/qfft.py with the wrapper kwargs drop), Grover
    amplitudes sin/cos((2t+1) theta), product states == numpy.kron, marginals == sums of
    |psi|^2 over the other tensor axes;
  * exhaustive brute force of bit interleaving on small widths.
A plausible slip (swapped group order, MSB/LSB confusion, wrong control block, dropped term, off
by one in the CDF walk) fails at least one test here.
"""
import random

import numpy
import pytest

from oracle import oracle
from oracle.oracle_backend import OracleBackend
from qip_amd import CMat, F, H, QFFT, Qubit, StochasticMeasure, SwapMat, run

RNG = numpy.random.default_rng(20240611)


# ---- independent dense constructions --------------------------------------------------------

def apply_dense_tensor(state: numpy.ndarray, n: int, qubits, D: numpy.ndarray) -> numpy.ndarray:
    """D (2^L x 2^L, first qubit = MSB of its index) on `qubits` by tensor contraction."""
    L = len(qubits)
    t = state.reshape([2] * n)
    Dt = D.reshape([2] * (2 * L))
    out = numpy.tensordot(Dt, t, axes=(list(range(L, 2 * L)), list(qubits)))
    # tensordot puts the L new axes first; move them back to the qubits' places
    return numpy.moveaxis(out, list(range(L)), list(qubits)).reshape(-1)


def controlled(M: numpy.ndarray) -> numpy.ndarray:
    d = M.shape[0]
    return numpy.block([[numpy.eye(d), numpy.zeros((d, d))], [numpy.zeros((d, d)), M]])


def swap_registers_dense(w: int) -> numpy.ndarray:
    """Operator exchanging two w-qubit registers, by transposing tensor axes of basis vectors."""
    dim = 1 << (2 * w)
    out = numpy.zeros((dim, dim), dtype=numpy.complex128)
    for j in range(dim):
        e = numpy.zeros(dim, dtype=numpy.complex128)
        e[j] = 1
        t = e.reshape([2] * (2 * w)).transpose(list(range(w, 2 * w)) + list(range(w)))
        out[:, j] = t.reshape(-1)
    return out


def rand_unitary(d: int) -> numpy.ndarray:
    z = (RNG.standard_normal((d, d)) + 1j * RNG.standard_normal((d, d))) / numpy.sqrt(2)
    q, r = numpy.linalg.qr(z)
    return q * (numpy.diagonal(r) / numpy.abs(numpy.diagonal(r)))


def rand_state(n: int) -> numpy.ndarray:
    v = RNG.standard_normal(2 ** n) + 1j * RNG.standard_normal(2 ** n)
    return v / numpy.linalg.norm(v)


def oracle_apply(mats, vec, n):
    out = numpy.zeros(2 ** n, dtype=numpy.complex128)
    oracle.kronselect_dot(mats, vec, n, out)
    return out


# ---- tests/utiltest.py known answers ----------------------------------------------------------

I2 = numpy.array([[1, 0], [0, 1]])


@pytest.mark.parametrize("a,b", [(numpy.array([[1, 1], [1, -1]]), numpy.array([[1, 1], [1, -1]])),
                                 (numpy.array([[1, 2], [3, 4]]), numpy.array([[1, 2], [3, 4]])),
                                 (numpy.array([[1, 2], [3, 4]]), numpy.array([[5, 6], [7, 8]]))])
def test_utiltest_kron_columns(a, b):
    """utiltest.py:9-66: {0:a, 2:b} and {(0,2): kron(a,b)} and {(2,0): kron(b,a)} give the
    columns of kron(a, kron(I, b)) exactly."""
    ref = numpy.kron(a, numpy.kron(I2, b))
    for i in range(8):
        e = numpy.zeros(8, dtype=numpy.complex128)
        e[i] = 1
        for mats in ({0: a, 2: b}, {(0, 2): numpy.kron(a, b)}, {(2, 0): numpy.kron(b, a)}):
            assert numpy.array_equal(oracle_apply(mats, e, 3), ref[:, i])


def test_utiltest_output_offset():
    """utiltest.py:68-92."""
    h = numpy.array([[1, 1], [1, -1]])
    ref = numpy.kron(h, numpy.kron(I2, h))
    for i in range(8):
        e = numpy.zeros(8, dtype=numpy.complex128)
        e[i] = 1
        for off in range(4):
            for mats in ({0: h, 2: h}, {(0, 2): numpy.kron(h, h)}):
                t = numpy.zeros(4, dtype=numpy.complex128)
                oracle.kronselect_dot(mats, e, 3, t, output_offset=off)
                assert numpy.array_equal(ref[:, i][off:off + 4], t)


def test_utiltest_input_offset():
    """utiltest.py:94-125: windows of the input summed back give the whole column."""
    h = numpy.array([[1, 1], [1, -1]])
    ref = numpy.kron(h, numpy.kron(I2, h))
    for i in range(8):
        e = numpy.zeros(8, dtype=numpy.complex128)
        e[i] = 1
        for mats in ({0: h, 2: h}, {(0, 2): numpy.kron(h, h)}):
            acc = numpy.zeros(8, dtype=numpy.complex128)
            for off in range(0, 8, 2):
                t = numpy.zeros(8, dtype=numpy.complex128)
                oracle.kronselect_dot(mats, numpy.ascontiguousarray(e[off:off + 2]), 3, t, input_offset=off)
                acc = acc + t
            assert numpy.array_equal(ref[:, i], acc)


def test_utiltest_func_apply():
    """utiltest.py:127-153: f(x) = x+1 mod 4 on |x>|0>."""
    xs, ys = numpy.array([0, 1], dtype=numpy.int32), numpy.array([2, 3], dtype=numpy.int32)
    state = numpy.zeros(16, dtype=numpy.complex128)
    state[0], state[8], state[4], state[12] = 1, 2, 3, 4
    out = numpy.zeros(16, dtype=numpy.complex128)
    oracle.func_apply(xs, ys, lambda x: (x + 1) % 4, state, 4, out)
    assert out[1] == 1 and out[11] == 2 and out[6] == 3 and out[12] == 4


def test_kronselect_errors():
    """qip/util.py:29-33,56-58 and kronprod.pyx:114-116 messages."""
    v = numpy.zeros(8, dtype=numpy.complex128)
    o = numpy.zeros(8, dtype=numpy.complex128)
    with pytest.raises(ValueError, match="Input vector size plus offset"):
        oracle.kronselect_dot({0: numpy.eye(2)}, v, 3, o, input_offset=1)
    with pytest.raises(ValueError, match="Output vector size plus offset"):
        oracle.kronselect_dot({0: numpy.eye(2)}, v, 3, o, output_offset=1)
    with pytest.raises(Exception, match="Shape of square submatrix"):
        oracle.kronselect_dot({(0, 1): numpy.eye(2)}, v, 3, o)
    with pytest.raises(ValueError, match="Cannot pass matrices"):
        oracle._mat_structs([object()])


# ---- brute force against tensor contractions ------------------------------------------------

@pytest.mark.parametrize("trial", range(12))
def test_cdot_random_disjoint_groups(trial):
    n = int(RNG.integers(3, 8))
    qs = list(RNG.permutation(n))
    mats, groups = {}, []
    while len(qs) >= 1 and len(groups) < 3:
        L = int(RNG.integers(1, min(3, len(qs)) + 1))
        g, qs = tuple(int(q) for q in qs[:L]), qs[L:]
        kind = int(RNG.integers(0, 3))
        if kind == 0 or L == 1:
            M = rand_unitary(2 ** L)
            mats[g if L > 1 else g[0]] = M
            groups.append((g, M))
        elif kind == 1:
            M = rand_unitary(2 ** (L - 1))
            mats[g] = CMat(M)
            groups.append((g, controlled(M)))
        else:
            if L % 2:
                M = rand_unitary(2 ** L)
                mats[g] = M
                groups.append((g, M))
            else:
                mats[g] = SwapMat(L // 2)
                groups.append((g, swap_registers_dense(L // 2)))
    psi = rand_state(n)
    expect = psi
    for g, D in groups:
        expect = apply_dense_tensor(expect, n, g, D)
    assert numpy.max(numpy.abs(oracle_apply(mats, psi, n) - expect)) < 1e-12


def test_cdot_nested_cmat_and_cswap():
    """C(C(X)) = Toffoli; C(SwapMat(2)) on 5 qubits = controlled register swap."""
    n = 5
    psi = rand_state(n)
    X = numpy.array([[0, 1], [1, 0]], dtype=complex)
    got = oracle_apply({(4, 1, 2): CMat(CMat(X))}, psi, n)
    expect = apply_dense_tensor(psi, n, (4, 1, 2), controlled(controlled(X)))
    assert numpy.max(numpy.abs(got - expect)) < 1e-15
    got = oracle_apply({(2, 0, 4, 1, 3): CMat(SwapMat(2))}, psi, n)
    expect = apply_dense_tensor(psi, n, (2, 0, 4, 1, 3), controlled(swap_registers_dense(2)))
    assert numpy.max(numpy.abs(got - expect)) < 1e-15


def test_cdot_shared_control_index():
    """C(Not) on a 2-qubit register: {(c,t1): CMat(X), (c,t2): CMat(X)} is controlled-(X kron X)
    (qip/operators.py:41-43,221-227)."""
    n = 4
    psi = rand_state(n)
    X = numpy.array([[0, 1], [1, 0]], dtype=complex)
    got = oracle_apply({(1, 3): CMat(X), (1, 0): CMat(X)}, psi, n)
    expect = apply_dense_tensor(psi, n, (1, 3, 0), controlled(numpy.kron(X, X)))
    assert numpy.max(numpy.abs(got - expect)) < 1e-15


def test_entwine_bit_is_interleave():
    """util.pyx:2-23 equals pdep(m, mask) | pdep(r, ~mask), exhaustively for n <= 5."""
    for n in range(1, 6):
        for mask in range(1 << n):
            k = bin(mask).count("1")
            ones = [b for b in range(n) if mask >> b & 1]
            zeros = [b for b in range(n) if not mask >> b & 1]
            for m in range(1 << k):
                for r in range(1 << (n - k)):
                    want = sum(((m >> j) & 1) << ones[j] for j in range(k)) | \
                        sum(((r >> j) & 1) << zeros[j] for j in range(n - k))
                    assert oracle.entwine_bit(k, n - k, m, r, mask) == want


# ---- func_apply ------------------------------------------------------------------------------

def func_apply_expected(reg1, reg2, table, psi, n):
    """|x>|q>|r> -> |x>|f(x) xor q>|r> on the [2]*n tensor, by explicit basis relabelling."""
    out = numpy.zeros_like(psi)
    n1, n2 = len(reg1), len(reg2)
    for i in range(2 ** n):
        bits = [(i >> (n - 1 - q)) & 1 for q in range(n)]
        x = int("".join(str(bits[q]) for q in reg1), 2)
        qv = int("".join(str(bits[q]) for q in reg2), 2)
        y = (int(table[x]) ^ qv) & ((1 << n2) - 1)
        nb = list(bits)
        for j, q in enumerate(reg2):
            nb[q] = (y >> (n2 - 1 - j)) & 1
        out[int("".join(map(str, nb)), 2)] = psi[i]
    return out


@pytest.mark.parametrize("trial", range(8))
def test_func_apply_permutation(trial):
    n = int(RNG.integers(3, 8))
    perm = [int(q) for q in RNG.permutation(n)]
    n1 = int(RNG.integers(1, n))
    n2 = int(RNG.integers(1, n - n1 + 1))
    reg1, reg2 = perm[:n1], perm[n1:n1 + n2]
    table = RNG.integers(-50, 50, size=2 ** n1).astype(numpy.int32)
    psi = rand_state(n)
    expect = func_apply_expected(reg1, reg2, table, psi, n)
    out = numpy.zeros_like(psi)
    oracle.func_apply(reg1, reg2, None, psi, n, out, faithful=False, table=table)
    assert numpy.array_equal(out, expect)
    n_other = n - n1 - n2
    out_f = numpy.zeros_like(psi)
    oracle.func_apply(reg1, reg2, None, psi, n, out_f, faithful=True, table=table)
    assert numpy.array_equal(out_f, reference_bound_expected(reg1, reg2, expect, n))


def reference_bound_expected(reg1, reg2, expect, n):
    """The reference's output under its rest-qubit loop bound (func_apply.pyx:107-111): only the
    first len(reg2) rest qubits in qubit order are iterated, so the full permutation's amplitudes
    with any later rest qubit at 1 stay at the 0 the output was cleared to (func_apply.pyx:42-44).
    With no more rest qubits than reg2 qubits this is the full permutation."""
    rest = [q for q in range(n) if q not in reg1 and q not in reg2]
    dropped = sum(1 << (n - 1 - q) for q in rest[len(reg2):])
    keep = (numpy.arange(2 ** n) & dropped) == 0
    return numpy.where(keep, expect, 0)


@pytest.mark.parametrize("n,reg1,reg2", [
    (7, [1, 4], [6], ),               # rest 0, 2, 3, 5: 2, 3, 5 dropped
    (8, [7, 0], [3, 5], ),            # rest 1, 2, 4, 6: 4, 6 dropped
    (6, [2], [0], ),                  # rest 1, 3, 4, 5: 3, 4, 5 dropped
    (6, [5, 4, 3], [0, 1, 2], ),      # no rest qubits
])
def test_func_apply_reference_bound(n, reg1, reg2):
    """faithful=True against the full permutation with the dropped rest qubits' amplitudes cleared,
    on cases with more rest qubits than reg2 qubits (amplitude is lost) and with none."""
    table = RNG.integers(-9, 9, size=2 ** len(reg1)).astype(numpy.int32)
    psi = rand_state(n)
    expect = func_apply_expected(reg1, reg2, table, psi, n)
    out_f = numpy.zeros_like(psi)
    oracle.func_apply(reg1, reg2, None, psi, n, out_f, faithful=True, table=table)
    bound = reference_bound_expected(reg1, reg2, expect, n)
    assert numpy.array_equal(out_f, bound)
    rest = n - len(reg1) - len(reg2)
    if rest > len(reg2):
        assert numpy.sum(numpy.abs(out_f) ** 2) < numpy.sum(numpy.abs(psi) ** 2) - 1e-3


def test_func_apply_overflow():
    with pytest.raises(OverflowError):
        oracle.func_table(lambda x: 2 ** 40, 2)


# ---- measurement family ---------------------------------------------------------------------

def marginal_tensor(psi, n, indices):
    """P over `indices` with indices[j] on bit j of the outcome (measure_probabilities order)."""
    p = (numpy.abs(psi) ** 2).reshape([2] * n)
    others = [q for q in range(n) if q not in indices]
    p = p.sum(axis=tuple(others)) if others else p
    # remaining axes are the measured qubits in ascending qubit order; outcome bit j <- indices[j]
    asc = sorted(indices)
    p = numpy.transpose(p, [asc.index(q) for q in reversed(indices)])
    return p.reshape(-1)


@pytest.mark.parametrize("trial", range(6))
def test_measure_probabilities(trial):
    n = int(RNG.integers(2, 8))
    k = int(RNG.integers(1, n + 1))
    idx = [int(q) for q in RNG.permutation(n)[:k]]
    psi = rand_state(n)
    got = oracle.measure_probabilities(idx, n, psi)
    assert numpy.allclose(got, marginal_tensor(psi, n, idx), atol=1e-15, rtol=1e-12)


def test_measure_probabilities_bit_order():
    """Q9: indices = [0..n-1] gives the bit-reversed order of |psi|^2."""
    n = 4
    psi = rand_state(n)
    got = oracle.measure_probabilities(list(range(n)), n, psi)
    rev = [int(format(i, "04b")[::-1], 2) for i in range(16)]
    assert numpy.allclose(got, (numpy.abs(psi) ** 2)[rev])


def entwine_marginals(psi, n, indices):
    """Marginals with the smallest measured qubit as the MSB of the outcome."""
    return marginal_tensor(psi, n, sorted(indices, reverse=True))


@pytest.mark.parametrize("trial", range(6))
def test_soft_measure_walk(trial):
    n = int(RNG.integers(2, 7))
    k = int(RNG.integers(1, n + 1))
    idx = [int(q) for q in RNG.permutation(n)[:k]]
    psi = rand_state(n) * 1.7            # unnormalised on purpose: r is scaled by the norm
    marg = entwine_marginals(psi, n, idx)
    cdf = numpy.cumsum(marg) / numpy.sum(marg)
    for u in RNG.uniform(0, 1, size=20):
        if numpy.min(numpy.abs(cdf - u)) < 1e-9:
            continue
        want = int(numpy.searchsorted(cdf, u))
        m, p = oracle.soft_measure(idx, n, psi, u=float(u))
        assert m == want
        assert p == pytest.approx(marg[want], rel=1e-12)
    m, p = oracle.soft_measure(idx, n, psi, measured=len(marg) - 1, u=0.999)
    assert m == len(marg) - 1 and p == pytest.approx(marg[-1], rel=1e-12)


def test_soft_measure_draws_python_random():
    """kronprod.pyx:362: one random.random() per call from the global generator."""
    n, psi = 3, rand_state(3)
    random.seed(7)
    a = oracle.soft_measure([0, 2], n, psi)
    random.seed(7)
    u = random.random()
    b = oracle.soft_measure([0, 2], n, psi, u=u)
    assert a == b


def test_measure_collapse_and_reduce():
    n = 5
    psi = rand_state(n)
    idx = [3, 1]
    marg = entwine_marginals(psi, n, idx)
    for m in range(4):
        out = numpy.zeros_like(psi)
        mm, p = oracle.measure(idx, n, psi, out, measured=m, measured_prob=float(marg[m]))
        t = psi.reshape([2] * n).copy()
        keep = numpy.zeros_like(t)
        sl = [slice(None)] * n
        sl[1], sl[3] = (m >> 1) & 1, m & 1          # qubit 1 is the MSB of m
        keep[tuple(sl)] = t[tuple(sl)] / numpy.sqrt(marg[m])
        assert numpy.allclose(out, keep.reshape(-1), atol=1e-15)
        red = numpy.zeros_like(psi)
        oracle.reduce_measure(idx, n, psi, red, measured=m, measured_prob=float(marg[m]))
        reduced = t[tuple(sl)].reshape(-1) / numpy.sqrt(marg[m])
        assert numpy.allclose(red, numpy.tile(reduced, 4), atol=1e-15)


def test_measure_errors():
    psi = rand_state(3)
    out = numpy.zeros_like(psi)
    with pytest.raises(ValueError, match="Measured value must be less"):
        oracle.measure([0], 3, psi, out, measured=2)
    with pytest.raises(ValueError, match="measured_prob must be"):
        oracle.measure([0], 3, psi, out, measured=1, measured_prob=1.5)


def test_top_probabilities():
    n = 5
    psi = rand_state(n)
    idx = [4, 0, 2]
    marg = entwine_marginals(psi, n, idx)
    order = sorted(range(8), key=lambda m: (-marg[m], m))
    ms, ps = oracle.measure_top_probabilities(idx, n, 3, psi, faithful=False)
    assert ms == order[:3]
    assert numpy.allclose(ps, marg[order[:3]])
    # the reference test's case, through its heap (qiptest.py:304-315)
    psi2 = numpy.kron(numpy.array([0.6, 0.8, 0.0, 0.0]), numpy.array([0, 0, 1.0, 0])).astype(complex)
    ms, ps = oracle.measure_top_probabilities([0, 1], 4, 3, psi2, faithful=True)
    assert ms[:2] == [1, 0] and numpy.allclose(ps, [0.64, 0.36, 0.0])


# ---- initial state and whole circuits --------------------------------------------------------

def test_make_state_is_kron():
    """backend.py:73-104: feeds in qid order form a Kronecker product; unfed qubits are |0>."""
    a, b = rand_state(2), rand_state(1)
    got = oracle.make_state_vector(4, [[0, 1], [3]], [a, b])
    expect = numpy.kron(numpy.kron(a, numpy.array([1, 0])), b)
    assert numpy.allclose(got, expect, atol=0)
    got = oracle.make_state_vector(3, [[2, 0]], [numpy.array([1, 2, 3, 4])])   # joint feed, tuple key order
    expect = numpy.zeros(8, dtype=complex)
    for v in range(4):
        expect[((v >> 1) & 1) << 0 | (v & 1) << 2] = v + 1
    assert numpy.array_equal(got, expect)
    assert numpy.array_equal(oracle.make_state_vector(2, [], []), [1, 0, 0, 0])


@pytest.mark.parametrize("n", range(2, 9))
def test_qft_closed_form(n):
    """Q7: QFFT(q, rev=True) as built by qip/qfft.py equals ifft(norm='ortho')."""
    psi = rand_state(n)
    q = Qubit(n=n)
    o, _ = QFFT(q).run(feed={q: psi}, backend_constructor=OracleBackend.make_state)
    assert numpy.max(numpy.abs(o - numpy.fft.ifft(psi, norm="ortho"))) < 1e-13


def grover_graph(ns, x0, iters):
    """grovers_iterative.py:20-26 chained `iters` times: oracle F(x==x0), diffusion H F(x==0) H."""
    q = Qubit(n=ns, default=numpy.ones(2 ** ns) / numpy.sqrt(2 ** ns))
    anc = Qubit(n=1, default=[1 / numpy.sqrt(2), -1 / numpy.sqrt(2)])
    s, a = q, anc
    for _ in range(iters):
        s, a = F(lambda x: int(x == x0), s, a)
        hs = H(s)
        hs, a = F(lambda x: int(x == 0), hs, a)
        s = H(hs)
    return s, a


@pytest.mark.parametrize("iters", [1, 2, 3])
def test_grover_closed_form(iters):
    ns, x0 = 6, 37
    s, a = grover_graph(ns, x0, iters)
    o, _ = run(s, a, backend_constructor=OracleBackend.make_state)
    N = 2 ** ns
    theta = numpy.arcsin(1 / numpy.sqrt(N))
    sign = (-1) ** iters                       # H F(x==0) H = -(2|s><s| - I)
    marked = sign * numpy.sin((2 * iters + 1) * theta)
    other = sign * numpy.cos((2 * iters + 1) * theta) / numpy.sqrt(N - 1)
    search = numpy.full(N, other)
    search[x0] = marked
    expect = numpy.kron(search, numpy.array([1, -1]) / numpy.sqrt(2))
    assert numpy.max(numpy.abs(o - expect)) < 1e-13


def test_stochastic_measure_graph():
    """StochasticMeasure on a product state equals the feed's |amplitude|^2 in LSB-first order."""
    q1, q2 = Qubit(n=2), Qubit(n=3)
    a = rand_state(3)
    sm = StochasticMeasure(q2)
    _, c = run(sm, q1, feed={q1: [1, 0, 0, 0], q2: a}, backend_constructor=OracleBackend.make_state)
    rev = [int(format(i, "03b")[::-1], 2) for i in range(8)]
    assert numpy.allclose(c[sm], (numpy.abs(a) ** 2)[rev])


@pytest.mark.parametrize("diffusion", ["f", "mcz"])
def test_grover_workload_closed_form(diffusion):
    """The Grover workload's operation list (both diffusion forms) on the oracle equals the closed
    form after the full floor(pi/4 sqrt(N)) iterations."""
    from qip_amd.workloads import apply_ops, grover_closed_form, grover_initial, grover_iteration_ops, grover_iterations
    ns, xs = 7, 90
    groups, feeds = grover_initial(ns)
    be = OracleBackend.make_state(ns + 1, groups, feeds)
    t = grover_iterations(ns)
    for _ in range(t):
        apply_ops(be, grover_iteration_ops(ns, xs, diffusion))
    assert numpy.max(numpy.abs(be.get_state() - grover_closed_form(ns, xs, t))) < 1e-13
