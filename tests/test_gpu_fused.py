"""GPU parity of the fused split-BF16 MFMA kernel (complex64) against the plain complex128
engine on the same inputs (marked gpu). Bars: infidelity <= 1e-9 per fused block, amplitude
error within FP32-level relative accuracy, FP64 norm bookkeeping exact to rounding."""
import numpy
import pytest

from qip_amd import GpuBackend, infidelity, max_abs_diff

pytestmark = pytest.mark.gpu
RNG = numpy.random.default_rng(777)


def rand_state(n):
    v = RNG.standard_normal(2 ** n) + 1j * RNG.standard_normal(2 ** n)
    return v / numpy.linalg.norm(v)


def rand_u(d):
    z = (RNG.standard_normal((d, d)) + 1j * RNG.standard_normal((d, d))) / numpy.sqrt(2)
    q, r = numpy.linalg.qr(z)
    return q * (numpy.diagonal(r) / numpy.abs(numpy.diagonal(r)))


def pair(n):
    psi = rand_state(n)
    a = GpuBackend.make_state(n, [list(range(n))], [psi])
    b = GpuBackend.make_state(n, [list(range(n))], [psi], statetype=numpy.complex64)
    return a, b


@pytest.mark.parametrize("n", [12, 13, 16, 20])
@pytest.mark.parametrize("k", [6, 3, 1])
def test_fused_block_matches_plain(n, k):
    ref, fz = pair(n)
    qs = [int(q) for q in RNG.permutation(n)[:k]]
    U = rand_u(2 ** k)
    ref.apply_matrix(qs, U)
    fz.apply_fused(qs, U)
    fz.fused_finish()
    inf = infidelity(ref, fz)
    err = max_abs_diff(ref, fz)
    amp = 2.0 ** (-n / 2)
    assert inf < 1e-9, inf
    assert err < 1e-4 * amp * 8, (err, amp)


@pytest.mark.parametrize("targets", [[0, 1, 2, 3, 4, 5], [13, 12, 11, 10, 9, 8], [0, 13, 3, 9, 6, 11],
                                     [7, 2, 12, 5, 10, 1]])
def test_fused_target_layouts(targets):
    """High bits, low bits and interleaved target sets all exercise different tile shapes."""
    n = 14
    ref, fz = pair(n)
    U = rand_u(64)
    ref.apply_matrix(targets, U)
    fz.apply_fused(targets, U)
    assert infidelity(ref, fz) < 1e-9


def test_fused_sequence_and_renormalisation():
    n = 16
    ref, fz = pair(n)
    for _ in range(20):
        qs = [int(q) for q in RNG.permutation(n)[:6]]
        U = rand_u(64)
        ref.apply_matrix(qs, U)
        fz.apply_fused(qs, U)
        last = fz.fused_last_norm()
        assert abs(last - 1.0) < 1e-5
    fz.fused_finish()
    assert abs(fz.total_prob() - 1.0) < 1e-6
    assert infidelity(ref, fz) < 20 * 1e-9


def test_fused_then_plain_settles():
    """A plain call after fused blocks sees the renormalised state."""
    n = 12
    ref, fz = pair(n)
    U = rand_u(64)
    qs = [0, 1, 2, 3, 4, 5]
    ref.apply_matrix(qs, U)
    fz.apply_fused(qs, U)
    H = numpy.array([[1, 1], [1, -1]]) / numpy.sqrt(2)
    ref.kronselect_dot({7: H})
    fz.kronselect_dot({7: H})
    assert abs(fz.total_prob() - 1.0) < 1e-6
    assert infidelity(ref, fz) < 1e-9


def test_pending_scale_through_linear_sweeps():
    """Each fused block leaves its 1/sqrt(norm) pending. Gate, product, diagonal and func_apply sweeps
    are linear and leave it pending; the next fused block applies it as it loads, a read applies it
    to the whole state. Starting from 2 psi makes a dropped or doubled scale visible."""
    n = 14
    psi = rand_state(n)
    ref = GpuBackend.make_state(n, [list(range(n))], [psi])
    fz = GpuBackend.make_state(n, [list(range(n))], [2.0 * psi], statetype=numpy.complex64)
    H = numpy.array([[1, 1], [1, -1]]) / numpy.sqrt(2)
    S = numpy.diag([1.0, 1j])
    CX = numpy.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=complex)
    f = lambda x: (3 * x + 1) % 8
    U1, U2 = rand_u(64), rand_u(8)
    for be, fused in ((ref, False), (fz, True)):
        (be.apply_fused if fused else be.apply_matrix)([0, 3, 5, 8, 9, 13], U1)
        be.kronselect_dot({q: H for q in range(1, 11)})    # product passes
        be.kronselect_dot({4: S})                          # diagonal
        be.kronselect_dot({(2, 6): CX})                    # controlled gate
        be.func_apply([7, 10, 11], [12, 1, 0], f)
        (be.apply_fused if fused else be.apply_matrix)([13, 2, 6], U2)
        be.kronselect_dot({12: H})
    assert abs(fz.total_prob() - 1.0) < 1e-6
    assert infidelity(ref, fz) < 1e-8
    assert max_abs_diff(ref, fz) < 1e-4 * 2.0 ** (-n / 2) * 16


def test_fused_rejects():
    a = GpuBackend.make_state(12, [], [])
    with pytest.raises(ValueError):
        a.apply_fused([0], numpy.eye(2))          # complex128 state
    b = GpuBackend.make_state(8, [], [], statetype=numpy.complex64)
    with pytest.raises(NotImplementedError):
        b.apply_fused([0], numpy.eye(2))          # n < 12


def test_fused_backend_layered_circuit():
    """The precision-form backend on a layered random circuit vs the plain complex128 backend."""
    from qip_amd.fused import FusedBackend
    from qip_amd.workloads import layered_random_circuit
    n = 16
    ops = layered_random_circuit(n, 12, seed=3)
    pl = GpuBackend.make_state(n, [], [])
    fz = FusedBackend.make_state(n, [], [])
    deferred = FusedBackend.make_state(n, [], [])
    deferred.streaming = False
    for d in ops:
        pl.kronselect_dot(d)
        fz.kronselect_dot(d)
        deferred.kronselect_dot(d)
    assert deferred.pending_gates() == len(ops) and deferred.blocks_applied == 0
    assert fz.blocks_applied > 0 and fz.pending_gates() < len(ops)     # streamed before any read
    got = fz.get_state()                  # flushes the rest
    assert fz.blocks_applied < len(ops) // 4
    assert numpy.array_equal(got, deferred.get_state())    # same blocks, same order, same kernels
    assert deferred.blocks_applied == fz.blocks_applied
    assert infidelity(pl, fz) < 1e-8
    assert numpy.max(numpy.abs(got - pl.get_state())) < 1e-5
    assert abs(fz.total_prob() - 1.0) < 1e-6


def _apply_python_plan(ref, n, blocks):
    """The recorder's launches restated: each block composed by qipgpu_compose and applied (MFMA
    kernel at n >= 12, plain dense sweep below); at n >= 12 a run of consecutive single-qubit-only
    blocks is one product of per-qubit 2x2s (each composed by qipgpu_compose) on the product sweep."""
    from qip_amd.fused import compose
    i = 0
    while i < len(blocks):
        if n >= 12 and all(len(gq) == 1 for gq, _ in blocks[i].gates):
            per = {}
            while i < len(blocks) and all(len(gq) == 1 for gq, _ in blocks[i].gates):
                for gq, G in blocks[i].gates:
                    per.setdefault(gq[0], []).append((gq, G))
                i += 1
            qs = sorted(per)
            ref.apply_product(qs, [compose([q], per[q]) for q in qs])
            continue
        qs, M = blocks[i].matrix()
        if n >= 12:
            ref.apply_fused(qs, M)
        else:
            ref.apply_matrix(qs, M)
        i += 1


@pytest.mark.parametrize("n,streaming", [(16, True), (16, False), (10, True)])
def test_recorder_blocks_equal_python_plan(n, streaming):
    """The library's recorder launches exactly the blocks of the Python statement of the light-cone
    rule, in the same launches: a bitwise equal state, streamed or deferred."""
    from qip_amd.fused import FusedBackend, descriptor_gates, fuse
    from qip_amd.workloads import layered_random_circuit
    ops = layered_random_circuit(n, 10, seed=n)
    fz = FusedBackend.make_state(n, [], [])
    fz.streaming = streaming
    for d in ops:
        fz.kronselect_dot(d)
    got = fz.get_state()
    blocks = fuse([g for d in ops for g in descriptor_gates(d)], 6)
    assert fz.blocks_applied == len(blocks)
    ref = GpuBackend.make_state(n, [], [], statetype=numpy.complex64)
    _apply_python_plan(ref, n, blocks)
    assert numpy.array_equal(got, ref.get_state())


@pytest.mark.parametrize("n,seed", [(14, 1), (17, 2)])
def test_recorder_random_gates(n, seed):
    """Gates of 1..4 qubits on random qubit sets (the look-ahead of the plan at work): streamed and
    deferred recording give the same bits, so do the Python plan's blocks applied one by one, and
    the result stays within 1e-9 infidelity per block of complex128."""
    from qip_amd.fused import FusedBackend, fuse
    rng = numpy.random.default_rng(seed)
    gates = []
    for _ in range(160):
        L = int(rng.integers(1, 5))
        gates.append((tuple(int(q) for q in rng.choice(n, L, replace=False)), rand_u(2 ** L)))
    psi = rand_state(n)
    outs = []
    for streaming in (True, False):
        fz = FusedBackend.make_state(n, [list(range(n))], [psi])
        fz.streaming = streaming
        for gq, G in gates:
            fz.kronselect_dot({gq: G})
        outs.append(fz.get_state())
    assert numpy.array_equal(outs[0], outs[1])
    blocks = fuse(gates, 6)
    assert fz.blocks_applied == len(blocks)
    ref = GpuBackend.make_state(n, [list(range(n))], [psi], statetype=numpy.complex64)
    _apply_python_plan(ref, n, blocks)
    assert numpy.array_equal(outs[0], ref.get_state())
    exact = GpuBackend.make_state(n, [list(range(n))], [psi])
    for gq, G in gates:
        exact.kronselect_dot({gq: G})
    assert infidelity(exact, fz) < 1e-9 * len(blocks)
    assert numpy.max(numpy.abs(exact.get_state() - outs[0])) < 1e-5


def test_recorder_repeated_pattern():
    """The same gate pattern recorded again (new matrices, a flush after each repetition, as in an
    iterative algorithm) is planned from the recorder's memo of recent plans: bitwise the blocks of
    the Python plan of each repetition."""
    from qip_amd.fused import FusedBackend, descriptor_gates, fuse
    from qip_amd.workloads import layered_random_circuit
    n = 15
    psi = rand_state(n)
    fz = FusedBackend.make_state(n, [list(range(n))], [psi])
    ref = GpuBackend.make_state(n, [list(range(n))], [psi], statetype=numpy.complex64)
    for rep in range(3):
        ops = layered_random_circuit(n, 6, seed=40 + rep)      # same pattern, new unitaries
        for d in ops:
            fz.kronselect_dot(d)
        fz.flush()
        _apply_python_plan(ref, n, fuse([g for d in ops for g in descriptor_gates(d)], 6))
    assert numpy.array_equal(fz.get_state(), ref.get_state())


def test_single_qubit_runs_take_the_product_sweep():
    """Consecutive blocks of single-qubit gates only are one product of per-qubit 2x2s applied with
    FP64 arithmetic: the same bits as the Python restatement, and (on random single-qubit layers)
    much closer to the complex128 result than the same blocks on split-BF16 MFMA."""
    from qip_amd.fused import FusedBackend, descriptor_gates, fuse
    n = 18
    psi = rand_state(n)
    H = numpy.array([[1, 1], [1, -1]]) / numpy.sqrt(2)
    ops = [{(3, 4): rand_u(4)}, {q: H for q in range(17)}, {q: rand_u(2) for q in range(0, 17, 2)},
           {q: H for q in range(17)}]
    fz = FusedBackend.make_state(n, [list(range(n))], [psi])
    for d in ops:
        fz.kronselect_dot(d)
    got = fz.get_state()
    blocks = fuse([g for d in ops for g in descriptor_gates(d)], 6)
    kinds = [all(len(gq) == 1 for gq, _ in b.gates) for b in blocks]
    assert kinds == [False, True, True]      # the 2-qubit gate lands in the first block
    ref = GpuBackend.make_state(n, [list(range(n))], [psi], statetype=numpy.complex64)
    _apply_python_plan(ref, n, blocks)
    assert numpy.array_equal(got, ref.get_state())
    exact = GpuBackend.make_state(n, [list(range(n))], [psi])
    for d in ops:
        exact.kronselect_dot(d)
    assert numpy.max(numpy.abs(exact.get_state() - got)) < 2e-7
    # random single-qubit layers only: every block a product; against the blocks on MFMA
    ops = [{q: rand_u(2) for q in range(n)} for _ in range(3)]
    fz.reinit([list(range(n))], [psi])
    mf = GpuBackend.make_state(n, [list(range(n))], [psi], statetype=numpy.complex64)
    exact.reinit([list(range(n))], [psi])
    for d in ops:
        fz.kronselect_dot(d)
        exact.kronselect_dot(d)
    for b in fuse([g for d in ops for g in descriptor_gates(d)], 6):
        mf.apply_fused(*b.matrix())
    e = exact.get_state()
    err_product = numpy.max(numpy.abs(fz.get_state() - e))
    err_mfma = numpy.max(numpy.abs(mf.get_state() - e))
    assert err_product < err_mfma / 10, (err_product, err_mfma)


def test_recorder_orders_wide_and_direct_calls():
    """Gates wider than the fusion width (here a 7-qubit multi-controlled Z, built from CMat, never
    expanded to a dense matrix) and direct calls between recorded gates act in program order."""
    from qip_amd.fused import FusedBackend
    from qip_amd.operators import CMat
    n = 13
    psi = rand_state(n)
    pl = GpuBackend.make_state(n, [list(range(n))], [psi])
    fz = FusedBackend.make_state(n, [list(range(n))], [psi])
    H = numpy.array([[1, 1], [1, -1]]) / numpy.sqrt(2)
    m = numpy.diag([1.0, -1.0])
    for _ in range(6):
        m = CMat(m)
    U = rand_u(4)
    for be in (pl, fz):
        be.kronselect_dot({q: H for q in range(n)})
        be.kronselect_dot({tuple(range(7)): m})           # 6 controls + Z: applied unfused
        be.kronselect_dot({(3, 9): U})
        be.apply_product([0, 5], [H, H])
        be.func_apply([1, 2], [11, 12], lambda x: x ^ 1)  # direct call between recorded gates
        be.kronselect_dot({12: H, 4: H})
    assert fz.pending_gates() > 0 or fz.blocks_applied > 0
    assert infidelity(pl, fz) < 1e-8
    assert fz.pending_gates() == 0


def test_recorder_edges():
    """Recorded gates and host calls: reinit drops the record; range writes, measurement and a
    change of streaming launch it first; closing with gates recorded is fine; a bad gate is
    rejected with nothing recorded; a complex128 state cannot record."""
    import random
    from qip_amd import _capi
    from qip_amd.fused import FusedBackend
    n = 12
    H = numpy.array([[1, 1], [1, -1]]) / numpy.sqrt(2)
    psi = rand_state(n)
    # reinit drops what was recorded
    fz = FusedBackend.make_state(n, [], [])
    fz.streaming = False
    fz.kronselect_dot({q: H for q in range(n)})
    assert fz.pending_gates() == n
    fz.reinit([list(range(n))], [psi])
    assert fz.pending_gates() == 0
    assert numpy.max(numpy.abs(fz.get_state() - psi)) < 1e-7
    # a range write after recorded gates sees them applied first
    ref = GpuBackend.make_state(n, [list(range(n))], [psi], statetype=numpy.complex64)
    U37 = rand_u(4)
    for be in (ref, fz):
        be.kronselect_dot({q: H for q in range(0, n, 2)})
        be.overwrite_relative_range(5, 9, numpy.arange(4) * 0.01)
        be.addto_relative_range(100, 102, numpy.array([0.5, -0.5]))
        be.kronselect_dot({(3, 7): U37})
    want = ref.get_state().astype(numpy.complex128)
    want /= numpy.linalg.norm(want)       # every fused block leaves the state normalised
    assert numpy.max(numpy.abs(fz.get_state() - want)) < 1e-5
    # measurement launches the record first: the outcome the plain state gives for the same uniform
    # (the outcome depends only on the relative probabilities)
    for be in (ref, fz):
        be.kronselect_dot({1: H, 2: H})
    random.seed(11)
    m_ref = ref.measure([1, 2])
    random.seed(11)
    m_fz = fz.measure([1, 2])
    assert m_ref[0] == m_fz[0]
    # switching streaming back on launches the record
    fz.kronselect_dot({0: H})
    assert fz.pending_gates() == 1
    fz.streaming = True
    assert fz.pending_gates() == 0
    fz.kronselect_dot({4: H})
    fz.close()                              # gates still recorded: dropped with the state
    # all or nothing: the second gate is wider than the width, the first is not recorded either
    fz = FusedBackend.make_state(n, [], [])
    fz.streaming = False
    k = numpy.array([1, 3], dtype=numpy.int32)
    q = numpy.zeros(12, dtype=numpy.int32)
    q[6:9] = [0, 1, 2]
    m = numpy.zeros(4 + 64, dtype=numpy.complex128)
    lib = _capi.load()
    assert lib.qipgpu_fused_set_recording(fz.handle, 2, 0) == 0
    assert lib.qipgpu_fused_record(fz.handle, 2, _capi.ptr(k), _capi.ptr(q), _capi.ptr(m)) == _capi.EINVAL
    assert fz.pending_gates() == 0
    fz.close()
    # a complex128 state records too (its blocks are applied in FP64); recording nothing is a no-op
    c128 = GpuBackend.make_state(n, [], [])
    assert lib.qipgpu_fused_record(c128.handle, 0, None, None, None) == 0
    assert numpy.array_equal(c128.get_state(), GpuBackend.make_state(n, [], []).get_state())
    c128.close()


def test_fused_backend_graph_qft():
    """A qip graph run through FusedBackend: the 14-qubit QFT equals ifft(norm='ortho')."""
    from qip_amd import QFFT, Qubit, run
    from qip_amd.fused import FusedBackend
    n = 14
    psi = rand_state(n)
    q = Qubit(n=n)
    o, _ = run(QFFT(q), feed={q: psi}, backend_constructor=FusedBackend.make_state)
    assert o.dtype == numpy.complex64
    o = o.astype(numpy.complex128)         # norms in FP64: a complex64 vdot is good to ~1e-7 only
    expect = numpy.fft.ifft(psi, norm="ortho")
    ov = numpy.vdot(expect, o)
    assert 1 - abs(ov) ** 2 / numpy.vdot(o, o).real < 1e-8


def test_fused_backend_24_qubits_against_oracle():
    """The precision form at 24 qubits against the oracle's FP64 restatement of cdot_loop (one
    out-of-place pass per descriptor on the CPU): layered random circuit, depth 4, 142 gates."""
    import os
    from oracle import oracle
    from qip_amd.fused import FusedBackend
    from qip_amd.workloads import layered_random_circuit
    n = 24
    ops = layered_random_circuit(n, 4, seed=5)
    fz = FusedBackend.make_state(n, [], [])
    for d in ops:
        fz.kronselect_dot(d)
    got = fz.get_state().astype(numpy.complex128)
    fz.close()
    v = numpy.zeros(2 ** n, dtype=numpy.complex128)
    v[0] = 1.0
    out = numpy.zeros_like(v)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    for d in ops:
        oracle.kronselect_dot(d, v, n, out, fast=True, nthreads=threads)
        v, out = out, v
    ov = numpy.vdot(v, got)
    inf = 1.0 - abs(ov) ** 2 / (numpy.vdot(v, v).real * numpy.vdot(got, got).real)
    print("24-qubit fused vs oracle: 1-F", inf, flush=True)
    assert inf < 1e-8
    assert numpy.max(numpy.abs(got - v)) < 1e-6


def test_pattern_sweep_leaves_state_unchanged():
    """The pattern-floor diagnostic moves every amplitude in place without changing it, for target
    sets of every kind (low, high, mixed, fewer than six)."""
    n = 14
    psi = rand_state(n)
    fz = GpuBackend.make_state(n, [list(range(n))], [psi], statetype=numpy.complex64)
    before = fz.get_state()
    for t in ([0, 1, 2, 3, 4, 5], [13, 12, 11, 10, 9, 8], [0, 13, 3, 9, 6, 11], [7], [2, 11]):
        fz.fused_pattern_sweep(t)
    assert numpy.array_equal(fz.get_state(), before)
    with pytest.raises(ValueError):
        fz.fused_pattern_sweep([0, 0])
    fz.close()


@pytest.mark.parametrize("width", [2, 4])
def test_recorder_narrower_width(width):
    """A FusedBackend with a narrower fusion width: the library fuses to that width (the blocks of the
    Python rule at that width, bitwise), and gates wider than it go to the plain path in order."""
    from qip_amd.fused import FusedBackend, descriptor_gates, fuse
    from qip_amd.workloads import layered_random_circuit

    class Narrow(FusedBackend):
        fusion_width = width

    n = 14
    ops = layered_random_circuit(n, 6, seed=width)
    fz = Narrow.make_state(n, [], [])
    fz.streaming = False
    for d in ops:
        fz.kronselect_dot(d)
    got = fz.get_state()
    blocks = fuse([g for d in ops for g in descriptor_gates(d)], width)
    assert fz.blocks_applied == len(blocks)
    ref = GpuBackend.make_state(n, [], [], statetype=numpy.complex64)
    _apply_python_plan(ref, n, blocks)
    assert numpy.array_equal(got, ref.get_state())
    if width == 2:                        # a 3-qubit gate is wider than the width: applied unfused
        U = rand_u(8)
        pl = GpuBackend.make_state(n, [], [])
        for d in ops:
            pl.kronselect_dot(d)
        for be in (pl, fz):
            be.kronselect_dot({(1, 5, 9): U})
            be.kronselect_dot({2: numpy.array([[0, 1], [1, 0]])})
        assert infidelity(pl, fz) < 1e-8


def test_one_gate_fast_path_equals_planner():
    """A one-gate descriptor with a numpy matrix skips the planner; the same circuit written with list
    matrices (which go through the planner) records the same gates: bitwise equal states. A bad
    descriptor still raises as the planner does."""
    from qip_amd.fused import FusedBackend
    from qip_amd.workloads import layered_random_circuit
    n = 13
    ops = layered_random_circuit(n, 6, seed=21)
    a = FusedBackend.make_state(n, [], [])
    b = FusedBackend.make_state(n, [], [])
    for d in ops:
        a.kronselect_dot(d)
        b.kronselect_dot({k: numpy.asarray(v).tolist() for k, v in d.items()})
    assert numpy.array_equal(a.get_state(), b.get_state())
    with pytest.raises(Exception):
        a.kronselect_dot({(0, 1): numpy.eye(2)})          # shape does not match the qubit count
    with pytest.raises(ValueError):
        a.kronselect_dot({n: numpy.eye(2)})               # qubit out of range
