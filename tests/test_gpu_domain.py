"""cdot_loop's whole operator domain on the GPU (marked gpu), against the oracle's restatement of
qip/ext/kronprod.pyx:41-235 on the same inputs.

* controls shared by every group of a descriptor (C(Op) on a register, qip/operators.py:221-227),
  peeled off and applied as controlled sweeps: C(Not) on 12 qubits, C(C(H)) on 8;
* dense operators on 4..6 targets (LDS tiles) and 7..12 targets (FP64 GEMM), with and without
  controls, in both precisions, including states smaller than one tile;
* diagonal operators on more than 6 qubits;
* components that nothing specialised covers (groups overlapping over more than 10 qubits, a dense
  matrix on 13 qubits, C_MAT chains and SWAP_MAT inside), through the engine's restatement of the
  reference's column sum.
Bars: complex128 within 1e-12 of the oracle, complex64 within 2e-6 relative to the largest
amplitude (FP64 arithmetic, one rounding per amplitude per call).
"""
import numpy
import pytest

from oracle import oracle
from qip_amd import C, CMat, GpuBackend, H, MatrixOp, Not, Qubit, SwapMat, planner, run
from oracle.oracle_backend import OracleBackend

pytestmark = pytest.mark.gpu
RNG = numpy.random.default_rng(4242)


def rand_state(n):
    v = RNG.standard_normal(2 ** n) + 1j * RNG.standard_normal(2 ** n)
    return v / numpy.linalg.norm(v)


def rand_u(d):
    z = (RNG.standard_normal((d, d)) + 1j * RNG.standard_normal((d, d))) / numpy.sqrt(2)
    if d >= 1024:                  # a QR of this size costs seconds: a scaled Gaussian operator is as good a test
        return z / numpy.sqrt(d)
    q, r = numpy.linalg.qr(z)
    return q * (numpy.diagonal(r) / numpy.abs(numpy.diagonal(r)))


def oracle_apply(mats, psi, n):
    out = numpy.zeros_like(psi)
    oracle.kronselect_dot(mats, psi, n, out, fast=True, nthreads=8)
    return out


def check(mats, n, dtype=numpy.complex128, kinds=None):
    psi = rand_state(n)
    if kinds is not None:
        got_kinds = sorted(type(c).__name__ for c in planner.plan(mats))
        assert got_kinds == sorted(kinds), got_kinds
    be = GpuBackend.make_state(n, [list(range(n))], [psi], statetype=dtype)
    be.kronselect_dot(mats)
    got = be.get_state().astype(numpy.complex128)
    be.close()
    want = oracle_apply(mats, psi, n)
    err = float(numpy.max(numpy.abs(got - want)))
    bar = 1e-12 if dtype == numpy.complex128 else 2e-6 * float(numpy.max(numpy.abs(want)))
    assert err < bar, (err, bar)


X = numpy.array([[0.0, 1.0], [1.0, 0.0]])
HM = numpy.array([[1.0, 1.0], [1.0, -1.0]]) / numpy.sqrt(2.0)


@pytest.mark.parametrize("dtype", [numpy.complex128, numpy.complex64])
def test_controlled_not_on_twelve_qubits(dtype):
    print("C(Not) on 12 qubits", numpy.dtype(dtype).name, flush=True)
    mats = {(0, i): CMat(X) for i in range(1, 13)}
    check(mats, 14, dtype, kinds=["MatrixCall"] * 4)


@pytest.mark.parametrize("dtype", [numpy.complex128, numpy.complex64])
def test_doubly_controlled_h_on_eight_qubits(dtype):
    print("C(C(H)) on 8 qubits", numpy.dtype(dtype).name, flush=True)
    mats = {(3, 0, q): CMat(CMat(HM)) for q in (1, 2, 4, 5, 6, 7, 8, 9)}
    check(mats, 11, dtype, kinds=["MatrixCall"] * 3)


def test_graph_controlled_ops_on_registers():
    """The same operators built by the graph frontend, GPU against the oracle backend."""
    print("graph C(Not), C(C(H))", flush=True)
    psi_r, psi_q = rand_state(12), rand_state(8)
    results = []
    for bc in (GpuBackend.make_state, OracleBackend.make_state):
        c, r = Qubit(n=1), Qubit(n=12)
        a, b = C(Not)(c, r)
        c2, c3, q = Qubit(n=1), Qubit(n=1), Qubit(n=8)
        x, y, z = C(C(H))(c2, c3, q)
        st, _ = run(a, b, x, y, z, feed={c: [0.6, 0.8], r: psi_r, c2: [0.0, 1.0], c3: [0.8, 0.6], q: psi_q},
                    backend_constructor=bc)
        results.append(numpy.asarray(st))
    assert numpy.max(numpy.abs(results[0] - results[1])) < 1e-12


@pytest.mark.parametrize("k", [4, 5, 6])
@pytest.mark.parametrize("n,controls", [(6, 0), (9, 1), (16, 0), (17, 2)])
@pytest.mark.parametrize("dtype", [numpy.complex128, numpy.complex64])
def test_dense_tiles(k, n, controls, dtype):
    if k + controls > n:
        pytest.skip("operator wider than the state")
    qs = [int(q) for q in RNG.permutation(n)[:k + controls]]
    print("dense tile k", k, "n", n, "controls", qs[:controls], "targets", qs[controls:], flush=True)
    m = rand_u(2 ** k)
    for _ in range(controls):
        m = CMat(m)
    check({tuple(qs): m}, n, dtype)


@pytest.mark.parametrize("k", [4, 5, 6])
@pytest.mark.parametrize("where", ["low", "high", "low_controlled"])
@pytest.mark.parametrize("dtype", [numpy.complex128, numpy.complex64])
def test_dense_tiles_both_layouts(k, where, dtype):
    """4 to 6 targets (the FP64 MFMA tiles) with state bit 0 among the targets (lanes along the
    operator's rows) and without it (lanes along the bases), over enough tiles that every workgroup
    takes several; once with a control on the highest qubit."""
    n = 15
    if where == "high":
        qs = list(range(k))                                   # qubits 0.. = the highest state bits
    else:
        qs = list(range(n - k, n))[::-1]                      # state bits 0..k-1, in scrambled order
        qs = qs[1::2] + qs[0::2]
    if where == "low_controlled":
        qs = [0] + qs
    print("dense tile layout", where, "k", k, "qubits", qs, flush=True)
    m = rand_u(2 ** k)
    if where == "low_controlled":
        m = CMat(m)
    check({tuple(qs): m}, n, dtype)


@pytest.mark.parametrize("k,n,controls", [(7, 7, 0), (7, 12, 1), (8, 12, 0), (8, 14, 2), (10, 13, 0), (12, 14, 0)])
@pytest.mark.parametrize("dtype", [numpy.complex128, numpy.complex64])
def test_dense_gemm(k, n, controls, dtype):
    qs = [int(q) for q in RNG.permutation(n)[:k + controls]]
    print("dense gemm k", k, "n", n, "qubits", qs, flush=True)
    m = rand_u(2 ** k)
    for _ in range(controls):
        m = CMat(m)
    check({tuple(qs): m}, n, dtype)


@pytest.mark.parametrize("dtype", [numpy.complex128, numpy.complex64])
def test_dense_gemm_calls_in_a_row(dtype):
    """Several wide operators on one state: the operator staging and the kept temporary are reused,
    grown (8 -> 10 targets) and reused smaller, each call against the oracle applied in turn."""
    n = 13
    psi = RNG.standard_normal(2 ** n) + 1j * RNG.standard_normal(2 ** n)
    psi /= numpy.linalg.norm(psi)
    be = GpuBackend.make_state(n, [list(range(n))], [psi], statetype=dtype)
    ref = be.get_state().astype(numpy.complex128)
    tol = 1e-12 if dtype == numpy.complex128 else 2e-6
    for k, controls in ((8, 0), (7, 0), (10, 0), (8, 1), (9, 0)):
        qs = [int(q) for q in RNG.permutation(n)[:k + controls]]
        print("gemm in a row: k", k, "controls", controls, "qubits", qs, flush=True)
        m = rand_u(2 ** k)
        for _ in range(controls):
            m = CMat(m)
        be.kronselect_dot({tuple(qs): m})
        out = numpy.zeros_like(ref)
        oracle.kronselect_dot({tuple(qs): m}, ref, n, out)
        ref = out
        assert numpy.max(numpy.abs(be.get_state() - ref)) < tol
    be.close()


def test_dense_eight_qubit_matrix_op():
    """A user MatrixOp on 8 qubits (a random unitary), through the graph frontend."""
    U = rand_u(256)

    class Haar8(MatrixOp):
        def makemats(self, index_groups):
            return {tuple(q for g in index_groups for q in g): U}

    psi = rand_state(11)
    outs = []
    for bc in (GpuBackend.make_state, OracleBackend.make_state):
        a, b = Qubit(n=3), Qubit(n=8)
        h = Haar8(b)
        st, _ = run(a, h, feed={(a, b): psi}, backend_constructor=bc)
        outs.append(numpy.asarray(st))
    assert numpy.max(numpy.abs(outs[0] - outs[1])) < 1e-12


@pytest.mark.parametrize("dtype", [numpy.complex128, numpy.complex64])
def test_diagonal_wide(dtype):
    n, k = 13, 9
    qs = [int(q) for q in RNG.permutation(n)[:k + 1]]
    d = numpy.exp(1j * RNG.uniform(0, 2 * numpy.pi, 2 ** k))
    psi = rand_state(n)
    be = GpuBackend.make_state(n, [list(range(n))], [psi], statetype=dtype)
    from qip_amd import _capi
    t, c = _capi.i32(qs[1:]), _capi.i32(qs[:1])
    dd = numpy.ascontiguousarray(d, dtype=numpy.complex128)
    _capi.check(be._lib.qipgpu_apply_diagonal(be.handle, k, _capi.ptr(t), _capi.ptr(dd.view(numpy.float64)), 1,
                                              _capi.ptr(c)), "diag")
    got = be.get_state().astype(numpy.complex128)
    be.close()
    want = oracle_apply({tuple(qs): CMat(numpy.diag(d))}, psi, n)
    bar = 1e-12 if dtype == numpy.complex128 else 2e-6 * float(numpy.max(numpy.abs(want)))
    assert numpy.max(numpy.abs(got - want)) < bar


@pytest.mark.parametrize("dtype", [numpy.complex128, numpy.complex64])
def test_general_overlapping_chain(dtype):
    """Two-qubit groups overlapping in a chain over 13 qubits: one component, no shared control."""
    mats = {(i, i + 1): rand_u(4) for i in range(0, 12, 1)}
    print("general chain", flush=True)
    check(mats, 14, dtype, kinds=["GeneralCall"])


def test_general_mixed_structures():
    """C_MAT chains, SWAP_MAT and a 7-qubit dense matrix in one component over 13 qubits."""
    mats = {(0, 1, 2, 3, 4): CMat(SwapMat(2)), (4, 5, 6, 7, 8, 9, 10): rand_u(128),
            (10, 11, 12): CMat(CMat(HM))}
    print("general mixed", flush=True)
    check(mats, 13, numpy.complex128, kinds=["GeneralCall"])


def test_general_thirteen_qubit_dense():
    print("general dense 13", flush=True)
    m = rand_u(2 ** 13)
    check({tuple(range(13)): m}, 13, numpy.complex128, kinds=["GeneralCall"])


def test_general_call_on_fused_backend():
    """FusedBackend applies a component it cannot fuse through the plain engine, in program order."""
    from qip_amd.fused import FusedBackend
    from qip_amd.workloads import layered_random_circuit
    n = 14
    ops = layered_random_circuit(n, 2, seed=3)
    chain = {(i, i + 1): rand_u(4) for i in range(12)}
    ops = ops[:20] + [chain] + ops[20:]
    fz = FusedBackend.make_state(n, [], [])
    pl = GpuBackend.make_state(n, [], [])
    for d in ops:
        fz.kronselect_dot(d)
        pl.kronselect_dot(d)
    from qip_amd import infidelity
    assert infidelity(pl, fz) < 1e-8
    fz.close()
    pl.close()
