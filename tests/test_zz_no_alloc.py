"""qipgpu_probabilities_noalloc (no device allocation) against qipgpu_probabilities, through the C ABI
(marked gpu): the same outcome distribution to FP64 rounding on two calls in a row, the second on a
smaller state with fewer measured qubits, then at the size limit, on a complex64 state and on a
precision-form state with a pending renormalisation; the documented refusals."""
import ctypes

import numpy
import pytest

from qip_amd import _capi

pytestmark = pytest.mark.gpu
RNG = numpy.random.default_rng(77)


def make_state(n, dtype_code):
    lib = _capi.load()
    h = ctypes.c_void_p()
    _capi.check(lib.qipgpu_create(n, dtype_code, ctypes.byref(h)), "create")
    v = RNG.standard_normal(2 ** n) + 1j * RNG.standard_normal(2 ** n)
    v /= numpy.linalg.norm(v)
    host = numpy.ascontiguousarray(v, dtype=numpy.complex128 if dtype_code == _capi.C128 else numpy.complex64)
    _capi.check(lib.qipgpu_set_range(h, 0, 2 ** n, _capi.ptr(host)), "set_range")
    return h


def both(h, idx):
    """(qipgpu_probabilities, qipgpu_probabilities_noalloc) on the same state."""
    lib = _capi.load()
    q = _capi.i32(idx)
    want = numpy.full(2 ** len(q), numpy.nan)
    got = numpy.full(2 ** len(q), numpy.nan)
    _capi.check(lib.qipgpu_probabilities(h, len(q), _capi.ptr(q), _capi.ptr(want)), "probabilities")
    _capi.check(lib.qipgpu_probabilities_noalloc(h, len(q), _capi.ptr(q), _capi.ptr(got)), "probabilities_noalloc")
    return want, got


def test_two_calls_in_a_row_then_smaller():
    lib = _capi.load()
    for n, k in ((22, 16), (14, 6), (24, 20), (11, 3)):
        print("no-alloc probabilities n", n, "k", k, flush=True)
        h = make_state(n, _capi.C128)
        idx = [int(x) for x in RNG.permutation(n)[:k]]
        want, got = both(h, idx)
        assert not numpy.isnan(got).any()
        assert numpy.max(numpy.abs(got - want)) < 1e-14
        assert abs(got.sum() - 1.0) < 1e-12
        lib.qipgpu_destroy(h)


def test_complex64_state():
    lib = _capi.load()
    h = make_state(18, _capi.C64)
    idx = [17, 0, 9, 4, 12, 3, 8]
    want, got = both(h, idx)
    assert numpy.max(numpy.abs(got - want)) < 1e-14
    lib.qipgpu_destroy(h)


def test_precision_form_state_and_refusals():
    from qip_amd.fused import FusedBackend
    from qip_amd.workloads import layered_random_circuit
    lib = _capi.load()
    n = 16
    fz = FusedBackend.make_state(n, [], [])
    fz.streaming = False
    for d in layered_random_circuit(n, 2, seed=4):
        fz.kronselect_dot(d)
    q = _capi.i32([0, 5, 10, 15])
    out = numpy.zeros(16)
    assert fz.pending_gates() > 0
    with pytest.raises(ValueError):        # recorded gates not launched yet
        _capi.check(lib.qipgpu_probabilities_noalloc(fz.handle, len(q), _capi.ptr(q), _capi.ptr(out)), "noalloc")
    fz.flush()                              # blocks launched; the last block's renormalisation pends
    got = numpy.zeros(16)
    _capi.check(lib.qipgpu_probabilities_noalloc(fz.handle, len(q), _capi.ptr(q), _capi.ptr(got)), "noalloc")
    want = numpy.zeros(16)
    _capi.check(lib.qipgpu_probabilities(fz.handle, len(q), _capi.ptr(q), _capi.ptr(want)), "probabilities")
    assert numpy.max(numpy.abs(got - want)) < 1e-14
    assert abs(got.sum() - 1.0) < 1e-6
    fz.close()
    h = make_state(22, _capi.C128)
    big = _capi.i32(list(range(21)))
    out = numpy.zeros(2 ** 21)
    with pytest.raises(ValueError):        # past the fixed array's limit
        _capi.check(lib.qipgpu_probabilities_noalloc(h, len(big), _capi.ptr(big), _capi.ptr(out)), "noalloc")
    lib.qipgpu_destroy(h)
