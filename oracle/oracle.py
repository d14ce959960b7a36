"""CPU oracle for the QIP state-vector hot path.

TEST INFRASTRUCTURE ONLY: this module is the checker the HIP engine is compared against, and
its C library (built a second time with the reference's own compiler flags) is the CPU baseline
timed beside the GPU. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import it; nothing under qip_amd/ does.

Layers:
  * ``_lib`` / ``_lib_fast``: ctypes handles on ``oracle/lib/liboracle.so`` (checker build,
    -O2, no fast-math) and ``oracle/lib/liboracle_fast.so`` (baseline build with the flags of
    the reference's setup.py:42-43: -O3 -ffast-math -fopenmp, portable -march).
  * Python restatements of the marshalling the reference does in Python around its kernels:
    qip/util.py:15-67 (kronselect_dot validation and packing), qip/ext/kronprod.pyx:54-132
    (MatStruct building), qip/ext/func_apply.pyx:26-69 (checks and f table),
    qip/backend.py:73-104 + qip/util.py:108-124 (initial product state).
  * ``OracleBackend``: a StateType whose every method runs the restated reference kernel,
    mirroring CythonBackend (qip/backend.py:68-175) method for method.

Pinning (tests/test_oracle.py): the reference's own known-answer tests (tests/utiltest.py,
tests/qiptest.py, tests/qfttest.py), brute force against independently built dense numpy
operators (tensor contractions, numpy.kron, numpy.block), the QFT closed form
ifft(norm="ortho"), numpy marginals, and the Grover closed form.
"""
import ctypes
import os
import random
from typing import Callable, Mapping, Optional, Sequence

import numpy

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

OR_OTHER, OR_NUMPY, OR_CMAT, OR_SWAP = 0, 1, 2, 3   # qip/ext/kronprod.pyx:18-19


class _OrMat(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("param", ctypes.c_int32), ("child", ctypes.c_int32),
                ("pad_", ctypes.c_int32), ("data", ctypes.c_void_p)]


_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_D = ctypes.c_double


def _load(name: str):
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise OSError("oracle library {} is not built; run __graft_entry__.build()".format(path))
    lib = ctypes.CDLL(path)
    lib.or_entwine_bit.restype = _I64
    lib.or_entwine_bit.argtypes = [ctypes.c_int, ctypes.c_int, _I64, _I64, _I64]
    lib.or_cdot.restype = ctypes.c_int
    lib.or_cdot.argtypes = [_I32, _P, _P, _P, _P, _I64, _I32, _P, _I64, _I64, _I64, ctypes.c_int]
    lib.or_func_apply.restype = ctypes.c_int
    lib.or_func_apply.argtypes = [_I32, _P, _I32, _P, _I32, _P, _P, _P, _I32, _P, ctypes.c_int,
                                  ctypes.c_int]
    lib.or_prob_magnitude.restype = _D
    lib.or_prob_magnitude.argtypes = [_P, _I64]
    lib.or_measure_probabilities.restype = None
    lib.or_measure_probabilities.argtypes = [_I32, _P, _I32, _P, _P]
    lib.or_measure_top_probabilities.restype = _I32
    lib.or_measure_top_probabilities.argtypes = [_I32, _P, _I32, _I32, _P, ctypes.c_int, _P, _P]
    lib.or_soft_measure.restype = None
    lib.or_soft_measure.argtypes = [_I32, _P, _I32, _P, _I64, _I64, _I64, _D, _P, _P]
    lib.or_collapse.restype = None
    lib.or_collapse.argtypes = [_I32, _P, _I32, _P, _P, _I64, _D]
    lib.or_reduce.restype = None
    lib.or_reduce.argtypes = [_I32, _P, _I32, _P, _P, _I64, _D]
    return lib


_lib_cache = {}


def lib(fast: bool = False):
    """The checker build (default) or the reference-flags baseline build (fast=True)."""
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name not in _lib_cache:
        _lib_cache[name] = _load(name)
    return _lib_cache[name]


def _ptr(a: numpy.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def entwine_bit(n_indices: int, n_non_indices: int, v_indices: int, v_non_indices: int, mask: int) -> int:
    """qip/ext/util.pyx:2-23."""
    return int(lib().or_entwine_bit(n_indices, n_non_indices, v_indices, v_non_indices, mask))


# ---------------------------------------------------------------------------------------------
# kronselect_dot: qip/util.py:15-67 (validation + packing) -> cdot_loop (qip/ext/kronprod.pyx)
# ---------------------------------------------------------------------------------------------

def _mat_structs(matrices: Sequence[object]):
    """qip/ext/kronprod.pyx:54-132: one struct per top-level matrix, nested CMat children
    appended after them and linked by index. Returns (struct array, keep-alive list)."""
    top = []
    extra = []
    keep = []

    def make(mat):
        struct_val = OR_OTHER
        if hasattr(mat, "_kron_struct"):
            struct_val = mat._kron_struct
        elif type(mat) == numpy.ndarray:
            struct_val = OR_NUMPY
        s = _OrMat()
        s.type = struct_val
        s.child = -1
        if struct_val == OR_NUMPY:
            arr = numpy.ascontiguousarray(mat, numpy.complex128)
            keep.append(arr)
            s.data = arr.ctypes.data
            s.param = mat.shape[0]
        elif struct_val == OR_SWAP:
            s.param = mat.n
        elif struct_val == OR_CMAT:
            s.param = mat.shape[0]
        else:
            raise ValueError("Cannot pass matrices which are not numpy, SwapMat, or CMat")
        return s, struct_val

    for mat in matrices:
        s, sv = make(mat)
        top.append(s)
        parent = ("top", len(top) - 1)
        while sv == OR_CMAT:
            mat = mat.m
            child, sv = make(mat)
            extra.append(child)
            idx = len(extra) - 1
            if parent[0] == "top":
                top[parent[1]].child = idx  # fixed up below (offset by len(top))
            else:
                extra[parent[1]].child = idx
            parent = ("extra", idx)
    ntop = len(top)
    for s in top:
        if s.child >= 0:
            s.child += ntop
    for s in extra:
        if s.child >= 0:
            s.child += ntop
    arr = (_OrMat * max(1, ntop + len(extra)))(*(top + extra))
    return arr, keep


def normalise_mats(mats: Mapping[object, object]):
    """qip/util.py:35-58: keys int|tuple, list matrices to numpy, square 2^m shape check."""
    newmats = {}
    for indices in mats:
        if type(indices) != tuple and type(indices) != int:
            raise Exception("Type of indices must be tuple: {}".format(indices))
        m = mats[indices]
        if type(m) == list:
            m = numpy.array(m)
        if type(indices) == tuple:
            key, nindices = indices, len(indices)
        else:
            key, nindices = (indices,), 1
        newmats[key] = m
        if 2 ** nindices != m.shape[0] or 2 ** nindices != m.shape[1]:
            raise Exception("Shape of square submatrix must equal 2**(number of indices): "
                            "{}: {}".format(indices, m))
    return newmats


def kronselect_dot(mats, vec: numpy.ndarray, n: int, outputarray: numpy.ndarray,
                   input_offset: int = 0, output_offset: int = 0, fast: bool = False,
                   nthreads: int = 0) -> None:
    """qip/util.py:15-67 with dot_impl=cdot_loop (qip/ext/kronprod.pyx:41-200)."""
    if len(vec) + input_offset > 2 ** n:
        raise ValueError("Input vector size plus offset may be no larger than the total number of qubit states (2^n)")
    if len(outputarray) + output_offset > 2 ** n:
        raise ValueError("Output vector size plus offset may be no larger than the total number of qubit states (2^n)")
    newmats = normalise_mats(mats)
    keys = list(newmats.keys())
    group_len = numpy.array([len(k) for k in keys], dtype=numpy.int32)
    group_idx = numpy.array([i for k in keys for i in k], dtype=numpy.int32)
    cmats = [newmats[k] if type(newmats[k]) != numpy.ndarray else newmats[k].astype(numpy.complex128)
             for k in keys]
    structs, keep = _mat_structs(cmats)
    vec_c = numpy.ascontiguousarray(vec, dtype=numpy.complex128)
    if outputarray.dtype != numpy.complex128 or not outputarray.flags.c_contiguous:
        raise ValueError("output must be a contiguous complex128 array")
    rc = lib(fast).or_cdot(len(keys), _ptr(group_len), _ptr(group_idx), ctypes.cast(structs, ctypes.c_void_p),
                           _ptr(vec_c), len(vec_c), n, _ptr(outputarray), len(outputarray),
                           input_offset, output_offset, nthreads)
    if rc != 0:
        raise MemoryError("or_cdot failed")


# ---------------------------------------------------------------------------------------------
# func_apply: qip/ext/func_apply.pyx:11-112
# ---------------------------------------------------------------------------------------------

def func_table(func: Callable[[int], int], n_reg1: int) -> numpy.ndarray:
    """func_apply.pyx:66-69: f tabulated into C ints (OverflowError past int32, as Cython). An
    integer array is taken as the table itself."""
    if isinstance(func, numpy.ndarray):
        return numpy.ascontiguousarray(func, dtype=numpy.int32)
    if isinstance(getattr(func, "values", None), numpy.ndarray):
        return numpy.ascontiguousarray(func.values, dtype=numpy.int32)
    table = numpy.zeros(2 ** n_reg1, dtype=numpy.int32)
    for x in range(2 ** n_reg1):
        v = int(func(x))
        if not (-2 ** 31 <= v < 2 ** 31):
            raise OverflowError("value too large to convert to int")
        table[x] = v
    return table


def func_apply(reg1_indices, reg2_indices, func, vec: numpy.ndarray, n: int, output: numpy.ndarray,
               faithful: bool = True, table: Optional[numpy.ndarray] = None, fast: bool = False,
               nthreads: int = 0) -> None:
    """func_apply.pyx:11-112. `table` may be passed pre-tabulated (then `func` is unused)."""
    if vec.shape[0] != output.shape[0] or 2 ** n != vec.shape[0]:
        raise ValueError("Both input and output vectors must be of size 2**n")
    reg1 = numpy.ascontiguousarray(reg1_indices, dtype=numpy.int32)
    reg2 = numpy.ascontiguousarray(reg2_indices, dtype=numpy.int32)
    other = numpy.array([i for i in numpy.arange(0, n) if (i not in reg1 and i not in reg2)],
                        dtype=numpy.int32)
    if table is None:
        table = func_table(func, len(reg1))
    table = numpy.ascontiguousarray(table, dtype=numpy.int32)
    vec_c = numpy.ascontiguousarray(vec, dtype=numpy.complex128)
    lib(fast).or_func_apply(len(reg1), _ptr(reg1), len(reg2), _ptr(reg2), len(other),
                            _ptr(other if len(other) else numpy.zeros(1, numpy.int32)), _ptr(table),
                            _ptr(vec_c), n, _ptr(output), 1 if faithful else 0, nthreads)


# ---------------------------------------------------------------------------------------------
# measurement family: qip/ext/kronprod.pyx:238-490
# ---------------------------------------------------------------------------------------------

def prob_magnitude(vec: numpy.ndarray) -> float:
    v = numpy.ascontiguousarray(vec, dtype=numpy.complex128)
    return float(lib().or_prob_magnitude(_ptr(v), len(v)))


def measure_probabilities(indices, n: int, vec: numpy.ndarray) -> numpy.ndarray:
    idx = numpy.ascontiguousarray(indices, dtype=numpy.int32)
    v = numpy.ascontiguousarray(vec, dtype=numpy.complex128)
    out = numpy.zeros(2 ** len(idx), dtype=numpy.float64)
    lib().or_measure_probabilities(len(idx), _ptr(idx), n, _ptr(v), _ptr(out))
    return out


def measure_top_probabilities(indices, n: int, top_k: int, vec: numpy.ndarray, faithful: bool = False):
    idx = numpy.ascontiguousarray(indices, dtype=numpy.int32)
    v = numpy.ascontiguousarray(vec, dtype=numpy.complex128)
    kk = min(top_k, 2 ** len(idx))
    ms = numpy.zeros(max(kk, 1), dtype=numpy.int64)
    ps = numpy.zeros(max(kk, 1), dtype=numpy.float64)
    got = lib().or_measure_top_probabilities(len(idx), _ptr(idx), n, top_k, _ptr(v), 1 if faithful else 0,
                                             _ptr(ms), _ptr(ps))
    return [int(m) for m in ms[:got]], [float(p) for p in ps[:got]]


def soft_measure(indices, n: int, vec: numpy.ndarray, measured: Optional[int] = None,
                 input_offset: int = 0, u: Optional[float] = None):
    """kronprod.pyx:331-383; draws u = random.random() when not given (kronprod.pyx:362)."""
    if u is None:
        u = random.random()
    idx = numpy.ascontiguousarray(indices, dtype=numpy.int32)
    v = numpy.ascontiguousarray(vec, dtype=numpy.complex128)
    m = ctypes.c_int64()
    p = ctypes.c_double()
    lib().or_soft_measure(len(idx), _ptr(idx), n, _ptr(v), len(v), -1 if measured is None else measured,
                          input_offset, u, ctypes.byref(m), ctypes.byref(p))
    return int(m.value), float(p.value)


def _check_measured(k, measured, measured_prob):
    if measured is not None:
        if not (0 <= measured < 2 ** k):
            raise ValueError("Measured value must be less than 2**len(indices)")
    if measured_prob is not None:
        if not 0.0 < measured_prob <= 1.0:
            raise ValueError("measured_prob must be 0 < p <= 1")


def measure(indices, n: int, vec: numpy.ndarray, out: numpy.ndarray, measured=None, measured_prob=None,
            u: Optional[float] = None):
    """kronprod.pyx:388-446 (offsets 0)."""
    idx = numpy.ascontiguousarray(indices, dtype=numpy.int32)
    _check_measured(len(idx), measured, measured_prob)
    if measured is None or measured_prob is None:
        m, p = soft_measure(idx, n, vec, measured=measured, u=u)
    else:
        m, p = measured, measured_prob
    mult_p = numpy.sqrt(1.0 / p)
    v = numpy.ascontiguousarray(vec, dtype=numpy.complex128)
    lib().or_collapse(len(idx), _ptr(idx), n, _ptr(v), _ptr(out), m, float(mult_p))
    return m, p


def reduce_measure(indices, n: int, vec: numpy.ndarray, out: numpy.ndarray, measured=None,
                   measured_prob=None, u: Optional[float] = None):
    """kronprod.pyx:451-490 (offsets 0): output is the reduced state tiled over the 2^n arena."""
    idx = numpy.ascontiguousarray(indices, dtype=numpy.int32)
    _check_measured(len(idx), measured, measured_prob)
    if measured is None or measured_prob is None:
        m, p = soft_measure(idx, n, vec, measured=measured, u=u)
    else:
        m, p = measured, measured_prob
    mult_p = numpy.sqrt(1.0 / p)
    v = numpy.ascontiguousarray(vec, dtype=numpy.complex128)
    lib().or_reduce(len(idx), _ptr(idx), n, _ptr(v), _ptr(out), m, float(mult_p))
    return m, p


# ---------------------------------------------------------------------------------------------
# initial state: qip/backend.py:73-104 with qip/util.py:108-124 (gen_edit_indices)
# ---------------------------------------------------------------------------------------------

def make_state_vector(n: int, index_groups, feed_list, statetype=numpy.complex128) -> numpy.ndarray:
    """Tensor product of the feeds over their index groups, unfed qubits |0>, |0..0> when there is
    no feed. Enumerates the 2^K fed-bit combinations like gen_edit_indices (util.py:113-124), with
    numpy doing each step over all combinations at once; the product is taken group by group in
    feed order starting from 1.0, as backend.py:99-101 does."""
    state = numpy.zeros(2 ** n, dtype=statetype)
    if len(feed_list) == 0:
        state[0] = 1
        return state
    allindices = [i for g in index_groups for i in g]
    K = len(allindices)
    combos = numpy.arange(2 ** K, dtype=numpy.int64)
    index = numpy.zeros(2 ** K, dtype=numpy.int64)
    for j, q in enumerate(allindices):
        bit = (combos >> (K - 1 - j)) & 1            # flips[j], MSB-first (util.py:114)
        index |= bit << (n - 1 - q)                  # bits[allindices[j]] (util.py:116)
    vals = numpy.ones(2 ** K, dtype=statetype)
    pos = 0
    for g, grp in enumerate(index_groups):
        sub = numpy.zeros(2 ** K, dtype=numpy.int64)
        for j in range(len(grp)):
            sub = (sub << 1) | ((combos >> (K - 1 - (pos + j))) & 1)
        feed = numpy.asarray(feed_list[g])
        vals = (vals * feed[sub]).astype(statetype)
        pos += len(grp)
    state[index] = vals
    return state
