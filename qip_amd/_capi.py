"""ctypes binding of libqipgpu.so (declared in include/qipgpu.h).

The library is built in-tree by ``__graft_entry__.build()`` into ``qip_amd/lib/``. There is no
CPU fallback: if the library or a GPU is missing, every backend call raises.
"""
import ctypes
import struct
import os

import numpy

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libqipgpu.so")

C128, C64 = 0, 1
FUNC_REFERENCE_BOUND = 1   # QIPGPU_FUNC_REFERENCE_BOUND
OK, EINVAL, ENOMEM, EHIP, EZERODIV, EUNSUP = 0, -1, -2, -3, -4, -5

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_D = ctypes.c_double
_PP = ctypes.POINTER(ctypes.c_void_p)
_PI32 = ctypes.POINTER(ctypes.c_int32)
_PI64 = ctypes.POINTER(ctypes.c_int64)
_PD = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); mirrors include/qipgpu.h
SIGNATURES = {
    "qipgpu_last_error": (ctypes.c_char_p, []),
    "qipgpu_device_count": (ctypes.c_int, [_PI32]),
    "qipgpu_version": (ctypes.c_char_p, []),
    "qipgpu_create": (ctypes.c_int, [_I32, _I32, _PP]),
    "qipgpu_destroy": (ctypes.c_int, [_P]),
    "qipgpu_info": (ctypes.c_int, [_P, _PI32, _PI32]),
    "qipgpu_init_product": (ctypes.c_int, [_P, _I32, _P, _P, _P]),
    "qipgpu_get_range": (ctypes.c_int, [_P, _I64, _I64, _P]),
    "qipgpu_set_range": (ctypes.c_int, [_P, _I64, _I64, _P]),
    "qipgpu_add_range": (ctypes.c_int, [_P, _I64, _I64, _P]),
    "qipgpu_apply_matrix": (ctypes.c_int, [_P, _I32, _P, _P, _I32, _P]),
    "qipgpu_apply_product": (ctypes.c_int, [_P, _I32, _P, _P]),
    "qipgpu_apply_diagonal": (ctypes.c_int, [_P, _I32, _P, _P, _I32, _P]),
    "qipgpu_apply_kron": (ctypes.c_int, [_P, _I32, _P, _P, _P, _P, _P, _P]),
    "qipgpu_func_apply": (ctypes.c_int, [_P, _I32, _P, _I32, _P, _P]),
    "qipgpu_table_create": (ctypes.c_int, [_I32, _P, _PP]),
    "qipgpu_table_destroy": (ctypes.c_int, [_P]),
    "qipgpu_func_apply_table": (ctypes.c_int, [_P, _I32, _P, _I32, _P, _P]),
    "qipgpu_func_apply_table_ex": (ctypes.c_int, [_P, _I32, _P, _I32, _P, _P, _I32]),
    "qipgpu_total_prob": (ctypes.c_int, [_P, _PD]),
    "qipgpu_probabilities": (ctypes.c_int, [_P, _I32, _P, _P]),
    "qipgpu_probabilities_noalloc": (ctypes.c_int, [_P, _I32, _P, _P]),
    "qipgpu_marginals": (ctypes.c_int, [_P, _I32, _P, _P]),
    "qipgpu_top_probabilities": (ctypes.c_int, [_P, _I32, _P, _I32, _P, _P, _PI32]),
    "qipgpu_soft_measure": (ctypes.c_int, [_P, _I32, _P, _D, _I64, _PI64, _PD]),
    "qipgpu_measure": (ctypes.c_int, [_P, _I32, _P, _D, _I64, _D, _PI64, _PD]),
    "qipgpu_reduce_measure": (ctypes.c_int, [_P, _I32, _P, _D, _I64, _D, _PI64, _PD]),
    "qipgpu_reduce_measure_shrink": (ctypes.c_int, [_P, _I32, _P, _D, _I64, _D, _PI64, _PD]),
    "qipgpu_fused_apply": (ctypes.c_int, [_P, _I32, _P, _P]),
    "qipgpu_fused_pattern_sweep": (ctypes.c_int, [_P, _I32, _P]),
    "qipgpu_compose": (ctypes.c_int, [_I32, _P, _I32, _P, _P, _P, _P]),
    "qipgpu_fused_record": (ctypes.c_int, [_P, _I32, _P, _P, _P]),
    "qipgpu_fused_set_recording": (ctypes.c_int, [_P, _I32, _I32]),
    "qipgpu_fused_flush": (ctypes.c_int, [_P]),
    "qipgpu_fused_recorded": (ctypes.c_int, [_P, _PI64, _PI64]),
    "qipgpu_fuse_online": (ctypes.c_int, [_I32, _I32, _I32, _P, _P, _P, _P]),
    "qipgpu_fuse_launches": (ctypes.c_int, [_I32, _I32, _I32, _I32, _P, _P, _P, _PI64]),
    "qipgpu_fused_finish": (ctypes.c_int, [_P]),
    "qipgpu_fused_last_norm": (ctypes.c_int, [_P, _PD]),
    "qipgpu_overlap": (ctypes.c_int, [_P, _P, _PD, _PD, _PD, _PD]),
    "qipgpu_max_abs_diff": (ctypes.c_int, [_P, _P, _PD]),
    "qipgpu_synchronize": (ctypes.c_int, [_P]),
    "qipgpu_capture_begin": (ctypes.c_int, [_P]),
    "qipgpu_capture_end": (ctypes.c_int, [_P, _PP]),
    "qipgpu_graph_matches": (ctypes.c_int, [_P, _P]),
    "qipgpu_graph_launch": (ctypes.c_int, [_P, _P]),
    "qipgpu_graph_destroy": (ctypes.c_int, [_P]),
    "qipgpu_get_stream": (ctypes.c_int, [_P, _PP]),
    "qipgpu_batch_plan": (ctypes.c_int, [_I32, _P, _I64, _P, _P, _P, _I32, _I32, _P, _P]),
    "qipgpu_batch_create": (ctypes.c_int, [_I32, _P, _PP]),
    "qipgpu_batch_destroy": (ctypes.c_int, [_P]),
    "qipgpu_batch_load": (ctypes.c_int, [_P, _I64, _P, _P, _P, _P, _P]),
    "qipgpu_batch_set_tuning": (ctypes.c_int, [_P, _P]),
    "qipgpu_batch_get_tuning": (ctypes.c_int, [_P, _P]),
    "qipgpu_batch_autotune": (ctypes.c_int, [_P, _I32, _P, _PD]),
    "qipgpu_batch_reset": (ctypes.c_int, [_P]),
    "qipgpu_batch_run": (ctypes.c_int, [_P]),
    "qipgpu_batch_synchronize": (ctypes.c_int, [_P]),
    "qipgpu_batch_get_state": (ctypes.c_int, [_P, _I32, _P]),
    "qipgpu_batch_set_state": (ctypes.c_int, [_P, _I32, _P]),
    "qipgpu_batch_info": (ctypes.c_int, [_P, _P, _I32]),
    "qipgpu_batch_get_stream": (ctypes.c_int, [_P, _PP]),
}


class BatchTuning(ctypes.Structure):
    """qipgpu_batch_tuning"""
    _fields_ = [("fusion_width", ctypes.c_int32), ("lds_max_qubits", ctypes.c_int32),
                ("chunk_tiles", ctypes.c_int32), ("use_graph", ctypes.c_int32)]

_lib = None


def load(path: str = LIB_PATH):
    """Load libqipgpu.so and bind every entry point; raises OSError if it is not built."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise OSError("libqipgpu.so is not built at {} (run __graft_entry__.build())".format(path))
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib


def last_error() -> str:
    msg = load().qipgpu_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str = "") -> None:
    """Turn a status code into the exception the reference would raise for the same fault."""
    if rc == OK:
        return
    msg = last_error()
    if what:
        msg = "{}: {}".format(what, msg) if msg else what
    if rc == EINVAL:
        raise ValueError(msg)
    if rc == ENOMEM:
        raise MemoryError(msg)
    if rc == EZERODIV:
        raise ZeroDivisionError("float division by zero")
    if rc == EUNSUP:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)


def device_count() -> int:
    c = ctypes.c_int32(0)
    rc = load().qipgpu_device_count(ctypes.byref(c))
    return int(c.value) if rc == OK else 0


def i32(seq) -> numpy.ndarray:
    return numpy.ascontiguousarray(numpy.asarray(seq, dtype=numpy.int64).astype(numpy.int32))


def i32_bytes(seq) -> bytes:
    """A short int sequence packed as int32 for a `void*` argument (ctypes passes a bytes object's
    own buffer, read-only, for the duration of the call)."""
    return struct.pack("%di" % len(seq), *seq)


# matrices up to this size are handed over as a bytes copy, which ctypes passes faster than a
# numpy array's address; larger ones by address
SMALL_MATRIX_BYTES = 1 << 16


def c128_arg(m: numpy.ndarray):
    """A complex128 matrix for a `const double*` argument: a bytes copy when small, else the address
    of a C-contiguous complex128 array. The caller keeps the returned object until the call returns."""
    if m.dtype != numpy.complex128 or not m.flags.c_contiguous:
        m = numpy.ascontiguousarray(m, dtype=numpy.complex128)
    if m.nbytes <= SMALL_MATRIX_BYTES:
        return m.tobytes(), m
    return m.ctypes.data, m


def ptr(a: numpy.ndarray) -> int:
    """Address of `a`'s data for a `void*` argument. The int holds no reference to `a`: the caller
    keeps the array (or the array a view comes from) alive until the call returns."""
    return a.ctypes.data
