"""State backends: the StateType interface and the MI355X plain backend.

``StateType`` restates the reference's backend interface (qip/backend.py:14-65) method for
method; graphs call it through ``run(..., backend_constructor=...)`` (qip/pipeline.py:74,95-96,
133). ``GpuBackend`` is the drop-in for ``CythonBackend`` (qip/backend.py:68-175): the state
lives in HBM behind a libqipgpu handle and every method is one or a few HIP kernels:

  make_state            -> qipgpu_init_product / qipgpu_set_range        (backend.py:73-104)
  kronselect_dot        -> planner.plan + qipgpu_apply_matrix             (backend.py:112-116)
  func_apply            -> qipgpu_func_apply_table_ex                     (backend.py:118-121)
  total_prob            -> qipgpu_total_prob                              (backend.py:123-124)
  measure               -> qipgpu_measure                                 (backend.py:126-134)
  reduce_measure        -> qipgpu_reduce_measure                          (backend.py:136-151)
  soft_measure          -> qipgpu_soft_measure                            (backend.py:153-156)
  measure_probabilities -> qipgpu_probabilities / qipgpu_top_probabilities (backend.py:158-163)
  get_state / ranges    -> qipgpu_get_range / set_range / add_range       (backend.py:106-175)

Differences from CythonBackend, all deliberate:
  * gates are applied in place (no state/arena double buffer), so memory is one state;
  * get_state() returns a fresh host copy, not the live buffer;
  * statetype complex64 works (the reference's memoryviews accept complex128 only);
  * the uniform draw of soft_measure comes from the same Python `random` module, once per
    soft_measure, as kronprod.pyx:362 does, and is handed to the engine;
  * measure_probabilities(top_k) returns the true top-k (the reference's heap writes past its
    end, kronprod.pyx:298-311), ties ordered by ascending outcome;
  * func_apply applies the intended permutation on every rest register bit (the reference loops
    j < n_reg2 over the rest qubits, func_apply.pyx:109, and loses amplitudes when there are more
    rest qubits than reg2 qubits); set ``reference_func_bound = True`` for the reference's output
    there.
"""
import ctypes
import random
import weakref
from typing import Callable, Mapping, Optional, Sequence, Tuple, Union

import numpy

from . import _capi
from . import planner
from .util import IndexType, InitialState, MatrixType


class StateType(object):
    """The backend interface (qip/backend.py:14-65)."""

    def __init__(self, n: int, state: object):
        self.n = n
        self.state = state

    def get_state(self) -> object:
        return self.state

    def kronselect_dot(self, mats: Mapping[IndexType, MatrixType],
                       input_offset: int = 0, output_offset: int = 0) -> None:
        raise NotImplementedError("kronselect_dot not implemented by base class")

    def func_apply(self, reg1_indices: Sequence[int], reg2_indices: Sequence[int], func: Callable[[int], int],
                   input_offset: int = 0, output_offset: int = 0) -> None:
        raise NotImplementedError("func_apply not implemented by base class")

    def total_prob(self) -> float:
        raise NotImplementedError("total_prob not implemented by base class")

    def measure(self, indices: Sequence[int], measured: Optional[int] = None,
                measured_prob: Optional[float] = None,
                input_offset: int = 0, output_offset: int = 0) -> Tuple[int, float]:
        raise NotImplementedError("measure not implemented by base class")

    def reduce_measure(self, indices: Sequence[int], measured: Optional[int] = None,
                       measured_prob: Optional[float] = None,
                       input_offset: int = 0, output_offset: int = 0) -> Tuple[int, float]:
        raise NotImplementedError("reduce_measure not implemented by base class")

    def soft_measure(self, indices: Sequence[int], measured: Optional[int] = None,
                     input_offset: int = 0) -> Tuple[int, float]:
        raise NotImplementedError("soft_measure not implemented by base class")

    def measure_probabilities(self, indices: Sequence[int],
                              top_k: int = 0) -> Union[Sequence[float], Tuple[Sequence[int], Sequence[float]]]:
        """If top_k is 0 output all probabilities, else (indices, probs) of the top_k."""
        raise NotImplementedError("measure_probabilities not implemented by base class")

    def get_state_size(self) -> int:
        raise NotImplementedError("get_state_size not implemented by base class")

    def get_relative_range(self, start: int, end: int) -> Sequence[complex]:
        raise NotImplementedError("get_relative_range not implemented by base class")

    def overwrite_relative_range(self, start: int, end: int, data: Sequence[complex]):
        raise NotImplementedError("overwrite_relative_range not implemented by base class")

    def addto_relative_range(self, start: int, end: int, data: Sequence[complex]):
        raise NotImplementedError("addto_relative_range not implemented by base class")

    def close(self):
        pass


def dtype_code(statetype) -> int:
    st = numpy.dtype(statetype)
    if st == numpy.complex128:
        return _capi.C128
    if st == numpy.complex64:
        return _capi.C64
    raise ValueError("statetype must be numpy.complex128 or numpy.complex64, not {}".format(statetype))


def validate_measured(k: int, measured, measured_prob) -> None:
    """The argument checks of kronprod.pyx:402-407 / 453-458, same messages."""
    if measured is not None:
        if not (0 <= measured < 2 ** k):
            raise ValueError("Measured value must be less than 2**len(indices)")
    if measured_prob is not None:
        if not 0.0 < measured_prob <= 1.0:
            raise ValueError("measured_prob must be 0 < p <= 1")


def check_windows(n: int, state_len: int, input_offset: int, output_offset: int) -> None:
    """qip/util.py:29-33: a full-state backend accepts only zero windows offsets."""
    if state_len + input_offset > 2 ** n:
        raise ValueError("Input vector size plus offset may be no larger than the total number of qubit states (2^n)")
    if state_len + output_offset > 2 ** n:
        raise ValueError("Output vector size plus offset may be no larger than the total number of qubit states (2^n)")


class _FuncTable:
    """A function's f table resident on the device (func_apply.pyx:66-69 tabulates it per call)."""

    def __init__(self, n_reg1: int, host: numpy.ndarray):
        self.n_reg1 = n_reg1
        self.host = host
        h = ctypes.c_void_p()
        _capi.check(_capi.load().qipgpu_table_create(n_reg1, _capi.ptr(host), ctypes.byref(h)), "table_create")
        self.handle = h

    def __del__(self):
        try:
            if self.handle:
                _capi.load().qipgpu_table_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


_func_tables = weakref.WeakKeyDictionary()


class FuncTable:
    """A classical function given by its table of values, f(x) = values[x].

    It is callable like any f, so every backend accepts it, and GpuBackend keeps its device copy
    for the object's lifetime: pass the same FuncTable to every func_apply of a loop (a Grover
    oracle, say) and the table crosses PCIe once."""

    def __init__(self, values):
        v = numpy.asarray(values)
        if v.ndim != 1 or (v.size & (v.size - 1)) or v.size == 0:
            raise ValueError("a function table needs 2**k entries")
        vi = v.astype(numpy.int64)
        if vi.min() < -2 ** 31 or vi.max() >= 2 ** 31:
            raise OverflowError("value too large to convert to int")
        self.values = numpy.ascontiguousarray(vi.astype(numpy.int32))
        self._device = None

    def __call__(self, x: int) -> int:
        return int(self.values[x])

    def device_table(self, n_reg1: int) -> "_FuncTable":
        if self.values.shape[0] != 2 ** n_reg1:
            raise ValueError("function table has {} entries for a {}-qubit register".format(self.values.shape[0], n_reg1))
        if self._device is None:
            self._device = _FuncTable(n_reg1, self.values)
        return self._device


def tabulate(func, n_reg1: int) -> numpy.ndarray:
    """f(x) for every x of the register as C int, like func_apply.pyx:66-69 (OverflowError past
    int32, as Cython's conversion raises). An integer ndarray is taken as the table itself."""
    if isinstance(func, numpy.ndarray):
        table = numpy.ascontiguousarray(func.astype(numpy.int64))
        if table.shape != (2 ** n_reg1,):
            raise ValueError("function table must have 2**len(reg1) entries")
        if table.size and (table.min() < -2 ** 31 or table.max() >= 2 ** 31):
            raise OverflowError("value too large to convert to int")
        return table.astype(numpy.int32)
    table = numpy.empty(2 ** n_reg1, dtype=numpy.int32)
    for x in range(2 ** n_reg1):
        v = int(func(x))
        if not (-2 ** 31 <= v < 2 ** 31):
            raise OverflowError("value too large to convert to int")
        table[x] = v
    return table


def func_table(func, n_reg1: int, cache: bool = False) -> _FuncTable:
    """The device table of `func` on an n_reg1-qubit register: a FuncTable's own, a fresh one, or
    (cache=True) one kept per function object for later calls."""
    if isinstance(func, FuncTable):
        return func.device_table(n_reg1)
    if not cache or isinstance(func, numpy.ndarray):
        return _FuncTable(n_reg1, tabulate(func, n_reg1))
    try:
        per_func = _func_tables.get(func)
    except TypeError:
        per_func = None
        cache = False
    if per_func is not None and n_reg1 in per_func:
        return per_func[n_reg1]
    t = _FuncTable(n_reg1, tabulate(func, n_reg1))
    if cache:
        try:
            _func_tables.setdefault(func, {})[n_reg1] = t
        except TypeError:
            pass
    return t


PRODUCT_MAX = 12            # single-qubit factors per product sweep (QIPGPU_MAX_PRODUCT)
PRODUCT_LOW_BITS = 4        # lowest state bits every product sweep keeps in its tile
PRODUCT_MIN_QUBITS = 12     # the product sweep tiles 12 state bits


def product_passes(n: int, targets: Sequence[int]):
    """Split the factors of a product of single-qubit gates (on distinct qubits; they commute)
    into sweeps: every sweep keeps the PRODUCT_LOW_BITS lowest state bits in its tile, so it takes
    the factors on those bits and at most PRODUCT_MAX - PRODUCT_LOW_BITS others. Returns lists of
    indices into `targets`, the first holding every low-bit factor."""
    low = [j for j, q in enumerate(targets) if n - 1 - q < PRODUCT_LOW_BITS]
    high = [j for j, q in enumerate(targets) if n - 1 - q >= PRODUCT_LOW_BITS]
    step = PRODUCT_MAX - PRODUCT_LOW_BITS
    chunks = [low + high[:step]] + [high[i:i + step] for i in range(step, len(high), step)]
    return [c for c in chunks if c]


class GateGraph:
    """A captured gate sequence (qipgpu_graph): the kernels of the gates applied between
    GpuBackend.capture_begin and capture_end, launched again with launch_graph on any state that
    holds the buffer it was captured on. Freed with the object."""

    def __init__(self, lib, handle: ctypes.c_void_p):
        self.handle = handle
        self._finalizer = weakref.finalize(self, lib.qipgpu_graph_destroy, handle)


def pack_matrix(targets: Sequence[int], mat: numpy.ndarray, controls: Sequence[int]) -> tuple:
    """qipgpu_apply_matrix's arguments after the state handle: (k, targets, matrix, number of
    controls, controls), then the array the matrix argument points into (kept alive with them)."""
    k = len(targets)
    if mat.shape != (2 ** k, 2 ** k):
        raise ValueError("matrix shape {} does not match {} targets".format(mat.shape, k))
    m, keep = _capi.c128_arg(mat)
    return (k, _capi.i32_bytes(targets), m, len(controls), _capi.i32_bytes(controls) if len(controls) else None,
            keep)


class GpuBackend(StateType):
    """Plain MI355X backend: one in-place HIP sweep per connected component of each operator,
    complex128 (the drop-in) or complex64 state, FP64 reductions."""

    # A plain callable is tabulated again on every func_apply, as the reference does
    # (func_apply.pyx:66-69), so a function whose values change between calls is followed. Pass a
    # FuncTable (or an integer ndarray) to keep a table on the device across calls; set this to True
    # to keep one device table per function object instead (only for functions whose values never
    # change).
    cache_func_tables = False
    # True: func_apply keeps the reference's rest-qubit loop bound (func_apply.pyx:107-111). With
    # more rest qubits than reg2 qubits, the reference sets only the first n_reg2 rest qubits (in
    # qubit order) of its output index, so every amplitude with one of the others at 1 ends 0. The
    # default applies the permutation to the whole state.
    reference_func_bound = False

    def __init__(self, n: int, handle: ctypes.c_void_p, dtype: int):
        super().__init__(n, None)
        self._h = handle
        self._dtype = dtype
        self._np_dtype = numpy.complex128 if dtype == _capi.C128 else numpy.complex64
        self._lib = _capi.load()

    # A destroyed state of up to 256 MiB leaves its buffer to the next state of its size, so a gate
    # graph captured on one run's state can be launched on the next run's (pipeline.Schedule.play)
    GRAPH_STATE_BYTES = 256 << 20

    @property
    def graphs_reusable(self) -> bool:
        """Whether gate graphs captured on this state can serve the states made after it."""
        return type(self).supports_gate_graphs and (2 ** self.n) * numpy.dtype(self._np_dtype).itemsize <= \
            self.GRAPH_STATE_BYTES

    supports_gate_graphs = True

    def capture_begin(self) -> None:
        """Record the following gate calls into a graph instead of running them (only gates whose
        operands are kernel arguments: dense on up to 3 targets, diagonal on up to 6, products of
        single-qubit factors; anything else raises NotImplementedError and drops the capture)."""
        _capi.check(self._lib.qipgpu_capture_begin(self._h), "capture_begin")

    def capture_end(self) -> Optional[GateGraph]:
        """The graph of the gates recorded since capture_begin (none of them has run yet), or None
        when the capture was dropped."""
        g = ctypes.c_void_p()
        rc = self._lib.qipgpu_capture_end(self._h, ctypes.byref(g))
        if rc == _capi.EUNSUP:
            return None
        _capi.check(rc, "capture_end")
        return GateGraph(self._lib, g)

    def graph_matches(self, graph: GateGraph) -> bool:
        return bool(self._lib.qipgpu_graph_matches(self._h, graph.handle))

    def launch_graph(self, graph: GateGraph) -> None:
        _capi.check(self._lib.qipgpu_graph_launch(self._h, graph.handle), "graph_launch")

    # -- construction: qip/backend.py:73-104 --------------------------------------------------
    @classmethod
    def make_state(cls, n: int, index_groups: Sequence[Sequence[int]], feed_list: Sequence[InitialState],
                   inputstartindex: Optional[int] = None, inputendindex: Optional[int] = None,
                   outputstartindex: Optional[int] = None, outputendindex: Optional[int] = None,
                   statetype: type = numpy.complex128) -> "GpuBackend":
        for name, v, full in (("inputstartindex", inputstartindex, 0), ("inputendindex", inputendindex, 2 ** n),
                              ("outputstartindex", outputstartindex, 0), ("outputendindex", outputendindex, 2 ** n)):
            if v is not None and v != full:
                raise NotImplementedError("{} windows are a distributed-worker feature; this backend holds "
                                          "the whole state".format(name))
        dtype = dtype_code(statetype)
        lib = _capi.load()
        h = ctypes.c_void_p()
        _capi.check(lib.qipgpu_create(int(n), dtype, ctypes.byref(h)), "create")
        be = cls(n, h, dtype)
        try:
            be._init(index_groups, feed_list)
        except Exception:
            be.close()
            raise
        return be

    def _init(self, index_groups, feed_list) -> None:
        n = self.n
        if len(feed_list) == 0:
            _capi.check(self._lib.qipgpu_init_product(self._h, 0, None, None, None), "init")
            return
        groups = [list(g) for g in index_groups]
        feeds = []
        for g, f in zip(groups, feed_list):
            v = numpy.asarray(f).astype(numpy.complex128).ravel()
            if v.shape[0] < 2 ** len(g):
                raise IndexError("feed of length {} for {} qubits".format(v.shape[0], len(g)))
            feeds.append(v[: 2 ** len(g)])
        if len(groups) == 1 and groups[0] == list(range(n)):
            # a full-state feed (the deprecated run(state=...), qip/pipeline.py:129-131): copy it
            self.overwrite_relative_range(0, 2 ** n, feeds[0])
            return
        glen = _capi.i32([len(g) for g in groups])
        gidx = _capi.i32([q for g in groups for q in g])
        flat = numpy.ascontiguousarray(numpy.concatenate(feeds)).view(numpy.float64)
        _capi.check(self._lib.qipgpu_init_product(self._h, len(groups), _capi.ptr(glen), _capi.ptr(gidx),
                                                  _capi.ptr(flat)), "init")

    def reinit(self, index_groups: Sequence[Sequence[int]] = (), feed_list: Sequence[InitialState] = ()) -> None:
        """Rewrite the state in place as make_state would build it (no reallocation)."""
        self._init(list(index_groups), list(feed_list))

    @property
    def handle(self):
        return self._h

    @property
    def stream(self) -> int:
        """The hipStream_t (as an integer) the engine launches this state's kernels on."""
        st = ctypes.c_void_p()
        _capi.check(self._lib.qipgpu_get_stream(self._h, ctypes.byref(st)), "get_stream")
        return st.value or 0

    @property
    def dtype(self):
        return self._np_dtype

    # -- state access ---------------------------------------------------------------------------
    def get_state(self) -> numpy.ndarray:
        return self.get_relative_range(0, 2 ** self.n)

    def get_state_size(self) -> int:
        return 2 ** self.n

    def get_relative_range(self, start: int, end: int) -> numpy.ndarray:
        start, end = self._slice(start, end)
        out = numpy.empty(end - start, dtype=self._np_dtype)
        _capi.check(self._lib.qipgpu_get_range(self._h, start, end, _capi.ptr(out)), "get_range")
        return out

    def overwrite_relative_range(self, start: int, end: int, data: Sequence[complex]):
        start, end = self._slice(start, end)
        buf = numpy.ascontiguousarray(numpy.broadcast_to(numpy.asarray(data, dtype=self._np_dtype), (end - start,)))
        _capi.check(self._lib.qipgpu_set_range(self._h, start, end, _capi.ptr(buf)), "set_range")

    def addto_relative_range(self, start: int, end: int, data: Sequence[complex]):
        start, end = self._slice(start, end)
        buf = numpy.ascontiguousarray(numpy.broadcast_to(numpy.asarray(data, dtype=self._np_dtype), (end - start,)))
        _capi.check(self._lib.qipgpu_add_range(self._h, start, end, _capi.ptr(buf)), "add_range")

    def _slice(self, start: int, end: int) -> Tuple[int, int]:
        # numpy slice semantics of state[start:end] (qip/backend.py:168-175)
        s, e, _ = slice(start, end).indices(2 ** self.n)
        return s, max(s, e)

    # -- operators ------------------------------------------------------------------------------
    def kronselect_dot(self, mats: Mapping[IndexType, MatrixType],
                       input_offset: int = 0, output_offset: int = 0) -> None:
        if input_offset or output_offset:
            check_windows(self.n, 2 ** self.n, input_offset, output_offset)
        for call in planner.plan(mats):
            if isinstance(call, planner.MatrixCall):
                # the call's arguments are packed once and kept with it (a planned Descriptor's calls
                # serve every run of its gate node)
                p = call.packed
                if p is None:
                    p = call.packed = pack_matrix(call.targets, call.mat, call.controls)
                rc = self._lib.qipgpu_apply_matrix(self._h, p[0], p[1], p[2], p[3], p[4])
                if rc:
                    _capi.check(rc, "apply_matrix")
            elif isinstance(call, planner.ProductCall):
                self._apply_product_call(call)
            else:
                self.apply_general(call)

    def apply_general(self, call: "planner.GeneralCall") -> None:
        """A descriptor component by the reference's column-sum formula (qipgpu_apply_kron)."""
        lens, idx, levels, kinds, params, mats = call.packed()
        a = [_capi.i32(v) for v in (lens, idx, levels, kinds, params)]
        m = numpy.ascontiguousarray(mats, dtype=numpy.complex128)
        _capi.check(self._lib.qipgpu_apply_kron(self._h, len(lens), *[_capi.ptr(v) for v in a],
                                                _capi.ptr(m.view(numpy.float64))), "apply_kron")

    def apply_product(self, targets: Sequence[int], mats: Sequence[numpy.ndarray]) -> None:
        """Single-qubit operators mats[j] on distinct qubits targets[j], up to 12 per sweep
        (n >= 12; smaller states take Kronecker blocks of 3 through apply_matrix)."""
        self._apply_product_call(planner.ProductCall(targets, mats))

    def _apply_product_call(self, call: "planner.ProductCall") -> None:
        if self.n < PRODUCT_MIN_QUBITS:
            for c in call.matrix_calls():
                self.apply_matrix(c.targets, c.mat)
            return
        # the passes depend on n: packed per state size and kept with the call
        if call.packed is None or call.packed[0] != self.n:
            passes = []
            for chunk in product_passes(self.n, call.targets):
                m = b"".join(numpy.asarray(call.mats[j], dtype=numpy.complex128).reshape(2, 2).tobytes()
                             for j in chunk)
                passes.append((len(chunk), _capi.i32_bytes([call.targets[j] for j in chunk]), m))
            call.packed = (self.n, passes)
        for k, t, m in call.packed[1]:
            rc = self._lib.qipgpu_apply_product(self._h, k, t, m)
            if rc:
                _capi.check(rc, "apply_product")

    def apply_matrix(self, targets: Sequence[int], mat: numpy.ndarray, controls: Sequence[int] = ()) -> None:
        """One engine call: `mat` on `targets` (first = MSB) where all `controls` are 1."""
        p = pack_matrix(targets, mat, controls)
        rc = self._lib.qipgpu_apply_matrix(self._h, p[0], p[1], p[2], p[3], p[4])
        if rc:
            _capi.check(rc, "apply_matrix")

    def apply_fused(self, targets: Sequence[int], mat: numpy.ndarray) -> None:
        """Dense operator on <= 6 qubits through the split-BF16 MFMA kernel (complex64 states,
        n >= 12). The result carries a pending FP64 renormalisation that the next fused call or
        any other operation applies."""
        t = _capi.i32(targets)
        m = numpy.ascontiguousarray(mat, dtype=numpy.complex128)
        if m.shape != (2 ** len(t), 2 ** len(t)):
            raise ValueError("matrix shape {} does not match {} targets".format(m.shape, len(t)))
        _capi.check(self._lib.qipgpu_fused_apply(self._h, len(t), _capi.ptr(t), _capi.ptr(m.view(numpy.float64))),
                    "fused_apply")

    def fused_pattern_sweep(self, targets: Sequence[int]) -> None:
        """Diagnostic: load and store the state in place with the tile pattern apply_fused uses for
        these targets, without arithmetic (the state is unchanged)."""
        t = _capi.i32(targets)
        _capi.check(self._lib.qipgpu_fused_pattern_sweep(self._h, len(t), _capi.ptr(t)), "fused_pattern_sweep")

    def fused_finish(self) -> None:
        """Apply the pending renormalisation factor of the fused path now."""
        _capi.check(self._lib.qipgpu_fused_finish(self._h), "fused_finish")

    def fused_last_norm(self) -> float:
        """FP64 norm of the state produced by the last fused call, before renormalisation."""
        out = ctypes.c_double()
        _capi.check(self._lib.qipgpu_fused_last_norm(self._h, ctypes.byref(out)), "fused_last_norm")
        return float(out.value)

    def func_apply(self, reg1_indices: Sequence[int], reg2_indices: Sequence[int], func: Callable[[int], int],
                   input_offset: int = 0, output_offset: int = 0) -> None:
        # CythonBackend ignores both offsets (qip/backend.py:118-121); FOp passes n as the
        # fourth positional argument (qip/operators.py:291)
        reg1 = _capi.i32(reg1_indices)
        reg2 = _capi.i32(reg2_indices)
        table = func_table(func, len(reg1), cache=self.cache_func_tables)
        flags = _capi.FUNC_REFERENCE_BOUND if self.reference_func_bound else 0
        _capi.check(self._lib.qipgpu_func_apply_table_ex(self._h, len(reg1), _capi.ptr(reg1), len(reg2),
                                                         _capi.ptr(reg2), table.handle, flags), "func_apply")

    # -- reductions and measurement -------------------------------------------------------------
    def total_prob(self) -> float:
        out = ctypes.c_double()
        _capi.check(self._lib.qipgpu_total_prob(self._h, ctypes.byref(out)), "total_prob")
        return float(out.value)

    def measure(self, indices: Sequence[int], measured: Optional[int] = None,
                measured_prob: Optional[float] = None,
                input_offset: int = 0, output_offset: int = 0) -> Tuple[int, float]:
        idx = _capi.i32(indices)
        validate_measured(len(idx), measured, measured_prob)
        u = random.random() if (measured is None or measured_prob is None) else 0.0
        m = ctypes.c_int64()
        p = ctypes.c_double()
        _capi.check(self._lib.qipgpu_measure(self._h, len(idx), _capi.ptr(idx), u,
                                             -1 if measured is None else int(measured),
                                             -1.0 if measured_prob is None else float(measured_prob),
                                             ctypes.byref(m), ctypes.byref(p)), "measure")
        return int(m.value), float(p.value)

    def reduce_measure(self, indices: Sequence[int], measured: Optional[int] = None,
                       measured_prob: Optional[float] = None,
                       input_offset: int = 0, output_offset: int = 0, *, shrink: bool = False) -> Tuple[int, float]:
        """Measure, then drop the measured qubits. By default the result is the reference's: the
        reduced state repeated 2^k times over the 2^n buffer (qip/backend.py:139-142). With
        shrink=True the state becomes the 2^(n-k) amplitudes of the unmeasured qubits and `n`
        drops by k."""
        idx = _capi.i32(indices)
        validate_measured(len(idx), measured, measured_prob)
        u = random.random() if (measured is None or measured_prob is None) else 0.0
        m = ctypes.c_int64()
        p = ctypes.c_double()
        fn = self._lib.qipgpu_reduce_measure_shrink if shrink else self._lib.qipgpu_reduce_measure
        _capi.check(fn(self._h, len(idx), _capi.ptr(idx), u,
                       -1 if measured is None else int(measured),
                       -1.0 if measured_prob is None else float(measured_prob),
                       ctypes.byref(m), ctypes.byref(p)), "reduce_measure")
        if shrink:
            self.n -= len(idx)
        return int(m.value), float(p.value)

    def soft_measure(self, indices: Sequence[int], measured: Optional[int] = None,
                     input_offset: int = 0) -> Tuple[int, float]:
        idx = _capi.i32(indices)
        u = random.random()
        if measured is not None and not (0 <= measured < 2 ** len(idx)):
            raise ValueError("Measured value must be less than 2**len(indices)")
        m = ctypes.c_int64()
        p = ctypes.c_double()
        _capi.check(self._lib.qipgpu_soft_measure(self._h, len(idx), _capi.ptr(idx), u,
                                                  -1 if measured is None else int(measured),
                                                  ctypes.byref(m), ctypes.byref(p)), "soft_measure")
        return int(m.value), float(p.value)

    def measure_probabilities(self, indices: Sequence[int],
                              top_k: int = 0) -> Union[Sequence[float], Tuple[Sequence[int], Sequence[float]]]:
        idx = _capi.i32(indices)
        if top_k:
            kk = min(int(top_k), 2 ** len(idx))
            ms = numpy.zeros(max(kk, 1), dtype=numpy.int64)
            ps = numpy.zeros(max(kk, 1), dtype=numpy.float64)
            cnt = ctypes.c_int32()
            _capi.check(self._lib.qipgpu_top_probabilities(self._h, len(idx), _capi.ptr(idx), int(top_k),
                                                           _capi.ptr(ms), _capi.ptr(ps), ctypes.byref(cnt)),
                        "top_probabilities")
            return [int(v) for v in ms[:cnt.value]], [float(v) for v in ps[:cnt.value]]
        out = numpy.zeros(2 ** len(idx), dtype=numpy.float64)
        _capi.check(self._lib.qipgpu_probabilities(self._h, len(idx), _capi.ptr(idx), _capi.ptr(out)),
                    "probabilities")
        return out

    def marginals(self, indices: Sequence[int]) -> numpy.ndarray:
        """Marginal distribution in measurement (entwine) order, FP64."""
        idx = _capi.i32(indices)
        out = numpy.zeros(2 ** len(idx), dtype=numpy.float64)
        _capi.check(self._lib.qipgpu_marginals(self._h, len(idx), _capi.ptr(idx), _capi.ptr(out)), "marginals")
        return out

    def synchronize(self) -> None:
        _capi.check(self._lib.qipgpu_synchronize(self._h), "synchronize")

    def flush(self) -> None:
        """Launch what the engine holds back on this state, without waiting for it: runs of diagonal
        gates and of single-qubit gates wait for the next operation that does not join them, so that
        a run is one sweep (diagonals) or one pass per tile (single-qubit gates)."""
        _capi.check(self._lib.qipgpu_fused_flush(self._h), "flush")

    def close(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._lib.qipgpu_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _settled(*states):
    for st in states:
        if hasattr(st, "flush"):
            st.flush()


def overlap(a: GpuBackend, b: GpuBackend) -> Tuple[complex, float, float]:
    """<a|b> in FP64 and the two squared norms, computed on the device."""
    _settled(a, b)
    re, im, na, nb = (ctypes.c_double() for _ in range(4))
    _capi.check(_capi.load().qipgpu_overlap(a.handle, b.handle, ctypes.byref(re), ctypes.byref(im),
                                            ctypes.byref(na), ctypes.byref(nb)), "overlap")
    return complex(re.value, im.value), float(na.value), float(nb.value)


def infidelity(a: GpuBackend, b: GpuBackend) -> float:
    """1 - |<a|b>|^2 / (<a|a><b|b>) in FP64."""
    ov, na, nb = overlap(a, b)
    return float(1.0 - (abs(ov) ** 2) / (na * nb))


def max_abs_diff(a: GpuBackend, b: GpuBackend) -> float:
    _settled(a, b)
    out = ctypes.c_double()
    _capi.check(_capi.load().qipgpu_max_abs_diff(a.handle, b.handle, ctypes.byref(out)), "max_abs_diff")
    return float(out.value)
