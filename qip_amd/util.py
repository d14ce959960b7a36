"""Host-side helpers of the QIP interface (counterparts of qip/util.py).

Types and small utilities the graph frontend and the backends share. The operator-descriptor
handling of ``kronselect_dot`` lives in ``qip_amd.planner``.
"""
import collections.abc
from typing import Any, Iterable, List, MutableSequence, Sequence, Tuple, TypeVar, Union

import numpy

InitialState = Union[MutableSequence[complex], int, numpy.ndarray]   # qip/util.py:8
IndexType = Union[int, Tuple[int, ...]]                              # qip/util.py:10
MatrixType = Any                                                     # qip/util.py:11
T = TypeVar("T")


def flatten(lst: Iterable[Any]) -> List[Any]:
    """One level of flattening; non-iterable items count as 1-tuples (qip/util.py:152-154).
    Uses collections.abc.Iterable: the reference's collections.Iterable no longer exists on
    Python >= 3.10."""
    out = []
    for item in lst:
        t = type(item)
        if t is list or t is tuple or t is range:          # the common cases, without the ABC check
            out.extend(item)
        elif t is int:
            out.append(item)
        elif isinstance(item, collections.abc.Iterable):
            out.extend(item)
        else:
            out.append(item)
    return out


def uint_to_bitarray(num: int, n: int) -> List[int]:
    """MSB-first bits of num (qip/util.py:137-142)."""
    return [(num >> (n - 1 - i)) & 1 for i in range(n)]


def bitarray_to_uint(bits: Sequence[int]) -> int:
    """Inverse of uint_to_bitarray (qip/util.py:145-149)."""
    v = 0
    for b in bits:
        v = (v << 1) | (1 if b else 0)
    return v


def qubit_index_notation(i: int, *qns: int, n: int = None) -> List[int]:
    """Split basis index i into per-register values (qip/util.py:157-167)."""
    if n is None:
        n = sum(qns)
    out = []
    used = 0
    for size in qns:
        out.append((i >> (n - used - size)) & ((1 << size) - 1))
        used += size
    return out


def gen_qubit_prints(state: Sequence[complex], *qs: Any):
    """Yield "|a,b,...> = amplitude" for the non-zero amplitudes (qip/util.py:170-184)."""
    sizes = [q if type(q) == int else q.n for q in qs]
    n = sum(sizes)
    for i in range(len(state)):
        if state[i] == 0:
            continue
        yield "|" + ",".join(map(str, qubit_index_notation(i, *sizes, n=n))) + "> = {}".format(state[i])
