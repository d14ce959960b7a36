"""CPU tests of the boundary: the library loads and exports every declared entry point, the
ctypes signatures cover exactly the header, and the descriptor planner produces the operator
the reference's kronselect_dot applies (checked against the oracle on dense expansions)."""
import os
import re

import numpy
import pytest

from oracle import oracle
from qip_amd import CMat, SwapMat, _capi, planner

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qipgpu.h")
RNG = numpy.random.default_rng(99)


def declared_symbols():
    with open(HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(qipgpu_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header():
    lib = _capi.load()
    names = declared_symbols()
    assert len(names) >= 29
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(_capi.SIGNATURES) == names


def test_library_reports_version():
    assert b"gfx950" in _capi.load().qipgpu_version()


def expand_calls(calls, n):
    """Dense 2^n operator of a list of planner calls, applied in order (numpy only)."""
    dim = 2 ** n
    total = numpy.eye(dim, dtype=complex)
    expanded = []
    for call in calls:
        expanded.extend(call.matrix_calls() if isinstance(call, planner.ProductCall) else [call])
    for call in expanded:
        op = numpy.zeros((dim, dim), dtype=complex)
        t, c = call.targets, call.controls
        for col in range(dim):
            bits = [(col >> (n - 1 - q)) & 1 for q in range(n)]
            if not all(bits[q] for q in c):
                op[col, col] = 1
                continue
            j = int("".join(str(bits[q]) for q in t), 2) if t else 0
            for i in range(2 ** len(t)):
                nb = list(bits)
                for pos, q in enumerate(t):
                    nb[q] = (i >> (len(t) - 1 - pos)) & 1
                op[int("".join(map(str, nb)), 2), col] += call.mat[i, j]
        total = op @ total
    return total


def oracle_dense(mats, n):
    dim = 2 ** n
    out = numpy.zeros((dim, dim), dtype=complex)
    for col in range(dim):
        e = numpy.zeros(dim, dtype=complex)
        e[col] = 1
        o = numpy.zeros(dim, dtype=complex)
        oracle.kronselect_dot(mats, e, n, o)
        out[:, col] = o
    return out


def rand_u(d):
    z = RNG.standard_normal((d, d)) + 1j * RNG.standard_normal((d, d))
    return numpy.linalg.qr(z)[0]


X = numpy.array([[0, 1], [1, 0]], dtype=complex)


@pytest.mark.parametrize("mats", [
    {0: rand_u(2), 2: rand_u(2)},
    {(2, 0): rand_u(4)},
    {(3, 1): CMat(rand_u(2))},
    {(0, 3, 1): CMat(CMat(X))},
    {(1, 0, 3, 2): SwapMat(2)},
    {(2, 0, 1): CMat(SwapMat(1))},
    {(1, 3): CMat(X), (1, 0): CMat(X)},              # C(Not) on a 2-qubit register
    {(0, 1): CMat(X), (1, 2): rand_u(4), 3: rand_u(2)},
    {0: [[0, 1], [1, 0]]},
])
def test_planner_matches_kronselect(mats):
    n = 4
    calls = planner.plan(mats)
    got = expand_calls(calls, n)
    assert numpy.max(numpy.abs(got - oracle_dense(mats, n))) < 1e-14


def test_planner_errors_match_reference():
    with pytest.raises(Exception, match="Shape of square submatrix"):
        planner.plan({(0, 1): numpy.eye(2)})
    with pytest.raises(Exception, match="Type of indices must be tuple"):
        planner.plan({"a": numpy.eye(2)})
    with pytest.raises(ValueError, match="Cannot pass matrices"), pytest.warns(PendingDeprecationWarning):
        planner.plan({0: numpy.matrix(numpy.eye(2))})   # an ndarray subclass is not an ndarray here


def test_planner_swap_is_pairwise():
    calls = planner.plan({(0, 1, 2, 3, 4, 5): SwapMat(3)})
    assert [c.targets for c in calls] == [[0, 3], [1, 4], [2, 5]]
    calls = planner.plan({(7, 0, 1, 2, 3): CMat(SwapMat(2))})
    assert all(c.controls == [7] for c in calls)


def test_check_windows():
    from qip_amd.backend import check_windows
    check_windows(3, 8, 0, 0)
    with pytest.raises(ValueError, match="Input vector size plus offset"):
        check_windows(3, 8, 1, 0)
    with pytest.raises(ValueError, match="Output vector size plus offset"):
        check_windows(3, 8, 0, 2)


def test_tabulate():
    from qip_amd.backend import tabulate
    t = tabulate(lambda x: (x * 3) % 5, 3)
    assert t.dtype == numpy.int32 and list(t) == [(x * 3) % 5 for x in range(8)]
    with pytest.raises(OverflowError):
        tabulate(lambda x: 2 ** 33, 2)
    assert list(tabulate(numpy.arange(4)[::-1], 2)) == [3, 2, 1, 0]


@pytest.mark.parametrize("n,targets", [(24, list(range(23))), (30, list(range(30))), (12, [0, 11, 5]),
                                       (16, [15, 14, 13, 12]), (20, [0, 1, 2]), (13, list(range(13)))])
def test_product_passes_keep_low_bits(n, targets):
    """Every product sweep holds at most 8 factors above state bit 3 and at most 12 in all; the
    passes cover each factor once; the low-bit factors share the first pass."""
    from qip_amd.backend import PRODUCT_LOW_BITS, PRODUCT_MAX, product_passes
    passes = product_passes(n, targets)
    seen = sorted(j for p in passes for j in p)
    assert seen == list(range(len(targets)))
    low = {j for j, q in enumerate(targets) if n - 1 - q < PRODUCT_LOW_BITS}
    assert low <= set(passes[0])
    for p in passes:
        assert len(p) <= PRODUCT_MAX
        assert sum(1 for j in p if j not in low) <= PRODUCT_MAX - PRODUCT_LOW_BITS
    high = len(targets) - len(low)
    assert len(passes) == max(1, -(-high // (PRODUCT_MAX - PRODUCT_LOW_BITS)))
