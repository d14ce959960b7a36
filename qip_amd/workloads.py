"""Seeded synthetic workloads, expressed as kronselect_dot descriptors.

layered_random_circuit(n, depth, seed): every layer is a Haar-random single-qubit unitary on each
qubit followed by a brickwork of two-qubit gates, CNOT on (2j, 2j+1) in even layers and CZ on
(2j+1, 2j+2) in odd layers (control first). The Haar unitaries are U = Q diag(R_ii / |R_ii|) for
Q, R = qr(Z), Z = (randn + i randn) / sqrt(2) of shape 2x2 (Mezzadri 2007), drawn from
numpy.random.default_rng(seed) layer by layer, qubit by qubit, real part then imaginary part.
Gate counts: n * depth single-qubit gates, ceil(depth/2) * floor(n/2) + floor(depth/2) *
floor((n-1)/2) two-qubit gates (1780 in all for n = 30, depth = 40; 830 for n = 28, depth = 20).

The descriptors use the same CMat objects the graph frontend builds for C(X) / C(Z)
(qip/operators.py:182-268), so every backend receives exactly what a qip graph would send.
"""
from typing import Dict, List, Tuple

import numpy

from .backend import FuncTable
from .operators import CMat

X = numpy.array([[0.0, 1.0], [1.0, 0.0]], dtype=numpy.complex128)
Z = numpy.array([[1.0, 0.0], [0.0, -1.0]], dtype=numpy.complex128)
H = numpy.array([[1.0, 1.0], [1.0, -1.0]], dtype=numpy.complex128) / numpy.sqrt(2.0)


def haar_2x2(rng: numpy.random.Generator) -> numpy.ndarray:
    z = (rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))) / numpy.sqrt(2.0)
    q, r = numpy.linalg.qr(z)
    d = numpy.diagonal(r)
    return q * (d / numpy.abs(d))


def layered_random_circuit(n: int, depth: int, seed: int) -> List[Dict]:
    """The descriptors of the layered random circuit, in program order."""
    rng = numpy.random.default_rng(seed)
    ops = []
    for layer in range(depth):
        for q in range(n):
            ops.append({q: haar_2x2(rng)})
        if layer % 2 == 0:
            for j in range(0, n - 1, 2):
                ops.append({(j, j + 1): CMat(X)})
        else:
            for j in range(1, n - 1, 2):
                ops.append({(j, j + 1): CMat(Z)})
    return ops


def gate_counts(ops: List[Dict]) -> Tuple[int, int]:
    """(single-qubit, multi-qubit) gate counts of a descriptor list."""
    one = 0
    for d in ops:
        key = next(iter(d))
        if isinstance(key, int) or len(key) == 1:
            one += 1
    return one, len(ops) - one


def algorithmic_bytes_plain(ops: List[Dict], n: int, amp_bytes: int) -> int:
    """Algorithmic bytes of a plain run: read + write of every amplitude a gate can change,
    2 * S * 2^(n - c) for a gate with c control qubits as written (a single-qubit gate 2^n, CNOT and
    CZ, each one control, 2^(n-1)), with no credit for diagonal or permutation structure."""
    total = 0
    for d in ops:
        key = next(iter(d))
        m = d[key]
        controls = 0
        while isinstance(m, CMat):
            controls += 1
            m = m.m
        total += 2 * amp_bytes * 2 ** (n - controls)
    return total


def touched_bytes_plain(ops: List[Dict], n: int, amp_bytes: int) -> int:
    """Bytes the plain engine's kernels actually read and write: as algorithmic_bytes_plain, except
    that CZ, which the engine applies as a phase on the amplitudes with both qubits 1, touches
    2^(n-2) amplitudes."""
    total = 0
    for d in ops:
        key = next(iter(d))
        m = d[key]
        if isinstance(key, int) or len(key) == 1:
            total += 2 * amp_bytes * 2 ** n
        elif isinstance(m, CMat) and numpy.array_equal(numpy.asarray(m.m), Z):
            total += 2 * amp_bytes * 2 ** (n - 2)
        else:
            total += 2 * amp_bytes * 2 ** (n - 1)
    return total


def random_product_unitaries(n: int, seed: int) -> List[Dict]:
    """One Haar single-qubit unitary per qubit (a layer without the brickwork)."""
    rng = numpy.random.default_rng(seed)
    return [{q: haar_2x2(rng)} for q in range(n)]


# ---- iterative Grover search --------------------------------------------------------------------

def grover_iterations(n_search: int) -> int:
    """floor(pi/4 * sqrt(2^n_search))."""
    return int(numpy.floor(numpy.pi / 4 * numpy.sqrt(2.0 ** n_search)))


def grover_initial(n_search: int):
    """(index_groups, feeds) of the search register in uniform superposition and the ancilla in
    |-> (the defaults of the reference's examples/grovers_iterative.py:33-35)."""
    N = 2 ** n_search
    return ([list(range(n_search)), [n_search]],
            [numpy.ones(N) / numpy.sqrt(N), numpy.array([1.0, -1.0]) / numpy.sqrt(2.0)])


def grover_iteration_ops(n_search: int, x_star: int, diffusion: str = "f") -> List[Tuple]:
    """One Grover iteration on search qubits 0..n_search-1 and ancilla qubit n_search.

    Oracle: F(x == x_star) into the |-> ancilla (phase kickback), as grovers_iterative.py:20-21.
    Diffusion "f": H on the register, F(x == 0), H (grovers_iterative.py:24-26);
    diffusion "mcz": H, X, the (n_search-1)-controlled Z, X, H. Both equal
    -(2|s><s| - I) on the search register. Operations are ("kron", mats) for
    backend.kronselect_dot and ("func", reg1, reg2, f) for backend.func_apply, with f a FuncTable
    (callable, and kept on the device across iterations)."""
    N = 2 ** n_search
    search = list(range(n_search))
    anc = [n_search]
    xs = numpy.arange(N)
    oracle = FuncTable(xs == x_star)
    hs = {q: H for q in search}
    ops = [("func", search, anc, oracle)]
    if diffusion == "f":
        ops += [("kron", hs), ("func", search, anc, FuncTable(xs == 0)), ("kron", hs)]
    elif diffusion == "mcz":
        m = Z
        for _ in range(n_search - 1):
            m = CMat(m)
        xx = {q: X for q in search}
        ops += [("kron", hs), ("kron", xx), ("kron", {tuple(search): m}), ("kron", xx), ("kron", hs)]
    else:
        raise ValueError("diffusion must be 'f' or 'mcz'")
    return ops


def apply_ops(backend, ops) -> None:
    for op in ops:
        if op[0] == "kron":
            backend.kronselect_dot(op[1])
        else:
            backend.func_apply(op[1], op[2], op[3])


def grover_closed_form(n_search: int, x_star: int, iterations: int) -> numpy.ndarray:
    """State after `iterations` Grover iterations (either diffusion form): marked amplitude
    (-1)^t sin((2t+1) theta), others (-1)^t cos((2t+1) theta) / sqrt(N-1), sin theta = 2^(-n/2),
    times the |-> ancilla."""
    N = 2 ** n_search
    theta = numpy.arcsin(1.0 / numpy.sqrt(N))
    sign = (-1.0) ** iterations
    search = numpy.full(N, sign * numpy.cos((2 * iterations + 1) * theta) / numpy.sqrt(N - 1))
    search[x_star] = sign * numpy.sin((2 * iterations + 1) * theta)
    return numpy.kron(search, numpy.array([1.0, -1.0]) / numpy.sqrt(2.0))


# ---- batches of uneven layered random circuits -----------------------------------------------

CNOT4 = numpy.eye(4, dtype=numpy.complex128)[[0, 1, 3, 2]]        # control = first qubit (MSB)
CZ4 = numpy.diag([1.0, 1.0, 1.0, -1.0]).astype(numpy.complex128)


def layered_circuit_gates(n: int, depth: int, seed: int) -> List[Tuple[Tuple[int, ...], numpy.ndarray]]:
    """layered_random_circuit(n, depth, seed) as dense gates (controls first), in program order,
    with the Haar unitaries drawn in one call and factored by one stacked QR: the same random
    stream and the same LAPACK factorisation per matrix as haar_2x2."""
    rng = numpy.random.default_rng(seed)
    z = rng.standard_normal((depth * n, 2, 2, 2))          # per gate: real 2x2, then imaginary 2x2
    zc = (z[:, 0] + 1j * z[:, 1]) / numpy.sqrt(2.0)
    q, r = numpy.linalg.qr(zc)
    d = numpy.diagonal(r, axis1=1, axis2=2)
    us = q * (d / numpy.abs(d))[:, None, :]
    gates = []
    for layer in range(depth):
        for qb in range(n):
            gates.append(((qb,), us[layer * n + qb]))
        if layer % 2 == 0:
            gates.extend(((j, j + 1), CNOT4) for j in range(0, n - 1, 2))
        else:
            gates.extend(((j, j + 1), CZ4) for j in range(1, n - 1, 2))
    return gates


def batch_layered_circuits(count: int, seed: int, n_range: Tuple[int, int] = (10, 22),
                           depth_range: Tuple[int, int] = (8, 32)):
    """`count` layered random circuits with qubit counts and depths drawn uniformly from the
    inclusive ranges and one circuit seed each, all from numpy.random.default_rng(seed) (counts,
    then depths, then seeds). Returns (n list, depth list, seed list, PackedCircuits)."""
    from .batch import PackedCircuits
    rng = numpy.random.default_rng(seed)
    ns = rng.integers(n_range[0], n_range[1] + 1, size=count)
    depths = rng.integers(depth_range[0], depth_range[1] + 1, size=count)
    seeds = rng.integers(0, 2 ** 31 - 1, size=count)
    circuits = [layered_circuit_gates(int(n), int(d), int(s)) for n, d, s in zip(ns, depths, seeds)]
    packed = PackedCircuits.from_gates([int(n) for n in ns], circuits)
    return [int(n) for n in ns], [int(d) for d in depths], [int(s) for s in seeds], packed
