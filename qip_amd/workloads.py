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
    """Bytes a plain sweep must move: read + write of every amplitude a gate can change. A
    single-qubit gate changes all 2^n; CNOT (control 1) changes 2^(n-1); CZ changes only the
    amplitudes with both qubits 1, 2^(n-2)."""
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
