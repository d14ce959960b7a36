"""Quantum Fourier transform circuit (counterpart of qip/qfft.py).

For qubits q_0..q_{n-1} (q_0 the most significant): for each offset t, H on q_t then, for every
i > t, a controlled phase R_{1+i-t} with control q_i and target q_t; with rev=True the register
is then reversed by swaps. Because the op wrapper drops call-time keyword arguments (see
qip_amd.qubit_util), makeR's negate=True is lost and the phases are e^{+2 pi i / 2^m}: on a
register fed with x the circuit yields numpy.fft.ifft(x, norm="ortho").
"""
from typing import MutableSequence, Tuple, Union

from .operators import C, H, Rm, Swap
from .qip import Qubit
from .qubit_util import QubitFuncWrapper
from .util import flatten


@QubitFuncWrapper.wrap
def QFFT(*inputs: Qubit, rev: bool = True) -> Union[Qubit, Tuple[Qubit, ...]]:
    qarr = flatten([inp.split(range(inp.n)) for inp in inputs])
    recQFFT(qarr)
    if rev:
        for i in range(len(qarr) // 2):
            qarr[i], qarr[-1 - i] = Swap(qarr[i], qarr[-1 - i])
    if len(inputs) < len(qarr):
        outputs = []
        pos = 0
        for qubit in inputs:
            outputs.append(Qubit(*(qarr[pos:pos + qubit.n])))
            pos += qubit.n
    else:
        outputs = qarr
    return tuple(outputs) if len(outputs) > 1 else outputs[0]


def recQFFT(qarr: MutableSequence[Qubit], offset: int = 0):
    # iterative form of the reference's tail recursion (qip/qfft.py:33-39)
    while len(qarr) > offset:
        qarr[offset] = H(qarr[offset])
        for i in range(offset + 1, len(qarr)):
            qarr[i], qarr[offset] = C(makeR(1 + i - offset))(qarr[i], qarr[offset])
        offset += 1


def makeR(m: int):
    return lambda *inputs, **kwargs: Rm(m, *inputs, negate=True, **kwargs)
