"""OracleBackend: a StateType running the CPU restatement of the reference kernels.

TEST INFRASTRUCTURE ONLY (see oracle/oracle.py). Mirrors CythonBackend (qip/backend.py:68-175)
method for method, including its state/arena double buffer and live get_state() buffer, so the
graph frontend can be tested on the CPU and GPU results can be compared with the reference's
algorithm on the same inputs.
"""
import os
import sys
from typing import Optional, Sequence

import numpy

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from qip_amd.backend import StateType  # noqa: E402

from . import oracle  # noqa: E402


class OracleBackend(StateType):
    # faithful=True restates the reference's func_apply rest-register bound and top-k heap
    faithful = False

    def __init__(self, n: int, state: numpy.ndarray, arena: numpy.ndarray):
        super().__init__(n, state)
        self.arena = arena

    @staticmethod
    def make_state(n, index_groups, feed_list, inputstartindex=None, inputendindex=None, outputstartindex=None,
                   outputendindex=None, statetype=numpy.complex128) -> "OracleBackend":
        state = oracle.make_state_vector(n, index_groups, feed_list, statetype=numpy.complex128)
        arena = numpy.zeros(2 ** n, dtype=numpy.complex128)
        return OracleBackend(n, state, arena)

    def get_state(self) -> numpy.ndarray:
        return self.state

    def _swap(self):
        self.state, self.arena = self.arena, self.state

    def kronselect_dot(self, mats, input_offset: int = 0, output_offset: int = 0) -> None:
        oracle.kronselect_dot(mats, self.state, self.n, self.arena, input_offset=input_offset,
                              output_offset=output_offset)
        self._swap()

    def func_apply(self, reg1_indices, reg2_indices, func, input_offset: int = 0, output_offset: int = 0) -> None:
        oracle.func_apply(reg1_indices, reg2_indices, func, self.state, self.n, self.arena,
                          faithful=self.faithful)
        self._swap()

    def total_prob(self) -> float:
        return oracle.prob_magnitude(self.state)

    def measure(self, indices, measured=None, measured_prob=None, input_offset=0, output_offset=0):
        bits, prob = oracle.measure(numpy.asarray(indices, dtype=numpy.int32), self.n, self.state, self.arena,
                                    measured=measured, measured_prob=measured_prob)
        self._swap()
        return bits, prob

    def reduce_measure(self, indices, measured=None, measured_prob=None, input_offset=0, output_offset=0):
        bits, prob = oracle.reduce_measure(numpy.asarray(indices, dtype=numpy.int32), self.n, self.state,
                                           self.arena, measured=measured, measured_prob=measured_prob)
        self._swap()
        return bits, prob

    def soft_measure(self, indices, measured: Optional[int] = None, input_offset: int = 0):
        return oracle.soft_measure(numpy.asarray(indices, dtype=numpy.int32), self.n, self.state,
                                   measured=measured, input_offset=input_offset)

    def measure_probabilities(self, indices: Sequence[int], top_k: int = 0):
        idx = numpy.asarray(indices, dtype=numpy.int32)
        if top_k:
            return oracle.measure_top_probabilities(idx, self.n, top_k, self.state, faithful=self.faithful)
        return oracle.measure_probabilities(idx, self.n, self.state)

    def get_state_size(self) -> int:
        return len(self.state)

    def get_relative_range(self, start: int, end: int):
        return self.state[start:end]

    def overwrite_relative_range(self, start: int, end: int, data):
        self.state[start:end] = data

    def addto_relative_range(self, start: int, end: int, data):
        self.state[start:end] += data
