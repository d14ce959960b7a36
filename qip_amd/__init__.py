"""qip_amd: an MI355X (gfx950) state-vector engine for QIP graphs.

The HIP engine (qip_amd/csrc, C ABI in include/qipgpu.h) sits behind the StateType backend
interface of QIP; the graph frontend modules here (qip, operators, pipeline, qubit_util, qfft)
build the same graphs the reference does and run them on it:

    from qip_amd import Qubit, H, C, Swap, QFFT, run
    q = Qubit(n=16)
    out, classic = QFFT(q).run(feed={q: psi})          # GpuBackend, complex128

Backends: GpuBackend (plain complex128 / complex64 drop-in), FusedBackend (complex64 state,
gates fused into <= 6-qubit blocks applied with split-BF16 MFMA).
"""
from .backend import FuncTable, GpuBackend, StateType, infidelity, max_abs_diff, overlap
from .operators import (C, CMat, COp, F, FOp, H, HOp, MatrixOp, Not, NotOp, R, Rm, RmOp, ROp, Swap, SwapMat,
                        SwapOp, X, XOp, Y, YOp, Z, ZOp)
from .pipeline import PipelineObject, get_deps, make_circuit_mat, run, run_graph
from .qfft import QFFT
from .qip import Measure, OpConstructor, Q, Qubit, SplitQubit, StochasticMeasure
from .qubit_util import QubitFuncWrapper, QubitOpWrapper, QubitWrapperContext

__all__ = [
    "FuncTable", "GpuBackend", "StateType", "infidelity", "max_abs_diff", "overlap",
    "C", "CMat", "COp", "F", "FOp", "H", "HOp", "MatrixOp", "Not", "NotOp", "R", "Rm", "RmOp", "ROp",
    "Swap", "SwapMat", "SwapOp", "X", "XOp", "Y", "YOp", "Z", "ZOp",
    "PipelineObject", "get_deps", "make_circuit_mat", "run", "run_graph", "QFFT",
    "Measure", "OpConstructor", "Q", "Qubit", "SplitQubit", "StochasticMeasure",
    "QubitFuncWrapper", "QubitOpWrapper", "QubitWrapperContext",
]
