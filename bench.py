#!/usr/bin/env python3
"""Benchmark: 30-qubit gates/s and HBM GB/s, fused complex64 vs plain complex128, with the
infidelity of the fused result against the complex128 one.

Workload (one "step"): the 30-qubit depth-40 layered random circuit (Haar single-qubit unitary
on every qubit, then a CNOT / CZ brickwork; 1780 gates, seed fixed) applied to |0...0> by the
precision-form FusedBackend as a program drives it: reset the complex64 state, record the 1780
descriptors (they are fused into <= 6-qubit blocks as they arrive, and each block launches on the
split-BF16 MFMA kernel as soon as it is final), flush the rest, apply the final FP64
renormalisation, synchronise. `value` = 1780 * steps / timed seconds.

Beside it, in the same run:
  * the plain complex128 GpuBackend (the drop-in) on the same circuit, once, with its gates/s,
    effective GB/s (the gates' algorithmic bytes over the time) and the HBM roofline of its dominant
    kernel (k_tile_run: a held run of single-qubit gates in one pass, each gate applied by itself);
  * the complex128 fused form (Fused128Backend: the same blocks applied in FP64) on the same circuit,
    once, with its gates/s and its difference from the plain complex128 result;
  * 1 - F and max |amplitude difference| of the fused result against the complex128 result;
  * the HBM roofline of the fused kernel from HIP events on the state's stream, over one more run
    of the same circuit's blocks, composed up front and launched back to back;
  * the memory-pattern floor of those blocks: the same tile patterns swept in place with no
    arithmetic (qipgpu_fused_pattern_sweep), and the kernel's time over that floor;
  * cpu_baseline: the oracle's restatement of the reference's cdot_loop (oracle/liboracle_fast.so,
    the reference's -O3 -ffast-math -fopenmp flags, all visible cores) on a time-bounded sample of
    the circuit's own gates at full size, extrapolated to the circuit's gate mix.

The default run then measures every other configuration in the same process and nests each under
"configs" in the same JSON line, with its own rate, the roofline of its dominant kernel, a bounded
CPU baseline (--config-cpu-seconds of work) and its error figures: qft16 (16-qubit QFT graph),
circuit28 (28-qubit depth-20 circuit, plain complex128 vs unfused complex64), grover24 (24-qubit
Grover search in both forms), batch512 (512 uneven circuits in the batching form), measure28
(marginal probabilities, norm and measurement on a 28-qubit complex128 state), func28 (func_apply
on a 28-qubit complex128 state) and dense28 (dense operators on 4..8 targets). --workload NAME runs
one of them alone as its own JSON line (circuit30: the headline alone); --workload product24 times
the product sweep's pass shapes alone (not part of the default run).

--no-cpu-baseline leaves out the CPU leg (for profiler runs); it does not change the timed work.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0     # MI355X HBM3E, published
EFFECTIVE_NOTE = ("the gates' algorithmic bytes as written (2 S 2^(n-c) per gate) over the run time; held runs "
                  "of single-qubit or diagonal gates share HBM passes, so this can exceed the HBM peak")
MFMA_BF16_PEAK_TFLOPS = 2500.0   # MI355X dense BF16 matrix, published
FP64_PEAK_TFLOPS = 78.6          # MI355X FP64 (vector and matrix), published

# HBM traffic of one k_fused6 launch on a 30-qubit complex64 state (8 GiB), measured by two
# rocprofv3 --pmc runs of `bench.py --workload circuit30 --steps 1 --warmup 0 --no-cpu-baseline
# --no-plain --copy-calibration-gib 1` on the current kernel (FETCH_SIZE, then WRITE_SIZE; KiB per
# dispatch, mean over the 198 block launches of the run). Each counter is scaled by what it reads for
# the 1 GiB device-to-device copy in the same run: FETCH_SIZE reads 524314.6 KiB for 1048576 KiB moved
# (x2.0), WRITE_SIZE 1048576 KiB (x1.0). A PMC pass cannot run inside the timed process, so the
# figure is a constant here. Commands and per-kernel figures: profiles/k_fused6_circuit30_depth40.txt.
FUSED6_PMC_30Q = {"fetch_kib": 4196076.3, "fetch_scale": 1048576.0 / 524314.6,
                  "write_kib": 8388628.6, "write_scale": 1.0}


# HBM traffic of one k_marg_pass launch on a 28-qubit complex128 state (4 GiB), measured the same way
# by two rocprofv3 --pmc runs of `bench.py --workload measure28 --no-cpu-baseline
# --copy-calibration-gib 1` (mean over the 42 launches of the run; FETCH_SIZE reads 524313.75 KiB for
# the 1 GiB copy, WRITE_SIZE 1048576 KiB). Commands and figures: profiles/k_marg_pass_measure28.txt.
MARG_PMC_28Q = {"fetch_kib": 2097182.76, "fetch_scale": 1048576.0 / 524313.75,
                "write_kib": 3026.64, "write_scale": 1.0}


# HBM traffic of one k_tile_run pass (a held run of single-qubit gates) on a 28-qubit complex128 state
# (4 GiB), measured by two rocprofv3 --pmc runs of `bench.py --workload circuit28 --no-cpu-baseline
# --copy-calibration-gib 1` (mean over the run's 102 complex128 passes; FETCH_SIZE reads 524315.2 KiB
# for the 1 GiB copy, WRITE_SIZE 1048576 KiB): 1.0031x the pass's algorithmic bytes, the excess being the
# last prefetch of each workgroup, which re-reads a tile. Commands and figures:
# profiles/k_tile_run_circuit28_depth20.txt.
TILE_RUN_PMC_28Q = {"fetch_kib": 2110371.9, "fetch_scale": 1048576.0 / 524315.2,
                    "write_kib": 4194310.4, "write_scale": 1.0}


def pmc_bytes(p) -> int:
    return int((p["fetch_kib"] * p["fetch_scale"] + p["write_kib"] * p["write_scale"]) * 1024)


def fused6_traffic_bytes(n: int):
    """Measured HBM bytes per k_fused6 launch at n = 30 (None for other sizes)."""
    if n != 30:
        return None
    p = FUSED6_PMC_30Q
    return int((p["fetch_kib"] * p["fetch_scale"] + p["write_kib"] * p["write_scale"]) * 1024)


class HipEvents:
    """hipEvent timing on an arbitrary stream, through libamdhip64."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def create(self):
        ev = ctypes.c_void_p()
        if self.hip.hipEventCreate(ctypes.byref(ev)) != 0:
            raise RuntimeError("hipEventCreate failed")
        return ev

    def record(self, ev, stream: int):
        if self.hip.hipEventRecord(ev, ctypes.c_void_p(stream)) != 0:
            raise RuntimeError("hipEventRecord failed")

    def elapsed_ms(self, a, b) -> float:
        self.hip.hipEventSynchronize(b)
        ms = ctypes.c_float()
        if self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b) != 0:
            raise RuntimeError("hipEventElapsedTime failed")
        return float(ms.value)


def run_fused_step(be, ops):
    """One step: reset to |0..0>, record the circuit (blocks launch while the host records the
    rest), flush, finish, synchronise. Returns the number of blocks applied."""
    before = be.blocks_applied
    be.reinit()
    for d in ops:
        be.kronselect_dot(d)
    be.flush()
    be.fused_finish()
    be.synchronize()
    return be.blocks_applied - before


def fused_kernel_ms(be, ops, ev):
    """Average time per block pass of the fused kernel: the circuit's blocks (the recorder's light-cone
    rule restated by qip_amd.fused.fuse, composed up front) launched back to back between HIP events
    on the state's stream; the final scaling sweep counts as one more pass of the same bytes.
    Returns (ms per pass, blocks)."""
    from qip_amd.fused import descriptor_gates, fuse
    gates = [g for d in ops for g in descriptor_gates(d)]
    blocks = [b.matrix() for b in fuse(gates, be.fusion_width)]
    be.reinit()
    be.synchronize()
    t0, t1 = ev.create(), ev.create()
    ev.record(t0, be.stream)
    for qs, M in blocks:
        be.apply_fused(qs, M)
    be.fused_finish()
    ev.record(t1, be.stream)
    be.synchronize()
    return ev.elapsed_ms(t0, t1) / (len(blocks) + 1), len(blocks)


def fused_pattern_floor_ms(be, ops, ev):
    """Average time per block of the same blocks' memory pattern alone: every amplitude loaded and
    stored in place with each block's tile pattern and k_fused6's grid, no arithmetic
    (qipgpu_fused_pattern_sweep), launched back to back between HIP events."""
    from qip_amd.fused import descriptor_gates, fuse
    blocks = fuse([g for d in ops for g in descriptor_gates(d)], be.fusion_width)
    be.synchronize()
    t0, t1 = ev.create(), ev.create()
    ev.record(t0, be.stream)
    for b in blocks:
        be.fused_pattern_sweep(sorted(b.qubits))
    ev.record(t1, be.stream)
    be.synchronize()
    return ev.elapsed_ms(t0, t1) / len(blocks)


def plain_dominant_roofline(be, n, ev, amp: int = 16):
    """Average duration of the plain path's dominant kernel from HIP events on the state's stream:
    k_tile_run, one HBM pass that applies a held run of single-qubit gates (here 8 Haar gates on the
    8 highest state bits, the run a 28- or 30-qubit layer is cut into), each gate by itself on the
    tile in LDS. Algorithmic bytes per launch: one read and one write of the state."""
    import numpy
    from qip_amd.workloads import haar_2x2
    rng = numpy.random.default_rng(77)
    run = [{q: haar_2x2(rng)} for q in range(8)]
    reps = 5
    for d in run:                                  # warm-up pass
        be.kronselect_dot(d)
    be.flush()
    be.synchronize()
    a, b = ev.create(), ev.create()
    ev.record(a, be.stream)
    for _ in range(reps):
        for d in run:
            be.kronselect_dot(d)
        be.flush()
    ev.record(b, be.stream)
    ms = ev.elapsed_ms(a, b) / reps
    bytes_per = 2 * amp * 2 ** n
    kind = "double, complex128" if amp == 16 else "float, complex64"
    return {"kernel": "k_tile_run<{}> (one pass, 8 held single-qubit gates)".format(kind), "avg_ms": ms,
            "bytes_per_launch": bytes_per, "gates_per_launch": len(run),
            "achieved_gbs": bytes_per / (ms * 1e-3) / 1e9}


def cpu_baseline(ops, n, seconds: float):
    """Oracle cdot restatement (reference flags, all visible cores) on the circuit's first gates at
    full size until `seconds` of work, alternating single- and two-qubit gates; extrapolated to
    the whole circuit's mix."""
    import numpy
    from oracle import oracle
    from qip_amd.workloads import gate_counts
    cores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    vec = numpy.zeros(2 ** n, dtype=numpy.complex128)
    out = numpy.zeros(2 ** n, dtype=numpy.complex128)
    vec.fill(0)          # touch every page before timing
    out.fill(0)
    vec[0] = 1.0
    one = [d for d in ops if isinstance(next(iter(d)), int)]
    two = [d for d in ops if not isinstance(next(iter(d)), int)]
    t_one, t_two = [], []
    start = time.perf_counter()
    i = 0
    while (time.perf_counter() - start < seconds or not t_one or not t_two) and i < 64:
        d = one[i // 2] if i % 2 == 0 else two[i // 2]
        t = time.perf_counter()
        oracle.kronselect_dot(d, vec, n, out, fast=True, nthreads=cores)
        (t_one if i % 2 == 0 else t_two).append(time.perf_counter() - t)
        vec, out = out, vec
        i += 1
    n1, n2 = gate_counts(ops)
    circuit_s = n1 * (sum(t_one) / len(t_one)) + n2 * (sum(t_two) / len(t_two))
    return {
        "value": (n1 + n2) / circuit_s, "unit": "gates/s", "cores": cores, "kind": "port",
        "sample": "{} single-qubit + {} two-qubit gates of this {}-qubit circuit at full size on the oracle's "
                  "cdot_loop restatement (-O3 -ffast-math -fopenmp), {:.1f} s of work; rate extrapolated to "
                  "the circuit's {} + {} gate mix".format(len(t_one), len(t_two), n, sum(t_one) + sum(t_two), n1, n2),
    }


def copy_calibration(ev, nbytes: int) -> None:
    hip = ev.hip
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), nbytes) or hip.hipMalloc(ctypes.byref(b), nbytes):
        raise MemoryError("calibration buffers")
    if hip.hipMemcpy(b, a, nbytes, 3):          # hipMemcpyDeviceToDevice
        raise RuntimeError("calibration copy failed")
    hip.hipFree(a)
    hip.hipFree(b)


def _line(metric, value, unit, steps, warmup, ms, dtype, config, **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
           "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
           "data": "synthetic", "config": config}
    out.update(extra)
    return out


def bench_circuit28(args):
    """28-qubit, 20-layer random circuit: plain complex128 vs plain complex64 (unfused: every gate
    applied by itself, with the bits of a sweep per gate); 1 - F of complex64 against complex128."""
    import numpy
    from qip_amd import GpuBackend, infidelity, max_abs_diff
    from qip_amd.workloads import algorithmic_bytes_plain, layered_random_circuit
    n, depth = 28, 20
    ops = layered_random_circuit(n, depth, args.seed)
    res = {}
    states = {}
    for name, st, amp in (("c128", numpy.complex128, 16), ("c64", numpy.complex64, 8)):
        be = GpuBackend.make_state(n, [], [], statetype=st)
        for d in ops[:45]:                     # warm-up: first layer
            be.kronselect_dot(d)
        be.reinit()
        be.synchronize()
        t = time.perf_counter()
        for d in ops:
            be.kronselect_dot(d)
        be.synchronize()
        dt = time.perf_counter() - t
        gb = algorithmic_bytes_plain(ops, n, amp) / 1e9
        res[name] = {"gates_per_s": len(ops) / dt, "s_per_circuit": dt, "effective_gbs": gb / dt,
                     "effective_gbs_note": EFFECTIVE_NOTE}
        states[name] = be
    res["infidelity_c64_vs_c128"] = infidelity(states["c128"], states["c64"])
    res["max_abs_err_c64_vs_c128"] = max_abs_diff(states["c128"], states["c64"])
    ev = HipEvents()
    dom = {name: plain_dominant_roofline(states[name], n, ev, amp) for name, amp in (("c128", 16), ("c64", 8))}
    for be in states.values():
        be.close()
    cfg = {"workload": "28-qubit depth-20 layered random circuit, plain complex128 vs plain complex64 (unfused: "
                       "each gate applied by itself, bit for bit a sweep per gate; held runs of single-qubit gates "
                       "share one HBM pass per tile, runs of diagonal gates one sweep)", "n_qubits": n,
           "depth": depth, "gates": len(ops), "seed": args.seed}
    d = dom["c128"]
    out = _line("28-qubit gates/s, plain c128 vs unfused c64", res["c128"]["gates_per_s"], "gates/s", 1, 1,
                res["c128"]["s_per_circuit"] * 1e3, "c128", cfg, plain_c128=res["c128"], plain_c64=res["c64"],
                infidelity_c64_vs_c128=res["infidelity_c64_vs_c128"],
                max_abs_err_c64_vs_c128=res["max_abs_err_c64_vs_c128"],
                roofline={"bound": "hbm", "kernel": d["kernel"], "achieved": d["achieved_gbs"], "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": d["achieved_gbs"] / HBM_PEAK_GBS,
                          "traffic": pmc_bytes(TILE_RUN_PMC_28Q) if n == 28 else None,
                          "bytes_per_launch": d["bytes_per_launch"], "avg_launch_ms": d["avg_ms"]},
                roofline_c64=dict(dom["c64"], frac=dom["c64"]["achieved_gbs"] / HBM_PEAK_GBS))
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(ops, n, args.cpu_seconds)
    return out


def bench_grover24(args):
    """24-qubit iterative Grover (23 search qubits + ancilla, 2274 iterations): plain complex128
    with H/X/multi-controlled-Z diffusion, the fused complex64 backend with F diffusion, and the
    complex128 fused form with H/X/multi-controlled-Z diffusion; then a seeded measurement of the
    register. Amplitude error against the closed form."""
    import random
    import numpy
    from qip_amd import GpuBackend
    from qip_amd.fused import FusedBackend
    from qip_amd.workloads import (apply_ops, grover_closed_form, grover_initial, grover_iteration_ops,
                                   grover_iterations)
    ns = 23
    x_star = int(numpy.random.default_rng(args.seed).integers(0, 2 ** ns))
    t_iter = grover_iterations(ns)
    expect = grover_closed_form(ns, x_star, t_iter)
    groups, feeds = grover_initial(ns)
    res = {}
    from qip_amd.fused import Fused128Backend
    for name, make, diff in (("plain_c128", GpuBackend.make_state, "mcz"),
                             ("fused_c64", FusedBackend.make_state, "f"),
                             ("fused_c128", Fused128Backend.make_state, "mcz")):
        be = make(ns + 1, groups, feeds)
        ops = grover_iteration_ops(ns, x_star, diff)
        apply_ops(be, ops)                      # warm-up iteration (tables, kernels)
        be.reinit(groups, feeds)
        be.synchronize()
        t = time.perf_counter()
        for _ in range(t_iter):
            apply_ops(be, ops)
        be.synchronize()
        dt = time.perf_counter() - t
        err = float(numpy.max(numpy.abs(be.get_state() - expect)))
        random.seed(args.seed)
        m, p = be.measure(list(range(ns)))
        amp = 8 if name == "fused_c64" else 16
        alg, touched = grover_iteration_bytes(ns, amp, diff)
        res[name] = {"iterations_per_s": t_iter / dt, "s_total": dt, "diffusion": diff,
                     "max_abs_err_vs_closed_form": err, "measured": m, "measured_prob": p,
                     "found_marked": m == x_star,
                     "algorithmic_gbs": alg * t_iter / dt / 1e9, "hbm_frac": alg * t_iter / dt / 1e9 / HBM_PEAK_GBS,
                     "touched_gbs": touched * t_iter / dt / 1e9}
        if name != "fused_c128":
            res[name]["product_pass"] = product_pass_roofline(be, ns, amp)
        else:
            res[name]["note"] = ("the same H/X/MCZ iteration on Fused128Backend: each run of H and X layers "
                                 "is one product of per-qubit 2x2s (3 product passes instead of 6)")
        be.close()
    cfg = {"workload": "24-qubit iterative Grover: 23 search qubits + |-> ancilla, F(x==x*) oracle, {} iterations, "
                       "seeded measurement".format(t_iter), "n_qubits": ns + 1, "iterations": t_iter,
           "x_star": x_star, "seed": args.seed}
    r = res["plain_c128"]["product_pass"]
    out = _line("24-qubit Grover iterations/s (plain c128, H/X/MCZ diffusion)", res["plain_c128"]["iterations_per_s"],
                "iterations/s", 1, 1, res["plain_c128"]["s_total"] * 1e3, "c128", cfg, roofline=dict(r, traffic=None),
                **res)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_grover(ns, x_star, args.cpu_seconds)
    return out


def product_pass_roofline(be, ns: int, amp: int):
    """The Grover iteration's dominant kernel, the product sweep (k_product): an H layer on the ns
    search qubits is three passes over the whole state; HIP events on the state's stream around ten
    such layers (an even count, so the state is unchanged)."""
    import numpy
    from qip_amd import GpuBackend
    H = numpy.array([[1, 1], [1, -1]]) / numpy.sqrt(2)
    targets, mats = list(range(ns)), [H] * ns
    ev = HipEvents()
    be.flush() if hasattr(be, "flush") else None
    GpuBackend.apply_product(be, targets, mats)        # launched directly, also on a recording backend
    GpuBackend.apply_product(be, targets, mats)
    be.synchronize()
    a, b = ev.create(), ev.create()
    ev.record(a, be.stream)
    for _ in range(10):
        GpuBackend.apply_product(be, targets, mats)
    ev.record(b, be.stream)
    passes = 10 * -(-(ns - 3) // 8)
    ms = ev.elapsed_ms(a, b) / passes
    bytes_per = 2 * amp * 2 ** (ns + 1)
    gbs = bytes_per / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": "k_product<{}> (single-qubit factors of an H layer, up to 12 per pass)".format(
                "double" if amp == 16 else "float"), "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": gbs / HBM_PEAK_GBS, "bytes_per_launch": bytes_per, "avg_launch_ms": ms}


def grover_iteration_bytes(ns: int, amp: int, diffusion: str):
    """(algorithmic, touched) HBM bytes of one Grover iteration on ns search qubits + 1 ancilla.
    Algorithmic, per descriptor as the reference sweeps it: read + write of every amplitude a
    descriptor can change (2 S 2^n; the 22-controlled Z changes 2^2), plus the 4 B x 2^ns f table per
    func_apply. Touched, what the kernels here move: each H or X layer on 23 qubits is three product
    passes over the whole state (every pass keeps the four lowest state bits in its tile), func_apply
    reads the table and swaps only the pairs where f(x) = 1 (one, for these oracles)."""
    n = ns + 1
    sweep = 2 * amp * 2 ** n
    table = 4 * 2 ** ns
    layers = 4 if diffusion == "mcz" else 2           # H, X, X, H  or  H, H
    funcs = 1 if diffusion == "mcz" else 2            # the oracle (+ F(x == 0) for F diffusion)
    mcz = 2 * amp * 2 ** (n - (ns - 1)) if diffusion == "mcz" else 0
    algorithmic = funcs * (sweep + table) + layers * sweep + mcz
    passes = -(-(ns - 3) // 8)                        # three low qubits ride in the first pass
    touched = funcs * (table + 2 * 2 * amp) + layers * passes * sweep + mcz
    return algorithmic, touched


def cpu_baseline_grover(ns, x_star, seconds):
    """Oracle restatement of the reference on Grover iterations at full size with the operations
    the reference can run (per-qubit H descriptors and F-based diffusion: one H descriptor on the
    whole register would cost 2^(n+23) in cdot_loop); per-iteration rate."""
    import numpy
    from oracle import oracle
    n = ns + 1
    cores = len(os.sched_getaffinity(0))
    from qip_amd.workloads import H, grover_initial
    groups, feeds = grover_initial(ns)
    vec = oracle.make_state_vector(n, groups, feeds)
    out = numpy.zeros_like(vec)
    xs = numpy.arange(2 ** ns)
    t_oracle = (xs == x_star).astype(numpy.int32)
    t_zero = (xs == 0).astype(numpy.int32)
    search, anc = list(range(ns)), [ns]
    done, start = 0, time.perf_counter()
    while time.perf_counter() - start < seconds or done == 0:
        for table in (t_oracle, None, t_zero, None):
            if table is None:
                for q in search:
                    oracle.kronselect_dot({q: H}, vec, n, out, fast=True, nthreads=cores)
                    vec, out = out, vec
            else:
                oracle.func_apply(search, anc, None, vec, n, out, table=table, fast=True, nthreads=cores)
                vec, out = out, vec
        done += 1
    dt = time.perf_counter() - start
    return {"value": done / dt, "unit": "iterations/s", "cores": cores, "kind": "port",
            "sample": "{} Grover iteration(s) at full size (24 qubits) on the oracle restatement: func_apply oracle, "
                      "23 per-qubit H descriptors, func_apply F(x==0), 23 per-qubit H, {:.1f} s".format(done, dt)}


def bench_qft16(args):
    """16-qubit QFT (qip/qfft.py circuit, 144 gates) through the graph frontend on a seeded random
    state: plain complex128; error against ifft(norm='ortho') and the probabilities."""
    import numpy
    from qip_amd import GpuBackend, QFFT, Qubit, StochasticMeasure, run
    n = 16
    rng = numpy.random.default_rng(args.seed)
    psi = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    psi /= numpy.linalg.norm(psi)
    q = Qubit(n=n)
    sm = StochasticMeasure(QFFT(q))
    # a run takes about a millisecond: 5 untimed runs (clocks, first uses), then 100 timed ones
    warm, reps = 5, 100
    for _ in range(warm):
        run(sm, feed={q: psi}, backend_constructor=GpuBackend.make_state)
    t = time.perf_counter()
    for _ in range(reps):
        state, classic = run(sm, feed={q: psi}, backend_constructor=GpuBackend.make_state)
    dt = (time.perf_counter() - t) / reps
    # the same run after retiring the stored schedule: the graph walked again, as on a first run
    from qip_amd.pipeline import graph_changed
    graph_changed()
    t = time.perf_counter()
    run(sm, feed={q: psi}, backend_constructor=GpuBackend.make_state)
    walk_ms = (time.perf_counter() - t) * 1e3
    expect = numpy.fft.ifft(psi, norm="ortho")
    err = float(numpy.max(numpy.abs(state - expect)))
    # the same runs on the complex128 fused form (blocks of up to 6 qubits applied in FP64)
    from qip_amd.fused import Fused128Backend
    for _ in range(warm):
        run(sm, feed={q: psi}, backend_constructor=Fused128Backend.make_state)
    t = time.perf_counter()
    for _ in range(reps):
        f_state, _ = run(sm, feed={q: psi}, backend_constructor=Fused128Backend.make_state)
    f_dt = (time.perf_counter() - t) / reps
    fused128 = {"circuits_per_s": 1.0 / f_dt, "ms_per_run": f_dt * 1e3,
                "max_abs_err_vs_ifft": float(numpy.max(numpy.abs(f_state - expect)))}
    # the circuit's most frequent gate: 120 controlled phases, applied back to back between HIP events on a
    # state of the same size (the engine applies each run of them sharing a control in one k_diag_chain
    # sweep over the half of the state where that control is 1)
    be = GpuBackend.make_state(n, [list(range(n))], [psi])
    ev = HipEvents()
    from qip_amd import CMat
    descs = [{(i, t): CMat(numpy.diag([1.0, numpy.exp(2j * numpy.pi / 2.0 ** (1 + i - t))]))}
             for t in range(n) for i in range(t + 1, n)]
    for d in descs[:8]:
        be.kronselect_dot(d)
    be.synchronize()
    a, b = ev.create(), ev.create()
    ev.record(a, be.stream)
    for d in descs:
        be.kronselect_dot(d)
    be.flush()
    ev.record(b, be.stream)
    diag_ms = ev.elapsed_ms(a, b) / len(descs)
    be.close()
    diag_bytes = 2 * 16 * 2 ** (n - 2)
    perm = [int(format(i, "016b")[::-1], 2) for i in range(2 ** n)]
    perr = float(numpy.max(numpy.abs(classic[sm] - (numpy.abs(expect) ** 2)[perm])))
    cfg = {"workload": "16-qubit QFT (qip/qfft.py: 16 H, 120 controlled phases, 8 swaps) + StochasticMeasure, graph "
                       "frontend, seeded random complex128 state", "n_qubits": n, "gates": 144, "seed": args.seed,
           "frontend": "visiting order and index groups stored by the warm-up run and replayed; the 144 gates "
                       "launched as one HIP graph recorded on an earlier run (the state buffer is reused); the "
                       "state upload, every gate, the measurement and the read-back run each time"}
    diag_gbs = diag_bytes / (diag_ms * 1e-3) / 1e9
    out = _line("16-qubit QFT circuits/s (graph run incl. state upload and read-back)", 1.0 / dt, "circuits/s", reps,
                warm, dt * 1e3, "c128", cfg, max_abs_err_vs_ifft=err, max_abs_err_probabilities=perr,
                ms_first_run_walking_graph=walk_ms, fused_c128=fused128,
                roofline={"bound": "launch", "kernel": "k_diag_chain<double> (controlled phases, 120 of the 144 gates, per gate)",
                          "achieved": diag_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": diag_gbs / HBM_PEAK_GBS,
                          "traffic": None, "bytes_per_launch": diag_bytes, "avg_launch_ms": diag_ms,
                          "note": "a 1 MiB state stays in cache: each launch is bound by launch latency, not HBM"})
    if not args.no_cpu_baseline:
        from oracle.oracle_backend import OracleBackend
        t = time.perf_counter()
        run(sm, feed={q: psi}, backend_constructor=OracleBackend.make_state)
        cdt = time.perf_counter() - t
        out["cpu_baseline"] = {"value": 1.0 / cdt, "unit": "circuits/s", "cores": len(os.sched_getaffinity(0)),
                               "kind": "port", "sample": "one full run of the same graph on the oracle backend"}
    return out


def bench_measure28(args):
    """Measurement reductions on a 28-qubit complex128 state (4 GiB; H then one Haar unitary on every
    qubit): marginal probabilities over 8 high, 8 low and 8 scattered qubits, 12 low, one high and one
    low qubit (measure_probabilities, one pass over the state each); prob_magnitude (total_prob); and
    measure and then reduce_measure on the 8 scattered qubits (norm, marginals, collapse or tiled
    reduction). Algorithmic bytes per call as defined for these functions: S 2^n + 8 2^k for the
    probabilities, S 2^n for the norm, 4 S 2^n for measure and reduce_measure. Last, the initial state
    (make_state's product of fed vectors, built on the device): unfed |0...0>, 28 one-qubit feeds, four
    seven-qubit feeds; S 2^n bytes written."""
    import random
    import numpy
    from qip_amd import GpuBackend
    from qip_amd.workloads import random_product_unitaries
    n, amp = 28, 16
    H = numpy.array([[1, 1], [1, -1]]) / numpy.sqrt(2)
    be = GpuBackend.make_state(n, [], [])
    be.kronselect_dot({q: H for q in range(n)})
    for d in random_product_unitaries(n, seed=args.seed):
        be.kronselect_dot(d)
    be.synchronize()
    ev = HipEvents()
    state_bytes = amp * 2 ** n
    reps = max(3, args.steps)

    def timed(call):
        call()
        be.synchronize()
        t0, t1 = ev.create(), ev.create()
        w = time.perf_counter()
        ev.record(t0, be.stream)
        for _ in range(reps):
            call()
        ev.record(t1, be.stream)
        be.synchronize()
        return ev.elapsed_ms(t0, t1) / reps, (time.perf_counter() - w) * 1e3 / reps

    cases = {"high8": list(range(8)), "low8": list(range(n - 8, n)), "scattered8": [0, 5, 9, 13, 17, 21, 25, 27],
             "low12": list(range(n - 12, n)), "high1": [0], "low1": [n - 1]}
    probs = {}
    for name, idx in cases.items():
        ms, wall = timed(lambda idx=idx: be.measure_probabilities(idx))
        b = state_bytes + 8 * 2 ** len(idx)
        probs[name] = {"qubits": idx, "ms": ms, "wall_ms": wall, "algorithmic_gbs": b / ms / 1e6,
                       "hbm_frac": b / ms / 1e6 / HBM_PEAK_GBS}
    ms, wall = timed(lambda: be.total_prob())
    norm = {"ms": ms, "wall_ms": wall, "algorithmic_gbs": state_bytes / ms / 1e6,
            "hbm_frac": state_bytes / ms / 1e6 / HBM_PEAK_GBS}
    random.seed(args.seed)
    ms, wall = timed(lambda: be.measure(cases["scattered8"]))
    meas = {"qubits": cases["scattered8"], "ms": ms, "wall_ms": wall, "algorithmic_gbs": 4 * state_bytes / ms / 1e6,
            "hbm_frac": 4 * state_bytes / ms / 1e6 / HBM_PEAK_GBS}
    # reduce_measure last (it rewrites the state with the reduced one, tiled over the arena)
    ms, wall = timed(lambda: be.reduce_measure(cases["scattered8"]))
    red = {"qubits": cases["scattered8"], "ms": ms, "wall_ms": wall, "algorithmic_gbs": 4 * state_bytes / ms / 1e6,
           "hbm_frac": 4 * state_bytes / ms / 1e6 / HBM_PEAK_GBS}
    # the initial state (make_state's product of fed vectors, built on the device in place)
    rng = numpy.random.default_rng(args.seed)

    def rand_vec(k):
        v = rng.standard_normal(2 ** k) + 1j * rng.standard_normal(2 ** k)
        return v / numpy.linalg.norm(v)
    inits = {"zero": ([], []),
             "28_one_qubit_feeds": ([[q] for q in range(n)], [rand_vec(1) for _ in range(n)]),
             "4_seven_qubit_feeds": ([list(range(7 * g, 7 * g + 7)) for g in range(4)], [rand_vec(7) for _ in range(4)])}
    init = {}
    for name, (groups, feeds) in inits.items():
        ms, wall = timed(lambda groups=groups, feeds=feeds: be.reinit(groups, feeds))
        init[name] = {"ms": ms, "wall_ms": wall, "algorithmic_gbs": state_bytes / ms / 1e6,
                      "hbm_frac": state_bytes / ms / 1e6 / HBM_PEAK_GBS}
    eight = [probs[c] for c in ("high8", "low8", "scattered8")]
    value = sum(c["algorithmic_gbs"] for c in eight) / 3
    avg_ms = sum(c["ms"] for c in eight) / 3
    cfg = {"workload": "28-qubit complex128 state (H, then a Haar unitary on every qubit): measure_probabilities "
                       "over 8 high / 8 low / 8 scattered qubits (value: their mean), 12 low, 1 high, 1 low; "
                       "total_prob; measure and reduce_measure on 8 scattered qubits", "n_qubits": n,
           "seed": args.seed}
    out = _line("28-qubit c128 marginal probabilities over 8 qubits: algorithmic GB/s", value, "GB/s", reps, 1,
                avg_ms, "f64", cfg, probabilities=probs, total_prob=norm, measure=meas, reduce_measure=red,
                initial_state=init,
                roofline={"bound": "hbm", "kernel": "k_marg_pass (one pass over the state, FP64 sums)",
                          "achieved": value, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": value / HBM_PEAK_GBS,
                          "traffic": int((MARG_PMC_28Q["fetch_kib"] * MARG_PMC_28Q["fetch_scale"] +
                                          MARG_PMC_28Q["write_kib"] * MARG_PMC_28Q["write_scale"]) * 1024),
                          "bytes_per_launch": state_bytes + 8 * 2 ** 8, "avg_launch_ms": avg_ms,
                          "note": "time per call: k_marg_pass, the fold of its parts and the copy of the 2^k "
                                  "results to the host (the kernel alone is in profiles/k_marg_pass_measure28.txt)"})
    if not args.no_cpu_baseline:
        from oracle import oracle
        v = be.get_state()
        t = time.perf_counter()
        oracle.measure_probabilities(cases["scattered8"], n, v)
        cdt = time.perf_counter() - t
        out["cpu_baseline"] = {"value": (state_bytes + 8 * 2 ** 8) / cdt / 1e9, "unit": "GB/s", "cores": 1,
                               "kind": "port", "sample": "one full-size call of the oracle's measure_probabilities "
                                                         "(kronprod.pyx:238-261 restated, serial) on the scattered "
                                                         "8 qubits, {:.1f} s".format(cdt)}
    return out


def bench_func28(args):
    """func_apply on a 28-qubit complex128 state (4 GiB; H on every qubit): f(x) = x mod 251 (dense:
    almost every amplitude moves) with reg1 on the 20 highest state bits and reg2 on the 8 lowest,
    the reverse, and scattered registers; and a Grover-style oracle (reg1 = 27 qubits, f = [x == x*]:
    one pair moves). Algorithmic bytes per call as defined for func_apply: 2 S 2^n + 4 2^n1 (no
    credit for amplitudes that do not move)."""
    import numpy
    from qip_amd import GpuBackend
    from qip_amd.backend import FuncTable
    n, amp = 28, 16
    H = numpy.array([[1, 1], [1, -1]]) / numpy.sqrt(2)
    be = GpuBackend.make_state(n, [], [])
    be.kronselect_dot({q: H for q in range(n)})
    be.synchronize()
    ev = HipEvents()
    reps = max(3, args.steps)
    cases = {"reg1_high20_reg2_low8": (list(range(20)), list(range(20, 28)), "mod"),
             "reg1_low20_reg2_high8": (list(range(8, 28)), list(range(8)), "mod"),
             "scattered": (list(range(4, 24)), list(range(0, 4)) + list(range(24, 28)), "mod"),
             "grover_oracle_27_1": (list(range(27)), [27], "one")}
    out_cases = {}
    for name, (r1, r2, kind) in cases.items():
        x = numpy.arange(2 ** len(r1), dtype=numpy.int64)
        vals = x % 251 if kind == "mod" else (x == 12345).astype(numpy.int64)
        tab = FuncTable(vals.astype(numpy.int32))
        be.func_apply(r1, r2, tab)
        be.synchronize()
        t0, t1 = ev.create(), ev.create()
        ev.record(t0, be.stream)
        for _ in range(reps):
            be.func_apply(r1, r2, tab)
        ev.record(t1, be.stream)
        be.synchronize()
        ms = ev.elapsed_ms(t0, t1) / reps
        b = 2 * amp * 2 ** n + 4 * 2 ** len(r1)
        out_cases[name] = {"reg1": r1, "reg2": r2, "ms": ms, "algorithmic_gbs": b / ms / 1e6,
                           "hbm_frac": b / ms / 1e6 / HBM_PEAK_GBS}
    dense = [out_cases[c] for c in ("reg1_high20_reg2_low8", "reg1_low20_reg2_high8", "scattered")]
    value = sum(c["algorithmic_gbs"] for c in dense) / 3
    cfg = {"workload": "func_apply on a 28-qubit complex128 state: f(x) = x mod 251 with three register layouts "
                       "(value: their mean), and a Grover-style oracle on 27 + 1 qubits", "n_qubits": n}
    out = _line("28-qubit c128 func_apply (dense f): algorithmic GB/s", value, "GB/s", reps, 1,
                sum(c["ms"] for c in dense) / 3, "c128", cfg, cases=out_cases,
                roofline={"bound": "hbm", "kernel": "k_func_tile (LDS-tiled XOR permutation)", "achieved": value,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": value / HBM_PEAK_GBS, "traffic": None})
    if not args.no_cpu_baseline:
        from oracle import oracle
        r1, r2, _ = cases["reg1_high20_reg2_low8"]
        v = be.get_state()
        o = numpy.zeros_like(v)
        tabv = (numpy.arange(2 ** 20) % 251).astype(numpy.int32)
        t = time.perf_counter()
        oracle.func_apply(r1, r2, None, v, n, o, faithful=False, table=tabv, fast=True)
        cdt = time.perf_counter() - t
        out["cpu_baseline"] = {"value": (2 * amp * 2 ** n + 4 * 2 ** 20) / cdt / 1e9, "unit": "GB/s",
                               "cores": len(os.sched_getaffinity(0)), "kind": "port",
                               "sample": "one full-size call of the oracle's func_apply restatement "
                                         "(func_apply.pyx:11-112, -O3 -fopenmp build, all cores) on the first "
                                         "layout, {:.2f} s".format(cdt)}
    return out


def bench_dense28(args):
    """Dense operators on 4..6 targets (k_dense_mfma) on a 28-qubit complex128 state (4 GiB): the
    targets on the lowest, the highest and scattered state bits, and one 6-target operator under a
    control; then 8 targets (the FP64 MFMA GEMM, k_dense_gemm + scatter) on 26 qubits. Per call: HIP-event
    time on the state's stream, algorithmic bytes 2 S 2^(n-c), FP64 flops 8 2^K 2^(n-c) (a complex
    matrix-vector product counted the usual way); the MFMA tiles and the GEMM execute 6 2^K 2^(n-c)
    of them, with three real products per complex one."""
    import numpy
    from qip_amd import CMat, GpuBackend
    ev = HipEvents()
    rng = numpy.random.default_rng(args.seed)

    def haar(d):
        z = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / numpy.sqrt(2)
        q, r = numpy.linalg.qr(z)
        return q * (numpy.diagonal(r) / numpy.abs(numpy.diagonal(r)))

    reps = max(3, args.steps)
    cases = {}

    def run_case(be, n, name, qubits, controls, K):
        m = haar(2 ** K)
        for _ in range(controls):
            m = CMat(m)
        d = {tuple(qubits): m}
        be.kronselect_dot(d)
        be.synchronize()
        a, b = ev.create(), ev.create()
        ev.record(a, be.stream)
        for _ in range(reps):
            be.kronselect_dot(d)
        ev.record(b, be.stream)
        ms = ev.elapsed_ms(a, b) / reps
        nb = 2 * 16 * 2 ** (n - controls)
        fl = 8 * 2 ** K * 2 ** (n - controls)
        ex = 6 * 2 ** K * 2 ** (n - controls)
        cases[name] = {"n": n, "qubits": list(qubits), "controls": controls, "targets": K, "ms": ms,
                       "algorithmic_gbs": nb / ms / 1e6, "hbm_frac": nb / ms / 1e6 / HBM_PEAK_GBS,
                       "fp64_tflops": fl / ms / 1e9, "fp64_frac": fl / ms / 1e9 / FP64_PEAK_TFLOPS,
                       "fp64_executed_tflops": ex / ms / 1e9}

    n = 28
    be = GpuBackend.make_state(n, [], [])
    be.kronselect_dot({q: numpy.array([[1, 1], [1, -1]]) / numpy.sqrt(2) for q in range(n)})
    for K in (4, 5, 6):
        run_case(be, n, "k{}_low".format(K), list(range(n - K, n)), 0, K)
        run_case(be, n, "k{}_high".format(K), list(range(K)), 0, K)
        run_case(be, n, "k{}_scattered".format(K), [int(q) for q in numpy.linspace(1, n - 2, K).round()], 0, K)
    run_case(be, n, "k6_controlled", [3, 7, 11, 15, 19, 23, 26], 1, 6)
    be.close()
    be = GpuBackend.make_state(26, [], [])
    be.kronselect_dot({q: numpy.array([[1, 1], [1, -1]]) / numpy.sqrt(2) for q in range(26)})
    run_case(be, 26, "k8_gemm_scattered", [0, 3, 6, 9, 12, 15, 18, 25], 0, 8)
    be.close()
    tiles = [c for k, c in cases.items() if k.startswith(("k4", "k5", "k6")) and c["controls"] == 0]
    value = sum(c["algorithmic_gbs"] for c in tiles) / len(tiles)
    cfg = {"workload": "dense operators on a 28-qubit complex128 state: 4, 5 and 6 targets on low, high and "
                       "scattered qubits (value: their mean), 6 targets under one control; 8 targets on 26 qubits",
           "seed": args.seed}
    k6 = [cases[k] for k in ("k6_low", "k6_high", "k6_scattered")]
    k6_gbs = sum(c["algorithmic_gbs"] for c in k6) / 3
    out = _line("28-qubit c128 dense 4-6 target operators: algorithmic GB/s", value, "GB/s", reps, 1,
                sum(c["ms"] for c in tiles) / len(tiles), "c128", cfg, cases=cases,
                roofline={"bound": "hbm", "kernel": "k_dense_mfma<double,K> (K = 4..6: FP64 MFMA, three real "
                                                    "products per complex one)", "achieved": value,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": value / HBM_PEAK_GBS, "traffic": None,
                          "k6_fp64": {"achieved_gbs": k6_gbs,
                                      "fp64_tflops": sum(c["fp64_tflops"] for c in k6) / 3,
                                      "fp64_frac": sum(c["fp64_frac"] for c in k6) / 3,
                                      "fp64_executed_tflops": sum(c["fp64_executed_tflops"] for c in k6) / 3,
                                      "note": "at 6 targets the FP64 arithmetic (512 flop per 32 B moved counted "
                                              "the usual way, 384 executed) is the bound, not HBM"}})
    if not args.no_cpu_baseline:
        from oracle import oracle
        nc = 22
        v = numpy.zeros(2 ** nc, dtype=numpy.complex128)
        v[0] = 1.0
        o = numpy.zeros_like(v)
        d = {tuple(range(nc - 4, nc)): haar(16)}
        cores = len(os.sched_getaffinity(0))
        t = time.perf_counter()
        oracle.kronselect_dot(d, v, nc, o, fast=True, nthreads=cores)
        cdt = time.perf_counter() - t
        out["cpu_baseline"] = {"value": 2 * 16 * 2 ** nc / cdt / 1e9, "unit": "GB/s", "cores": cores, "kind": "port",
                               "sample": "one 4-target dense descriptor on a 22-qubit state on the oracle's cdot_loop "
                                         "restatement (-O3 -ffast-math -fopenmp, all cores), {:.2f} s".format(cdt)}
    return out


PRODUCT_CASES = {
    # name: (low state bits with a factor, number of high targets from the top down, factor)
    "h4": ((), 4, "H"),            # tile bits 8..11 only: the variant without LDS
    "h8": ((), 8, "H"),            # layouts A and B
    "h3l8": ((1, 2, 3), 8, "H"),   # layouts A, B and C (the first pass of a 23-qubit H layer)
    "u4l8": ((0, 1, 2, 3), 8, "U"),  # complex factors on every layout
    "x3l8": ((1, 2, 3), 8, "X"),   # X only: flipped store addresses, no LDS
}


def bench_product24(args):
    """The product sweep (k_product) alone: one pass of single-qubit factors per case (PRODUCT_CASES)
    on 24- and 28-qubit states in both precisions, HIP events around repeated calls on the state's
    stream, algorithmic bytes 2 S 2^n per pass. --case NAME runs one case at 24 qubits in complex64
    only (for counter runs). Not part of the default run."""
    import numpy
    from qip_amd import GpuBackend
    ev = HipEvents()
    mats = {"H": numpy.array([[1, 1], [1, -1]]) / numpy.sqrt(2), "X": numpy.array([[0.0, 1.0], [1.0, 0.0]]),
            "U": numpy.array([[0.6, 0.8j], [0.8j, 0.6]])}
    reps = max(10, 4 * args.steps)
    names = [args.case] if args.case else list(PRODUCT_CASES)
    sizes = (24,) if args.case else (24, 28)
    dtypes = (("c64", numpy.complex64, 8),) if args.case else (("c64", numpy.complex64, 8),
                                                                ("c128", numpy.complex128, 16))
    cases = {}
    for n in sizes:
        for label, st, amp in dtypes:
            be = GpuBackend.make_state(n, [], [], statetype=st)
            for name in names:
                low, nhigh, f = PRODUCT_CASES[name]
                bits = list(low) + list(range(n - nhigh, n))
                targets, ms_ = [n - 1 - b for b in bits], [mats[f]] * len(bits)
                be.apply_product(targets, ms_)
                be.synchronize()
                a, b = ev.create(), ev.create()
                ev.record(a, be.stream)
                for _ in range(reps):
                    be.apply_product(targets, ms_)
                ev.record(b, be.stream)
                ms = ev.elapsed_ms(a, b) / reps
                gbs = 2 * amp * 2 ** n / ms / 1e6
                cases["{}_{}_n{}".format(name, label, n)] = {"ms": ms, "algorithmic_gbs": gbs,
                                                             "hbm_frac": gbs / HBM_PEAK_GBS}
            be.close()
    c64 = [c for k, c in cases.items() if "_c64_n24" in k]
    value = sum(c["algorithmic_gbs"] for c in c64) / len(c64)
    cfg = {"workload": "product sweep passes on 24- and 28-qubit states (value: mean over the complex64 "
                       "24-qubit cases)", "cases": {k: list(v[:2]) + [v[2]] for k, v in PRODUCT_CASES.items()}}
    return _line("24-qubit c64 product sweep: algorithmic GB/s", value, "GB/s", reps, 1,
                 sum(c["ms"] for c in c64) / len(c64), "c64", cfg, cases=cases,
                 roofline={"bound": "hbm", "kernel": "k_product<float, lds>", "achieved": value,
                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": value / HBM_PEAK_GBS, "traffic": None})


def bench_batch512(args):
    """512 independent layered random circuits, n ~ U{10..22}, depth ~ U{8..32} (seeded), complex64:
    the batching form (one captured graph per run, autotuned) against the plain complex64 path with
    one engine call per gate (from 12 qubits on, held runs of small gates share passes), and per-circuit
    1 - F against plain complex128.
    A step is: reset all 512 states to |0...0>, run every circuit, synchronise."""
    import numpy
    from qip_amd import GpuBackend
    from qip_amd.batch import GpuBatch
    from qip_amd.workloads import batch_layered_circuits
    log = lambda msg: print("[bench] " + msg, file=sys.stderr, flush=True)  # noqa: E731
    count = 512
    t = time.perf_counter()
    ns, depths, seeds, packed = batch_layered_circuits(count, args.seed)
    gen_s = time.perf_counter() - t
    b = GpuBatch(ns)
    t = time.perf_counter()
    b.load(packed)
    load_s = time.perf_counter() - t
    t = time.perf_counter()
    tuning, tuned_ms = b.autotune(reps=2)
    tune_s = time.perf_counter() - t
    info = b.info()
    log("batch: {} gates, {} blocks in {} waves, tuning {} ({:.2f} ms)".format(
        info["gates"], info["blocks"], info["waves"], tuning, tuned_ms))
    ev = HipEvents()
    e0, e1 = ev.create(), ev.create()
    for _ in range(args.warmup):
        b.reset()
        b.run()
        b.synchronize()
    gpu_ms = []
    start = time.perf_counter()
    for _ in range(args.steps):
        b.reset()
        ev.record(e0, b.stream)
        b.run()
        ev.record(e1, b.stream)
        gpu_ms.append(ev.elapsed_ms(e0, e1))
    b.synchronize()
    elapsed = time.perf_counter() - start
    circuits_s = count * args.steps / elapsed
    run_ms = sum(gpu_ms) / len(gpu_ms)
    eff_gbs = info["algorithmic_bytes"] / (run_ms * 1e-3) / 1e9
    log("batch: {} steps in {:.3f} s, run {:.2f} ms".format(args.steps, elapsed, run_ms))

    # the plain path, one engine call per gate (every gate applied by itself; runs of held small gates
    # share passes): complex64 (timed) and complex128 (the fidelity reference)
    gates = [packed.gates_of(i) for i in range(count)]
    plain = {}
    fids, errs = [], []
    for name, st in (() if args.no_plain else (("c64", numpy.complex64), ("c128", numpy.complex128))):
        warm = GpuBackend.make_state(ns[0], [], [], statetype=st)
        for gq, G in gates[0]:
            warm.apply_matrix(list(gq), G)
        warm.close()
        t = time.perf_counter()
        for i in range(count):
            be = GpuBackend.make_state(ns[i], [], [], statetype=st)
            for gq, G in gates[i]:
                be.apply_matrix(list(gq), G)
            be.synchronize()
            if name == "c128":
                want = be.get_state()
                got = b.get_state(i).astype(numpy.complex128)
                ov = numpy.vdot(want, got)
                fids.append(1.0 - abs(ov) ** 2 / (numpy.vdot(want, want).real * numpy.vdot(got, got).real))
                errs.append(float(numpy.max(numpy.abs(want - got))))
            be.close()
        plain[name] = time.perf_counter() - t
        log("plain {}: {:.3f} s".format(name, plain[name]))
    b.close()
    cfg = {"workload": "512 independent layered random circuits (Haar 1q + CNOT/CZ brickwork), n ~ U{10..22}, "
                       "depth ~ U{8..32}, complex64, batched (LDS-resident small states + fused MFMA waves, one "
                       "hipGraph per run)", "circuits": count, "seed": args.seed, "gates": info["gates"],
           "qubits_total": int(sum(ns)), "lds_states": info["lds_states"], "tiled_states": info["tiled_states"],
           "blocks": info["blocks"], "waves": info["waves"], "tuning": tuning}
    out = _line("512 uneven circuits/s, batched c64 vs the plain path", circuits_s, "circuits/s", args.steps,
                args.warmup, elapsed * 1e3 / args.steps, "c64", cfg,
                batch={"circuits_per_s": circuits_s, "gpu_ms_per_run": run_ms, "algorithmic_bytes_per_run":
                       info["algorithmic_bytes"], "effective_gbs": eff_gbs, "hbm_frac": eff_gbs / HBM_PEAK_GBS,
                       "generate_s": gen_s, "load_plan_capture_s": load_s, "autotune_s": tune_s,
                       "autotuned_run_ms": tuned_ms},
                roofline={"bound": "hbm", "kernel": "one batch run (k_batch6 waves + LDS-resident small states)",
                          "achieved": eff_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": eff_gbs / HBM_PEAK_GBS,
                          "traffic": None, "bytes_per_launch": info["algorithmic_bytes"], "avg_launch_ms": run_ms,
                          "note": "the whole captured graph as one unit: algorithmic bytes of every state's passes "
                                  "over the run's GPU time (per-kernel figures: profiles/k_batch6_batch512.txt)"})
    if plain:
        out.update(plain_c64={"circuits_per_s": count / plain["c64"], "s_per_batch": plain["c64"]},
                   plain_c128={"circuits_per_s": count / plain["c128"], "s_per_batch": plain["c128"]},
                   speedup_vs_plain_c64=plain["c64"] / (elapsed / args.steps),
                   infidelity_vs_c128={"max": float(max(fids)), "median": float(numpy.median(fids))},
                   max_abs_err_vs_c128=float(max(errs)))
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_batch(ns, gates, args.cpu_seconds)
    return out


def cpu_baseline_batch(ns, gates, seconds: float):
    """The oracle's cdot_loop restatement (complex128, one out-of-place pass per gate, all visible
    cores) on whole circuits of the batch, in batch order, until `seconds` of work; the rate in
    amplitude-gate updates (sum over gates of 2^n) is extrapolated to the whole batch."""
    import numpy
    from oracle import oracle
    cores = len(os.sched_getaffinity(0))
    total_work = sum(len(g) * 2 ** n for n, g in zip(ns, gates))
    done_work, spent, circuits = 0, 0.0, 0
    for n, gs in zip(ns, gates):
        v = numpy.zeros(2 ** n, dtype=numpy.complex128)
        v[0] = 1.0
        out = numpy.zeros_like(v)
        t = time.perf_counter()
        for gq, G in gs:
            oracle.kronselect_dot({tuple(gq): G}, v, n, out, fast=True, nthreads=cores)
            v, out = out, v
        spent += time.perf_counter() - t
        done_work += len(gs) * 2 ** n
        circuits += 1
        if spent >= seconds:
            break
    full_s = spent * total_work / done_work
    return {"value": len(ns) / full_s, "unit": "circuits/s", "cores": cores, "kind": "port",
            "sample": "the first {} circuits of the batch run in full on the oracle's cdot_loop restatement "
                      "(-O3 -ffast-math -fopenmp), {:.1f} s; rate per amplitude-gate update extrapolated to the "
                      "whole batch".format(circuits, spent)}


def mfma_fraction(n: int, block_ms: float):
    """The fused kernel against the dense BF16 MFMA peak: algorithmic flops of a 6-qubit block, 8 * 64
    real flops per amplitude (a 64 x 64 complex matrix per group of 64 amplitudes), and the flops the
    split-BF16 3-product form executes (three times as many)."""
    flops = 8 * 64 * 2 ** n
    tf = flops / (block_ms * 1e-3) / 1e12
    return {"bound": "mfma (secondary: the kernel is HBM-bound)", "achieved": tf, "peak": MFMA_BF16_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": tf / MFMA_BF16_PEAK_TFLOPS, "flops_per_launch": flops,
            "executed_tflops": 3 * tf, "executed_frac": 3 * tf / MFMA_BF16_PEAK_TFLOPS}


def bench_circuit30(args):
    """The headline (module docstring): fused complex64 against plain complex128 at 30 qubits."""
    import numpy
    from qip_amd import GpuBackend, infidelity, max_abs_diff
    from qip_amd.fused import FusedBackend
    from qip_amd.workloads import algorithmic_bytes_plain, gate_counts, layered_random_circuit

    n, depth = args.n, args.depth
    ops = layered_random_circuit(n, depth, args.seed)
    n1, n2 = gate_counts(ops)
    ngates = len(ops)
    ev = HipEvents()
    log = lambda msg: print("[bench] " + msg, file=sys.stderr, flush=True)  # noqa: E731

    # ---- headline: fused complex64 ------------------------------------------------------------
    fz = FusedBackend.make_state(n, [], [])
    for w in range(args.warmup):
        nb = run_fused_step(fz, ops)
        log("warmup {} done: {} fused blocks".format(w, nb))
    fz.synchronize()
    start = time.perf_counter()
    for s in range(args.steps):
        nblocks = run_fused_step(fz, ops)
    fz.synchronize()
    elapsed = time.perf_counter() - start
    log("fused: {} steps in {:.3f} s ({} blocks/step)".format(args.steps, elapsed, nblocks))
    ms_per_step = elapsed * 1e3 / args.steps
    fused_gates_s = ngates * args.steps / elapsed
    block_bytes = 2 * 8 * 2 ** n
    # per block pass (k_fused6, whose last workgroup also sums the norm), from two back-to-back runs
    timings = [fused_kernel_ms(fz, ops, ev) for _ in range(2)]
    avg_block_ms = sum(t[0] for t in timings) / len(timings)
    fused_gbs = block_bytes / (avg_block_ms * 1e-3) / 1e9
    floor_ms = fused_pattern_floor_ms(fz, ops, ev)

    result = {
        "metric": "30-qubit gates/s + HBM GB/s: fused c64 vs plain c128 at 1-F<=1e-8",
        "value": fused_gates_s,
        "unit": "gates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {
            "workload": "{}-qubit depth-{} layered random circuit (Haar 1q on every qubit + CNOT/CZ brickwork), "
                        "fused complex64 split-BF16 MFMA vs plain complex128".format(n, depth),
            "n_qubits": n, "depth": depth, "gates": ngates, "one_qubit_gates": n1, "two_qubit_gates": n2,
            "seed": args.seed, "fusion_width": FusedBackend.fusion_width, "fused_blocks": nblocks,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm", "kernel": "k_fused6 (six-qubit split-BF16 MFMA block, complex64)",
            "achieved": fused_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": fused_gbs / HBM_PEAK_GBS,
            "traffic": fused6_traffic_bytes(n), "bytes_per_launch": block_bytes, "avg_launch_ms": avg_block_ms,
        },
        "roofline_mfma": mfma_fraction(n, avg_block_ms),
        "fused_c64": {"gates_per_s": fused_gates_s, "s_per_circuit": elapsed / args.steps,
                      "gpu_ms_per_circuit": avg_block_ms * (nblocks + 1), "blocks": nblocks,
                      "algorithmic_gb_per_circuit": block_bytes * (nblocks + 1) / 1e9,
                      "pattern_floor": {"what": "the same blocks' tile patterns swept in place with no arithmetic",
                                        "avg_launch_ms": floor_ms,
                                        "gbs": block_bytes / (floor_ms * 1e-3) / 1e9,
                                        "kernel_over_floor": avg_block_ms / floor_ms}},
    }

    # ---- plain complex128 on the same circuit ----------------------------------------------------
    if not args.no_plain:
        pl = GpuBackend.make_state(n, [], [])
        pl.synchronize()
        start = time.perf_counter()
        for d in ops:
            pl.kronselect_dot(d)
        pl.synchronize()
        plain_s = time.perf_counter() - start
        plain_bytes = algorithmic_bytes_plain(ops, n, 16)
        log("plain c128: {:.3f} s".format(plain_s))
        fz.flush()
        inf = infidelity(pl, fz)
        err = max_abs_diff(pl, fz)
        # the complex128 fused form: the same blocks applied in FP64 (dense sweeps, FP64 MFMA tiles)
        from qip_amd.fused import Fused128Backend
        f128 = Fused128Backend.make_state(n, [], [])
        f128.synchronize()
        start = time.perf_counter()
        for d in ops:
            f128.kronselect_dot(d)
        f128.synchronize()
        f128_s = time.perf_counter() - start
        log("fused c128: {:.3f} s ({} blocks)".format(f128_s, f128.blocks_applied))
        result["fused_c128"] = {
            "gates_per_s": ngates / f128_s, "s_per_circuit": f128_s, "blocks": f128.blocks_applied,
            "speedup_vs_plain_c128": plain_s / f128_s, "max_abs_err_vs_plain_c128": max_abs_diff(pl, f128),
            "infidelity_vs_plain_c128": infidelity(pl, f128),
            "note": "Fused128Backend: the precision form's blocks applied in FP64 on a complex128 state"}
        f128.close()
        dom = plain_dominant_roofline(pl, n, ev)     # applies further gates to pl: after every comparison
        result["plain_c128"] = {
            "gates_per_s": ngates / plain_s, "s_per_circuit": plain_s,
            "effective_gbs": plain_bytes / plain_s / 1e9, "effective_gbs_note": EFFECTIVE_NOTE,
            "dominant_kernel": dict(dom, frac=dom["achieved_gbs"] / HBM_PEAK_GBS),
        }
        result["speedup_fused_vs_plain"] = plain_s / (elapsed / args.steps)
        result["infidelity_fused_vs_c128"] = inf
        result["max_abs_err_fused_vs_c128"] = err
        pl.close()
    fz.close()

    # ---- CPU baseline ---------------------------------------------------------------------------
    if not args.no_cpu_baseline:
        log("cpu baseline: sampling")
        result["cpu_baseline"] = cpu_baseline(ops, n, args.cpu_seconds)

    return result


# the other workloads, run after the headline in the default invocation, nested under "configs"
SECONDARY = (("qft16", "bench_qft16"), ("circuit28", "bench_circuit28"), ("grover24", "bench_grover24"),
             ("batch512", "bench_batch512"), ("measure28", "bench_measure28"), ("func28", "bench_func28"),
             ("dense28", "bench_dense28"))
_LINE_ONLY = ("n_gpus", "higher_is_better", "scaling", "vs_baseline", "data")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=30)
    ap.add_argument("--depth", type=int, default=40)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--config-cpu-seconds", type=float, default=5.0,
                    help="CPU-baseline work per secondary configuration in the default run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="leave out the plain complex128 comparison run")
    ap.add_argument("--copy-calibration-gib", type=float, default=0.0,
                    help="after the benchmark, one device-to-device copy of this many GiB (a kernel of known "
                         "traffic for calibrating hardware byte counters)")
    ap.add_argument("--workload", default="all",
                    choices=["all", "circuit30", "circuit28", "grover24", "qft16", "batch512", "measure28", "func28",
                             "dense28", "product24"],
                    help="all (default): the headline, then every other configuration nested under 'configs' in the "
                         "same JSON line; or one workload alone, as its own JSON line")
    ap.add_argument("--case", choices=sorted(PRODUCT_CASES), default=None,
                    help="product24 only: one case, complex64 at 24 qubits (for counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the engine holds the whole state on a single MI355X; "
              "a multi-GPU run is not part of this build (--gpus {} requested)".format(args.gpus), file=sys.stderr)
        sys.exit(2)
    log = lambda msg: print("[bench] " + msg, file=sys.stderr, flush=True)  # noqa: E731
    funcs = dict(SECONDARY, circuit30="bench_circuit30", product24="bench_product24")
    if args.workload != "all":
        result = globals()[funcs[args.workload]](args)
    else:
        t = time.perf_counter()
        result = bench_circuit30(args)
        result["wall_s"] = time.perf_counter() - t
        configs = {}
        sub = argparse.Namespace(**vars(args))
        sub.cpu_seconds = args.config_cpu_seconds
        for name, fn in SECONDARY:
            log("config {}: start".format(name))
            t = time.perf_counter()
            line = globals()[fn](sub)
            entry = {k: v for k, v in line.items() if k not in _LINE_ONLY}
            entry["wall_s"] = time.perf_counter() - t
            configs[name] = entry
            log("config {}: {:.4g} {} in {:.1f} s".format(name, line["value"], line["unit"], entry["wall_s"]))
        result["configs"] = configs
    if args.copy_calibration_gib > 0:
        copy_calibration(HipEvents(), int(args.copy_calibration_gib * 2 ** 30))
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
