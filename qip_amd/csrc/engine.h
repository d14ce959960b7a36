// engine.h -- internal declarations shared by the translation units of libqipgpu.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "qipgpu.h"

namespace qipgpu {

// Interleaved complex amplitude, aligned so one amplitude is one vector load
// (16 B dwordx4 for complex128, 8 B dwordx2 for complex64).
template <typename R>
struct __attribute__((aligned(2 * sizeof(R)))) cx {
    R re, im;
};

// Ascending list of state-bit positions at which zero bits are inserted into a compressed
// index (the positions of target and control qubits). Lives in kernel arguments, four 8-bit
// positions per 32-bit word: a loop over them then reads the kernel arguments with scalar loads
// (byte elements would need vector loads, whose waits also drain every load and store the thread
// has in flight), and the arguments stay small (launch cost matters for small states).
struct PosList {
    int32_t count;
    uint32_t w[(QIPGPU_MAX_QUBITS + 3) / 4];

    __host__ __device__ __forceinline__ uint32_t at(int i) const { return (w[i >> 2] >> (8 * (i & 3))) & 0xffu; }
    void set(int i, uint32_t p) { w[i >> 2] = (w[i >> 2] & ~(0xffu << (8 * (i & 3)))) | (p << (8 * (i & 3))); }
};

// Insert a zero bit at each listed position (ascending), turning an enumeration index over
// the free bits into a state index with those bits clear.
__device__ __forceinline__ uint64_t insert_zeros(uint64_t g, const PosList &pl) {
    for (int i = 0; i < pl.count; ++i) {
        const uint32_t p = pl.at(i);
        const uint64_t lo = g & ((1ull << p) - 1ull);
        g = ((g ^ lo) << 1) | lo;
    }
    return g;
}

struct Recorder;   // recorder.hip: gates recorded for online fusion

// Product sweeps (product.hip): every pass keeps the kProductLowBits lowest state bits in its tile and
// takes at most kProductHighPerPass factors above them.
constexpr int kProductLowBits = 4;
constexpr int kProductHighPerPass = 8;

constexpr int kRingSlots = 8;
constexpr size_t kRingSlotBytes = 64 * 64 * 16 + 4096;  // a 64x64 complex128 matrix + offsets

// A diagonal gate on at most two targets, held back on the state (qipgpu_state::chain) so that a run
// of them is applied by one kernel: the factor d[j] multiplies the amplitudes whose ctrl_mask bits are
// all 1, j being the value of state bits tpos[0] (MSB) .. tpos[k-1].
constexpr int kDiagChainMax = 16;
struct ChainGate {
    uint64_t ctrl_mask;
    cx<double> d[4];
    uint8_t k;
    uint8_t tpos[2];
};

// A single-qubit gate (2x2 U on state bit tpos where the bits of ctrl_mask are 1), held back on the
// state (qipgpu_state::run) so that a run of them whose bits share one 4096-amplitude tile is applied
// in one pass, each gate in turn on the tile in LDS.
constexpr int kTileRunMax = 32;
struct RunGate {
    uint64_t ctrl_mask;
    cx<double> U[4];
    uint8_t tpos;
};

}  // namespace qipgpu

struct qipgpu_state {
    int32_t n = 0;
    int32_t dtype = 0;
    int32_t device = 0;
    uint64_t len = 0;          // 2^n amplitudes
    size_t amp_bytes = 0;      // bytes per amplitude
    void *amps = nullptr;
    hipStream_t stream = nullptr;
    // growable device scratch (reduction partials, staging)
    void *scratch = nullptr;
    size_t scratch_cap = 0;
    // device staging ring for operators too large for kernel arguments
    void *d_ring = nullptr;
    void *h_ring = nullptr;      // pinned
    void *h_ring_dev = nullptr;  // the pinned ring as the device addresses it
    hipEvent_t ring_ev[qipgpu::kRingSlots] = {};
    bool ring_used[qipgpu::kRingSlots] = {};
    int ring_next = 0;
    // fused path: d_fused[0] = renormalisation factor and d_fused[1] = norm of the last block (written
    // when the state is settled or the norm asked for), then two arrays of `fused_partials` per-workgroup
    // norm partials used by consecutive blocks in turn. A block's output stays unnormalised until the
    // next block or a settle applies 1/sqrt(sum of its partials): fused_prev_count > 0 marks that
    // pending factor, with the partials in array fused_parity ^ 1.
    double *d_fused = nullptr;
    size_t fused_partials = 0;
    uint32_t fused_prev_count = 0;
    int fused_parity = 0;
    // fused path operator staging (split BF16 planes), device
    void *d_fmat = nullptr;
    size_t d_fmat_bytes = 0;
    // dense operators on 7..12 targets (dense.hip launch_gemm): the operator and its offsets pass
    // through pinned host memory to a device buffer on the state's queue (gemm_ev marks the last
    // copy out of the pinned buffer), and the compact temporary is kept while it is small
    void *h_gemm_op = nullptr;   // pinned
    size_t h_gemm_op_bytes = 0;
    void *d_gemm_op = nullptr;
    size_t d_gemm_op_bytes = 0;
    hipEvent_t gemm_ev = nullptr;
    void *gemm_tmp = nullptr;
    size_t gemm_tmp_bytes = 0;
    // a gate graph being recorded (qipgpu_capture_begin .. qipgpu_capture_end): gate kernels whose
    // operands are all kernel arguments become graph nodes, each after the previous one, instead of
    // launches (launch_gate); any other operation is refused and marks the recording refused
    bool capturing = false;
    bool capture_refused = false;
    hipGraph_t rec_graph = nullptr;
    hipGraphNode_t rec_last = nullptr;
    // gates recorded for online fusion and not launched yet (complex64 states)
    qipgpu::Recorder *recorder = nullptr;
    // gates held back (flush_held applies them before anything else touches the state): a run of
    // diagonal gates or a run of single-qubit gates, never both at once
    qipgpu::ChainGate chain[qipgpu::kDiagChainMax];
    int32_t chain_len = 0;
    qipgpu::RunGate run[qipgpu::kTileRunMax];
    int32_t run_len = 0;
    uint64_t run_bits = 0;     // target and control bits of the held single-qubit gates
};

namespace qipgpu {

int set_error(int code, const char *fmt, ...);

#define QIP_HIP(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::qipgpu::set_error(QIPGPU_EHIP, "%s failed: %s", #expr,                \
                                       hipGetErrorString(e_));                             \
    } while (0)

#define QIP_LAUNCHED()                                                                     \
    do {                                                                                   \
        hipError_t e_ = hipGetLastError();                                                 \
        if (e_ != hipSuccess)                                                              \
            return ::qipgpu::set_error(QIPGPU_EHIP, "kernel launch failed: %s",            \
                                       hipGetErrorString(e_));                             \
    } while (0)

// A device buffer for one call, freed on every exit path once the state's stream is done with it.
struct TempBuffer {
    hipStream_t stream;
    void *p = nullptr;
    explicit TempBuffer(hipStream_t st) : stream(st) {}
    TempBuffer(const TempBuffer &) = delete;
    TempBuffer &operator=(const TempBuffer &) = delete;
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes); }
    // hand the buffer over (it is not freed here any more)
    void *release() {
        void *q = p;
        p = nullptr;
        return q;
    }
    ~TempBuffer() {
        if (p) {
            (void)hipStreamSynchronize(stream);
            (void)hipFree(p);
        }
    }
};

// Make sure s->scratch holds at least `bytes`.
int ensure_scratch(qipgpu_state *s, size_t bytes);
// Compute units of the current device, queried once per device (persistent grids are sized by it).
int device_cus();
// Validate a list of qubit indices against n (range, distinctness); returns QIPGPU_OK or EINVAL.
int check_qubits(const qipgpu_state *s, int32_t k, const int32_t *q, const char *what);
// Reserve a staging ring slot, copy `bytes` from `src` to the device, return device pointer.
int ring_upload(qipgpu_state *s, const void *src, size_t bytes, void **dev);
// Apply the gates held back on the state, in order (engine.hip): every entry point that touches the
// state does so first, through flush_recorded.
int flush_held(qipgpu_state *s);
// Apply the held-back gates, then launch every recorded gate (recorder.hip); the recorded gates are
// left alone while the recorder itself is launching.
int flush_recorded(qipgpu_state *s);
// While s is capturing: mark the capture refused and return QIPGPU_EUNSUP naming `what`.
int refuse_in_capture(qipgpu_state *s, const char *what);
#define QIP_NOT_CAPTURING(s, what)                                                        \
    do {                                                                                   \
        if ((s)->capturing) return ::qipgpu::refuse_in_capture((s), (what));               \
    } while (0)

// A gate kernel whose operands all travel as kernel arguments: launched on the state's queue, or,
// while a gate graph is recorded, added to it after the previously recorded kernel (the node copies
// the arguments).
template <typename... P>
int launch_gate(qipgpu_state *s, void (*kernel)(P...), dim3 grid, dim3 block, P... args) {
    if (!s->capturing) {
        hipLaunchKernelGGL(kernel, grid, block, 0, s->stream, args...);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return set_error(QIPGPU_EHIP, "kernel launch failed: %s", hipGetErrorString(e));
        return QIPGPU_OK;
    }
    void *params[] = {static_cast<void *>(&args)...};
    hipKernelNodeParams p;
    std::memset(&p, 0, sizeof(p));
    p.func = reinterpret_cast<void *>(kernel);
    p.gridDim = grid;
    p.blockDim = block;
    p.sharedMemBytes = 0;
    p.kernelParams = params;
    p.extra = nullptr;
    hipGraphNode_t node = nullptr;
    const hipGraphNode_t prev = s->rec_last;
    const hipError_t e = hipGraphAddKernelNode(&node, s->rec_graph, prev ? &prev : nullptr, prev ? 1 : 0, &p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        s->capture_refused = true;
        return set_error(QIPGPU_EHIP, "hipGraphAddKernelNode failed: %s", hipGetErrorString(e));
    }
    s->rec_last = node;
    return QIPGPU_OK;
}
// Drop recorded gates (the state is about to be overwritten) / free the recorder.
void recorder_discard(qipgpu_state *s);
void recorder_free(qipgpu_state *s);
// Gates recorded and not launched yet.
int64_t recorder_pending(const qipgpu_state *s);
// Launch recorded gates, then apply a fused path's pending renormalisation factor: before the state
// is read or written from the host.
int settle(qipgpu_state *s);
// Sum of |psi|^2 in FP64, result left on the device at `dst` (enqueued on s->stream).
int norm_to_device(qipgpu_state *s, double *dst);
// Dense operator U (2^K x 2^K, row-major) on 4 <= K <= QIPGPU_MAX_DENSE targets at state bits `tbits`
// (first = MSB of U's index), where the bits of ctrl_mask are 1; `allpos` = control and target
// bits (dense.hip).
int dense_large(qipgpu_state *s, const std::vector<int> &tbits, uint64_t ctrl_mask, const std::vector<int> &allpos,
                const std::complex<double> *U);
// Diagonal d (2^K entries) on 7 <= K <= QIPGPU_MAX_DENSE targets, where the bits of ctrl_mask are 1;
// `cpos` = control bits (dense.hip).
int diag_large(qipgpu_state *s, const std::vector<int> &tbits, uint64_t ctrl_mask, const std::vector<int> &cpos,
               const std::complex<double> *d);

}  // namespace qipgpu
