// engine.hip -- plain state-vector engine: state lifecycle, gate application, func_apply,
// reductions and measurement, for complex128 and complex64 states on gfx950.
//
// Every kernel here is a streaming sweep bound by HBM bandwidth: each amplitude a gate can
// change is read once and written once, 16 B (complex128) or 8 B (complex64) per access, with
// consecutive lanes on consecutive enumeration indices so loads coalesce whenever the lowest
// state bits are free. Reductions accumulate in FP64 and combine partials in a fixed order, so
// results do not depend on scheduling.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "engine.h"

#include <mutex>

using qipgpu::cx;
using qipgpu::insert_zeros;
using qipgpu::PosList;

namespace qipgpu {

static thread_local std::string g_last_error;

int set_error(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int device_cus() {
    static std::atomic<int> cached[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
        (void)hipGetLastError();
        return 256;
    }
    int cus = cached[dev].load(std::memory_order_relaxed);
    if (cus > 0) return cus;
    hipDeviceProp_t p;
    cus = (hipGetDeviceProperties(&p, dev) == hipSuccess && p.multiProcessorCount > 0) ? p.multiProcessorCount : 256;
    (void)hipGetLastError();
    cached[dev].store(cus, std::memory_order_relaxed);
    return cus;
}

int refuse_in_capture(qipgpu_state *s, const char *what) {
    s->capture_refused = true;
    return set_error(QIPGPU_EUNSUP, "%s cannot be recorded into a gate graph (only gates whose operands are "
                                    "kernel arguments can)", what);
}

int ensure_scratch(qipgpu_state *s, size_t bytes) {
    QIP_NOT_CAPTURING(s, "an operation that uses device scratch");
    if (bytes <= s->scratch_cap) return QIPGPU_OK;
    if (s->scratch) {
        QIP_HIP(hipStreamSynchronize(s->stream));
        QIP_HIP(hipFree(s->scratch));
        s->scratch = nullptr;
        s->scratch_cap = 0;
    }
    size_t cap = std::max<size_t>(bytes, 1 << 20);
    if (hipMalloc(&s->scratch, cap) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(QIPGPU_ENOMEM, "cannot allocate %zu bytes of device scratch", cap);
    }
    s->scratch_cap = cap;
    return QIPGPU_OK;
}

int check_qubits(const qipgpu_state *s, int32_t k, const int32_t *q, const char *what) {
    if (k < 0 || k > s->n) return set_error(QIPGPU_EINVAL, "%s: bad qubit count %d for n=%d", what, k, s->n);
    uint64_t seen = 0;
    for (int i = 0; i < k; ++i) {
        if (q[i] < 0 || q[i] >= s->n)
            return set_error(QIPGPU_EINVAL, "%s: qubit index %d out of range for n=%d", what, q[i], s->n);
        uint64_t b = 1ull << q[i];
        if (seen & b) return set_error(QIPGPU_EINVAL, "%s: qubit index %d repeated", what, q[i]);
        seen |= b;
    }
    return QIPGPU_OK;
}

// Copy a staging slot out of pinned host memory on the state's own queue. A small kernel reads the
// slot over the host link (system-scope loads, so no cache holds an older copy of a reused slot).
// A copy-engine transfer would instead add a cross-queue wait before every kernel that uses the slot.
__global__ __launch_bounds__(256) void k_stage(const uint64_t *src, uint64_t *__restrict__ dst, uint32_t words) {
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < words; i += gridDim.x * 256)
        dst[i] = __hip_atomic_load(src + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

int ring_upload(qipgpu_state *s, const void *src, size_t bytes, void **dev) {
    QIP_NOT_CAPTURING(s, "an operator staged through the pinned ring");
    if (bytes > kRingSlotBytes) return set_error(QIPGPU_EINVAL, "operator of %zu bytes exceeds staging slot", bytes);
    if (!s->d_ring) {
        QIP_HIP(hipMalloc(&s->d_ring, kRingSlots * kRingSlotBytes));
        QIP_HIP(hipHostMalloc(&s->h_ring, kRingSlots * kRingSlotBytes, hipHostMallocDefault));
        QIP_HIP(hipHostGetDevicePointer(&s->h_ring_dev, s->h_ring, 0));
        for (int i = 0; i < kRingSlots; ++i) QIP_HIP(hipEventCreateWithFlags(&s->ring_ev[i], hipEventDisableTiming));
    }
    const int slot = s->ring_next;
    s->ring_next = (slot + 1) % kRingSlots;
    if (s->ring_used[slot]) QIP_HIP(hipEventSynchronize(s->ring_ev[slot]));   // its last stage has run
    char *h = static_cast<char *>(s->h_ring) + slot * kRingSlotBytes;
    char *d = static_cast<char *>(s->d_ring) + slot * kRingSlotBytes;
    std::memcpy(h, src, bytes);
    // whole 8-byte words; the slot size is a multiple of 8, so the last word stays inside the slot
    const uint32_t words = (uint32_t)((bytes + 7) / 8);
    const uint32_t grid = std::min<uint32_t>((words + 255) / 256, 32u);
    hipLaunchKernelGGL(k_stage, dim3(grid), dim3(256), 0, s->stream,
                       (const uint64_t *)(static_cast<char *>(s->h_ring_dev) + slot * kRingSlotBytes), (uint64_t *)d,
                       words);
    QIP_LAUNCHED();
    QIP_HIP(hipEventRecord(s->ring_ev[slot], s->stream));
    s->ring_used[slot] = true;
    *dev = d;
    return QIPGPU_OK;
}

}  // namespace qipgpu

using namespace qipgpu;

// ------------------------------------------------------------------------------------------------
// small device helpers
// ------------------------------------------------------------------------------------------------

// Gate arithmetic runs in FP64 for both state precisions: a complex64 state is widened on load
// and rounded once on store, so its only error per sweep is the unbiased storage rounding (an FP32
// matrix would not be exactly unitary and would drift the state coherently over many sweeps).
template <typename R>
__device__ __forceinline__ cx<double> widen(cx<R> a) {
    cx<double> r;
    r.re = (double)a.re;
    r.im = (double)a.im;
    return r;
}

template <typename R>
__device__ __forceinline__ cx<R> narrow(cx<double> a) {
    cx<R> r;
    r.re = (R)a.re;
    r.im = (R)a.im;
    return r;
}

template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
    cx<R> r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}

template <typename R>
__device__ __forceinline__ void cfma(cx<R> &acc, cx<R> a, cx<R> b) {
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(-a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(a.im, b.re, acc.im);
}

// Scatter the low `count` bits of v (bit b of v) onto state bits pos[b].
__device__ __forceinline__ uint64_t scatter_bits(uint64_t v, const uint8_t *pos, int count) {
    uint64_t r = 0;
    for (int b = 0; b < count; ++b) r |= ((v >> b) & 1ull) << pos[b];
    return r;
}

template <typename T>
__device__ __forceinline__ T block_sum(T v, T *lds) {
    // 256-thread block: wave-level shuffle then LDS across the 4 waves, fixed order.
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    T r = 0;
    if (threadIdx.x == 0) r = (lds[0] + lds[1]) + (lds[2] + lds[3]);
    __syncthreads();
    return r;
}

static inline uint32_t blocks_for(uint64_t work, uint64_t per_block) {
    uint64_t b = (work + per_block - 1) / per_block;
    return (uint32_t)std::max<uint64_t>(b, 1);
}

// A launch may hold fewer than 2^32 threads along x; per-element kernels loop over their blocks
// when a state needs more (2^32 bases at n = 33).
constexpr uint32_t kMaxGridBlocks = 1u << 22;   // 2^30 threads at 256 per workgroup
static inline uint32_t capped_grid(uint32_t nblocks) { return std::min(nblocks, kMaxGridBlocks); }

// Workgroup order of the gate sweeps. With S = 2^split_bits segments, workgroup b takes block
// (b mod S) * (nblocks / S) + b / S: the workgroups in flight at any moment work on S regions of
// the state far apart instead of one narrow window, which keeps more HBM channels and banks busy
// at once. Measured on MI355X, single-qubit sweep, harmonic mean over every target bit, TB/s:
//   complex128: n=30 (16 GiB) S=1 5.52, S=8 5.71; n=28 (4 GiB) S=1 5.54, S=8 5.76;
//               n=26 (1 GiB) S=1 5.54, S=4 5.69, S=8 5.56
//   complex64:  n=30 (8 GiB) S=1 5.59, S=4 6.17; n=28 S=1 5.88, S=4 6.16; n=26 S=1 5.93, S=4 5.99
struct BlockOrder {
    uint32_t split_bits = 0;   // log2 S
    uint32_t seg_shift = 0;    // log2(nblocks / S)
};

__device__ __forceinline__ uint64_t ordered_block(uint32_t b, const BlockOrder &o) {
    return ((uint64_t)(b & ((1u << o.split_bits) - 1u)) << o.seg_shift) | (b >> o.split_bits);
}

// S from the measurements above: 8 for complex128 states of 4 GiB and more, 4 for other states of
// 512 MiB and more, 1 below (unmeasured there; such states largely sit in the 256 MB last-level
// cache) and whenever nblocks is not a power of two that S divides.
static inline BlockOrder block_order(uint64_t state_bytes, size_t amp_bytes, uint32_t nblocks) {
    BlockOrder o;
    if (amp_bytes == 16 && state_bytes >= (4ull << 30)) o.split_bits = 3;
    else if (state_bytes >= (512ull << 20)) o.split_bits = 2;
    if (nblocks & (nblocks - 1u) || nblocks < (1u << o.split_bits)) o.split_bits = 0;
    uint32_t lg = 0;
    while ((1u << lg) < nblocks) ++lg;
    o.seg_shift = lg - o.split_bits;
    return o;
}

// ------------------------------------------------------------------------------------------------
// state lifecycle
// ------------------------------------------------------------------------------------------------

struct InitArgs {
    uint64_t len;
    uint64_t unfed_mask;  // state bits of qubits no feed covers: amplitude is 0 unless they are 0
    int32_t n_groups;
    const int32_t *glen;   // device
    const int32_t *gpos;   // device, state-bit position of each group index, MSB-first per group
    const int64_t *goff;   // device, offset of each feed in `feeds`
    const cx<double> *feeds;
};

// qip/backend.py:88-101: amplitude = prod_g feed_g[value of the index on group g], computed in
// group order from 1.0 with numpy's complex product (no fused multiply-add), rounded to the
// state's precision once at the end.
// The same product in the same order, chunk by chunk (n >= 12): the leading `m_prefix` groups have all
// their bits above bit 11, so their partial product is the same for the 4096 amplitudes of a chunk and
// is formed once per chunk; each amplitude continues it with the other groups in order. Every
// rounding is the one k_init_product makes.
template <typename R>
__global__ __launch_bounds__(256) void k_init_chunked(cx<R> *amps, InitArgs a, int32_t m_prefix) {
#pragma clang fp contract(off)
    const uint64_t nchunks = a.len >> 12;
    for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const uint64_t base = c << 12;
        double pre = 1.0, pim = 0.0;
        int pos = 0;
        for (int g = 0; g < m_prefix; ++g) {
            const int L = a.glen[g];
            uint64_t sub = 0;
            for (int j = 0; j < L; ++j) sub = (sub << 1) | ((base >> a.gpos[pos + j]) & 1ull);
            pos += L;
            const cx<double> f = a.feeds[a.goff[g] + (int64_t)sub];
            const double nre = pre * f.re - pim * f.im;
            const double nim = pre * f.im + pim * f.re;
            pre = nre;
            pim = nim;
        }
        for (int e = 0; e < 16; ++e) {
            const uint64_t i = base + threadIdx.x + 256u * e;
            double re = 0.0, im = 0.0;
            if ((i & a.unfed_mask) == 0) {
                re = pre;
                im = pim;
                int p2 = pos;
                for (int g = m_prefix; g < a.n_groups; ++g) {
                    const int L = a.glen[g];
                    uint64_t sub = 0;
                    for (int j = 0; j < L; ++j) sub = (sub << 1) | ((i >> a.gpos[p2 + j]) & 1ull);
                    p2 += L;
                    const cx<double> f = a.feeds[a.goff[g] + (int64_t)sub];
                    const double nre = re * f.re - im * f.im;
                    const double nim = re * f.im + im * f.re;
                    re = nre;
                    im = nim;
                }
            }
            cx<R> o;
            o.re = (R)re;
            o.im = (R)im;
            amps[i] = o;
        }
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_init_product(cx<R> *amps, InitArgs a) {
#pragma clang fp contract(off)
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.len; i += stride) {
        double re = 0.0, im = 0.0;
        if ((i & a.unfed_mask) == 0) {
            re = 1.0;
            int pos = 0;
            for (int g = 0; g < a.n_groups; ++g) {
                const int L = a.glen[g];
                uint64_t sub = 0;
                for (int j = 0; j < L; ++j) sub = (sub << 1) | ((i >> a.gpos[pos + j]) & 1ull);
                pos += L;
                const cx<double> f = a.feeds[a.goff[g] + (int64_t)sub];
                const double nre = re * f.re - im * f.im;
                const double nim = re * f.im + im * f.re;
                re = nre;
                im = nim;
            }
        }
        cx<R> o;
        o.re = (R)re;
        o.im = (R)im;
        amps[i] = o;
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_add_range(cx<R> *dst, const cx<R> *src, uint64_t count) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        cx<R> a = dst[i], b = src[i];
        a.re += b.re;
        a.im += b.im;
        dst[i] = a;
    }
}

extern "C" {

const char *qipgpu_last_error(void) { return g_last_error.c_str(); }

const char *qipgpu_version(void) { return "qipgpu 0.1 gfx950"; }

int qipgpu_device_count(int32_t *count) {
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *count = 0;
        return set_error(QIPGPU_EHIP, "hipGetDeviceCount: %s", hipGetErrorString(e));
    }
    *count = c;
    return QIPGPU_OK;
}

}  // extern "C"

namespace {

// Amplitude buffers and queues of destroyed small states, kept for the next state of the same size on
// the same device: creating a state then costs no allocation, and a state made again after another
// was destroyed gets the same buffer, so gate graphs captured on the first replay on the second. At
// most kPoolEntries buffers of up to kPoolMaxBytes each are held (2 GiB), until the process ends.
struct PoolEntry {
    int device;
    size_t bytes;
    void *amps;
    hipStream_t stream;
};
constexpr size_t kPoolMaxBytes = size_t(256) << 20;
constexpr size_t kPoolEntries = 8;
std::mutex g_pool_mu;
std::vector<PoolEntry> g_pool;

// the most recently given buffer of the size first
bool pool_take(int device, size_t bytes, void **amps, hipStream_t *stream) {
    std::lock_guard<std::mutex> lock(g_pool_mu);
    for (size_t i = g_pool.size(); i-- > 0;)
        if (g_pool[i].device == device && g_pool[i].bytes == bytes) {
            *amps = g_pool[i].amps;
            *stream = g_pool[i].stream;
            g_pool.erase(g_pool.begin() + (long)i);
            return true;
        }
    return false;
}

bool pool_give(int device, size_t bytes, void *amps, hipStream_t stream) {
    if (bytes > kPoolMaxBytes) return false;
    std::lock_guard<std::mutex> lock(g_pool_mu);
    if (g_pool.size() >= kPoolEntries) return false;
    g_pool.push_back(PoolEntry{device, bytes, amps, stream});
    return true;
}

}  // namespace

extern "C" {

int qipgpu_create(int32_t n, int32_t dtype, qipgpu_state **out) {
    *out = nullptr;
    if (n < 1 || n > QIPGPU_MAX_QUBITS) return set_error(QIPGPU_EINVAL, "n=%d outside 1..%d", n, QIPGPU_MAX_QUBITS);
    if (dtype != QIPGPU_C128 && dtype != QIPGPU_C64) return set_error(QIPGPU_EINVAL, "unknown dtype %d", dtype);
    qipgpu_state *s = new qipgpu_state();
    s->n = n;
    s->dtype = dtype;
    s->len = 1ull << n;
    s->amp_bytes = dtype == QIPGPU_C128 ? 16 : 8;
    if (hipGetDevice(&s->device) != hipSuccess) {
        delete s;
        return set_error(QIPGPU_EHIP, "hipGetDevice failed");
    }
    if (pool_take(s->device, s->len * s->amp_bytes, &s->amps, &s->stream)) {
        *out = s;
        return QIPGPU_OK;
    }
    if (hipMalloc(&s->amps, s->len * s->amp_bytes) != hipSuccess) {
        (void)hipGetLastError();
        delete s;
        return set_error(QIPGPU_ENOMEM, "cannot allocate a %d-qubit state (%llu bytes)", n,
                         (unsigned long long)(s->len * (dtype == QIPGPU_C128 ? 16 : 8)));
    }
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipFree(s->amps);
        delete s;
        return set_error(QIPGPU_EHIP, "hipStreamCreate failed");
    }
    *out = s;
    return QIPGPU_OK;
}

int qipgpu_destroy(qipgpu_state *s) {
    if (!s) return QIPGPU_OK;
    if (s->rec_graph) (void)hipGraphDestroy(s->rec_graph);   // an unfinished recording is dropped
    (void)hipStreamSynchronize(s->stream);
    const bool pooled = s->amps && pool_give(s->device, s->len * s->amp_bytes, s->amps, s->stream);
    if (s->amps && !pooled) (void)hipFree(s->amps);
    if (s->scratch) (void)hipFree(s->scratch);
    if (s->d_ring) (void)hipFree(s->d_ring);
    if (s->h_ring) (void)hipHostFree(s->h_ring);
    for (int i = 0; i < kRingSlots; ++i)
        if (s->ring_ev[i]) (void)hipEventDestroy(s->ring_ev[i]);
    if (s->d_fused) (void)hipFree(s->d_fused);
    if (s->d_fmat) (void)hipFree(s->d_fmat);
    if (s->h_gemm_op) (void)hipHostFree(s->h_gemm_op);
    if (s->d_gemm_op) (void)hipFree(s->d_gemm_op);
    if (s->gemm_ev) (void)hipEventDestroy(s->gemm_ev);
    if (s->gemm_tmp) (void)hipFree(s->gemm_tmp);
    recorder_free(s);
    if (!pooled) (void)hipStreamDestroy(s->stream);
    delete s;
    return QIPGPU_OK;
}

// ---- gate graphs (no counterpart in the reference) ----------------------------------------------

struct qipgpu_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;  // recorded after each launch
    bool launched = false;
    void *amps = nullptr;       // the state buffer the graph's kernels update
    uint64_t len = 0;
    int32_t dtype = 0;
    int device = 0;
};

// The graph is built node by node (launch_gate), not by stream capture: no queue enters a capture
// mode, so other states and threads go on as usual while one state records.
int qipgpu_capture_begin(qipgpu_state *s) {
    if (s->capturing) return set_error(QIPGPU_EINVAL, "the state is already capturing");
    int rc = flush_recorded(s);
    if (rc) return rc;
    QIP_HIP(hipGraphCreate(&s->rec_graph, 0));
    s->rec_last = nullptr;
    s->capturing = true;
    s->capture_refused = false;
    return QIPGPU_OK;
}

int qipgpu_capture_end(qipgpu_state *s, qipgpu_graph **out) {
    *out = nullptr;
    if (!s->capturing) return set_error(QIPGPU_EINVAL, "the state is not capturing");
    // gates still held back belong to the recording: their node goes last
    if (!s->capture_refused && flush_held(s) != QIPGPU_OK) s->capture_refused = true;
    s->chain_len = 0;
    s->run_len = 0;
    s->run_bits = 0;
    s->capturing = false;
    hipGraph_t g = s->rec_graph;
    s->rec_graph = nullptr;
    s->rec_last = nullptr;
    if (s->capture_refused) {
        (void)hipGraphDestroy(g);
        return set_error(QIPGPU_EUNSUP, "capture dropped: an operation was refused while recording");
    }
    qipgpu_graph *q = new qipgpu_graph();
    q->graph = g;
    if (hipGraphInstantiate(&q->exec, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipGraphDestroy(g);
        delete q;
        return set_error(QIPGPU_EHIP, "hipGraphInstantiate failed");
    }
    if (hipEventCreateWithFlags(&q->done, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipGraphExecDestroy(q->exec);
        (void)hipGraphDestroy(g);
        delete q;
        return set_error(QIPGPU_EHIP, "hipEventCreate failed");
    }
    q->amps = s->amps;
    q->len = s->len;
    q->dtype = s->dtype;
    q->device = s->device;
    *out = q;
    return QIPGPU_OK;
}

int qipgpu_graph_launch(qipgpu_state *s, qipgpu_graph *g) {
    if (!g) return set_error(QIPGPU_EINVAL, "null graph");
    if (s->capturing) return refuse_in_capture(s, "a graph launch");
    if (g->amps != s->amps || g->len != s->len || g->dtype != s->dtype || g->device != s->device)
        return set_error(QIPGPU_EINVAL, "the graph was captured on another state buffer");
    int rc = flush_recorded(s);
    if (rc) return rc;
    QIP_HIP(hipGraphLaunch(g->exec, s->stream));
    QIP_HIP(hipEventRecord(g->done, s->stream));
    g->launched = true;
    return QIPGPU_OK;
}

int qipgpu_graph_matches(const qipgpu_state *s, const qipgpu_graph *g) {
    return g && g->amps == s->amps && g->len == s->len && g->dtype == s->dtype && g->device == s->device;
}

int qipgpu_graph_destroy(qipgpu_graph *g) {
    if (!g) return QIPGPU_OK;
    if (g->launched) (void)hipEventSynchronize(g->done);   // its last launch has run
    if (g->done) (void)hipEventDestroy(g->done);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return QIPGPU_OK;
}

int qipgpu_info(const qipgpu_state *s, int32_t *n, int32_t *dtype) {
    *n = s->n;
    *dtype = s->dtype;
    return QIPGPU_OK;
}

int qipgpu_synchronize(qipgpu_state *s) {
    QIP_NOT_CAPTURING(s, "a synchronisation");
    const int rc = flush_recorded(s);
    if (rc) return rc;
    QIP_HIP(hipStreamSynchronize(s->stream));
    return QIPGPU_OK;
}

int qipgpu_get_stream(qipgpu_state *s, void **stream) {
    *stream = (void *)s->stream;
    return QIPGPU_OK;
}

int qipgpu_init_product(qipgpu_state *s, int32_t n_groups, const int32_t *group_len,
                        const int32_t *group_idx, const double *feeds) {
    QIP_NOT_CAPTURING(s, "a state initialisation");
    // every amplitude is rewritten, so recorded gates, held-back diagonals and the factor are moot
    s->fused_prev_count = 0;
    s->chain_len = 0;
    s->run_len = 0;
    s->run_bits = 0;
    recorder_discard(s);
    if (n_groups < 0) return set_error(QIPGPU_EINVAL, "negative group count");
    uint64_t fed = 0;
    int64_t total_idx = 0, total_feed = 0;
    std::vector<int64_t> goff(std::max(n_groups, 1));
    for (int g = 0; g < n_groups; ++g) {
        if (group_len[g] < 1 || group_len[g] > s->n) return set_error(QIPGPU_EINVAL, "bad feed group length %d", group_len[g]);
        goff[g] = total_feed;
        total_feed += 1ll << group_len[g];
        for (int j = 0; j < group_len[g]; ++j) {
            int q = group_idx[total_idx + j];
            if (q < 0 || q >= s->n) return set_error(QIPGPU_EINVAL, "feed qubit index %d out of range", q);
            fed |= 1ull << (s->n - 1 - q);
        }
        total_idx += group_len[g];
    }
    if (n_groups == 0) {
        // |0...0> (qip/backend.py:90-91)
        QIP_HIP(hipMemsetAsync(s->amps, 0, s->len * s->amp_bytes, s->stream));
        if (s->dtype == QIPGPU_C128) {
            const double one[2] = {1.0, 0.0};
            QIP_HIP(hipMemcpyAsync(s->amps, one, sizeof(one), hipMemcpyHostToDevice, s->stream));
        } else {
            const float one[2] = {1.0f, 0.0f};
            QIP_HIP(hipMemcpyAsync(s->amps, one, sizeof(one), hipMemcpyHostToDevice, s->stream));
        }
        QIP_HIP(hipStreamSynchronize(s->stream));
        return QIPGPU_OK;
    }
    // device copies of the descriptor: [glen | gpos | pad | goff | feeds]
    std::vector<int32_t> gpos(total_idx);
    for (int64_t i = 0; i < total_idx; ++i) gpos[i] = s->n - 1 - group_idx[i];
    const size_t b_glen = sizeof(int32_t) * n_groups, b_gpos = sizeof(int32_t) * total_idx;
    const size_t o_goff = ((b_glen + b_gpos + 15) / 16) * 16;
    const size_t o_feed = o_goff + ((sizeof(int64_t) * n_groups + 15) / 16) * 16;
    const size_t bytes = o_feed + sizeof(double) * 2 * total_feed;
    int rc = ensure_scratch(s, bytes);
    if (rc) return rc;
    char *d = static_cast<char *>(s->scratch);
    std::vector<char> h(bytes);
    std::memcpy(h.data(), group_len, b_glen);
    std::memcpy(h.data() + b_glen, gpos.data(), b_gpos);
    std::memcpy(h.data() + o_goff, goff.data(), sizeof(int64_t) * n_groups);
    std::memcpy(h.data() + o_feed, feeds, sizeof(double) * 2 * total_feed);
    QIP_HIP(hipMemcpyAsync(d, h.data(), bytes, hipMemcpyHostToDevice, s->stream));
    InitArgs a;
    a.len = s->len;
    a.unfed_mask = (s->len - 1) & ~fed;
    a.n_groups = n_groups;
    a.glen = reinterpret_cast<const int32_t *>(d);
    a.gpos = reinterpret_cast<const int32_t *>(d + b_glen);
    a.goff = reinterpret_cast<const int64_t *>(d + o_goff);
    a.feeds = reinterpret_cast<const cx<double> *>(d + o_feed);
    // leading groups whose bits are all above bit 11: formed once per 4096-amplitude chunk
    int32_t m_prefix = 0;
    if (s->n >= 12) {
        int64_t pos = 0;
        for (int g = 0; g < n_groups; ++g) {
            bool high = true;
            for (int j = 0; j < group_len[g]; ++j) high = high && gpos[pos + j] >= 12;
            pos += group_len[g];
            if (!high) break;
            ++m_prefix;
        }
    }
    if (m_prefix > 0) {
        const uint32_t nb = (uint32_t)std::min<uint64_t>(s->len >> 12, (uint64_t)device_cus() * 8);
        if (s->dtype == QIPGPU_C128)
            hipLaunchKernelGGL(k_init_chunked<double>, dim3(nb), dim3(256), 0, s->stream, (cx<double> *)s->amps, a,
                               m_prefix);
        else
            hipLaunchKernelGGL(k_init_chunked<float>, dim3(nb), dim3(256), 0, s->stream, (cx<float> *)s->amps, a,
                               m_prefix);
    } else {
        const uint32_t nb = std::min<uint32_t>(blocks_for(s->len, 256), 1u << 20);
        if (s->dtype == QIPGPU_C128)
            hipLaunchKernelGGL(k_init_product<double>, dim3(nb), dim3(256), 0, s->stream, (cx<double> *)s->amps, a);
        else
            hipLaunchKernelGGL(k_init_product<float>, dim3(nb), dim3(256), 0, s->stream, (cx<float> *)s->amps, a);
    }
    QIP_LAUNCHED();
    QIP_HIP(hipStreamSynchronize(s->stream));  // host descriptor buffer goes out of scope
    return QIPGPU_OK;
}

static int check_range(qipgpu_state *s, int64_t start, int64_t end) {
    if (start < 0 || end < start || (uint64_t)end > s->len)
        return set_error(QIPGPU_EINVAL, "range [%lld, %lld) outside the %llu-amplitude state", (long long)start,
                         (long long)end, (unsigned long long)s->len);
    return QIPGPU_OK;
}

int qipgpu_get_range(qipgpu_state *s, int64_t start, int64_t end, void *host) {
    QIP_NOT_CAPTURING(s, "a host copy");
    int rc = check_range(s, start, end);
    if (rc) return rc;
    if (end == start) return QIPGPU_OK;
    rc = settle(s);
    if (rc) return rc;
    // on the state's own queue, not the legacy one, which would wait for every blocking queue
    QIP_HIP(hipMemcpyAsync(host, static_cast<char *>(s->amps) + start * s->amp_bytes, (end - start) * s->amp_bytes,
                           hipMemcpyDeviceToHost, s->stream));
    QIP_HIP(hipStreamSynchronize(s->stream));
    return QIPGPU_OK;
}

int qipgpu_set_range(qipgpu_state *s, int64_t start, int64_t end, const void *host) {
    QIP_NOT_CAPTURING(s, "a host copy");
    int rc = check_range(s, start, end);
    if (rc) return rc;
    if (end == start) return QIPGPU_OK;
    rc = settle(s);
    if (rc) return rc;
    QIP_HIP(hipMemcpyAsync(static_cast<char *>(s->amps) + start * s->amp_bytes, host, (end - start) * s->amp_bytes,
                           hipMemcpyHostToDevice, s->stream));
    QIP_HIP(hipStreamSynchronize(s->stream));   // the host data is read before return
    return QIPGPU_OK;
}

int qipgpu_add_range(qipgpu_state *s, int64_t start, int64_t end, const void *host) {
    QIP_NOT_CAPTURING(s, "a host copy");
    int rc = check_range(s, start, end);
    if (rc) return rc;
    if (end == start) return QIPGPU_OK;
    const uint64_t count = end - start;
    rc = settle(s);
    if (rc) return rc;
    rc = ensure_scratch(s, count * s->amp_bytes);
    if (rc) return rc;
    QIP_HIP(hipMemcpyAsync(s->scratch, host, count * s->amp_bytes, hipMemcpyHostToDevice, s->stream));
    const uint32_t nb = std::min<uint32_t>(blocks_for(count, 256), 1u << 20);
    if (s->dtype == QIPGPU_C128)
        hipLaunchKernelGGL(k_add_range<double>, dim3(nb), dim3(256), 0, s->stream,
                           (cx<double> *)s->amps + start, (const cx<double> *)s->scratch, count);
    else
        hipLaunchKernelGGL(k_add_range<float>, dim3(nb), dim3(256), 0, s->stream, (cx<float> *)s->amps + start,
                           (const cx<float> *)s->scratch, count);
    QIP_LAUNCHED();
    QIP_HIP(hipStreamSynchronize(s->stream));
    return QIPGPU_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------
// gate application
// ------------------------------------------------------------------------------------------------

// Dense operator on K <= 3 target qubits, matrix in kernel arguments. One enumeration index g
// ("base") per item: the state index with all target bits 0 and all control bits 1; the 2^K
// amplitudes base + off[j] form one closed group the operator mixes. Each thread handles ITEMS
// bases spaced one block-width apart so every load instruction of a wave is contiguous when the
// low state bits are free; all loads of a thread are issued before any arithmetic.
template <typename R, int K>
struct DenseArgs {
    cx<R> *amps;
    uint64_t nbase;
    uint64_t ctrl_mask;
    PosList pl;
    uint64_t off[1 << K];
    cx<double> U[(1 << K) * (1 << K)];
    BlockOrder order;
    uint32_t nblocks;
};

template <typename R, int K, int ITEMS>
__global__ __launch_bounds__(256) void k_dense(const DenseArgs<R, K> a) {
    constexpr int D = 1 << K;
    for (uint32_t blk = blockIdx.x; blk < a.nblocks; blk += gridDim.x) {
    const uint64_t g0 = ordered_block(blk, a.order) * (256 * ITEMS) + threadIdx.x;
    cx<R> v[ITEMS][D];
    uint64_t idx[ITEMS];
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
        const uint64_t g = g0 + (uint64_t)it * 256;
        idx[it] = insert_zeros(g, a.pl) | a.ctrl_mask;
        if (g < a.nbase) {
#pragma unroll
            for (int j = 0; j < D; ++j) v[it][j] = a.amps[idx[it] + a.off[j]];
        }
    }
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
        if (g0 + (uint64_t)it * 256 < a.nbase) {
#pragma unroll
            for (int i = 0; i < D; ++i) {
                cx<double> acc;
                acc.re = 0;
                acc.im = 0;
#pragma unroll
                for (int j = 0; j < D; ++j) cfma(acc, a.U[i * D + j], widen(v[it][j]));
                a.amps[idx[it] + a.off[i]] = narrow<R>(acc);
            }
        }
    }
    }
}

// Diagonal operator on K <= 6 targets (K = 0: a phase on the controlled subspace). One item per
// amplitude of the controlled subspace.
template <typename R, int K>
struct DiagArgs {
    cx<R> *amps;
    uint64_t count;
    uint64_t ctrl_mask;
    PosList pl;          // control positions only
    uint8_t tpos[8];     // state bit of target j (j = 0 is the MSB of the diagonal index)
    cx<double> d[1 << K];
    BlockOrder order;
    uint32_t nblocks;
};

template <typename R, int K, int ITEMS>
__global__ __launch_bounds__(256) void k_diag(const DiagArgs<R, K> a) {
    for (uint32_t blk = blockIdx.x; blk < a.nblocks; blk += gridDim.x) {
    const uint64_t g0 = ordered_block(blk, a.order) * (256 * ITEMS) + threadIdx.x;
    cx<R> v[ITEMS];
    uint64_t idx[ITEMS];
    // items past the end load a valid amplitude and store nothing: unconditional loads let the
    // compiler keep all of them in flight (a load under a branch makes it drain every earlier one)
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
        const uint64_t g = g0 + (uint64_t)it * 256;
        idx[it] = insert_zeros(g < a.count ? g : a.count - 1, a.pl) | a.ctrl_mask;
        v[it] = a.amps[idx[it]];
    }
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
        if (g0 + (uint64_t)it * 256 < a.count) {
            uint32_t j = 0;
#pragma unroll
            for (int b = 0; b < K; ++b) j = (j << 1) | (uint32_t)((idx[it] >> a.tpos[b]) & 1ull);
            a.amps[idx[it]] = narrow<R>(cmul(widen(v[it]), a.d[j]));
        }
    }
    }
}

namespace {

using cd = std::complex<double>;

// Structure of a dense operator: control qubits hidden in it, and whether the rest is diagonal.
struct Analysed {
    std::vector<int32_t> targets;   // remaining target qubits, MSB first
    std::vector<int32_t> controls;  // extra control qubits found
    std::vector<cd> U;              // operator on `targets`
    bool diagonal = false;
};

// Bit b (counted from the LSB of the matrix index) is a control when the operator never changes
// it and acts as the identity where it is 0; the operator is then diag(I, M') and M' is the
// bit-1 block (the C_MAT structure of kronprod.pyx:215-225).
static bool is_control_bit(const std::vector<cd> &U, int k, int b) {
    const int D = 1 << k;
    const int m = 1 << b;
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) {
            const cd e = U[r * D + c];
            if ((r & m) != (c & m)) {
                if (e != 0.0) return false;
            } else if (!(r & m)) {
                if (e != (r == c ? 1.0 : 0.0)) return false;
            }
        }
    return true;
}

static std::vector<cd> take_block(const std::vector<cd> &U, int k, int b) {
    const int D = 1 << k, Dn = D >> 1;
    std::vector<cd> out(Dn * Dn);
    auto expand = [b](int x) {
        const int lo = x & ((1 << b) - 1);
        return ((x >> b) << (b + 1)) | (1 << b) | lo;
    };
    for (int r = 0; r < Dn; ++r)
        for (int c = 0; c < Dn; ++c) out[r * Dn + c] = U[expand(r) * D + expand(c)];
    return out;
}

static Analysed analyse(int k, const int32_t *targets, const double *mat) {
    Analysed a;
    a.targets.assign(targets, targets + k);
    const int D = 1 << k;
    a.U.resize((size_t)D * D);
    for (int i = 0; i < D * D; ++i) a.U[i] = cd(mat[2 * i], mat[2 * i + 1]);
    bool found = true;
    while (found && !a.targets.empty()) {
        found = false;
        const int kk = (int)a.targets.size();
        for (int b = 0; b < kk; ++b) {
            if (is_control_bit(a.U, kk, b)) {
                a.U = take_block(a.U, kk, b);
                const int j = kk - 1 - b;  // targets[j] carries matrix bit b
                a.controls.push_back(a.targets[j]);
                a.targets.erase(a.targets.begin() + j);
                found = true;
                break;
            }
        }
    }
    const int kk = (int)a.targets.size();
    const int Dk = 1 << kk;
    a.diagonal = true;
    for (int r = 0; r < Dk && a.diagonal; ++r)
        for (int c = 0; c < Dk; ++c)
            if (r != c && a.U[r * Dk + c] != 0.0) {
                a.diagonal = false;
                break;
            }
    return a;
}

static void fill_poslist(PosList &pl, std::vector<int> pos) {
    std::sort(pos.begin(), pos.end());
    pl.count = (int32_t)pos.size();
    for (size_t i = 0; i < pos.size(); ++i) pl.set((int)i, (uint32_t)pos[i]);
}

template <typename R, int K>
static int launch_dense_reg(qipgpu_state *s, const std::vector<int> &tbits, uint64_t ctrl_mask,
                            const std::vector<int> &allpos, const std::vector<cd> &U) {
    DenseArgs<R, K> a;
    a.amps = (cx<R> *)s->amps;
    a.nbase = s->len >> allpos.size();
    a.ctrl_mask = ctrl_mask;
    fill_poslist(a.pl, allpos);
    constexpr int D = 1 << K;
    for (int j = 0; j < D; ++j) {
        uint64_t o = 0;
        for (int t = 0; t < K; ++t)
            if ((j >> (K - 1 - t)) & 1) o |= 1ull << tbits[t];
        a.off[j] = o;
    }
    for (int i = 0; i < D * D; ++i) {
        a.U[i].re = U[i].real();
        a.U[i].im = U[i].imag();
    }
    constexpr int ITEMS = 1;  // one base per thread measured fastest (large grids keep HBM busy)
    const uint32_t nb = blocks_for(a.nbase, 256 * ITEMS);
    a.order = block_order(s->len * s->amp_bytes, s->amp_bytes, nb);
    a.nblocks = nb;
    return launch_gate(s, &k_dense<R, K, ITEMS>, dim3(capped_grid(nb)), dim3(256), a);
}

template <typename R, int K>
static int launch_diag(qipgpu_state *s, const std::vector<int> &tbits, uint64_t ctrl_mask,
                       const std::vector<int> &cpos, const std::vector<cd> &d) {
    DiagArgs<R, K> a;
    a.amps = (cx<R> *)s->amps;
    a.count = s->len >> cpos.size();
    a.ctrl_mask = ctrl_mask;
    fill_poslist(a.pl, cpos);
    for (int b = 0; b < K; ++b) a.tpos[b] = (uint8_t)tbits[b];
    for (int j = 0; j < (1 << K); ++j) {
        a.d[j].re = d[j].real();
        a.d[j].im = d[j].imag();
    }
    constexpr int ITEMS = 4;
    const uint32_t nb = blocks_for(a.count, 256 * ITEMS);
    a.order = block_order(s->len * s->amp_bytes, s->amp_bytes, nb);
    a.nblocks = nb;
    return launch_gate(s, &k_diag<R, K, ITEMS>, dim3(capped_grid(nb)), dim3(256), a);
}

// A run of held-back diagonal gates in one sweep over the amplitudes whose bits common_mask (the
// controls every gate of the run has) are 1. Each amplitude meets the gates in order and is rounded
// after each one, with k_diag's arithmetic, so the result is bit for bit that of launching the gates
// one by one.
template <typename R>
struct DiagChainArgs {
    cx<R> *amps;
    uint64_t count;
    uint64_t common_mask;
    PosList pl;          // common control positions
    int32_t ngates;
    ChainGate g[kDiagChainMax];
    BlockOrder order;
    uint32_t nblocks;
};

template <typename R, int ITEMS>
__global__ __launch_bounds__(256) void k_diag_chain(const DiagChainArgs<R> a) {
    for (uint32_t blk = blockIdx.x; blk < a.nblocks; blk += gridDim.x) {
    const uint64_t g0 = ordered_block(blk, a.order) * (256 * ITEMS) + threadIdx.x;
    cx<R> v[ITEMS];
    uint64_t idx[ITEMS];
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
        const uint64_t g = g0 + (uint64_t)it * 256;
        idx[it] = insert_zeros(g < a.count ? g : a.count - 1, a.pl) | a.common_mask;
        v[it] = a.amps[idx[it]];
    }
#pragma unroll
    for (int it = 0; it < ITEMS; ++it) {
        if (g0 + (uint64_t)it * 256 < a.count) {
            cx<R> x = v[it];
            bool touched = false;
            for (int q = 0; q < a.ngates; ++q) {
                if ((idx[it] & a.g[q].ctrl_mask) != a.g[q].ctrl_mask) continue;
                uint32_t j = 0;
                for (int b = 0; b < a.g[q].k; ++b) j = (j << 1) | (uint32_t)((idx[it] >> a.g[q].tpos[b]) & 1ull);
                x = narrow<R>(cmul(widen(x), a.g[q].d[j]));
                touched = true;
            }
            if (touched) a.amps[idx[it]] = x;
        }
    }
    }
}

// Amplitude-sized cost of one launch (about 4 us of HBM time) when choosing between one sweep for a
// run of diagonal gates and a launch per gate.
constexpr uint64_t kLaunchCostAmps = 1ull << 19;

template <typename R>
static int launch_chain_gate(qipgpu_state *s, const ChainGate &c) {
    std::vector<int> tbits, cpos;
    for (int b = 0; b < c.k; ++b) tbits.push_back(c.tpos[b]);
    for (int p = 0; p < 64; ++p)
        if (c.ctrl_mask >> p & 1) cpos.push_back(p);
    std::vector<cd> d((size_t)1 << c.k);
    for (size_t j = 0; j < d.size(); ++j) d[j] = cd(c.d[j].re, c.d[j].im);
    switch (c.k) {
        case 0: return launch_diag<R, 0>(s, tbits, c.ctrl_mask, cpos, d);
        case 1: return launch_diag<R, 1>(s, tbits, c.ctrl_mask, cpos, d);
        default: return launch_diag<R, 2>(s, tbits, c.ctrl_mask, cpos, d);
    }
}

template <typename R>
static int flush_chain_typed(qipgpu_state *s) {
    const int m = s->chain_len;
    s->chain_len = 0;
    uint64_t common = ~0ull, separate = 0;
    for (int q = 0; q < m; ++q) {
        common &= s->chain[q].ctrl_mask;
        separate += (s->len >> __builtin_popcountll(s->chain[q].ctrl_mask)) + kLaunchCostAmps;
    }
    const uint64_t swept = (s->len >> __builtin_popcountll(common)) + kLaunchCostAmps;
    if (m == 1 || swept > separate) {   // one gate, or gates on small disjoint subspaces of a large state
        for (int q = 0; q < m; ++q) {
            const int rc = launch_chain_gate<R>(s, s->chain[q]);
            if (rc) return rc;
        }
        return QIPGPU_OK;
    }
    DiagChainArgs<R> a;
    std::memset(&a, 0, sizeof(a));
    a.amps = (cx<R> *)s->amps;
    a.common_mask = common;
    std::vector<int> cpos;
    for (int p = 0; p < s->n; ++p)
        if (common >> p & 1) cpos.push_back(p);
    a.count = s->len >> cpos.size();
    fill_poslist(a.pl, cpos);
    a.ngates = m;
    for (int q = 0; q < m; ++q) a.g[q] = s->chain[q];
    constexpr int ITEMS = 4;
    const uint32_t nb = blocks_for(a.count, 256 * ITEMS);
    a.order = block_order(s->len * s->amp_bytes, s->amp_bytes, nb);
    a.nblocks = nb;
    return launch_gate(s, &k_diag_chain<R, ITEMS>, dim3(capped_grid(nb)), dim3(256), a);
}

// A run of held-back single-qubit gates whose target and control bits all lie in one tile of 4096
// amplitudes (tile bits P: those bits, state bits 0..3 and free bits, ascending; u bit i <-> state
// bit P[i]). A workgroup loads a tile into LDS and applies the gates to it one after the other, each
// to every pair of the tile with k_dense's arithmetic (FP64 from zero, rounded to the state's
// precision), then stores the tile: the bits of a launch per gate, for one HBM pass.
struct RunSlot {
    cx<double> U[4];
    uint32_t cmask;     // control bits, in tile (u) coordinates
    uint32_t tb;        // target bit, in tile coordinates
};

template <typename R>
struct TileRunArgs {
    cx<R> *amps;
    uint64_t ntiles;
    int32_t ngates;
    uint8_t pbits[12];
    RunSlot g[kTileRunMax];
};

template <typename R>
__device__ __forceinline__ uint64_t run_tile_base(uint64_t t, const TileRunArgs<R> &a) {
#pragma unroll
    for (int i = 0; i < 12; ++i) {
        const uint32_t p = a.pbits[i];
        const uint64_t lo = t & ((1ull << p) - 1ull);
        t = ((t ^ lo) << 1) | lo;
    }
    return t;
}

// 512 threads per workgroup (two waves per SIMD from each workgroup, so one wave's LDS and FP64
// latency overlaps another's). Thread t holds the tile's amplitudes u = t + 512 i, i = 0..7: u bits
// 0..8 are t, bits 9..11 are i.
constexpr uint32_t kRunThreads = 512;

template <typename R>
__device__ __forceinline__ uint64_t run_off_hi(int i, const TileRunArgs<R> &a) {
    uint64_t o = 0;
#pragma unroll
    for (int b = 0; b < 3; ++b) o |= (uint64_t)((i >> b) & 1) << a.pbits[9 + b];
    return o;
}

// One gate on one pair (x0: target bit 0, x1: target bit 1), as k_dense<R,1> computes it: each output
// an FP64 sum from zero of U's row times the pair (cfma), rounded to the state's precision.
template <typename R>
__device__ __forceinline__ void run_gate(const RunSlot &g, cx<R> &x0, cx<R> &x1) {
    cx<double> acc0, acc1;
    acc0.re = acc0.im = acc1.re = acc1.im = 0.0;
    cfma(acc0, g.U[0], widen(x0));
    cfma(acc0, g.U[1], widen(x1));
    cfma(acc1, g.U[2], widen(x0));
    cfma(acc1, g.U[3], widen(x1));
    x0 = narrow<R>(acc0);
    x1 = narrow<R>(acc1);
}

// A thread's 8 amplitudes of a tile in registers, as separate real and imaginary parts (plain scalar
// arrays stay in registers where arrays of the complex struct would go to scratch memory).
template <typename R>
struct RunRegs {
    R re[8], im[8];
};

template <typename R>
__device__ __forceinline__ void run_load(RunRegs<R> &v, uint64_t base, const TileRunArgs<R> &a) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const cx<R> x = a.amps[base + run_off_hi(i, a)];
        v.re[i] = x.re;
        v.im[i] = x.im;
    }
}

template <typename R>
__device__ __forceinline__ void run_tile(const RunRegs<R> &v, uint64_t base, cx<R> *t, const TileRunArgs<R> &a) {
    const uint32_t tid = threadIdx.x;
    __syncthreads();  // the previous tile has left LDS
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        cx<R> x;
        x.re = v.re[i];
        x.im = v.im[i];
        t[tid + kRunThreads * i] = x;
    }
    __syncthreads();
    int gi = 0;
    // two gates per LDS round trip: each thread owns whole quads {u, u^b1, u^b2, u^b1^b2} (or pairs,
    // when both gates act on one bit) and applies the first gate to them, then the second
    for (; gi + 1 < a.ngates; gi += 2) {
        const RunSlot &g1 = a.g[gi], &g2 = a.g[gi + 1];
        const uint32_t b1 = 1u << g1.tb, b2 = 1u << g2.tb;
        if (b1 == b2) {
            const uint32_t lo = b1 - 1u;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t p = tid + kRunThreads * j;    // pair index 0..2047
                const uint32_t u0 = ((p & ~lo) << 1) | (p & lo);
                cx<R> x0 = t[u0], x1 = t[u0 | b1];
                if ((u0 & g1.cmask) == g1.cmask) run_gate(g1, x0, x1);
                if ((u0 & g2.cmask) == g2.cmask) run_gate(g2, x0, x1);
                t[u0] = x0;
                t[u0 | b1] = x1;
            }
        } else {
            const uint32_t bl = b1 < b2 ? b1 : b2, bh = b1 < b2 ? b2 : b1;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                uint32_t q = tid + kRunThreads * j;          // quad index 0..1023
                q = ((q & ~(bl - 1u)) << 1) | (q & (bl - 1u));
                const uint32_t u = ((q & ~(bh - 1u)) << 1) | (q & (bh - 1u));
                cx<R> a00 = t[u], a10 = t[u | b1], a01 = t[u | b2], a11 = t[u | b1 | b2];
                if ((u & g1.cmask) == g1.cmask) run_gate(g1, a00, a10);
                if (((u | b2) & g1.cmask) == g1.cmask) run_gate(g1, a01, a11);
                if ((u & g2.cmask) == g2.cmask) run_gate(g2, a00, a01);
                if (((u | b1) & g2.cmask) == g2.cmask) run_gate(g2, a10, a11);
                t[u] = a00;
                t[u | b1] = a10;
                t[u | b2] = a01;
                t[u | b1 | b2] = a11;
            }
        }
        __syncthreads();
    }
    if (gi < a.ngates) {
        const RunSlot &g = a.g[gi];
        const uint32_t b = 1u << g.tb, lo = b - 1u;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t p = tid + kRunThreads * j;
            const uint32_t u0 = ((p & ~lo) << 1) | (p & lo);
            if ((u0 & g.cmask) != g.cmask) continue;
            cx<R> x0 = t[u0], x1 = t[u0 | b];
            run_gate(g, x0, x1);
            t[u0] = x0;
            t[u0 | b] = x1;
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) a.amps[base + run_off_hi(i, a)] = t[tid + kRunThreads * i];
}

// Contiguous tile ranges per workgroup, two (complex128) or four (complex64) workgroups per CU by LDS.
template <typename R>
__global__ __launch_bounds__(kRunThreads, 2) void k_tile_run(const TileRunArgs<R> a) {
    __shared__ cx<R> t[4096];
    const uint32_t tid = threadIdx.x;
    uint64_t off_lo = 0;
#pragma unroll
    for (int b = 0; b < 9; ++b) off_lo |= (uint64_t)((tid >> b) & 1u) << a.pbits[b];
    const uint64_t per = (a.ntiles + gridDim.x - 1) / gridDim.x;
    uint64_t tile = (uint64_t)blockIdx.x * per;
    const uint64_t end = min(tile + per, a.ntiles);
    if (tile >= end) return;   // the same for every thread of the workgroup
    const uint64_t last = end - 1;
    // two register buffers alternate: the next tile's loads are in flight while one is processed
    // (the last prefetch re-reads an in-range tile)
    RunRegs<R> va, vb;
    run_load(va, run_tile_base(tile, a) + off_lo, a);
    for (;;) {
        run_load(vb, run_tile_base(min(tile + 1, last), a) + off_lo, a);
        run_tile(va, run_tile_base(tile, a) + off_lo, t, a);
        if (++tile >= end) break;
        run_load(va, run_tile_base(min(tile + 1, last), a) + off_lo, a);
        run_tile(vb, run_tile_base(tile, a) + off_lo, t, a);
        if (++tile >= end) break;
    }
}

template <typename R>
static int launch_run_gate(qipgpu_state *s, const RunGate &g) {
    std::vector<int> tbits(1, g.tpos), allpos;
    for (int p = 0; p < 64; ++p)
        if (g.ctrl_mask >> p & 1) allpos.push_back(p);
    allpos.push_back(g.tpos);
    std::vector<cd> U(4);
    for (int i = 0; i < 4; ++i) U[i] = cd(g.U[i].re, g.U[i].im);
    return launch_dense_reg<R, 1>(s, tbits, g.ctrl_mask, allpos, U);
}

template <typename R>
static int flush_run_typed(qipgpu_state *s) {
    const int m = s->run_len;
    const uint64_t bits = s->run_bits;
    s->run_len = 0;
    s->run_bits = 0;
    uint64_t separate = 0;
    for (int q = 0; q < m; ++q) separate += (s->len >> __builtin_popcountll(s->run[q].ctrl_mask)) + kLaunchCostAmps;
    if (m == 1 || s->len + kLaunchCostAmps > separate) {   // one gate, or gates on small subspaces
        for (int q = 0; q < m; ++q) {
            const int rc = launch_run_gate<R>(s, s->run[q]);
            if (rc) return rc;
        }
        return QIPGPU_OK;
    }
    // tile bits: the gates' bits, state bits 0..3, then free bits in the order 4, 5, 12..17, 6..11, 18..
    uint64_t pmask = bits | 0xfull;
    int order[64], no = 0;
    for (int b = 4; b < 6; ++b) order[no++] = b;
    for (int b = 12; b < 18; ++b) order[no++] = b;
    for (int b = 6; b < 12; ++b) order[no++] = b;
    for (int b = 18; b < 64; ++b) order[no++] = b;
    for (int i = 0; i < no && __builtin_popcountll(pmask) < 12; ++i)
        if (order[i] < s->n) pmask |= 1ull << order[i];
    TileRunArgs<R> a;
    std::memset(&a, 0, sizeof(a));
    int ub[64];
    for (int p = 0, i = 0; p < 64; ++p)
        if (pmask >> p & 1) {
            a.pbits[i] = (uint8_t)p;
            ub[p] = i++;
        }
    a.amps = (cx<R> *)s->amps;
    a.ntiles = s->len >> 12;
    a.ngates = m;
    for (int q = 0; q < m; ++q) {
        const RunGate &g = s->run[q];
        for (int i = 0; i < 4; ++i) a.g[q].U[i] = g.U[i];
        a.g[q].tb = (uint32_t)ub[g.tpos];
        uint32_t cm = 0;
        for (int p = 0; p < 64; ++p)
            if (g.ctrl_mask >> p & 1) cm |= 1u << ub[p];
        a.g[q].cmask = cm;
    }
    const uint64_t per_cu = sizeof(R) == 8 ? 2 : 4;   // 64 KB / 32 KB of LDS per workgroup
    const uint32_t grid = (uint32_t)std::min<uint64_t>(a.ntiles, (uint64_t)device_cus() * per_cu);
    return launch_gate(s, &k_tile_run<R>, dim3(grid), dim3(kRunThreads), a);
}

// Hold a single-qubit gate (target bit tpos, under ctrl_mask) back on the state: held diagonal gates
// and a run it does not fit (more than 8 bits above bit 3, or full) are applied first.
static int hold_single(qipgpu_state *s, int tpos, uint64_t ctrl_mask, const std::vector<cd> &U) {
    const uint64_t gb = ctrl_mask | (1ull << tpos);
    if (s->chain_len || s->run_len == kTileRunMax || __builtin_popcountll((s->run_bits | gb) & ~0xfull) > 8) {
        const int rc = flush_held(s);
        if (rc) return rc;
    }
    RunGate &g = s->run[s->run_len++];
    std::memset(&g, 0, sizeof(g));
    g.ctrl_mask = ctrl_mask;
    g.tpos = (uint8_t)tpos;
    for (int i = 0; i < 4; ++i) {
        g.U[i].re = U[i].real();
        g.U[i].im = U[i].imag();
    }
    s->run_bits |= gb;
    return QIPGPU_OK;
}

// Hold a diagonal gate (k <= 2 targets) back on the state; held single-qubit gates and a full run
// are applied first.
static int hold_diagonal(qipgpu_state *s, const std::vector<int> &tbits, uint64_t ctrl_mask, const std::vector<cd> &d) {
    if (s->run_len || s->chain_len == kDiagChainMax) {
        const int rc = flush_held(s);
        if (rc) return rc;
    }
    ChainGate &c = s->chain[s->chain_len++];
    std::memset(&c, 0, sizeof(c));
    c.ctrl_mask = ctrl_mask;
    c.k = (uint8_t)tbits.size();
    for (size_t b = 0; b < tbits.size(); ++b) c.tpos[b] = (uint8_t)tbits[b];
    for (size_t j = 0; j < d.size(); ++j) {
        c.d[j].re = d[j].real();
        c.d[j].im = d[j].imag();
    }
    return QIPGPU_OK;
}

template <typename R>
static int dispatch(qipgpu_state *s, const Analysed &an, const std::vector<int32_t> &controls_in) {
    const int n = s->n;
    std::vector<int> tbits, allpos, cpos;
    uint64_t ctrl_mask = 0;
    for (int32_t q : an.targets) tbits.push_back(n - 1 - q);
    std::vector<int32_t> controls(controls_in);
    controls.insert(controls.end(), an.controls.begin(), an.controls.end());
    for (int32_t q : controls) {
        cpos.push_back(n - 1 - q);
        ctrl_mask |= 1ull << (n - 1 - q);
    }
    allpos = cpos;
    allpos.insert(allpos.end(), tbits.begin(), tbits.end());
    const int k = (int)an.targets.size();
    if (an.diagonal) {
        const int D = 1 << k;
        std::vector<cd> d(D);
        bool identity = true;
        for (int j = 0; j < D; ++j) {
            d[j] = an.U[j * D + j];
            identity = identity && d[j] == 1.0;
        }
        if (identity) return QIPGPU_OK;
        if (k <= 2) return hold_diagonal(s, tbits, ctrl_mask, d);
        const int rc = flush_held(s);
        if (rc) return rc;
        switch (k) {
            case 0: return launch_diag<R, 0>(s, tbits, ctrl_mask, cpos, d);
            case 1: return launch_diag<R, 1>(s, tbits, ctrl_mask, cpos, d);
            case 2: return launch_diag<R, 2>(s, tbits, ctrl_mask, cpos, d);
            case 3: return launch_diag<R, 3>(s, tbits, ctrl_mask, cpos, d);
            case 4: return launch_diag<R, 4>(s, tbits, ctrl_mask, cpos, d);
            case 5: return launch_diag<R, 5>(s, tbits, ctrl_mask, cpos, d);
            case 6: return launch_diag<R, 6>(s, tbits, ctrl_mask, cpos, d);
            default: break;
        }
        return diag_large(s, tbits, ctrl_mask, cpos, d.data());
    }
    if (k == 1 && s->n >= 12 && __builtin_popcountll((ctrl_mask | (1ull << tbits[0])) & ~0xfull) <= 8)
        return hold_single(s, tbits[0], ctrl_mask, an.U);
    const int rc = flush_held(s);
    if (rc) return rc;
    switch (k) {
        case 1: return launch_dense_reg<R, 1>(s, tbits, ctrl_mask, allpos, an.U);
        case 2: return launch_dense_reg<R, 2>(s, tbits, ctrl_mask, allpos, an.U);
        case 3: return launch_dense_reg<R, 3>(s, tbits, ctrl_mask, allpos, an.U);
        default: break;
    }
    return dense_large(s, tbits, ctrl_mask, allpos, an.U.data());
}

static int check_disjoint(const qipgpu_state *s, int32_t k, const int32_t *t, int32_t c, const int32_t *ctl) {
    std::vector<int32_t> all(t, t + k);
    all.insert(all.end(), ctl, ctl + c);
    return check_qubits(s, (int32_t)all.size(), all.data(), "operator qubits");
}

}  // namespace

namespace qipgpu {

int flush_held(qipgpu_state *s) {
    const bool c128 = s->dtype == QIPGPU_C128;
    if (s->chain_len) {
        const int rc = c128 ? flush_chain_typed<double>(s) : flush_chain_typed<float>(s);
        if (rc) return rc;
    }
    if (s->run_len) return c128 ? flush_run_typed<double>(s) : flush_run_typed<float>(s);
    return QIPGPU_OK;
}

}  // namespace qipgpu

extern "C" {

int qipgpu_apply_matrix(qipgpu_state *s, int32_t k, const int32_t *targets, const double *mat, int32_t n_ctrl,
                        const int32_t *controls) {
    if (k < 0 || k > 12) return set_error(QIPGPU_EINVAL, "operator on %d qubits: at most 12 accepted", k);
    if (n_ctrl < 0) return set_error(QIPGPU_EINVAL, "negative control count");
    int rc = check_disjoint(s, k, targets, n_ctrl, controls);
    if (rc) return rc;
    // linear in the state: a pending fused renormalisation (a scalar) stays pending; recorded gates
    // go first (with the diagonal gates held back before them); held-back diagonal gates otherwise
    // stay held, and dispatch applies them before anything that is not held
    rc = recorder_pending(s) ? flush_recorded(s) : QIPGPU_OK;
    if (rc) return rc;
    Analysed an = analyse(k, targets, mat);
    std::vector<int32_t> ctl(controls, controls + n_ctrl);
    return s->dtype == QIPGPU_C128 ? dispatch<double>(s, an, ctl) : dispatch<float>(s, an, ctl);
}

int qipgpu_apply_diagonal(qipgpu_state *s, int32_t k, const int32_t *targets, const double *diag, int32_t n_ctrl,
                          const int32_t *controls) {
    if (k < 0 || k > QIPGPU_MAX_DENSE) return set_error(QIPGPU_EINVAL, "diagonal on %d qubits", k);
    int rc = check_disjoint(s, k, targets, n_ctrl, controls);
    if (rc) return rc;
    // as in qipgpu_apply_matrix
    rc = recorder_pending(s) ? flush_recorded(s) : QIPGPU_OK;
    if (rc) return rc;
    Analysed an;
    an.targets.assign(targets, targets + k);
    const int D = 1 << k;
    an.U.assign((size_t)D * D, cd(0.0, 0.0));
    for (int j = 0; j < D; ++j) an.U[j * D + j] = cd(diag[2 * j], diag[2 * j + 1]);
    an.diagonal = true;
    std::vector<int32_t> ctl(controls, controls + n_ctrl);
    return s->dtype == QIPGPU_C128 ? dispatch<double>(s, an, ctl) : dispatch<float>(s, an, ctl);
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------
// func_apply (qip/ext/func_apply.pyx:11-112), in place
// ------------------------------------------------------------------------------------------------

struct qipgpu_table {
    int32_t n_reg1;
    int32_t *d;
    uint64_t nonzero;   // entries f(x) != 0
};

namespace {

uint64_t count_nonzero(const int32_t *table, uint64_t count) {
    uint64_t nz = 0;
    for (uint64_t i = 0; i < count; ++i) nz += table[i] != 0;
    return nz;
}

}  // namespace

// For x with y = f(x) & (2^n2 - 1) != 0 the map q -> q xor y pairs the reg2 values; each thread
// owns one pair (q with bit hb(y) clear, q xor y) for one rest value r and swaps the two
// amplitudes, which is exactly output[x, q^y, r] = input[x, q, r]. x with y == 0 is untouched.
struct FuncArgs {
    uint64_t nthreads;
    int32_t n1, n2, nother, rbits;
    uint32_t ymask;
    uint8_t xpos[QIPGPU_MAX_QUBITS];  // value bit b -> state bit
    uint8_t qpos[QIPGPU_MAX_QUBITS];
    uint8_t rpos[QIPGPU_MAX_QUBITS];
};

template <typename R>
__global__ __launch_bounds__(256) void k_func_apply(cx<R> *amps, const int32_t *__restrict__ table, FuncArgs a) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < a.nthreads; t += stride) {
    const uint64_t x = t >> a.rbits;
    const uint32_t y = (uint32_t)table[x] & a.ymask;
    if (y == 0) continue;
    const uint64_t rest = t & ((1ull << a.rbits) - 1ull);
    const int hb = 31 - __clz(y);
    const uint64_t qlow = rest & ((1ull << (a.n2 - 1)) - 1ull);
    const uint64_t r = rest >> (a.n2 - 1);
    const uint64_t q = (((qlow >> hb) << (hb + 1)) | (qlow & ((1ull << hb) - 1ull)));
    const uint64_t ia = scatter_bits(x, a.xpos, a.n1) | scatter_bits(q, a.qpos, a.n2) | scatter_bits(r, a.rpos, a.nother);
    const uint64_t ib = ia ^ scatter_bits(y, a.qpos, a.n2);
    const cx<R> va = amps[ia], vb = amps[ib];
    amps[ia] = vb;
    amps[ib] = va;
    }
}

// The same permutation tile by tile (n >= 12, reg2 of at most 8 qubits, reg1 of at most 31). A tile is the 4096
// amplitudes spanned by the reg2 state bits (its rows, q) and the lowest other state bits (its
// columns); each column has one x, so its rows move by q -> q xor y(x). The block reads the y of
// every column; a tile where all of them are 0 is left untouched, else it is read in address order
// into LDS and written back in address order with out[q, c] = in[q xor y(c), c]. Every amplitude of a
// moved tile is read and written once, in runs of at least 2^(lowest non-reg2 bits) amplitudes.
constexpr int kFuncTileBits = 12;

struct FuncTileArgs {
    uint64_t ntiles;
    uint32_t ymask;
    int32_t n1, n2, ncol_bits;             // ncol_bits = 12 - n2
    uint8_t pbits[kFuncTileBits];          // u bit i <-> state bit pbits[i], ascending
    uint8_t ru[8];                         // u bit of reg2 value bit b
    uint8_t cu[kFuncTileBits];             // u bit of column bit b, ascending
    uint8_t xpos[QIPGPU_MAX_QUBITS];       // reg1 value bit b -> state bit
};

__device__ __forceinline__ uint64_t func_tile_base(uint64_t t, const FuncTileArgs &a) {
#pragma unroll
    for (int i = 0; i < kFuncTileBits; ++i) {
        const uint32_t p = a.pbits[i];
        const uint64_t lo = t & ((1ull << p) - 1ull);
        t = ((t ^ lo) << 1) | lo;
    }
    return t;
}

template <int LO, int HI>
__device__ __forceinline__ uint64_t func_u_offset(uint32_t u, const FuncTileArgs &a) {
    uint64_t o = 0;
#pragma unroll
    for (int i = LO; i < HI; ++i) o |= (uint64_t)((u >> i) & 1u) << a.pbits[i];
    return o;
}

// column index of tile position u (its bits at the column positions, in order)
__device__ __forceinline__ uint32_t func_col(uint32_t u, const FuncTileArgs &a) {
    uint32_t c = 0;
    for (int b = 0; b < a.ncol_bits; ++b) c |= ((u >> a.cu[b]) & 1u) << b;
    return c;
}

template <typename R>
__global__ __launch_bounds__(256) void k_func_tile(cx<R> *amps, const int32_t *__restrict__ table, FuncTileArgs a) {
    __shared__ cx<R> t[1 << kFuncTileBits];
    __shared__ uint32_t xcol[1 << (kFuncTileBits - 1)];   // the x bits of each column (at most 2048)
    __shared__ uint8_t ys[1 << (kFuncTileBits - 1)];      // y of each column of the tile (< 2^8)
    __shared__ uint16_t rowx[256];                        // y -> its bits on the row positions of u
    const uint32_t tid = threadIdx.x;
    const uint32_t ncol = 1u << a.ncol_bits;
    {
        uint32_t ry = 0;
        for (int b = 0; b < a.n2; ++b) ry |= ((tid >> b) & 1u) << a.ru[b];
        rowx[tid] = (uint16_t)ry;
    }
    // x = (x bits of the tile) | (x bits of the column): the column part once per kernel
    for (uint32_t c = tid; c < ncol; c += 256) {
        uint32_t u = 0;
        for (int b = 0; b < a.ncol_bits; ++b) u |= ((c >> b) & 1u) << a.cu[b];
        const uint64_t st = func_u_offset<0, kFuncTileBits>(u, a);
        uint32_t x = 0;
        for (int b = 0; b < a.n1; ++b) x |= (uint32_t)((st >> a.xpos[b]) & 1ull) << b;
        xcol[c] = x;
    }
    const uint64_t off_t = func_u_offset<0, 8>(tid, a);
    const uint32_t col_t = func_col(tid, a);            // column bits among u bits 0..7
    uint32_t col_j[16];                                 // column bits among u bits 8..11, per j
#pragma unroll
    for (int j = 0; j < 16; ++j) col_j[j] = col_t | func_col(256u * j, a);
    const uint64_t per = (a.ntiles + gridDim.x - 1) / gridDim.x;
    uint64_t tile = (uint64_t)blockIdx.x * per;
    const uint64_t end = min(tile + per, a.ntiles);
    __syncthreads();
    for (; tile < end; ++tile) {
        const uint64_t tb = func_tile_base(tile, a);
        uint32_t xt = 0;
        for (int b = 0; b < a.n1; ++b) xt |= (uint32_t)((tb >> a.xpos[b]) & 1ull) << b;
        int any = 0;
        for (uint32_t c = tid; c < ncol; c += 256) {
            const uint32_t y = (uint32_t)table[xt | xcol[c]] & a.ymask;
            ys[c] = (uint8_t)y;
            any |= y != 0u;
        }
        if (!__syncthreads_or(any)) continue;   // nothing moves in this tile
        cx<R> v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = amps[tb + off_t + func_u_offset<8, kFuncTileBits>(256u * j, a)];
#pragma unroll
        for (int j = 0; j < 16; ++j) t[tid + 256u * j] = v[j];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = t[(tid + 256u * j) ^ rowx[ys[col_j[j]]]];
#pragma unroll
        for (int j = 0; j < 16; ++j) amps[tb + off_t + func_u_offset<8, kFuncTileBits>(256u * j, a)] = v[j];
        __syncthreads();   // t and ys are rewritten for the next tile
    }
}

// Collapse (kronprod.pyx:439-444), in place: amplitudes whose measured bits equal `target` are
// scaled, the rest are zeroed without being read.
template <typename R>
__global__ __launch_bounds__(256) void k_collapse(cx<R> *amps, uint64_t len, uint64_t mask, uint64_t target, double mult) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
        cx<R> o;
        if ((i & mask) == target) {
            const cx<R> v = amps[i];
            o.re = (R)((double)v.re * mult);
            o.im = (R)((double)v.im * mult);
        } else {
            o.re = 0;
            o.im = 0;
        }
        amps[i] = o;
    }
}

namespace {

// The reference's rest-qubit loop (func_apply.pyx:107-111) sets only the first n_reg2 rest qubits,
// in qubit order, into its output index; the others stay 0 there, so every output amplitude with
// one of them at 1 keeps the 0 it was cleared to (func_apply.pyx:42-44). The permutation leaves the
// rest bits where they are, so that output is the whole permutation with those amplitudes zeroed:
// k_collapse with the dropped rest bits as the mask, target 0 and factor 1 (x * 1.0 == x).
static int zero_dropped_rest(qipgpu_state *s, const std::vector<int> &other, int32_t n_reg2) {
    uint64_t mask = 0;
    for (size_t j = (size_t)n_reg2; j < other.size(); ++j) mask |= 1ull << (s->n - 1 - other[j]);
    if (!mask) return QIPGPU_OK;
    const uint32_t nb = std::min<uint32_t>(blocks_for(s->len, 256 * 4), 1u << 20);
    if (s->dtype == QIPGPU_C128)
        hipLaunchKernelGGL(k_collapse<double>, dim3(nb), dim3(256), 0, s->stream, (cx<double> *)s->amps, s->len, mask,
                           (uint64_t)0, 1.0);
    else
        hipLaunchKernelGGL(k_collapse<float>, dim3(nb), dim3(256), 0, s->stream, (cx<float> *)s->amps, s->len, mask,
                           (uint64_t)0, 1.0);
    QIP_LAUNCHED();
    return QIPGPU_OK;
}

static int permute_reg2(qipgpu_state *s, int32_t n_reg1, int32_t n_reg2, const int32_t *d_table, uint64_t nonzero,
                        const FuncArgs &a);

// `nonzero` = entries of the table that are not 0: where fewer than one x in 64 moves anything, the
// per-pair kernel (which touches only the moved pairs) runs instead of the tiles.
static int func_apply_dev(qipgpu_state *s, int32_t n_reg1, const int32_t *reg1, int32_t n_reg2, const int32_t *reg2,
                          const int32_t *d_table, uint64_t nonzero, int32_t flags) {
    QIP_NOT_CAPTURING(s, "func_apply (its f table lives outside the graph)");
    const int n = s->n;
    if (n_reg1 < 1 || n_reg2 < 1 || n_reg1 + n_reg2 > n)
        return set_error(QIPGPU_EINVAL, "func_apply: register sizes %d + %d for n=%d", n_reg1, n_reg2, n);
    if (flags & ~QIPGPU_FUNC_REFERENCE_BOUND) return set_error(QIPGPU_EINVAL, "func_apply: unknown flags %d", flags);
    std::vector<int32_t> all(reg1, reg1 + n_reg1);
    all.insert(all.end(), reg2, reg2 + n_reg2);
    int rc = check_qubits(s, (int32_t)all.size(), all.data(), "func_apply registers");
    if (rc) return rc;
    // linear in the state: a pending fused renormalisation (a scalar) stays pending; recorded gates
    // go first
    rc = flush_recorded(s);
    if (rc) return rc;
    FuncArgs a;
    std::memset(&a, 0, sizeof(a));
    a.n1 = n_reg1;
    a.n2 = n_reg2;
    a.nother = n - n_reg1 - n_reg2;
    a.rbits = n - n_reg1 - 1;
    a.ymask = n_reg2 >= 32 ? 0xffffffffu : ((1u << n_reg2) - 1u);
    a.nthreads = 1ull << (n - 1);
    // x is read MSB-first from reg1 (func_apply.pyx:82-85): value bit b <- reg1[n1-1-b]
    for (int b = 0; b < n_reg1; ++b) a.xpos[b] = (uint8_t)(n - 1 - reg1[n_reg1 - 1 - b]);
    for (int b = 0; b < n_reg2; ++b) a.qpos[b] = (uint8_t)(n - 1 - reg2[n_reg2 - 1 - b]);
    // rest qubits ascending (func_apply.pyx:58-60), value MSB-first over them
    std::vector<int> other;
    uint64_t used = 0;
    for (int32_t q : all) used |= 1ull << q;
    for (int q = 0; q < n; ++q)
        if (!(used & (1ull << q))) other.push_back(q);
    for (int b = 0; b < a.nother; ++b) a.rpos[b] = (uint8_t)(n - 1 - other[a.nother - 1 - b]);
    rc = permute_reg2(s, n_reg1, n_reg2, d_table, nonzero, a);
    if (rc) return rc;
    return (flags & QIPGPU_FUNC_REFERENCE_BOUND) ? zero_dropped_rest(s, other, n_reg2) : QIPGPU_OK;
}

// The permutation itself: |x>|q>|r> -> |x>|f(x) xor q>|r> over the whole state, on the tile kernel or
// the per-pair kernel.
static int permute_reg2(qipgpu_state *s, int32_t n_reg1, int32_t n_reg2, const int32_t *d_table, uint64_t nonzero,
                        const FuncArgs &a) {
    const int n = s->n;
    if (n >= kFuncTileBits && n_reg2 <= 8 && n_reg1 <= 31 && nonzero * 64 >= (1ull << n_reg1)) {
        FuncTileArgs f;
        std::memset(&f, 0, sizeof(f));
        f.ntiles = 1ull << (n - kFuncTileBits);
        f.ymask = a.ymask;
        f.n1 = n_reg1;
        f.n2 = n_reg2;
        f.ncol_bits = kFuncTileBits - n_reg2;
        std::memcpy(f.xpos, a.xpos, sizeof(f.xpos));
        // tile bits: the reg2 bits and the lowest other bits
        uint64_t rmask = 0;
        for (int b = 0; b < n_reg2; ++b) rmask |= 1ull << a.qpos[b];
        std::vector<int> pb;
        for (int b = 0; b < n_reg2; ++b) pb.push_back(a.qpos[b]);
        for (int b = 0; b < n && (int)pb.size() < kFuncTileBits; ++b)
            if (!(rmask >> b & 1)) pb.push_back(b);
        std::sort(pb.begin(), pb.end());
        for (int i = 0; i < kFuncTileBits; ++i) f.pbits[i] = (uint8_t)pb[i];
        for (int b = 0; b < n_reg2; ++b)
            f.ru[b] = (uint8_t)(std::find(pb.begin(), pb.end(), (int)a.qpos[b]) - pb.begin());
        int nc = 0;
        for (int i = 0; i < kFuncTileBits; ++i)
            if (!(rmask >> pb[i] & 1)) f.cu[nc++] = (uint8_t)i;
        const uint32_t nb = (uint32_t)std::min<uint64_t>(f.ntiles, (uint64_t)device_cus() * 2);
        if (s->dtype == QIPGPU_C128)
            hipLaunchKernelGGL(k_func_tile<double>, dim3(nb), dim3(256), 0, s->stream, (cx<double> *)s->amps, d_table, f);
        else
            hipLaunchKernelGGL(k_func_tile<float>, dim3(nb), dim3(256), 0, s->stream, (cx<float> *)s->amps, d_table, f);
        QIP_LAUNCHED();
        return QIPGPU_OK;
    }
    const uint32_t nb = capped_grid(blocks_for(a.nthreads, 256));
    if (s->dtype == QIPGPU_C128)
        hipLaunchKernelGGL(k_func_apply<double>, dim3(nb), dim3(256), 0, s->stream, (cx<double> *)s->amps, d_table, a);
    else
        hipLaunchKernelGGL(k_func_apply<float>, dim3(nb), dim3(256), 0, s->stream, (cx<float> *)s->amps, d_table, a);
    QIP_LAUNCHED();
    return QIPGPU_OK;
}

}  // namespace

extern "C" {

int qipgpu_table_create(int32_t n_reg1, const int32_t *table, qipgpu_table **out) {
    *out = nullptr;
    if (n_reg1 < 1 || n_reg1 > QIPGPU_MAX_QUBITS) return set_error(QIPGPU_EINVAL, "table for %d-qubit register", n_reg1);
    qipgpu_table *t = new qipgpu_table();
    t->n_reg1 = n_reg1;
    t->nonzero = count_nonzero(table, 1ull << n_reg1);
    const size_t bytes = sizeof(int32_t) << n_reg1;
    if (hipMalloc(&t->d, bytes) != hipSuccess) {
        (void)hipGetLastError();
        delete t;
        return set_error(QIPGPU_ENOMEM, "cannot allocate %zu-byte function table", bytes);
    }
    if (hipMemcpy(t->d, table, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(t->d);
        delete t;
        return set_error(QIPGPU_EHIP, "function table upload failed");
    }
    *out = t;
    return QIPGPU_OK;
}

int qipgpu_table_destroy(qipgpu_table *t) {
    if (!t) return QIPGPU_OK;
    (void)hipDeviceSynchronize();
    (void)hipFree(t->d);
    delete t;
    return QIPGPU_OK;
}

int qipgpu_func_apply_table(qipgpu_state *s, int32_t n_reg1, const int32_t *reg1, int32_t n_reg2, const int32_t *reg2,
                            const qipgpu_table *t) {
    if (!t || t->n_reg1 != n_reg1)
        return set_error(QIPGPU_EINVAL, "function table is for a %d-qubit register, not %d", t ? t->n_reg1 : -1, n_reg1);
    return func_apply_dev(s, n_reg1, reg1, n_reg2, reg2, t->d, t->nonzero, 0);
}

int qipgpu_func_apply_table_ex(qipgpu_state *s, int32_t n_reg1, const int32_t *reg1, int32_t n_reg2,
                               const int32_t *reg2, const qipgpu_table *t, int32_t flags) {
    if (!t || t->n_reg1 != n_reg1)
        return set_error(QIPGPU_EINVAL, "function table is for a %d-qubit register, not %d", t ? t->n_reg1 : -1, n_reg1);
    return func_apply_dev(s, n_reg1, reg1, n_reg2, reg2, t->d, t->nonzero, flags);
}

int qipgpu_func_apply(qipgpu_state *s, int32_t n_reg1, const int32_t *reg1, int32_t n_reg2, const int32_t *reg2,
                      const int32_t *table) {
    if (n_reg1 < 1 || n_reg1 > s->n) return set_error(QIPGPU_EINVAL, "func_apply: register size %d", n_reg1);
    const size_t bytes = sizeof(int32_t) << n_reg1;
    int rc = ensure_scratch(s, bytes);
    if (rc) return rc;
    QIP_HIP(hipMemcpyAsync(s->scratch, table, bytes, hipMemcpyHostToDevice, s->stream));
    rc = func_apply_dev(s, n_reg1, reg1, n_reg2, reg2, (const int32_t *)s->scratch,
                        count_nonzero(table, 1ull << n_reg1), 0);
    const hipError_t e = hipStreamSynchronize(s->stream);   // the host table is read before return
    if (rc) return rc;
    if (e != hipSuccess) return set_error(QIPGPU_EHIP, "hipStreamSynchronize failed: %s", hipGetErrorString(e));
    return QIPGPU_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------
// reductions
// ------------------------------------------------------------------------------------------------

constexpr uint32_t kNormBlocks = 2048;

template <typename R>
__global__ __launch_bounds__(256) void k_norm_partial(const cx<R> *__restrict__ amps, uint64_t len, double *partial) {
    __shared__ double lds[4];
    const uint64_t per = (len + gridDim.x - 1) / gridDim.x;
    const uint64_t lo = (uint64_t)blockIdx.x * per, hi = min(len, lo + per);
    double acc = 0.0;
    uint64_t i = lo + threadIdx.x;
    // sixteen independent loads per step keep enough bytes in flight to fill HBM
    for (; i + 256 * 15 < hi; i += 256 * 16) {
        cx<R> y[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) y[j] = amps[i + 256 * j];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const double re = y[j].re, im = y[j].im;
            acc = fma(re, re, acc);
            acc = fma(im, im, acc);
        }
    }
    for (; i < hi; i += 256) {
        const cx<R> v = amps[i];
        const double re = v.re, im = v.im;
        acc = fma(re, re, acc);
        acc = fma(im, im, acc);
    }
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// Sum `count` partials in a fixed order with one block; result to dst[0].
__global__ __launch_bounds__(256) void k_sum_final(const double *__restrict__ partial, uint32_t count, double *dst) {
    __shared__ double lds[4];
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < count; i += 256) acc += partial[i];
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) dst[0] = acc;
}

namespace qipgpu {

int norm_to_device(qipgpu_state *s, double *dst) {
    const uint32_t nb = (uint32_t)std::min<uint64_t>(kNormBlocks, std::max<uint64_t>(1, s->len / 256));
    int rc = ensure_scratch(s, sizeof(double) * (nb + 8));
    if (rc) return rc;
    double *partial = static_cast<double *>(s->scratch) + 8;
    if (s->dtype == QIPGPU_C128)
        hipLaunchKernelGGL(k_norm_partial<double>, dim3(nb), dim3(256), 0, s->stream, (const cx<double> *)s->amps,
                           s->len, partial);
    else
        hipLaunchKernelGGL(k_norm_partial<float>, dim3(nb), dim3(256), 0, s->stream, (const cx<float> *)s->amps, s->len,
                           partial);
    QIP_LAUNCHED();
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, s->stream, partial, nb, dst);
    QIP_LAUNCHED();
    return QIPGPU_OK;
}

}  // namespace qipgpu

// Marginals: out[m] (m in [m_lo, m_lo + m_count)) = sum over r of |amps[S(m) | P(r)]|^2 where S
// scatters the k output bits onto state bits opos[] and P inserts zero bits at the measured
// positions (so r enumerates the unmeasured bits in ascending order).
struct MargArgs {
    uint64_t m_lo, m_count, terms, chunk, nchunks;
    int32_t k;
    uint8_t opos[QIPGPU_MAX_QUBITS];
    PosList mpl;  // measured state-bit positions, ascending
};

template <typename R>
__global__ __launch_bounds__(256) void k_marg_serial(const cx<R> *__restrict__ amps, MargArgs a, double *out) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.m_count) return;
    const uint64_t m = a.m_lo + t;
    const uint64_t base = scatter_bits(m, a.opos, a.k);
    double acc = 0.0;
    for (uint64_t r = 0; r < a.terms; ++r) {
        const cx<R> v = amps[base | insert_zeros(r, a.mpl)];
        const double re = v.re, im = v.im;
        acc = fma(re, re, acc);
        acc = fma(im, im, acc);
    }
    out[t] = acc;
}

template <typename R>
__global__ __launch_bounds__(256) void k_marg_chunk(const cx<R> *__restrict__ amps, MargArgs a, double *partial) {
    __shared__ double lds[4];
    const uint64_t c = blockIdx.x % a.nchunks;
    const uint64_t mt = blockIdx.x / a.nchunks;
    const uint64_t base = scatter_bits(a.m_lo + mt, a.opos, a.k);
    const uint64_t r0 = c * a.chunk, r1 = min(a.terms, r0 + a.chunk);
    double acc = 0.0;
    for (uint64_t r = r0 + threadIdx.x; r < r1; r += 256) {
        const cx<R> v = amps[base | insert_zeros(r, a.mpl)];
        const double re = v.re, im = v.im;
        acc = fma(re, re, acc);
        acc = fma(im, im, acc);
    }
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ __launch_bounds__(256) void k_marg_final(const double *__restrict__ partial, uint64_t m_count, uint64_t nchunks,
                                                    double *out) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m_count) return;
    double acc = 0.0;
    for (uint64_t c = 0; c < nchunks; ++c) acc += partial[t * nchunks + c];
    out[t] = acc;
}

// Marginals of every outcome in one pass over the state (n >= 12). The state is read in chunks of
// 4096 amplitudes, the state bits 0..11; the other bits are the k_hi measured bits above 11 and the
// free (unmeasured) bits above 11. Block b takes outcome bits m_hi = b / nparts and the part
// b % nparts of the free-bit combinations, sums |amp|^2 per chunk position over its chunks (thread t
// keeps positions t + 256 j in registers, so every load is contiguous across the block), then folds
// the 4096 position sums into the 2^k_lo outcomes of the measured bits among 0..11, in a fixed order.
// k_marg_pass_final adds the parts of each outcome in order and writes it at its output index.
constexpr int kMargChunkBits = 12;

struct MargPassArgs {
    uint64_t per;                // free-bit combinations per block
    uint64_t free_mask;          // unmeasured state bits above 11
    uint64_t m_total;            // 2^k outcomes
    uint32_t nparts;
    int32_t k, klo, khi, nfree_lo;
    uint8_t hi_meas[QIPGPU_MAX_QUBITS];   // measured state bits above 11, ascending (m_hi bit i)
    uint8_t lo_meas[kMargChunkBits];      // measured state bits among 0..11, ascending (m_lo bit i)
    uint8_t lo_free[kMargChunkBits];      // unmeasured state bits among 0..11, ascending
    uint8_t opos[QIPGPU_MAX_QUBITS];      // state bit of output bit j
};

template <typename R>
__global__ __launch_bounds__(256) void k_marg_pass(const cx<R> *__restrict__ amps, MargPassArgs a, double *partial) {
    __shared__ double sums[1 << kMargChunkBits];
    __shared__ double red[256];
    const uint32_t tid = threadIdx.x;
    const uint64_t m_hi = blockIdx.x / a.nparts, part = blockIdx.x % a.nparts;
    const uint64_t hi_base = scatter_bits(m_hi, a.hi_meas, a.khi);
    // free bits of the first combination of this part; the next combination in ascending order is
    // (x - free_mask) & free_mask
    uint64_t x = 0;
    {
        const uint64_t v0 = part * a.per;
        uint64_t m = a.free_mask;
        for (int b = 0; m; ++b, m &= m - 1)
            if (v0 >> b & 1) x |= m & (~m + 1);
    }
    double acc[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = 0.0;
    for (uint64_t v = 0; v < a.per; ++v) {
        const cx<R> *p = amps + (hi_base | x) + tid;
        cx<R> y[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) y[j] = p[256 * j];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const double re = y[j].re, im = y[j].im;
            acc[j] = fma(re, re, acc[j]);
            acc[j] = fma(im, im, acc[j]);
        }
        x = (x - a.free_mask) & a.free_mask;
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) sums[tid + 256 * j] = acc[j];
    __syncthreads();
    // outcome m_lo sums the positions with its measured bits, the free ones ascending; with fewer
    // than 256 outcomes, g = 256 / 2^klo threads share one (free combinations u = g', g' + g, ...)
    const uint32_t nlo = 1u << a.klo;
    const uint32_t g = nlo >= 256 ? 1u : 256u / nlo;
    for (uint32_t idx = tid; idx < nlo * g; idx += 256) {
        const uint32_t mlo = idx / g, gi = idx % g;
        const uint32_t lb = (uint32_t)scatter_bits(mlo, a.lo_meas, a.klo);
        double s = 0.0;
        for (uint32_t u = gi; u < (uint32_t)a.nfree_lo; u += g) s += sums[lb | (uint32_t)scatter_bits(u, a.lo_free, kMargChunkBits - a.klo)];
        if (g == 1) partial[(uint64_t)blockIdx.x * nlo + mlo] = s;
        else red[idx] = s;
    }
    if (g > 1) {
        __syncthreads();
        if (tid < nlo) {
            double s = 0.0;
            for (uint32_t gi = 0; gi < g; ++gi) s += red[tid * g + gi];
            partial[(uint64_t)blockIdx.x * nlo + tid] = s;
        }
    }
}

__device__ __forceinline__ uint64_t marg_out_index(uint64_t m_hi, uint64_t mlo, const MargPassArgs &a) {
    const uint64_t sb = scatter_bits(m_hi, a.hi_meas, a.khi) | scatter_bits(mlo, a.lo_meas, a.klo);
    uint64_t oi = 0;
    for (int j = 0; j < a.k; ++j) oi |= ((sb >> a.opos[j]) & 1ull) << j;
    return oi;
}

// one thread per outcome, the parts in order (few parts)
__global__ __launch_bounds__(256) void k_marg_pass_final(const double *__restrict__ partial, MargPassArgs a, double *out) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.m_total) return;
    const uint64_t m_hi = t >> a.klo, mlo = t & ((1ull << a.klo) - 1ull);
    double s = 0.0;
    for (uint32_t p = 0; p < a.nparts; ++p) s += partial[((m_hi * a.nparts + p) << a.klo) + mlo];
    out[marg_out_index(m_hi, mlo, a)] = s;
}

// one wavefront per outcome (many parts): lane l adds parts l, l + 64, ... in order, then a butterfly
__global__ __launch_bounds__(64) void k_marg_pass_final_wave(const double *__restrict__ partial, MargPassArgs a,
                                                             double *out) {
    const uint64_t t = blockIdx.x;
    const uint64_t m_hi = t >> a.klo, mlo = t & ((1ull << a.klo) - 1ull);
    double s = 0.0;
    for (uint32_t p = threadIdx.x; p < a.nparts; p += 64) s += partial[((m_hi * a.nparts + p) << a.klo) + mlo];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (threadIdx.x == 0) out[marg_out_index(m_hi, mlo, a)] = s;
}

namespace {

enum class Order { LsbFirst, Entwine };

// Where a reduction keeps its results and partials: the state's growable scratch (area == nullptr),
// or a fixed device area that is never reallocated.
struct FixedArea {
    char *base;
    size_t cap;
};

static int reduction_space(qipgpu_state *s, const FixedArea *area, size_t bytes, char **out) {
    if (!area) {
        int rc = ensure_scratch(s, bytes);
        if (rc) return rc;
        *out = static_cast<char *>(s->scratch);
        return QIPGPU_OK;
    }
    if (bytes > area->cap)
        return set_error(QIPGPU_EINVAL, "the reduction needs %zu bytes of a %zu-byte fixed area", bytes, area->cap);
    *out = area->base;
    return QIPGPU_OK;
}

// Every outcome's marginal in one pass over the state (k_marg_pass), left on the device at the start
// of the scratch (2^k values in the output order given by opos). The scratch also holds `extra` bytes
// from the first 256-byte boundary after the results on, free once this returns.
static int marginals_pass(qipgpu_state *s, const std::vector<int> &mpos, const uint8_t *opos, size_t extra,
                          const FixedArea *area) {
    const int n = s->n, k = (int)mpos.size();
    MargPassArgs a;
    std::memset(&a, 0, sizeof(a));
    a.k = k;
    a.m_total = 1ull << k;
    std::vector<int> sorted(mpos);
    std::sort(sorted.begin(), sorted.end());
    uint64_t meas_mask = 0;
    for (int p : sorted) {
        meas_mask |= 1ull << p;
        if (p < kMargChunkBits) a.lo_meas[a.klo++] = (uint8_t)p;
        else a.hi_meas[a.khi++] = (uint8_t)p;
    }
    for (int b = 0; b < kMargChunkBits; ++b)
        if (!(meas_mask >> b & 1)) a.lo_free[a.nfree_lo++] = (uint8_t)b;
    a.nfree_lo = 1 << a.nfree_lo;   // combinations of the free low bits
    for (int b = kMargChunkBits; b < n; ++b)
        if (!(meas_mask >> b & 1)) a.free_mask |= 1ull << b;
    for (int j = 0; j < k; ++j) a.opos[j] = opos[j];
    // about eight blocks per CU, each summing a power-of-two share of the free combinations. The
    // partials (2^k per part, read back strided) stay within 8 MiB unless that leaves fewer than two
    // blocks per CU, and within 256 MiB.
    const int nfree_hi = n - kMargChunkBits - a.khi;
    const size_t res_bytes = ((sizeof(double) * a.m_total + 255) / 256) * 256;
    const size_t part_cap = area ? (area->cap > res_bytes ? area->cap - res_bytes : 0) : (256ull << 20);
    uint64_t nparts = 1;
    auto blocks_of = [&](uint64_t parts) { return (1ull << a.khi) * parts; };
    while (nparts < (1ull << nfree_hi) && blocks_of(nparts) < 2048 &&
           (sizeof(double) * a.m_total * nparts * 2 <= (8ull << 20) || blocks_of(nparts) < 512) &&
           sizeof(double) * a.m_total * nparts * 2 <= part_cap)
        nparts *= 2;
    a.nparts = (uint32_t)nparts;
    a.per = (1ull << nfree_hi) / nparts;
    const uint64_t nblk = (1ull << a.khi) * nparts;
    if (nblk > 0x7fffffffull) return set_error(QIPGPU_EUNSUP, "marginal grid too large");
    const size_t part_off = res_bytes;
    char *space = nullptr;
    int rc = reduction_space(s, area, part_off + std::max(sizeof(double) * a.m_total * nparts, extra), &space);
    if (rc) return rc;
    double *d_out = reinterpret_cast<double *>(space);
    double *d_part = reinterpret_cast<double *>(space + part_off);
    if (s->dtype == QIPGPU_C128)
        hipLaunchKernelGGL(k_marg_pass<double>, dim3((uint32_t)nblk), dim3(256), 0, s->stream, (const cx<double> *)s->amps,
                           a, d_part);
    else
        hipLaunchKernelGGL(k_marg_pass<float>, dim3((uint32_t)nblk), dim3(256), 0, s->stream, (const cx<float> *)s->amps,
                           a, d_part);
    QIP_LAUNCHED();
    if (nparts >= 64)   // then 2^k_hi <= 32 and m_total <= 2^17 wavefronts
        hipLaunchKernelGGL(k_marg_pass_final_wave, dim3((uint32_t)a.m_total), dim3(64), 0, s->stream, d_part, a, d_out);
    else
        hipLaunchKernelGGL(k_marg_pass_final, dim3(blocks_for(a.m_total, 256)), dim3(256), 0, s->stream, d_part, a, d_out);
    QIP_LAUNCHED();
    return QIPGPU_OK;
}

// Marginals over outcomes [m_lo, m_lo + m_count), left on the device at the start of the scratch;
// the scratch holds `extra` more bytes from the first 256-byte boundary after them.
static int marginals_dev(qipgpu_state *s, int32_t k, const int32_t *indices, Order order, uint64_t m_lo,
                         uint64_t m_count, size_t extra, const FixedArea *area = nullptr) {
    int rc = check_qubits(s, k, indices, "measured qubits");
    if (rc) return rc;
    rc = settle(s);
    if (rc) return rc;
    const int n = s->n;
    MargArgs a;
    std::memset(&a, 0, sizeof(a));
    a.k = k;
    a.m_lo = m_lo;
    a.m_count = m_count;
    a.terms = 1ull << (n - k);
    std::vector<int> mpos;
    for (int j = 0; j < k; ++j) mpos.push_back(n - 1 - indices[j]);
    if (order == Order::LsbFirst) {
        // kronprod.pyx:258-259: qubit indices[j] on output bit j
        for (int j = 0; j < k; ++j) a.opos[j] = (uint8_t)(n - 1 - indices[j]);
    } else {
        // entwine_bit: output bits in the order of the measured state bits
        std::vector<int> sorted(mpos);
        std::sort(sorted.begin(), sorted.end());
        for (int j = 0; j < k; ++j) a.opos[j] = (uint8_t)sorted[j];
    }
    fill_poslist(a.mpl, mpos);
    // every outcome of up to 2^26: one pass over the state (its partials stay within 1 GiB)
    if (n >= kMargChunkBits && k >= 1 && k <= 26 && m_lo == 0 && m_count == (1ull << k))
        return marginals_pass(s, mpos, a.opos, extra, area);
    const bool serial = a.terms <= 64;
    a.chunk = 256 * 64;
    a.nchunks = serial ? 1 : (a.terms + a.chunk - 1) / a.chunk;
    const size_t out_bytes = sizeof(double) * m_count;
    const size_t part_bytes = serial ? 0 : sizeof(double) * m_count * a.nchunks;
    const size_t part_off = ((out_bytes + 255) / 256) * 256;
    char *space = nullptr;
    rc = reduction_space(s, area, part_off + std::max(part_bytes, extra), &space);
    if (rc) return rc;
    double *d_out = reinterpret_cast<double *>(space);
    double *d_part = reinterpret_cast<double *>(space + part_off);
    if (serial) {
        const uint32_t nb = blocks_for(m_count, 256);
        if (s->dtype == QIPGPU_C128)
            hipLaunchKernelGGL(k_marg_serial<double>, dim3(nb), dim3(256), 0, s->stream, (const cx<double> *)s->amps, a, d_out);
        else
            hipLaunchKernelGGL(k_marg_serial<float>, dim3(nb), dim3(256), 0, s->stream, (const cx<float> *)s->amps, a, d_out);
        QIP_LAUNCHED();
    } else {
        const uint64_t nblk = m_count * a.nchunks;
        if (nblk > 0x7fffffffull) return set_error(QIPGPU_EUNSUP, "marginal grid too large");
        if (s->dtype == QIPGPU_C128)
            hipLaunchKernelGGL(k_marg_chunk<double>, dim3((uint32_t)nblk), dim3(256), 0, s->stream, (const cx<double> *)s->amps,
                               a, d_part);
        else
            hipLaunchKernelGGL(k_marg_chunk<float>, dim3((uint32_t)nblk), dim3(256), 0, s->stream, (const cx<float> *)s->amps, a,
                               d_part);
        QIP_LAUNCHED();
        hipLaunchKernelGGL(k_marg_final, dim3(blocks_for(m_count, 256)), dim3(256), 0, s->stream, d_part, m_count, a.nchunks,
                           d_out);
        QIP_LAUNCHED();
    }
    return QIPGPU_OK;
}

// The same into host `out`.
static int marginals_host(qipgpu_state *s, int32_t k, const int32_t *indices, Order order, uint64_t m_lo,
                          uint64_t m_count, double *out) {
    int rc = marginals_dev(s, k, indices, order, m_lo, m_count, 0);
    if (rc) return rc;
    QIP_HIP(hipMemcpyAsync(out, s->scratch, sizeof(double) * m_count, hipMemcpyDeviceToHost, s->stream));
    QIP_HIP(hipStreamSynchronize(s->stream));
    return QIPGPU_OK;
}

}  // namespace

// reduce_measure (kronprod.pyx:482-489): out[i] = in[measured bits m, unmeasured bits = low
// n-k bits of i] * mult, out of place.
// In place, in two steps: k_reduce_gather copies the kept slice, scaled (slice[j] = amps[measured
// bits | j with zeros inserted at the measured positions] * mult), into a compact buffer of 2^(n-k)
// amplitudes; k_reduce_tile then writes amps[i] = slice[i mod 2^(n-k)] over the whole state in address
// order. Peak memory is the state plus the slice, not two states.
template <typename R>
__global__ __launch_bounds__(256) void k_reduce_gather(const cx<R> *__restrict__ amps, cx<R> *slice, uint64_t count,
                                                       uint64_t mbits, PosList mpl, double mult) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < count; j += stride) {
        const cx<R> v = amps[mbits | insert_zeros(j, mpl)];
        cx<R> o;
        o.re = (R)((double)v.re * mult);
        o.im = (R)((double)v.im * mult);
        slice[j] = o;
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_reduce_tile(cx<R> *amps, const cx<R> *__restrict__ slice, uint64_t len,
                                                     uint64_t lowmask) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) amps[i] = slice[i & lowmask];
}

// ---- top-k on the device --------------------------------------------------------------------------
// Order: larger probability first, equal probabilities by ascending outcome (a strict total order on
// distinct outcomes). A round sorts chunks of kTopChunk candidates in LDS (bitonic network) and keeps
// the first T of each; rounds repeat until one chunk is left. Slots past the end hold (-1, ~0), which
// sort after every probability.
constexpr int kTopChunk = 2048;
constexpr int kTopMaxDevice = 1024;

__device__ __forceinline__ bool top_before(double va, uint64_t ia, double vb, uint64_t ib) {
    return va > vb || (va == vb && ia < ib);
}

__global__ __launch_bounds__(256) void k_topk_round(const double *__restrict__ vin, const uint64_t *__restrict__ iin,
                                                    uint64_t len, uint32_t T, double *__restrict__ vout,
                                                    uint64_t *__restrict__ iout) {
    __shared__ double sv[kTopChunk];
    __shared__ uint64_t si[kTopChunk];
    const uint64_t c0 = (uint64_t)blockIdx.x * kTopChunk;
    for (int e = threadIdx.x; e < kTopChunk; e += 256) {
        const uint64_t g = c0 + e;
        if (g < len) {
            sv[e] = vin[g];
            si[e] = iin ? iin[g] : g;
        } else {
            sv[e] = -1.0;
            si[e] = ~0ull;
        }
    }
    __syncthreads();
    for (int size = 2; size <= kTopChunk; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = threadIdx.x; t < kTopChunk / 2; t += 256) {
                const int lo = 2 * t - (t & (stride - 1)), hi = lo + stride;
                const bool asc = (lo & size) == 0;
                const bool out_of_order = asc ? top_before(sv[hi], si[hi], sv[lo], si[lo])
                                              : top_before(sv[lo], si[lo], sv[hi], si[hi]);
                if (out_of_order) {
                    const double tv = sv[lo];
                    sv[lo] = sv[hi];
                    sv[hi] = tv;
                    const uint64_t ti = si[lo];
                    si[lo] = si[hi];
                    si[hi] = ti;
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t e = threadIdx.x; e < T; e += 256) {
        vout[(uint64_t)blockIdx.x * T + e] = sv[e];
        iout[(uint64_t)blockIdx.x * T + e] = si[e];
    }
}

namespace {

static uint64_t measured_bits(int n, int32_t k, const int32_t *indices, uint64_t m, uint64_t *mask_out) {
    std::vector<int> pos;
    uint64_t mask = 0;
    for (int j = 0; j < k; ++j) {
        pos.push_back(n - 1 - indices[j]);
        mask |= 1ull << (n - 1 - indices[j]);
    }
    std::sort(pos.begin(), pos.end());
    uint64_t bits = 0;
    for (int j = 0; j < k; ++j) bits |= ((m >> j) & 1ull) << pos[j];
    *mask_out = mask;
    return bits;
}

static int soft_measure_impl(qipgpu_state *s, int32_t k, const int32_t *indices, double u, int64_t measured, int64_t *m_out,
                             double *p_out) {
    int rc = check_qubits(s, k, indices, "measured qubits");
    if (rc) return rc;
    if (k < 1) return set_error(QIPGPU_EINVAL, "measure needs at least one qubit");
    const uint64_t nm = 1ull << k;
    if (measured >= (int64_t)nm) return set_error(QIPGPU_EINVAL, "Measured value must be less than 2**len(indices)");
    if (measured >= 0) {
        // the draw is made by the caller and unused (kronprod.pyx:362); only the outcome's probability
        // is needed, which reads the 2^(n-k) amplitudes of that outcome
        double acc = 0.0;
        rc = marginals_host(s, k, indices, Order::Entwine, (uint64_t)measured, 1, &acc);
        if (rc) return rc;
        *m_out = measured;
        *p_out = acc;
        return QIPGPU_OK;
    }
    // one pass over the state for every outcome's probability; prob_magnitude (kronprod.pyx:362-365)
    // is their sum, in ascending outcome order, in FP64
    std::vector<double> marg(nm);
    rc = marginals_host(s, k, indices, Order::Entwine, 0, nm, marg.data());
    if (rc) return rc;
    double total = 0.0;
    for (uint64_t i = 0; i < nm; ++i) total += marg[i];
    double r = u * total;
    // kronprod.pyx:371-381: sequential subtraction, first outcome with r <= 0; if none, the last.
    uint64_t p_index = 0;
    double acc = 0.0;
    for (p_index = 0; p_index < nm; ++p_index) {
        acc = marg[p_index];
        r -= acc;
        if (r <= 0) break;
    }
    if (p_index == nm) p_index = nm - 1;
    *m_out = (int64_t)p_index;
    *p_out = acc;
    return QIPGPU_OK;
}

static int resolve_outcome(qipgpu_state *s, int32_t k, const int32_t *indices, double u, int64_t measured,
                           double measured_prob, int64_t *m, double *p) {
    if (k < 1) return set_error(QIPGPU_EINVAL, "measure needs at least one qubit");
    if (measured >= 0 && (uint64_t)measured >= (1ull << k))
        return set_error(QIPGPU_EINVAL, "Measured value must be less than 2**len(indices)");
    if (measured_prob > 1.0) return set_error(QIPGPU_EINVAL, "measured_prob must be 0 < p <= 1");
    if (measured < 0 || !(measured_prob > 0.0)) {
        int rc = soft_measure_impl(s, k, indices, u, measured, m, p);
        if (rc) return rc;
    } else {
        *m = measured;
        *p = measured_prob;
    }
    if (*p == 0.0) return set_error(QIPGPU_EZERODIV, "float division by zero");
    return QIPGPU_OK;
}

}  // namespace

extern "C" {

int qipgpu_total_prob(qipgpu_state *s, double *out) {
    int rc = settle(s);
    if (rc) return rc;
    // size the scratch first: norm_to_device keeps its partials behind the result slot
    rc = ensure_scratch(s, sizeof(double) * (kNormBlocks + 8));
    if (rc) return rc;
    rc = norm_to_device(s, static_cast<double *>(s->scratch));
    if (rc) return rc;
    QIP_HIP(hipMemcpyAsync(out, s->scratch, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    QIP_HIP(hipStreamSynchronize(s->stream));
    return QIPGPU_OK;
}

int qipgpu_probabilities(qipgpu_state *s, int32_t k, const int32_t *indices, double *out) {
    if (k < 0 || k > s->n) return set_error(QIPGPU_EINVAL, "bad qubit count %d", k);
    return marginals_host(s, k, indices, Order::LsbFirst, 0, 1ull << k, out);
}

int qipgpu_marginals(qipgpu_state *s, int32_t k, const int32_t *indices, double *out) {
    if (k < 0 || k > s->n) return set_error(QIPGPU_EINVAL, "bad qubit count %d", k);
    return marginals_host(s, k, indices, Order::Entwine, 0, 1ull << k, out);
}

int qipgpu_top_probabilities(qipgpu_state *s, int32_t k, const int32_t *indices, int32_t top_k, int64_t *ms, double *ps,
                             int32_t *count) {
    if (k < 0 || k > s->n) return set_error(QIPGPU_EINVAL, "bad qubit count %d", k);
    if (top_k < 0) return set_error(QIPGPU_EINVAL, "negative top_k");
    const uint64_t nm = 1ull << k;
    const uint64_t kk = std::min<uint64_t>((uint64_t)top_k, nm);
    *count = (int32_t)kk;
    if (kk == 0) return check_qubits(s, k, indices, "measured qubits");
    if (kk > (uint64_t)kTopMaxDevice) {
        // wide selections: every marginal to the host, selected there
        std::vector<double> marg(nm);
        int rc = marginals_host(s, k, indices, Order::Entwine, 0, nm, marg.data());
        if (rc) return rc;
        std::vector<uint64_t> order(nm);
        for (uint64_t i = 0; i < nm; ++i) order[i] = i;
        std::partial_sort(order.begin(), order.begin() + kk, order.end(), [&](uint64_t a, uint64_t b) {
            return marg[a] > marg[b] || (marg[a] == marg[b] && a < b);
        });
        for (uint64_t i = 0; i < kk; ++i) {
            ms[i] = (int64_t)order[i];
            ps[i] = marg[order[i]];
        }
        return QIPGPU_OK;
    }
    // candidates of the first round: ceil(2^k / chunk) * T pairs, ping-pong with the next round's
    const uint32_t T = (uint32_t)kk;
    const uint64_t blocks0 = (nm + kTopChunk - 1) / kTopChunk;
    const size_t pairs = (size_t)blocks0 * T;
    const size_t region = pairs * (sizeof(double) + sizeof(uint64_t));
    const size_t out_off = ((sizeof(double) * nm + 255) / 256) * 256;
    int rc = marginals_dev(s, k, indices, Order::Entwine, 0, nm, 2 * region + 256);
    if (rc) return rc;
    double *marg = static_cast<double *>(s->scratch);
    char *base = static_cast<char *>(s->scratch) + out_off;
    double *va = reinterpret_cast<double *>(base);
    uint64_t *ia = reinterpret_cast<uint64_t *>(base + pairs * sizeof(double));
    double *vb = reinterpret_cast<double *>(base + region);
    uint64_t *ib = reinterpret_cast<uint64_t *>(base + region + pairs * sizeof(double));
    const double *vin = marg;
    const uint64_t *iin = nullptr;
    uint64_t len = nm;
    bool into_a = true;
    for (;;) {
        const uint64_t blocks = (len + kTopChunk - 1) / kTopChunk;
        double *vo = into_a ? va : vb;
        uint64_t *io = into_a ? ia : ib;
        hipLaunchKernelGGL(k_topk_round, dim3((uint32_t)blocks), dim3(256), 0, s->stream, vin, iin, len, T, vo, io);
        QIP_LAUNCHED();
        vin = vo;
        iin = io;
        into_a = !into_a;
        if (blocks == 1) break;
        len = blocks * T;
    }
    std::vector<uint64_t> idx(T);
    QIP_HIP(hipMemcpyAsync(ps, vin, sizeof(double) * T, hipMemcpyDeviceToHost, s->stream));
    QIP_HIP(hipMemcpyAsync(idx.data(), iin, sizeof(uint64_t) * T, hipMemcpyDeviceToHost, s->stream));
    QIP_HIP(hipStreamSynchronize(s->stream));
    for (uint32_t i = 0; i < T; ++i) ms[i] = (int64_t)idx[i];
    return QIPGPU_OK;
}

int qipgpu_soft_measure(qipgpu_state *s, int32_t k, const int32_t *indices, double u, int64_t measured, int64_t *m,
                        double *p) {
    return soft_measure_impl(s, k, indices, u, measured, m, p);
}

int qipgpu_measure(qipgpu_state *s, int32_t k, const int32_t *indices, double u, int64_t measured, double measured_prob,
                   int64_t *m, double *p) {
    int rc = check_qubits(s, k, indices, "measured qubits");
    if (rc) return rc;
    rc = resolve_outcome(s, k, indices, u, measured, measured_prob, m, p);
    if (rc) return rc;
    rc = settle(s);
    if (rc) return rc;
    const double mult = std::sqrt(1.0 / *p);
    uint64_t mask = 0;
    const uint64_t target = measured_bits(s->n, k, indices, (uint64_t)*m, &mask);
    const uint32_t nb = std::min<uint32_t>(blocks_for(s->len, 256 * 4), 1u << 20);
    if (s->dtype == QIPGPU_C128)
        hipLaunchKernelGGL(k_collapse<double>, dim3(nb), dim3(256), 0, s->stream, (cx<double> *)s->amps, s->len, mask, target, mult);
    else
        hipLaunchKernelGGL(k_collapse<float>, dim3(nb), dim3(256), 0, s->stream, (cx<float> *)s->amps, s->len, mask, target, mult);
    QIP_LAUNCHED();
    QIP_HIP(hipStreamSynchronize(s->stream));
    return QIPGPU_OK;
}

// reduce_measure's kept slice: outcome, then slice[j] = amps[measured bits | j with zeros inserted at
// the measured positions] * sqrt(1/p) in a new buffer of 2^(n-k) amplitudes (handed to the caller).
static int reduced_slice(qipgpu_state *s, int32_t k, const int32_t *indices, double u, int64_t measured,
                         double measured_prob, int64_t *m, double *p, TempBuffer &slice) {
    int rc = check_qubits(s, k, indices, "measured qubits");
    if (rc) return rc;
    rc = resolve_outcome(s, k, indices, u, measured, measured_prob, m, p);
    if (rc) return rc;
    rc = settle(s);
    if (rc) return rc;
    const double mult = std::sqrt(1.0 / *p);
    uint64_t mask = 0;
    const uint64_t mbits = measured_bits(s->n, k, indices, (uint64_t)*m, &mask);
    std::vector<int> mpos;
    for (int j = 0; j < k; ++j) mpos.push_back(s->n - 1 - indices[j]);
    PosList mpl;
    fill_poslist(mpl, mpos);
    const uint64_t count = 1ull << (s->n - k);
    if (slice.alloc(count * s->amp_bytes) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(QIPGPU_ENOMEM, "reduce_measure: cannot allocate the %llu-amplitude slice",
                         (unsigned long long)count);
    }
    const uint32_t ng = std::min<uint32_t>(blocks_for(count, 256 * 4), 1u << 20);
    if (s->dtype == QIPGPU_C128)
        hipLaunchKernelGGL(k_reduce_gather<double>, dim3(ng), dim3(256), 0, s->stream, (const cx<double> *)s->amps,
                           (cx<double> *)slice.p, count, mbits, mpl, mult);
    else
        hipLaunchKernelGGL(k_reduce_gather<float>, dim3(ng), dim3(256), 0, s->stream, (const cx<float> *)s->amps,
                           (cx<float> *)slice.p, count, mbits, mpl, mult);
    QIP_LAUNCHED();
    return QIPGPU_OK;
}

int qipgpu_reduce_measure(qipgpu_state *s, int32_t k, const int32_t *indices, double u, int64_t measured,
                          double measured_prob, int64_t *m, double *p) {
    TempBuffer slice(s->stream);
    int rc = reduced_slice(s, k, indices, u, measured, measured_prob, m, p, slice);
    if (rc) return rc;
    // the reference's arena keeps its 2^n size: the slice repeats over it (backend.py:139-142)
    const uint64_t count = 1ull << (s->n - k);
    const uint32_t nt = std::min<uint32_t>(blocks_for(s->len, 256 * 4), 1u << 20);
    if (s->dtype == QIPGPU_C128)
        hipLaunchKernelGGL(k_reduce_tile<double>, dim3(nt), dim3(256), 0, s->stream, (cx<double> *)s->amps,
                           (const cx<double> *)slice.p, s->len, count - 1);
    else
        hipLaunchKernelGGL(k_reduce_tile<float>, dim3(nt), dim3(256), 0, s->stream, (cx<float> *)s->amps,
                           (const cx<float> *)slice.p, s->len, count - 1);
    QIP_LAUNCHED();
    QIP_HIP(hipStreamSynchronize(s->stream));
    return QIPGPU_OK;
}

int qipgpu_reduce_measure_shrink(qipgpu_state *s, int32_t k, const int32_t *indices, double u, int64_t measured,
                                 double measured_prob, int64_t *m, double *p) {
    if (k >= s->n) return set_error(QIPGPU_EINVAL, "shrinking by %d of %d qubits leaves none", k, s->n);
    TempBuffer slice(s->stream);
    int rc = reduced_slice(s, k, indices, u, measured, measured_prob, m, p, slice);
    if (rc) return rc;
    QIP_HIP(hipStreamSynchronize(s->stream));
    QIP_HIP(hipFree(s->amps));
    s->amps = slice.release();
    s->n -= k;
    s->len = 1ull << s->n;
    recorder_discard(s);      // recorded nothing (settled); the record now counts n - k qubits
    return QIPGPU_OK;
}

}  // extern "C"

// ---- measure_probabilities without allocation ------------------------------------------------------
// The reduction's results and partials live in `scratch`, a file-scope device array of
// kNoAllocScratchBytes that every GPU the library runs on holds until the process ends. Host code
// reaches it only through the address hipGetSymbolAddress gives on each device (the array's name is
// a host-side stand-in), fetched on the first call with that device current.
constexpr size_t kNoAllocScratchBytes = 16ull << 20;
constexpr int kNoAllocMaxK = 20;   // 2^20 results (8 MiB) plus at least one part of partials

static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

namespace {
std::mutex g_noalloc_mu;
double *g_noalloc_area[64] = {};

int noalloc_area(int device, double **out) {
    if (device < 0 || device >= 64) return set_error(QIPGPU_EINVAL, "device %d", device);
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_area[device]) {
        int cur = -1;
        QIP_HIP(hipGetDevice(&cur));
        if (cur != device)
            return set_error(QIPGPU_EINVAL, "call with the state's device (%d) current, not %d", device, cur);
        void *p = nullptr;
        QIP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_noalloc_area[device] = static_cast<double *>(p);
    }
    *out = g_noalloc_area[device];
    return QIPGPU_OK;
}
}  // namespace

extern "C" {

int qipgpu_probabilities_noalloc(qipgpu_state *s, int32_t k, const int32_t *indices, double *out) {
    if (!s) return set_error(QIPGPU_EINVAL, "null state");
    if (k < 0 || k > s->n) return set_error(QIPGPU_EINVAL, "bad qubit count %d", k);
    if (k > kNoAllocMaxK)
        return set_error(QIPGPU_EINVAL, "%d measured qubits exceed the no-allocation limit of %d", k, kNoAllocMaxK);
    int rc = check_qubits(s, k, indices, "measured qubits");
    if (rc) return rc;
    // launching recorded gates may allocate the fused path's buffers: the caller flushes them first
    if (recorder_pending(s) > 0)
        return set_error(QIPGPU_EINVAL, "%lld recorded gates are not launched yet (qipgpu_fused_flush first)",
                         (long long)recorder_pending(s));
    double *area = nullptr;
    rc = noalloc_area(s->device, &area);
    if (rc) return rc;
    rc = settle(s);                 // no recorded gates: at most the pending scale, no allocation
    if (rc) return rc;
    // every result and partial is written before it is read, so nothing needs clearing between calls
    const FixedArea fixed = {reinterpret_cast<char *>(area), kNoAllocScratchBytes};
    const uint64_t nm = 1ull << k;
    rc = marginals_dev(s, k, indices, Order::LsbFirst, 0, nm, 0, &fixed);
    if (rc) return rc;
    QIP_HIP(hipMemcpyAsync(out, area, sizeof(double) * nm, hipMemcpyDeviceToHost, s->stream));
    QIP_HIP(hipStreamSynchronize(s->stream));
    return QIPGPU_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------
// comparison helpers: <a|b> and max |a - b| in FP64
// ------------------------------------------------------------------------------------------------

template <typename Ra, typename Rb>
__global__ __launch_bounds__(256) void k_overlap_partial(const cx<Ra> *__restrict__ a, const cx<Rb> *__restrict__ b, uint64_t len,
                                                         double *partial) {
    __shared__ double lds[4];
    const uint64_t per = (len + gridDim.x - 1) / gridDim.x;
    const uint64_t lo = (uint64_t)blockIdx.x * per, hi = min(len, lo + per);
    double re = 0, im = 0, na = 0, nb = 0;
    for (uint64_t i = lo + threadIdx.x; i < hi; i += 256) {
        const cx<Ra> x = a[i];
        const cx<Rb> y = b[i];
        const double xr = x.re, xi = x.im, yr = y.re, yi = y.im;
        re += xr * yr + xi * yi;
        im += xr * yi - xi * yr;
        na += xr * xr + xi * xi;
        nb += yr * yr + yi * yi;
    }
    re = block_sum(re, lds);
    im = block_sum(im, lds);
    na = block_sum(na, lds);
    nb = block_sum(nb, lds);
    if (threadIdx.x == 0) {
        partial[4 * blockIdx.x + 0] = re;
        partial[4 * blockIdx.x + 1] = im;
        partial[4 * blockIdx.x + 2] = na;
        partial[4 * blockIdx.x + 3] = nb;
    }
}

template <typename Ra, typename Rb>
__global__ __launch_bounds__(256) void k_maxdiff_partial(const cx<Ra> *__restrict__ a, const cx<Rb> *__restrict__ b, uint64_t len,
                                                         double *partial) {
    __shared__ double lds[256];
    const uint64_t per = (len + gridDim.x - 1) / gridDim.x;
    const uint64_t lo = (uint64_t)blockIdx.x * per, hi = min(len, lo + per);
    double m = 0;
    for (uint64_t i = lo + threadIdx.x; i < hi; i += 256) {
        const cx<Ra> x = a[i];
        const cx<Rb> y = b[i];
        const double dr = (double)x.re - (double)y.re, di = (double)x.im - (double)y.im;
        m = fmax(m, sqrt(dr * dr + di * di));
    }
    lds[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) lds[threadIdx.x] = fmax(lds[threadIdx.x], lds[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = lds[0];
}

namespace {

template <typename Ra, typename Rb>
static void launch_overlap(qipgpu_state *a, qipgpu_state *b, uint32_t nb, double *partial) {
    hipLaunchKernelGGL((k_overlap_partial<Ra, Rb>), dim3(nb), dim3(256), 0, a->stream, (const cx<Ra> *)a->amps,
                       (const cx<Rb> *)b->amps, a->len, partial);
}

template <typename Ra, typename Rb>
static void launch_maxdiff(qipgpu_state *a, qipgpu_state *b, uint32_t nb, double *partial) {
    hipLaunchKernelGGL((k_maxdiff_partial<Ra, Rb>), dim3(nb), dim3(256), 0, a->stream, (const cx<Ra> *)a->amps,
                       (const cx<Rb> *)b->amps, a->len, partial);
}

}  // namespace

extern "C" {

int qipgpu_overlap(qipgpu_state *a, qipgpu_state *b, double *re, double *im, double *norm_a, double *norm_b) {
    if (a->n != b->n) return set_error(QIPGPU_EINVAL, "overlap of states with %d and %d qubits", a->n, b->n);
    int rc = settle(a);
    if (rc) return rc;
    rc = settle(b);
    if (rc) return rc;
    QIP_HIP(hipStreamSynchronize(b->stream));
    const uint32_t nb = (uint32_t)std::min<uint64_t>(kNormBlocks, std::max<uint64_t>(1, a->len / 256));
    rc = ensure_scratch(a, sizeof(double) * 4 * nb);
    if (rc) return rc;
    double *partial = static_cast<double *>(a->scratch);
    const bool a128 = a->dtype == QIPGPU_C128, b128 = b->dtype == QIPGPU_C128;
    if (a128 && b128) launch_overlap<double, double>(a, b, nb, partial);
    else if (a128) launch_overlap<double, float>(a, b, nb, partial);
    else if (b128) launch_overlap<float, double>(a, b, nb, partial);
    else launch_overlap<float, float>(a, b, nb, partial);
    QIP_LAUNCHED();
    std::vector<double> h(4 * nb);
    QIP_HIP(hipMemcpyAsync(h.data(), partial, sizeof(double) * 4 * nb, hipMemcpyDeviceToHost, a->stream));
    QIP_HIP(hipStreamSynchronize(a->stream));
    double r[4] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < nb; ++i)
        for (int c = 0; c < 4; ++c) r[c] += h[4 * i + c];
    *re = r[0];
    *im = r[1];
    *norm_a = r[2];
    *norm_b = r[3];
    return QIPGPU_OK;
}

int qipgpu_max_abs_diff(qipgpu_state *a, qipgpu_state *b, double *out) {
    if (a->n != b->n) return set_error(QIPGPU_EINVAL, "comparison of states with %d and %d qubits", a->n, b->n);
    int rc = settle(a);
    if (rc) return rc;
    rc = settle(b);
    if (rc) return rc;
    QIP_HIP(hipStreamSynchronize(b->stream));
    const uint32_t nb = (uint32_t)std::min<uint64_t>(kNormBlocks, std::max<uint64_t>(1, a->len / 256));
    rc = ensure_scratch(a, sizeof(double) * nb);
    if (rc) return rc;
    double *partial = static_cast<double *>(a->scratch);
    const bool a128 = a->dtype == QIPGPU_C128, b128 = b->dtype == QIPGPU_C128;
    if (a128 && b128) launch_maxdiff<double, double>(a, b, nb, partial);
    else if (a128) launch_maxdiff<double, float>(a, b, nb, partial);
    else if (b128) launch_maxdiff<float, double>(a, b, nb, partial);
    else launch_maxdiff<float, float>(a, b, nb, partial);
    QIP_LAUNCHED();
    std::vector<double> h(nb);
    QIP_HIP(hipMemcpyAsync(h.data(), partial, sizeof(double) * nb, hipMemcpyDeviceToHost, a->stream));
    QIP_HIP(hipStreamSynchronize(a->stream));
    double m = 0;
    for (uint32_t i = 0; i < nb; ++i) m = std::max(m, h[i]);
    *out = m;
    return QIPGPU_OK;
}

}  // extern "C"
