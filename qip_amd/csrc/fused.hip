// fused.hip -- precision form: a 64x64 complex operator (six fused qubits) applied to a complex64
// state on MFMA with split-BF16 compensated products, FP32 accumulation and FP64 norms. The tile,
// its data flow and the operator fragment layout are in tile6.h; this file holds the persistent
// single-state kernel, the FP64 renormalisation and the C entry points.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.h"
#include "tile6.h"

using qipgpu::cx;
using namespace qipgpu;
using namespace qipgpu::tile;

namespace {

constexpr int kFusedBlocksPerCU = 2;   // 69 KB of LDS and 212 VGPRs per workgroup

__global__ __launch_bounds__(256, 2) void k_fused6(cx<float> *amps, const bf16x8 *__restrict__ wfrag, FusedArgs a,
                                                   const double *__restrict__ fstate, double *partial) {
    typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f4;
    __shared__ TileLds lds;
    __shared__ double red[4];
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int row0 = wave * 16;

    bf16x8 ah[8], al[8];
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
        ah[ks] = wfrag[((wave * 8 + ks) * 2 + 0) * 64 + lane];
        al[ks] = wfrag[((wave * 8 + ks) * 2 + 1) * 64 + lane];
    }
    const double scale = fstate[0];
    const uint64_t off_t = u_offset<0, 9>(2u * (uint32_t)tid, a);     // this thread's u-pair offset
    const uint32_t ucol_gather = u_of_col(lane, a);
    const uint32_t ucol_out0 = u_of_col(lane & 31, a), ucol_out1 = u_of_col(32 + (lane & 31), a);
    const uint32_t urow_half = (lane >> 5) ? u_of_row(4, a) : 0u;    // row bit 2 from the lane half

    double nacc = 0.0;
    TileCtx ctx;
    ctx.tid = tid;
    ctx.lane = lane;
    ctx.row0 = row0;
    ctx.ucol_gather = ucol_gather;
    ctx.ucol_out0 = ucol_out0;
    ctx.ucol_out1 = ucol_out1;
    ctx.urow_half = urow_half;
    ctx.scale = scale;

    // Each workgroup takes a contiguous range of tiles (neighbouring tiles share DRAM pages more
    // often than tiles a grid apart). Two register buffers alternate: while one tile is processed
    // the next one's loads are in flight in the other buffer, with no copy of in-flight registers.
    // The prefetch is issued unconditionally (the last one re-reads an in-range tile) so the wait
    // before each tile covers only its own loads.
    const uint64_t per = (a.ntiles + gridDim.x - 1) / gridDim.x;
    uint64_t tile = (uint64_t)blockIdx.x * per;
    const uint64_t end = min(tile + per, a.ntiles);
    if (tile < end) {
        const uint64_t last = end - 1;
        cx<float> va[16], vb[16];
        load_tile(va, amps, tile_base(tile, a) + off_t, a);
        for (;;) {
            load_tile(vb, amps, tile_base(min(tile + 1, last), a) + off_t, a);
            process_tile(va, tile_base(tile, a) + off_t, amps, lds, ah, al, ctx, a, nacc);
            if (++tile >= end) break;
            load_tile(va, amps, tile_base(min(tile + 1, last), a) + off_t, a);
            process_tile(vb, tile_base(tile, a) + off_t, amps, lds, ah, al, ctx, a, nacc);
            if (++tile >= end) break;
        }
    }
    // block reduction of the FP64 norm partial, fixed order
    for (int off = 32; off > 0; off >>= 1) nacc += __shfl_down(nacc, off, 64);
    if (lane == 0) red[wave] = nacc;
    __syncthreads();
    if (tid == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Sum the partials in a fixed order; fstate[1] = norm, fstate[0] = 1 / sqrt(norm).
__global__ __launch_bounds__(256) void k_fused_norm(const double *__restrict__ partial, uint32_t count, double *fstate) {
    __shared__ double red[256];
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < count; i += 256) acc += partial[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        fstate[1] = red[0];
        fstate[0] = 1.0 / sqrt(red[0]);
    }
}

__global__ void k_set_scale(double *fstate, double v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) fstate[0] = v;
}

template <typename R>
__global__ __launch_bounds__(256) void k_scale(cx<R> *amps, uint64_t len, const double *__restrict__ fstate) {
    const double s = fstate[0];
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
        cx<R> v = amps[i];
        v.re = (R)((double)v.re * s);
        v.im = (R)((double)v.im * s);
        amps[i] = v;
    }
}

int ensure_fused_state(qipgpu_state *s, uint32_t nblocks) {
    if (s->d_fused && s->fused_partials >= nblocks) return QIPGPU_OK;
    if (s->d_fused) {
        QIP_HIP(hipStreamSynchronize(s->stream));
        QIP_HIP(hipFree(s->d_fused));
        s->d_fused = nullptr;
    }
    QIP_HIP(hipMalloc(&s->d_fused, sizeof(double) * (4 + nblocks)));
    s->fused_partials = nblocks;
    hipLaunchKernelGGL(k_set_scale, dim3(1), dim3(64), 0, s->stream, s->d_fused, 1.0);
    QIP_LAUNCHED();
    return QIPGPU_OK;
}

uint32_t fused_grid(uint64_t ntiles) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) == hipSuccess && p.multiProcessorCount > 0) cus = p.multiProcessorCount;
    }
    (void)hipGetLastError();
    return (uint32_t)std::min<uint64_t>(ntiles, (uint64_t)cus * kFusedBlocksPerCU);
}

}  // namespace

namespace qipgpu {

// Apply a pending renormalisation factor before any non-fused use of the state.
int settle(qipgpu_state *s) {
    if (!s->fused_pending) return QIPGPU_OK;
    return qipgpu_fused_finish(s);
}

}  // namespace qipgpu

extern "C" {

int qipgpu_fused_apply(qipgpu_state *s, int32_t k, const int32_t *targets, const double *mat) {
    if (s->dtype != QIPGPU_C64) return set_error(QIPGPU_EINVAL, "fused blocks apply to complex64 states");
    if (s->n < kTileBits) return set_error(QIPGPU_EUNSUP, "fused blocks need n >= %d (n=%d)", kTileBits, s->n);
    if (k < 1 || k > 6) return set_error(QIPGPU_EINVAL, "fused block on %d qubits (1..6)", k);
    int rc = check_qubits(s, k, targets, "fused block qubits");
    if (rc) return rc;
    const int n = s->n;
    // tile bits: the targets padded to six, and six column bits (choose_tile_bits)
    std::vector<int> tbits;  // state bits in the caller's matrix order (MSB first), then padding
    uint64_t tmask = 0;
    for (int j = 0; j < k; ++j) {
        tbits.push_back(n - 1 - targets[j]);
        tmask |= 1ull << (n - 1 - targets[j]);
    }
    int ts[6], ls[6];
    choose_tile_bits(n, tmask, ts, ls);
    for (int b = 0; b < 6; ++b)
        if (!(tmask >> ts[b] & 1)) tbits.push_back(ts[b]);
    // row index r: bit b on the b-th lowest T bit. Map r -> index of the padded caller matrix
    // (caller bits MSB-first over tbits[0..k-1], padding bits are identity).
    std::vector<int> tsorted(ts, ts + 6);
    int rowbit_of_tbit[64];
    for (int b = 0; b < 6; ++b) rowbit_of_tbit[tsorted[b]] = b;
    auto caller_index = [&](int r, int *pad) {
        int ci = 0, pv = 0;
        for (int j = 0; j < k; ++j) ci = (ci << 1) | ((r >> rowbit_of_tbit[tbits[j]]) & 1);
        for (int j = k; j < 6; ++j) pv = (pv << 1) | ((r >> rowbit_of_tbit[tbits[j]]) & 1);
        *pad = pv;
        return ci;
    };
    const int D = 1 << k;
    std::vector<double> ur(64 * 64), ui(64 * 64);
    for (int r = 0; r < 64; ++r) {
        int pr;
        const int cr = caller_index(r, &pr);
        for (int c = 0; c < 64; ++c) {
            int pc;
            const int cc = caller_index(c, &pc);
            const bool same = pr == pc;
            ur[r * 64 + c] = same ? mat[2 * (cr * D + cc)] : 0.0;
            ui[r * 64 + c] = same ? mat[2 * (cr * D + cc) + 1] : 0.0;
        }
    }
    std::vector<uint16_t> frag(kFragBytes / 2);
    split_fragments(ur.data(), ui.data(), frag.data());
    void *dfrag = nullptr;
    rc = ring_upload(s, frag.data(), kFragBytes, &dfrag);
    if (rc) return rc;

    const FusedArgs a = make_tile_args(n, ts, ls);

    const uint32_t grid = fused_grid(a.ntiles);
    rc = ensure_fused_state(s, grid);
    if (rc) return rc;
    if (!s->fused_pending) {
        hipLaunchKernelGGL(k_set_scale, dim3(1), dim3(64), 0, s->stream, s->d_fused, 1.0);
        QIP_LAUNCHED();
    }
    hipLaunchKernelGGL(k_fused6, dim3(grid), dim3(256), 0, s->stream, (cx<float> *)s->amps, (const bf16x8 *)dfrag, a,
                       (const double *)s->d_fused, s->d_fused + 4);
    QIP_LAUNCHED();
    hipLaunchKernelGGL(k_fused_norm, dim3(1), dim3(256), 0, s->stream, (const double *)(s->d_fused + 4), grid, s->d_fused);
    QIP_LAUNCHED();
    s->fused_pending = true;
    return QIPGPU_OK;
}

int qipgpu_fused_finish(qipgpu_state *s) {
    if (!s->fused_pending) return QIPGPU_OK;
    s->fused_pending = false;
    const uint32_t nb = (uint32_t)std::min<uint64_t>((s->len + 1023) / 1024, 1u << 20);
    if (s->dtype == QIPGPU_C128)
        hipLaunchKernelGGL(k_scale<double>, dim3(nb), dim3(256), 0, s->stream, (cx<double> *)s->amps, s->len,
                           (const double *)s->d_fused);
    else
        hipLaunchKernelGGL(k_scale<float>, dim3(nb), dim3(256), 0, s->stream, (cx<float> *)s->amps, s->len,
                           (const double *)s->d_fused);
    QIP_LAUNCHED();
    return QIPGPU_OK;
}

int qipgpu_fused_last_norm(qipgpu_state *s, double *out) {
    if (!s->d_fused) return set_error(QIPGPU_EINVAL, "no fused block has been applied to this state");
    QIP_HIP(hipMemcpyAsync(out, s->d_fused + 1, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    QIP_HIP(hipStreamSynchronize(s->stream));
    return QIPGPU_OK;
}

}  // extern "C"
