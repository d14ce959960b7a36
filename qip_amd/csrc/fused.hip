// fused.hip -- precision form: a 64x64 complex operator (six fused qubits) applied to a complex64
// state on MFMA with split-BF16 compensated products, FP32 accumulation and FP64 norms.
//
// Tile. The state is cut into tiles of 4096 amplitudes: the 6 target bits T plus the 6 lowest
// bits outside T (L); together they are the 12 tile bits P, which always include state bit 0.
// Inside a tile, amplitude (r, c) has its T bits = r and L bits = c; the other n-12 bits select the
// tile. Enumerating a tile by u = its P bits in ascending order visits memory in address order:
// u and u+1 are adjacent, and runs are at least 64 amplitudes (512 B) long.
//
// GEMM. out(:, c) = U x(:, c) for the 64 columns c, written as a real 128x128 by 128x64 product
//     [Re o; Im o] = [[Ur, -Ui], [Ui, Ur]] [Re x; Im x]
// on v_mfma_f32_32x32x16_bf16. Both operands are split into BF16 hi + lo (x = xh + xl with
// |x - xh - xl| <= 2^-16 |x|; the matrix is split on the host) and each 32x32x16 block takes the
// three products hi*hi + hi*lo + lo*hi, accumulated in FP32.
//
// Data flow per tile (256 threads = 4 waves, two workgroups per CU, persistent grid):
//   1. HBM -> registers in u order, 16 B (two amplitudes) per lane, so every load instruction is
//      contiguous whatever the target bits; the next tile is prefetched while this one is
//      processed;
//   2. registers -> LDS "raw" tile in u order (rows of 64 amplitudes padded to 65 slots so column
//      gathers do not collide on banks) -> each wave gathers its 16 rows at column = lane;
//   3. split into BF16 hi/lo planes laid out [plane][column][row] (8 consecutive rows = one 16 B
//      B-fragment);
//   4. MFMA: wave w owns a 32-row M-block of the real matrix whose rows are permuted to hold Re
//      then Im of complex output rows 16w..16w+15 (A fragments, hi and lo, 64 VGPRs, stay in
//      registers for the whole loop), so each lane ends with complete complex outputs;
//   5. outputs, scaled by the pending FP64 renormalisation factor and summed into an FP64 norm,
//      go back through the raw LDS tile and leave in u order with 16 B contiguous stores.
// Per-block norm partials are summed in a fixed order by a one-block kernel that also writes the
// next renormalisation factor.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.h"

using qipgpu::cx;
using namespace qipgpu;

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

namespace {

constexpr int kTileBits = 12;
constexpr int kRowPad = 72;              // 64 rows + 8 bf16 of padding per column (bank spread)
constexpr int kPlaneElems = 64 * kRowPad;
constexpr int kRawStride = 65;           // 64 amplitudes + 1 slot of padding per raw row
constexpr size_t kFragBytes = 4 * 8 * 2 * 64 * 16;   // [wave][ks][hi/lo][lane] x 16 B = 64 KB
constexpr int kFusedBlocksPerCU = 2;   // 69 KB of LDS per workgroup

struct FusedArgs {
    uint64_t ntiles;
    uint8_t pbits[12];     // the 12 tile bits (T and L), ascending: u bit i <-> state bit pbits[i]
    uint8_t tu[6];         // u bit of row bit b (rows: b-th lowest T bit)
    uint8_t lu[6];         // u bit of column bit b (columns: b-th lowest L bit)
};

// MFMA 32x32x16 operand and result layouts (per lane l, 64 lanes):
//   A (32 x 16): A[l % 32][8 * (l / 32) + j], j = 0..7
//   B (16 x 32): B[8 * (l / 32) + j][l % 32]
//   D (32 x 32): D[8 * (v / 4) + 4 * (l / 32) + v % 4][l % 32], v = 0..15
// With M-block rows 0..15 = Re and 16..31 = Im of the same complex rows, acc[v] (v < 8) and
// acc[v + 8] are Re and Im of complex row 8 * (v / 4) + 4 * (l / 32) + v % 4, column l % 32.

// Wave-uniform address arithmetic stays on the scalar unit with constant kernel-argument
// offsets: dynamic indexing into kernel arguments would become vector loads whose waits also
// drain the prefetched tile and the previous tile's stores.
__device__ __forceinline__ uint64_t tile_base(uint64_t t, const FusedArgs &a) {
#pragma unroll
    for (int i = 0; i < 12; ++i) {
        const uint32_t p = a.pbits[i];
        const uint64_t lo = t & ((1ull << p) - 1ull);
        t = ((t ^ lo) << 1) | lo;
    }
    return t;
}

// state offset of tile-local index u (u bit i on state bit pbits[i]), for u bits [lo, hi)
template <int LO, int HI>
__device__ __forceinline__ uint64_t u_offset(uint32_t u, const FusedArgs &a) {
    uint64_t o = 0;
#pragma unroll
    for (int i = LO; i < HI; ++i) o |= (uint64_t)((u >> i) & 1u) << a.pbits[i];
    return o;
}

__device__ __forceinline__ uint32_t u_of_row(uint32_t r, const FusedArgs &a) {
    uint32_t u = 0;
#pragma unroll
    for (int b = 0; b < 6; ++b) u |= ((r >> b) & 1u) << a.tu[b];
    return u;
}

__device__ __forceinline__ uint32_t u_of_col(uint32_t c, const FusedArgs &a) {
    uint32_t u = 0;
#pragma unroll
    for (int b = 0; b < 6; ++b) u |= ((c >> b) & 1u) << a.lu[b];
    return u;
}

__device__ __forceinline__ int rawpos(uint32_t u) { return (int)(u + (u >> 6)); }

struct __attribute__((aligned(16))) TileLds {
    __bf16 planes[4][kPlaneElems];       // re_hi, re_lo, im_hi, im_lo ; [column][row]
    cx<float> raw[64 * kRawStride];      // tile in u order, 65-slot rows (input, then output)
};

// Thread t owns the amplitude pairs u = 2 (t + 256 i), i = 0..7: u bits 1..8 are t (a per-thread
// offset), u bits 9..11 are i (wave-uniform).
__device__ __forceinline__ void load_tile(cx<float> (&v)[16], const cx<float> *amps, uint64_t base,
                                          const FusedArgs &a) {
    typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f4;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const f4 p = *reinterpret_cast<const f4 *>(amps + base + u_offset<9, 12>(512u * i, a));
        v[2 * i].re = p[0];
        v[2 * i].im = p[1];
        v[2 * i + 1].re = p[2];
        v[2 * i + 1].im = p[3];
    }
}

struct TileCtx {
    uint32_t tid, lane, row0;
    uint32_t ucol_gather, ucol_out0, ucol_out1, urow_half;
    double scale;
};

// One tile: raw LDS staging, gather, BF16 split, MFMA, renormalised output back through LDS and
// out to HBM in u order. `v` holds the tile's amplitudes in u order, `tb` its base + this thread's
// u-pair offset.
__device__ __forceinline__ void process_tile(const cx<float> (&v_in)[16], uint64_t tb, cx<float> *amps, TileLds &lds,
                                             const bf16x8 (&ah)[8], const bf16x8 (&al)[8], const TileCtx &c,
                                             const FusedArgs &a, double &nacc) {
    typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f4;
    const uint32_t tid = c.tid, lane = c.lane, row0 = c.row0;
    const uint32_t ucol_gather = c.ucol_gather, ucol_out0 = c.ucol_out0, ucol_out1 = c.ucol_out1;
    const uint32_t urow_half = c.urow_half;
    const double scale = c.scale;
    cx<float> v[16];
    __syncthreads();  // (1) the previous tile's output has left the raw tile
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint32_t u = 2u * (uint32_t)tid + 512u * i;
        lds.raw[rawpos(u)] = v_in[2 * i];
        lds.raw[rawpos(u) + 1] = v_in[2 * i + 1];
    }
    __syncthreads();  // (2) raw tile complete
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = lds.raw[rawpos(u_of_row(row0 + i, a) | ucol_gather)];
    {
        bf16x8 rh[2], rl[2], ih[2], il[2];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const __bf16 xh = (__bf16)v[i].re;
            const __bf16 xl = (__bf16)(v[i].re - (float)xh);
            const __bf16 yh = (__bf16)v[i].im;
            const __bf16 yl = (__bf16)(v[i].im - (float)yh);
            rh[i >> 3][i & 7] = xh;
            rl[i >> 3][i & 7] = xl;
            ih[i >> 3][i & 7] = yh;
            il[i >> 3][i & 7] = yl;
        }
        const int o = lane * kRowPad + row0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            *reinterpret_cast<bf16x8 *>(&lds.planes[0][o + 8 * h]) = rh[h];
            *reinterpret_cast<bf16x8 *>(&lds.planes[1][o + 8 * h]) = rl[h];
            *reinterpret_cast<bf16x8 *>(&lds.planes[2][o + 8 * h]) = ih[h];
            *reinterpret_cast<bf16x8 *>(&lds.planes[3][o + 8 * h]) = il[h];
        }
    }
    __syncthreads();  // (3) planes complete; every gather from the raw tile is done

#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
        f32x16 acc = f32x16{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        const int col = 32 * nb + (lane & 31);
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) {
            const int p = ks < 4 ? 0 : 2;                      // Re rows of x, then Im rows
            const int r0 = ((16 * ks) & 63) + 8 * (lane >> 5);
            const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(&lds.planes[p][col * kRowPad + r0]);
            const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(&lds.planes[p + 1][col * kRowPad + r0]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[ks], bh, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[ks], bl, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[ks], bh, acc, 0, 0, 0);
        }
        const uint32_t uc = (nb ? ucol_out1 : ucol_out0) | urow_half;
#pragma unroll
        for (int vv = 0; vv < 8; ++vv) {
            const uint32_t r = row0 + 8 * (vv >> 2) + (vv & 3);   // row bit 2 is in urow_half
            cx<float> o;
            o.re = (float)((double)acc[vv] * scale);
            o.im = (float)((double)acc[vv + 8] * scale);
            const double ore = o.re, oim = o.im;
            nacc = fma(ore, ore, nacc);
            nacc = fma(oim, oim, nacc);
            lds.raw[rawpos(u_of_row(r, a) | uc)] = o;
        }
    }
    __syncthreads();  // (4) output tile complete in u order
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint32_t u = 2u * (uint32_t)tid + 512u * i;
        const cx<float> o0 = lds.raw[rawpos(u)], o1 = lds.raw[rawpos(u) + 1];
        f4 p;
        p[0] = o0.re;
        p[1] = o0.im;
        p[2] = o1.re;
        p[3] = o1.im;
        *reinterpret_cast<f4 *>(amps + tb + u_offset<9, 12>(512u * i, a)) = p;
    }
}

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

    // Two register buffers alternate: while one tile is processed the next one's loads are in
    // flight in the other buffer, with no copy of in-flight registers. The grid never exceeds
    // the tile count, so the first tile exists; the prefetch is issued unconditionally (the last
    // one re-reads an in-range tile) so the wait before each tile covers only its own loads.
    const uint64_t last = a.ntiles - 1;
    cx<float> va[16], vb[16];
    uint64_t tile = blockIdx.x;
    load_tile(va, amps, tile_base(tile, a) + off_t, a);
    for (;;) {
        load_tile(vb, amps, tile_base(min(tile + gridDim.x, last), a) + off_t, a);
        process_tile(va, tile_base(tile, a) + off_t, amps, lds, ah, al, ctx, a, nacc);
        tile += gridDim.x;
        if (tile >= a.ntiles) break;
        load_tile(va, amps, tile_base(min(tile + gridDim.x, last), a) + off_t, a);
        process_tile(vb, tile_base(tile, a) + off_t, amps, lds, ah, al, ctx, a, nacc);
        tile += gridDim.x;
        if (tile >= a.ntiles) break;
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

uint16_t bf16_bits_rne(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

float bf16_value(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
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
    // pad the block to six target bits with the lowest free state bits
    std::vector<int> tbits;  // state bits in the caller's matrix order (MSB first)
    uint64_t tmask = 0;
    for (int j = 0; j < k; ++j) {
        tbits.push_back(n - 1 - targets[j]);
        tmask |= 1ull << (n - 1 - targets[j]);
    }
    for (int b = 0; (int)tbits.size() < 6; ++b)
        if (!(tmask >> b & 1)) {
            tbits.push_back(b);
            tmask |= 1ull << b;
        }
    // row index r: bit b on the b-th lowest T bit. Map r -> index of the padded caller matrix
    // (caller bits MSB-first over tbits[0..k-1], padding bits are identity).
    std::vector<int> tsorted(tbits);
    std::sort(tsorted.begin(), tsorted.end());
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
    // real 128x128 operator W = [[Ur, -Ui], [Ui, Ur]] with its rows permuted so that M-block w
    // holds the Re rows then the Im rows of complex rows 16w..16w+15, split into hi/lo BF16 A
    // fragments ordered [wave][ks][hi/lo][lane][j]
    std::vector<uint16_t> frag(kFragBytes / 2);
    for (int w = 0; w < 4; ++w)
        for (int ks = 0; ks < 8; ++ks)
            for (int lane = 0; lane < 64; ++lane)
                for (int j = 0; j < 8; ++j) {
                    const int i = lane & 31;
                    const bool im_row = i >= 16;
                    const int rr = 16 * w + (i & 15);
                    const int col = 16 * ks + 8 * (lane >> 5) + j;
                    const int cc = col & 63;
                    double val;
                    if (!im_row) val = col < 64 ? ur[rr * 64 + cc] : -ui[rr * 64 + cc];
                    else val = col < 64 ? ui[rr * 64 + cc] : ur[rr * 64 + cc];
                    const float vf = (float)val;
                    const uint16_t hi = bf16_bits_rne(vf);
                    const uint16_t lo = bf16_bits_rne((float)(val - (double)bf16_value(hi)));
                    const size_t base = ((size_t)((w * 8 + ks) * 2) * 64 + lane) * 8 + j;
                    frag[base] = hi;
                    frag[base + 64 * 8] = lo;
                }
    void *dfrag = nullptr;
    rc = ring_upload(s, frag.data(), kFragBytes, &dfrag);
    if (rc) return rc;

    FusedArgs a;
    std::memset(&a, 0, sizeof(a));
    a.ntiles = 1ull << (n - kTileBits);
    std::vector<int> lbits;
    for (int b = 0; (int)lbits.size() < 6; ++b)
        if (!(tmask >> b & 1)) lbits.push_back(b);
    std::vector<int> pbits(tsorted);
    pbits.insert(pbits.end(), lbits.begin(), lbits.end());
    std::sort(pbits.begin(), pbits.end());
    for (int i = 0; i < 12; ++i) a.pbits[i] = (uint8_t)pbits[i];
    for (int b = 0; b < 6; ++b) {
        a.tu[b] = (uint8_t)(std::find(pbits.begin(), pbits.end(), tsorted[b]) - pbits.begin());
        a.lu[b] = (uint8_t)(std::find(pbits.begin(), pbits.end(), lbits[b]) - pbits.begin());
    }

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
