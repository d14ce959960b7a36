// batch.hip -- batching form: many independent circuits of uneven qubit counts, complex64 states
// in one device arena, executed together.
//
// Plan (host, at load / re-tune):
//   * "LDS states" (n <= lds_max_qubits): one workgroup per state keeps the whole state in LDS
//     and applies its gate list in order, FP64 arithmetic per gate, then renormalises in FP64 and
//     writes the state back (k_batch_small), on a stream of its own beside the waves.
//   * "Tiled states": gates fused into <= fusion_width-qubit blocks by the precision form's
//     light-cone rule (plan.h). Block j of every tiled state belongs to wave j. Per wave:
//       k_batch_build -- one workgroup per block: the 64x64 operator (block qubits padded to six
//         with free state bits, tile6.h) composed from its gates in FP64 in LDS, split into BF16
//         hi/lo MFMA fragments. Builds run on a second stream, one wave ahead, into two
//         alternating fragment sets, so they overlap the previous wave's tile passes;
//       k_batch_scales -- the renormalisation factor of each block from the FP64 norm partials of
//         the same state's previous block;
//       k_batch6 -- one workgroup per chunk of `chunk_tiles` tiles of one block (tile6.h),
//         double-buffered loads within the chunk, FP64 norm partial per chunk.
//   * k_batch_finish applies every tiled state's last renormalisation.
// Matrices are stored with their index bits in ascending state-bit order, so a gate is a bit mask
// (its qubits) on the device: sub-index = the mask's bits of an index, gathered in order.
// A run is captured once as a hipGraph and replayed (tuning.use_graph), or launched eagerly.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "engine.h"
#include "plan.h"
#include "tile6.h"

using qipgpu::cx;
using namespace qipgpu;
using namespace qipgpu::tile;

namespace {

constexpr int kMaxBatchQubits = 30;
constexpr int kLdsMaxQubits = 13;        // 2 x 2^13 complex64 = 128 KB of LDS
constexpr int kSmallThreads = 512;
constexpr uint32_t kFinishSegment = 1u << 16;   // amplitudes per finishing workgroup
constexpr size_t kStateAlign = 512;             // amplitudes (4 KB) between state slices

struct BGate {
    uint32_t mat;    // offset of the 2^k x 2^k complex128 matrix (ascending-bit order)
    uint32_t k;
    uint32_t mask;   // the gate's bits: state bits (LDS states) or tile row bits (blocks)
    uint32_t pad;
};

struct BBlock {
    uint64_t amp_off;       // state slice offset in the arena (amplitudes)
    FusedArgs a;            // tile arguments of the padded six target bits
    uint32_t gate0, ngates; // range in the block gate list
    uint32_t frag;          // fragment slot within the wave
    uint32_t prev_part, prev_count;   // norm partials of this state's previous block (0: none)
    uint32_t pad;
};

struct BChunk {
    uint32_t block;         // index into the block list
    uint32_t tile0, ntiles;
    uint32_t part;          // norm partial slot
};

struct BSmall {
    uint64_t amp_off;
    uint32_t n, gate0, ngates, pad;
};

struct BFinish {
    uint64_t amp_off;       // segment start (amplitudes)
    uint32_t count;         // amplitudes in the segment
    uint32_t part, npart;   // norm partials of the state's last block
    uint32_t pad;
};

__device__ __forceinline__ uint32_t gather_bits(uint32_t x, uint32_t m) {
    uint32_t r = 0, bit = 1;
    while (m) {
        const uint32_t low = m & (0u - m);
        if (x & low) r |= bit;
        bit <<= 1;
        m ^= low;
    }
    return r;
}

__device__ __forceinline__ cx<double> cmul_acc(cx<double> acc, cx<double> g, cx<double> x) {
    acc.re = fma(g.re, x.re, acc.re);
    acc.re = fma(-g.im, x.im, acc.re);
    acc.im = fma(g.re, x.im, acc.im);
    acc.im = fma(g.im, x.re, acc.im);
    return acc;
}

// Fixed-order sum of partials[p0 .. p0 + count) over the workgroup (256 threads); every thread
// returns the same value.
__device__ double group_sum_256(const double *__restrict__ partials, uint32_t p0, uint32_t count, double *red) {
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < count; i += 256) s += partials[p0 + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const double total = red[0];
    __syncthreads();
    return total;
}

__device__ __forceinline__ double renorm_factor(double norm) { return norm > 0.0 ? 1.0 / sqrt(norm) : 1.0; }

// ---- reset ----------------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_batch_zero(cx<float> *arena, uint64_t len) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
        cx<float> z;
        z.re = 0.0f;
        z.im = 0.0f;
        arena[i] = z;
    }
}

__global__ __launch_bounds__(256) void k_batch_ones(cx<float> *arena, const uint64_t *__restrict__ offsets,
                                                    uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        cx<float> one;
        one.re = 1.0f;
        one.im = 0.0f;
        arena[offsets[i]] = one;
    }
}

// ---- LDS states -----------------------------------------------------------------------------

__global__ __launch_bounds__(kSmallThreads) void k_batch_small(cx<float> *arena, const BSmall *__restrict__ smalls,
                                                               const BGate *__restrict__ gates,
                                                               const cx<double> *__restrict__ mats) {
    __shared__ cx<float> buf[2 << kLdsMaxQubits];
    __shared__ double red[kSmallThreads];
    const BSmall st = smalls[blockIdx.x];
    const uint32_t len = 1u << st.n;
    const uint32_t tid = threadIdx.x;
    cx<float> *amps = arena + st.amp_off;
    cx<float> *cur = buf, *oth = buf + len;
    for (uint32_t i = tid; i < len; i += kSmallThreads) cur[i] = amps[i];
    __syncthreads();
    for (uint32_t g = 0; g < st.ngates; ++g) {
        const BGate gt = gates[st.gate0 + g];
        const cx<double> *G = mats + gt.mat;
        if (gt.k == 1) {
            const uint32_t m = gt.mask, lowm = m - 1u;
            const cx<double> g00 = G[0], g01 = G[1], g10 = G[2], g11 = G[3];
            for (uint32_t p = tid; p < (len >> 1); p += kSmallThreads) {
                const uint32_t i0 = ((p & ~lowm) << 1) | (p & lowm), i1 = i0 | m;
                const cx<float> a = cur[i0], b = cur[i1];
                cx<double> x0, x1, y0 = {0.0, 0.0}, y1 = {0.0, 0.0};
                x0.re = a.re;
                x0.im = a.im;
                x1.re = b.re;
                x1.im = b.im;
                y0 = cmul_acc(cmul_acc(y0, g00, x0), g01, x1);
                y1 = cmul_acc(cmul_acc(y1, g10, x0), g11, x1);
                cx<float> o0, o1;
                o0.re = (float)y0.re;
                o0.im = (float)y0.im;
                o1.re = (float)y1.re;
                o1.im = (float)y1.im;
                cur[i0] = o0;
                cur[i1] = o1;
            }
        } else if (gt.k == 2) {
            const uint32_t mlo = gt.mask & (0u - gt.mask), mhi = gt.mask ^ mlo;
            for (uint32_t p = tid; p < (len >> 2); p += kSmallThreads) {
                uint32_t i = ((p & ~(mlo - 1u)) << 1) | (p & (mlo - 1u));
                i = ((i & ~(mhi - 1u)) << 1) | (i & (mhi - 1u));
                const uint32_t idx[4] = {i, i | mlo, i | mhi, i | mlo | mhi};
                cx<double> x[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const cx<float> v = cur[idx[t]];
                    x[t].re = v.re;
                    x[t].im = v.im;
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    cx<double> y = {0.0, 0.0};
#pragma unroll
                    for (int t = 0; t < 4; ++t) y = cmul_acc(y, G[4 * r + t], x[t]);
                    cx<float> o;
                    o.re = (float)y.re;
                    o.im = (float)y.im;
                    cur[idx[r]] = o;
                }
            }
        } else {
            // k >= 3: each output from the 2^k inputs of its group, into the other buffer
            const uint32_t m = gt.mask, D = 1u << gt.k;
            for (uint32_t e = tid; e < len; e += kSmallThreads) {
                const uint32_t s = gather_bits(e, m), base = e & ~m;
                cx<double> y = {0.0, 0.0};
                uint32_t d = 0;
                for (uint32_t t = 0; t < D; ++t) {
                    const cx<float> v = cur[base | d];
                    cx<double> xv;
                    xv.re = v.re;
                    xv.im = v.im;
                    y = cmul_acc(y, G[s * D + t], xv);
                    d = (d - m) & m;
                }
                cx<float> o;
                o.re = (float)y.re;
                o.im = (float)y.im;
                oth[e] = o;
            }
            cx<float> *t = cur;
            cur = oth;
            oth = t;
        }
        __syncthreads();
    }
    // FP64 norm in a fixed order, then renormalise on the way out
    double s = 0.0;
    for (uint32_t i = tid; i < len; i += kSmallThreads) {
        const double re = cur[i].re, im = cur[i].im;
        s = fma(re, re, s);
        s = fma(im, im, s);
    }
    red[tid] = s;
    __syncthreads();
    for (int o = kSmallThreads / 2; o > 0; o >>= 1) {
        if ((int)tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    const double scale = renorm_factor(red[0]);
    for (uint32_t i = tid; i < len; i += kSmallThreads) {
        const cx<float> v = cur[i];
        cx<float> o;
        o.re = (float)((double)v.re * scale);
        o.im = (float)((double)v.im * scale);
        amps[i] = o;
    }
}

// ---- tiled states ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_batch_build(const BBlock *__restrict__ blocks, uint32_t block0,
                                                     const BGate *__restrict__ gates,
                                                     const cx<double> *__restrict__ mats, bf16x8 *frags) {
    __shared__ cx<double> m[2][4096];   // [row][column], row = tile row index (row bits)
    const BBlock bk = blocks[block0 + blockIdx.x];
    const uint32_t tid = threadIdx.x;
    for (uint32_t e = tid; e < 4096; e += 256) {
        cx<double> v;
        v.re = (e >> 6) == (e & 63) ? 1.0 : 0.0;
        v.im = 0.0;
        m[0][e] = v;
    }
    int cur = 0;
    __syncthreads();
    // Lane = column throughout, so rows, gate sub-indices and matrix entries are wave-uniform and
    // come through the scalar cache. One- and two-qubit gates (the common case) update groups of
    // rows in place with the matrix in registers; wider gates go out of place into the other
    // buffer, one output row at a time.
    const uint32_t wv = __builtin_amdgcn_readfirstlane(tid >> 6), c = tid & 63;
    for (uint32_t g = 0; g < bk.ngates; ++g) {
        const BGate gt = gates[bk.gate0 + g];
        const cx<double> *G = mats + gt.mat;
        const uint32_t msk = gt.mask;
        if (gt.k == 1) {
            const cx<double> g00 = G[0], g01 = G[1], g10 = G[2], g11 = G[3];
            const uint32_t lowm = msk - 1u;
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q) {          // 32 row pairs, 8 per wave
                const uint32_t j = wv + 4 * q;
                const uint32_t r0 = ((j & ~lowm) << 1) | (j & lowm), r1 = r0 | msk;
                const cx<double> x0 = m[cur][(r0 << 6) | c], x1 = m[cur][(r1 << 6) | c];
                cx<double> y0 = {0.0, 0.0}, y1 = {0.0, 0.0};
                y0 = cmul_acc(cmul_acc(y0, g00, x0), g01, x1);
                y1 = cmul_acc(cmul_acc(y1, g10, x0), g11, x1);
                m[cur][(r0 << 6) | c] = y0;
                m[cur][(r1 << 6) | c] = y1;
            }
        } else if (gt.k == 2) {
            cx<double> gm[16];
#pragma unroll
            for (int t = 0; t < 16; ++t) gm[t] = G[t];
            const uint32_t mlo = msk & (0u - msk), mhi = msk ^ mlo;
#pragma unroll
            for (uint32_t q = 0; q < 4; ++q) {          // 16 row quads, 4 per wave
                const uint32_t j = wv + 4 * q;
                uint32_t r = ((j & ~(mlo - 1u)) << 1) | (j & (mlo - 1u));
                r = ((r & ~(mhi - 1u)) << 1) | (r & (mhi - 1u));
                const uint32_t rows[4] = {r, r | mlo, r | mhi, r | mlo | mhi};
                cx<double> x[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) x[t] = m[cur][(rows[t] << 6) | c];
#pragma unroll
                for (int o = 0; o < 4; ++o) {
                    cx<double> y = {0.0, 0.0};
#pragma unroll
                    for (int t = 0; t < 4; ++t) y = cmul_acc(y, gm[4 * o + t], x[t]);
                    m[cur][(rows[o] << 6) | c] = y;
                }
            }
        } else {
            const uint32_t D = 1u << gt.k;
            for (uint32_t i = 0; i < 16; ++i) {
                const uint32_t r = wv + 4 * i;
                const uint32_t s = gather_bits(r, msk), rb = r & ~msk;
                cx<double> y = {0.0, 0.0};
                uint32_t d = 0;
                for (uint32_t t = 0; t < D; ++t) {
                    y = cmul_acc(y, G[s * D + t], m[cur][((rb | d) << 6) | c]);
                    d = (d - msk) & msk;
                }
                m[cur ^ 1][(r << 6) | c] = y;
            }
            cur ^= 1;
        }
        __syncthreads();
    }
    // split-BF16 A fragments, [wave][ks][hi/lo][lane] x 8
    bf16x8 *out = frags + (size_t)bk.frag * (kFragBytes / 16);
    for (uint32_t q = tid; q < 2048; q += 256) {
        const int w = (int)(q >> 9), ks = (int)((q >> 6) & 7), lane = (int)(q & 63);
        bf16x8 hi, lo;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const FragEntry fe = frag_entry(w, ks, lane, j);
            const cx<double> u = m[cur][fe.rr * 64 + fe.cc];
            double val = fe.imag ? u.im : u.re;
            if (fe.negate) val = -val;
            const __bf16 h = (__bf16)(float)val;
            hi[j] = h;
            lo[j] = (__bf16)(float)(val - (double)(float)h);
        }
        out[((w * 8 + ks) * 2 + 0) * 64 + lane] = hi;
        out[((w * 8 + ks) * 2 + 1) * 64 + lane] = lo;
    }
}

// Renormalisation factor of each block of a wave: 1 / sqrt of the FP64 norm of the same state's
// previous block output (sum of its chunk partials, fixed order).
__global__ __launch_bounds__(256) void k_batch_scales(const BBlock *__restrict__ blocks, uint32_t block0,
                                                      const double *__restrict__ partials, double *scales) {
    __shared__ double red[256];
    const uint32_t bi = block0 + blockIdx.x;
    const BBlock bk = blocks[bi];
    if (!bk.prev_count) {
        if (threadIdx.x == 0) scales[bi] = 1.0;
        return;
    }
    const double norm = group_sum_256(partials, bk.prev_part, bk.prev_count, red);
    if (threadIdx.x == 0) scales[bi] = renorm_factor(norm);
}

__global__ __launch_bounds__(256, 2) void k_batch6(cx<float> *arena, const BBlock *__restrict__ blocks,
                                                   const BChunk *__restrict__ chunks, uint32_t chunk0,
                                                   const bf16x8 *__restrict__ frags,
                                                   const double *__restrict__ scales, double *partials) {
    __shared__ TileLds lds;
    __shared__ double red[4];
    const BChunk ch = chunks[chunk0 + blockIdx.x];
    const BBlock bk = blocks[ch.block];
    const FusedArgs a = bk.a;
    cx<float> *amps = arena + bk.amp_off;
    const bf16x8 *wfrag = frags + (size_t)bk.frag * (kFragBytes / 16);
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;

    bf16x8 ah[8], al[8];
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
        ah[ks] = wfrag[((wave * 8 + ks) * 2 + 0) * 64 + lane];
        al[ks] = wfrag[((wave * 8 + ks) * 2 + 1) * 64 + lane];
    }
    TileCtx ctx;
    ctx.tid = tid;
    ctx.lane = lane;
    ctx.row0 = wave * 16;
    ctx.ucol_gather = u_of_col(lane, a);
    ctx.ucol_out0 = u_of_col(lane & 31, a);
    ctx.ucol_out1 = u_of_col(32 + (lane & 31), a);
    ctx.urow_half = (lane >> 5) ? u_of_row(4, a) : 0u;
    ctx.scale = scales[ch.block];
    const uint64_t off_t = u_offset<0, 9>(2u * (uint32_t)tid, a);

    // Same two-buffer pipeline as the single-state kernel, over this chunk's tiles; the last
    // prefetch re-reads the chunk's last tile (in range, never used).
    double nacc = 0.0;
    const uint64_t last = (uint64_t)ch.tile0 + ch.ntiles - 1;
    cx<float> va[16], vb[16];
    uint64_t tile = ch.tile0;
    load_tile(va, amps, tile_base(tile, a) + off_t, a);
    for (;;) {
        load_tile(vb, amps, tile_base(min(tile + 1, last), a) + off_t, a);
        process_tile(va, tile_base(tile, a) + off_t, amps, lds, ah, al, ctx, a, nacc);
        if (++tile > last) break;
        load_tile(va, amps, tile_base(min(tile + 1, last), a) + off_t, a);
        process_tile(vb, tile_base(tile, a) + off_t, amps, lds, ah, al, ctx, a, nacc);
        if (++tile > last) break;
    }
    for (int off = 32; off > 0; off >>= 1) nacc += __shfl_down(nacc, off, 64);
    if (lane == 0) red[wave] = nacc;
    __syncthreads();
    if (tid == 0) partials[ch.part] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void k_batch_finish(cx<float> *arena, const BFinish *__restrict__ segs,
                                                      const double *__restrict__ partials) {
    __shared__ double red[256];
    const BFinish f = segs[blockIdx.x];
    const double scale = renorm_factor(group_sum_256(partials, f.part, f.npart, red));
    cx<float> *amps = arena + f.amp_off;
    for (uint32_t i = threadIdx.x; i < f.count; i += 256) {
        const cx<float> v = amps[i];
        cx<float> o;
        o.re = (float)((double)v.re * scale);
        o.im = (float)((double)v.im * scale);
        amps[i] = o;
    }
}

// ---- host planning ----------------------------------------------------------------------------

struct Tuning {
    int32_t fusion_width = 6;
    int32_t lds_max_qubits = kLdsMaxQubits;
    int32_t chunk_tiles = 8;
    int32_t use_graph = 1;
};

int check_tuning(const Tuning &t) {
    if (t.fusion_width < 1 || t.fusion_width > 6)
        return set_error(QIPGPU_EINVAL, "fusion_width %d outside 1..6", t.fusion_width);
    if (t.lds_max_qubits < kTileBits - 1 || t.lds_max_qubits > kLdsMaxQubits)
        return set_error(QIPGPU_EINVAL, "lds_max_qubits %d outside %d..%d", t.lds_max_qubits, kTileBits - 1,
                         kLdsMaxQubits);
    if (t.chunk_tiles < 1 || t.chunk_tiles > 64)
        return set_error(QIPGPU_EINVAL, "chunk_tiles %d outside 1..64", t.chunk_tiles);
    if (t.use_graph != 0 && t.use_graph != 1) return set_error(QIPGPU_EINVAL, "use_graph must be 0 or 1");
    return QIPGPU_OK;
}

Tuning tuning_from(const qipgpu_batch_tuning *t) {
    Tuning r;
    if (t) {
        r.fusion_width = t->fusion_width;
        r.lds_max_qubits = t->lds_max_qubits;
        r.chunk_tiles = t->chunk_tiles;
        r.use_graph = t->use_graph;
    }
    return r;
}

// Validated gate list: per gate its state, k, qubits and the offset of its matrix in the packed
// input (complex units).
struct GateList {
    std::vector<int32_t> state, k;
    std::vector<int32_t> qubits;     // 6 per gate
    std::vector<uint64_t> mat_in;    // offset of the caller's matrix
    uint64_t mat_total = 0;
};

int validate_gates(int32_t count, const std::vector<int32_t> &ns, int64_t n_gates, const int32_t *gate_state,
                   const int32_t *gate_k, const int32_t *gate_qubits, GateList &gl) {
    if (n_gates < 0) return set_error(QIPGPU_EINVAL, "negative gate count");
    if (n_gates > 0 && (!gate_state || !gate_k || !gate_qubits)) return set_error(QIPGPU_EINVAL, "null gate arrays");
    gl.state.assign(gate_state, gate_state + n_gates);
    gl.k.assign(gate_k, gate_k + n_gates);
    gl.qubits.assign(gate_qubits, gate_qubits + 6 * n_gates);
    gl.mat_in.resize(n_gates);
    uint64_t off = 0;
    for (int64_t g = 0; g < n_gates; ++g) {
        const int32_t st = gl.state[g], k = gl.k[g];
        if (st < 0 || st >= count) return set_error(QIPGPU_EINVAL, "gate %lld: state %d outside 0..%d", (long long)g, st, count - 1);
        if (k < 1 || k > 6) return set_error(QIPGPU_EINVAL, "gate %lld: %d qubits (1..6)", (long long)g, k);
        const int n = ns[st];
        uint64_t seen = 0;
        for (int j = 0; j < k; ++j) {
            const int32_t q = gl.qubits[6 * g + j];
            if (q < 0 || q >= n)
                return set_error(QIPGPU_EINVAL, "gate %lld: qubit %d outside 0..%d of state %d", (long long)g, q, n - 1, st);
            if (seen >> q & 1) return set_error(QIPGPU_EINVAL, "gate %lld: qubit %d repeated", (long long)g, q);
            seen |= 1ull << q;
        }
        gl.mat_in[g] = off;
        off += 1ull << (2 * k);
    }
    gl.mat_total = off;
    return QIPGPU_OK;
}

// Fusion of one state's gates (program order) by the light-cone rule of plan.h: the block qubit masks
// in launch order and each gate's block.
void fuse_state(const std::vector<int64_t> &gids, const GateList &gl, int width, std::vector<uint64_t> &block_mask,
                std::vector<int32_t> &gate_block) {
    block_mask.clear();
    std::vector<uint64_t> masks(gids.size(), 0);
    for (size_t i = 0; i < gids.size(); ++i)
        for (int t = 0; t < gl.k[gids[i]]; ++t) masks[i] |= 1ull << gl.qubits[6 * gids[i] + t];
    std::vector<PlannedBlock> blocks;
    plan_blocks(masks.data(), (int64_t)masks.size(), width, 64, false, blocks);
    for (size_t b = 0; b < blocks.size(); ++b) {
        block_mask.push_back(blocks[b].mask);
        for (int64_t i : blocks[b].gates) gate_block[gids[i]] = (int32_t)b;
    }
}

struct Plan {
    std::vector<BGate> sgates, bgates;
    std::vector<BSmall> smalls;
    std::vector<BBlock> blocks;
    std::vector<BChunk> chunks;
    std::vector<BFinish> finish;
    std::vector<uint32_t> wave_block0, wave_chunk0;   // size waves + 1
    uint32_t max_wave_blocks = 0;
    uint32_t nparts = 0;
    int64_t n_small = 0, n_tiled = 0;
    uint64_t algorithmic_bytes = 0;
};

// mask over `bits` (state bits), the gate's qubits mapped through `rank` (state bit -> row bit,
// or identity)
template <typename F>
uint32_t gate_mask(const GateList &gl, int64_t g, int n, F rank) {
    uint32_t m = 0;
    for (int t = 0; t < gl.k[g]; ++t) m |= 1u << rank(n - 1 - gl.qubits[6 * g + t]);
    return m;
}

int build_plan(const std::vector<int32_t> &ns, const std::vector<uint64_t> &offs, const GateList &gl,
               const std::vector<uint64_t> &mat_dev, const Tuning &t, Plan &p) {
    const int32_t count = (int32_t)ns.size();
    std::vector<std::vector<int64_t>> per_state(count);
    for (int64_t g = 0; g < (int64_t)gl.state.size(); ++g) per_state[gl.state[g]].push_back(g);
    p = Plan();
    std::vector<int32_t> gate_block(gl.state.size(), -1);
    std::vector<std::vector<uint64_t>> bmask(count);
    std::vector<int32_t> tiled;
    for (int32_t i = 0; i < count; ++i) {
        const uint64_t len = 1ull << ns[i];
        if (ns[i] <= t.lds_max_qubits) {
            ++p.n_small;
            if (per_state[i].empty()) continue;   // no gates: the state is left as it is
            BSmall s;
            s.amp_off = offs[i];
            s.n = (uint32_t)ns[i];
            s.gate0 = (uint32_t)p.sgates.size();
            s.ngates = (uint32_t)per_state[i].size();
            s.pad = 0;
            for (int64_t g : per_state[i]) {
                BGate bg;
                bg.mat = (uint32_t)mat_dev[g];
                bg.k = (uint32_t)gl.k[g];
                bg.mask = gate_mask(gl, g, ns[i], [](int b) { return b; });
                bg.pad = 0;
                p.sgates.push_back(bg);
            }
            p.smalls.push_back(s);
            p.algorithmic_bytes += 2 * 8 * len;
        } else {
            fuse_state(per_state[i], gl, t.fusion_width, bmask[i], gate_block);
            if (!bmask[i].empty()) tiled.push_back(i);
            ++p.n_tiled;
        }
    }
    // larger states first inside each wave
    std::stable_sort(tiled.begin(), tiled.end(), [&](int32_t x, int32_t y) { return ns[x] > ns[y]; });
    size_t waves = 0;
    for (int32_t i : tiled) waves = std::max(waves, bmask[i].size());
    // gates of each (state, block), in program order
    std::vector<std::vector<std::vector<int64_t>>> bgl(count);
    for (int32_t i : tiled) {
        bgl[i].resize(bmask[i].size());
        for (int64_t g : per_state[i]) bgl[i][gate_block[g]].push_back(g);
    }
    std::vector<uint32_t> prev_part(count, 0), prev_count(count, 0);
    for (size_t w = 0; w < waves; ++w) {
        p.wave_block0.push_back((uint32_t)p.blocks.size());
        p.wave_chunk0.push_back((uint32_t)p.chunks.size());
        uint32_t slot = 0;
        for (int32_t i : tiled) {
            if (w >= bmask[i].size()) continue;
            const int n = ns[i];
            // tile bits: the block's qubits padded to six target bits, and six column bits
            uint64_t tmask = 0;
            const uint64_t qm = bmask[i][w];
            for (int q = 0; q < n; ++q)
                if (qm >> q & 1) tmask |= 1ull << (n - 1 - q);
            int ts[6], ls[6], rank_of[64];
            choose_tile_bits(n, tmask, ts, ls);
            for (int b = 0; b < 6; ++b) rank_of[ts[b]] = b;
            BBlock bk;
            std::memset(&bk, 0, sizeof(bk));
            bk.amp_off = offs[i];
            bk.a = make_tile_args(n, ts, ls);
            bk.gate0 = (uint32_t)p.bgates.size();
            bk.ngates = (uint32_t)bgl[i][w].size();
            bk.frag = slot++;
            bk.prev_part = prev_part[i];
            bk.prev_count = prev_count[i];
            for (int64_t g : bgl[i][w]) {
                BGate bg;
                bg.mat = (uint32_t)mat_dev[g];
                bg.k = (uint32_t)gl.k[g];
                bg.mask = gate_mask(gl, g, n, [&](int b) { return rank_of[b]; });
                bg.pad = 0;
                p.bgates.push_back(bg);
            }
            const uint32_t bidx = (uint32_t)p.blocks.size();
            p.blocks.push_back(bk);
            const uint64_t ntiles = bk.a.ntiles;
            prev_part[i] = p.nparts;
            prev_count[i] = 0;
            for (uint64_t t0 = 0; t0 < ntiles; t0 += (uint64_t)t.chunk_tiles) {
                BChunk c;
                c.block = bidx;
                c.tile0 = (uint32_t)t0;
                c.ntiles = (uint32_t)std::min<uint64_t>((uint64_t)t.chunk_tiles, ntiles - t0);
                c.part = p.nparts++;
                p.chunks.push_back(c);
                ++prev_count[i];
            }
            p.algorithmic_bytes += 2 * 8 * (1ull << n);
        }
        p.max_wave_blocks = std::max(p.max_wave_blocks, slot);
    }
    p.wave_block0.push_back((uint32_t)p.blocks.size());
    p.wave_chunk0.push_back((uint32_t)p.chunks.size());
    for (int32_t i : tiled) {
        const uint64_t len = 1ull << ns[i];
        for (uint64_t s0 = 0; s0 < len; s0 += kFinishSegment) {
            BFinish f;
            f.amp_off = offs[i] + s0;
            f.count = (uint32_t)std::min<uint64_t>(kFinishSegment, len - s0);
            f.part = prev_part[i];
            f.npart = prev_count[i];
            f.pad = 0;
            p.finish.push_back(f);
        }
        p.algorithmic_bytes += 2 * 8 * len;
    }
    return QIPGPU_OK;
}

// Matrix of gate g re-indexed so that index bit t is the gate's t-th lowest state bit (caller's
// order: bit k-1-j is gate qubit j).
void ascending_matrix(const GateList &gl, int64_t g, int n, const double *in, std::vector<double> &out) {
    const int k = gl.k[g], D = 1 << k;
    int bits[6], order[6];
    for (int j = 0; j < k; ++j) {
        bits[j] = n - 1 - gl.qubits[6 * g + j];
        order[j] = j;
    }
    std::sort(order, order + k, [&](int x, int y) { return bits[x] < bits[y]; });
    // caller index of ascending index v: ascending bit t is gate qubit order[t] = caller bit k-1-order[t]
    std::vector<int> cidx(D);
    for (int v = 0; v < D; ++v) {
        int ci = 0;
        for (int t = 0; t < k; ++t)
            if (v >> t & 1) ci |= 1 << (k - 1 - order[t]);
        cidx[v] = ci;
    }
    const size_t base = out.size();
    out.resize(base + 2 * (size_t)D * D);
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) {
            out[base + 2 * (r * D + c)] = in[2 * (cidx[r] * D + cidx[c])];
            out[base + 2 * (r * D + c) + 1] = in[2 * (cidx[r] * D + cidx[c]) + 1];
        }
}

template <typename T>
int upload(std::vector<void *> &owned, const std::vector<T> &v, T **dev) {
    *dev = nullptr;
    if (v.empty()) return QIPGPU_OK;
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(T) * v.size()) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(QIPGPU_ENOMEM, "batch plan: device allocation of %zu bytes failed", sizeof(T) * v.size());
    }
    owned.push_back(p);
    QIP_HIP(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    *dev = (T *)p;
    return QIPGPU_OK;
}

}  // namespace

struct qipgpu_batch {
    int32_t count = 0;
    std::vector<int32_t> ns;
    std::vector<uint64_t> offs;
    uint64_t arena_len = 0;
    cx<float> *arena = nullptr;
    uint64_t *d_offs = nullptr;
    hipStream_t stream = nullptr;   // waves, finish, resets, copies
    hipStream_t side = nullptr;     // fragment builds, one wave ahead
    hipStream_t lds = nullptr;      // LDS states
    hipEvent_t fork = nullptr, join = nullptr;
    std::vector<hipEvent_t> ev_built, ev_applied;   // per wave
    // loaded gates
    GateList gl;
    std::vector<uint64_t> mat_dev;      // per gate: offset in d_mats (complex units)
    cx<double> *d_mats = nullptr;
    bool loaded = false;
    // plan
    Tuning tuning;
    Plan plan;
    std::vector<void *> plan_mem;
    BGate *d_sgates = nullptr, *d_bgates = nullptr;
    BSmall *d_smalls = nullptr;
    BBlock *d_blocks = nullptr;
    BChunk *d_chunks = nullptr;
    BFinish *d_finish = nullptr;
    double *d_partials = nullptr, *d_scales = nullptr;
    bf16x8 *d_frags = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void free_plan(qipgpu_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    for (void *p : b->plan_mem) (void)hipFree(p);
    b->plan_mem.clear();
    for (hipEvent_t e : b->ev_built) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->ev_applied) (void)hipEventDestroy(e);
    b->ev_built.clear();
    b->ev_applied.clear();
    b->d_sgates = b->d_bgates = nullptr;
    b->d_smalls = nullptr;
    b->d_blocks = nullptr;
    b->d_chunks = nullptr;
    b->d_finish = nullptr;
    b->d_partials = b->d_scales = nullptr;
    b->d_frags = nullptr;
    (void)hipGetLastError();
}

// Enqueue one run of the plan. The run starts and ends on `stream`, which carries the tile passes,
// the scale kernels and the finish. `side` builds wave w's fragments into set w % 2 once wave
// w - 2's tile pass has released that set, one wave ahead of the tile passes; `lds` runs the LDS
// states. Both join back into `stream` before the run ends.
int enqueue_run(qipgpu_batch *b) {
    const Plan &p = b->plan;
    const size_t waves = p.wave_block0.empty() ? 0 : p.wave_block0.size() - 1;
    const bool small = !p.smalls.empty();
    if (!small && waves == 0) return QIPGPU_OK;
    QIP_HIP(hipEventRecord(b->fork, b->stream));
    if (small) {
        QIP_HIP(hipStreamWaitEvent(b->lds, b->fork, 0));
        hipLaunchKernelGGL(k_batch_small, dim3((uint32_t)p.smalls.size()), dim3(kSmallThreads), 0, b->lds, b->arena,
                           (const BSmall *)b->d_smalls, (const BGate *)b->d_sgates, (const cx<double> *)b->d_mats);
        QIP_LAUNCHED();
        QIP_HIP(hipEventRecord(b->join, b->lds));
    }
    if (waves) QIP_HIP(hipStreamWaitEvent(b->side, b->fork, 0));
    const size_t set = (size_t)p.max_wave_blocks * (kFragBytes / 16);
    for (size_t w = 0; w < waves; ++w) {
        const uint32_t nb = p.wave_block0[w + 1] - p.wave_block0[w];
        const uint32_t nc = p.wave_chunk0[w + 1] - p.wave_chunk0[w];
        bf16x8 *frags = b->d_frags + (w & 1) * set;
        if (w >= 2) QIP_HIP(hipStreamWaitEvent(b->side, b->ev_applied[w - 2], 0));
        hipLaunchKernelGGL(k_batch_build, dim3(nb), dim3(256), 0, b->side, (const BBlock *)b->d_blocks,
                           p.wave_block0[w], (const BGate *)b->d_bgates, (const cx<double> *)b->d_mats, frags);
        QIP_LAUNCHED();
        QIP_HIP(hipEventRecord(b->ev_built[w], b->side));
        if (w > 0) {   // wave 0 blocks have no previous block: their factor stays 1
            hipLaunchKernelGGL(k_batch_scales, dim3(nb), dim3(256), 0, b->stream, (const BBlock *)b->d_blocks,
                               p.wave_block0[w], (const double *)b->d_partials, b->d_scales);
            QIP_LAUNCHED();
        }
        QIP_HIP(hipStreamWaitEvent(b->stream, b->ev_built[w], 0));
        hipLaunchKernelGGL(k_batch6, dim3(nc), dim3(256), 0, b->stream, b->arena, (const BBlock *)b->d_blocks,
                           (const BChunk *)b->d_chunks, p.wave_chunk0[w], (const bf16x8 *)frags,
                           (const double *)b->d_scales, b->d_partials);
        QIP_LAUNCHED();
        QIP_HIP(hipEventRecord(b->ev_applied[w], b->stream));
    }
    if (!p.finish.empty()) {
        hipLaunchKernelGGL(k_batch_finish, dim3((uint32_t)p.finish.size()), dim3(256), 0, b->stream, b->arena,
                           (const BFinish *)b->d_finish, (const double *)b->d_partials);
        QIP_LAUNCHED();
    }
    if (small) QIP_HIP(hipStreamWaitEvent(b->stream, b->join, 0));
    return QIPGPU_OK;
}

int sync_streams(qipgpu_batch *b) {
    QIP_HIP(hipStreamSynchronize(b->stream));
    QIP_HIP(hipStreamSynchronize(b->side));
    QIP_HIP(hipStreamSynchronize(b->lds));
    return QIPGPU_OK;
}

int make_plan(qipgpu_batch *b, const Tuning &t) {
    int rc = check_tuning(t);
    if (rc) return rc;
    if ((rc = sync_streams(b))) return rc;
    free_plan(b);
    b->tuning = t;
    rc = build_plan(b->ns, b->offs, b->gl, b->mat_dev, t, b->plan);
    if (rc) return rc;
    const Plan &p = b->plan;
    if ((rc = upload(b->plan_mem, p.sgates, &b->d_sgates))) return rc;
    if ((rc = upload(b->plan_mem, p.bgates, &b->d_bgates))) return rc;
    if ((rc = upload(b->plan_mem, p.smalls, &b->d_smalls))) return rc;
    if ((rc = upload(b->plan_mem, p.blocks, &b->d_blocks))) return rc;
    if ((rc = upload(b->plan_mem, p.chunks, &b->d_chunks))) return rc;
    if ((rc = upload(b->plan_mem, p.finish, &b->d_finish))) return rc;
    std::vector<double> zeros(std::max<uint32_t>(p.nparts, 1), 0.0);
    if ((rc = upload(b->plan_mem, zeros, &b->d_partials))) return rc;
    std::vector<double> ones(std::max<size_t>(p.blocks.size(), 1), 1.0);
    if ((rc = upload(b->plan_mem, ones, &b->d_scales))) return rc;
    const size_t waves = p.wave_block0.empty() ? 0 : p.wave_block0.size() - 1;
    for (size_t w = 0; w < waves; ++w) {
        hipEvent_t e1 = nullptr, e2 = nullptr;
        QIP_HIP(hipEventCreateWithFlags(&e1, hipEventDisableTiming));
        b->ev_built.push_back(e1);
        QIP_HIP(hipEventCreateWithFlags(&e2, hipEventDisableTiming));
        b->ev_applied.push_back(e2);
    }
    if (p.max_wave_blocks) {   // two alternating fragment sets
        void *f = nullptr;
        if (hipMalloc(&f, 2 * kFragBytes * (size_t)p.max_wave_blocks) != hipSuccess) {
            (void)hipGetLastError();
            return set_error(QIPGPU_ENOMEM, "batch plan: fragment buffer of %u blocks", p.max_wave_blocks);
        }
        b->plan_mem.push_back(f);
        b->d_frags = (bf16x8 *)f;
    }
    const bool work = !p.smalls.empty() || !p.blocks.empty();
    if (t.use_graph && work) {
        QIP_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_run(b);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return set_error(QIPGPU_EHIP, "batch graph capture failed: %s", hipGetErrorString(e));
        b->graph = g;
        QIP_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    return QIPGPU_OK;
}

int run_batch(qipgpu_batch *b) {
    if (!b->loaded) {
        if (b->gl.state.empty()) return QIPGPU_OK;   // never given gates: nothing to run
        return set_error(QIPGPU_EINVAL, "the batch has no usable plan (its last load or re-tune failed)");
    }
    if (b->exec) {
        QIP_HIP(hipGraphLaunch(b->exec, b->stream));
        return QIPGPU_OK;
    }
    return enqueue_run(b);
}

int reset_batch(qipgpu_batch *b) {
    const uint32_t nb = (uint32_t)std::min<uint64_t>((b->arena_len + 255) / 256, 1u << 16);
    hipLaunchKernelGGL(k_batch_zero, dim3(nb), dim3(256), 0, b->stream, b->arena, b->arena_len);
    QIP_LAUNCHED();
    hipLaunchKernelGGL(k_batch_ones, dim3((b->count + 255) / 256), dim3(256), 0, b->stream, b->arena,
                       (const uint64_t *)b->d_offs, (uint32_t)b->count);
    QIP_LAUNCHED();
    return QIPGPU_OK;
}

int validate_ns(int32_t count, const int32_t *n_qubits) {
    if (count < 1) return set_error(QIPGPU_EINVAL, "a batch needs at least one state (count=%d)", count);
    if (!n_qubits) return set_error(QIPGPU_EINVAL, "null qubit counts");
    for (int32_t i = 0; i < count; ++i)
        if (n_qubits[i] < 1 || n_qubits[i] > kMaxBatchQubits)
            return set_error(QIPGPU_EINVAL, "state %d: %d qubits (1..%d)", i, n_qubits[i], kMaxBatchQubits);
    return QIPGPU_OK;
}

}  // namespace

extern "C" {

int qipgpu_batch_plan(int32_t count, const int32_t *n_qubits, int64_t n_gates, const int32_t *gate_state,
                      const int32_t *gate_k, const int32_t *gate_qubits, int32_t fusion_width, int32_t lds_max_qubits,
                      int32_t *gate_block, int32_t *state_blocks) {
    int rc = validate_ns(count, n_qubits);
    if (rc) return rc;
    Tuning t;
    t.fusion_width = fusion_width;
    t.lds_max_qubits = lds_max_qubits;
    if ((rc = check_tuning(t))) return rc;
    std::vector<int32_t> ns(n_qubits, n_qubits + count);
    GateList gl;
    if ((rc = validate_gates(count, ns, n_gates, gate_state, gate_k, gate_qubits, gl))) return rc;
    std::vector<std::vector<int64_t>> per_state(count);
    for (int64_t g = 0; g < n_gates; ++g) per_state[gl.state[g]].push_back(g);
    std::vector<int32_t> gb(n_gates, -1);
    for (int32_t i = 0; i < count; ++i) {
        std::vector<uint64_t> masks;
        if (ns[i] > lds_max_qubits) fuse_state(per_state[i], gl, fusion_width, masks, gb);
        if (state_blocks) state_blocks[i] = (int32_t)masks.size();
    }
    if (gate_block) std::copy(gb.begin(), gb.end(), gate_block);
    return QIPGPU_OK;
}

int qipgpu_batch_create(int32_t count, const int32_t *n_qubits, qipgpu_batch **out) {
    if (!out) return set_error(QIPGPU_EINVAL, "null output handle");
    *out = nullptr;
    int rc = validate_ns(count, n_qubits);
    if (rc) return rc;
    qipgpu_batch *b = new qipgpu_batch();
    b->count = count;
    b->ns.assign(n_qubits, n_qubits + count);
    uint64_t off = 0;
    for (int32_t i = 0; i < count; ++i) {
        b->offs.push_back(off);
        const uint64_t len = 1ull << b->ns[i];
        off += (len + kStateAlign - 1) / kStateAlign * kStateAlign;
    }
    b->arena_len = off;
    auto fail = [&](int code) {
        qipgpu_batch_destroy(b);
        return code;
    };
    if (hipMalloc((void **)&b->arena, sizeof(cx<float>) * off) != hipSuccess) {
        (void)hipGetLastError();
        return fail(set_error(QIPGPU_ENOMEM, "batch arena of %llu amplitudes", (unsigned long long)off));
    }
    if (hipMalloc((void **)&b->d_offs, sizeof(uint64_t) * count) != hipSuccess) {
        (void)hipGetLastError();
        return fail(set_error(QIPGPU_ENOMEM, "batch offsets"));
    }
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b->lds, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        return fail(set_error(QIPGPU_EHIP, "batch stream / event creation failed"));
    }
    if (hipMemcpy(b->d_offs, b->offs.data(), sizeof(uint64_t) * count, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        return fail(set_error(QIPGPU_EHIP, "batch offsets upload failed"));
    }
    if ((rc = reset_batch(b))) return fail(rc);
    if (hipStreamSynchronize(b->stream) != hipSuccess) {
        (void)hipGetLastError();
        return fail(set_error(QIPGPU_EHIP, "batch reset failed"));
    }
    *out = b;
    return QIPGPU_OK;
}

int qipgpu_batch_destroy(qipgpu_batch *b) {
    if (!b) return QIPGPU_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->side) (void)hipStreamSynchronize(b->side);
    if (b->lds) (void)hipStreamSynchronize(b->lds);
    free_plan(b);
    if (b->d_mats) (void)hipFree(b->d_mats);
    if (b->d_offs) (void)hipFree(b->d_offs);
    if (b->arena) (void)hipFree(b->arena);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->lds) (void)hipStreamDestroy(b->lds);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    (void)hipGetLastError();
    delete b;
    return QIPGPU_OK;
}

int qipgpu_batch_load(qipgpu_batch *b, int64_t n_gates, const int32_t *gate_state, const int32_t *gate_k,
                      const int32_t *gate_qubits, const double *mats, const qipgpu_batch_tuning *tuning) {
    if (!b) return set_error(QIPGPU_EINVAL, "null batch");
    const Tuning t = tuning_from(tuning);
    int rc = check_tuning(t);
    if (rc) return rc;
    GateList gl;
    if ((rc = validate_gates(b->count, b->ns, n_gates, gate_state, gate_k, gate_qubits, gl))) return rc;
    if (n_gates > 0 && !mats) return set_error(QIPGPU_EINVAL, "null matrices");
    if (gl.mat_total >= (1ull << 32)) return set_error(QIPGPU_EINVAL, "matrix data too large for one batch");
    std::vector<double> dev;
    dev.reserve(2 * gl.mat_total);
    std::vector<uint64_t> mat_dev(n_gates);
    for (int64_t g = 0; g < n_gates; ++g) {
        mat_dev[g] = dev.size() / 2;
        ascending_matrix(gl, g, b->ns[gl.state[g]], mats + 2 * gl.mat_in[g], dev);
    }
    if ((rc = sync_streams(b))) return rc;
    free_plan(b);
    if (b->d_mats) {
        (void)hipFree(b->d_mats);
        b->d_mats = nullptr;
    }
    b->loaded = false;
    if (!dev.empty()) {
        if (hipMalloc((void **)&b->d_mats, sizeof(double) * dev.size()) != hipSuccess) {
            (void)hipGetLastError();
            return set_error(QIPGPU_ENOMEM, "batch matrices (%zu bytes)", sizeof(double) * dev.size());
        }
        QIP_HIP(hipMemcpy(b->d_mats, dev.data(), sizeof(double) * dev.size(), hipMemcpyHostToDevice));
    }
    b->gl = std::move(gl);
    b->mat_dev = std::move(mat_dev);
    if ((rc = make_plan(b, t))) return rc;
    b->loaded = true;
    return QIPGPU_OK;
}

int qipgpu_batch_set_tuning(qipgpu_batch *b, const qipgpu_batch_tuning *tuning) {
    if (!b) return set_error(QIPGPU_EINVAL, "null batch");
    if (!b->loaded) return set_error(QIPGPU_EINVAL, "no gates loaded");
    const Tuning t = tuning_from(tuning);
    int rc = check_tuning(t);
    if (rc) return rc;
    b->loaded = false;
    if ((rc = make_plan(b, t))) return rc;
    b->loaded = true;
    return QIPGPU_OK;
}

int qipgpu_batch_get_tuning(qipgpu_batch *b, qipgpu_batch_tuning *out) {
    if (!b || !out) return set_error(QIPGPU_EINVAL, "null argument");
    out->fusion_width = b->tuning.fusion_width;
    out->lds_max_qubits = b->tuning.lds_max_qubits;
    out->chunk_tiles = b->tuning.chunk_tiles;
    out->use_graph = b->tuning.use_graph;
    return QIPGPU_OK;
}

int qipgpu_batch_reset(qipgpu_batch *b) {
    if (!b) return set_error(QIPGPU_EINVAL, "null batch");
    return reset_batch(b);
}

int qipgpu_batch_run(qipgpu_batch *b) {
    if (!b) return set_error(QIPGPU_EINVAL, "null batch");
    return run_batch(b);
}

int qipgpu_batch_synchronize(qipgpu_batch *b) {
    if (!b) return set_error(QIPGPU_EINVAL, "null batch");
    return sync_streams(b);
}

int qipgpu_batch_autotune(qipgpu_batch *b, int32_t reps, qipgpu_batch_tuning *chosen, double *best_ms) {
    if (!b) return set_error(QIPGPU_EINVAL, "null batch");
    if (!b->loaded) return set_error(QIPGPU_EINVAL, "no gates loaded");
    if (reps < 1 || reps > 100) return set_error(QIPGPU_EINVAL, "reps %d outside 1..100", reps);
    hipEvent_t e0, e1;
    QIP_HIP(hipEventCreate(&e0));
    QIP_HIP(hipEventCreate(&e1));
    int rc = QIPGPU_OK;
    auto time_tuning = [&](const Tuning &t, double *ms) -> int {
        int r = make_plan(b, t);
        if (r) return r;
        b->loaded = true;
        if ((r = reset_batch(b)) || (r = run_batch(b))) return r;
        QIP_HIP(hipStreamSynchronize(b->stream));
        float total = 0.0f;
        for (int i = 0; i < reps; ++i) {
            if ((r = reset_batch(b))) return r;
            QIP_HIP(hipEventRecord(e0, b->stream));
            if ((r = run_batch(b))) return r;
            QIP_HIP(hipEventRecord(e1, b->stream));
            QIP_HIP(hipEventSynchronize(e1));
            float f = 0.0f;
            QIP_HIP(hipEventElapsedTime(&f, e0, e1));
            total += f;
        }
        *ms = total / reps;
        return QIPGPU_OK;
    };
    Tuning best = b->tuning;
    double best_t = 0.0;
    rc = time_tuning(best, &best_t);
    const int chunks[] = {2, 4, 8, 16, 32};
    const int ldss[] = {11, 12, 13};
    const int widths[] = {5, 6};
    for (int c : chunks) {
        if (rc || c == best.chunk_tiles) continue;
        Tuning t = best;
        t.chunk_tiles = c;
        double ms = 0.0;
        rc = time_tuning(t, &ms);
        if (!rc && ms < best_t) best_t = ms, best = t;
    }
    for (int l : ldss) {
        if (rc || l == best.lds_max_qubits) continue;
        Tuning t = best;
        t.lds_max_qubits = l;
        double ms = 0.0;
        rc = time_tuning(t, &ms);
        if (!rc && ms < best_t) best_t = ms, best = t;
    }
    for (int w : widths) {
        if (rc || w == best.fusion_width) continue;
        Tuning t = best;
        t.fusion_width = w;
        double ms = 0.0;
        rc = time_tuning(t, &ms);
        if (!rc && ms < best_t) best_t = ms, best = t;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    b->loaded = false;
    if ((rc = make_plan(b, best))) return rc;
    b->loaded = true;
    if ((rc = reset_batch(b))) return rc;
    QIP_HIP(hipStreamSynchronize(b->stream));
    if (chosen) qipgpu_batch_get_tuning(b, chosen);
    if (best_ms) *best_ms = best_t;
    return QIPGPU_OK;
}

int qipgpu_batch_get_state(qipgpu_batch *b, int32_t i, void *out) {
    if (!b || !out) return set_error(QIPGPU_EINVAL, "null argument");
    if (i < 0 || i >= b->count) return set_error(QIPGPU_EINVAL, "state %d outside 0..%d", i, b->count - 1);
    int rc = sync_streams(b);
    if (rc) return rc;
    QIP_HIP(hipMemcpyAsync(out, b->arena + b->offs[i], sizeof(cx<float>) << b->ns[i], hipMemcpyDeviceToHost,
                           b->stream));
    QIP_HIP(hipStreamSynchronize(b->stream));
    return QIPGPU_OK;
}

int qipgpu_batch_set_state(qipgpu_batch *b, int32_t i, const void *in) {
    if (!b || !in) return set_error(QIPGPU_EINVAL, "null argument");
    if (i < 0 || i >= b->count) return set_error(QIPGPU_EINVAL, "state %d outside 0..%d", i, b->count - 1);
    int rc = sync_streams(b);
    if (rc) return rc;
    QIP_HIP(hipMemcpyAsync(b->arena + b->offs[i], in, sizeof(cx<float>) << b->ns[i], hipMemcpyHostToDevice,
                           b->stream));
    QIP_HIP(hipStreamSynchronize(b->stream));
    return QIPGPU_OK;
}

int qipgpu_batch_info(qipgpu_batch *b, int64_t *out, int32_t len) {
    if (!b || !out) return set_error(QIPGPU_EINVAL, "null argument");
    const Plan &p = b->plan;
    const int64_t v[] = {b->count,
                         p.n_small,
                         p.n_tiled,
                         (int64_t)b->gl.state.size(),
                         (int64_t)p.blocks.size(),
                         p.wave_block0.empty() ? 0 : (int64_t)p.wave_block0.size() - 1,
                         (int64_t)p.chunks.size(),
                         (int64_t)p.max_wave_blocks,
                         (int64_t)(b->arena_len * sizeof(cx<float>)),
                         (int64_t)p.algorithmic_bytes};
    const int32_t nv = (int32_t)(sizeof(v) / sizeof(v[0]));
    for (int32_t i = 0; i < len; ++i) out[i] = i < nv ? v[i] : 0;
    return QIPGPU_OK;
}

int qipgpu_batch_get_stream(qipgpu_batch *b, void **stream) {
    if (!b || !stream) return set_error(QIPGPU_EINVAL, "null argument");
    *stream = (void *)b->stream;
    return QIPGPU_OK;
}

}  // extern "C"
