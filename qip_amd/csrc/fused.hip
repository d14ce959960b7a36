// fused.hip -- precision form: a 64x64 complex operator (six fused qubits) applied to a complex64
// state on MFMA with split-BF16 compensated products, FP32 accumulation and FP64 norms. The tile,
// its data flow and the operator fragment layout are in tile6.h; this file holds the persistent
// single-state kernel, the FP64 renormalisation and the C entry points.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "engine.h"
#include "tile6.h"

using qipgpu::cx;
using namespace qipgpu;
using namespace qipgpu::tile;

namespace {

constexpr int kFusedBlocksPerCU = 2;   // 69 KB of LDS and 212 VGPRs per workgroup

// Sum `count` FP64 partials with one wavefront: lane l adds partials l, l + 64, ... in order, then a
// butterfly over the lanes. Each butterfly step adds the same two values on both lanes of a pair, so
// every lane, every wavefront and every workgroup that sums the same partials gets the same bits.
// The first 512 partials are loaded up front (partials_first) so a caller can issue other loads
// before the sum waits for these; partials_sum adds any beyond 512 and reduces.
__device__ __forceinline__ void partials_first(const double *__restrict__ partial, uint32_t count, int lane,
                                               double (&pv)[8]) {
    // unconditional loads (index clamped into the array, which holds at least one partial), so
    // none of them waits for another; partials_sum drops the clamped ones
    const uint32_t last = count ? count - 1 : 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) pv[j] = partial[min((uint32_t)lane + 64u * j, last)];
}

__device__ __forceinline__ double partials_sum(const double (&pv)[8], const double *__restrict__ partial,
                                               uint32_t count, int lane) {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j) acc += (uint32_t)lane + 64u * j < count ? pv[j] : 0.0;
    for (uint32_t i = (uint32_t)lane + 512u; i < count; i += 64) acc += partial[i];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    return acc;
}

// One fused block. Every amplitude is scaled as it is loaded by 1/sqrt(norm of the previous block's
// output): each workgroup sums the previous block's `prev_count` partials itself (the same sum in all
// of them; prev_count = 0 means no pending factor), so a block is one launch with no cross-workgroup
// synchronisation. partial[grid] receives this block's per-workgroup FP64 norm partials.
__global__ __launch_bounds__(256, 2) void k_fused6(cx<float> *amps, const bf16x8 *__restrict__ wfrag, FusedArgs a,
                                                   const double *__restrict__ prev_partial, uint32_t prev_count,
                                                   double *partial) {
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

    // Each workgroup takes a contiguous range of tiles (neighbouring tiles share DRAM pages more
    // often than tiles a grid apart). Two register buffers alternate: while one tile is processed
    // the next one's loads are in flight in the other buffer, with no copy of in-flight registers.
    // The prefetch is issued unconditionally (the last one re-reads an in-range tile) so the wait
    // before each tile covers only its own loads.
    const uint64_t per = (a.ntiles + gridDim.x - 1) / gridDim.x;
    uint64_t tile = (uint64_t)blockIdx.x * per;
    const uint64_t end = min(tile + per, a.ntiles);
    cx<float> va[16], vb[16];
    // the scale, per wavefront: the partials' loads go out first, so the sum waits for them while
    // the first tile's loads stay in flight
    double pv[8];
    partials_first(prev_partial, prev_count, lane, pv);
    // unconditional (a workgroup without tiles reads the last tile and never uses it), so the wait
    // before the sum covers the partials only on every path
    load_tile(va, amps, tile_base(min(tile, a.ntiles - 1), a) + off_t, a);
    ctx.scale = prev_count ? 1.0 / sqrt(partials_sum(pv, prev_partial, prev_count, lane)) : 1.0;
    if (tile < end) {
        const uint64_t last = end - 1;
        for (;;) {
            __builtin_amdgcn_s_setprio(2);
            load_tile(vb, amps, tile_base(min(tile + 1, last), a) + off_t, a);
            __builtin_amdgcn_s_setprio(0);
            process_tile(va, tile_base(tile, a) + off_t, amps, lds, ah, al, ctx, a, nacc);
            if (++tile >= end) break;
            __builtin_amdgcn_s_setprio(2);
            load_tile(va, amps, tile_base(min(tile + 1, last), a) + off_t, a);
            __builtin_amdgcn_s_setprio(0);
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

// fstate[1] = norm of the last block's output (its partials in the same fixed order as k_fused6),
// fstate[0] = 1 / sqrt(norm).
__global__ __launch_bounds__(64) void k_fused_norm(const double *__restrict__ partial, uint32_t count, double *fstate) {
    double pv[8];
    partials_first(partial, count, threadIdx.x, pv);
    const double nrm = partials_sum(pv, partial, count, threadIdx.x);
    if (threadIdx.x == 0) {
        fstate[1] = nrm;
        fstate[0] = 1.0 / sqrt(nrm);
    }
}

// The memory-pattern floor of a fused block: every amplitude is loaded and stored back in place
// with exactly k_fused6's tile pattern, grid, tile ranges and double buffering, and no arithmetic, no
// LDS. Its time against k_fused6's on the same targets separates what the tile's DRAM pattern costs
// from what the kernel's own work adds.
__global__ __launch_bounds__(256, 2) void k_tile_sweep(cx<float> *amps, FusedArgs a) {
    typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f4;
    const int tid = threadIdx.x;
    const uint64_t off_t = u_offset<0, 9>(2u * (uint32_t)tid, a);
    const uint64_t per = (a.ntiles + gridDim.x - 1) / gridDim.x;
    uint64_t tile = (uint64_t)blockIdx.x * per;
    const uint64_t end = min(tile + per, a.ntiles);
    if (tile >= end) return;
    const uint64_t last = end - 1;
    auto store = [&](const cx<float> (&v)[16], uint64_t base) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            f4 p;
            p[0] = v[2 * i].re;
            p[1] = v[2 * i].im;
            p[2] = v[2 * i + 1].re;
            p[3] = v[2 * i + 1].im;
            *reinterpret_cast<f4 *>(amps + base + u_offset<9, 12>(512u * i, a)) = p;
        }
    };
    cx<float> va[16], vb[16];
    load_tile(va, amps, tile_base(tile, a) + off_t, a);
    for (;;) {
        load_tile(vb, amps, tile_base(min(tile + 1, last), a) + off_t, a);
        store(va, tile_base(tile, a) + off_t);
        if (++tile >= end) break;
        load_tile(va, amps, tile_base(min(tile + 1, last), a) + off_t, a);
        store(vb, tile_base(tile, a) + off_t);
        if (++tile >= end) break;
    }
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

double *fused_partials_of(qipgpu_state *s, int parity) {
    return s->d_fused + 4 + (size_t)parity * s->fused_partials;
}

// d_fused: [0] scale and [1] norm written by k_fused_norm, then two partial arrays of `cap` doubles
// that alternate between consecutive blocks (a block reads the other array while it writes its own).
// A state's fused grid depends only on its size and the device, so the arrays are sized once.
int ensure_fused_state(qipgpu_state *s, uint32_t nblocks) {
    if (s->d_fused) {
        if (s->fused_partials >= nblocks) return QIPGPU_OK;
        return set_error(QIPGPU_EHIP, "fused grid of %u workgroups exceeds the %zu partials sized for this state",
                         nblocks, s->fused_partials);
    }
    QIP_HIP(hipMalloc(&s->d_fused, sizeof(double) * (4 + 2 * (size_t)nblocks)));
    s->fused_partials = nblocks;
    return QIPGPU_OK;
}

__attribute__((always_inline)) inline int compose_impl(int32_t k, const int32_t *qubits, int32_t n_gates,
                                                      const int32_t *gate_k, const int32_t *gate_qubits,
                                                      const double *mats, double *out) {
    if (k < 1 || k > 6) return set_error(QIPGPU_EINVAL, "block of %d qubits (1..6)", k);
    if (n_gates < 0) return set_error(QIPGPU_EINVAL, "negative gate count");
    if (!qubits || !out || (n_gates > 0 && (!gate_k || !gate_qubits || !mats)))
        return set_error(QIPGPU_EINVAL, "null argument");
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < i; ++j)
            if (qubits[i] == qubits[j]) return set_error(QIPGPU_EINVAL, "block qubit %d repeated", qubits[i]);
    // bit position of each gate qubit in the block index (block qubit i is bit k-1-i)
    std::vector<int8_t> gbits((size_t)n_gates * 6);
    for (int32_t g = 0; g < n_gates; ++g) {
        const int kg = gate_k[g];
        if (kg < 1 || kg > k) return set_error(QIPGPU_EINVAL, "gate %d on %d qubits of a %d-qubit block", g, kg, k);
        uint32_t mask = 0;
        for (int j = 0; j < kg; ++j) {
            const int q = gate_qubits[6 * g + j];
            int pos = -1;
            for (int i = 0; i < k; ++i)
                if (qubits[i] == q) pos = i;
            if (pos < 0) return set_error(QIPGPU_EINVAL, "gate %d: qubit %d is not a block qubit", g, q);
            if (mask >> (k - 1 - pos) & 1) return set_error(QIPGPU_EINVAL, "gate %d: qubit %d repeated", g, q);
            mask |= 1u << (k - 1 - pos);
            gbits[6 * g + j] = (int8_t)(k - 1 - pos);
        }
    }
    // Single-qubit gates commute with every gate on other qubits, so each one waits on its bit (the
    // product of the waiting ones kept as a 2x2) until a wider gate touches that bit, where it is
    // folded into that gate's small matrix; the rest meet the product at the end. While the product is
    // still the identity, a gate is written into it directly and the final single-qubit gates as
    // their Kronecker product, so a block of single-qubit gates costs D^2 products and not k 2 D^2.
    const int D = 1 << k;
    std::vector<double> mre(D * D, 0.0), mim(D * D, 0.0), tre(D * D), tim(D * D);
    bool ident = true;
    double pw[6][8];      // waiting 2x2 per bit, interleaved complex row-major
    bool has[6] = {false, false, false, false, false, false};
    double gre[4096], gim[4096];   // a wider gate with its waiting single-qubit gates folded in
    size_t moff = 0;
    for (int32_t g = 0; g < n_gates; ++g) {
        const int kg = gate_k[g], dg = 1 << kg;
        const double *G = mats + 2 * moff;   // interleaved complex, row-major dg x dg
        moff += (size_t)dg * dg;
        const int8_t *bits = &gbits[6 * g];
        if (kg == 1) {
            double *P = pw[bits[0]];
            if (!has[bits[0]]) {
                std::memcpy(P, G, sizeof(double) * 8);
                has[bits[0]] = true;
            } else {   // P = G P
                double r[8];
                for (int i = 0; i < 2; ++i)
                    for (int j = 0; j < 2; ++j) {
                        double re = 0.0, im = 0.0;
                        for (int e = 0; e < 2; ++e) {
                            const double ar = G[2 * (2 * i + e)], ai = G[2 * (2 * i + e) + 1];
                            const double br = P[2 * (2 * e + j)], bi = P[2 * (2 * e + j) + 1];
                            re += ar * br - ai * bi;
                            im += ar * bi + ai * br;
                        }
                        r[2 * (2 * i + j)] = re;
                        r[2 * (2 * i + j) + 1] = im;
                    }
                std::memcpy(P, r, sizeof(r));
            }
            continue;
        }
        for (int i = 0; i < dg * dg; ++i) {
            gre[i] = G[2 * i];
            gim[i] = G[2 * i + 1];
        }
        uint32_t mask = 0;
        for (int j = 0; j < kg; ++j) {
            mask |= 1u << bits[j];
            if (!has[bits[j]]) continue;
            // G <- G (I x P x I) with P on gate index bit kg-1-j: mixes column pairs c0 / c1
            const double *P = pw[bits[j]];
            const int cb = 1 << (kg - 1 - j);
            for (int r = 0; r < dg; ++r)
                for (int c0 = 0; c0 < dg; ++c0) {
                    if (c0 & cb) continue;
                    const int c1 = c0 | cb;
                    const double ar = gre[r * dg + c0], ai = gim[r * dg + c0];
                    const double br = gre[r * dg + c1], bi = gim[r * dg + c1];
                    gre[r * dg + c0] = ar * P[0] - ai * P[1] + br * P[4] - bi * P[5];
                    gim[r * dg + c0] = ar * P[1] + ai * P[0] + br * P[5] + bi * P[4];
                    gre[r * dg + c1] = ar * P[2] - ai * P[3] + br * P[6] - bi * P[7];
                    gim[r * dg + c1] = ar * P[3] + ai * P[2] + br * P[7] + bi * P[6];
                }
            has[bits[j]] = false;
        }
        int dep[64];   // gate index e -> its bits placed in the block index
        for (int e = 0; e < dg; ++e) {
            int d = 0;
            for (int j = 0; j < kg; ++j)
                if ((e >> (kg - 1 - j)) & 1) d |= 1 << bits[j];
            dep[e] = d;
        }
        auto sub = [&](int x) {
            int v = 0;
            for (int j = 0; j < kg; ++j) v = (v << 1) | ((x >> bits[j]) & 1);
            return v;
        };
        if (ident) {   // product = this gate embedded in the block
            for (int r = 0; r < D; ++r) {
                const int sr = sub(r), base = r & ~(int)mask;
                for (int e = 0; e < dg; ++e) {
                    mre[r * D + (base | dep[e])] = gre[sr * dg + e];
                    mim[r * D + (base | dep[e])] = gim[sr * dg + e];
                }
            }
            ident = false;
            continue;
        }
        // row r of the product takes gate row sub(r) against the rows base(r) | dep[e] of m
        for (int r = 0; r < D; ++r) {
            const int sr = sub(r), base = r & ~(int)mask;
            double *__restrict__ ore = &tre[r * D], *__restrict__ oim = &tim[r * D];
            for (int c = 0; c < D; ++c) ore[c] = oim[c] = 0.0;
            for (int e = 0; e < dg; ++e) {
                const double gr = gre[sr * dg + e], gi = gim[sr * dg + e];
                if (gr == 0.0 && gi == 0.0) continue;
                const double *are = &mre[(base | dep[e]) * D], *aim = &mim[(base | dep[e]) * D];
                for (int c = 0; c < D; ++c) {
                    ore[c] += gr * are[c] - gi * aim[c];
                    oim[c] += gr * aim[c] + gi * are[c];
                }
            }
        }
        mre.swap(tre);
        mim.swap(tim);
    }
    if (ident) {   // Kronecker product of the waiting single-qubit gates, MSB first, by doubling
        mre[0] = 1.0;
        mim[0] = 0.0;
        for (int b = k - 1, d = 1; b >= 0; --b, d *= 2) {
            double P[8] = {1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0};
            if (has[b]) std::memcpy(P, pw[b], sizeof(P));
            const int d2 = 2 * d;
            for (int r = 0; r < d; ++r)
                for (int c = 0; c < d; ++c) {
                    const double xr = mre[r * d + c], xi = mim[r * d + c];
                    for (int i = 0; i < 2; ++i)
                        for (int j = 0; j < 2; ++j) {
                            const double pr = P[2 * (2 * i + j)], pi = P[2 * (2 * i + j) + 1];
                            tre[(2 * r + i) * d2 + 2 * c + j] = xr * pr - xi * pi;
                            tim[(2 * r + i) * d2 + 2 * c + j] = xr * pi + xi * pr;
                        }
                }
            mre.swap(tre);
            mim.swap(tim);
        }
    } else {
        for (int b = 0; b < k; ++b) {   // product <- P_b product: mixes row pairs r0 / r1
            if (!has[b]) continue;
            const double *P = pw[b];
            for (int r0 = 0; r0 < D; ++r0) {
                if (r0 >> b & 1) continue;
                const int r1 = r0 | (1 << b);
                double *__restrict__ a_re = &mre[r0 * D], *__restrict__ a_im = &mim[r0 * D];
                double *__restrict__ b_re = &mre[r1 * D], *__restrict__ b_im = &mim[r1 * D];
                for (int c = 0; c < D; ++c) {
                    const double ar = a_re[c], ai = a_im[c], br = b_re[c], bi = b_im[c];
                    a_re[c] = P[0] * ar - P[1] * ai + P[2] * br - P[3] * bi;
                    a_im[c] = P[0] * ai + P[1] * ar + P[2] * bi + P[3] * br;
                    b_re[c] = P[4] * ar - P[5] * ai + P[6] * br - P[7] * bi;
                    b_im[c] = P[4] * ai + P[5] * ar + P[6] * bi + P[7] * br;
                }
            }
        }
    }
    for (int i = 0; i < D * D; ++i) {
        out[2 * i] = mre[i];
        out[2 * i + 1] = mim[i];
    }
    return QIPGPU_OK;
}

// The composition is host work on the launch path of every block: where the host CPU has AVX2 and FMA
// it runs a copy built for them (the same loops, wider vectors), else the baseline build.
__attribute__((target("avx2,fma"))) int compose_avx2(int32_t k, const int32_t *qubits, int32_t n_gates,
                                                     const int32_t *gate_k, const int32_t *gate_qubits,
                                                     const double *mats, double *out) {
    return compose_impl(k, qubits, n_gates, gate_k, gate_qubits, mats, out);
}

int compose_base(int32_t k, const int32_t *qubits, int32_t n_gates, const int32_t *gate_k, const int32_t *gate_qubits,
                 const double *mats, double *out) {
    return compose_impl(k, qubits, n_gates, gate_k, gate_qubits, mats, out);
}

uint32_t fused_grid(uint64_t ntiles) {
    return (uint32_t)std::min<uint64_t>(ntiles, (uint64_t)device_cus() * kFusedBlocksPerCU);
}

}  // namespace

namespace qipgpu {

// Launch recorded gates and apply a pending renormalisation factor before the state is read or written
// from the host. Gate,
// diagonal, product and func_apply sweeps are linear, so they commute with the scalar and leave it
// pending for the next fused block (which applies it as it loads) or the next read.
int settle(qipgpu_state *s) {
    return qipgpu_fused_finish(s);
}

}  // namespace qipgpu

extern "C" {

int qipgpu_fused_apply(qipgpu_state *s, int32_t k, const int32_t *targets, const double *mat) {
    QIP_NOT_CAPTURING(s, "a fused block");
    if (s->dtype != QIPGPU_C64) return set_error(QIPGPU_EINVAL, "fused blocks apply to complex64 states");
    if (s->n < kTileBits) return set_error(QIPGPU_EUNSUP, "fused blocks need n >= %d (n=%d)", kTileBits, s->n);
    if (k < 1 || k > 6) return set_error(QIPGPU_EINVAL, "fused block on %d qubits (1..6)", k);
    int rc = check_qubits(s, k, targets, "fused block qubits");
    if (rc) return rc;
    rc = flush_recorded(s);   // recorded gates go first (a no-op when the recorder is launching)
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
    int cidx[64], pad[64];
    for (int r = 0; r < 64; ++r) {
        int ci = 0, pv = 0;
        for (int j = 0; j < k; ++j) ci = (ci << 1) | ((r >> rowbit_of_tbit[tbits[j]]) & 1);
        for (int j = k; j < 6; ++j) pv = (pv << 1) | ((r >> rowbit_of_tbit[tbits[j]]) & 1);
        cidx[r] = ci;
        pad[r] = pv;
    }
    const int D = 1 << k;
    std::vector<double> ur(64 * 64), ui(64 * 64);
    for (int r = 0; r < 64; ++r)
        for (int c = 0; c < 64; ++c) {
            const bool same = pad[r] == pad[c];
            ur[r * 64 + c] = same ? mat[2 * (cidx[r] * D + cidx[c])] : 0.0;
            ui[r * 64 + c] = same ? mat[2 * (cidx[r] * D + cidx[c]) + 1] : 0.0;
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
    hipLaunchKernelGGL(k_fused6, dim3(grid), dim3(256), 0, s->stream, (cx<float> *)s->amps, (const bf16x8 *)dfrag, a,
                       (const double *)fused_partials_of(s, s->fused_parity ^ 1), s->fused_prev_count,
                       fused_partials_of(s, s->fused_parity));
    QIP_LAUNCHED();
    s->fused_prev_count = grid;
    s->fused_parity ^= 1;
    return QIPGPU_OK;
}

int qipgpu_fused_pattern_sweep(qipgpu_state *s, int32_t k, const int32_t *targets) {
    QIP_NOT_CAPTURING(s, "a fused pattern sweep");
    if (s->dtype != QIPGPU_C64) return set_error(QIPGPU_EINVAL, "fused blocks apply to complex64 states");
    if (s->n < kTileBits) return set_error(QIPGPU_EUNSUP, "fused blocks need n >= %d (n=%d)", kTileBits, s->n);
    if (k < 1 || k > 6) return set_error(QIPGPU_EINVAL, "fused block on %d qubits (1..6)", k);
    int rc = check_qubits(s, k, targets, "fused block qubits");
    if (rc) return rc;
    rc = flush_recorded(s);
    if (rc) return rc;
    uint64_t tmask = 0;
    for (int j = 0; j < k; ++j) tmask |= 1ull << (s->n - 1 - targets[j]);
    int ts[6], ls[6];
    choose_tile_bits(s->n, tmask, ts, ls);
    const FusedArgs a = make_tile_args(s->n, ts, ls);
    hipLaunchKernelGGL(k_tile_sweep, dim3(fused_grid(a.ntiles)), dim3(256), 0, s->stream, (cx<float> *)s->amps, a);
    QIP_LAUNCHED();
    return QIPGPU_OK;
}

int qipgpu_compose(int32_t k, const int32_t *qubits, int32_t n_gates, const int32_t *gate_k,
                   const int32_t *gate_qubits, const double *mats, double *out) {
    static const bool avx2 = __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
    return avx2 ? compose_avx2(k, qubits, n_gates, gate_k, gate_qubits, mats, out)
                : compose_base(k, qubits, n_gates, gate_k, gate_qubits, mats, out);
}

int qipgpu_fused_finish(qipgpu_state *s) {
    const int rc = flush_recorded(s);
    if (rc) return rc;
    if (!s->fused_prev_count) return QIPGPU_OK;
    QIP_NOT_CAPTURING(s, "a pending fused renormalisation");
    hipLaunchKernelGGL(k_fused_norm, dim3(1), dim3(64), 0, s->stream,
                       (const double *)fused_partials_of(s, s->fused_parity ^ 1), s->fused_prev_count, s->d_fused);
    QIP_LAUNCHED();
    s->fused_prev_count = 0;
    const uint32_t nb = (uint32_t)std::min<uint64_t>((s->len + 1023) / 1024, 1u << 20);
    hipLaunchKernelGGL(k_scale<float>, dim3(nb), dim3(256), 0, s->stream, (cx<float> *)s->amps, s->len,
                       (const double *)s->d_fused);
    QIP_LAUNCHED();
    return QIPGPU_OK;
}

int qipgpu_fused_last_norm(qipgpu_state *s, double *out) {
    if (!s->d_fused) return set_error(QIPGPU_EINVAL, "no fused block has been applied to this state");
    if (s->fused_prev_count) {   // still pending: sum the last block's partials now
        hipLaunchKernelGGL(k_fused_norm, dim3(1), dim3(64), 0, s->stream,
                           (const double *)fused_partials_of(s, s->fused_parity ^ 1), s->fused_prev_count, s->d_fused);
        QIP_LAUNCHED();
    }
    QIP_HIP(hipMemcpyAsync(out, s->d_fused + 1, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    QIP_HIP(hipStreamSynchronize(s->stream));
    return QIPGPU_OK;
}

}  // extern "C"
