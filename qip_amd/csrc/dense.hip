// dense.hip -- dense operators on 4..12 target qubits, and any kronselect_dot descriptor component
// (qip/ext/kronprod.pyx:41-235) that no specialised sweep covers.
//
// Terms: a dense operator U (D x D, D = 2^K, row-major, first target = MSB of the row index) acts
// on the state's "bases": a base is a state index with every target bit 0 and every control bit 1;
// the D amplitudes base + off[j] form one closed group, and out[base + off[i]] =
// sum_j U[i][j] * in[base + off[j]]. That is the matrix product Y = U X with X[j][b] the amplitudes
// of base b. FP64 arithmetic for both state precisions, rounded once on store.
//
//   k_dense_mfma<R, K>   K = 4..6, in place. A persistent workgroup takes tiles of B bases: it stages
//                        X (D x B = 2048 amplitudes, 32 KB in FP64) in LDS as Re / Im planes, forms
//                        Y = U X with three real products on FP64 MFMA (v_mfma_f64_16x16x4f64), each
//                        wave holding its share of Re U and Im U in registers for the launch, and
//                        writes Y back through the same planes. A tile holds every amplitude of its
//                        bases before any is written, so in place is safe.
//   k_dense_gemm<R>      K = 7..12, out of place: a 64 x 64 output tile per workgroup over 16-column
//                        steps of j through LDS, into a compact [D][nbase] buffer; k_dense_scatter
//                        writes it back to the bases' positions.
//   k_kron_general<R>    the reference's own formula, out of place: for every output row r, the sum
//                        over the 2^u columns c that agree with r outside the union of the groups'
//                        qubits, in increasing order, of prod_g entry_g(r|_g, c|_g) * in[c] (the
//                        product stops at a zero factor), with entry_g following C_MAT levels and a
//                        NUMPY or SWAP_MAT inner matrix (calc_mat_entry, kronprod.pyx:203-235).
#include "engine.h"

#include <algorithm>
#include <complex>
#include <cstring>
#include <vector>

using namespace qipgpu;

namespace {

template <typename R>
__device__ __forceinline__ cx<double> to_d(cx<R> v) {
    cx<double> o;
    o.re = (double)v.re;
    o.im = (double)v.im;
    return o;
}

template <typename R>
__device__ __forceinline__ cx<R> from_d(cx<double> v) {
    cx<R> o;
    o.re = (R)v.re;
    o.im = (R)v.im;
    return o;
}

__device__ __forceinline__ void mac(cx<double> &acc, const cx<double> a, const cx<double> b) {
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(-a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(a.im, b.re, acc.im);
}

__device__ __forceinline__ uint64_t base_of(uint64_t g, const PosList &pl, uint64_t ctrl_mask) {
    return insert_zeros(g, pl) | ctrl_mask;
}

constexpr int kTileThreads = 256;

// ---- K = 4..6: tiles in place on FP64 MFMA -------------------------------------------------------
//
// Y = U X with U = A + i Bm and X = P + i Q takes three real D x D products (the 3M method, Higham,
// "Stability of a method for multiplying complex matrices with three real matrix multiplications",
// SIAM J. Matrix Anal. Appl. 13 (1992)): T1 = A P, T2 = Bm Q, T3 = (A + Bm)(P + Q), then
// Re Y = T1 - T2 and Im Y = T3 - T1 - T2. That is 6 D^2 real FMAs per base instead of the 8 of the
// real 2D x 2D form. Re Y rounds as the usual form does; the rounding error of Im Y is bounded in
// terms of (|A| + |Bm|)(|P| + |Q|) instead of |A||Q| + |Bm||P|, at most a small constant factor more
// in FP64 (the paper's bound).
// v_mfma_f64_16x16x4f64 computes D[16 x 16] += A[16 x 4] B[4 x 16]; lane l holds A[l % 16][l / 16]
// and B[l / 16][l % 16], and D[4 i + l / 16][l % 16] for its items i = 0..3, so the three products of
// one output block land in the same lanes and are combined in registers.
//
// Wave w owns output row block w % RB (RB = D / 16 blocks of 16 complex rows) and CW consecutive
// 16-base column blocks starting at (w / RB) CW; its fragments of A and Bm (2 x D / 4 doubles per
// lane) are gathered from U once per launch and stay in registers, and A + Bm is formed per k-step.
// X goes through LDS as two FP64 planes (Re, Im) with one slot of padding per row (or per base, when
// lanes run along the rows), and so does Y on its way out, so loads and stores stay coalesced. Per
// tile: 3 D x D x B FMAs in 3 x (D / 4) x CW MFMAs per wave.
template <int K>
struct MfmaShape {
    static constexpr int D = 1 << K;
    // 2048 amplitudes per tile: 34 KB of LDS and at most 168 VGPRs, so three workgroups share a CU
    // (three waves per SIMD); with 4096 (67 KB) only two fit, which left the MFMA units idle while
    // a workgroup loads or stores
    static constexpr int TE = 2048;
    static constexpr int WGS = 3;                // workgroups per CU
    static constexpr int B = TE / D;             // bases per tile
    static constexpr int LOADS = TE / kTileThreads;   // amplitudes per thread
    static constexpr int RB = D / 16;            // 16-row output blocks
    static constexpr int CB = B / 16;            // 16-column blocks
    static constexpr int CW = CB * RB / 4;       // column blocks per wave (4 waves)
    static constexpr int KS = D / 4;             // 4-column k-steps over the D complex columns
    static constexpr int XS = B + 1;             // padded plane row, row-major planes (doubles)
    static constexpr int XSJ = D + 1;            // padded plane column, base-major planes
    static constexpr int PLANE = D * B + (D > B ? D : B);
    static_assert(K >= 4 && K <= 6, "MFMA tiles for 4 to 6 targets");
    static_assert(CW * 4 == CB * RB, "column blocks split evenly over the waves");
};

typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename R, int K, bool ROWS_FAST>
__global__ __launch_bounds__(kTileThreads, MfmaShape<K>::WGS) void k_dense_mfma(cx<R> *amps, uint64_t nbase, uint64_t ctrl_mask,
                                                                PosList pl, const uint64_t *__restrict__ off_g,
                                                                const cx<double> *__restrict__ U_g, uint64_t ntiles) {
    using M = MfmaShape<K>;
    constexpr int D = M::D, B = M::B, XS = M::XS;
    // Re and Im planes of the tile: row-major (row * XS + base) when lanes run along the bases,
    // base-major (base * XSJ + row) when they run along the rows, so that consecutive lanes touch
    // consecutive slots in the loads and stores and the MFMA operand reads spread over the banks
    constexpr int XSJ = M::XSJ;
    __shared__ double sX[2][M::PLANE];
    auto idx = [](int j, int b) { return ROWS_FAST ? b * XSJ + j : j * XS + b; };
    __shared__ uint64_t sOff[D];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    for (int e = tid; e < D; e += kTileThreads) sOff[e] = off_g[e];

    // this wave's fragments: lane -> output row rho = lane % 16 of its block, k = lane / 16 of each
    // k-step; afr[0] = Re U, afr[1] = Im U (Re U + Im U is formed per k-step)
    const int rho = lane & 15, kap = lane >> 4;
    const int row0 = 16 * (wave % M::RB);        // first output row of the wave's block
    const int col0 = 16 * M::CW * (wave / M::RB); // first base of the wave's column blocks
    double afr[2][M::KS];
#pragma unroll
    for (int s = 0; s < M::KS; ++s) {
        const cx<double> u = U_g[(row0 + rho) * D + 4 * s + kap];
        afr[0][s] = u.re;
        afr[1][s] = u.im;
    }
    __syncthreads();

    // loads and stores: element i of this thread is (row j, base b); lanes run along the bases, or
    // along the rows when ROWS_FAST
    constexpr int BSTEP = kTileThreads / D;
    constexpr int JSTEP = kTileThreads / B;
    const int b_ld = ROWS_FAST ? tid / D : tid % B;
    const int j_ld = ROWS_FAST ? tid % D : tid / B;
    cx<R> pre[M::LOADS];
    // position of base t B + v (v < B; a base past the end reads the last one, whose products are
    // not stored)
    auto base_at = [&](uint64_t t, int v) {
        const uint64_t g = t * B + v;
        return base_of(g < nbase ? g : nbase - 1, pl, ctrl_mask);
    };
    auto load_tile = [&](uint64_t t) {
        if constexpr (ROWS_FAST) {
#pragma unroll
            for (int i = 0; i < M::LOADS; ++i) pre[i] = amps[base_at(t, b_ld + BSTEP * i) + sOff[j_ld]];
        } else {
            const uint64_t base = base_at(t, b_ld);
#pragma unroll
            for (int i = 0; i < M::LOADS; ++i) pre[i] = amps[base + sOff[j_ld + i * JSTEP]];
        }
    };
    // plane slot of element i of this thread
    auto slot = [&](int i) { return ROWS_FAST ? idx(j_ld, b_ld + BSTEP * i) : idx(j_ld + i * JSTEP, b_ld); };

    // the tile's amplitudes (pre) into LDS as Re and Im planes
    auto write_x = [&]() {
#pragma unroll
        for (int i = 0; i < M::LOADS; ++i) {
            const cx<double> v = to_d(pre[i]);
            sX[0][slot(i)] = v.re;
            sX[1][slot(i)] = v.im;
        }
    };
    f64x4 acc[3][M::CW];
    // the three products of the tile in LDS: lanes read P and Q of row 4 s + kap, base col0 + 16 c + rho
    auto products = [&]() {
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int c = 0; c < M::CW; ++c) acc[m][c] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < M::KS; ++s) {
            // the empty asm makes the copies opaque, so the sum below stays in the tile loop instead
            // of being hoisted out of it as a third set of fragment registers
            double a0 = afr[0][s], a1 = afr[1][s];
            asm volatile("" : "+v"(a0), "+v"(a1));
            const double a2 = a0 + a1;
            const int at = idx(4 * s + kap, col0 + rho);
            double p[M::CW], q[M::CW];
#pragma unroll
            for (int c = 0; c < M::CW; ++c) {
                p[c] = sX[0][at + idx(0, 16 * c)];
                q[c] = sX[1][at + idx(0, 16 * c)];
            }
#pragma unroll
            for (int c = 0; c < M::CW; ++c) {
                acc[0][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, p[c], acc[0][c], 0, 0, 0);
                acc[1][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, q[c], acc[1][c], 0, 0, 0);
                acc[2][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, p[c] + q[c], acc[2][c], 0, 0, 0);
            }
        }
    };
    // item i of a block is output row row0 + 4 i + kap, base col0 + 16 c + rho
    auto write_y = [&]() {
#pragma unroll
        for (int c = 0; c < M::CW; ++c)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int at = idx(row0 + 4 * i + kap, col0 + 16 * c + rho);
                const double t1 = acc[0][c][i], t2 = acc[1][c][i];
                sX[0][at] = t1 - t2;
                sX[1][at] = acc[2][c][i] - t1 - t2;
            }
    };
    // the tile's outputs from LDS to the state, with the load mapping
    auto store_y = [&](uint64_t t) {
#pragma unroll
        for (int i = 0; i < M::LOADS; ++i) {
            const int v = ROWS_FAST ? b_ld + BSTEP * i : b_ld;
            if (t * B + v < nbase) {
                cx<double> y;
                y.re = sX[0][slot(i)];
                y.im = sX[1][slot(i)];
                const int j = ROWS_FAST ? j_ld : j_ld + i * JSTEP;
                amps[base_at(t, v) + sOff[j]] = from_d<R>(y);
            }
        }
    };

    const uint64_t G = gridDim.x;
    uint64_t t = blockIdx.x;
    if (t >= ntiles) return;                   // the whole workgroup: no barrier is skipped by part of it
    // the next tile's loads in registers while the products run, where they fit beside the
    // fragments and accumulators at three workgroups per CU (all but K = 6 with lanes along the rows,
    // whose extra address registers would spill); otherwise the other workgroups on the CU cover
    // this one's loads
    constexpr bool kPrefetch = K <= 5 || !ROWS_FAST;
    if constexpr (kPrefetch) {
        load_tile(t);
        for (; t < ntiles; t += G) {
            __syncthreads();                   // the previous tile's outputs have been read back
            write_x();
            __syncthreads();
            if (t + G < ntiles) load_tile(t + G);   // in flight during the products
            products();
            __syncthreads();                   // every wave has read X
            write_y();
            __syncthreads();
            store_y(t);
        }
    } else {
        // no prefetch registers: the other workgroup on the CU covers this one's loads
        for (; t < ntiles; t += G) {
            __syncthreads();
            load_tile(t);
            write_x();
            __syncthreads();
            products();
            __syncthreads();
            write_y();
            __syncthreads();
            store_y(t);
        }
    }
}

// ---- K = 7..12: tiled GEMM out of place ---------------------------------------------------------
//
// A workgroup computes a 64-row x 64-base block of Y = U X into a compact [D][nbase] buffer, over
// the D input rows in steps of 16 through LDS (Re / Im planes of the U block and of the X block,
// padded by one slot per row). On FP64 MFMA with the same three real products per complex one as
// the tiles above: wave w owns output rows 16 w .. 16 w + 15 and all four 16-base column blocks,
// 12 accumulators; per step it reads its A fragments (Re U, Im U; Re U + Im U formed in registers)
// and the B fragments of X from LDS. Bases past nbase read the last base and are not stored.

constexpr int kGemmTile = 64;
constexpr int kGemmStep = 16;

template <typename R>
__global__ __launch_bounds__(256, 2) void k_dense_gemm(const cx<R> *__restrict__ amps, cx<R> *__restrict__ out,
                                                       uint64_t nbase, int D, uint64_t ctrl_mask, PosList pl,
                                                       const uint64_t *__restrict__ off,
                                                       const cx<double> *__restrict__ U) {
    constexpr int US = kGemmStep + 1;          // padded U row (doubles)
    constexpr int XS = kGemmTile + 1;          // padded X row
    __shared__ double sU[2][kGemmTile * US];
    __shared__ double sX[2][kGemmStep * XS];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int rho = lane & 15, kap = lane >> 4;
    const uint64_t b0 = (uint64_t)blockIdx.x * kGemmTile;   // first base of the block
    const int i0 = blockIdx.y * kGemmTile;                    // first output row of the block
    // loader roles: X row jj0 + 4 s of base b_ld; U row u_row, columns u_col + 4 s
    const int b_ld = tid % kGemmTile;
    const int jj0 = tid / kGemmTile;
    const uint64_t gb = b0 + b_ld;
    const uint64_t base = base_of(gb < nbase ? gb : nbase - 1, pl, ctrl_mask);
    const int u_row = tid / 4, u_col = tid % 4;
    f64x4 acc[3][4];
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[m][c] = f64x4{0.0, 0.0, 0.0, 0.0};
    for (int jc = 0; jc < D; jc += kGemmStep) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int jj = jj0 + 4 * s;
            const cx<double> v = to_d(amps[base + off[jc + jj]]);
            sX[0][jj * XS + b_ld] = v.re;
            sX[1][jj * XS + b_ld] = v.im;
            const cx<double> u = U[(size_t)(i0 + u_row) * D + jc + u_col + 4 * s];
            sU[0][u_row * US + u_col + 4 * s] = u.re;
            sU[1][u_row * US + u_col + 4 * s] = u.im;
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < kGemmStep / 4; ++s) {
            const int ua = (16 * wave + rho) * US + 4 * s + kap;
            const double a0 = sU[0][ua], a1 = sU[1][ua];
            const double a2 = a0 + a1;
            const int xa = (4 * s + kap) * XS + rho;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const double p = sX[0][xa + 16 * c], q = sX[1][xa + 16 * c];
                acc[0][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, p, acc[0][c], 0, 0, 0);
                acc[1][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, q, acc[1][c], 0, 0, 0);
                acc[2][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, p + q, acc[2][c], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // item i of column block c: output row i0 + 16 wave + 4 i + kap, base b0 + 16 c + rho
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const uint64_t b = b0 + 16 * c + rho;
        if (b < nbase) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const double t1 = acc[0][c][i], t2 = acc[1][c][i];
                cx<double> y;
                y.re = t1 - t2;
                y.im = acc[2][c][i] - t1 - t2;
                out[(size_t)(i0 + 16 * wave + 4 * i + kap) * nbase + b] = from_d<R>(y);
            }
        }
    }
}

// out[i][b] (compact) -> amps[base(b) + off[i]]
template <typename R>
__global__ __launch_bounds__(256) void k_dense_scatter(cx<R> *amps, const cx<R> *__restrict__ out, uint64_t nbase,
                                                       int D, uint64_t ctrl_mask, PosList pl,
                                                       const uint64_t *__restrict__ off) {
    const uint64_t total = nbase * (uint64_t)D;
    for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (uint64_t)gridDim.x * 256) {
        const uint64_t i = e / nbase, b = e % nbase;
        amps[base_of(b, pl, ctrl_mask) + off[i]] = out[e];
    }
}

// ---- any descriptor component: the reference's column sum ---------------------------------------

struct KronGroupDev {
    int32_t len;        // indices in the group
    int32_t levels;     // C_MAT levels: the first `levels` indices are controls
    int32_t kind;       // inner matrix: 1 NUMPY (dense), 3 SWAP_MAT
    int32_t param;      // SWAP_MAT block width
    int32_t pos_at;     // first entry in the positions array (state bits, first = MSB of the index)
    int32_t pad_;
    uint64_t mat_at;    // first complex entry of a NUMPY inner matrix
};

struct KronArgs {
    int32_t n_groups;
    int32_t u;                  // union size
    uint8_t upos[QIPGPU_MAX_QUBITS];   // union state bits, enumeration bit u-1-j -> upos[j]
    uint64_t umask;
};

__device__ __forceinline__ cx<double> group_entry(const KronGroupDev &g, const int32_t *__restrict__ pos,
                                                 const cx<double> *__restrict__ mats, uint64_t row, uint64_t col) {
    uint64_t ri = 0, ci = 0;
    for (int m = 0; m < g.len; ++m) {
        const int p = pos[g.pos_at + m];
        ri = (ri << 1) | ((row >> p) & 1ull);
        ci = (ci << 1) | ((col >> p) & 1ull);
    }
    cx<double> one, zero;
    one.re = 1.0;
    one.im = 0.0;
    zero.re = zero.im = 0.0;
    int len = g.len;
    for (int lv = 0; lv < g.levels; ++lv) {
        const uint64_t top = 1ull << (len - 1);
        const bool rb = (ri & top) != 0, cb = (ci & top) != 0;
        if (rb != cb) return zero;
        ri &= top - 1;
        ci &= top - 1;
        --len;
        if (!rb) return ri == ci ? one : zero;
    }
    if (g.kind == 3) {
        const uint64_t half = 1ull << g.param;
        const uint64_t partner = (ri % half) * half + ri / half;
        return partner == ci ? one : zero;
    }
    return mats[g.mat_at + (ri << len) + ci];
}

template <typename R>
__global__ __launch_bounds__(256) void k_kron_general(const cx<R> *__restrict__ in, cx<R> *__restrict__ out,
                                                      uint64_t len, KronArgs a,
                                                      const KronGroupDev *__restrict__ groups,
                                                      const int32_t *__restrict__ pos,
                                                      const cx<double> *__restrict__ mats) {
    const uint64_t cols = 1ull << a.u;
    for (uint64_t row = (uint64_t)blockIdx.x * 256 + threadIdx.x; row < len; row += (uint64_t)gridDim.x * 256) {
        cx<double> s;
        s.re = s.im = 0.0;
        const uint64_t keep = row & ~a.umask;
        for (uint64_t e = 0; e < cols; ++e) {
            uint64_t col = keep;
            for (int j = 0; j < a.u; ++j) col |= ((e >> (a.u - 1 - j)) & 1ull) << a.upos[j];
            cx<double> p;
            p.re = 1.0;
            p.im = 0.0;
            for (int gi = 0; gi < a.n_groups; ++gi) {
                const cx<double> m = group_entry(groups[gi], pos, mats, row, col);
                cx<double> np;
                np.re = p.re * m.re - p.im * m.im;
                np.im = p.re * m.im + p.im * m.re;
                p = np;
                if (p.re == 0.0 && p.im == 0.0) break;
            }
            const cx<double> v = to_d(in[col]);
            s.re += p.re * v.re - p.im * v.im;
            s.im += p.re * v.im + p.im * v.re;
        }
        out[row] = from_d<R>(s);
    }
}

// Diagonal on up to QIPGPU_MAX_DENSE targets, entries in device memory; one item per amplitude of the
// controlled subspace.
template <typename R>
__global__ __launch_bounds__(256) void k_diag_table(cx<R> *amps, uint64_t count, uint64_t ctrl_mask, PosList pl,
                                                    int K, const uint8_t *__restrict__ tpos_g,
                                                    const cx<double> *__restrict__ d) {
    __shared__ uint8_t tpos[QIPGPU_MAX_DENSE];
    if (threadIdx.x < K) tpos[threadIdx.x] = tpos_g[threadIdx.x];
    __syncthreads();
    for (uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x; g < count; g += (uint64_t)gridDim.x * 256) {
        const uint64_t i = base_of(g, pl, ctrl_mask);
        uint32_t j = 0;
        for (int b = 0; b < K; ++b) j = (j << 1) | (uint32_t)((i >> tpos[b]) & 1ull);
        const cx<double> v = to_d(amps[i]), f = d[j];
        cx<double> o;
        o.re = f.re * v.re - f.im * v.im;
        o.im = f.re * v.im + f.im * v.re;
        amps[i] = from_d<R>(o);
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_copy_amps(cx<R> *dst, const cx<R> *__restrict__ src, uint64_t len) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < len; i += (uint64_t)gridDim.x * 256) dst[i] = src[i];
}

// ---- host side -----------------------------------------------------------------------------------

using StreamBuffer = TempBuffer;

std::vector<uint64_t> offsets(const std::vector<int> &tbits) {
    const int K = (int)tbits.size();
    std::vector<uint64_t> off((size_t)1 << K);
    for (size_t j = 0; j < off.size(); ++j) {
        uint64_t o = 0;
        for (int t = 0; t < K; ++t)
            if ((j >> (K - 1 - t)) & 1) o |= 1ull << tbits[t];
        off[j] = o;
    }
    return off;
}

PosList poslist(std::vector<int> pos) {
    std::sort(pos.begin(), pos.end());
    PosList pl;
    pl.count = (int32_t)pos.size();
    for (size_t i = 0; i < pos.size(); ++i) pl.set((int)i, (uint32_t)pos[i]);
    return pl;
}

template <typename R, int K>
int launch_tile(qipgpu_state *s, const std::vector<int> &tbits, uint64_t ctrl_mask, const std::vector<int> &allpos,
                const std::complex<double> *U) {
    constexpr int D = 1 << K;
    const std::vector<uint64_t> off = offsets(tbits);
    std::vector<char> host(sizeof(uint64_t) * D + sizeof(cx<double>) * D * D);
    std::memcpy(host.data(), off.data(), sizeof(uint64_t) * D);
    std::memcpy(host.data() + sizeof(uint64_t) * D, U, sizeof(cx<double>) * D * D);
    void *d = nullptr;
    int rc = ring_upload(s, host.data(), host.size(), &d);
    if (rc) return rc;
    const uint64_t nbase = s->len >> allpos.size();
    const uint64_t ntiles = (nbase + MfmaShape<K>::B - 1) / MfmaShape<K>::B;
    const bool rows_fast = std::find(tbits.begin(), tbits.end(), 0) != tbits.end();
    const cx<double> *dU = (const cx<double> *)((char *)d + sizeof(uint64_t) * D);
    // FP64 MFMA: the waves hold their operator fragments in registers; two workgroups per CU
    const uint64_t grid = std::min<uint64_t>(ntiles, (uint64_t)device_cus() * MfmaShape<K>::WGS);
    if (rows_fast)
        hipLaunchKernelGGL((k_dense_mfma<R, K, true>), dim3((uint32_t)grid), dim3(kTileThreads), 0, s->stream,
                           (cx<R> *)s->amps, nbase, ctrl_mask, poslist(allpos), (const uint64_t *)d, dU, ntiles);
    else
        hipLaunchKernelGGL((k_dense_mfma<R, K, false>), dim3((uint32_t)grid), dim3(kTileThreads), 0, s->stream,
                           (cx<R> *)s->amps, nbase, ctrl_mask, poslist(allpos), (const uint64_t *)d, dU, ntiles);
    QIP_LAUNCHED();
    return QIPGPU_OK;
}

// Temporaries of wide dense operators up to this size stay with the state for the next call.
constexpr size_t kKeepGemmTmpBytes = size_t(1) << 30;

// The offsets and U of a wide dense operator into the state's operator buffer, on its queue: through
// pinned memory, so the copy is asynchronous and the queue is not drained first. The pinned buffer is
// rewritten only once the previous copy out of it has run; the device buffer is rewritten in queue
// order, after the kernels that read it.
int stage_gemm_operator(qipgpu_state *s, const std::vector<uint64_t> &off, const std::complex<double> *U, size_t D,
                        void **dev) {
    const size_t ob = off.size() * sizeof(uint64_t), ub = D * D * sizeof(cx<double>), need = ob + ub;
    if (!s->gemm_ev) QIP_HIP(hipEventCreateWithFlags(&s->gemm_ev, hipEventDisableTiming));
    else QIP_HIP(hipEventSynchronize(s->gemm_ev));
    if (s->h_gemm_op_bytes < need) {
        if (s->h_gemm_op) QIP_HIP(hipHostFree(s->h_gemm_op));
        s->h_gemm_op = nullptr;
        s->h_gemm_op_bytes = 0;
        QIP_HIP(hipHostMalloc(&s->h_gemm_op, need, hipHostMallocDefault));
        s->h_gemm_op_bytes = need;
    }
    if (s->d_gemm_op_bytes < need) {
        QIP_HIP(hipStreamSynchronize(s->stream));   // queued kernels may still read the old buffer
        if (s->d_gemm_op) QIP_HIP(hipFree(s->d_gemm_op));
        s->d_gemm_op = nullptr;
        s->d_gemm_op_bytes = 0;
        QIP_HIP(hipMalloc(&s->d_gemm_op, need));
        s->d_gemm_op_bytes = need;
    }
    std::memcpy(s->h_gemm_op, off.data(), ob);
    std::memcpy((char *)s->h_gemm_op + ob, U, ub);
    QIP_HIP(hipMemcpyAsync(s->d_gemm_op, s->h_gemm_op, need, hipMemcpyHostToDevice, s->stream));
    QIP_HIP(hipEventRecord(s->gemm_ev, s->stream));
    *dev = s->d_gemm_op;
    return QIPGPU_OK;
}

template <typename R>
int launch_gemm(qipgpu_state *s, const std::vector<int> &tbits, uint64_t ctrl_mask, const std::vector<int> &allpos,
                const std::complex<double> *U) {
    const int K = (int)tbits.size();
    const int D = 1 << K;
    const std::vector<uint64_t> off = offsets(tbits);
    void *op = nullptr;
    int rc = stage_gemm_operator(s, off, U, (size_t)D, &op);
    if (rc) return rc;
    const uint64_t nbase = s->len >> allpos.size();
    // the compact temporary: the state's kept one when small, else one for this call
    const size_t tb = nbase * (size_t)D * sizeof(cx<R>);
    StreamBuffer big(s->stream);
    void *tmp = nullptr;
    if (tb <= kKeepGemmTmpBytes) {
        if (s->gemm_tmp_bytes < tb) {
            QIP_HIP(hipStreamSynchronize(s->stream));   // queued kernels may still use the old one
            if (s->gemm_tmp) QIP_HIP(hipFree(s->gemm_tmp));
            s->gemm_tmp = nullptr;
            s->gemm_tmp_bytes = 0;
            QIP_HIP(hipMalloc(&s->gemm_tmp, tb));
            s->gemm_tmp_bytes = tb;
        }
        tmp = s->gemm_tmp;
    } else {
        QIP_HIP(big.alloc(tb));
        tmp = big.p;
    }
    const PosList pl = poslist(allpos);
    const uint64_t *doff = (const uint64_t *)op;
    const cx<double> *dU = (const cx<double> *)((char *)op + off.size() * sizeof(uint64_t));
    const dim3 grid((uint32_t)((nbase + kGemmTile - 1) / kGemmTile), (uint32_t)(D / kGemmTile));
    hipLaunchKernelGGL((k_dense_gemm<R>), grid, dim3(256), 0, s->stream, (const cx<R> *)s->amps, (cx<R> *)tmp,
                       nbase, D, ctrl_mask, pl, doff, dU);
    QIP_LAUNCHED();
    const uint64_t total = nbase * (uint64_t)D;
    const uint32_t sb = (uint32_t)std::min<uint64_t>((total + 255) / 256, (uint64_t)device_cus() * 64);
    hipLaunchKernelGGL((k_dense_scatter<R>), dim3(sb), dim3(256), 0, s->stream, (cx<R> *)s->amps,
                       (const cx<R> *)tmp, nbase, D, ctrl_mask, pl, doff);
    QIP_LAUNCHED();
    return QIPGPU_OK;
}

template <typename R>
int kron_general(qipgpu_state *s, int32_t n_groups, const int32_t *group_len, const int32_t *group_idx,
                 const int32_t *levels, const int32_t *kinds, const int32_t *params, const double *mats) {
    const int n = s->n;
    std::vector<KronGroupDev> groups((size_t)n_groups);
    std::vector<int32_t> pos;
    std::vector<int32_t> uq;
    size_t mat_entries = 0;
    int32_t at = 0;
    for (int g = 0; g < n_groups; ++g) {
        KronGroupDev &G = groups[g];
        G.len = group_len[g];
        G.levels = levels[g];
        G.kind = kinds[g];
        G.param = params[g];
        G.pos_at = (int32_t)pos.size();
        G.pad_ = 0;
        G.mat_at = mat_entries;
        for (int m = 0; m < G.len; ++m) {
            const int32_t q = group_idx[at + m];
            pos.push_back(n - 1 - q);
            if (std::find(uq.begin(), uq.end(), q) == uq.end()) uq.push_back(q);
        }
        at += G.len;
        if (G.kind == 1) {
            const int inner = G.len - G.levels;
            mat_entries += (size_t)1 << (2 * inner);
        }
    }
    std::sort(uq.begin(), uq.end());    // ascending qubits: the enumeration's MSB is the smallest qubit
    KronArgs a;
    a.n_groups = n_groups;
    a.u = (int32_t)uq.size();
    a.umask = 0;
    for (size_t j = 0; j < uq.size(); ++j) {
        a.upos[j] = (uint8_t)(n - 1 - uq[j]);
        a.umask |= 1ull << (n - 1 - uq[j]);
    }
    const size_t gb = groups.size() * sizeof(KronGroupDev), pb = std::max<size_t>(pos.size(), 1) * sizeof(int32_t);
    const size_t mb = std::max<size_t>(mat_entries, 1) * sizeof(cx<double>);
    StreamBuffer desc(s->stream), tmp(s->stream);
    QIP_HIP(desc.alloc(gb + pb + mb));
    char *d = (char *)desc.p;
    QIP_HIP(hipMemcpyAsync(d, groups.data(), gb, hipMemcpyHostToDevice, s->stream));
    if (!pos.empty()) QIP_HIP(hipMemcpyAsync(d + gb, pos.data(), pos.size() * sizeof(int32_t), hipMemcpyHostToDevice, s->stream));
    if (mat_entries) QIP_HIP(hipMemcpyAsync(d + gb + pb, mats, mat_entries * sizeof(cx<double>), hipMemcpyHostToDevice, s->stream));
    QIP_HIP(tmp.alloc(s->len * sizeof(cx<R>)));
    const uint32_t nb = (uint32_t)std::min<uint64_t>((s->len + 255) / 256, (uint64_t)device_cus() * 16);
    hipLaunchKernelGGL((k_kron_general<R>), dim3(nb), dim3(256), 0, s->stream, (const cx<R> *)s->amps,
                       (cx<R> *)tmp.p, s->len, a, (const KronGroupDev *)d, (const int32_t *)(d + gb),
                       (const cx<double> *)(d + gb + pb));
    QIP_LAUNCHED();
    const uint32_t cb = (uint32_t)std::min<uint64_t>((s->len + 255) / 256, (uint64_t)device_cus() * 64);
    hipLaunchKernelGGL((k_copy_amps<R>), dim3(cb), dim3(256), 0, s->stream, (cx<R> *)s->amps, (const cx<R> *)tmp.p, s->len);
    QIP_LAUNCHED();
    QIP_HIP(hipStreamSynchronize(s->stream));   // the host descriptor arrays may be pageable
    return QIPGPU_OK;
}

template <typename R>
int launch_diag_table(qipgpu_state *s, const std::vector<int> &tbits, uint64_t ctrl_mask, const std::vector<int> &cpos,
                      const std::complex<double> *d) {
    const int K = (int)tbits.size();
    const size_t D = (size_t)1 << K;
    std::vector<char> host(D * sizeof(cx<double>) + 64);
    std::memcpy(host.data(), d, D * sizeof(cx<double>));
    uint8_t *tp = (uint8_t *)(host.data() + D * sizeof(cx<double>));
    for (int b = 0; b < K; ++b) tp[b] = (uint8_t)tbits[b];
    StreamBuffer buf(s->stream);
    QIP_HIP(buf.alloc(host.size()));
    QIP_HIP(hipMemcpyAsync(buf.p, host.data(), host.size(), hipMemcpyHostToDevice, s->stream));
    const uint64_t count = s->len >> cpos.size();
    const uint32_t nb = (uint32_t)std::min<uint64_t>((count + 255) / 256, (uint64_t)device_cus() * 64);
    hipLaunchKernelGGL((k_diag_table<R>), dim3(nb), dim3(256), 0, s->stream, (cx<R> *)s->amps, count, ctrl_mask,
                       poslist(cpos), K, (const uint8_t *)buf.p + D * sizeof(cx<double>), (const cx<double> *)buf.p);
    QIP_LAUNCHED();
    QIP_HIP(hipStreamSynchronize(s->stream));   // the host copy may be pageable
    return QIPGPU_OK;
}

}  // namespace

namespace qipgpu {

int diag_large(qipgpu_state *s, const std::vector<int> &tbits, uint64_t ctrl_mask, const std::vector<int> &cpos,
               const std::complex<double> *d) {
    QIP_NOT_CAPTURING(s, "a diagonal on more than 6 targets");
    if ((int)tbits.size() > QIPGPU_MAX_DENSE)
        return set_error(QIPGPU_EUNSUP, "diagonal operator on %d qubits exceeds %d", (int)tbits.size(), QIPGPU_MAX_DENSE);
    return s->dtype == QIPGPU_C128 ? launch_diag_table<double>(s, tbits, ctrl_mask, cpos, d)
                                   : launch_diag_table<float>(s, tbits, ctrl_mask, cpos, d);
}

int dense_large(qipgpu_state *s, const std::vector<int> &tbits, uint64_t ctrl_mask, const std::vector<int> &allpos,
                const std::complex<double> *U) {
    QIP_NOT_CAPTURING(s, "a dense operator on more than 3 targets");
    const int K = (int)tbits.size();
    const bool c128 = s->dtype == QIPGPU_C128;
    switch (K) {
        case 4: return c128 ? launch_tile<double, 4>(s, tbits, ctrl_mask, allpos, U) : launch_tile<float, 4>(s, tbits, ctrl_mask, allpos, U);
        case 5: return c128 ? launch_tile<double, 5>(s, tbits, ctrl_mask, allpos, U) : launch_tile<float, 5>(s, tbits, ctrl_mask, allpos, U);
        case 6: return c128 ? launch_tile<double, 6>(s, tbits, ctrl_mask, allpos, U) : launch_tile<float, 6>(s, tbits, ctrl_mask, allpos, U);
        default: break;
    }
    if (K >= 7 && K <= QIPGPU_MAX_DENSE)
        return c128 ? launch_gemm<double>(s, tbits, ctrl_mask, allpos, U) : launch_gemm<float>(s, tbits, ctrl_mask, allpos, U);
    return set_error(QIPGPU_EUNSUP, "dense operator on %d target qubits exceeds %d", K, QIPGPU_MAX_DENSE);
}

}  // namespace qipgpu

extern "C" {

int qipgpu_apply_kron(qipgpu_state *s, int32_t n_groups, const int32_t *group_len, const int32_t *group_idx,
                      const int32_t *levels, const int32_t *kinds, const int32_t *params, const double *mats) {
    if (!s) return set_error(QIPGPU_EINVAL, "null state");
    QIP_NOT_CAPTURING(s, "a descriptor by the column sum");
    if (n_groups < 1) return set_error(QIPGPU_EINVAL, "a descriptor needs at least one group");
    std::vector<int32_t> all;
    int32_t at = 0;
    for (int g = 0; g < n_groups; ++g) {
        const int32_t L = group_len[g];
        if (L < 1 || L > s->n) return set_error(QIPGPU_EINVAL, "group %d has %d indices", g, L);
        if (levels[g] < 0 || levels[g] > L) return set_error(QIPGPU_EINVAL, "group %d: %d control levels", g, levels[g]);
        const int inner = L - levels[g];
        if (kinds[g] == 1) {
            if (inner > 14) return set_error(QIPGPU_EUNSUP, "group %d: a %d-qubit dense matrix", g, inner);
        } else if (kinds[g] == 3) {
            if (params[g] < 0 || 2 * params[g] != inner)
                return set_error(QIPGPU_EINVAL, "group %d: SwapMat(%d) on %d indices", g, params[g], inner);
        } else {
            return set_error(QIPGPU_EINVAL, "group %d: matrix kind %d", g, kinds[g]);
        }
        // indices within one group must differ; groups may share indices (kronprod.pyx:140-141)
        int rc = check_qubits(s, L, group_idx + at, "group indices");
        if (rc) return rc;
        for (int m = 0; m < L; ++m)
            if (std::find(all.begin(), all.end(), group_idx[at + m]) == all.end()) all.push_back(group_idx[at + m]);
        at += L;
    }
    if ((int)all.size() > 30) return set_error(QIPGPU_EUNSUP, "descriptor spans %d qubits", (int)all.size());
    int rc = flush_recorded(s);
    if (rc) return rc;
    return s->dtype == QIPGPU_C128 ? kron_general<double>(s, n_groups, group_len, group_idx, levels, kinds, params, mats)
                                   : kron_general<float>(s, n_groups, group_len, group_idx, levels, kinds, params, mats);
}

}  // extern "C"
