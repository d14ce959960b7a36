// tile6.h -- the six-target-qubit tile of a complex64 state and its split-BF16 MFMA update,
// shared by the single-state fused path (fused.hip) and the batched path (batch.hip).
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
// Data flow per tile (256 threads = 4 waves, two workgroups per CU):
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
// The caller sums the per-thread FP64 norms into per-workgroup partials; the next block's scale is
// 1 / sqrt of their fixed-order sum.
#pragma once

#include <algorithm>
#include <cstring>
#include <vector>

#include "engine.h"

namespace qipgpu {
namespace tile {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

constexpr int kTileBits = 12;
constexpr int kRowPad = 72;              // 64 rows + 8 bf16 of padding per column (bank spread)
constexpr int kPlaneElems = 64 * kRowPad;
constexpr int kRawStride = 65;           // 64 amplitudes + 1 slot of padding per raw row
constexpr size_t kFragBytes = 4 * 8 * 2 * 64 * 16;   // [wave][ks][hi/lo][lane] x 16 B = 64 KB

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

// ---- host side ---------------------------------------------------------------------------------

inline uint16_t bf16_bits_rne(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

inline float bf16_value(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Tile bits of a block whose real target state bits are `tmask` (1 to 6 bits, n >= 12). The 12 tile
// bits P are the targets plus the first free bits in the order 0..5, 12..17, 6..11, 18..n-1; the
// targets are padded to six with the first of those extra bits (identity rows of the operator) and
// the other six are the column bits L. Keeping the six lowest state bits in P makes every run of
// the tile 512 B; the next preference puts the remaining bits 2^12 amplitudes apart, the pattern
// measured fastest at n = 30 (per block: P = 0..5 + 12..17 3.08 ms, P = 0..11 3.45 ms,
// P = 0..5 + 24..29 3.58 ms; targets {0,1,20..23}: columns {2..7} 3.63 ms, {2..5,12,13} 3.25 ms).
inline void choose_tile_bits(int n, uint64_t tmask, int (&T)[6], int (&L)[6]) {
    int order[64], no = 0;
    for (int b = 0; b < 6; ++b) order[no++] = b;
    for (int b = 12; b < 18; ++b) order[no++] = b;
    for (int b = 6; b < 12; ++b) order[no++] = b;
    for (int b = 18; b < 64; ++b) order[no++] = b;
    int nt = 0, nl = 0;
    for (int b = 0; b < 64; ++b)
        if (tmask >> b & 1) T[nt++] = b;
    for (int i = 0; i < no && (nt < 6 || nl < 6); ++i) {
        const int b = order[i];
        if (b >= n || (tmask >> b & 1)) continue;
        if (nt < 6) T[nt++] = b;
        else L[nl++] = b;
    }
    std::sort(T, T + 6);
    std::sort(L, L + 6);
}

// Tile arguments for target bits T and column bits L (each six distinct bits, ascending, disjoint,
// below n): P = T u L ascending, u bit i <-> state bit P[i].
inline FusedArgs make_tile_args(int n, const int (&T)[6], const int (&L)[6]) {
    FusedArgs a;
    std::memset(&a, 0, sizeof(a));
    a.ntiles = 1ull << (n - kTileBits);
    int pbits[12];
    for (int b = 0; b < 6; ++b) {
        pbits[b] = T[b];
        pbits[6 + b] = L[b];
    }
    std::sort(pbits, pbits + 12);
    for (int i = 0; i < 12; ++i) a.pbits[i] = (uint8_t)pbits[i];
    for (int b = 0; b < 6; ++b) {
        a.tu[b] = (uint8_t)(std::find(pbits, pbits + 12, T[b]) - pbits);
        a.lu[b] = (uint8_t)(std::find(pbits, pbits + 12, L[b]) - pbits);
    }
    return a;
}

// Position of BF16 element j of the A fragment (wave w, k-step ks, lane) in the real operator
// W = [[Ur, -Ui], [Ui, Ur]] with M-block rows permuted so that wave w's 32 rows are the Re rows
// then the Im rows of complex rows 16w..16w+15: complex row rr, complex column cc, and which
// component of U (and with which sign) the entry is.
struct FragEntry {
    int rr, cc;
    bool imag;      // take Im U (else Re U)
    bool negate;
};

__host__ __device__ inline FragEntry frag_entry(int w, int ks, int lane, int j) {
    const int i = lane & 31;
    const bool im_row = i >= 16;
    const int col = 16 * ks + 8 * (lane >> 5) + j;
    FragEntry e;
    e.rr = 16 * w + (i & 15);
    e.cc = col & 63;
    if (!im_row) {
        e.imag = col >= 64;       // Re rows: [Ur, -Ui]
        e.negate = col >= 64;
    } else {
        e.imag = col < 64;        // Im rows: [Ui, Ur]
        e.negate = false;
    }
    return e;
}

// Split the 64x64 operator (row-bit order, Ur/Ui row-major) into hi/lo BF16 A fragments ordered
// [wave][ks][hi/lo][lane][j] (kFragBytes in all). Each of the 8192 distinct entries is split once;
// the negated copies of Ui flip the sign bit of both halves (round-to-nearest-even is symmetric),
// except a zero low half, which stays +0 as x - x does.
inline void split_fragments(const double *ur, const double *ui, uint16_t *frag) {
    std::vector<uint16_t> hl(4 * 4096);   // [Ur hi | Ur lo | Ui hi | Ui lo]
    for (int p = 0; p < 2; ++p) {
        const double *src = p ? ui : ur;
        uint16_t *hi = hl.data() + p * 8192, *lo = hi + 4096;
        for (int i = 0; i < 4096; ++i) {
            hi[i] = bf16_bits_rne((float)src[i]);
            lo[i] = bf16_bits_rne((float)(src[i] - (double)bf16_value(hi[i])));
        }
    }
    // frag_entry for the 8 entries j of one (wave, k-step, lane): one row, 8 consecutive columns
    for (int w = 0; w < 4; ++w)
        for (int ks = 0; ks < 8; ++ks)
            for (int lane = 0; lane < 64; ++lane) {
                const FragEntry e = frag_entry(w, ks, lane, 0);
                const uint16_t *h = hl.data() + (e.imag ? 8192 : 0) + e.rr * 64 + e.cc;
                const uint16_t *l = h + 4096;
                uint16_t *fh = frag + ((size_t)((w * 8 + ks) * 2) * 64 + lane) * 8, *fl = fh + 64 * 8;
                if (e.negate) {
                    for (int j = 0; j < 8; ++j) {
                        fh[j] = (uint16_t)(h[j] ^ 0x8000u);
                        fl[j] = (l[j] & 0x7fffu) ? (uint16_t)(l[j] ^ 0x8000u) : l[j];
                    }
                } else {
                    for (int j = 0; j < 8; ++j) {
                        fh[j] = h[j];
                        fl[j] = l[j];
                    }
                }
            }
}

}  // namespace tile
}  // namespace qipgpu
