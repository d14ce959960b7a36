// product.hip -- product sweep: a product of single-qubit operators on up to 12 distinct qubits
// applied in one pass over the state (plain path, complex128 or complex64 storage, FP64 math).
//
// A descriptor such as H on a whole register (qip/operators.py:97-99) is one operator in the
// reference, applied in one cdot_loop sweep. Here its single-qubit factors commute and are applied
// tile by tile: the state is cut into tiles of 4096 amplitudes spanning 12 tile bits, numbered u.
// u bits 0..3 are always state bits 0..3, so a tile is read in runs of at least 16 amplitudes: a
// pass takes at most 8 targets above bit 3, and a product with more is applied in several passes.
// The other targets and the lowest free state bits fill u bits 4..11, placed by what each factor
// needs (product_pass).
//
// Each thread holds 16 amplitudes of the tile in FP64 registers, and a factor on u bit b is applied
// where b is one of the four bits that vary across those 16:
//   layout A (as loaded and stored, address order): u = tid + 256 e, u bits 8..11 = e;
//   layout B: u bits 4..7 = e;  layout C: u bits 0..3 = e.
// The tile is loaded in layout A and moves through LDS only as far as the factors need: factors on
// u bits 8..11 cost no LDS traffic at all; A -> B (one transpose, stored from B), A -> C -> A or
// A -> C -> B (two) for the others. Layout B stores whole runs of 16 amplitudes per 16 lanes, as A
// does; C does not (its lanes run 16 amplitudes apart), so a pass never ends there. One HBM read and
// one write per amplitude for up to 12 gates, and for a complex64 state one rounding instead of one
// per gate. LDS rows of 16 amplitudes are padded by one slot, so that the 16-amplitude-stride reads
// of layout C spread over the banks.
//
// A pass whose factors all sit on u bits 8..11 needs no LDS and runs a variant without the tile,
// three workgroups per CU for a complex64 state. Passes that need transposes are latency-bound
// rather than bandwidth-bound: the 68 KB FP64 tile and 178 VGPRs hold the LDS variant to 8 waves
// per CU, which do not cover the barrier-separated phases (on a 24-qubit complex64 pass over all
// three layouts, LDS is busy about an eighth of the time and the FP64 units about a third).
#include <algorithm>
#include <cstring>
#include <vector>

#include "engine.h"

using qipgpu::cx;
using namespace qipgpu;

namespace {

constexpr int kTileBits = 12;
constexpr int kTile = 1 << kTileBits;
constexpr int kPadded = kTile + kTile / 16;

// The factor on a tile bit is applied in one of three ways, by what its 2x2 matrix is:
//   X, exactly [[0, 1], [1, 0]]: no arithmetic. It flips the bit of every amplitude's index and
//     commutes with the other factors, so the tile is stored with that bit of its addresses flipped.
//     On u bits 6, 7 (the wave index in layout A) a wave then stores amplitudes another wave loaded:
//     a barrier after every wave's loads of the tile orders that (the LDS variant's first barrier,
//     or one per tile in the variant without LDS). A pass of X factors needs no LDS transposes;
//   real: two multiplies per output component;  complex: the full complex 2x2.
struct ProductArgs {
    uint64_t ntiles;
    uint8_t pbits[12];                 // u bit i <-> state bit pbits[i]
    uint8_t zbits[12];                 // the same state bits, ascending (tile bases)
    uint32_t fmask;                    // u bits whose factor needs arithmetic (all but X)
    uint32_t rmask;                    // ... of which real matrices
    uint32_t xmask;                    // u bits whose factor is X (flipped at the store)
    uint32_t xsync;                    // 1: a barrier per tile before its stores (X on u bit 6 or 7, no LDS)
    cx<double> g[12][4];               // the factor on u bit i: row-major 2x2
};

// The tile between layouts, in LDS: FP64 for both state precisions (68 KB, 2 workgroups per CU).
struct TileStore {
    cx<double> v[kPadded];
    // A view of the slots base, base + 1, ...: put/get take a constant offset from base, which the
    // LDS instructions carry as an immediate (one address register per layout, not one per slot)
    struct View {
        cx<double> *p;
        __device__ __forceinline__ void put(int k, cx<double> x) const { p[k] = x; }
        __device__ __forceinline__ cx<double> get(int k) const { return p[k]; }
    };
    __device__ __forceinline__ View at(int base) { return View{v + base}; }
};

// Workgroups per CU: the LDS variant is bounded by its tile, the variant without LDS by registers
// (16 FP64 amplitudes and one raw prefetch buffer per thread: 3 fit with complex64 raw data).
template <typename R, bool kLds>
constexpr int blocks_per_cu() { return (kLds || sizeof(R) == 8) ? 2 : 3; }

__device__ __forceinline__ uint64_t tile_base(uint64_t t, const ProductArgs &a) {
#pragma unroll
    for (int i = 0; i < kTileBits; ++i) {
        const uint32_t p = a.zbits[i];
        const uint64_t lo = t & ((1ull << p) - 1ull);
        t = ((t ^ lo) << 1) | lo;
    }
    return t;
}

template <int LO, int HI>
__device__ __forceinline__ uint64_t u_offset(uint32_t u, const ProductArgs &a) {
    uint64_t o = 0;
#pragma unroll
    for (int i = LO; i < HI; ++i) o |= (uint64_t)((u >> i) & 1u) << a.pbits[i];
    return o;
}

__device__ __forceinline__ cx<double> mac2(cx<double> m0, cx<double> x0, cx<double> m1, cx<double> x1) {
    cx<double> r;
    r.re = m0.re * x0.re - m0.im * x0.im + m1.re * x1.re - m1.im * x1.im;
    r.im = m0.re * x0.im + m0.im * x0.re + m1.re * x1.im + m1.im * x1.re;
    return r;
}

// The factors on u bits B..B+3 applied to 16 registers whose index e is u bits B..B+3.
template <int B>
__device__ __forceinline__ void apply4(cx<double> (&x)[16], const ProductArgs &a) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (!((a.fmask >> (B + j)) & 1u)) continue;
        const cx<double> m00 = a.g[B + j][0], m01 = a.g[B + j][1], m10 = a.g[B + j][2], m11 = a.g[B + j][3];
        if ((a.rmask >> (B + j)) & 1u) {   // real matrix: half the arithmetic
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                if ((e >> j) & 1) continue;
                const int f = e | (1 << j);
                const cx<double> x0 = x[e], x1 = x[f];
                x[e].re = m00.re * x0.re + m01.re * x1.re;
                x[e].im = m00.re * x0.im + m01.re * x1.im;
                x[f].re = m10.re * x0.re + m11.re * x1.re;
                x[f].im = m10.re * x0.im + m11.re * x1.im;
            }
            continue;
        }
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            if ((e >> j) & 1) continue;
            const int f = e | (1 << j);
            const cx<double> x0 = x[e], x1 = x[f];
            x[e] = mac2(m00, x0, m01, x1);
            x[f] = mac2(m10, x0, m11, x1);
        }
    }
}

template <typename R>
__device__ __forceinline__ void load_raw(cx<R> (&v)[16], const cx<R> *amps, uint64_t base, const ProductArgs &a) {
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = amps[base + u_offset<8, 12>(256u * i, a)];
}

template <typename R>
__device__ __forceinline__ cx<R> narrow(cx<double> x) {
    cx<R> o;
    o.re = (R)x.re;
    o.im = (R)x.im;
    return o;
}

// One tile: its 16 amplitudes per thread (layout A, widened to FP64) in, the factors applied, rounded
// and stored back. The tile visits only the layouts that hold factors, loaded in A and stored from the
// last one: A; A -> B; A -> C -> B; A -> C -> A (no factor on u bits 4..7: C cannot be stored from,
// since its lanes run 16 amplitudes apart). Layouts A and B both store whole runs of 16 amplitudes
// per 16 lanes. The store addresses have the X bits flipped: off_sa / off_sb are this thread's part
// of the tile offset in A / B with them flipped.
template <typename R, bool kLds>
__device__ __forceinline__ void product_tile(cx<double> (&x)[16], cx<R> *amps, uint64_t tbase, uint64_t off_sa,
                                             uint64_t off_sb, TileStore *lds, const ProductArgs &a, bool use_b,
                                             bool use_c, uint32_t tid) {
    apply4<8>(x, a);
    if (kLds && (use_b || use_c)) {
        TileStore &t = *lds;
        // Tile position u sits in LDS slot u + u / 16 (rows of 16 padded by one). Register e of thread
        // tid holds u = tid + 256 e in layout A, (tid & 15) | e << 4 | (tid >> 4) << 8 in B and
        // e | tid << 4 in C, so each layout's slots are a per-thread base plus a constant per register:
        // A: tid + tid / 16 + 272 e;  B: (tid & 15) + 272 (tid >> 4) + 17 e;  C: 17 tid + e
        const auto ta = t.at((int)(tid + (tid >> 4)));
        const auto tb = t.at((int)((tid & 15u) + 272u * (tid >> 4)));
        const auto tc = t.at((int)(17u * tid));
        // A phase reads slots other threads wrote in the phase before it, behind a barrier, and writes
        // back only where it read. A tile that ends in A (its own slots) leaves nothing another thread
        // still reads; one that ends in B does, so the next tile's A writes wait for a barrier.
        if (use_b) __syncthreads();
#pragma unroll
        for (int i = 0; i < 16; ++i) ta.put(272 * i, x[i]);
        __syncthreads();
        if (use_c) {
#pragma unroll
            for (int e = 0; e < 16; ++e) x[e] = tc.get(e);
            apply4<0>(x, a);
#pragma unroll
            for (int e = 0; e < 16; ++e) tc.put(e, x[e]);
            __syncthreads();
        }
        if (use_b) {
#pragma unroll
            for (int e = 0; e < 16; ++e) x[e] = tb.get(17 * e);
            apply4<4>(x, a);
            // register e holds u bits 4..7 = e; it goes to those bits xor the X bits
            const uint32_t xmid = (a.xmask >> 4) & 15u;
#pragma unroll
            for (int e = 0; e < 16; ++e)
                amps[tbase + off_sb + u_offset<4, 8>(((uint32_t)e ^ xmid) << 4, a)] = narrow<R>(x[e]);
            return;
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = ta.get(272 * i);
    }
    // register i holds u = tid + 256 i; it goes to u xor xmask (off_sa has the thread bits flipped)
    const uint32_t xhi = (a.xmask >> 8) & 15u;
#pragma unroll
    for (int i = 0; i < 16; ++i) amps[tbase + off_sa + u_offset<8, 12>(256u * ((uint32_t)i ^ xhi), a)] = narrow<R>(x[i]);
}

// Persistent grid: each workgroup takes a contiguous range of tiles. A tile's raw amplitudes are
// widened to FP64 registers, and the raw registers at once take the next tile's loads, which stay
// in flight while this tile is processed (the last prefetch re-reads an in-range tile and is not
// used).
template <typename R, bool kLds>
__device__ __forceinline__ void product_loop(cx<R> *amps, const ProductArgs &a, TileStore *t) {
    const uint32_t tid = threadIdx.x;
    const uint64_t off_t = u_offset<0, 8>(tid, a);       // layout A: u bits 0..7 are the thread index
    const uint64_t off_sa = u_offset<0, 8>(tid ^ (a.xmask & 0xffu), a);
    // layout B: u bits 0..3 and 8..11 are the thread index
    const uint64_t off_sb = u_offset<0, 12>(((tid & 15u) | (tid >> 4) << 8) ^ (a.xmask & 0xf0fu), a);
    const bool use_b = (a.fmask & 0xf0u) != 0, use_c = (a.fmask & 0x0fu) != 0;
    const uint64_t per = (a.ntiles + gridDim.x - 1) / gridDim.x;
    uint64_t tile = (uint64_t)blockIdx.x * per;
    const uint64_t end = min(tile + per, a.ntiles);
    if (tile >= end) return;
    const uint64_t last = end - 1;
    cx<R> v[16];
    load_raw(v, amps, tile_base(tile, a) + off_t, a);
    for (;;) {
        cx<double> x[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            x[i].re = (double)v[i].re;
            x[i].im = (double)v[i].im;
        }
        if (!kLds && a.xsync) __syncthreads();   // every wave holds the tile before any wave stores
        load_raw(v, amps, tile_base(min(tile + 1, last), a) + off_t, a);
        product_tile<R, kLds>(x, amps, tile_base(tile, a), off_sa, off_sb, t, a, use_b, use_c, tid);
        if (++tile >= end) return;
    }
}

template <typename R, bool kLds>
__global__ __launch_bounds__(256, (blocks_per_cu<R, kLds>())) void k_product(cx<R> *amps, ProductArgs a) {
    if constexpr (kLds) {
        __shared__ TileStore t;
        product_loop<R, true>(amps, a, &t);
    } else {
        product_loop<R, false>(amps, a, nullptr);
    }
}

template <typename R, bool kLds>
int launch_product(qipgpu_state *s, const ProductArgs &a) {
    const uint32_t grid = (uint32_t)std::min<uint64_t>(a.ntiles, (uint64_t)device_cus() * blocks_per_cu<R, kLds>());
    return launch_gate(s, &k_product<R, kLds>, dim3(grid), dim3(256), (cx<R> *)s->amps, a);
}

constexpr int kLowBits = kProductLowBits;           // state bits every pass keeps in its tile
constexpr int kHighPerPass = kProductHighPerPass;   // = kTileBits - kLowBits
static_assert(kLowBits + kHighPerPass == kTileBits, "a pass tile holds 12 state bits");

// One pass: the k single-qubit gates (k <= 12, all targets of at most kHighPerPass above the low
// bits) in one sweep.
int product_pass(qipgpu_state *s, int32_t k, const int32_t *targets, const double *mats) {
    const int n = s->n;
    uint64_t tmask = 0;
    std::vector<int> tbits(k);
    for (int j = 0; j < k; ++j) {
        tbits[j] = n - 1 - targets[j];
        tmask |= 1ull << tbits[j];
    }
    auto is_x = [&](int j) {
        const double *m = mats + 8 * j;
        return m[0] == 0.0 && m[1] == 0.0 && m[2] == 1.0 && m[3] == 0.0 && m[4] == 1.0 && m[5] == 0.0 &&
               m[6] == 0.0 && m[7] == 0.0;
    };
    // u bits 0..3 are state bits 0..3, which every pass keeps. The targets above them go on u bits
    // 4..11 by kind: a factor that needs arithmetic takes the register bits 8..11 of layout A first
    // (applied without a transpose), then 6, 7, 4, 5; an X takes the lane bits 4, 5 first, then
    // 8..11, and 6, 7 last (flipped at the store everywhere; on 6, 7 at the cost of a barrier). The
    // free u bits take the lowest free state bits.
    int ubit[kTileBits];
    std::fill(ubit, ubit + kTileBits, -1);
    for (int b = 0; b < kLowBits; ++b) ubit[b] = b;
    std::vector<int> place(k, -1);
    for (int j = 0; j < k; ++j)
        if (tbits[j] < kLowBits) place[j] = tbits[j];
    auto take = [&](int j, std::initializer_list<int> order) {
        for (int u : order)
            if (ubit[u] < 0) {
                ubit[u] = tbits[j];
                place[j] = u;
                return;
            }
    };
    for (int j = 0; j < k; ++j)
        if (place[j] < 0 && !is_x(j)) take(j, {8, 9, 10, 11});
    for (int j = 0; j < k; ++j)
        if (place[j] < 0 && is_x(j)) take(j, {4, 5, 8, 9, 10, 11});
    for (int j = 0; j < k; ++j)
        if (place[j] < 0) take(j, {6, 7, 4, 5});
    // at most kHighPerPass = 8 targets above the low bits share the 8 u bits 4..11: after the
    // rounds above every one has a place (arithmetic overflow finds 8..11 full and 4..7 free enough,
    // X overflow finds only 6, 7 free)
    for (int j = 0; j < k; ++j)
        if (place[j] < 0) return set_error(QIPGPU_EINVAL, "product pass: target %d has no tile bit", targets[j]);
    for (int b = kLowBits, u = kLowBits; u < kTileBits; ++u) {
        if (ubit[u] >= 0) continue;
        while (tmask >> b & 1) ++b;
        ubit[u] = b++;
    }
    ProductArgs a;
    std::memset(&a, 0, sizeof(a));
    a.ntiles = 1ull << (n - kTileBits);
    std::vector<int> sorted(ubit, ubit + kTileBits);
    std::sort(sorted.begin(), sorted.end());
    for (int i = 0; i < kTileBits; ++i) {
        a.pbits[i] = (uint8_t)ubit[i];
        a.zbits[i] = (uint8_t)sorted[i];
    }
    // each factor's kind by its matrix (the comment at ProductArgs)
    for (int j = 0; j < k; ++j) {
        const int ub = place[j];
        const double *m = mats + 8 * j;
        const bool real = m[1] == 0.0 && m[3] == 0.0 && m[5] == 0.0 && m[7] == 0.0;
        if (is_x(j)) {
            a.xmask |= 1u << ub;
            continue;
        }
        a.fmask |= 1u << ub;
        if (real) a.rmask |= 1u << ub;
        for (int e = 0; e < 4; ++e) {
            a.g[ub][e].re = m[2 * e];
            a.g[ub][e].im = m[2 * e + 1];
        }
    }
    const bool lds = (a.fmask & 0xffu) != 0;   // a factor below u bit 8 needs a transpose
    a.xsync = !lds && (a.xmask & 0xc0u) != 0;
    if (s->dtype == QIPGPU_C128) return lds ? launch_product<double, true>(s, a) : launch_product<double, false>(s, a);
    return lds ? launch_product<float, true>(s, a) : launch_product<float, false>(s, a);
}

}  // namespace

extern "C" int qipgpu_apply_product(qipgpu_state *s, int32_t k, const int32_t *targets, const double *mats) {
    if (k < 1 || k > QIPGPU_MAX_PRODUCT) return set_error(QIPGPU_EINVAL, "product of %d single-qubit gates (1..%d)", k,
                                                         QIPGPU_MAX_PRODUCT);
    if (s->n < kTileBits) return set_error(QIPGPU_EUNSUP, "product sweep needs n >= %d (n=%d)", kTileBits, s->n);
    int rc = check_qubits(s, k, targets, "product qubits");
    if (rc) return rc;
    // linear in the state: a pending fused renormalisation (a scalar) stays pending; recorded gates
    // go first
    rc = flush_recorded(s);
    if (rc) return rc;
    // the factors commute: the first pass takes every factor on the low bits and up to
    // kHighPerPass others, each further pass up to kHighPerPass more
    std::vector<int> low, high;
    for (int j = 0; j < k; ++j) (s->n - 1 - targets[j] < kLowBits ? low : high).push_back(j);
    size_t next = 0;
    bool first = true;
    while (first || next < high.size()) {
        std::vector<int> pick(first ? low : std::vector<int>());
        for (int c = 0; c < kHighPerPass && next < high.size(); ++c) pick.push_back(high[next++]);
        first = false;
        if (pick.empty()) continue;
        std::vector<int32_t> t;
        std::vector<double> m;
        for (int j : pick) {
            t.push_back(targets[j]);
            m.insert(m.end(), mats + 8 * j, mats + 8 * j + 8);
        }
        if ((rc = product_pass(s, (int32_t)t.size(), t.data(), m.data()))) return rc;
    }
    return QIPGPU_OK;
}
