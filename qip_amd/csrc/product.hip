// product.hip -- product sweep: a product of single-qubit operators on up to 12 distinct qubits
// applied in one pass over the state (plain path, complex128 or complex64 storage, FP64 math).
//
// A descriptor such as H on a whole register (qip/operators.py:97-99) is one operator in the
// reference, applied in one cdot_loop sweep. Here its single-qubit factors commute and are applied
// tile by tile: the state is cut into tiles of 4096 amplitudes spanning the target bits plus the
// lowest free bits (12 tile bits P). Every pass keeps the four lowest state bits in P, so a tile is
// read in runs of at least 16 amplitudes (256 B for complex128): a pass takes at most 8 targets
// above bit 3, and a product with more is applied in several passes. A workgroup loads a tile in
// address order (u = tile bits ascending) into LDS as FP64, applies the gates three target bits
// per round (each thread takes groups of 8 amplitudes that differ only in the round's bits and
// applies the round's 2x2 gates in registers), and writes the tile back once: one HBM read and one
// write per amplitude for up to 12 gates, and for a complex64 state one rounding instead of one
// per gate. LDS rows of 8 amplitudes are padded by one slot so that every access pattern of a
// round spreads over the banks.
#include <algorithm>
#include <cstring>
#include <vector>

#include "engine.h"

using qipgpu::cx;
using namespace qipgpu;

namespace {

constexpr int kTileBits = 12;
constexpr int kTile = 1 << kTileBits;
constexpr int kPadded = kTile + kTile / 8;
constexpr int kMaxRounds = 4;
constexpr int kBlocksPerCU = 2;   // 72 KB of LDS per workgroup

struct ProductArgs {
    uint64_t ntiles;
    uint8_t pbits[12];                 // u bit i <-> state bit pbits[i], ascending
    int32_t nrounds;
    int32_t rcount[kMaxRounds];        // target bits in each round (1..3)
    uint8_t rbits[kMaxRounds][3];      // their u bit positions, ascending
    cx<double> g[kMaxRounds][3][4];    // the gate on rbits[r][j]: row-major 2x2
};

__device__ __forceinline__ int ppos(uint32_t u) { return (int)(u + (u >> 3)); }

__device__ __forceinline__ uint64_t tile_base(uint64_t t, const ProductArgs &a) {
#pragma unroll
    for (int i = 0; i < kTileBits; ++i) {
        const uint32_t p = a.pbits[i];
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

// One round over the tile in LDS: CNT target bits, 8-amplitude (or 4, 2) groups per thread.
template <int CNT>
__device__ __forceinline__ void product_round(cx<double> *t, const ProductArgs &a, int r) {
    constexpr int G = 1 << CNT;                 // amplitudes per group
    constexpr int PER = (kTile / G) / 256;      // groups per thread
    uint32_t b[3];
#pragma unroll
    for (int j = 0; j < CNT; ++j) b[j] = a.rbits[r][j];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        uint32_t gi = threadIdx.x + 256u * q;
#pragma unroll
        for (int j = 0; j < CNT; ++j) {           // insert zeros at the round's bits, ascending
            const uint32_t lo = gi & ((1u << b[j]) - 1u);
            gi = ((gi ^ lo) << 1) | lo;
        }
        cx<double> x[G];
#pragma unroll
        for (int e = 0; e < G; ++e) {
            uint32_t u = gi;
#pragma unroll
            for (int j = 0; j < CNT; ++j) u |= (uint32_t)((e >> j) & 1) << b[j];
            x[e] = t[ppos(u)];
        }
#pragma unroll
        for (int j = 0; j < CNT; ++j) {
            const cx<double> m00 = a.g[r][j][0], m01 = a.g[r][j][1], m10 = a.g[r][j][2], m11 = a.g[r][j][3];
#pragma unroll
            for (int e = 0; e < G; ++e) {
                if ((e >> j) & 1) continue;
                const int f = e | (1 << j);
                const cx<double> x0 = x[e], x1 = x[f];
                x[e] = mac2(m00, x0, m01, x1);
                x[f] = mac2(m10, x0, m11, x1);
            }
        }
#pragma unroll
        for (int e = 0; e < G; ++e) {
            uint32_t u = gi;
#pragma unroll
            for (int j = 0; j < CNT; ++j) u |= (uint32_t)((e >> j) & 1) << b[j];
            t[ppos(u)] = x[e];
        }
    }
}

template <typename R>
__global__ __launch_bounds__(256, 2) void k_product(cx<R> *amps, ProductArgs a) {
    __shared__ cx<double> t[kPadded];
    const uint32_t tid = threadIdx.x;
    const uint64_t off_t = u_offset<0, 8>(tid, a);       // u bits 0..7 are the thread index
    for (uint64_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
        const uint64_t base = tile_base(tile, a) + off_t;
        cx<R> v[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = amps[base + u_offset<8, 12>(256u * i, a)];
        __syncthreads();   // the previous tile has left LDS
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            cx<double> w;
            w.re = (double)v[i].re;
            w.im = (double)v[i].im;
            t[ppos(tid + 256u * i)] = w;
        }
        __syncthreads();
        for (int r = 0; r < a.nrounds; ++r) {
            if (a.rcount[r] == 3) product_round<3>(t, a, r);
            else if (a.rcount[r] == 2) product_round<2>(t, a, r);
            else product_round<1>(t, a, r);
            __syncthreads();
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const cx<double> w = t[ppos(tid + 256u * i)];
            cx<R> o;
            o.re = (R)w.re;
            o.im = (R)w.im;
            amps[base + u_offset<8, 12>(256u * i, a)] = o;
        }
    }
}

uint32_t product_grid(uint64_t ntiles) {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) == hipSuccess && p.multiProcessorCount > 0) cus = p.multiProcessorCount;
    }
    (void)hipGetLastError();
    return (uint32_t)std::min<uint64_t>(ntiles, (uint64_t)cus * kBlocksPerCU * 4);
}

constexpr int kLowBits = 4;         // state bits every pass keeps in its tile
constexpr int kHighPerPass = kTileBits - kLowBits;

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
    std::vector<int> pbits;
    for (int b = 0; b < n && (int)pbits.size() < kTileBits; ++b)
        if (!(tmask >> b & 1)) pbits.push_back(b);
    // keep the k targets, fill with the lowest free bits, sort
    pbits.resize(kTileBits - k);
    for (int j = 0; j < k; ++j) pbits.push_back(tbits[j]);
    std::sort(pbits.begin(), pbits.end());
    ProductArgs a;
    std::memset(&a, 0, sizeof(a));
    a.ntiles = 1ull << (n - kTileBits);
    for (int i = 0; i < kTileBits; ++i) a.pbits[i] = (uint8_t)pbits[i];
    // rounds of up to three targets, in u-bit order
    std::vector<std::pair<int, int>> tu;   // (u bit, gate index)
    for (int j = 0; j < k; ++j)
        tu.push_back({(int)(std::find(pbits.begin(), pbits.end(), tbits[j]) - pbits.begin()), j});
    std::sort(tu.begin(), tu.end());
    a.nrounds = (k + 2) / 3;
    for (int r = 0; r < a.nrounds; ++r) {
        const int lo = 3 * r, hi = std::min(k, 3 * r + 3);
        a.rcount[r] = hi - lo;
        for (int j = lo; j < hi; ++j) {
            a.rbits[r][j - lo] = (uint8_t)tu[j].first;
            const double *m = mats + 8 * tu[j].second;
            for (int e = 0; e < 4; ++e) {
                a.g[r][j - lo][e].re = m[2 * e];
                a.g[r][j - lo][e].im = m[2 * e + 1];
            }
        }
    }
    const uint32_t grid = product_grid(a.ntiles);
    if (s->dtype == QIPGPU_C128)
        hipLaunchKernelGGL(k_product<double>, dim3(grid), dim3(256), 0, s->stream, (cx<double> *)s->amps, a);
    else
        hipLaunchKernelGGL(k_product<float>, dim3(grid), dim3(256), 0, s->stream, (cx<float> *)s->amps, a);
    QIP_LAUNCHED();
    return QIPGPU_OK;
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
