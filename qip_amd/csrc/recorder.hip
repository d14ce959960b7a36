// recorder.hip -- precision form, host side: gates recorded on a complex64 state are fused online
// into blocks of at most `width` qubits, and each block is composed (qipgpu_compose) and launched
// (qipgpu_fused_apply, or the plain dense sweep below 12 qubits) as soon as no later gate can join it.
//
// Greedy rule, one gate at a time in program order: for a gate on qubit set S let j be the last block
// that touches S. The gate joins the first block from j on (those after j do not touch S, so the gate
// commutes past them) whose qubit set stays within the width, else it opens a block at the end.
// Since a later gate scans from the last block touching one of its qubits, once every qubit of the
// state has been touched by a block after block i, no later gate can reach block i: it is final.
// Launching final blocks while the caller records the rest gives the same blocks in the same order,
// and the same results bit for bit, as fusing the whole record at the next read.
//
// Every other entry point that touches the state first launches what is recorded (flush_recorded),
// so recorded gates always act in program order with direct calls.
#include <algorithm>
#include <deque>
#include <vector>

#include "engine.h"

namespace qipgpu {

namespace {

constexpr int kFusedMinQubits = 12;   // the MFMA tile spans 12 state bits

// The greedy rule over qubit masks (n <= 64), blocks numbered globally from 0; `base` is the global
// number of the oldest open block.
struct OnlineFuser {
    int n = 0, width = 6;
    int64_t base = 0;
    std::deque<uint64_t> masks;      // open blocks
    std::vector<int64_t> last;       // per qubit: global number of the last block touching it, or -1
    int touched = 0;                 // qubits with last >= 0

    void reset(int n_, int width_) {
        n = n_;
        width = width_;
        base = 0;
        masks.clear();
        last.assign(n, -1);
        touched = 0;
    }

    // Place a gate on qubits q[0..k-1]; returns the open-block position (0 = oldest) it joined.
    size_t add(const int32_t *q, int k) {
        uint64_t S = 0;
        int64_t j = -1;
        for (int t = 0; t < k; ++t) {
            S |= 1ull << q[t];
            j = std::max(j, last[q[t]]);
        }
        size_t pos = j > base ? (size_t)(j - base) : 0;
        for (; pos < masks.size(); ++pos)
            if (__builtin_popcountll(masks[pos] | S) <= width) break;
        if (pos == masks.size()) masks.push_back(0);
        masks[pos] |= S;
        const int64_t g = base + (int64_t)pos;
        for (int t = 0; t < k; ++t) {
            if (last[q[t]] < 0) ++touched;
            last[q[t]] = std::max(last[q[t]], g);
        }
        return pos;
    }

    // Leading open blocks that no later gate can join.
    int64_t final_count() const {
        if (touched < n) return 0;
        int64_t m = last[0];
        for (int q = 1; q < n; ++q) m = std::min(m, last[q]);
        return std::max<int64_t>(0, m - base);
    }

    void take(int64_t count) {
        for (int64_t i = 0; i < count; ++i) masks.pop_front();
        base += count;
    }
};

struct RecBlock {
    std::vector<int32_t> k, q;     // per gate: width, and its qubits (six slots)
    std::vector<double> mats;      // gate matrices in order, interleaved complex
    bool product = true;           // single-qubit gates only
};

// The launch decisions, shared by the recorder and its host-only restatement (qipgpu_fuse_launches);
// `blocks[i].product` is all they read.
//
// End of the launch that starts at open block i of the first `count`: a run of consecutive
// single-qubit-only blocks is one launch (where product runs apply), any other block one launch.
template <class Blocks>
size_t launch_end(const Blocks &blocks, size_t i, size_t count, bool runs) {
    if (!runs || !blocks[i].product) return i + 1;
    size_t j = i;
    while (j < count && blocks[j].product) ++j;
    return j;
}

// Final blocks that can launch now: all of them, except a trailing run of single-qubit-only blocks,
// which waits until the block after it is final too (or the next flush), so that a run is always
// launched whole and the launches do not depend on when blocks became final.
template <class Blocks>
int64_t launchable_count(const Blocks &blocks, int64_t final_count, bool runs) {
    int64_t c = final_count;
    if (!runs || c == 0 || c == (int64_t)blocks.size()) return c;
    while (c > 0 && blocks[(size_t)c - 1].product && blocks[(size_t)c].product) --c;
    return c;
}

}  // namespace

struct Recorder {
    OnlineFuser fuser;
    std::deque<RecBlock> blocks;   // parallel to fuser.masks
    bool streaming = true;
    bool in_flush = false;         // launching: the entry points called from here must not flush again
    int64_t pending = 0;           // recorded gates not launched
    int64_t launched = 0;          // blocks launched since the recorder was made
};

namespace {

int validate_gate(const qipgpu_state *s, int width, int32_t g, int32_t k, const int32_t *q) {
    if (k < 1 || k > width)
        return set_error(QIPGPU_EINVAL, "recorded gate %d on %d qubits (1..%d)", g, k, width);
    for (int t = 0; t < k; ++t) {
        if (q[t] < 0 || q[t] >= s->n)
            return set_error(QIPGPU_EINVAL, "recorded gate %d: qubit %d outside 0..%d", g, q[t], s->n - 1);
        for (int u = 0; u < t; ++u)
            if (q[u] == q[t]) return set_error(QIPGPU_EINVAL, "recorded gate %d: qubit %d repeated", g, q[t]);
    }
    return QIPGPU_OK;
}

// One launch: a dense block (composed matrix on its qubits) or a run of consecutive single-qubit-only
// blocks, which is a product of one 2x2 per qubit and goes to the product sweep (FP64 arithmetic,
// up to 4 low + 8 other factors per pass) instead of the split-BF16 MFMA kernel.
struct Launch {
    bool product = false;
    std::vector<int32_t> qubits;   // dense: the block's qubits ascending; product: the factors' qubits ascending
    std::vector<double> mat;       // dense: 2^k x 2^k; product: one 2x2 per qubit
};

int compose_dense(const RecBlock &b, uint64_t mask, Launch &c) {
    c.product = false;
    c.qubits.clear();
    for (int q = 0; q < 64; ++q)
        if (mask >> q & 1) c.qubits.push_back(q);
    const int k = (int)c.qubits.size();
    c.mat.resize((size_t)2 << (2 * k));
    return qipgpu_compose(k, c.qubits.data(), (int32_t)b.k.size(), b.k.data(), b.q.data(), b.mats.data(),
                          c.mat.data());
}

// The per-qubit products of blocks [i0, i1), all single-qubit-only, each qubit's gates in program
// order (blocks are in program order on every qubit), each product composed by qipgpu_compose.
int compose_product(const std::deque<RecBlock> &blocks, size_t i0, size_t i1, Launch &c) {
    c.product = true;
    c.qubits.clear();
    c.mat.clear();
    std::vector<std::vector<double>> per(64);
    for (size_t i = i0; i < i1; ++i) {
        const RecBlock &b = blocks[i];
        for (size_t g = 0; g < b.k.size(); ++g)
            per[b.q[6 * g]].insert(per[b.q[6 * g]].end(), b.mats.begin() + 8 * g, b.mats.begin() + 8 * g + 8);
    }
    for (int q = 0; q < 64; ++q) {
        if (per[q].empty()) continue;
        const int32_t ng = (int32_t)(per[q].size() / 8);
        std::vector<int32_t> gk(ng, 1), gq(6 * (size_t)ng, 0);
        for (int32_t g = 0; g < ng; ++g) gq[6 * g] = q;
        double out[8];
        const int rc = qipgpu_compose(1, &q, ng, gk.data(), gq.data(), per[q].data(), out);
        if (rc) return rc;
        c.qubits.push_back(q);
        c.mat.insert(c.mat.end(), out, out + 8);
    }
    return QIPGPU_OK;
}

int launch(qipgpu_state *s, const Launch &c) {
    const int k = (int)c.qubits.size();
    if (!c.product) {
        if (s->n >= kFusedMinQubits) return qipgpu_fused_apply(s, k, c.qubits.data(), c.mat.data());
        return qipgpu_apply_matrix(s, k, c.qubits.data(), c.mat.data(), 0, nullptr);
    }
    // one product call per pass: the factors on the kProductLowBits lowest state bits with up to
    // kProductHighPerPass others, then kProductHighPerPass more per call
    std::vector<int> low, high;
    for (int j = 0; j < k; ++j) (s->n - 1 - c.qubits[j] < kProductLowBits ? low : high).push_back(j);
    size_t next = 0;
    bool first = true;
    while (first || next < high.size()) {
        std::vector<int> pick(first ? low : std::vector<int>());
        for (int t = 0; t < kProductHighPerPass && next < high.size(); ++t) pick.push_back(high[next++]);
        first = false;
        if (pick.empty()) continue;
        std::vector<int32_t> t;
        std::vector<double> m;
        for (int j : pick) {
            t.push_back(c.qubits[j]);
            m.insert(m.end(), c.mat.begin() + 8 * j, c.mat.begin() + 8 * j + 8);
        }
        const int rc = qipgpu_apply_product(s, (int32_t)t.size(), t.data(), m.data());
        if (rc) return rc;
    }
    return QIPGPU_OK;
}

// Product runs exist only where the product sweep does (n >= 12).
bool product_runs(const qipgpu_state *s) { return s->n >= kFusedMinQubits; }

// Compose the first `count` open blocks, take them out, then launch them back to back. A run of
// consecutive single-qubit-only blocks is one product launch.
int launch_leading(qipgpu_state *s, int64_t count) {
    Recorder *r = s->recorder;
    if (count <= 0) return QIPGPU_OK;
    std::vector<Launch> ls;
    for (size_t i = 0; i < (size_t)count;) {
        ls.emplace_back();
        const size_t j = launch_end(r->blocks, i, (size_t)count, product_runs(s));
        const int rc = r->blocks[i].product && product_runs(s) ? compose_product(r->blocks, i, j, ls.back())
                                                                : compose_dense(r->blocks[i], r->fuser.masks[i], ls.back());
        if (rc) return rc;
        i = j;
    }
    for (int64_t i = 0; i < count; ++i) {
        r->pending -= (int64_t)r->blocks.front().k.size();
        r->blocks.pop_front();
    }
    r->fuser.take(count);
    r->launched += count;
    r->in_flush = true;
    int rc = QIPGPU_OK;
    for (size_t i = 0; i < ls.size() && rc == QIPGPU_OK; ++i) rc = launch(s, ls[i]);
    r->in_flush = false;
    return rc;
}

int64_t launchable(const qipgpu_state *s, int64_t final_count) {
    return launchable_count(s->recorder->blocks, final_count, product_runs(s));
}

int ensure_recorder(qipgpu_state *s) {
    if (s->dtype != QIPGPU_C64) return set_error(QIPGPU_EINVAL, "gates are recorded on complex64 states");
    if (!s->recorder) {
        s->recorder = new Recorder();
        s->recorder->fuser.reset(s->n, 6);
    }
    return QIPGPU_OK;
}

}  // namespace

int flush_recorded(qipgpu_state *s) {
    Recorder *r = s->recorder;
    if (!r || r->in_flush || r->blocks.empty()) return QIPGPU_OK;
    return launch_leading(s, (int64_t)r->blocks.size());
}

void recorder_discard(qipgpu_state *s) {
    Recorder *r = s->recorder;
    if (!r || r->in_flush) return;
    r->blocks.clear();
    r->fuser.reset(s->n, r->fuser.width);
    r->pending = 0;
}

void recorder_free(qipgpu_state *s) {
    delete s->recorder;
    s->recorder = nullptr;
}

}  // namespace qipgpu

using namespace qipgpu;

extern "C" {

int qipgpu_fused_set_recording(qipgpu_state *s, int32_t width, int32_t streaming) {
    if (width < 1 || width > 6) return set_error(QIPGPU_EINVAL, "fusion width %d outside 1..6", width);
    int rc = ensure_recorder(s);
    if (rc) return rc;
    rc = flush_recorded(s);
    if (rc) return rc;
    s->recorder->fuser.reset(s->n, width);
    s->recorder->streaming = streaming != 0;
    return QIPGPU_OK;
}

int qipgpu_fused_record(qipgpu_state *s, int32_t n_gates, const int32_t *gate_k, const int32_t *gate_qubits,
                        const double *mats) {
    int rc = ensure_recorder(s);
    if (rc) return rc;
    if (n_gates < 0) return set_error(QIPGPU_EINVAL, "negative gate count");
    if (n_gates > 0 && (!gate_k || !gate_qubits || !mats)) return set_error(QIPGPU_EINVAL, "null argument");
    Recorder *r = s->recorder;
    if (r->in_flush) return set_error(QIPGPU_EINVAL, "recording while recorded blocks launch");
    for (int32_t g = 0; g < n_gates; ++g) {   // all or nothing: check every gate first
        rc = validate_gate(s, r->fuser.width, g, gate_k[g], gate_qubits + 6 * (size_t)g);
        if (rc) return rc;
    }
    size_t moff = 0;
    for (int32_t g = 0; g < n_gates; ++g) {
        const int k = gate_k[g];
        const int32_t *q = gate_qubits + 6 * (size_t)g;
        const size_t pos = r->fuser.add(q, k);
        if (pos == r->blocks.size()) r->blocks.emplace_back();
        RecBlock &b = r->blocks[pos];
        b.k.push_back(k);
        b.q.insert(b.q.end(), q, q + 6);
        const size_t len = (size_t)2 << (2 * k);
        b.mats.insert(b.mats.end(), mats + moff, mats + moff + len);
        if (k > 1) b.product = false;
        moff += len;
        ++r->pending;
        if (r->streaming) {
            rc = launch_leading(s, launchable(s, r->fuser.final_count()));
            if (rc) return rc;
        }
    }
    return QIPGPU_OK;
}

int qipgpu_fused_flush(qipgpu_state *s) {
    return flush_recorded(s);
}

int qipgpu_fused_recorded(const qipgpu_state *s, int64_t *pending_gates, int64_t *blocks_launched) {
    *pending_gates = s->recorder ? s->recorder->pending : 0;
    *blocks_launched = s->recorder ? s->recorder->launched : 0;
    return QIPGPU_OK;
}

int qipgpu_fuse_online(int32_t n, int32_t width, int32_t n_gates, const int32_t *gate_k, const int32_t *gate_qubits,
                       int64_t *gate_block, int64_t *final_after) {
    if (n < 1 || n > 64) return set_error(QIPGPU_EINVAL, "%d qubits (1..64)", n);
    if (width < 1 || width > 6) return set_error(QIPGPU_EINVAL, "fusion width %d outside 1..6", width);
    if (n_gates < 0) return set_error(QIPGPU_EINVAL, "negative gate count");
    OnlineFuser f;
    f.reset(n, width);
    for (int32_t g = 0; g < n_gates; ++g) {
        const int32_t k = gate_k[g];
        const int32_t *q = gate_qubits + 6 * (size_t)g;
        if (k < 1 || k > width) return set_error(QIPGPU_EINVAL, "gate %d on %d qubits (1..%d)", g, k, width);
        for (int t = 0; t < k; ++t) {
            if (q[t] < 0 || q[t] >= n) return set_error(QIPGPU_EINVAL, "gate %d: qubit %d outside 0..%d", g, q[t], n - 1);
            for (int u = 0; u < t; ++u)
                if (q[u] == q[t]) return set_error(QIPGPU_EINVAL, "gate %d: qubit %d repeated", g, q[t]);
        }
        gate_block[g] = f.base + (int64_t)f.add(q, k);
        // as the streaming recorder does: take out every block that is final after this gate
        const int64_t fc = f.final_count();
        f.take(fc);
        final_after[g] = f.base;
    }
    return QIPGPU_OK;
}

int qipgpu_fuse_launches(int32_t n, int32_t width, int32_t streaming, int32_t n_gates, const int32_t *gate_k,
                         const int32_t *gate_qubits, int64_t *launch_of_block, int64_t *n_blocks) {
    if (n < 1 || n > 64) return set_error(QIPGPU_EINVAL, "%d qubits (1..64)", n);
    if (width < 1 || width > 6) return set_error(QIPGPU_EINVAL, "fusion width %d outside 1..6", width);
    if (n_gates < 0) return set_error(QIPGPU_EINVAL, "negative gate count");
    struct Flag {
        bool product = true;
    };
    const bool runs = n >= kFusedMinQubits;
    OnlineFuser f;
    f.reset(n, width);
    std::deque<Flag> open;
    int64_t done = 0, launches = 0;
    auto launch_first = [&](int64_t count) {
        for (size_t i = 0; i < (size_t)count;) {
            const size_t j = launch_end(open, i, (size_t)count, runs);
            for (size_t b = i; b < j; ++b) launch_of_block[done + (int64_t)b] = launches;
            ++launches;
            i = j;
        }
        for (int64_t i = 0; i < count; ++i) open.pop_front();
        f.take(count);
        done += count;
    };
    for (int32_t g = 0; g < n_gates; ++g) {
        const int32_t k = gate_k[g];
        const int32_t *q = gate_qubits + 6 * (size_t)g;
        if (k < 1 || k > width) return set_error(QIPGPU_EINVAL, "gate %d on %d qubits (1..%d)", g, k, width);
        for (int t = 0; t < k; ++t) {
            if (q[t] < 0 || q[t] >= n) return set_error(QIPGPU_EINVAL, "gate %d: qubit %d outside 0..%d", g, q[t], n - 1);
            for (int u = 0; u < t; ++u)
                if (q[u] == q[t]) return set_error(QIPGPU_EINVAL, "gate %d: qubit %d repeated", g, q[t]);
        }
        const size_t pos = f.add(q, k);
        if (pos == open.size()) open.emplace_back();
        if (k > 1) open[pos].product = false;
        if (streaming) launch_first(launchable_count(open, f.final_count(), runs));
    }
    launch_first((int64_t)open.size());   // the flush at the next read
    *n_blocks = done;
    return QIPGPU_OK;
}

}  // extern "C"
