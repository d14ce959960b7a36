// recorder.hip -- fused forms, host side: gates recorded on a state are planned into blocks of at
// most `width` qubits by the light-cone rule (plan.h), and each block is composed in FP64
// (qipgpu_compose) and launched as soon as it is final, i.e. no gate recorded later can change it.
// On a complex64 state from 12 qubits on a block goes to the split-BF16 MFMA kernel
// (qipgpu_fused_apply, the precision form); on a complex128 state, or below 12 qubits, to the FP64
// dense sweep (qipgpu_apply_matrix: registers, LDS tiles or FP64 MFMA tiles by width), so a
// complex128 state keeps FP64 arithmetic with one dense sweep per block instead of one per gate.
//
// Streaming. The recorded gates not launched yet are planned again every max(16, pending / 8)
// recorded gates (so planning costs a bounded number of steps per gate), and the final leading
// blocks are launched. Final blocks are exactly the blocks of planning the whole record at once, so
// streaming and deferred recording launch the same blocks in the same order, with the same results
// bit for bit, whenever the plans happen.
//
// Every other entry point that touches the state first launches what is recorded (flush_recorded),
// so recorded gates always act in program order with direct calls.
#include <algorithm>
#include <deque>
#include <vector>

#include "engine.h"
#include "plan.h"

namespace qipgpu {

namespace {

constexpr int kFusedMinQubits = 12;   // the MFMA tile spans 12 state bits

// After each recorded gate (streaming): plan again once this many gates have been recorded since the
// last plan.
bool plan_due(int64_t since_plan, int64_t pending) {
    return since_plan >= std::max<int64_t>(16, pending / 8);
}

// The launch decisions, shared by the recorder and its host-only restatement (qipgpu_fuse_launches);
// product[b] says block b holds single-qubit gates only.
//
// End of the launch that starts at block i of the first `count`: a run of consecutive single-qubit-
// only blocks is one launch (where product runs apply), any other block one launch.
size_t launch_end(const std::vector<char> &product, size_t i, size_t count, bool runs) {
    if (!runs || !product[i]) return i + 1;
    size_t j = i;
    while (j < count && product[j]) ++j;
    return j;
}

// Final blocks that can launch now: all of them, except a trailing run of single-qubit-only blocks,
// which waits until a final block that is not single-qubit-only follows it (or the next flush), so
// that a run is always launched whole and the launches do not depend on when the plans happen.
int64_t launchable_count(const std::vector<char> &product, int64_t final_count, bool runs) {
    int64_t c = final_count;
    if (runs)
        while (c > 0 && product[(size_t)c - 1]) --c;
    return c;
}

std::vector<char> product_flags(const std::vector<PlannedBlock> &blocks, const int32_t *gate_k) {
    std::vector<char> p(blocks.size(), 1);
    for (size_t b = 0; b < blocks.size(); ++b)
        for (int64_t g : blocks[b].gates)
            if (gate_k[g] > 1) p[b] = 0;
    return p;
}

}  // namespace

// A plan depends only on the gates' qubit masks, the width, n and whether it stops at the first open
// block, so programs that record the same pattern again (the layers of an iterative algorithm) reuse
// it instead of planning again.
struct PlanMemo {
    std::vector<uint64_t> masks;
    int width = 0, n = 0;
    bool stop_at_open = false;
    std::vector<PlannedBlock> blocks;
    int64_t final_count = 0;
};
constexpr size_t kPlanMemos = 4;
constexpr size_t kPlanMemoMaxGates = 4096;

struct Recorder {
    int n = 0, width = 6;
    std::deque<PlanMemo> memos;       // most recent first
    // recorded gates not launched yet, in program order
    std::vector<int32_t> k, q;        // width and qubits (six slots) per gate
    std::vector<uint64_t> masks;      // qubit mask per gate
    std::vector<size_t> moff;         // offset of each gate's matrix in mats
    std::vector<double> mats;         // interleaved complex
    bool streaming = true;
    bool in_flush = false;            // launching: the entry points called from here must not flush again
    int64_t since_plan = 0;           // gates recorded since the last plan
    int64_t launched = 0;             // blocks launched since the recorder was made

    int64_t pending() const { return (int64_t)k.size(); }

    void reset(int n_, int width_) {
        n = n_;
        width = width_;
        k.clear();
        q.clear();
        masks.clear();
        moff.clear();
        mats.clear();
        since_plan = 0;
    }
};

namespace {

// plan_blocks over the recorded gates, through the recorder's memo of recent plans
int64_t plan_recorded(Recorder &r, bool stop_at_open, std::vector<PlannedBlock> &blocks) {
    for (size_t i = 0; i < r.memos.size(); ++i) {
        const PlanMemo &m = r.memos[i];
        if (m.width == r.width && m.n == r.n && m.stop_at_open == stop_at_open && m.masks == r.masks) {
            blocks = m.blocks;
            const int64_t fc = m.final_count;
            if (i) std::rotate(r.memos.begin(), r.memos.begin() + (long)i, r.memos.begin() + (long)i + 1);
            return fc;
        }
    }
    const int64_t fc = plan_blocks(r.masks.data(), r.pending(), r.width, r.n, stop_at_open, blocks);
    if (r.masks.size() <= kPlanMemoMaxGates) {
        if (r.memos.size() == kPlanMemos) r.memos.pop_back();
        PlanMemo m;
        m.masks = r.masks;
        m.width = r.width;
        m.n = r.n;
        m.stop_at_open = stop_at_open;
        m.blocks = blocks;
        m.final_count = fc;
        r.memos.push_front(std::move(m));
    }
    return fc;
}

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

size_t mat_len(int32_t k) { return (size_t)2 << (2 * k); }

int compose_dense(const Recorder &r, const PlannedBlock &b, Launch &c) {
    c.product = false;
    c.qubits.clear();
    for (int q = 0; q < 64; ++q)
        if (b.mask >> q & 1) c.qubits.push_back(q);
    const int k = (int)c.qubits.size();
    std::vector<int32_t> gk, gq;
    std::vector<double> gm;
    for (int64_t g : b.gates) {
        gk.push_back(r.k[g]);
        gq.insert(gq.end(), r.q.begin() + 6 * g, r.q.begin() + 6 * g + 6);
        gm.insert(gm.end(), r.mats.begin() + r.moff[g], r.mats.begin() + r.moff[g] + mat_len(r.k[g]));
    }
    c.mat.resize(mat_len(k));
    return qipgpu_compose(k, c.qubits.data(), (int32_t)gk.size(), gk.data(), gq.data(), gm.data(), c.mat.data());
}

// The per-qubit products of blocks [i0, i1), all single-qubit-only. Dependencies keep each qubit's
// gates in block order, and a block's gates are in program order, so each qubit's list is in program
// order; each product is composed by qipgpu_compose.
int compose_product(const Recorder &r, const std::vector<PlannedBlock> &blocks, size_t i0, size_t i1, Launch &c) {
    c.product = true;
    c.qubits.clear();
    c.mat.clear();
    std::vector<std::vector<double>> per(64);
    for (size_t i = i0; i < i1; ++i)
        for (int64_t g : blocks[i].gates) {
            const int32_t q = r.q[6 * g];
            per[q].insert(per[q].end(), r.mats.begin() + r.moff[g], r.mats.begin() + r.moff[g] + 8);
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
        if (s->dtype == QIPGPU_C64 && s->n >= kFusedMinQubits)
            return qipgpu_fused_apply(s, k, c.qubits.data(), c.mat.data());
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

// Drop the gates of blocks [0, count) from the pending record (the rest keep program order).
void remove_planned(Recorder &r, const std::vector<PlannedBlock> &blocks, size_t count) {
    std::vector<char> used(r.k.size(), 0);
    for (size_t b = 0; b < count; ++b)
        for (int64_t g : blocks[b].gates) used[g] = 1;
    size_t w = 0;
    std::vector<double> mats;
    std::vector<size_t> moff;
    for (size_t g = 0; g < r.k.size(); ++g) {
        if (used[g]) continue;
        moff.push_back(mats.size());
        mats.insert(mats.end(), r.mats.begin() + r.moff[g], r.mats.begin() + r.moff[g] + mat_len(r.k[g]));
        r.k[w] = r.k[g];
        std::copy(r.q.begin() + 6 * g, r.q.begin() + 6 * g + 6, r.q.begin() + 6 * w);
        r.masks[w] = r.masks[g];
        ++w;
    }
    r.k.resize(w);
    r.q.resize(6 * w);
    r.masks.resize(w);
    r.moff.swap(moff);
    r.mats.swap(mats);
}

// Compose blocks [0, count) of the plan, take their gates out of the record, then launch them back
// to back. A run of consecutive single-qubit-only blocks is one product launch.
int launch_planned(qipgpu_state *s, const std::vector<PlannedBlock> &blocks, int64_t count) {
    Recorder *r = s->recorder;
    if (count <= 0) return QIPGPU_OK;
    const std::vector<char> product = product_flags(blocks, r->k.data());
    std::vector<Launch> ls;
    for (size_t i = 0; i < (size_t)count;) {
        ls.emplace_back();
        const size_t j = launch_end(product, i, (size_t)count, product_runs(s));
        const int rc = product[i] && product_runs(s) ? compose_product(*r, blocks, i, j, ls.back())
                                                     : compose_dense(*r, blocks[i], ls.back());
        if (rc) return rc;
        i = j;
    }
    remove_planned(*r, blocks, (size_t)count);
    r->launched += count;
    r->in_flush = true;
    int rc = QIPGPU_OK;
    for (size_t i = 0; i < ls.size() && rc == QIPGPU_OK; ++i) rc = launch(s, ls[i]);
    r->in_flush = false;
    return rc;
}

// Streaming: plan what is recorded and launch the final leading blocks that can launch now.
int launch_final(qipgpu_state *s) {
    Recorder *r = s->recorder;
    r->since_plan = 0;
    std::vector<PlannedBlock> blocks;
    const int64_t fc = plan_recorded(*r, true, blocks);
    const std::vector<char> product = product_flags(blocks, r->k.data());
    return launch_planned(s, blocks, launchable_count(product, fc, product_runs(s)));
}

int ensure_recorder(qipgpu_state *s) {
    if (!s->recorder) {
        s->recorder = new Recorder();
        s->recorder->reset(s->n, 6);
    }
    return QIPGPU_OK;
}

int check_plan_args(int32_t n, int32_t width, int32_t n_gates) {
    if (n < 1 || n > 64) return set_error(QIPGPU_EINVAL, "%d qubits (1..64)", n);
    if (width < 1 || width > 6) return set_error(QIPGPU_EINVAL, "fusion width %d outside 1..6", width);
    if (n_gates < 0) return set_error(QIPGPU_EINVAL, "negative gate count");
    return QIPGPU_OK;
}

// Qubit masks of a host-only gate list, checked as the recorder checks recorded gates.
int gate_masks(int32_t n, int32_t width, int32_t n_gates, const int32_t *gate_k, const int32_t *gate_qubits,
               std::vector<uint64_t> &masks) {
    masks.assign((size_t)n_gates, 0);
    for (int32_t g = 0; g < n_gates; ++g) {
        const int32_t k = gate_k[g];
        const int32_t *q = gate_qubits + 6 * (size_t)g;
        if (k < 1 || k > width) return set_error(QIPGPU_EINVAL, "gate %d on %d qubits (1..%d)", g, k, width);
        for (int t = 0; t < k; ++t) {
            if (q[t] < 0 || q[t] >= n) return set_error(QIPGPU_EINVAL, "gate %d: qubit %d outside 0..%d", g, q[t], n - 1);
            if (masks[g] >> q[t] & 1) return set_error(QIPGPU_EINVAL, "gate %d: qubit %d repeated", g, q[t]);
            masks[g] |= 1ull << q[t];
        }
    }
    return QIPGPU_OK;
}

// The recorder's schedule over a host-only gate list: gates recorded one at a time, planned when
// plan_due says so (when streaming), flushed at the end. For every taken block, `take` receives the
// global ids of its gates and whether it starts a new launch.
template <class Take>
void simulate_recorder(int32_t n, int32_t width, bool streaming, bool use_launchable, int32_t n_gates,
                       const int32_t *gate_k, const std::vector<uint64_t> &all_masks, Take take,
                       int64_t *final_after) {
    const bool runs = n >= kFusedMinQubits;
    std::vector<uint64_t> masks;
    std::vector<int32_t> ks;
    std::vector<int64_t> ids;
    std::vector<PlannedBlock> blocks;
    int64_t since = 0, taken = 0;
    auto take_first = [&](int64_t count) {
        if (count <= 0) return;
        const std::vector<char> product = product_flags(blocks, ks.data());
        std::vector<char> used(ids.size(), 0);
        for (size_t i = 0; i < (size_t)count;) {
            const size_t j = launch_end(product, i, (size_t)count, runs);
            for (size_t b = i; b < j; ++b) {
                std::vector<int64_t> gids;
                for (int64_t g : blocks[b].gates) {
                    gids.push_back(ids[g]);
                    used[g] = 1;
                }
                take(gids, b == i);
            }
            i = j;
        }
        size_t w = 0;
        for (size_t g = 0; g < ids.size(); ++g) {
            if (used[g]) continue;
            masks[w] = masks[g];
            ks[w] = ks[g];
            ids[w] = ids[g];
            ++w;
        }
        masks.resize(w);
        ks.resize(w);
        ids.resize(w);
        taken += count;
    };
    for (int32_t g = 0; g < n_gates; ++g) {
        masks.push_back(all_masks[g]);
        ks.push_back(gate_k[g]);
        ids.push_back(g);
        ++since;
        if (streaming && plan_due(since, (int64_t)masks.size())) {
            since = 0;
            const int64_t fc = plan_blocks(masks.data(), (int64_t)masks.size(), width, n, true, blocks);
            take_first(use_launchable ? launchable_count(product_flags(blocks, ks.data()), fc, runs) : fc);
        }
        if (final_after) final_after[g] = taken;
    }
    plan_blocks(masks.data(), (int64_t)masks.size(), width, n, false, blocks);
    take_first((int64_t)blocks.size());
}

}  // namespace

int flush_recorded(qipgpu_state *s) {
    int rc = flush_held(s);   // held-back diagonal gates came first
    if (rc) return rc;
    Recorder *r = s->recorder;
    if (!r || r->in_flush || r->k.empty()) return QIPGPU_OK;
    std::vector<PlannedBlock> blocks;
    plan_recorded(*r, false, blocks);
    r->since_plan = 0;
    rc = launch_planned(s, blocks, (int64_t)blocks.size());
    if (rc) return rc;
    return flush_held(s);     // a complex128 block that is diagonal is held like any such gate
}

void recorder_discard(qipgpu_state *s) {
    Recorder *r = s->recorder;
    if (!r || r->in_flush) return;
    r->reset(s->n, r->width);
}

void recorder_free(qipgpu_state *s) {
    delete s->recorder;
    s->recorder = nullptr;
}

int64_t recorder_pending(const qipgpu_state *s) { return s->recorder ? s->recorder->pending() : 0; }

}  // namespace qipgpu

using namespace qipgpu;

extern "C" {

int qipgpu_fused_set_recording(qipgpu_state *s, int32_t width, int32_t streaming) {
    if (width < 1 || width > 6) return set_error(QIPGPU_EINVAL, "fusion width %d outside 1..6", width);
    int rc = ensure_recorder(s);
    if (rc) return rc;
    rc = flush_recorded(s);
    if (rc) return rc;
    s->recorder->reset(s->n, width);
    s->recorder->streaming = streaming != 0;
    return QIPGPU_OK;
}

int qipgpu_fused_record(qipgpu_state *s, int32_t n_gates, const int32_t *gate_k, const int32_t *gate_qubits,
                        const double *mats) {
    QIP_NOT_CAPTURING(s, "gate recording for fusion");
    int rc = flush_held(s);   // diagonal gates held back come before what is recorded now
    if (rc) return rc;
    rc = ensure_recorder(s);
    if (rc) return rc;
    if (n_gates < 0) return set_error(QIPGPU_EINVAL, "negative gate count");
    if (n_gates > 0 && (!gate_k || !gate_qubits || !mats)) return set_error(QIPGPU_EINVAL, "null argument");
    Recorder *r = s->recorder;
    if (r->in_flush) return set_error(QIPGPU_EINVAL, "recording while recorded blocks launch");
    for (int32_t g = 0; g < n_gates; ++g) {   // all or nothing: check every gate first
        rc = validate_gate(s, r->width, g, gate_k[g], gate_qubits + 6 * (size_t)g);
        if (rc) return rc;
    }
    size_t moff = 0;
    for (int32_t g = 0; g < n_gates; ++g) {
        const int k = gate_k[g];
        const int32_t *q = gate_qubits + 6 * (size_t)g;
        uint64_t mask = 0;
        for (int t = 0; t < k; ++t) mask |= 1ull << q[t];
        r->k.push_back(k);
        r->q.insert(r->q.end(), q, q + 6);
        r->masks.push_back(mask);
        r->moff.push_back(r->mats.size());
        r->mats.insert(r->mats.end(), mats + moff, mats + moff + mat_len(k));
        moff += mat_len(k);
        ++r->since_plan;
        if (r->streaming && plan_due(r->since_plan, r->pending())) {
            rc = launch_final(s);
            if (rc) return rc;
        }
    }
    return QIPGPU_OK;
}

int qipgpu_fused_flush(qipgpu_state *s) {
    return flush_recorded(s);
}

int qipgpu_fused_recorded(const qipgpu_state *s, int64_t *pending_gates, int64_t *blocks_launched) {
    *pending_gates = s->recorder ? s->recorder->pending() : 0;
    *blocks_launched = s->recorder ? s->recorder->launched : 0;
    return QIPGPU_OK;
}

int qipgpu_fuse_online(int32_t n, int32_t width, int32_t n_gates, const int32_t *gate_k, const int32_t *gate_qubits,
                       int64_t *gate_block, int64_t *final_after) {
    int rc = check_plan_args(n, width, n_gates);
    if (rc) return rc;
    std::vector<uint64_t> masks;
    rc = gate_masks(n, width, n_gates, gate_k, gate_qubits, masks);
    if (rc) return rc;
    int64_t block = 0;
    simulate_recorder(n, width, true, false, n_gates, gate_k, masks,
                      [&](const std::vector<int64_t> &gids, bool) {
                          for (int64_t g : gids) gate_block[g] = block;
                          ++block;
                      },
                      final_after);
    return QIPGPU_OK;
}

int qipgpu_fuse_launches(int32_t n, int32_t width, int32_t streaming, int32_t n_gates, const int32_t *gate_k,
                         const int32_t *gate_qubits, int64_t *launch_of_block, int64_t *n_blocks) {
    int rc = check_plan_args(n, width, n_gates);
    if (rc) return rc;
    std::vector<uint64_t> masks;
    rc = gate_masks(n, width, n_gates, gate_k, gate_qubits, masks);
    if (rc) return rc;
    int64_t block = 0, launches = -1;
    simulate_recorder(n, width, streaming != 0, true, n_gates, gate_k, masks,
                      [&](const std::vector<int64_t> &, bool starts) {
                          if (starts) ++launches;
                          launch_of_block[block++] = launches;
                      },
                      nullptr);
    *n_blocks = block;
    return QIPGPU_OK;
}

}  // extern "C"
