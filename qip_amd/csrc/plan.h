// plan.h -- the block plan of the precision form: gates in program order, grouped into blocks of at
// most `width` qubits that are applied one after another, each as one dense operator.
//
// Light-cone rule. A gate waits for the previous gate on each of its qubits; a gate is ready once
// all of those are in blocks. A block starts from the earliest ready gate, with qubit set S = its
// qubits, and alternates two steps until no ready gate fits beside S:
//   absorb: every ready gate whose qubits lie in S joins the block (which can make later gates ready);
//   grow:   each ready gate g that keeps S within the width is scored by a two-step look-ahead: the
//           number of gates the block would absorb with S u qubits(g) per qubit added, plus half of
//           the same for the best further gate after that. S takes the qubits of the best-scoring
//           gate that shares a qubit with S, unless the best of all scores more than 1.5 times as
//           high (or none shares one); ties go to the earliest gate. Scores are exact fractions.
// The next block starts from what is left. Each block's gates, in program order, compose to its
// operator; the order of the blocks respects every dependency. Growing along shared qubits follows
// the light cone of the gates already in the block, so on brickwork circuits a block covers a diamond
// of the circuit instead of a few layers of one window (on the 30-qubit depth-40 circuit: 66 blocks of
// six qubits, where adding each gate to the first compatible block in program order gives 128); the
// look-ahead lets blocks of QFT-like circuits take the qubits their next gates need (QFT on 24 qubits:
// 28 blocks instead of 48).
//
// Final blocks. A gate recorded later can change the plan only if it enters it: as a ready gate,
// which needs every one of its qubits to be without a recorded gate still outside the blocks, or
// inside a look-ahead that used up the recorded gates of one of the look-ahead's qubits. So a block
// is final (no later gate changes it or any block before it) if every qubit of the state still has a
// recorded gate outside the blocks and no look-ahead of the block used one up. Launching the final
// leading blocks while the rest is recorded gives the blocks, and the results, of planning the whole
// record at once.
#pragma once

#include <algorithm>
#include <cstdint>
#include <unordered_map>
#include <utility>
#include <vector>

namespace qipgpu {

struct PlannedBlock {
    uint64_t mask = 0;                // qubit set (bit q = qubit q)
    std::vector<int64_t> gates;       // indices of its gates, ascending (program order)
};

// Plan the gates [0, count) given by their qubit masks (n <= 64 qubits, each gate within `width`).
// Returns the number of leading final blocks; with stop_at_open the plan ends after the first block
// that is not final.
inline int64_t plan_blocks(const uint64_t *masks, int64_t count, int width, int n, bool stop_at_open,
                           std::vector<PlannedBlock> &blocks) {
    blocks.clear();
    if (count <= 0) return 0;
    // dependencies: up to six distinct predecessors per gate, successors in CSR form
    std::vector<int64_t> pred(6 * (size_t)count), sstart((size_t)count + 1, 0), succ;
    std::vector<int32_t> npred((size_t)count, 0);
    int64_t last[64], open_on[64];
    for (int q = 0; q < 64; ++q) {
        last[q] = -1;
        open_on[q] = 0;
    }
    for (int64_t g = 0; g < count; ++g) {
        int np = 0;
        for (uint64_t m = masks[g]; m; m &= m - 1) {
            const int q = __builtin_ctzll(m);
            const int64_t p = last[q];
            if (p >= 0 && std::find(&pred[6 * g], &pred[6 * g] + np, p) == &pred[6 * g] + np) {
                pred[6 * g + np++] = p;
                ++sstart[p + 1];
            }
            last[q] = g;
            ++open_on[q];
        }
        npred[g] = np;
    }
    for (int64_t g = 0; g < count; ++g) sstart[g + 1] += sstart[g];
    succ.resize((size_t)sstart[count]);
    {
        std::vector<int64_t> fill(sstart.begin(), sstart.end() - 1);
        for (int64_t g = 0; g < count; ++g)
            for (int j = 0; j < npred[g]; ++j) succ[fill[pred[6 * g + j]]++] = g;
    }
    // ready gates have pairwise disjoint qubits, so there are at most n of them: a sorted vector
    std::vector<int64_t> ready;
    for (int64_t g = 0; g < count; ++g)
        if (npred[g] == 0) ready.push_back(g);
    auto ready_insert = [&](int64_t g) { ready.insert(std::lower_bound(ready.begin(), ready.end(), g), g); };
    auto ready_erase = [&](int64_t g) { ready.erase(std::lower_bound(ready.begin(), ready.end(), g)); };
    std::vector<int64_t> work;
    // look-ahead scratch: a gate's pending predecessor count is sim_pred[g] when stamp[g] == cur
    std::vector<int32_t> sim_pred((size_t)count);
    std::vector<uint32_t> stamp((size_t)count, 0);
    uint32_t cur = 0;

    // absorb every ready gate within S, transitively, into the block
    auto absorb = [&](uint64_t S, PlannedBlock &b) {
        int64_t c = 0;
        work.clear();
        for (int64_t g : ready)
            if ((masks[g] & ~S) == 0) work.push_back(g);
        while (!work.empty()) {
            const int64_t g = work.back();
            work.pop_back();
            ready_erase(g);
            b.gates.push_back(g);
            ++c;
            for (uint64_t m = masks[g]; m; m &= m - 1) --open_on[__builtin_ctzll(m)];
            for (int64_t i = sstart[g]; i < sstart[g + 1]; ++i) {
                const int64_t t = succ[i];
                if (--npred[t] == 0) {
                    ready_insert(t);
                    if ((masks[t] & ~S) == 0) work.push_back(t);
                }
            }
        }
        return c;
    };
    // how many gates absorb(S) would take; `kept` = every qubit of S keeps a recorded gate outside;
    // `left` (if given) receives the gates that would be ready outside S afterwards
    auto look_ahead = [&](uint64_t S, bool &kept, std::vector<int64_t> *left) {
        ++cur;
        int64_t used[64];
        uint64_t touched = 0;     // qubits whose used[] count is live
        int64_t c = 0;
        work.clear();
        if (left) left->clear();
        for (int64_t g : ready) {
            if ((masks[g] & ~S) == 0) work.push_back(g);
            else if (left) left->push_back(g);
        }
        while (!work.empty()) {
            const int64_t g = work.back();
            work.pop_back();
            ++c;
            for (uint64_t m = masks[g]; m; m &= m - 1) {
                const int q = __builtin_ctzll(m);
                if (!(touched >> q & 1)) {
                    touched |= 1ull << q;
                    used[q] = 0;
                }
                ++used[q];
            }
            for (int64_t i = sstart[g]; i < sstart[g + 1]; ++i) {
                const int64_t t = succ[i];
                if (stamp[t] != cur) {
                    stamp[t] = cur;
                    sim_pred[t] = npred[t];
                }
                if (--sim_pred[t] == 0) {
                    if ((masks[t] & ~S) == 0) work.push_back(t);
                    else if (left) left->push_back(t);
                }
            }
        }
        kept = true;
        for (uint64_t m = S; m; m &= m - 1) {
            const int q = __builtin_ctzll(m);
            kept = kept && open_on[q] - (touched >> q & 1 ? used[q] : 0) > 0;
        }
        return c;
    };
    // look_ahead without `left`, memoised by S while the placed gates do not change (the second steps
    // of different candidates often reach the same set)
    std::unordered_map<uint64_t, std::pair<int64_t, bool>> memo;
    auto look_ahead_memo = [&](uint64_t S, bool &kept) {
        const auto it = memo.find(S);
        if (it != memo.end()) {
            kept = it->second.second;
            return it->second.first;
        }
        const int64_t c = look_ahead(S, kept, nullptr);
        memo.emplace(S, std::make_pair(c, kept));
        return c;
    };
    // Look-ahead score of growing S by gate g as the fraction num / den: the gates absorbed per qubit
    // added, plus half of the same for the best further step; `kept` as in look_ahead, over every
    // simulation made.
    std::vector<int64_t> left;
    auto score = [&](uint64_t S, int64_t g, int64_t &num, int64_t &den, bool &kept) {
        const uint64_t S2 = S | masks[g];
        const int64_t a = __builtin_popcountll(S2) - __builtin_popcountll(S);
        const int64_t c = look_ahead(S2, kept, &left);
        memo.emplace(S2, std::make_pair(c, kept));
        int64_t d = 0, a2 = 1;
        for (int64_t h : left) {
            const uint64_t S3 = S2 | masks[h];
            if (__builtin_popcountll(S3) > width) continue;
            bool kept3;
            const int64_t c3 = look_ahead_memo(S3, kept3);
            kept = kept && kept3;
            const int64_t a3 = __builtin_popcountll(S3) - __builtin_popcountll(S2);
            if ((c3 - c) * a2 > d * a3) {
                d = c3 - c;
                a2 = a3;
            }
        }
        num = 2 * c * a2 + a * d;
        den = 2 * a * a2;
    };

    int64_t absorbed = 0, final_prefix = 0;
    bool prefix_final = true;
    while (absorbed < count) {
        PlannedBlock b;
        uint64_t S = masks[ready.front()];
        bool exhausted = false;   // a look-ahead used up some qubit's recorded gates
        for (;;) {
            absorbed += absorb(S, b);
            memo.clear();
            int64_t ga = -1, na = 0, da = 1;   // best of all: gate, score numerator, denominator
            int64_t gs = -1, ns = 0, ds = 1;   // best sharing a qubit with S
            for (int64_t g : ready) {
                if (__builtin_popcountll(S | masks[g]) > width) continue;
                int64_t num, den;
                bool kept;
                score(S, g, num, den, kept);
                exhausted = exhausted || !kept;
                if (ga < 0 || num * da > na * den) {
                    ga = g;
                    na = num;
                    da = den;
                }
                if ((masks[g] & S) && (gs < 0 || num * ds > ns * den)) {
                    gs = g;
                    ns = num;
                    ds = den;
                }
            }
            if (ga < 0) break;
            S |= (gs < 0 || 2 * na * ds > 3 * ns * da) ? masks[ga] : masks[gs];
        }
        std::sort(b.gates.begin(), b.gates.end());
        b.mask = S;
        bool fin = !exhausted;
        for (int q = 0; q < n; ++q) fin = fin && open_on[q] > 0;
        blocks.push_back(std::move(b));
        prefix_final = prefix_final && fin;
        if (prefix_final) ++final_prefix;
        if (stop_at_open && !prefix_final) break;
    }
    return final_prefix;
}

}  // namespace qipgpu
