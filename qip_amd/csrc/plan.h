// plan.h -- the block plan of the precision form: gates in program order, grouped into blocks of at
// most `width` qubits that are applied one after another, each as one dense operator.
//
// Light-cone rule. A gate waits for the previous gate on each of its qubits; a gate is ready once
// all of those are in blocks. A block starts from the earliest ready gate, with qubit set S = its
// qubits, and alternates two steps until neither changes anything:
//   absorb: every ready gate whose qubits lie in S joins the block (which can make later gates ready);
//   grow:   S takes the qubits of the earliest ready gate that shares a qubit with S and keeps S within
//           the width, or, if no such gate exists, of the earliest ready gate that keeps S within it.
// The next block starts from what is left. Each block's gates, in program order, compose to its
// operator; the order of the blocks respects every dependency. Growing along shared qubits follows
// the light cone of the gates already in the block, so on brickwork circuits a block covers a diamond
// of the circuit instead of a few layers of one window (on the 30-qubit depth-40 circuit: 77 blocks of
// six qubits, where adding each gate to the first compatible block in program order gives 128).
//
// Final blocks. A gate recorded later can change the plan only if it becomes ready while a block is
// being planned, which needs every one of its qubits to be without a recorded gate still outside the
// blocks. So a block is final (no later gate changes it or any block before it) if all of its qubits
// still have such an open gate, and either it holds `width` qubits (no later gate fits its growth) or
// every qubit of the state has one (no later gate is ready at all). Launching the final leading blocks
// while the rest is recorded gives the blocks, and the results, of planning the whole record at once.
#pragma once

#include <algorithm>
#include <cstdint>
#include <set>
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
    const int full = std::min(width, n);
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
    // ready gates have pairwise disjoint qubits, so there are at most n of them
    std::set<int64_t> ready;
    for (int64_t g = 0; g < count; ++g)
        if (npred[g] == 0) ready.insert(g);
    int64_t absorbed = 0, final_prefix = 0;
    bool prefix_final = true;
    std::vector<int64_t> work;
    while (absorbed < count) {
        PlannedBlock b;
        uint64_t S = masks[*ready.begin()];
        for (;;) {
            work.clear();
            for (int64_t g : ready)
                if ((masks[g] & ~S) == 0) work.push_back(g);
            while (!work.empty()) {
                const int64_t g = work.back();
                work.pop_back();
                ready.erase(g);
                b.gates.push_back(g);
                ++absorbed;
                for (uint64_t m = masks[g]; m; m &= m - 1) --open_on[__builtin_ctzll(m)];
                for (int64_t i = sstart[g]; i < sstart[g + 1]; ++i) {
                    const int64_t t = succ[i];
                    if (--npred[t] == 0) {
                        ready.insert(t);
                        if ((masks[t] & ~S) == 0) work.push_back(t);
                    }
                }
            }
            int64_t pick = -1;
            for (int64_t g : ready)
                if ((masks[g] & S) && __builtin_popcountll(S | masks[g]) <= width) {
                    pick = g;
                    break;
                }
            if (pick < 0)
                for (int64_t g : ready)
                    if (__builtin_popcountll(S | masks[g]) <= width) {
                        pick = g;
                        break;
                    }
            if (pick < 0) break;
            S |= masks[pick];
        }
        std::sort(b.gates.begin(), b.gates.end());
        b.mask = S;
        bool fin = true;
        for (uint64_t m = S; m; m &= m - 1) fin = fin && open_on[__builtin_ctzll(m)] > 0;
        if (fin && __builtin_popcountll(S) != full)
            for (int q = 0; q < n; ++q) fin = fin && open_on[q] > 0;
        blocks.push_back(std::move(b));
        prefix_final = prefix_final && fin;
        if (prefix_final) ++final_prefix;
        if (stop_at_open && !prefix_final) break;
    }
    return final_prefix;
}

}  // namespace qipgpu
