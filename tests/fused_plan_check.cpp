// Host-side check of the solve work lists (BuildQueues and BuildSolveTiles
// in psp_amd/csrc/symbolic.cpp) for the grid given on the command line:
//   fused_plan_check nx ny nz [queues rows cut wgPerQueue persist maxSplit]
// For every panel class and phase it checks that
//  - every node at or below the cut has each of its row blocks in exactly
//    one queue, as one tile or as slices that partition the row block's
//    inner range, with distinct partial slots below maxSlots;
//  - every index a kernel reads (node, slot, reduce record, queue range)
//    is inside its array;
//  - a tile's dependencies are the nodes the method needs (forward: the
//    children with rows; backward: the nearest ancestor with pivots, when it
//    is at or below the cut), each with its number of row blocks, and every
//    row block of a dependency completes earlier in the same queue;
//  - the per-level accounting covers the same levels;
//  - the one-launch-per-level lists placed per XCD keep each split row
//    block's slices on one XCD (checkLevels).
// Prints lines per class and exits nonzero on the first violation.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "symbolic.hpp"

using namespace psp_amd;

static int fail(const char* what, int cls, int ph, int a = 0, int b = 0)
{
    std::fprintf(stderr, "FAIL class %d phase %d: %s (%d, %d)\n", cls, ph, what, a, b);
    return 1;
}

static int rowsOf(int ph, const Node& n) { return ph == 0 ? n.size + n.usize : n.size; }

static int check(const ClassPlan& C, int cls)
{
    if (C.qcut < 0) {
        std::printf("class %d: no fused launch\n", cls);
        return 0;
    }
    const int nn = (int)C.nodes.size(), nq = C.nqueues, R = C.queueRows;
    for (int ph = 0; ph < 2; ++ph) {
        const auto& Q = C.queue[ph];
        if ((int)Q.qstart.size() != nq + 1 || Q.qstart[0] != 0 || Q.qstart[nq] != (int)Q.tiles.size())
            return fail("queue starts", cls, ph);
        if ((int)Q.dep.size() != nn) return fail("dependency table size", cls, ph);
        std::vector<int> queueOfNode(nn, -1);
        std::map<std::pair<int, int>, int> blockDonePos;  // (node, row0) -> tile index completing it
        std::vector<int> seenSlot(C.maxSlots, 0);
        std::map<int, std::vector<std::pair<int, int>>> sliceCols;  // red -> col ranges
        for (int q = 0; q < nq; ++q) {
            if (Q.qstart[q] > Q.qstart[q + 1]) return fail("queue order", cls, ph, q);
            std::map<int, int> redArrived;
            for (int i = Q.qstart[q]; i < Q.qstart[q + 1]; ++i) {
                const QTile& t = Q.tiles[i];
                if (t.node < 0 || t.node >= nn) return fail("node index", cls, ph, i, t.node);
                const Node& n = C.nodes[t.node];
                if (n.depth < C.qcut) return fail("node above the cut", cls, ph, t.node);
                if (queueOfNode[t.node] >= 0 && queueOfNode[t.node] != q)
                    return fail("node in two queues", cls, ph, t.node);
                queueOfNode[t.node] = q;
                if (t.row0 < 0 || t.row0 % R || t.nrows < 1 || t.nrows > R ||
                    t.row0 + t.nrows > rowsOf(ph, n))
                    return fail("row block", cls, ph, i, t.row0);
                const int lo = ph == 0 ? 0 : t.row0;
                const int hi = ph == 0 ? std::min(n.size, t.row0 + R) : n.size + n.usize;
                if (t.col0 < lo || t.col1 > std::max(lo, hi) || t.col0 > t.col1)
                    return fail("inner range", cls, ph, i, t.col0);
                // dependencies come first in this queue
                const auto& d = Q.dep[t.node];
                for (int j = 0; j < 2; ++j) {
                    if (d[j] < 0) continue;
                    if (d[j] >= nn) return fail("dependency index", cls, ph, t.node, d[j]);
                    int blocks = 0;
                    for (int r0 = 0; r0 < rowsOf(ph, C.nodes[d[j]]); r0 += R) {
                        auto it = blockDonePos.find({d[j], r0});
                        if (it == blockDonePos.end() || it->second >= i)
                            return fail("dependency not done earlier in the queue", cls, ph,
                                        t.node, d[j]);
                        if (queueOfNode[d[j]] != q) return fail("dependency in another queue", cls, ph);
                        ++blocks;
                    }
                    if (blocks != d[2 + j]) return fail("dependency row-block count", cls, ph, t.node);
                }
                if (t.red < 0) {
                    if (t.slot != -1) return fail("slot without reduce", cls, ph, i);
                    if (t.col0 != lo || t.col1 != std::max(lo, hi)) return fail("unsplit range", cls, ph, i);
                    if (!blockDonePos.emplace(std::make_pair(t.node, t.row0), i).second)
                        return fail("row block twice", cls, ph, t.node, t.row0);
                } else {
                    if (t.red >= (int)Q.reds.size()) return fail("reduce index", cls, ph, i, t.red);
                    const QRed& r = Q.reds[t.red];
                    if (r.node != t.node || r.row0 != t.row0 || r.nrows != t.nrows)
                        return fail("reduce record", cls, ph, i);
                    if (t.slot < r.slot0 || t.slot >= r.slot0 + r.nslots || t.slot >= C.maxSlots)
                        return fail("slot", cls, ph, i, t.slot);
                    if (seenSlot[t.slot]++) return fail("slot twice", cls, ph, t.slot);
                    sliceCols[t.red].push_back({t.col0, t.col1});
                    if (++redArrived[t.red] == r.nslots)
                        if (!blockDonePos.emplace(std::make_pair(t.node, t.row0), i).second)
                            return fail("row block twice", cls, ph, t.node, t.row0);
                }
            }
        }
        for (auto& kv : sliceCols) {
            const QRed& r = Q.reds[kv.first];
            if ((int)kv.second.size() != r.nslots) return fail("slices per reduce", cls, ph, kv.first);
            const Node& n = C.nodes[r.node];
            const int lo = ph == 0 ? 0 : r.row0;
            const int hi = ph == 0 ? std::min(n.size, r.row0 + R) : n.size + n.usize;
            int at = lo;
            for (auto& c : kv.second) {
                if (c.first != at) return fail("slices do not tile the range", cls, ph, kv.first);
                at = c.second;
            }
            if (at != hi) return fail("slices end", cls, ph, kv.first);
        }
        // every row block of every fused node, and the expected dependencies
        for (int t = 0; t < nn; ++t) {
            const Node& n = C.nodes[t];
            if (n.depth < C.qcut) continue;
            const bool wanted = ph == 0 ? n.size + n.usize > 0 : n.size > 0;
            for (int r0 = 0; wanted && r0 < rowsOf(ph, n); r0 += R)
                if (!blockDonePos.count({t, r0})) return fail("row block missing", cls, ph, t, r0);
            std::set<int> want;
            if (ph == 0) {
                for (int c = 0; c < n.nchild; ++c)
                    if (n.child[c] >= 0 && rowsOf(0, C.nodes[n.child[c]]) > 0) want.insert(n.child[c]);
            } else {
                int a = n.parent;
                while (a >= 0 && C.nodes[a].size == 0) a = C.nodes[a].parent;
                if (a >= 0 && C.nodes[a].depth >= C.qcut) want.insert(a);
            }
            std::set<int> got;
            for (int j = 0; j < 2; ++j)
                if (Q.dep[t][j] >= 0) got.insert(Q.dep[t][j]);
            if (got != want) return fail("dependency set", cls, ph, t);
        }
    }
    std::printf("class %d: cut %d, %d queues, %zu/%zu tiles, %zu/%zu split row blocks, %d slots\n",
                cls, C.qcut, nq, C.queue[0].tiles.size(), C.queue[1].tiles.size(),
                C.queue[0].reds.size(), C.queue[1].reds.size(), C.maxSlots);
    return 0;
}

// The one-launch-per-level tile lists: a level whose tiles were placed per
// XCD has a multiple of 8 tiles, holes (no rows, no slot) only there, and the
// slices of each split row block at positions equal mod 8, each slot of the
// block exactly once; other levels have no holes.
static int checkLevels(const ClassPlan& C, int cls)
{
    int placed = 0;
    for (int d = 0; d <= C.maxDepth; ++d)
        for (int ph = 0; ph < 2; ++ph) {
            const ClassPlan::Range& r = ph == 0 ? C.solveLevels[d].f : C.solveLevels[d].b;
            if (r.start < 0 || r.start + r.count > (int)C.mv.size())
                return fail("level range", cls, ph, d);
            if (r.xcd && (r.persist || r.count % 8))
                return fail("placed level shape", cls, ph, d);
            placed += r.xcd;
            std::map<int, int> residue;
            std::map<int, std::set<int>> slotsSeen;
            for (int i = r.start; i < r.start + r.count; ++i) {
                const MvTile& m = C.mv[i];
                if (m.nrows <= 0) {
                    if (!r.xcd || m.slot != -1 || m.red != -1) return fail("hole", cls, ph, d, i);
                    continue;
                }
                if (m.red < 0) continue;
                if (m.red < r.rstart || m.red >= r.rstart + r.rcount)
                    return fail("reduce item out of the level", cls, ph, d, m.red);
                const MvTile& it = C.mvred[m.red];
                if (m.slot < it.col0 || m.slot >= it.col0 + it.col1 || m.node != it.node ||
                    m.row0 != it.row0)
                    return fail("slice of another row block", cls, ph, d, i);
                if (!slotsSeen[m.red].insert(m.slot).second) return fail("slot twice", cls, ph, d, m.slot);
                const int res = (i - r.start) % 8;
                if (r.xcd) {
                    auto ins = residue.emplace(m.red, res);
                    if (ins.first->second != res) return fail("slices on two XCDs", cls, ph, d, m.red);
                }
            }
            for (int j = r.rstart; j < r.rstart + r.rcount; ++j)
                if ((int)slotsSeen[j].size() != C.mvred[j].col1)
                    return fail("row block slices", cls, ph, d, j);
        }
    std::printf("class %d: %d levels placed per XCD\n", cls, placed);
    return 0;
}

// Schur-complement workspace (Schedule): the updates of each depth lie in
// the region of that depth's parity (even: [0, R0), odd: [R0, ustore)), the
// level's range [steps[d].uoff, + ucount) holds exactly its nodes' updates,
// without overlap, and the extend-add items name the child's and the
// parent's update offsets.  Two consecutive depths then never share
// entries, so a level's updates survive until its parent level reads them.
static int checkUpdates(const ClassPlan& C, int cls)
{
    const int64_t r0 = C.maxDepth >= 1 ? C.steps[1].uoff : C.ustore;
    for (int d = 0; d <= C.maxDepth; ++d) {
        const ClassPlan::LevelSteps& S = C.steps[d];
        const int64_t lo = (d & 1) ? r0 : 0, hi = (d & 1) ? C.ustore : r0;
        if (S.uoff != lo || S.uoff + S.ucount > hi) return fail("level update range", cls, 0, d);
        std::vector<std::pair<int64_t, int64_t>> rs;
        int64_t sum = 0;
        for (const Node& n : C.nodes)
            if (n.depth == d && n.usize > 0) {
                const int64_t e = (int64_t)n.usize * n.usize;
                rs.push_back({n.uoff, n.uoff + e});
                sum += e;
            }
        std::sort(rs.begin(), rs.end());
        for (size_t i = 0; i < rs.size(); ++i) {
            if (rs[i].first < S.uoff || rs[i].second > S.uoff + S.ucount)
                return fail("update outside its level", cls, 0, d, (int)i);
            if (i && rs[i].first < rs[i - 1].second) return fail("updates overlap", cls, 0, d, (int)i);
        }
        if (sum != S.ucount) return fail("level update size", cls, 0, d);
    }
    for (const ExtAddItem& it : C.extadd) {
        bool found = false;
        for (const Node& p : C.nodes)
            for (int sl = 0; sl < 2 && !found; ++sl) {
                const int c = p.child[sl];
                if (c >= 0 && C.nodes[c].uoff == it.uChild && p.uoff == it.uParent &&
                    p.loff == it.lParent && C.nodes[c].usize == it.uc)
                    found = true;
            }
        if (!found) return fail("extend-add offsets", cls, 0, it.col0);
    }
    std::printf("class %d: updates in two regions of %lld + %lld entries\n", cls, (long long)r0,
                (long long)(C.ustore - r0));
    return 0;
}

int main(int argc, char** argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s nx ny nz [queues rows cut wgPerQueue persist maxSplit]\n",
                     argv[0]);
        return 2;
    }
    psp_hip_desc d{};
    d.nx = std::atoi(argv[1]);
    d.ny = std::atoi(argv[2]);
    d.nz = std::atoi(argv[3]);
    const double h = 1.0 / (std::max(d.nx, std::max(d.ny, d.nz)) + 1);
    d.wx = (d.nx + 1) * h;
    d.wy = (d.ny + 1) * h;
    d.wz = (d.nz + 1) * h;
    d.omega = 30.0;
    d.bx = d.by = d.bz = 5;
    d.sigmax = d.sigmay = d.sigmaz = 1.5;
    d.damping = 7.0;
    d.planesPerPanel = 4;
    d.cutoff = 12;
    Plan plan;
    plan.fused = true;
    if (argc > 4) plan.fusedQueues = std::atoi(argv[4]);
    if (argc > 5) plan.fusedRows = std::atoi(argv[5]);
    if (argc > 6) plan.fusedCut = std::atoi(argv[6]);
    if (argc > 7) plan.fusedWgPerQueue = std::atoi(argv[7]);
    if (argc > 8) plan.persist = std::atoi(argv[8]) != 0;
    if (argc > 9) plan.maxSplit = std::atoi(argv[9]);
    std::string msg;
    if (BuildPlan(d, plan, msg) != PSP_HIP_OK) {
        std::fprintf(stderr, "plan: %s\n", msg.c_str());
        return 2;
    }
    for (int c = 0; c < 3; ++c)
        if (plan.present[c] && (check(plan.cls[c], c) || checkLevels(plan.cls[c], c) ||
                                checkUpdates(plan.cls[c], c)))
            return 1;
    return 0;
}
