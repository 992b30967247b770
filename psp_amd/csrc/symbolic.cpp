// Host symbolic analysis and static GPU work lists.  See symbolic.hpp.
#include "symbolic.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stack>

namespace psp_amd {

// --------------------------------------------------------------------------
// Geometry: DistHelmholtz ctor panel split (main.hpp:35-57) and the panel
// helpers PanelPadding/PanelDepth/PanelV (main.hpp:728-789).
// --------------------------------------------------------------------------
int Geometry::init(const psp_hip_desc& d)
{
    nx = d.nx;
    ny = d.ny;
    nz = d.nz;
    bz = d.bz;
    P = d.planesPerPanel;
    cutoff = d.cutoff;
    topPML = d.topBC == PSP_HIP_PML;
    bottomDepth = bz + P;
    topOrigDepth = topPML ? bz + P : P;
    if (nz <= bottomDepth + topOrigDepth) return PSP_HIP_ERR_SHALLOW;
    innerDepth = nz - (bottomDepth + topOrigDepth);
    leftoverDepth = innerDepth % P;
    haveLeftover = leftoverDepth != 0;
    numFullInner = innerDepth / P;
    numPanels = 1 + numFullInner + (haveLeftover ? 1 : 0) + 1;
    return PSP_HIP_OK;
}

int Geometry::depth(int p) const
{
    if (p == 0) return bottomDepth;
    if (p < 1 + numFullInner) return P;
    if (haveLeftover && p == 1 + numFullInner) return leftoverDepth;
    return topOrigDepth;
}

int Geometry::panelV(int p) const
{
    if (p == 0) return 0;
    if (p < 1 + numFullInner) return bottomDepth + P * (p - 1);
    if (haveLeftover && p == 1 + numFullInner) return bottomDepth + P * numFullInner;
    return bottomDepth + P * numFullInner + leftoverDepth;
}

int Geometry::cls(int p) const
{
    if (p < 1 + numFullInner) return 0;
    if (haveLeftover && p == 1 + numFullInner) return 1;
    return 2;
}

int Geometry::classVSize(int c) const
{
    if (c == 0) return bottomDepth;  // == P + bz for the full inner panels
    if (c == 1) return leftoverDepth + bz;
    return topOrigDepth + bz;
}

// --------------------------------------------------------------------------
// ReorderedIndexRecursion (main.hpp:1441-1477) with depthTilSerial = 0:
// leaves are x-fastest boxes; otherwise the longer side is cut at
// (len-1)/2 and the index is left part, right part, then the separator.
// --------------------------------------------------------------------------
int NestedIndex(int x, int y, int vl, int xs, int ys, int vs, int cutoff)
{
    if (xs * ys <= cutoff) return x + y * xs + vl * xs * ys;
    if (xs >= ys) {
        const int mid = (xs - 1) / 2;
        if (x < mid) return NestedIndex(x, y, vl, mid, ys, vs, cutoff);
        if (x == mid) return std::max(xs - 1, 0) * ys * vs + y + vl * ys;
        return mid * ys * vs +
               NestedIndex(x - mid - 1, y, vl, std::max(xs - mid - 1, 0), ys, vs, cutoff);
    }
    const int mid = (ys - 1) / 2;
    if (y < mid) return NestedIndex(x, y, vl, xs, mid, vs, cutoff);
    if (y == mid) return xs * std::max(ys - 1, 0) * vs + x + vl * xs;
    return xs * mid * vs +
           NestedIndex(x, y - mid - 1, vl, xs, std::max(ys - mid - 1, 0), vs, cutoff);
}

namespace {

struct BoxRec {  // environment.hpp:117-121
    int parent, nx, ny, xo, yo;
    bool left;
};

int CountNodes(int bnx, int bny, int cutoff)  // NumLocalNodesRecursion
{
    if (bnx * bny <= cutoff) return 1;
    if (bnx >= bny) {
        const int mid = (bnx - 1) / 2;
        return 1 + CountNodes(mid, bny, cutoff) +
               CountNodes(std::max(bnx - mid - 1, 0), bny, cutoff);
    }
    const int mid = (bny - 1) / 2;
    return 1 + CountNodes(bnx, mid, cutoff) +
           CountNodes(bnx, std::max(bny - mid - 1, 0), cutoff);
}

int64_t CeilDiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Elimination tree of one panel class (FillPanelLocalElimTree,
// main.hpp:1709-1900), built with the same box stack so node s runs from
// the root (last) down; this yields a post-order numbering.
void BuildTree(const Geometry& g, ClassPlan& C, std::vector<std::vector<int>>& orig)
{
    const int nx = g.nx, ny = g.ny, vs = C.vsize, cut = g.cutoff;
    const int nn = CountNodes(nx, ny, cut);
    C.nodes.assign(nn, Node());
    orig.assign(nn, {});
    auto nd = [&](int x, int y, int v) { return NestedIndex(x, y, v, nx, ny, vs, cut); };
    std::stack<BoxRec> boxes;
    boxes.push({-1, nx, ny, 0, 0, false});
    std::vector<int> depthOf(nn, 0);
    for (int s = nn - 1; s >= 0; --s) {
        const BoxRec b = boxes.top();
        boxes.pop();
        Node& n = C.nodes[s];
        n.parent = b.parent;
        n.depth = b.parent < 0 ? 0 : C.nodes[b.parent].depth + 1;
        if (b.parent >= 0) {
            Node& p = C.nodes[b.parent];
            p.child[b.left ? 0 : 1] = s;
        }
        std::vector<int>& os = orig[s];
        if (b.nx * b.ny <= cut) {
            n.size = b.nx * b.ny * vs;
            n.off = n.size ? nd(b.xo, b.yo, 0) : 0;
            if (b.xo > 0)
                for (int i = 0; i < vs; ++i)
                    for (int j = 0; j < b.ny; ++j) os.push_back(nd(b.xo - 1, b.yo + j, i));
            if (b.xo + b.nx < nx)
                for (int i = 0; i < vs; ++i)
                    for (int j = 0; j < b.ny; ++j) os.push_back(nd(b.xo + b.nx, b.yo + j, i));
            if (b.yo > 0)
                for (int i = 0; i < vs; ++i)
                    for (int j = 0; j < b.nx; ++j) os.push_back(nd(b.xo + j, b.yo - 1, i));
            if (b.yo + b.ny < ny)
                for (int i = 0; i < vs; ++i)
                    for (int j = 0; j < b.nx; ++j) os.push_back(nd(b.xo + j, b.yo + b.ny, i));
        } else if (b.nx >= b.ny) {
            const int mid = (b.nx - 1) / 2;
            n.size = b.ny * vs;
            n.off = nd(b.xo + mid, b.yo, 0);
            if (b.yo > 0)
                for (int i = 0; i < vs; ++i) os.push_back(nd(b.xo + mid, b.yo - 1, i));
            if (b.yo + b.ny < ny)
                for (int i = 0; i < vs; ++i) os.push_back(nd(b.xo + mid, b.yo + b.ny, i));
            boxes.push({s, mid, b.ny, b.xo, b.yo, true});
            boxes.push({s, std::max(b.nx - mid - 1, 0), b.ny, b.xo + mid + 1, b.yo, false});
        } else {
            const int mid = (b.ny - 1) / 2;
            n.size = b.nx * vs;
            n.off = nd(b.xo, b.yo + mid, 0);
            if (b.xo > 0)
                for (int i = 0; i < vs; ++i) os.push_back(nd(b.xo - 1, b.yo + mid, i));
            if (b.xo + b.nx < nx)
                for (int i = 0; i < vs; ++i) os.push_back(nd(b.xo + b.nx, b.yo + mid, i));
            boxes.push({s, b.nx, mid, b.xo, b.yo, true});
            boxes.push({s, b.nx, std::max(b.ny - mid - 1, 0), b.xo, b.yo + mid + 1, false});
        }
        std::sort(os.begin(), os.end());
    }
    for (Node& n : C.nodes) n.nchild = (n.child[0] >= 0) + (n.child[1] >= 0);
}

// Symbolic factorization: struct(t) = (orig(t) U struct(children)) \ own(t).
void Analyse(ClassPlan& C, const std::vector<std::vector<int>>& orig)
{
    const int nn = (int)C.nodes.size();
    std::vector<std::vector<int>> st(nn);
    for (int t = 0; t < nn; ++t) {
        Node& n = C.nodes[t];
        std::vector<int> acc = orig[t];
        for (int k = 0; k < 2; ++k) {
            if (n.child[k] < 0) continue;
            std::vector<int> merged;
            const std::vector<int>& cs = st[n.child[k]];
            std::set_union(acc.begin(), acc.end(), cs.begin(), cs.end(),
                           std::back_inserter(merged));
            acc.swap(merged);
        }
        acc.erase(std::unique(acc.begin(), acc.end()), acc.end());
        std::vector<int> out;
        out.reserve(acc.size());
        for (int gidx : acc)
            if (gidx < n.off || gidx >= n.off + n.size) out.push_back(gidx);
        st[t].swap(out);
    }
    int64_t total = 0;
    for (int t = 0; t < nn; ++t) {
        C.nodes[t].ustart = total;
        C.nodes[t].usize = (int)st[t].size();
        total += C.nodes[t].usize;
    }
    C.structs.resize(total);
    C.relinds.assign(total, -1);
    for (int t = 0; t < nn; ++t)
        std::copy(st[t].begin(), st[t].end(), C.structs.begin() + C.nodes[t].ustart);
    for (int t = 0; t < nn; ++t) {
        const Node& n = C.nodes[t];
        if (n.parent < 0) continue;
        const Node& p = C.nodes[n.parent];
        const int* ps = C.structs.data() + p.ustart;
        for (int a = 0; a < n.usize; ++a) {
            const int gidx = C.structs[n.ustart + a];
            int rel;
            if (gidx >= p.off && gidx < p.off + p.size) {
                rel = gidx - p.off;
            } else {
                const int* f = std::lower_bound(ps, ps + p.usize, gidx);
                rel = (f != ps + p.usize && *f == gidx) ? p.size + int(f - ps) : -1;
            }
            C.relinds[n.ustart + a] = rel;
        }
    }
    C.maxDepth = 0;
    for (const Node& n : C.nodes) C.maxDepth = std::max(C.maxDepth, n.depth);
    C.levels.assign(C.maxDepth + 1, {});
    for (int t = 0; t < nn; ++t) C.levels[C.nodes[t].depth].push_back(t);
}

// Front-assembly metadata: for every padded unknown its coordinates, node,
// column, and the front rows of its strictly-lower couplings
// (FormLowerColumnOfNode, initialize_finalize.hpp:589-806).
void AssemblyMeta(const Geometry& g, ClassPlan& C)
{
    const int nx = g.nx, ny = g.ny, vs = C.vsize, cut = g.cutoff;
    const int n = C.n;
    C.qxy.assign(n, 0);
    C.qv.assign(n, 0);
    C.qnode.assign(n, -1);
    C.qcol.assign(n, 0);
    C.qrows.assign((size_t)n * 6, -1);
    for (int v = 0; v < vs; ++v)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const int q = NestedIndex(x, y, v, nx, ny, vs, cut);
                C.qxy[q] = x | (y << 16);
                C.qv[q] = v;
            }
    for (int t = 0; t < (int)C.nodes.size(); ++t) {
        const Node& nd = C.nodes[t];
        for (int j = 0; j < nd.size; ++j) {
            C.qnode[nd.off + j] = t;
            C.qcol[nd.off + j] = j;
        }
    }
    for (int q = 0; q < n; ++q) {
        const int t = C.qnode[q];
        const Node& nd = C.nodes[t];
        const int x = C.qxy[q] & 0xffff, y = C.qxy[q] >> 16, v = C.qv[q];
        const int nbx[6] = {x - 1, x + 1, x, x, x, x};
        const int nby[6] = {y, y, y - 1, y + 1, y, y};
        const int nbv[6] = {v, v, v, v, v - 1, v + 1};
        for (int e = 0; e < 6; ++e) {
            if (nbx[e] < 0 || nbx[e] >= nx || nby[e] < 0 || nby[e] >= ny || nbv[e] < 0 ||
                nbv[e] >= vs)
                continue;
            const int ni = NestedIndex(nbx[e], nby[e], nbv[e], nx, ny, vs, cut);
            if (ni <= q) continue;
            int row;
            if (ni < nd.off + nd.size) {
                row = ni - nd.off;
            } else {
                const int* s0 = C.structs.data() + nd.ustart;
                const int* f = std::lower_bound(s0, s0 + nd.usize, ni);
                row = (f != s0 + nd.usize && *f == ni) ? nd.size + int(f - s0) : -1;
            }
            C.qrows[(size_t)q * 6 + e] = row;
        }
    }
}

void AddGemmTiles(ClassPlan& C, const GemmDesc& d, int tile)
{
    const int di = (int)C.gemms.size();
    C.gemms.push_back(d);
    for (int tn = 0; tn < d.N; tn += tile)
        for (int tm = 0; tm < d.M; tm += tile) {
            if ((d.flags & G_CLOWER) && tm + tile - 1 < tn) continue;
            C.tiles.push_back({di, tm, tn, 0});
        }
}

// Recursive combine steps of the unit-lower inverse; returns the height.
struct Combine {
    int lo, mid, hi;
};
int InvRecurse(int lo, int hi, std::vector<std::vector<Combine>>& byHeight)
{
    if (hi - lo <= 1) return 0;
    const int mid = (lo + hi) / 2;
    const int h = 1 + std::max(InvRecurse(lo, mid, byHeight), InvRecurse(mid, hi, byHeight));
    if ((int)byHeight.size() <= h) byHeight.resize(h + 1);
    byHeight[h].push_back({lo, mid, hi});
    return h;
}

void Schedule(const Plan& plan, ClassPlan& C)
{
    const int nb = plan.blockNb, tile = plan.tile;
    const int nn = (int)C.nodes.size();
    int64_t lo = 0, wo = 0;
    for (Node& n : C.nodes) {
        const int64_t fs = n.size + n.usize;
        n.loff = lo;
        lo += fs * n.size;
        n.woff = wo;
        wo += fs;
    }
    C.fstore = lo;
    C.wstore = wo;
    // Schur-complement workspace: the updates of depth d are written while
    // that level is factored (the children's extend-add into the rows past
    // the pivots, then the Schur GEMM) and read by the extend-add of depth
    // d - 1, after which they are dead.  So the levels of even depth share
    // one region and those of odd depth another; a level's part is zeroed
    // just before its extend-adds.
    C.steps.assign(C.maxDepth + 1, {});
    std::vector<int64_t> lvU(C.maxDepth + 1, 0);
    for (const Node& n : C.nodes) lvU[n.depth] += (int64_t)n.usize * n.usize;
    int64_t region[2] = {0, 0};
    for (int d = 0; d <= C.maxDepth; ++d) region[d & 1] = std::max(region[d & 1], lvU[d]);
    std::vector<int64_t> uat(C.maxDepth + 1);
    for (int d = 0; d <= C.maxDepth; ++d) {
        uat[d] = (d & 1) ? region[0] : 0;
        C.steps[d].uoff = uat[d];
        C.steps[d].ucount = lvU[d];
    }
    for (Node& n : C.nodes) {
        n.uoff = uat[n.depth];
        uat[n.depth] += (int64_t)n.usize * n.usize;
    }
    C.ustore = region[0] + region[1];

    for (int d = C.maxDepth; d >= 0; --d) {
        ClassPlan::LevelSteps& S = C.steps[d];
        const std::vector<int>& lv = C.levels[d];
        for (int slot = 0; slot < 2; ++slot) {
            const int b0 = (int)C.extadd.size();
            for (int t : lv) {
                const Node& p = C.nodes[t];
                const int c = p.child[slot];
                if (c < 0) continue;
                const Node& ch = C.nodes[c];
                if (ch.usize == 0) continue;
                for (int c0 = 0; c0 < ch.usize; c0 += 16) {
                    ExtAddItem it;
                    it.uChild = ch.uoff;
                    it.lParent = p.loff;
                    it.uParent = p.uoff;
                    it.relStart = (int)ch.ustart;
                    it.uc = ch.usize;
                    it.sp = p.size;
                    it.up = p.usize;
                    it.col0 = c0;
                    it.col1 = std::min(c0 + 16, ch.usize);
                    C.extadd.push_back(it);
                }
            }
            S.extadd[slot].push_back({b0, (int)C.extadd.size() - b0});
        }
        int maxs = 0;
        for (int t : lv) maxs = std::max(maxs, C.nodes[t].size);
        for (int j0 = 0; j0 < maxs; j0 += nb) {
            const int l0 = (int)C.ldl.size();
            for (int t : lv) {
                const Node& n = C.nodes[t];
                if (j0 >= n.size) continue;
                LdlItem it;
                it.l = n.loff;
                it.ld = n.size + n.usize;
                it.j0 = j0;
                it.j1 = std::min(j0 + nb, n.size);
                it.r0 = it.r1 = -1;
                C.ldl.push_back(it);
            }
            S.ldlDiag.push_back({l0, (int)C.ldl.size() - l0});
            const int l1 = (int)C.ldl.size();
            for (int t : lv) {
                const Node& n = C.nodes[t];
                if (j0 >= n.size) continue;
                const int j1 = std::min(j0 + nb, n.size);
                const int fs = n.size + n.usize;
                for (int r0 = j1; r0 < fs; r0 += 256) {
                    LdlItem it;
                    it.l = n.loff;
                    it.ld = fs;
                    it.j0 = j0;
                    it.j1 = j1;
                    it.r0 = r0;
                    it.r1 = std::min(r0 + 256, fs);
                    C.ldl.push_back(it);
                }
            }
            S.ldlTrsm.push_back({l1, (int)C.ldl.size() - l1});
            // Trailing updates in panels of two blocks: after the first block
            // only the second block's columns are updated; after the second,
            // every later pivot column at once with both blocks (K = 2 nb),
            // so the trailing matrix is read and written once per two blocks.
            const bool first = (j0 / nb) % 2 == 0;
            GemmLaunch gl;
            gl.tileStart = (int)C.tiles.size();
            for (int t : lv) {
                const Node& n = C.nodes[t];
                if (j0 >= n.size) continue;
                const int j1 = std::min(j0 + nb, n.size);
                if (j1 >= n.size) continue;
                const int fs = n.size + n.usize;
                const int p0 = first ? j0 : j0 - nb;  // the panel's first column
                GemmDesc g{};
                g.a = n.loff + j1 + (int64_t)p0 * fs;
                g.b = g.a;
                g.c = n.loff + j1 + (int64_t)j1 * fs;
                g.d = n.loff + p0 + (int64_t)p0 * fs;
                g.lda = g.ldb = g.ldc = fs;
                g.dinc = fs + 1;
                g.M = fs - j1;
                g.N = first ? std::min(j1 + nb, n.size) - j1 : n.size - j1;
                g.K = j1 - p0;
                g.abuf = g.bbuf = g.cbuf = g.dbuf = BUF_L;
                g.flags = G_BTRANS | G_HASD | G_NEG | G_CLOWER;
                AddGemmTiles(C, g, tile);
            }
            gl.tileCount = (int)C.tiles.size() - gl.tileStart;
            S.trail.push_back(gl);
        }
        S.schur.tileStart = (int)C.tiles.size();
        for (int t : lv) {
            const Node& n = C.nodes[t];
            if (n.usize == 0 || n.size == 0 || n.parent < 0) continue;
            const int fs = n.size + n.usize;
            GemmDesc g{};
            g.a = n.loff + n.size;
            g.b = g.a;
            g.c = n.uoff;
            g.d = n.loff;
            g.lda = g.ldb = fs;
            g.ldc = n.usize;
            g.dinc = fs + 1;
            g.M = g.N = n.usize;
            g.K = n.size;
            g.abuf = g.bbuf = g.dbuf = BUF_L;
            g.cbuf = BUF_U;
            g.flags = G_BTRANS | G_HASD | G_NEG | G_CLOWER;
            AddGemmTiles(C, g, tile);
        }
        S.schur.tileCount = (int)C.tiles.size() - S.schur.tileStart;
    }

    // Diagonal-block inversion: every nb block, then recursive combines.
    std::vector<std::vector<std::pair<int, Combine>>> comb;  // by height
    int64_t to = 0;
    std::vector<std::vector<int64_t>> tOffsets(nn);
    for (int t = 0; t < nn; ++t) {
        Node& n = C.nodes[t];
        n.toff = to;
        if (n.size == 0) continue;
        const int fs = n.size + n.usize;
        const int nblk = (int)CeilDiv(n.size, nb);
        for (int b = 0; b < nblk; ++b) {
            InvItem it;
            it.l = n.loff;
            it.ld = fs;
            it.j0 = b * nb;
            it.j1 = std::min((b + 1) * nb, n.size);
            it.pad_ = 0;
            C.inv.push_back(it);
        }
        std::vector<std::vector<Combine>> byH;
        InvRecurse(0, nblk, byH);
        if (comb.size() < byH.size()) comb.resize(byH.size());
        for (size_t h = 1; h < byH.size(); ++h)
            for (const Combine& c : byH[h]) {
                comb[h].push_back({t, c});
                const int r0 = c.mid * nb, r1 = std::min(c.hi * nb, n.size);
                const int c0 = c.lo * nb, c1 = c.mid * nb;
                to += (int64_t)(r1 - r0) * (c1 - c0);
            }
    }
    C.tstore = to;
    // assign T offsets per combine in the same order
    std::vector<int64_t> cursor(nn);
    for (int t = 0; t < nn; ++t) cursor[t] = C.nodes[t].toff;
    for (size_t h = 1; h < comb.size(); ++h) {
        GemmLaunch gT, gX;
        gT.tileStart = (int)C.tiles.size();
        std::vector<GemmDesc> xs;
        for (const auto& pc : comb[h]) {
            const Node& n = C.nodes[pc.first];
            const Combine& c = pc.second;
            const int fs = n.size + n.usize;
            const int r0 = c.mid * nb, r1 = std::min(c.hi * nb, n.size);
            const int c0 = c.lo * nb, c1 = c.mid * nb;
            const int64_t tOff = cursor[pc.first];
            cursor[pc.first] += (int64_t)(r1 - r0) * (c1 - c0);
            GemmDesc g{};  // T = C * Ainv
            g.a = n.loff + r0 + (int64_t)c0 * fs;
            g.lda = fs;
            g.b = n.loff + c0 + (int64_t)c0 * fs;
            g.ldb = fs;
            g.c = tOff;
            g.ldc = r1 - r0;
            g.M = r1 - r0;
            g.N = c1 - c0;
            g.K = c1 - c0;
            g.abuf = g.bbuf = BUF_L;
            g.cbuf = BUF_T;
            g.dbuf = BUF_L;
            g.flags = G_BUNIT | G_BETA0;
            AddGemmTiles(C, g, tile);
            GemmDesc x{};  // X = -Binv * T
            x.a = n.loff + r0 + (int64_t)r0 * fs;
            x.lda = fs;
            x.b = tOff;
            x.ldb = r1 - r0;
            x.c = n.loff + r0 + (int64_t)c0 * fs;
            x.ldc = fs;
            x.M = r1 - r0;
            x.N = c1 - c0;
            x.K = r1 - r0;
            x.abuf = BUF_L;
            x.bbuf = BUF_T;
            x.cbuf = BUF_L;
            x.dbuf = BUF_L;
            x.flags = G_AUNIT | G_NEG | G_BETA0;
            xs.push_back(x);
        }
        gT.tileCount = (int)C.tiles.size() - gT.tileStart;
        gX.tileStart = (int)C.tiles.size();
        for (const GemmDesc& x : xs) AddGemmTiles(C, x, tile);
        gX.tileCount = (int)C.tiles.size() - gX.tileStart;
        C.invGemm.push_back(gT);
        C.invGemm.push_back(gX);
    }

    // Solve schedule.
    {
        C.ginv.assign((size_t)2 * C.wstore, -1);
        for (const Node& n : C.nodes)
            for (int k = 0; k < 2; ++k) {
                if (n.child[k] < 0) continue;
                const Node& ch = C.nodes[n.child[k]];
                for (int a = 0; a < ch.usize; ++a) {
                    const int r = C.relinds[ch.ustart + a];
                    C.ginv[(size_t)2 * (n.woff + r) + k] = (int32_t)(ch.woff + ch.size + a);
                }
            }
        DefaultSolveShapes(C, plan.minTiles, plan.maxSplit, plan.wideWg, plan.persist);
        BuildSolveTiles(C);
        BuildQueues(C, plan.fused ? plan.fusedQueues : 0, plan.fusedRows, plan.fusedCut,
                    plan.fusedWgPerQueue);
        // Solve layout: LT[:, :s] = Linv^T D^-1 (32x32 blocks of L_TL),
        // LT[:, s:] = (-L_BL Linv)^T by one batched GEMM, then L_BL and the
        // diagonal of L from LT (an item with i0 == 0 per column block also
        // stores the unit diagonal, so nodes without an update get one).
        C.fold.tileStart = (int)C.tiles.size();
        for (int t = 0; t < nn; ++t) {
            const Node& n = C.nodes[t];
            const int fs = n.size + n.usize;
            for (int i0 = 0; i0 < n.size; i0 += 32)
                for (int j0 = 0; j0 < n.size; j0 += 32) C.trans.push_back({t, i0, j0, 0});
            for (int i0 = 0; i0 < std::max(n.usize, 1); i0 += 32)
                for (int j0 = 0; j0 < n.size; j0 += 32) C.foldBack.push_back({t, i0, j0, 0});
            if (n.size == 0 || n.usize == 0) continue;
            GemmDesc g{};
            g.a = n.loff + n.size;  // L_BL
            g.lda = fs;
            g.b = n.loff;           // Linv (unit lower)
            g.ldb = fs;
            g.c = n.loff + (int64_t)n.size * n.size;  // LT[:, s:], stored transposed
            g.ldc = n.size;
            g.d = 0;
            g.dinc = 0;
            g.M = n.usize;
            g.N = n.size;
            g.K = n.size;
            g.abuf = BUF_L;
            g.bbuf = BUF_L;
            g.cbuf = BUF_LT;
            g.dbuf = BUF_L;
            g.flags = G_BUNIT | G_NEG | G_BETA0 | G_CTRANS;
            AddGemmTiles(C, g, plan.tile);
        }
        C.fold.tileCount = (int)C.tiles.size() - C.fold.tileStart;
    }

    // Algorithmic counts (8 real flops per complex multiply-add).
    double fl = 0, sc = 0, sv = 0, fe = 0;
    for (const Node& n : C.nodes) {
        const double s = n.size, u = n.usize;
        fl += 8.0 * (s * s * s / 3.0 + u * s * s / 2.0 + u * (u + 1) * s / 2.0);
        sc += 8.0 * (u * (u + 1) * s / 2.0);
        sv += 8.0 * (s * s + 2.0 * u * s);
        fe += (s + u) * s;
    }
    C.flops = fl;
    C.schurFlops = sc;
    C.solveFlopsPerRhs = sv;
    C.factorEntries = fe;
}

}  // namespace

namespace {

// Phase 0 (F): rows [0, s+u), inner [0, min(s, r0 + rows)) (rows < s are
// lower triangular).  Phase 1 (B): rows [0, s), inner [r0, s+u) (upper
// triangle, then the update columns).
int SolveRowsOf(int ph, const Node& n) { return ph == 0 ? n.size + n.usize : n.size; }

bool SolveWanted(int ph, const Node& n) { return ph == 0 ? n.size + n.usize > 0 : n.size > 0; }

int SolveLo(int ph, int r0) { return ph == 1 ? r0 : 0; }
int SolveHi(int ph, const Node& n, int r0, int rows)
{
    if (ph == 0) return std::min(n.size, r0 + rows);
    return n.size + n.usize;
}

int64_t RowBlocks(const ClassPlan& C, int d, int ph, int rows)
{
    int64_t t = 0;
    for (int v : C.levels[d])
        if (SolveWanted(ph, C.nodes[v])) t += CeilDiv(SolveRowsOf(ph, C.nodes[v]), rows);
    return t;
}

}  // namespace

void DefaultSolveShapes(ClassPlan& C, int minTiles, int maxSplit, bool wideWg, bool persist)
{
    C.shapes.assign(C.maxDepth + 1, {});
    for (int d = 0; d <= C.maxDepth; ++d)
        for (int ph = 0; ph < 2; ++ph) {
            SolveShape sh{16, 1, 4, persist ? 1 : 0};
            for (int cand : {64, 32, 16}) {
                sh.rows = cand;
                if (RowBlocks(C, d, ph, cand) >= minTiles) break;
            }
            const int64_t t = RowBlocks(C, d, ph, sh.rows);
            if (t > 0 && t < minTiles) {
                if (wideWg && !persist)
                    sh.waves = 16;
                else
                    sh.split = (int)std::min<int64_t>(maxSplit, CeilDiv(minTiles, t));
            }
            C.shapes[d][ph] = sh;
        }
}

void BuildSolveTiles(ClassPlan& C)
{
    C.mv.clear();
    C.mvred.clear();
    C.maxSlots = 0;
    C.solveLevels.assign(C.maxDepth + 1, {});
    // One tile per (block of `rows` output rows, inner slice).  A row block
    // whose inner range [lo, hi) is long enough is cut into up to `split`
    // slices of at least kMinSlice indices (multiples of 16); its tiles then
    // write partial sums to consecutive slots, and one reduce item sums them
    // in slot order.
    constexpr int kMinSlice = 64;
    auto build = [&](int d, int ph, ClassPlan::Range& rg) {
        const SolveShape sh = C.shapes[d][ph];
        rg.rows = sh.rows;
        rg.waves = sh.waves;
        rg.persist = sh.persist;
        rg.start = (int)C.mv.size();
        rg.rstart = (int)C.mvred.size();
        int slots = 0;
        for (int t : C.levels[d]) {
            const Node& n = C.nodes[t];
            if (!SolveWanted(ph, n)) continue;
            const int nr = SolveRowsOf(ph, n);
            for (int r0 = 0; r0 < nr; r0 += sh.rows) {
                const int lo = SolveLo(ph, r0);
                const int hi = std::max(lo, SolveHi(ph, n, r0, sh.rows));
                const int len = hi - lo;
                int per = len, nsl = 1;
                if (sh.split > 1 && len >= 2 * kMinSlice) {
                    per = std::max(kMinSlice, (int)CeilDiv(CeilDiv(len, sh.split), 16) * 16);
                    nsl = (int)CeilDiv(len, per);
                }
                MvTile m;
                m.node = t;
                m.row0 = r0;
                m.nrows = std::min(sh.rows, nr - r0);
                if (nsl == 1) {
                    m.col0 = lo;
                    m.col1 = hi;
                    m.slot = -1;
                    C.mv.push_back(m);
                    continue;
                }
                MvTile red = m;
                red.col0 = slots;  // first slot
                red.col1 = nsl;    // slot count
                red.slot = 0;
                m.red = (int)C.mvred.size();
                C.mvred.push_back(red);
                for (int j = 0; j < nsl; ++j) {
                    m.col0 = lo + j * per;
                    m.col1 = std::min(hi, lo + (j + 1) * per);
                    m.slot = slots++;
                    C.mv.push_back(m);
                }
            }
        }
        rg.count = (int)C.mv.size() - rg.start;
        rg.rcount = (int)C.mvred.size() - rg.rstart;
        C.maxSlots = std::max(C.maxSlots, slots);
    };
    // XCD placement of a one-workgroup-per-tile level with split row blocks:
    // workgroup i of a launch runs on XCD i % 8, so every slice of a row
    // block goes to a position with the same residue mod 8 (a row block per
    // residue, least loaded residue first); the last slice to finish can then
    // sum the block in the launch, with hand-offs inside one XCD's L2.  The
    // positions a residue leaves free are holes (tiles with no rows).
    auto place = [&](ClassPlan::Range& rg) {
        rg.xcd = 0;
        if (rg.persist || rg.rcount <= 0) return;
        std::vector<std::vector<MvTile>> lane(8);
        for (int i = rg.start; i < rg.start + rg.count;) {
            int j = i + 1;
            if (C.mv[i].red >= 0)
                while (j < rg.start + rg.count && C.mv[j].red == C.mv[i].red) ++j;
            int best = 0;
            for (int q = 1; q < 8; ++q)
                if (lane[q].size() < lane[best].size()) best = q;
            for (int t = i; t < j; ++t) lane[best].push_back(C.mv[t]);
            i = j;
        }
        size_t len = 0;
        for (const auto& l : lane) len = std::max(len, l.size());
        MvTile hole = C.mv[rg.start];
        hole.nrows = 0;
        hole.slot = -1;
        hole.red = -1;
        std::vector<MvTile> placed(8 * len, hole);
        for (int q = 0; q < 8; ++q)
            for (size_t t = 0; t < lane[q].size(); ++t) placed[8 * t + q] = lane[q][t];
        C.mv.erase(C.mv.begin() + rg.start, C.mv.end());
        C.mv.insert(C.mv.end(), placed.begin(), placed.end());
        rg.count = (int)placed.size();
        rg.xcd = 1;
    };
    // Compulsory work of a phase: the factor entries it applies, and per
    // node its inner vector rows read once plus its output rows.
    auto account = [&](ClassPlan::SolveLevel& L, int ph, const ClassPlan::Range& r) {
        int node = -1, lo = 0, hi = 0, rows = 0, lastRow = -1;
        auto flush = [&]() {
            if (node >= 0) L.vecRows[ph] += (hi - lo) + rows;
        };
        for (int i = r.start; i < r.start + r.count; ++i) {
            const MvTile& m = C.mv[i];
            double e = 0;
            for (int row = m.row0; row < m.row0 + m.nrows; ++row) {
                if (ph == 0) e += std::max(0, std::min(row + 1, m.col1) - m.col0);  // lower incl. diagonal
                else e += std::max(0, m.col1 - std::max(row, m.col0));             // upper incl. diagonal
            }
            L.entries[ph] += e;
            if (m.node != node) {
                flush();
                node = m.node;
                lo = m.col0;
                hi = m.col1;
                rows = 0;
                lastRow = -1;
            }
            lo = std::min(lo, m.col0);
            hi = std::max(hi, m.col1);
            if (m.row0 != lastRow) rows += m.nrows;
            lastRow = m.row0;
        }
        flush();
    };
    for (int d = 0; d <= C.maxDepth; ++d) {
        ClassPlan::SolveLevel& L = C.solveLevels[d];
        build(d, 0, L.f);  // [y; W_bot] = L g_top + [0; g_bot]
        account(L, 0, L.f);
        place(L.f);
        build(d, 1, L.b);  // x = LT [y; x_struct]
        account(L, 1, L.b);
        place(L.b);
    }
}

void BuildQueues(ClassPlan& C, int nqueues, int rows, int cut, int wgPerQueue)
{
    C.qcut = -1;
    C.nqueues = 0;
    C.queueRows = rows;
    for (auto& Q : C.queue) Q = ClassPlan::QueuePhase{};
    if (nqueues <= 0 || C.nodes.empty()) return;
    // cut: the first depth with nqueues nodes (or the requested one)
    int qc = -1;
    for (int d = 0; d <= C.maxDepth; ++d)
        if ((int)C.levels[d].size() >= std::min(nqueues, kQueues)) {
            qc = d;
            break;
        }
    if (cut >= 0) qc = qc < 0 ? -1 : std::max(qc, cut);
    if (qc < 0 || qc > C.maxDepth) return;
    const int nn = (int)C.nodes.size();
    // row blocks of a node per phase, and the node's factor entries (both
    // phases read the same count)
    auto rb = [&](int ph, int t) -> int {
        const Node& n = C.nodes[t];
        return SolveWanted(ph, n) ? (int)CeilDiv(SolveRowsOf(ph, n), rows) : 0;
    };
    std::vector<double> sub(nn, 0.0);  // subtree entries (post-order: children first)
    for (int t = 0; t < nn; ++t) {
        const Node& n = C.nodes[t];
        sub[t] += (double)n.size * (n.size + 1) / 2 + (double)n.usize * n.size;
        if (n.parent >= 0) sub[n.parent] += sub[t];
    }
    // the cut nodes' subtrees over the queues, largest first to the least
    // loaded queue
    std::vector<int> cutNodes = C.levels[qc];
    std::sort(cutNodes.begin(), cutNodes.end(),
              [&](int a, int b) { return sub[a] != sub[b] ? sub[a] > sub[b] : a < b; });
    std::vector<double> load(nqueues, 0.0);
    std::vector<int> queueOf(nn, -1);
    for (int t : cutNodes) {
        int best = 0;
        for (int q = 1; q < nqueues; ++q)
            if (load[q] < load[best]) best = q;
        load[best] += sub[t];
        queueOf[t] = best;
    }
    // post-order: a parent comes after its children, so walk the nodes from
    // the root down to give every fused node its cut ancestor's queue
    for (int t = nn - 1; t >= 0; --t) {
        const Node& n = C.nodes[t];
        if (n.depth > qc && n.parent >= 0) queueOf[t] = queueOf[n.parent];
    }
    constexpr int kMinSlice = 64;
    int slots = 0;
    for (int ph = 0; ph < 2; ++ph) {
        ClassPlan::QueuePhase& Q = C.queue[ph];
        Q.dep.assign(nn, {-1, -1, 0, 0});
        for (int t = 0; t < nn; ++t) {
            const Node& n = C.nodes[t];
            if (n.depth < qc) continue;
            auto& dp = Q.dep[t];
            if (ph == 0) {
                int j = 0;
                for (int c = 0; c < n.nchild; ++c) {
                    const int ch = n.child[c];
                    if (ch >= 0 && rb(0, ch) > 0) {
                        dp[j] = ch;
                        dp[2 + j] = rb(0, ch);
                        ++j;
                    }
                }
            } else {
                int a = n.parent;
                while (a >= 0 && rb(1, a) == 0) a = C.nodes[a].parent;
                if (a >= 0 && C.nodes[a].depth >= qc) {
                    dp[0] = a;
                    dp[2] = rb(1, a);
                }
            }
        }
        // inner slices per depth: enough tiles for the queue's workgroups
        std::vector<int> split(C.maxDepth + 1, 1);
        for (int d = qc; d <= C.maxDepth; ++d) {
            int64_t r = 0;
            for (int t : C.levels[d]) r += rb(ph, t);
            const double perQueue = (double)r / nqueues;
            if (perQueue > 0 && perQueue < wgPerQueue)
                split[d] = std::min(8, (int)std::ceil(wgPerQueue / perQueue));
        }
        Q.qstart.assign(nqueues + 1, 0);
        for (int q = 0; q < nqueues; ++q) {
            Q.qstart[q] = (int)Q.tiles.size();
            for (int i = 0; i <= C.maxDepth - qc; ++i) {
                const int d = ph == 0 ? C.maxDepth - i : qc + i;
                for (int t : C.levels[d]) {
                    if (queueOf[t] != q) continue;
                    const Node& n = C.nodes[t];
                    if (!SolveWanted(ph, n)) continue;
                    const int nr = SolveRowsOf(ph, n);
                    for (int r0 = 0; r0 < nr; r0 += rows) {
                        const int lo = SolveLo(ph, r0);
                        const int hi = std::max(lo, SolveHi(ph, n, r0, rows));
                        const int len = hi - lo;
                        int per = len, nsl = 1;
                        if (split[d] > 1 && len >= 2 * kMinSlice) {
                            per = std::max(kMinSlice, (int)CeilDiv(CeilDiv(len, split[d]), 16) * 16);
                            nsl = (int)CeilDiv(len, per);
                        }
                        QTile m{t, r0, std::min(rows, nr - r0), lo, hi, -1, -1, 0};
                        if (nsl > 1) {
                            m.red = (int)Q.reds.size();
                            Q.reds.push_back(QRed{t, r0, m.nrows, slots, nsl, {0, 0, 0}});
                        }
                        for (int j = 0; j < nsl; ++j) {
                            if (nsl > 1) {
                                m.col0 = lo + j * per;
                                m.col1 = std::min(hi, lo + (j + 1) * per);
                                m.slot = slots++;
                            }
                            Q.tiles.push_back(m);
                        }
                    }
                }
            }
        }
        Q.qstart[nqueues] = (int)Q.tiles.size();
        for (int d = qc; d <= C.maxDepth; ++d) {
            Q.entries += C.solveLevels[d].entries[ph];
            Q.vecRows += C.solveLevels[d].vecRows[ph];
        }
    }
    C.maxSlots = std::max(C.maxSlots, slots);
    C.qcut = qc;
    C.nqueues = nqueues;
}

int BuildPlan(const psp_hip_desc& d, Plan& plan, std::string& msg)
{
    if (d.nx < 1 || d.ny < 1 || d.nz < 1 || d.bz < 1 || d.planesPerPanel < 1 ||
        d.cutoff < 1 || d.nx > 32767 || d.ny > 32767) {
        msg = "invalid discretization sizes";
        return PSP_HIP_ERR_INVALID;
    }
    if (!(d.omega == d.omega) || !(d.wx > 0) || !(d.wy > 0) || !(d.wz > 0)) {
        msg = "invalid frequency or widths";
        return PSP_HIP_ERR_INVALID;
    }
    const int st = plan.geo.init(d);
    if (st != PSP_HIP_OK) {
        msg = "The domain is very shallow. Please run a sparse-direct factorization instead.";
        return st;
    }
    for (int c = 0; c < 3; ++c) {
        if (c == 1 && !plan.geo.haveLeftover) continue;
        ClassPlan& C = plan.cls[c];
        C.cls = c;
        C.vsize = plan.geo.classVSize(c);
        C.n = plan.geo.nx * plan.geo.ny * C.vsize;
        std::vector<std::vector<int>> orig;
        BuildTree(plan.geo, C, orig);
        Analyse(C, orig);
        AssemblyMeta(plan.geo, C);
        Schedule(plan, C);
        // Every index the kernels dereference must land inside its front.
        for (const Node& n : C.nodes)
            for (int a = 0; a < n.usize; ++a) {
                const int64_t e = n.ustart + a;
                if (C.structs[e] < 0 || C.structs[e] >= C.n ||
                    (n.parent >= 0 && C.relinds[e] < 0)) {
                    msg = "internal error: inconsistent front structure";
                    return PSP_HIP_ERR_INVALID;
                }
            }
        for (int q = 0; q < C.n; ++q)
            if (C.qnode[q] < 0) {
                msg = "internal error: unknown without a front";
                return PSP_HIP_ERR_INVALID;
            }
        plan.present[c] = true;
    }
    return PSP_HIP_OK;
}

}  // namespace psp_amd
