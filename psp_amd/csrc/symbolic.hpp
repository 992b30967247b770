// Host-side symbolic analysis for the MI355X sweep solver: PSP's panel
// split, analytic nested dissection, elimination tree and front structures,
// plus the static device work lists (front assembly, blocked LDL^T steps,
// Schur GEMMs, diagonal-block inversion, per-level solve phases).
//
// Everything here is integer work done once per DistHelmholtz object
// (PSP's constructor, dist_helmholtz/main.hpp:12-436).
#pragma once

#include <cstdint>
#include <string>
#include <array>
#include <vector>

#include "psp_hip.h"

namespace psp_amd {

// ---------------------------------------------------------------------------
// Panel geometry (dist_helmholtz/main.hpp:35-57, 728-821).
// ---------------------------------------------------------------------------
struct Geometry {
    int nx = 0, ny = 0, nz = 0, bz = 0, P = 0, cutoff = 0;
    bool topPML = true;
    int bottomDepth = 0, topOrigDepth = 0, innerDepth = 0, leftoverDepth = 0;
    int numFullInner = 0, numPanels = 0;
    bool haveLeftover = false;

    // Returns PSP_HIP_OK or PSP_HIP_ERR_SHALLOW.
    int init(const psp_hip_desc& d);
    int padding(int p) const { return p == 0 ? 0 : bz; }
    int depth(int p) const;
    int panelV(int p) const;
    int cls(int p) const;  // 0 bottom/full inner, 1 leftover, 2 top
    int vOffset(int p) const { return panelV(p) - padding(p); }
    int classVSize(int c) const;
};

// In-panel nested-dissection index (ReorderedIndexRecursion,
// dist_helmholtz/main.hpp:1425-1477, one process).
int NestedIndex(int x, int y, int vLocal, int xSize, int ySize, int vSize,
                int cutoff);

// ---------------------------------------------------------------------------
// Device work-list records.  All offsets are element offsets inside one
// panel's slice of the named buffer; the kernels add batch * stride.
// ---------------------------------------------------------------------------
enum BufId : int16_t { BUF_L = 0, BUF_U = 1, BUF_T = 2, BUF_LT = 3 };

enum GemmFlags : int32_t {
    G_BTRANS = 1,   // B(k,n) = Bp[n + k*ldb]
    G_AUNIT = 2,    // A is unit lower triangular (square)
    G_BUNIT = 4,    // B is unit lower triangular (square)
    G_CLOWER = 8,   // only C(i,j) with i >= j is touched
    G_HASD = 16,    // multiply by diag(D) between A and B
    G_NEG = 32,     // C -= ...
    G_BETA0 = 64,   // C is overwritten instead of accumulated
    G_CTRANS = 128, // C(i,j) is stored at Cp[j + i*ldc]
};

struct GemmDesc {
    int64_t a, b, c, d;
    int32_t lda, ldb, ldc, dinc;
    int32_t M, N, K;
    int16_t abuf, bbuf, cbuf, dbuf;
    int32_t flags;
    int32_t pad_;
};

struct GemmTile {
    int32_t desc;
    int32_t tm, tn;
    int32_t pad_;
};

// Extend-add of one child's Schur complement column range into its parent.
struct ExtAddItem {
    int64_t uChild;     // child's U offset (BUF_U)
    int64_t lParent;    // parent's L offset (BUF_L)
    int64_t uParent;    // parent's U offset (BUF_U)
    int32_t relStart;   // into relinds
    int32_t uc;         // child's update size
    int32_t sp, up;     // parent's pivot and update sizes
    int32_t col0, col1; // child's U column range handled by this item
};

// One block column step of the blocked LDL^T on one front: either the
// factorization of the diagonal block [j0,j1)^2 (r0 == r1 == -1) or the
// triangular solve of rows [r0, r1) against the factored diagonal block.
struct LdlItem {
    int64_t l;          // front's L offset
    int32_t ld;         // s + u
    int32_t j0, j1;     // block columns [j0, j1)
    int32_t r0, r1;     // rows [r0, r1) below the diagonal block
};

// Inversion of one diagonal block of a unit lower triangular L_TL.
struct InvItem {
    int64_t l;
    int32_t ld;
    int32_t j0, j1;
    int32_t pad_;
};

// One row block of a matrix-vector product of one node in a front solve:
// output rows [row0, row0 + nrows) (nrows <= 64), inner indices [col0, col1).
// slot >= 0: the tile is one inner slice of its row block and writes its
// partial sums to that slot of the partial buffer.  As a reduce item (same
// struct): col0 = first slot, col1 = number of slots.
struct MvTile {
    int32_t node;
    int32_t row0, nrows;  // nrows 0: a hole of the XCD placement (no work)
    int32_t col0, col1;
    int32_t slot;
    int32_t red = -1;     // reduce item of a slice's row block (into mvred), -1 if none
};

// Fused subtree solves: one persistent launch per phase runs every level of
// a panel's tree from the cut depth down, the subtrees under the cut split
// over kQueues queues (one per XCD of the device, each queue's tiles in
// dependency order).  A queue tile names its node (the node's fields come
// from the per-node records); red >= 0: the tile is one inner slice
// (partial sums in slot `slot`) of the split row block QRed[red].
constexpr int kQueues = 8;
struct QTile {
    int32_t node, row0, nrows, col0, col1, slot, red, pad_;
};
struct QRed {
    int32_t node, row0, nrows, slot0, nslots, pad_[3];
};
static_assert(sizeof(QTile) == 32 && sizeof(QRed) == 32, "queue records");

// Shape of the tiles of one (level, phase) of the front solves.
struct SolveShape {
    int rows = 64;  // output rows per tile: 64, 32 or 16
    int split = 1;  // inner slices per row block (partial sums + reduce)
    int waves = 4;  // waves per workgroup (4, or 16 to spread the inner range
                    // of a tile over more waves instead of slices), 8 <= k < 32
    int persist = 0;  // 1: persistent pipelined workgroups (8 <= k <= 16, 4 waves)
};

// 32x32 block (rows i0.., columns j0..) of a per-front copy between L and
// LT (see the solve layout in ClassPlan).
struct TransItem {
    int32_t node, i0, j0, pad_;
};

struct Node {
    int off = 0, size = 0, usize = 0, parent = -1;
    int child[2] = {-1, -1};
    int nchild = 0;
    int depth = 0;
    int64_t ustart = 0;   // into structs / relinds
    int64_t loff = 0;     // factor storage (BUF_L), (s+u) x s column-major
    int64_t uoff = 0;     // Schur complement workspace (BUF_U), u x u
    int64_t toff = 0;     // inversion temporary (BUF_T)
    int64_t woff = 0;     // solve work vector rows (s+u)
};

struct GemmLaunch {
    int32_t tileStart = 0, tileCount = 0;
};

struct ClassPlan {
    int cls = 0;
    int vsize = 0;
    int n = 0;          // padded unknowns nx*ny*vsize
    int maxDepth = 0;
    std::vector<Node> nodes;            // post-order, root last
    std::vector<int> structs;           // sorted lower structures
    std::vector<int> relinds;           // row of each struct entry in parent
    std::vector<std::vector<int>> levels;  // node ids per depth
    int64_t fstore = 0, ustore = 0, tstore = 0, wstore = 0;

    // Front assembly: per padded unknown q: packed x | y<<16, vLocal,
    // node, column within node, and six front rows (x-1,x+1,y-1,y+1,v-1,v+1)
    // or -1 when the coupling is absent or above the diagonal.
    std::vector<int32_t> qxy, qv, qnode, qcol;
    std::vector<int32_t> qrows;  // 6 per q

    // Factorization schedule.
    struct LevelSteps {
        int64_t uoff = 0, ucount = 0;                // the level's updates in BUF_U
        std::vector<std::pair<int, int>> extadd[2];  // item range per child slot
        std::vector<std::pair<int, int>> ldlDiag;    // per block step: item range
        std::vector<std::pair<int, int>> ldlTrsm;    // per block step: item range
        std::vector<GemmLaunch> trail;               // per block step
        GemmLaunch schur;
    };
    std::vector<ExtAddItem> extadd;
    std::vector<LdlItem> ldl;
    std::vector<GemmDesc> gemms;
    std::vector<GemmTile> tiles;
    std::vector<LevelSteps> steps;        // per depth (deepest first order)
    std::vector<InvItem> inv;
    std::vector<GemmLaunch> invGemm;      // alternating T = C*Ainv, X = -Binv*T

    // Solve layout of the factors, made after the selective inversion:
    //   L  ((s+u) x s, ld s+u): rows < s hold Linv (unit lower, ones stored
    //      on the diagonal), rows >= s hold -L_BL Linv;
    //   LT (s x (s+u), ld s):   columns < s hold Linv^T D^-1 (upper),
    //      columns >= s hold (-L_BL Linv)^T.
    // The forward solve of a node is then one product [y; W_bot] =
    // L g_top + [0; g_bot] and the backward solve one product
    // x = LT [y; x_struct]: one launch per level and direction.
    GemmLaunch fold;                 // LT[:, s:] = (-L_BL Linv)^T
    std::vector<TransItem> foldBack; // L_BL <- LT[:, s:]^T, diag(L) <- 1

    // Solve schedule.  Per depth: forward product (F) and backward product
    // (B).
    // ginv[2*r + c]: row of child c's work vector feeding front row r (or -1).
    std::vector<int32_t> ginv;
    std::vector<MvTile> mv;
    std::vector<MvTile> mvred;  // reduce items of split row blocks
    int maxSlots = 0;           // partial slots of the largest (level, phase)
    std::vector<std::array<SolveShape, 2>> shapes;  // per depth, per phase (F, B)
    struct Range {
        int start = 0, count = 0;    // tiles
        int rows = 64;               // output rows per tile (64, 32 or 16)
        int waves = 4;               // waves per workgroup of the matrix-core path
        int persist = 0;             // persistent pipelined workgroups
        int rstart = 0, rcount = 0;  // reduce items
        int xcd = 0;                 // tiles placed so each row block's slices share an XCD
    };
    struct SolveLevel {
        Range f, b;
        // per phase (F, B): factor entries read and vector rows (inner
        // inputs + outputs) moved, for algorithmic bytes and flops
        double entries[2] = {0, 0};
        double vecRows[2] = {0, 0};
    };
    std::vector<SolveLevel> solveLevels;  // index = depth
    std::vector<TransItem> trans;

    // Fused subtree solves (qcut < 0: none; every level is its own launch).
    // Per phase: the tiles of queue q are tiles[qstart[q] .. qstart[q+1]),
    // forward deepest level first, backward the cut level first; dep[t] =
    // {dep0, dep1, row blocks of dep0, of dep1} names the nodes whose row
    // blocks must be done before node t's tiles read their inputs (forward:
    // its children; backward: its nearest ancestor with pivots, when that
    // is inside the fused levels), -1 for none.
    int qcut = -1;
    int queueRows = 32;  // output rows per queue tile (64, 32 or 16)
    int nqueues = 0;
    struct QueuePhase {
        std::vector<QTile> tiles;
        std::vector<int32_t> qstart;
        std::vector<QRed> reds;
        std::vector<std::array<int32_t, 4>> dep;
        double entries = 0, vecRows = 0;  // algorithmic work of the launch
    };
    QueuePhase queue[2];

    // Algorithmic work per panel.
    double flops = 0, schurFlops = 0, solveFlopsPerRhs = 0;
    double factorEntries = 0;
};

struct Plan {
    Geometry geo;
    ClassPlan cls[3];
    bool present[3] = {false, false, false};
    int blockNb = 32;      // LDL^T block width
    int tile = 64;         // GEMM tile
    int minTiles = 256;    // solve tiles wanted per launch (default shapes)
    int maxSplit = 8;      // inner slices per row block at most (default shapes)
    bool wideWg = false;   // default shapes: 16-wave workgroups instead of slices
    bool persist = true;   // default shapes: persistent pipelined workgroups
    bool fused = false;    // fused subtree launches (8 <= k <= 16)
    int fusedCut = -1;     // cut depth of the fused launches (-1: the first with
                           // at least kQueues nodes)
    int fusedRows = 32;    // output rows per queue tile
    int fusedQueues = kQueues;  // queues (more than the device's XCDs leaves
                                // queues that the last workgroup runs alone)
    int fusedWgPerQueue = 64;   // workgroups a queue expects (inner slices are
                                // cut so a level has about that many tiles)
};

// Builds everything; returns PSP_HIP_OK or an error code with `msg` set.
int BuildPlan(const psp_hip_desc& desc, Plan& plan, std::string& msg);

// Solve tile shapes: the default per (level, phase) is the largest row
// block of 64, 32, 16 rows that gives minTiles tiles, and otherwise inner
// slices (at most maxSplit) to reach minTiles.
void DefaultSolveShapes(ClassPlan& C, int minTiles, int maxSplit, bool wideWg = false,
                        bool persist = false);
// (Re)builds the solve tiles, reduce items and their accounting from
// C.shapes.
void BuildSolveTiles(ClassPlan& C);
// (Re)builds the fused subtree launches of a class (after BuildSolveTiles:
// uses its per-level accounting and raises maxSlots to cover the queues'
// partial slots).  nqueues <= 0 or a tree with no level of nqueues nodes
// leaves qcut = -1.
void BuildQueues(ClassPlan& C, int nqueues, int rows, int cut, int wgPerQueue);

}  // namespace psp_amd
