// Device-side types and kernel launch wrappers shared by the host driver.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "symbolic.hpp"

namespace psp_amd {

template <typename T>
struct __attribute__((aligned(2 * sizeof(T)))) cx {
    T r, i;
};
using c64 = cx<float>;
using c128 = cx<double>;

// Stencil tables: PML profiles s1Inv(x), s2Inv(y), s3Inv(v) for the index
// range [-1, n] stored at offset +1 (initialize_finalize.hpp:415-525).
struct StencilArgs {
    const c128* s1;   // nx + 2
    const c128* s2;   // ny + 2
    const c128* s3;   // nz + 2 (global operator)
    // per-axis coupling factors (1/s(i-1) + 1/s(i)) / (2h^2) and
    // (1/s(i+1) + 1/s(i)) / (2h^2), interleaved L, R per index
    const c128* ax;   // 2 * nx
    const c128* ay;   // 2 * ny
    const c128* av;   // 2 * nz
    int nx, ny, nz;
    double hx, hy, hz;
    double omega;
};

// Per-node data the solve kernels read.
struct NodeDev {
    int64_t l, w;      // factor and work-vector offsets
    int32_t off, s, u; // first panel row, pivots, update size
    int32_t ustart;    // into structs / relinds
    int32_t pw;        // parent's work-vector offset (rows), 0 at a root
    int32_t cslot;     // this node's child slot in its parent (0 or 1), -1 at a root
};

// A front-solve tile (or reduce item) with its node's fields inlined, so a
// workgroup starts with one 64-byte load.
struct SolveTile {
    NodeDev nd;
    int32_t row0, nrows, col0, col1, slot;
    int32_t pad_;
};
static_assert(sizeof(SolveTile) == 64, "SolveTile is one 64-byte record");

// Padding of the solve buffers: the persistent solve kernels read up to this
// many columns (factor blocks) or rows (vectors, struct lists) past a tile,
// so the factor arrays, the panel vector, the work vectors, the slot copies
// and the struct list end in this much zeroed padding.
constexpr int kSolvePad = 64;

// Per-class device copy of the static plan.
struct ClassDev {
    int n = 0, vsize = 0, nnodes = 0, npanels = 0;
    int64_t fstore = 0, ustore = 0, tstore = 0, wstore = 0;
    int32_t *qxy = nullptr, *qv = nullptr, *qnode = nullptr, *qcol = nullptr, *qrows = nullptr;
    int32_t *nat2q = nullptr;  // (x + y*nx + vl*nx*ny) -> padded ND index
    int64_t *nodeL = nullptr, *nodeW = nullptr;
    int32_t *nodeS = nullptr, *nodeU = nullptr, *nodeUs = nullptr;
    int32_t *structs = nullptr, *relinds = nullptr;
    ExtAddItem* extadd = nullptr;
    LdlItem* ldl = nullptr;
    GemmDesc* gemms = nullptr;
    GemmTile* tiles = nullptr;
    InvItem* inv = nullptr;
    NodeDev* nodes = nullptr;
    SolveTile* mv = nullptr;
    SolveTile* mvred = nullptr;     // reduce items of split solve row blocks
    int32_t* redcBuf = nullptr;     // one arrival counter per reduce item (zero between launches)
    int32_t* redc = nullptr;        // redcBuf when split row blocks are summed in the launch
    TransItem* trans = nullptr;     // LT[:, :s] blocks
    TransItem* foldBack = nullptr;  // L_BL / diagonal blocks
    int32_t* ginv = nullptr;
    int ninv = 0, ntrans = 0, nfold = 0;
    int32_t* panelVOff = nullptr;   // per batch entry: panel vOffset
    c128* s3art = nullptr;          // per batch entry: vsize + 2 profile values
    void* L = nullptr;              // npanels * fstore
    void* U = nullptr;              // npanels * ustore
    void* T = nullptr;              // npanels * tstore
    void* LT = nullptr;             // npanels * fstore: per front L^T (s x (s+u), ld s)
    int cus = 0;                    // compute units of the device (persistent grids)
    int mvpGrid = 0;                // persistent solve launches: workgroup cap (0: none)
    // fused subtree solves, per phase (F, B): queue tiles, queue starts
    // (nqueues + 1), split row blocks, per-node dependencies {dep0, dep1,
    // row blocks of dep0, of dep1} (-1: none)
    QTile* qtiles[2] = {nullptr, nullptr};
    int32_t* qstart[2] = {nullptr, nullptr};
    QRed* qreds[2] = {nullptr, nullptr};
    int4* qdep[2] = {nullptr, nullptr};
    int nqueues = 0;                // 0: no fused launch for this class
    int nqtiles[2] = {0, 0}, nqreds[2] = {0, 0};
    int queueRows = 32;             // output rows per fused tile
    int nsync = 0;                  // ints of fused-launch state per class (both phases)
};

// Ints of fused-launch state of one phase: per-node done counters, queue
// heads, the exit counter and one arrival counter per split row block.
inline int QueueSyncInts(int nnodes, int nqueues, int nreds) { return nnodes + nqueues + 1 + nreds; }

// ---- batched sweeps of many problems (one launch per step of all sweeps) ----
// One workgroup of a batched launch: its tile or reduce item (products,
// reduces), its segment and problem, and its workgroup index inside the
// segment (copies).
struct BatchBlock {
    const SolveTile* item;
    int32_t seg, problem, blk, pad_;
};
// One problem's panel solve in a batched product or reduce launch (front
// precision erased: the kernels cast to their cx<T>).
struct BatchSolveSeg {
    const int32_t* structs;
    void* G;                 // slot copies of the panel's class, slot 1 at G + half elements
    int64_t half;
    const int32_t* relinds;
    const void *L, *LT;      // this panel's factors
    void *Xp, *W, *P;
    int32_t k, pad_;
};
enum BatchCopyOp : int { BATCH_COPY_IN = 0, BATCH_COPY_OUT = 1, BATCH_ZERO_SET_PLANE = 2 };
constexpr int kBatchCopyThreads = 256;  // elements per workgroup of a batched copy
// One problem's panel copy in a batched copy launch.
struct BatchCopySeg {
    StencilArgs st;
    const double* vel;
    const int32_t *qxy, *qv;
    c128* B;
    void* Xp;
    int32_t n, k, vOffset, padding, mode, v1, v2, vlLast;
};


// ---- launch wrappers (kernels.hip) ----
void LaunchFillZero(void* p, int64_t bytes, hipStream_t s);
// Zeroes rows ranges of `bytes` bytes, range r at p + r * strideBytes (both
// multiples of 8).
void LaunchFillZeroRows(void* p, int64_t bytes, int rows, int64_t strideBytes, hipStream_t s);
// True when workgroup b of a launch runs on XCD b % 8 (hardware register
// XCC_ID), checked with a 64-workgroup probe on stream s (not capturing).
bool XcdRoundRobin(hipStream_t s);
// Occupancy of the persistent solve kernels, queried once (outside any
// stream capture).
void PrimeSolveOccupancy();

template <typename T>
void LaunchScatterOrig(const ClassDev& c, const StencilArgs& st, const double* vel,
                       double damping, hipStream_t s);
template <typename T>
void LaunchExtendAdd(const ClassDev& c, int start, int count, hipStream_t s);
template <typename T>
void LaunchLdlDiag(const ClassDev& c, int start, int count, int nb, hipStream_t s);
template <typename T>
void LaunchLdlTrsm(const ClassDev& c, int start, int count, int nb, hipStream_t s);
template <typename T>
void LaunchGemm(const ClassDev& c, int tileStart, int tileCount, hipStream_t s);
template <typename T>
void LaunchInvDiag(const ClassDev& c, int nb, hipStream_t s);

// Optional timing of the front-solve product launches: HIP events around
// every launch (synchronizing after each), plus their algorithmic work.
struct MvProfile {
    hipEvent_t a = nullptr, b = nullptr;
    int launches = 0;
    double seconds = 0, bytes = 0, flops = 0;
};

// Solve of one panel (batch entry b of class c) on the padded ND vector Xp
// (n x k, k interleaved) with node work vectors W (wstore x k).
// sync: the class's fused-launch state (cleared by the copy-in or
// zero-set launch before the panel solve); null runs every level as its own
// launch.
template <typename T>
void LaunchPanelSolve(const ClassDev& c, const ClassPlan& plan, int b, void* Xp, void* W,
                      void* P, void* G,
                      int k, hipStream_t s, MvProfile* prof = nullptr, int onlyLevel = -1,
                      int onlyPhase = -1, int32_t* sync = nullptr);
template <typename T>
void LaunchTransposeScale(const ClassDev& c, hipStream_t s);
template <typename T>
void LaunchFoldBack(const ClassDev& c, hipStream_t s);

// Natural-order <-> padded panel vector.  mode 0: B = Xp, 1: B += Xp.
// sync (nsync ints, may be null): the class's fused-launch state, cleared
// by the same launch (the panel solve that follows starts from zero).
template <typename T>
void LaunchCopyIn(const ClassDev& c, const c128* B, void* Xp, int k, int nx, int ny, int nz,
                  int vOffset, int padding, hipStream_t s, int32_t* sync = nullptr);
template <typename T>
void LaunchCopyOut(const ClassDev& c, c128* B, const void* Xp, int k, int nx, int ny, int nz,
                   int vOffset, int padding, int mode, hipStream_t s,
                   const StencilArgs* st = nullptr, const double* vel = nullptr, int v1 = -1);
// Backward sweep, one pass: Xp = 0 except the superdiagonal coupling on the
// panel's last local plane (MultiplySuperdiagonal, solve.hpp:725-733).
template <typename T>
void LaunchZeroSetPlane(const ClassDev& c, const StencilArgs& st, const double* vel,
                        const c128* B, void* Xp, int k, int v2, int vlLast, hipStream_t s,
                        int32_t* sync = nullptr);
// Batched launches over nblocks BatchBlocks: products of phase ph with
// rows-per-tile rows, waves waves per workgroup and kb (8 or 16) right-hand
// sides per pass (k_mvm's tile work), reduce items, or panel copies (op).
template <typename T>
void LaunchSolveBatch(int ph, int rows, int waves, int kb, const BatchBlock* blocks, int nblocks,
                      const BatchSolveSeg* segs, const int32_t* active, hipStream_t s);
template <typename T>
void LaunchReduceBatch(int ph, const BatchBlock* blocks, int nblocks, const BatchSolveSeg* segs,
                       const int32_t* active, hipStream_t s);
template <typename T>
void LaunchCopyBatch(int op, const BatchBlock* blocks, int nblocks, const BatchCopySeg* segs,
                     const int32_t* active, hipStream_t s);

// y = A x (mode 0) or y = b - A x (mode 1), natural order.
void LaunchSpmv(const StencilArgs& st, const double* vel, const c128* x, const c128* b,
                c128* y, int k, int mode, hipStream_t s);

// Column reductions over N x k interleaved blocks (deterministic two-pass).
int ReducePartials(int64_t npoints);
void LaunchColNorm2(const c128* x, int64_t npoints, int k, double* part, double* out,
                    hipStream_t s);  // out[r] = ||x_r||_2
void LaunchColDot(const c128* v, const c128* w, int64_t npoints, int k, c128* part,
                  c128* out, hipStream_t s);  // out[r] = v_r^H w_r
// w -= alpha v, then out = conj(vn)^T w per column (one pass over w; the
// same sums as LaunchAxpyCols followed by LaunchColDot).
void LaunchAxpyDot(const c128* alpha, const c128* v, const c128* vn, c128* w, int64_t np, int k,
                   c128* part, c128* out, hipStream_t s);
void LaunchAxpyCols(const c128* alpha, int sign, const c128* v, c128* w, int64_t npoints,
                    int k, hipStream_t s);  // w += sign*alpha_r v_r
void LaunchScaleCols(const double* delta, const c128* w, c128* v, int64_t npoints, int k,
                     hipStream_t s);  // v = w / delta_r
void LaunchUpdateX(const c128* x0, const c128* V, const c128* y, int j, int64_t npoints, int k,
                   c128* x, hipStream_t s);  // x = x0 + sum_{i<=j} y(i,r) V_i
void LaunchNeg(const c128* x, c128* w, int64_t n, hipStream_t s);
void LaunchAdjointGrad(const StencilArgs& st, const c128* u, const c128* lam, int k,
                       double* grad, int accumulate, hipStream_t s);

}  // namespace psp_amd
