// C ABI implementation: DistHelmholtz's setup, Initialize and Solve on one
// MI355X.  The whole grid, every panel's factors and the Krylov basis stay in
// HBM; only the small GMRES least-squares work (Givens rotations, the
// triangular solve of the Hessenberg matrix) runs on the host, in FP64,
// exactly as InternalSolveWithGMRES (solve.hpp:241-611) does it.
#include <hip/hip_runtime.h>

#include <array>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <mutex>
#include <thread>
#include <algorithm>
#include <vector>

#include "kernels.hpp"
#include "psp_hip.h"
#include "symbolic.hpp"

using namespace psp_amd;
using zd = std::complex<double>;

namespace {

double Now()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct psp_hip_ctx {
    psp_hip_desc desc{};
    Plan plan;
    std::string err;
    bool devReady = false;
    bool factored = false;
    hipStream_t stream = nullptr;
    hipStream_t own = nullptr;
    hipStream_t side[2] = {nullptr, nullptr};  // factorization of panel classes 1 and 2
    std::vector<void*> allocs;  // plan + factor storage
    double* vel = nullptr;
    c128 *s1 = nullptr, *s2 = nullptr, *s3 = nullptr;
    c128 *ax = nullptr, *ay = nullptr, *av = nullptr;
    ClassDev dev[3];
    std::vector<int> panelBatch;   // panel -> batch index inside its class
    std::vector<std::vector<int>> classPanels;
    // solve workspace
    int wsK = 0, wsM = 0;
    std::vector<void*> wsAllocs;
    int64_t wsSlots = 0;  // partial-sum slots the workspace holds
    std::vector<void*> facAllocs;  // Schur-complement and inversion scratch, only while factoring
    std::vector<void*> facStore;   // L and LT of every panel, from Initialize to Finalize
    void *Xp = nullptr, *W = nullptr;
    void* P = nullptr;  // partial sums of split solve row blocks (64 rows x k per slot)
    void* G[3] = {nullptr, nullptr, nullptr};  // per class: the children's update rows, two slot copies
    int32_t* S[3] = {nullptr, nullptr, nullptr};  // per class: fused-launch state (counters)
    c128 *V = nullptr, *x = nullptr, *x0 = nullptr, *w = nullptr, *b = nullptr;
    double *partD = nullptr, *normD = nullptr;
    c128 *partC = nullptr, *Hcol = nullptr, *ydev = nullptr;
    // captured sweeps, keyed on (block pointer, k)
    bool useGraph = true;
    int mvpGrid = 0;  // persistent solve launches: workgroup cap (0: none)
    bool xcdReduce = true;  // split row blocks summed in the product launch (option)
    bool xcdOk = false;     // the device deals workgroups to its XCDs round-robin
    bool factorOverlap = true;  // panel classes 1 and 2 factored on side streams (option)
    struct SweepGraph {
        void* B = nullptr;
        int k = 0;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<SweepGraph> graphs;
    psp_hip_iterate_fn iterCb = nullptr;  // called with every unconverged iterate
    void* iterUser = nullptr;
    void* scratch = nullptr;  // device address of this device's no-allocation scratch
};

namespace {

#define HIPCHK(ctx, expr)                                                              \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);            \
            return e_ == hipErrorOutOfMemory ? PSP_HIP_ERR_MEMORY : PSP_HIP_ERR_HIP;   \
        }                                                                              \
    } while (0)

int Fail(psp_hip_ctx* c, int code, const std::string& m)
{
    c->err = m;
    return code;
}

template <typename X>
int Upload(psp_hip_ctx* c, X** dst, const std::vector<X>& src)
{
    *dst = nullptr;
    if (src.empty()) return PSP_HIP_OK;
    void* p = nullptr;
    HIPCHK(c, hipMalloc(&p, sizeof(X) * src.size()));
    c->allocs.push_back(p);
    HIPCHK(c, hipMemcpy(p, src.data(), sizeof(X) * src.size(), hipMemcpyHostToDevice));
    *dst = (X*)p;
    return PSP_HIP_OK;
}

int Alloc(psp_hip_ctx* c, void** p, size_t bytes, std::vector<void*>& list)
{
    *p = nullptr;
    if (bytes == 0) bytes = 16;
    HIPCHK(c, hipMalloc(p, bytes));
    list.push_back(*p);
    return PSP_HIP_OK;
}

// PML profiles (initialize_finalize.hpp:415-525).
zd ProfileValue(double sigma, int b, double h, double omega, double delta)
{
    const double eta = b * h;
    return zd(1.0, (sigma / eta) * (delta / b) * (delta / b) * (2 * M_PI / omega));
}

zd S1Inv(const psp_hip_desc& d, double hx, int x)
{
    if (x < d.bx - 1 && d.frontBC == PSP_HIP_PML)
        return ProfileValue(d.sigmax, d.bx, hx, d.omega, (d.bx - 1) - x);
    if (x > d.nx - d.bx && d.backBC == PSP_HIP_PML)
        return ProfileValue(d.sigmax, d.bx, hx, d.omega, x - (d.nx - d.bx));
    return 1.0;
}

zd S2Inv(const psp_hip_desc& d, double hy, int y)
{
    if (y < d.by - 1 && d.leftBC == PSP_HIP_PML)
        return ProfileValue(d.sigmay, d.by, hy, d.omega, (d.by - 1) - y);
    if (y > d.ny - d.by && d.rightBC == PSP_HIP_PML)
        return ProfileValue(d.sigmay, d.by, hy, d.omega, y - (d.ny - d.by));
    return 1.0;
}

zd S3Inv(const psp_hip_desc& d, double hz, int v, int vOffset, bool artificial)
{
    const int lim = artificial ? vOffset + d.bz - 1 : d.bz - 1;
    if (v < lim) return ProfileValue(d.sigmaz, d.bz, hz, d.omega, lim - v);
    if (v > d.nz - d.bz && d.topBC == PSP_HIP_PML)
        return ProfileValue(d.sigmaz, d.bz, hz, d.omega, v - (d.nz - d.bz));
    return 1.0;
}

std::vector<c128> ToC128(const std::vector<zd>& v)
{
    std::vector<c128> o(v.size());
    for (size_t i = 0; i < v.size(); ++i) o[i] = c128{v[i].real(), v[i].imag()};
    return o;
}

StencilArgs Stencil(const psp_hip_ctx* c)
{
    StencilArgs st;
    st.s1 = c->s1;
    st.s2 = c->s2;
    st.s3 = c->s3;
    st.ax = c->ax;
    st.ay = c->ay;
    st.av = c->av;
    st.nx = c->desc.nx;
    st.ny = c->desc.ny;
    st.nz = c->desc.nz;
    st.hx = c->desc.wx / (c->desc.nx + 1);
    st.hy = c->desc.wy / (c->desc.ny + 1);
    st.hz = c->desc.wz / (c->desc.nz + 1);
    st.omega = c->desc.omega;
    return st;
}

size_t FrontBytes(const psp_hip_ctx* c) { return c->desc.precision == PSP_HIP_FRONTS_FP32 ? 8 : 16; }

// One-time device setup: upload the static plan and allocate factor storage.
// The struct list with one trailing entry, so that a branch-free kernel
// may read structs[ustart] of a front without an update (the root).
std::vector<int32_t> WithSentinel(const std::vector<int>& v, int pad = 1)
{
    std::vector<int32_t> out(v.begin(), v.end());
    out.resize(out.size() + pad, 0);
    return out;
}

NodeDev MakeNodeDev(const ClassPlan& P, int t)
{
    const Node& n = P.nodes[t];
    int32_t pw = 0, cslot = -1;
    if (n.parent >= 0) {
        const Node& par = P.nodes[n.parent];
        pw = (int32_t)par.woff;
        for (int k = 0; k < 2; ++k)
            if (par.child[k] == t) cslot = k;
    }
    return NodeDev{n.loff, n.woff, n.off, n.size, n.usize, (int32_t)n.ustart, pw, cslot};
}

std::vector<SolveTile> SolveTiles(const ClassPlan& P, const std::vector<MvTile>& tiles)
{
    std::vector<SolveTile> out(tiles.size());
    for (size_t i = 0; i < tiles.size(); ++i) {
        const MvTile& m = tiles[i];
        SolveTile& t = out[i];
        t.nd = MakeNodeDev(P, m.node);
        t.row0 = m.row0;
        t.nrows = m.nrows;
        t.col0 = m.col0;
        t.col1 = m.col1;
        t.slot = m.slot;
        t.pad_ = m.red;  // the slice's reduce item (into mvred), -1 if none
    }
    return out;
}

// Zeroed arrival counters, one per reduce item of the class.
std::vector<int32_t> ReduceCounters(const ClassPlan& P)
{
    return std::vector<int32_t>(std::max<size_t>(1, P.mvred.size()), 0);
}

// Whether the product launches of a class sum split row blocks themselves
// (the option, and the device's round-robin XCD dealing, checked once per
// device).
bool XcdDealing(psp_hip_ctx* c)
{
    static std::mutex mu;
    static std::map<int, bool> known;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    std::lock_guard<std::mutex> lock(mu);
    auto it = known.find(dev);
    if (it == known.end()) it = known.emplace(dev, XcdRoundRobin(c->stream)).first;
    return it->second;
}

void SetSliceReduce(psp_hip_ctx* c, ClassDev& D)
{
    D.redc = c->xcdReduce && c->xcdOk ? D.redcBuf : nullptr;
}

int UploadQueues(psp_hip_ctx* c, int cl);

// Scratch of psp_hip_precondition_noalloc, the sweep entry point that
// allocates nothing: the sweep buffers (SweepBufferSizes) live in this
// file-scope device array, one copy per device, held until the process
// ends.  Host code never uses the array's name as an address: each device's
// copy is looked up once (ScratchAddress, at the context's device set-up)
// and that address is what kernels and fills receive.  Calls that overlap on
// one device share the array and race.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
void* gScratchAddr[kMaxDevices] = {};

// Device address of the current device's scratch copy (looked up once per
// device, with that device current).
int ScratchAddress(psp_hip_ctx* c, void** out)
{
    *out = nullptr;
    int dev = 0;
    HIPCHK(c, hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return Fail(c, PSP_HIP_ERR_HIP, "device index out of range");
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (!gScratchAddr[dev]) {
        void* p = nullptr;
        HIPCHK(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        gScratchAddr[dev] = p;
    }
    *out = gScratchAddr[dev];
    return PSP_HIP_OK;
}

int EnsureDevice(psp_hip_ctx* c)
{
    if (c->devReady) return PSP_HIP_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return Fail(c, PSP_HIP_ERR_HIP, "no HIP device available");
    if (!c->stream) {
        HIPCHK(c, hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking));
        c->stream = c->own;
    }
    c->xcdOk = XcdDealing(c);
    const Geometry& g = c->plan.geo;
    c->classPanels.assign(3, {});
    c->panelBatch.assign(g.numPanels, 0);
    for (int p = 0; p < g.numPanels; ++p) {
        const int cl = g.cls(p);
        c->panelBatch[p] = (int)c->classPanels[cl].size();
        c->classPanels[cl].push_back(p);
    }
    const psp_hip_desc& d = c->desc;
    const double hx = d.wx / (d.nx + 1), hy = d.wy / (d.ny + 1), hz = d.wz / (d.nz + 1);
    std::vector<zd> s1(d.nx + 2), s2(d.ny + 2), s3(d.nz + 2);
    for (int x = -1; x <= d.nx; ++x) s1[x + 1] = S1Inv(d, hx, x);
    for (int y = -1; y <= d.ny; ++y) s2[y + 1] = S2Inv(d, hy, y);
    for (int v = -1; v <= d.nz; ++v) s3[v + 1] = S3Inv(d, hz, v, 0, false);
    int st;
    if ((st = Upload(c, &c->s1, ToC128(s1)))) return st;
    if ((st = Upload(c, &c->s2, ToC128(s2)))) return st;
    if ((st = Upload(c, &c->s3, ToC128(s3)))) return st;
    auto axis = [](const std::vector<zd>& sv, int n, double h) {
        std::vector<zd> t(2 * (size_t)n);
        for (int i = 0; i < n; ++i) {
            const zd rm = 1.0 / sv[i + 1];
            t[2 * i] = (1.0 / sv[i] + rm) / (2 * h * h);
            t[2 * i + 1] = (1.0 / sv[i + 2] + rm) / (2 * h * h);
        }
        return t;
    };
    if ((st = Upload(c, &c->ax, ToC128(axis(s1, d.nx, hx))))) return st;
    if ((st = Upload(c, &c->ay, ToC128(axis(s2, d.ny, hy))))) return st;
    if ((st = Upload(c, &c->av, ToC128(axis(s3, d.nz, hz))))) return st;
    void* vp = nullptr;
    if ((st = Alloc(c, &vp, sizeof(double) * (size_t)d.nx * d.ny * d.nz, c->allocs))) return st;
    c->vel = (double*)vp;

    int dev = 0, cus = 0;
    HIPCHK(c, hipGetDevice(&dev));
    HIPCHK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if ((st = ScratchAddress(c, &c->scratch))) return st;
    PrimeSolveOccupancy();
    for (int cl = 0; cl < 3; ++cl) {
        if (!c->plan.present[cl]) continue;
        const ClassPlan& P = c->plan.cls[cl];
        ClassDev& D = c->dev[cl];
        D.cus = cus;
        D.mvpGrid = c->mvpGrid;
        D.n = P.n;
        D.vsize = P.vsize;
        D.nnodes = (int)P.nodes.size();
        D.npanels = (int)c->classPanels[cl].size();
        D.fstore = P.fstore;
        D.ustore = P.ustore;
        D.tstore = P.tstore;
        D.wstore = P.wstore;
        D.ninv = (int)P.inv.size();
        std::vector<int64_t> nL(D.nnodes), nW(D.nnodes);
        std::vector<int32_t> nS(D.nnodes), nU(D.nnodes), nUs(D.nnodes);
        for (int t = 0; t < D.nnodes; ++t) {
            nL[t] = P.nodes[t].loff;
            nW[t] = P.nodes[t].woff;
            nS[t] = P.nodes[t].size;
            nU[t] = P.nodes[t].usize;
            nUs[t] = (int32_t)P.nodes[t].ustart;
        }
        std::vector<int32_t> nat2q(P.n);
        for (int q = 0; q < P.n; ++q) {
            const int x = P.qxy[q] & 0xffff, y = P.qxy[q] >> 16;
            nat2q[x + y * d.nx + P.qv[q] * d.nx * d.ny] = q;
        }
        std::vector<int32_t> vOff(D.npanels);
        std::vector<zd> s3a((size_t)D.npanels * (P.vsize + 2));
        for (int b = 0; b < D.npanels; ++b) {
            const int p = c->classPanels[cl][b];
            vOff[b] = g.vOffset(p);
            for (int vl = -1; vl <= P.vsize; ++vl)
                s3a[(size_t)b * (P.vsize + 2) + vl + 1] =
                    S3Inv(d, hz, vOff[b] + vl, vOff[b], true);
        }
        if ((st = Upload(c, &D.qxy, P.qxy)) || (st = Upload(c, &D.qv, P.qv)) ||
            (st = Upload(c, &D.qnode, P.qnode)) || (st = Upload(c, &D.qcol, P.qcol)) ||
            (st = Upload(c, &D.qrows, P.qrows)) || (st = Upload(c, &D.nat2q, nat2q)) ||
            (st = Upload(c, &D.nodeL, nL)) || (st = Upload(c, &D.nodeW, nW)) ||
            (st = Upload(c, &D.nodeS, nS)) || (st = Upload(c, &D.nodeU, nU)) ||
            (st = Upload(c, &D.nodeUs, nUs)) || (st = Upload(c, &D.structs, WithSentinel(P.structs, kSolvePad + 1))) ||
            (st = Upload(c, &D.relinds, WithSentinel(P.relinds))) || (st = Upload(c, &D.extadd, P.extadd)) ||
            (st = Upload(c, &D.ldl, P.ldl)) || (st = Upload(c, &D.gemms, P.gemms)) ||
            (st = Upload(c, &D.tiles, P.tiles)) || (st = Upload(c, &D.inv, P.inv)) ||
            (st = Upload(c, &D.mv, SolveTiles(P, P.mv))) ||
            (st = Upload(c, &D.mvred, SolveTiles(P, P.mvred))) ||
            (st = Upload(c, &D.redcBuf, ReduceCounters(P))) ||
            (st = Upload(c, &D.trans, P.trans)) || (st = Upload(c, &D.ginv, P.ginv)) ||
            (st = Upload(c, &D.foldBack, P.foldBack)) ||
            (st = Upload(c, &D.panelVOff, vOff)))
            return st;
        std::vector<NodeDev> nd(D.nnodes);
        for (int t = 0; t < D.nnodes; ++t) nd[t] = MakeNodeDev(P, t);
        if ((st = Upload(c, &D.nodes, nd))) return st;
        SetSliceReduce(c, D);
        D.ntrans = (int)P.trans.size();
        D.nfold = (int)P.foldBack.size();
        std::vector<c128> s3c = ToC128(s3a);
        if ((st = Upload(c, &D.s3art, s3c))) return st;
        if ((st = UploadQueues(c, cl))) return st;
    }
    c->devReady = true;
    return PSP_HIP_OK;
}

// Factor storage (L and LT of every panel), allocated by the first
// Initialize after construction or Finalize and kept until Finalize.
int EnsureFactorStorage(psp_hip_ctx* c)
{
    const size_t eb = FrontBytes(c);
    for (int cl = 0; cl < 3; ++cl) {
        if (!c->plan.present[cl]) continue;
        ClassDev& D = c->dev[cl];
        if (D.L) continue;
        const ClassPlan& P = c->plan.cls[cl];
        // factor arrays end in kSolvePad columns of the tallest front (zeros)
        int64_t maxFs = 1;
        for (const Node& nd : P.nodes) maxFs = std::max<int64_t>(maxFs, nd.size + nd.usize);
        const size_t fb = eb * ((size_t)D.npanels * D.fstore + (size_t)kSolvePad * maxFs);
        int st;
        if ((st = Alloc(c, &D.L, fb, c->facStore))) return st;
        if ((st = Alloc(c, &D.LT, fb, c->facStore))) return st;
        // only the padding needs zeroing here: every factorization zeroes L
        // before assembly, and the fold writes every entry of LT
        const size_t body = eb * (size_t)D.npanels * D.fstore;
        HIPCHK(c, hipMemsetAsync((char*)D.L + body, 0, fb - body, c->stream));
        HIPCHK(c, hipMemsetAsync((char*)D.LT + body, 0, fb - body, c->stream));
    }
    return PSP_HIP_OK;
}

void FreeFactorStorage(psp_hip_ctx* c)
{
    for (void* p : c->facStore) (void)hipFree(p);
    c->facStore.clear();
    for (int cl = 0; cl < 3; ++cl) {
        c->dev[cl].L = nullptr;
        c->dev[cl].LT = nullptr;
    }
    c->factored = false;
}

void FreeGraphs(psp_hip_ctx* c)
{
    for (auto& g : c->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    c->graphs.clear();
}

void FreeFactorScratch(psp_hip_ctx* c)
{
    for (void* p : c->facAllocs) (void)hipFree(p);
    c->facAllocs.clear();
    for (int cl = 0; cl < 3; ++cl) {
        c->dev[cl].U = nullptr;
        c->dev[cl].T = nullptr;
    }
}

void FreeWorkspace(psp_hip_ctx* c)
{
    FreeGraphs(c);
    for (void* p : c->wsAllocs) (void)hipFree(p);
    c->wsAllocs.clear();
    c->wsK = c->wsM = 0;
}

int64_t NeededSlots(const psp_hip_ctx* c)
{
    int64_t maxs = 1;
    for (int cl = 0; cl < 3; ++cl)
        if (c->plan.present[cl]) maxs = std::max<int64_t>(maxs, c->plan.cls[cl].maxSlots);
    return maxs;
}

// Byte sizes of the sweep's buffers for k right-hand sides: the padded
// panel vector Xp, the node work vectors W, the partial sums P of split
// solve row blocks and, per panel class, the two slot copies G of the
// children's update rows.  Vectors and slot copies end in kSolvePad + 2
// rows that must read as zero (the solve kernels read past a tile's rows).
struct SweepSizes {
    size_t xp = 0, w = 0, p = 0, g[3] = {0, 0, 0}, s[3] = {0, 0, 0};
};

SweepSizes SweepBufferSizes(const psp_hip_ctx* c, int k)
{
    const size_t eb = FrontBytes(c);
    int64_t maxn = 0, maxw = 0, maxs = 1;
    for (int cl = 0; cl < 3; ++cl)
        if (c->plan.present[cl]) {
            maxn = std::max<int64_t>(maxn, c->plan.cls[cl].n);
            maxw = std::max<int64_t>(maxw, c->plan.cls[cl].wstore);
            maxs = std::max<int64_t>(maxs, c->plan.cls[cl].maxSlots);
        }
    const size_t padRows = (size_t)kSolvePad + 2;
    SweepSizes z;
    z.xp = eb * (maxn + padRows) * k;
    z.w = eb * (maxw + padRows) * k;
    z.p = eb * maxs * 64 * k;
    for (int cl = 0; cl < 3; ++cl)
        if (c->plan.present[cl]) {
            z.g[cl] = eb * (2 * (size_t)c->plan.cls[cl].wstore + padRows) * k;
            z.s[cl] = sizeof(int32_t) * (size_t)std::max(c->dev[cl].nsync, 1);
        }
    return z;
}

int EnsureWorkspace(psp_hip_ctx* c, int k, int m)
{
    if (c->wsK == k && c->wsM >= m && c->wsSlots >= NeededSlots(c)) return PSP_HIP_OK;
    FreeWorkspace(c);
    const int64_t N = (int64_t)c->desc.nx * c->desc.ny * c->desc.nz;
    const SweepSizes z = SweepBufferSizes(c, k);
    int st;
    void* p;
    if ((st = Alloc(c, &c->Xp, z.xp, c->wsAllocs))) return st;
    if ((st = Alloc(c, &c->W, z.w, c->wsAllocs))) return st;
    HIPCHK(c, hipMemsetAsync(c->Xp, 0, z.xp, c->stream));
    HIPCHK(c, hipMemsetAsync(c->W, 0, z.w, c->stream));
    if ((st = Alloc(c, &c->P, z.p, c->wsAllocs))) return st;
    for (int cl = 0; cl < 3; ++cl) {
        c->G[cl] = nullptr;
        if (!c->plan.present[cl]) continue;
        // Zero once: each class writes the same rows of its slot copies in
        // every solve, and the rows it never writes must read as zero.
        if ((st = Alloc(c, &c->G[cl], z.g[cl], c->wsAllocs))) return st;
        HIPCHK(c, hipMemsetAsync(c->G[cl], 0, z.g[cl], c->stream));
        void* sp = nullptr;
        if ((st = Alloc(c, &sp, z.s[cl], c->wsAllocs))) return st;
        c->S[cl] = (int32_t*)sp;
    }
    if ((st = Alloc(c, &p, sizeof(c128) * N * k * m, c->wsAllocs))) return st;
    c->V = (c128*)p;
    if ((st = Alloc(c, &p, sizeof(c128) * N * k, c->wsAllocs))) return st;
    c->x = (c128*)p;
    if ((st = Alloc(c, &p, sizeof(c128) * N * k, c->wsAllocs))) return st;
    c->x0 = (c128*)p;
    if ((st = Alloc(c, &p, sizeof(c128) * N * k, c->wsAllocs))) return st;
    c->w = (c128*)p;
    if ((st = Alloc(c, &p, sizeof(c128) * N * k, c->wsAllocs))) return st;
    c->b = (c128*)p;
    const int g = ReducePartials(N);
    if ((st = Alloc(c, &p, sizeof(double) * g * k, c->wsAllocs))) return st;
    c->partD = (double*)p;
    if ((st = Alloc(c, &p, sizeof(c128) * g * k, c->wsAllocs))) return st;
    c->partC = (c128*)p;
    if ((st = Alloc(c, &p, sizeof(double) * k, c->wsAllocs))) return st;
    c->normD = (double*)p;
    if ((st = Alloc(c, &p, sizeof(c128) * (m + 1) * k, c->wsAllocs))) return st;
    c->Hcol = (c128*)p;
    if ((st = Alloc(c, &p, sizeof(c128) * (m + 1) * k, c->wsAllocs))) return st;
    c->ydev = (c128*)p;
    c->wsK = k;
    c->wsM = m;
    c->wsSlots = NeededSlots(c);
    return PSP_HIP_OK;
}

// Replaces a device array uploaded by Upload with new contents.
template <typename X>
int ReUpload(psp_hip_ctx* c, X** dst, const std::vector<X>& src)
{
    if (*dst) {
        auto it = std::find(c->allocs.begin(), c->allocs.end(), (void*)*dst);
        if (it != c->allocs.end()) c->allocs.erase(it);
        HIPCHK(c, hipFree(*dst));
        *dst = nullptr;
    }
    return Upload(c, dst, src);
}

// Device copies of a class's fused subtree launches (none when the plan has
// no cut or the option is off).
int UploadQueues(psp_hip_ctx* c, int cl)
{
    const ClassPlan& P = c->plan.cls[cl];
    ClassDev& D = c->dev[cl];
    const bool on = P.qcut >= 0 && P.nqueues > 0;
    for (int ph = 0; ph < 2; ++ph) {
        const ClassPlan::QueuePhase& Q = P.queue[ph];
        std::vector<int4> dep;
        if (on) {
            dep.resize(Q.dep.size());
            for (size_t t = 0; t < Q.dep.size(); ++t)
                dep[t] = make_int4(Q.dep[t][0], Q.dep[t][1], Q.dep[t][2], Q.dep[t][3]);
        }
        int st;
        if ((st = ReUpload(c, &D.qtiles[ph], on ? Q.tiles : std::vector<QTile>{})) ||
            (st = ReUpload(c, &D.qstart[ph], on ? Q.qstart : std::vector<int32_t>{})) ||
            (st = ReUpload(c, &D.qreds[ph], on ? Q.reds : std::vector<QRed>{})) ||
            (st = ReUpload(c, &D.qdep[ph], dep)))
            return st;
        D.nqtiles[ph] = on ? (int)Q.tiles.size() : 0;
        D.nqreds[ph] = on ? (int)Q.reds.size() : 0;
    }
    D.nqueues = on ? P.nqueues : 0;
    D.queueRows = P.queueRows;
    D.nsync = on ? QueueSyncInts(D.nnodes, D.nqueues, D.nqreds[0]) +
                       QueueSyncInts(D.nnodes, D.nqueues, D.nqreds[1])
                 : 0;
    return PSP_HIP_OK;
}

void BuildClassQueues(psp_hip_ctx* c, ClassPlan& P)
{
    const Plan& pl = c->plan;
    BuildQueues(P, pl.fused ? pl.fusedQueues : 0, pl.fusedRows, pl.fusedCut, pl.fusedWgPerQueue);
}

// Rebuilds the front-solve tiles of every class from its shapes and, once
// the device is set up, replaces the device copies; the workspace (partial
// buffer sized by the slots) and the captured sweep graphs are dropped.
int RebuildSolveTiles(psp_hip_ctx* c)
{
    if (c->devReady) HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int cl = 0; cl < 3; ++cl) {
        if (!c->plan.present[cl]) continue;
        ClassPlan& P = c->plan.cls[cl];
        BuildSolveTiles(P);
        BuildClassQueues(c, P);
        if (!c->devReady) continue;
        ClassDev& D = c->dev[cl];
        int st;
        if ((st = ReUpload(c, &D.mv, SolveTiles(P, P.mv))) ||
            (st = ReUpload(c, &D.mvred, SolveTiles(P, P.mvred))) ||
            (st = ReUpload(c, &D.redcBuf, ReduceCounters(P))) || (st = UploadQueues(c, cl)))
            return st;
        SetSliceReduce(c, D);
    }
    FreeWorkspace(c);
    return PSP_HIP_OK;
}

template <typename T>
// The panel classes are independent.  Class 0 (the bottom and full inner
// panels, factored as one batch) runs on the context's stream; the leftover
// and top panels (classes 1 and 2, one panel each, whose launches are
// latency-bound) run on side streams at the same time (unless the context's
// factorOverlap option is off).  Phase times are summed over the classes.
int FactorClasses(psp_hip_ctx* c, double* tFill, double* tLdl, double* tInv, double* tSchur)
{
    const StencilArgs st = Stencil(c);
    for (hipStream_t& x : c->side)
        if (!x) HIPCHK(c, hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
    const hipStream_t ss[3] = {c->stream, c->factorOverlap ? c->side[0] : c->stream,
                               c->factorOverlap ? c->side[1] : c->stream};
    // every event of the factorization, destroyed on any return; an event is
    // named by its index (the list grows while events are in use)
    struct Events {
        std::vector<hipEvent_t> v;
        ~Events()
        {
            for (hipEvent_t x : v)
                if (x) (void)hipEventDestroy(x);
        }
    } evs;
    auto newEvent = [&](unsigned flags, int* idx) -> int {
        hipEvent_t x = nullptr;
        HIPCHK(c, hipEventCreateWithFlags(&x, flags));
        evs.v.push_back(x);
        *idx = (int)evs.v.size() - 1;
        return PSP_HIP_OK;
    };
    int rc, fork;
    // fork: the velocity is uploaded; per class: phase bounds e[0..3]
    if ((rc = newEvent(hipEventDisableTiming, &fork))) return rc;
    HIPCHK(c, hipEventRecord(evs.v[fork], c->stream));
    int ce[3][4] = {{-1}, {-1}, {-1}};
    std::vector<std::array<int, 2>> se;  // around every level's Schur-complement GEMM
    for (int cl = 0; cl < 3; ++cl) {
        if (!c->plan.present[cl]) continue;
        const ClassPlan& P = c->plan.cls[cl];
        const ClassDev& D = c->dev[cl];
        const int nb = c->plan.blockNb;
        hipStream_t s = ss[cl];
        for (int i = 0; i < 4; ++i)
            if ((rc = newEvent(hipEventDefault, &ce[cl][i]))) return rc;
        auto mark = [&](int i) { return hipEventRecord(evs.v[ce[cl][i]], s); };
        if (cl > 0) HIPCHK(c, hipStreamWaitEvent(s, evs.v[fork], 0));
        HIPCHK(c, mark(0));
        LaunchFillZero(D.L, (int64_t)sizeof(cx<T>) * D.npanels * D.fstore, s);
        LaunchScatterOrig<T>(D, st, c->vel, c->desc.damping, s);
        HIPCHK(c, mark(1));
        for (int dpt = P.maxDepth; dpt >= 0; --dpt) {
            const ClassPlan::LevelSteps& S = P.steps[dpt];
            // this level's updates start at zero (its region held the updates
            // of depth dpt + 2, read by the extend-add of depth dpt + 1)
            LaunchFillZeroRows((cx<T>*)D.U + S.uoff, (int64_t)sizeof(cx<T>) * S.ucount, D.npanels,
                               (int64_t)sizeof(cx<T>) * D.ustore, s);
            for (int slot = 0; slot < 2; ++slot)
                for (const auto& r : S.extadd[slot]) LaunchExtendAdd<T>(D, r.first, r.second, s);
            for (size_t kb = 0; kb < S.ldlDiag.size(); ++kb) {
                LaunchLdlDiag<T>(D, S.ldlDiag[kb].first, S.ldlDiag[kb].second, nb, s);
                LaunchLdlTrsm<T>(D, S.ldlTrsm[kb].first, S.ldlTrsm[kb].second, nb, s);
                LaunchGemm<T>(D, S.trail[kb].tileStart, S.trail[kb].tileCount, s);
            }
            if (S.schur.tileCount > 0) {
                std::array<int, 2> g;
                if ((rc = newEvent(hipEventDefault, &g[0])) ||
                    (rc = newEvent(hipEventDefault, &g[1])))
                    return rc;
                HIPCHK(c, hipEventRecord(evs.v[g[0]], s));
                LaunchGemm<T>(D, S.schur.tileStart, S.schur.tileCount, s);
                HIPCHK(c, hipEventRecord(evs.v[g[1]], s));
                se.push_back(g);
            }
        }
        HIPCHK(c, mark(2));
        LaunchInvDiag<T>(D, nb, s);
        for (const GemmLaunch& gl : P.invGemm) LaunchGemm<T>(D, gl.tileStart, gl.tileCount, s);
        // solve layout: LT[:, :s] = Linv^T D^-1, LT[:, s:] = (-L_BL Linv)^T, then
        // L_BL = -L_BL Linv and a unit diagonal in L
        LaunchTransposeScale<T>(D, s);
        LaunchGemm<T>(D, P.fold.tileStart, P.fold.tileCount, s);
        LaunchFoldBack<T>(D, s);
        HIPCHK(c, mark(3));
        HIPCHK(c, hipGetLastError());
        if (cl > 0) HIPCHK(c, hipStreamWaitEvent(c->stream, evs.v[ce[cl][3]], 0));  // join
    }
    float acc[3] = {0, 0, 0};
    double schur = 0;
    for (int cl = 0; cl < 3; ++cl) {
        if (ce[cl][0] < 0) continue;
        HIPCHK(c, hipEventSynchronize(evs.v[ce[cl][3]]));
        float ms;
        for (int i = 0; i < 3; ++i) {
            HIPCHK(c, hipEventElapsedTime(&ms, evs.v[ce[cl][i]], evs.v[ce[cl][i + 1]]));
            acc[i] += ms;
        }
    }
    for (const auto& g : se) {
        float ms;
        HIPCHK(c, hipEventElapsedTime(&ms, evs.v[g[0]], evs.v[g[1]]));
        schur += ms * 1e-3;
    }
    *tSchur = schur;
    *tFill = acc[0] * 1e-3;
    *tLdl = acc[1] * 1e-3;
    *tInv = acc[2] * 1e-3;
    return PSP_HIP_OK;
}

// ---------------------------------------------------------------------------
// Sweep: Precondition (solve.hpp:681-734) on a natural-order device block.
// ---------------------------------------------------------------------------
// couple >= 0: the copy-out also applies the subdiagonal coupling into the
// next panel's first plane `couple` (SolvePanel + MultiplySubdiagonal).
template <typename T>
void PanelForward(psp_hip_ctx* c, c128* B, int k, int p, MvProfile* prof = nullptr,
                  int couple = -1)
{
    const Geometry& g = c->plan.geo;
    const int cl = g.cls(p);
    const ClassDev& D = c->dev[cl];
    hipStream_t s = c->stream;
    const StencilArgs st = Stencil(c);
    int32_t* sync = D.nqueues > 0 ? c->S[cl] : nullptr;
    LaunchCopyIn<T>(D, B, c->Xp, k, g.nx, g.ny, g.nz, g.vOffset(p), g.padding(p), s, sync);
    LaunchPanelSolve<T>(D, c->plan.cls[cl], c->panelBatch[p], c->Xp, c->W, c->P, c->G[cl], k, s,
                        prof, -1, -1, sync);
    LaunchCopyOut<T>(D, B, c->Xp, k, g.nx, g.ny, g.nz, g.vOffset(p), g.padding(p), 0, s,
                     couple >= 0 ? &st : nullptr, c->vel, couple);
}

template <typename T>
void PanelBackward(psp_hip_ctx* c, c128* B, int k, int p, MvProfile* prof = nullptr)
{
    // ExtractPanel + MultiplySuperdiagonal + SolvePanel + UpdatePanel
    // (solve.hpp:725-733): B_i still holds Z, so the solve of the coupling
    // term is added onto it.
    const Geometry& g = c->plan.geo;
    const int cl = g.cls(p);
    const ClassDev& D = c->dev[cl];
    hipStream_t s = c->stream;
    // Xp = 0 except the superdiagonal coupling on the panel's last plane
    const int v2 = g.panelV(p) + g.depth(p) - 1;
    const int vlLast = g.padding(p) + g.depth(p) - 1;
    int32_t* sync = D.nqueues > 0 ? c->S[cl] : nullptr;
    LaunchZeroSetPlane<T>(D, Stencil(c), c->vel, B, c->Xp, k, v2, vlLast, s, sync);
    LaunchPanelSolve<T>(D, c->plan.cls[cl], c->panelBatch[p], c->Xp, c->W, c->P, c->G[cl], k, s,
                        prof, -1, -1, sync);
    LaunchCopyOut<T>(D, B, c->Xp, k, g.nx, g.ny, g.nz, g.vOffset(p), g.padding(p), 1, s);
}

template <typename T>
void Sweep(psp_hip_ctx* c, c128* B, int k, MvProfile* prof = nullptr)
{
    const Geometry& g = c->plan.geo;
    for (int i = 0; i < g.numPanels - 1; ++i) PanelForward<T>(c, B, k, i, prof, g.panelV(i + 1));
    PanelForward<T>(c, B, k, g.numPanels - 1, prof);
    for (int i = g.numPanels - 2; i >= 0; --i) PanelBackward<T>(c, B, k, i, prof);
}

void SweepEager(psp_hip_ctx* c, c128* B, int k)
{
    if (c->desc.precision == PSP_HIP_FRONTS_FP32)
        Sweep<float>(c, B, k);
    else
        Sweep<double>(c, B, k);
}

// Precondition: replays a hipGraph of the whole sweep (captured on first use
// for this block and k) or launches it kernel by kernel.
int Precondition(psp_hip_ctx* c, c128* B, int k)
{
    if (!c->useGraph || c->stream == nullptr) {
        SweepEager(c, B, k);
        return PSP_HIP_OK;
    }
    for (auto& g : c->graphs)
        if (g.B == B && g.k == k) {
            HIPCHK(c, hipGraphLaunch(g.exec, c->stream));
            return PSP_HIP_OK;
        }
    psp_hip_ctx::SweepGraph g;
    g.B = B;
    g.k = k;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    SweepEager(c, B, k);
    HIPCHK(c, hipStreamEndCapture(c->stream, &g.graph));
    HIPCHK(c, hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
    c->graphs.push_back(g);
    HIPCHK(c, hipGraphLaunch(g.exec, c->stream));
    return PSP_HIP_OK;
}

int Norms(psp_hip_ctx* c, const c128* X, int k, std::vector<double>& out)
{
    const int64_t N = (int64_t)c->desc.nx * c->desc.ny * c->desc.nz;
    LaunchColNorm2(X, N, k, c->partD, c->normD, c->stream);
    out.resize(k);
    HIPCHK(c, hipMemcpyAsync(out.data(), c->normD, sizeof(double) * k, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSP_HIP_OK;
}

bool HasNaN(const std::vector<double>& v)
{
    for (double a : v)
        if (a != a) return true;
    return false;
}

// lapack::ComputeGivens (solve.hpp:486): zlartg semantics,
// [c s; -conj(s) c] [f; g] = [r; 0].
void Givens(zd f, zd g, double& c, zd& s, zd& r)
{
    if (g == 0.0) {
        c = 1.0;
        s = 0.0;
        r = f;
        return;
    }
    if (f == 0.0) {
        c = 0.0;
        s = std::conj(g) / std::abs(g);
        r = std::abs(g);
        return;
    }
    const double af = std::abs(f), ag = std::abs(g);
    const double dd = std::hypot(af, ag);
    const zd ph = f / af;
    c = af / dd;
    s = ph * std::conj(g) / dd;
    r = ph * dd;
}

// Step j of right-hand side r's least-squares update (solve.hpp:427-470):
// column j of its Hessenberg matrix from the Arnoldi coefficients hc
// (element (i, r) at hc[i k + r]), the earlier rotations applied, a new
// rotation against delta, the rotated right-hand side z, and
// y = H(0:j, 0:j)^-1 z(0:j) into yh[i k + r].  Hr (H(i, j) = Hr[i + j m]),
// csr, snr and zr are this right-hand side's own arrays.  Returns nullptr,
// or the reason the step gave a NaN.
const char* LeastSquaresStep(int m, int k, int j, int r, const c128* hc, double delta, zd* Hr,
                             double* csr, zd* snr, zd* zr, c128* yh)
{
    for (int i = 0; i <= j; ++i) {
        const c128 a = hc[(size_t)i * k + r];
        Hr[i + (size_t)j * m] = zd(a.r, a.i);
    }
    for (int i = 0; i < j; ++i) {  // previous rotations (:442-456)
        const double cc = csr[i];
        const zd sv = snr[i];
        const zd e0 = Hr[i + (size_t)j * m], e1 = Hr[i + 1 + (size_t)j * m];
        Hr[i + (size_t)j * m] = cc * e0 + sv * e1;
        Hr[i + 1 + (size_t)j * m] = -std::conj(sv) * e0 + cc * e1;
    }
    const zd ejj = Hr[j + (size_t)j * m];
    if (ejj != ejj) return "H(j,j) was NaN";
    double cc;
    zd sv, rho;
    Givens(ejj, zd(delta), cc, sv, rho);
    if (cc != cc || sv != sv || rho != rho) return "Givens rotation was NaN";
    Hr[j + (size_t)j * m] = rho;
    csr[j] = cc;
    snr[j] = sv;
    const zd zj = zr[j], zj1 = zr[j + 1];
    zr[j] = cc * zj + sv * zj1;
    zr[j + 1] = -std::conj(sv) * zj + cc * zj1;
    std::vector<zd> y(zr, zr + j + 1);  // Trsv(UPPER, NORMAL, NON_UNIT)
    for (int i = j; i >= 0; --i) {
        zd a = y[i];
        for (int l = i + 1; l <= j; ++l) a -= Hr[i + (size_t)l * m] * y[l];
        y[i] = a / Hr[i + (size_t)i * m];
    }
    for (int i = 0; i <= j; ++i) yh[(size_t)i * k + r] = c128{y[i].real(), y[i].imag()};
    return nullptr;
}

struct GmresOut {
    int it = 0, restarts = 0;
    double maxRel = 0;
    double tP = 0, tM = 0, tA = 0;
    int nP = 0, nM = 0;
};

// InternalSolveWithGMRES (solve.hpp:241-611): k independent GMRES(m) sharing
// operator applications; x recomputed as x0 + V y every step; true residual
// every step; stop when the largest relative residual is below relTol.
// steps > 0: run exactly that many inner steps and record the history.
int Gmres(psp_hip_ctx* c, int k, int m, double relTol, int maxIts, int steps, double* hist,
          GmresOut& o)
{
    const int64_t N = (int64_t)c->desc.nx * c->desc.ny * c->desc.nz;
    const int64_t n = N * k;
    hipStream_t s = c->stream;
    const StencilArgs st = Stencil(c);
    std::vector<double> orig, beta, delta, resid;
    std::vector<zd> H((size_t)m * m * k), z((size_t)(m + 1) * k), sn((size_t)m * k);
    std::vector<double> cs((size_t)m * k);
    std::vector<c128> hc((size_t)(m + 1) * k), yh((size_t)(m + 1) * k);
    int rc;
    // Phase times from events on the stream, read at the host waits the
    // method needs (the Hessenberg column, the residual norms), so the
    // operator, the sweep and the Arnoldi steps queue back to back.
    struct Events {
        hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
        ~Events()
        {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    for (hipEvent_t& x : ev.e) HIPCHK(c, hipEventCreate(&x));
    auto span = [&](int a, int b) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ev.e[a], ev.e[b]);
        return ms * 1e-3;
    };

    LaunchFillZero(c->x, sizeof(c128) * n, s);
    HIPCHK(c, hipMemcpyAsync(c->w, c->b, sizeof(c128) * n, hipMemcpyDeviceToDevice, s));
    if ((rc = Norms(c, c->w, k, orig))) return rc;
    if (HasNaN(orig)) return Fail(c, PSP_HIP_ERR_NAN, "Original residual norms had a NaN");

    bool converged = false;
    while (!converged) {
        HIPCHK(c, hipMemcpyAsync(c->x0, c->x, sizeof(c128) * n, hipMemcpyDeviceToDevice, s));
        double t0 = Now();
        if ((rc = Precondition(c, c->w, k))) return rc;
        if ((rc = Norms(c, c->w, k, beta))) return rc;
        o.tP += Now() - t0;
        ++o.nP;
        if (HasNaN(beta)) return Fail(c, PSP_HIP_ERR_NAN, "beta list had a NaN");
        LaunchScaleCols(c->normD, c->w, c->V, N, k, s);
        std::fill(z.begin(), z.end(), zd(0));
        for (int r = 0; r < k; ++r) z[(size_t)r * (m + 1)] = beta[r];
        for (int j = 0; j < m; ++j) {
            c128* vj = c->V + (int64_t)j * n;
            HIPCHK(c, hipEventRecord(ev.e[0], s));
            LaunchSpmv(st, c->vel, vj, nullptr, c->w, k, 0, s);
            HIPCHK(c, hipEventRecord(ev.e[1], s));
            if ((rc = Precondition(c, c->w, k))) return rc;
            HIPCHK(c, hipEventRecord(ev.e[2], s));
            // Arnoldi by modified Gram-Schmidt (solve.hpp:382-399): step i's
            // update of w and step i + 1's dot product share one pass
            LaunchColDot(c->V, c->w, N, k, c->partC, c->Hcol, s);
            for (int i = 0; i < j; ++i)
                LaunchAxpyDot(c->Hcol + (int64_t)i * k, c->V + (int64_t)i * n,
                              c->V + (int64_t)(i + 1) * n, c->w, N, k, c->partC,
                              c->Hcol + (int64_t)(i + 1) * k, s);
            LaunchAxpyCols(c->Hcol + (int64_t)j * k, -1, c->V + (int64_t)j * n, c->w, N, k, s);
            if ((rc = Norms(c, c->w, k, delta))) return rc;
            HIPCHK(c, hipMemcpyAsync(hc.data(), c->Hcol, sizeof(c128) * (j + 1) * k,
                                     hipMemcpyDeviceToHost, s));
            HIPCHK(c, hipEventRecord(ev.e[3], s));
            HIPCHK(c, hipStreamSynchronize(s));
            o.tM += span(0, 1);
            ++o.nM;
            o.tP += span(1, 2);
            ++o.nP;
            o.tA += span(2, 3);
            if (HasNaN(delta)) return Fail(c, PSP_HIP_ERR_NAN, "delta list had a NaN");
            for (int r = 0; r < k; ++r)
                if (delta[r] == 0.0)
                    return Fail(c, PSP_HIP_ERR_BREAKDOWN,
                                "GMRES halted due to a (usually) lucky breakdown");
            if (j + 1 != m) LaunchScaleCols(c->normD, c->w, c->V + (int64_t)(j + 1) * n, N, k, s);
            for (int r = 0; r < k; ++r)
                if (const char* bad = LeastSquaresStep(
                        m, k, j, r, hc.data(), delta[r], H.data() + (size_t)r * m * m,
                        cs.data() + (size_t)r * m, sn.data() + (size_t)r * m,
                        z.data() + (size_t)r * (m + 1), yh.data()))
                    return Fail(c, PSP_HIP_ERR_NAN, bad);
            HIPCHK(c, hipMemcpyAsync(c->ydev, yh.data(), sizeof(c128) * (j + 1) * k,
                                     hipMemcpyHostToDevice, s));
            LaunchUpdateX(c->x0, c->V, c->ydev, j, N, k, c->x, s);  // x := x0 + V y
            HIPCHK(c, hipEventRecord(ev.e[0], s));
            LaunchSpmv(st, c->vel, c->x, c->b, c->w, k, 1, s);  // w := b - A x
            HIPCHK(c, hipEventRecord(ev.e[1], s));
            if ((rc = Norms(c, c->w, k, resid))) return rc;  // waits for the stream
            o.tM += span(0, 1);
            ++o.nM;
            if (HasNaN(resid)) return Fail(c, PSP_HIP_ERR_NAN, "resid norm list has NaN");
            double mx = 0;
            for (int r = 0; r < k; ++r) mx = std::max(mx, resid[r] / orig[r]);
            o.maxRel = mx;
            if (hist && o.it < steps) hist[o.it] = mx;
            ++o.it;
            if (steps > 0) {
                if (o.it >= steps) {
                    converged = true;
                    break;
                }
            } else if (mx < relTol) {
                converged = true;
                break;
            }
            if (c->iterCb) {  // viewIterates (solve.hpp:579-585): the unconverged iterate
                std::vector<c128> xh((size_t)n);
                HIPCHK(c, hipMemcpy(xh.data(), c->x, sizeof(c128) * n, hipMemcpyDeviceToHost));
                c->iterCb(c->iterUser, o.it - 1, xh.data(), k);
            }
            if (maxIts > 0 && o.it >= maxIts)
                return Fail(c, PSP_HIP_ERR_NOCONV, "GMRES reached the iteration cap");
        }
        ++o.restarts;
    }
    return PSP_HIP_OK;
}

int CheckSolveArgs(psp_hip_ctx* c, const void* B, int k, int m)
{
    if (!c) return PSP_HIP_ERR_INVALID;
    if (!c->factored) return Fail(c, PSP_HIP_ERR_STATE, "Solve called before Initialize");
    if (!B) return Fail(c, PSP_HIP_ERR_INVALID, "null right-hand side");
    if (k < 1 || k > 256) return Fail(c, PSP_HIP_ERR_INVALID, "number of RHS must be in [1, 256]");
    if (m < 1) return Fail(c, PSP_HIP_ERR_INVALID, "GMRES restart length must be >= 1");
    return PSP_HIP_OK;
}

int RunSolve(psp_hip_ctx* c, void* B, int k, int m, double relTol, int maxIts, int flags,
             int steps, double* hist, psp_hip_solve_report* rep)
{
    int rc = CheckSolveArgs(c, B, k, m);
    if (rc) return rc;
    const double t0 = Now();
    if ((rc = EnsureWorkspace(c, k, m))) return rc;
    const int64_t n = (int64_t)c->desc.nx * c->desc.ny * c->desc.nz * k;
    const bool dev = flags & PSP_HIP_DEVICE_POINTER;
    // PullRightHandSides: natural order is the working order on the GPU.
    HIPCHK(c, hipMemcpyAsync(c->b, B, sizeof(c128) * n,
                             dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
    GmresOut o;
    const int st = Gmres(c, k, m, relTol, maxIts, steps, hist, o);
    // PushRightHandSides: the solution (also on breakdown, cap or NaN).
    HIPCHK(c, hipMemcpyAsync(B, c->x, sizeof(c128) * n,
                             dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (rep) {
        rep->status = st;
        rep->iterations = o.it;
        rep->restarts = o.restarts;
        rep->final_max_rel = o.maxRel;
        rep->seconds_total = Now() - t0;
        rep->seconds_precond = o.tP;
        rep->seconds_multiply = o.tM;
        rep->seconds_arnoldi = o.tA;
        rep->num_precond = o.nP;
        rep->num_multiply = o.nM;
    }
    return st;
}

// ---------------------------------------------------------------------------
// Batched solves of independent problems (psp_hip_solve_batch): the sweeps of
// all problems advance together, one launch per step (a panel copy, or one
// tree level's products or reduce) for every problem at that step.
// ---------------------------------------------------------------------------
// The merged sweep: problem i's Precondition (solve.hpp:681-734) on its GMRES
// vector w, the steps of all problems aligned as
//   forward panel t = 0, 1, ...: copy-in, panel solve, copy-out with the
//   subdiagonal coupling into panel t + 1;
//   backward step t = 0, 1, ...: panel P_i - 2 - t: zero-set-plane with the
//   superdiagonal coupling, panel solve, additive copy-out;
// where a panel solve is the forward phase over the tree levels from the
// leaves (level d at step D - d, D the deepest level over all problems, so
// the roots line up) and the backward phase from the roots.
// A step's products go in one launch per tile shape (rows per tile, waves)
// present at that step.  Each problem's kernels see the tiles, factors and
// buffers of its own single-problem sweep, in the same order.
struct BatchSweep {
    enum Kind { PRODUCTS = 0, REDUCE = 1, COPY = 2 };
    struct Launch {
        int group = 0;  // problem group: its launches form one branch of the graph
        int kind = PRODUCTS, ph = 0, rows = 64, waves = 4, kb = 16, op = 0;
        int64_t blk0 = 0;
        int nblocks = 0;
    };
    std::vector<Launch> launches;
    int groups = 1;
    std::vector<BatchBlock> blocks;
    std::vector<BatchSolveSeg> solveSegs;
    std::vector<BatchCopySeg> copySegs;
    BatchBlock* dBlocks = nullptr;
    BatchSolveSeg* dSolve = nullptr;
    BatchCopySeg* dCopy = nullptr;
    int32_t* dActive = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~BatchSweep()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void* p : {(void*)dBlocks, (void*)dSolve, (void*)dCopy, (void*)dActive})
            if (p) (void)hipFree(p);
    }
};

// Appends the merged sweep of problems members (of cs: contexts with their
// workspaces for ks right-hand sides) on their GMRES vectors w, as the
// launches of problem group `group`.
void BuildBatchSweep(const std::vector<psp_hip_ctx*>& cs, const std::vector<int>& ks,
                     const std::vector<int>& members, int group, BatchSweep& bs)
{
    // stage key: sweep direction (0 forward, 1 backward), panel step,
    // sub-step (0: first copy; 1 .. 1 + D: forward phase, level D + 1 - sub;
    // 2 + D .. 2 + 2D: backward phase, level sub - 2 - D; 3 + 2D: copy-out),
    // part (0 products, 1 reduce), kernel shape (rows * 10000 + waves * 100 +
    // right-hand sides per pass: 8 for k <= 8, else 16; 0 for copies)
    using Key = std::array<int, 5>;
    struct Entry {
        int problem, seg;
        const SolveTile* items;
        int count;  // tiles / reduce items; copies: workgroups
    };
    std::map<Key, std::vector<Entry>> stages;
    int dmax = 0;
    for (int i : members)
        for (int cl = 0; cl < 3; ++cl)
            if (cs[i]->plan.present[cl]) dmax = std::max(dmax, cs[i]->plan.cls[cl].maxDepth);
    const int subOut = 2 * dmax + 3;
    for (int i : members) {
        psp_hip_ctx* c = cs[i];
        const int k = ks[i];
        const Geometry& g = c->plan.geo;
        const StencilArgs st = Stencil(c);
        const bool f32 = c->desc.precision == PSP_HIP_FRONTS_FP32;
        // the copy launch's op comes from its stage (first copy or copy-out)
        auto copySeg = [&](int p, int mode, int v1) {
            const ClassDev& D = c->dev[g.cls(p)];
            BatchCopySeg s{};
            s.st = st;
            s.vel = c->vel;
            s.qxy = D.qxy;
            s.qv = D.qv;
            s.B = c->w;
            s.Xp = c->Xp;
            s.n = D.n;
            s.k = k;
            s.vOffset = g.vOffset(p);
            s.padding = g.padding(p);
            s.mode = mode;
            s.v1 = v1;
            s.v2 = g.panelV(p) + g.depth(p) - 1;
            s.vlLast = g.padding(p) + g.depth(p) - 1;
            bs.copySegs.push_back(s);
            const int nb = (int)(((int64_t)D.n * k + kBatchCopyThreads - 1) / kBatchCopyThreads);
            return Entry{i, (int)bs.copySegs.size() - 1, nullptr, nb};
        };
        auto solveSeg = [&](int p) {
            const int cl = g.cls(p);
            const ClassDev& D = c->dev[cl];
            const size_t eb = f32 ? sizeof(c64) : sizeof(c128);
            const int64_t b = c->panelBatch[p];
            BatchSolveSeg s{};
            s.structs = D.structs;
            s.G = c->G[cl];
            s.half = D.wstore * k;
            s.relinds = D.relinds;
            s.L = (const char*)D.L + eb * b * D.fstore;
            s.LT = (const char*)D.LT + eb * b * D.fstore;
            s.Xp = c->Xp;
            s.W = c->W;
            s.P = c->P;
            s.k = k;
            bs.solveSegs.push_back(s);
            return (int)bs.solveSegs.size() - 1;
        };
        auto levels = [&](int dir, int t, int p) {
            const int cl = g.cls(p);
            const ClassPlan& plan = c->plan.cls[cl];
            const ClassDev& D = c->dev[cl];
            const int seg = solveSeg(p);
            const int kb = k <= 8 ? 8 : 16;
            for (int ph = 0; ph < 2; ++ph)
                for (int d = 0; d <= plan.maxDepth; ++d) {
                    const ClassPlan::SolveLevel& lv = plan.solveLevels[d];
                    const ClassPlan::Range& r = ph == 0 ? lv.f : lv.b;
                    const int sub = ph == 0 ? 1 + dmax - d : 2 + dmax + d;
                    if (r.count > 0)
                        stages[Key{dir, t, sub, 0, r.rows * 10000 + r.waves * 100 + kb}]
                            .push_back(Entry{i, seg, D.mv + r.start, r.count});
                    if (r.rcount > 0)
                        stages[Key{dir, t, sub, 1, 0}].push_back(
                            Entry{i, seg, D.mvred + r.rstart, r.rcount});
                }
        };
        for (int p = 0; p < g.numPanels; ++p) {
            stages[Key{0, p, 0, 0, 0}].push_back(copySeg(p, 0, -1));
            levels(0, p, p);
            const int couple = p + 1 < g.numPanels ? g.panelV(p + 1) : -1;
            stages[Key{0, p, subOut, 0, 0}].push_back(copySeg(p, 0, couple));
        }
        for (int t = 0; t + 1 < g.numPanels; ++t) {
            const int p = g.numPanels - 2 - t;
            stages[Key{1, t, 0, 0, 0}].push_back(copySeg(p, 0, -1));
            levels(1, t, p);
            stages[Key{1, t, subOut, 0, 0}].push_back(copySeg(p, 1, -1));
        }
    }
    for (const auto& kv : stages) {
        const Key& key = kv.first;
        BatchSweep::Launch L;
        L.group = group;
        L.ph = key[2] >= 2 + dmax ? 1 : 0;  // panel-solve phase: 0 forward, 1 backward
        const bool copy = key[2] == 0 || key[2] == subOut;
        L.kind = copy ? BatchSweep::COPY : key[3] == 0 ? BatchSweep::PRODUCTS : BatchSweep::REDUCE;
        if (copy)
            L.op = key[2] == 0 ? (key[0] == 0 ? BATCH_COPY_IN : BATCH_ZERO_SET_PLANE)
                               : BATCH_COPY_OUT;
        L.rows = key[4] / 10000;
        L.waves = key[4] / 100 % 100;
        L.kb = key[4] % 100;
        L.blk0 = (int64_t)bs.blocks.size();
        for (const Entry& e : kv.second)
            for (int j = 0; j < e.count; ++j) {
                BatchBlock b{};
                b.item = copy ? nullptr : e.items + j;
                b.seg = e.seg;
                b.problem = e.problem;
                b.blk = j;
                bs.blocks.push_back(b);
            }
        L.nblocks = (int)((int64_t)bs.blocks.size() - L.blk0);
        bs.launches.push_back(L);
    }
}

template <typename T>
void LaunchBatchSweep(const BatchSweep& bs, const hipStream_t* gs)
{
    for (const BatchSweep::Launch& L : bs.launches) {
        const hipStream_t s = gs[L.group];
        const BatchBlock* b = bs.dBlocks + L.blk0;
        if (L.kind == BatchSweep::PRODUCTS)
            LaunchSolveBatch<T>(L.ph, L.rows, L.waves, L.kb, b, L.nblocks, bs.dSolve, bs.dActive,
                                s);
        else if (L.kind == BatchSweep::REDUCE)
            LaunchReduceBatch<T>(L.ph, b, L.nblocks, bs.dSolve, bs.dActive, s);
        else
            LaunchCopyBatch<T>(L.op, b, L.nblocks, bs.dCopy, bs.dActive, s);
    }
}

// Device copies of the merged sweep's work lists and its captured graph:
// the groups' launches are parallel branches, forked from and joined to
// stream s (the other groups captured through streams of their own).
int UploadBatchSweep(psp_hip_ctx* c, BatchSweep& bs, int nproblems, bool f32, hipStream_t s)
{
    auto up = [&](auto** dst, const auto& v) -> int {
        const size_t bytes = sizeof(v[0]) * std::max<size_t>(v.size(), 1);
        HIPCHK(c, hipMalloc((void**)dst, bytes));
        if (!v.empty())
            HIPCHK(c, hipMemcpy(*dst, v.data(), sizeof(v[0]) * v.size(), hipMemcpyHostToDevice));
        return PSP_HIP_OK;
    };
    int rc;
    if ((rc = up(&bs.dBlocks, bs.blocks))) return rc;
    if ((rc = up(&bs.dSolve, bs.solveSegs))) return rc;
    if ((rc = up(&bs.dCopy, bs.copySegs))) return rc;
    if ((rc = up(&bs.dActive, std::vector<int32_t>(nproblems, 1)))) return rc;
    struct Scoped {
        std::vector<hipStream_t> streams;
        std::vector<hipEvent_t> events;
        ~Scoped()
        {
            for (hipStream_t x : streams) (void)hipStreamDestroy(x);
            for (hipEvent_t x : events) (void)hipEventDestroy(x);
        }
    } sc;
    std::vector<hipStream_t> gs(bs.groups, s);
    for (int g = 1; g < bs.groups; ++g) {
        HIPCHK(c, hipStreamCreateWithFlags(&gs[g], hipStreamNonBlocking));
        sc.streams.push_back(gs[g]);
    }
    for (int g = 0; g < bs.groups; ++g) {
        hipEvent_t e;
        HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        sc.events.push_back(e);
    }
    HIPCHK(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    HIPCHK(c, hipEventRecord(sc.events[0], s));  // fork
    for (int g = 1; g < bs.groups; ++g) HIPCHK(c, hipStreamWaitEvent(gs[g], sc.events[0], 0));
    if (f32)
        LaunchBatchSweep<float>(bs, gs.data());
    else
        LaunchBatchSweep<double>(bs, gs.data());
    for (int g = 1; g < bs.groups; ++g) {  // join
        HIPCHK(c, hipEventRecord(sc.events[g], gs[g]));
        HIPCHK(c, hipStreamWaitEvent(s, sc.events[g], 0));
    }
    HIPCHK(c, hipStreamEndCapture(s, &bs.graph));
    HIPCHK(c, hipGraphInstantiate(&bs.exec, bs.graph, nullptr, nullptr, 0));
    return PSP_HIP_OK;
}

// Host memory the device writes and reads between waits (pinned, so copies
// queue without blocking the host).
struct Pinned {
    void* p = nullptr;
    ~Pinned()
    {
        if (p) (void)hipHostFree(p);
    }
};

// InternalSolveWithGMRES (solve.hpp:241-611) for many problems in lockstep:
// problem i runs exactly the steps of Gmres on its own context, in the same
// order, with the sweeps of all unfinished problems merged (BatchSweep) and
// one host wait per step for all of them.  A problem that converges, fails
// or reaches maxIts stops; its workgroups in the merged sweep exit at once.
int GmresBatch(const std::vector<psp_hip_ctx*>& cs, const std::vector<int>& ks, int m,
               double relTol, int maxIts, BatchSweep& bs, hipStream_t s,
               std::vector<GmresOut>& outs, std::vector<int>& status)
{
    const int np = (int)cs.size();
    psp_hip_ctx* c0 = cs[0];
    struct Prob {
        int64_t N = 0, n = 0;
        int k = 0;
        bool active = true;
        std::vector<double> orig;
        std::vector<zd> H, z, sn;
        std::vector<double> cs;
        double *nrm = nullptr;    // pinned: k norms
        c128 *hc = nullptr, *yh = nullptr;  // pinned: (m + 1) k each
    };
    std::vector<Prob> P(np);
    // each piece starts on a 64-byte boundary (c128 is 16-byte aligned)
    auto round64 = [](size_t b) { return (b + 63) & ~(size_t)63; };
    size_t pinBytes = 0;
    for (int i = 0; i < np; ++i)
        pinBytes += round64(sizeof(double) * ks[i]) + 2 * round64(sizeof(c128) * (m + 1) * ks[i]);
    Pinned pin;
    HIPCHK(c0, hipHostMalloc(&pin.p, pinBytes));
    char* q = (char*)pin.p;
    for (int i = 0; i < np; ++i) {
        Prob& p = P[i];
        p.k = ks[i];
        p.N = (int64_t)cs[i]->desc.nx * cs[i]->desc.ny * cs[i]->desc.nz;
        p.n = p.N * p.k;
        p.H.assign((size_t)m * m * p.k, 0);
        p.z.assign((size_t)(m + 1) * p.k, 0);
        p.sn.assign((size_t)m * p.k, 0);
        p.cs.assign((size_t)m * p.k, 0);
        p.nrm = (double*)q;
        q += round64(sizeof(double) * p.k);
        p.hc = (c128*)q;
        q += round64(sizeof(c128) * (m + 1) * p.k);
        p.yh = (c128*)q;
        q += round64(sizeof(c128) * (m + 1) * p.k);
    }
    std::vector<int32_t> act(np, 1);
    bool actDirty = false;
    auto stop = [&](int i, int st) {
        P[i].active = false;
        act[i] = 0;
        actDirty = true;
        status[i] = st;
    };
    auto fail = [&](int i, int code, const char* msg) {
        cs[i]->err = msg;
        stop(i, code);
    };
    auto anyActive = [&]() {
        for (const Prob& p : P)
            if (p.active) return true;
        return false;
    };
    // norms of problem i's w into its pinned buffer
    auto norms = [&](int i) -> int {
        psp_hip_ctx* c = cs[i];
        LaunchColNorm2(c->w, P[i].N, P[i].k, c->partD, c->normD, s);
        HIPCHK(c, hipMemcpyAsync(P[i].nrm, c->normD, sizeof(double) * P[i].k,
                                 hipMemcpyDeviceToHost, s));
        return PSP_HIP_OK;
    };
    auto sweep = [&]() -> int {
        if (actDirty) {
            HIPCHK(c0, hipMemcpyAsync(bs.dActive, act.data(), sizeof(int32_t) * np,
                                      hipMemcpyHostToDevice, s));
            HIPCHK(c0, hipStreamSynchronize(s));  // act may change before the copy ran
            actDirty = false;
        }
        HIPCHK(c0, hipGraphLaunch(bs.exec, s));
        return PSP_HIP_OK;
    };
    struct Events {
        hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
        ~Events()
        {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    for (hipEvent_t& x : ev.e) HIPCHK(c0, hipEventCreate(&x));
    auto span = [&](int a, int b) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ev.e[a], ev.e[b]);
        return ms * 1e-3;
    };
    double tP = 0, tM = 0, tA = 0;
    int rc;

    for (int i = 0; i < np; ++i) {
        psp_hip_ctx* c = cs[i];
        LaunchFillZero(c->x, sizeof(c128) * P[i].n, s);
        HIPCHK(c, hipMemcpyAsync(c->w, c->b, sizeof(c128) * P[i].n, hipMemcpyDeviceToDevice, s));
        if ((rc = norms(i))) return rc;
    }
    HIPCHK(c0, hipStreamSynchronize(s));
    for (int i = 0; i < np; ++i) {
        P[i].orig.assign(P[i].nrm, P[i].nrm + P[i].k);
        if (HasNaN(P[i].orig)) fail(i, PSP_HIP_ERR_NAN, "Original residual norms had a NaN");
    }
    while (anyActive()) {
        std::vector<bool> inCycle(np);
        for (int i = 0; i < np; ++i) inCycle[i] = P[i].active;
        for (int i = 0; i < np; ++i)
            if (P[i].active)
                HIPCHK(c0, hipMemcpyAsync(cs[i]->x0, cs[i]->x, sizeof(c128) * P[i].n,
                                          hipMemcpyDeviceToDevice, s));
        double t0 = Now();
        if ((rc = sweep())) return rc;
        for (int i = 0; i < np; ++i)
            if (P[i].active && (rc = norms(i))) return rc;
        HIPCHK(c0, hipStreamSynchronize(s));
        tP += Now() - t0;
        for (int i = 0; i < np; ++i) {
            if (!P[i].active) continue;
            Prob& p = P[i];
            psp_hip_ctx* c = cs[i];
            ++outs[i].nP;
            std::vector<double> beta(p.nrm, p.nrm + p.k);
            if (HasNaN(beta)) {
                fail(i, PSP_HIP_ERR_NAN, "beta list had a NaN");
                continue;
            }
            LaunchScaleCols(c->normD, c->w, c->V, p.N, p.k, s);
            std::fill(p.z.begin(), p.z.end(), zd(0));
            for (int r = 0; r < p.k; ++r) p.z[(size_t)r * (m + 1)] = beta[r];
        }
        for (int j = 0; j < m && anyActive(); ++j) {
            HIPCHK(c0, hipEventRecord(ev.e[0], s));
            for (int i = 0; i < np; ++i)
                if (P[i].active)
                    LaunchSpmv(Stencil(cs[i]), cs[i]->vel, cs[i]->V + (int64_t)j * P[i].n, nullptr,
                               cs[i]->w, P[i].k, 0, s);
            HIPCHK(c0, hipEventRecord(ev.e[1], s));
            if ((rc = sweep())) return rc;
            HIPCHK(c0, hipEventRecord(ev.e[2], s));
            for (int i = 0; i < np; ++i) {
                if (!P[i].active) continue;
                psp_hip_ctx* c = cs[i];
                const int64_t N = P[i].N, n = P[i].n;
                const int k = P[i].k;
                LaunchColDot(c->V, c->w, N, k, c->partC, c->Hcol, s);
                for (int l = 0; l < j; ++l)
                    LaunchAxpyDot(c->Hcol + (int64_t)l * k, c->V + (int64_t)l * n,
                                  c->V + (int64_t)(l + 1) * n, c->w, N, k, c->partC,
                                  c->Hcol + (int64_t)(l + 1) * k, s);
                LaunchAxpyCols(c->Hcol + (int64_t)j * k, -1, c->V + (int64_t)j * n, c->w, N, k, s);
                if ((rc = norms(i))) return rc;
                HIPCHK(c, hipMemcpyAsync(P[i].hc, c->Hcol, sizeof(c128) * (j + 1) * k,
                                         hipMemcpyDeviceToHost, s));
            }
            HIPCHK(c0, hipEventRecord(ev.e[3], s));
            HIPCHK(c0, hipStreamSynchronize(s));
            tM += span(0, 1);
            tP += span(1, 2);
            tA += span(2, 3);
            for (int i = 0; i < np; ++i) {
                if (!P[i].active) continue;
                Prob& p = P[i];
                psp_hip_ctx* c = cs[i];
                const int k = p.k;
                const int64_t N = p.N, n = p.n;
                ++outs[i].nM;
                ++outs[i].nP;
                std::vector<double> delta(p.nrm, p.nrm + k);
                if (HasNaN(delta)) {
                    fail(i, PSP_HIP_ERR_NAN, "delta list had a NaN");
                    continue;
                }
                bool brk = false;
                for (int r = 0; r < k; ++r) brk = brk || delta[r] == 0.0;
                if (brk) {
                    fail(i, PSP_HIP_ERR_BREAKDOWN, "GMRES halted due to a (usually) lucky breakdown");
                    continue;
                }
                if (j + 1 != m) LaunchScaleCols(c->normD, c->w, c->V + (int64_t)(j + 1) * n, N, k, s);
                const char* bad = nullptr;
                for (int r = 0; r < k && !bad; ++r)
                    bad = LeastSquaresStep(m, k, j, r, p.hc, delta[r], p.H.data() + (size_t)r * m * m,
                                           p.cs.data() + (size_t)r * m, p.sn.data() + (size_t)r * m,
                                           p.z.data() + (size_t)r * (m + 1), p.yh);
                if (bad) {
                    fail(i, PSP_HIP_ERR_NAN, bad);
                    continue;
                }
                HIPCHK(c, hipMemcpyAsync(c->ydev, p.yh, sizeof(c128) * (j + 1) * k,
                                         hipMemcpyHostToDevice, s));
                LaunchUpdateX(c->x0, c->V, c->ydev, j, N, k, c->x, s);  // x := x0 + V y
                LaunchSpmv(Stencil(c), c->vel, c->x, c->b, c->w, k, 1, s);  // w := b - A x
                if ((rc = norms(i))) return rc;
            }
            HIPCHK(c0, hipStreamSynchronize(s));
            for (int i = 0; i < np; ++i) {
                if (!P[i].active) continue;
                Prob& p = P[i];
                psp_hip_ctx* c = cs[i];
                GmresOut& o = outs[i];
                ++o.nM;
                std::vector<double> resid(p.nrm, p.nrm + p.k);
                if (HasNaN(resid)) {
                    fail(i, PSP_HIP_ERR_NAN, "resid norm list has NaN");
                    continue;
                }
                double mx = 0;
                for (int r = 0; r < p.k; ++r) mx = std::max(mx, resid[r] / p.orig[r]);
                o.maxRel = mx;
                ++o.it;
                if (mx < relTol) {
                    stop(i, PSP_HIP_OK);
                    continue;
                }
                if (c->iterCb) {  // viewIterates (solve.hpp:579-585)
                    std::vector<c128> xh((size_t)p.n);
                    HIPCHK(c, hipMemcpy(xh.data(), c->x, sizeof(c128) * p.n, hipMemcpyDeviceToHost));
                    c->iterCb(c->iterUser, o.it - 1, xh.data(), p.k);
                }
                if (maxIts > 0 && o.it >= maxIts)
                    fail(i, PSP_HIP_ERR_NOCONV, "GMRES reached the iteration cap");
            }
        }
        // a cycle counts as a restart for the problems that ran it to its end
        // or converged in it (as Gmres counts them)
        for (int i = 0; i < np; ++i)
            if (inCycle[i] && (P[i].active || status[i] == PSP_HIP_OK)) ++outs[i].restarts;
    }
    for (GmresOut& o : outs) {
        o.tP = tP;
        o.tM = tM;
        o.tA = tA;
    }
    return PSP_HIP_OK;
}

// psp_hip_solve_batch with the merged sweep (problems of one front
// precision): checks, workspaces, the right-hand sides in, lockstep GMRES,
// the solutions out.
int SolveBatchMerged(psp_hip_ctx* const* ctxs, void* const* Bs, const int* ks, int n, int m,
                     double relTol, int maxIts, int flags, psp_hip_solve_report* reports)
{
    const double t0 = Now();
    std::vector<psp_hip_ctx*> cs(ctxs, ctxs + n);
    std::vector<int> kv(ks, ks + n);
    psp_hip_ctx* c0 = cs[0];
    int rc;
    for (int i = 0; i < n; ++i) {
        if ((rc = CheckSolveArgs(cs[i], Bs[i], kv[i], m))) return rc;
        if ((rc = EnsureWorkspace(cs[i], kv[i], m))) return rc;
    }
    for (psp_hip_ctx* c : cs) HIPCHK(c, hipStreamSynchronize(c->stream));
    hipStream_t s = c0->stream;
    const bool dev = flags & PSP_HIP_DEVICE_POINTER;
    for (int i = 0; i < n; ++i) {
        const int64_t e = (int64_t)cs[i]->desc.nx * cs[i]->desc.ny * cs[i]->desc.nz * kv[i];
        HIPCHK(cs[i], hipMemcpyAsync(cs[i]->b, Bs[i], sizeof(c128) * e,
                                     dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    }
    // Two problem groups of about equal work (grid points times right-hand
    // sides, dealt largest first), each group's merged sweep one branch of
    // the graph: the two chains of dependent launches overlap.  Measured on
    // 32 problems of 40^3..96^3: 1, 2, 4 and 8 groups gave 56.7, 61.2, 59.6
    // and 57.0 RHS/s.
    constexpr int kBatchGroups = 2;
    BatchSweep bs;
    bs.groups = std::min(n, kBatchGroups);
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    auto work = [&](int i) {
        return (double)cs[i]->desc.nx * cs[i]->desc.ny * cs[i]->desc.nz * kv[i];
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
    std::vector<std::vector<int>> members(bs.groups);
    std::vector<double> load(bs.groups, 0);
    for (int i : order) {
        const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        members[g].push_back(i);
        load[g] += work(i);
    }
    for (int g = 0; g < bs.groups; ++g) {
        std::sort(members[g].begin(), members[g].end());
        BuildBatchSweep(cs, kv, members[g], g, bs);
    }
    if ((rc = UploadBatchSweep(c0, bs, n, c0->desc.precision == PSP_HIP_FRONTS_FP32, s)))
        return rc;
    std::vector<GmresOut> outs(n);
    std::vector<int> st(n, PSP_HIP_OK);
    if ((rc = GmresBatch(cs, kv, m, relTol, maxIts, bs, s, outs, st))) return rc;
    for (int i = 0; i < n; ++i) {
        const int64_t e = (int64_t)cs[i]->desc.nx * cs[i]->desc.ny * cs[i]->desc.nz * kv[i];
        HIPCHK(cs[i], hipMemcpyAsync(Bs[i], cs[i]->x, sizeof(c128) * e,
                                     dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    }
    HIPCHK(c0, hipStreamSynchronize(s));
    const double t1 = Now();
    for (int i = 0; i < n && reports; ++i) {
        psp_hip_solve_report* rep = &reports[i];
        const GmresOut& o = outs[i];
        rep->status = st[i];
        rep->iterations = o.it;
        rep->restarts = o.restarts;
        rep->final_max_rel = o.maxRel;
        rep->seconds_total = t1 - t0;
        rep->seconds_precond = o.tP;
        rep->seconds_multiply = o.tM;
        rep->seconds_arnoldi = o.tA;
        rep->num_precond = o.nP;
        rep->num_multiply = o.nM;
    }
    for (int i = 0; i < n; ++i)
        if (st[i] != PSP_HIP_OK) return st[i];
    return PSP_HIP_OK;
}

}  // namespace

namespace {

// Tile-shape tuner for the front solves: for every panel class, level and
// phase, times the launch (products + reduce) of each candidate shape on the
// class's panels (one launch per panel, cold factors) and keeps the fastest.  The estimate of one sweep's
// product time weighs each class by its panel solves per sweep (2P - 1 in
// all: every panel forward, every panel but the last backward).
template <typename T>
int Autotune(psp_hip_ctx* c, int k, double* before, double* after)
{
    // Candidate shapes replace the device tile lists that captured sweeps
    // read: drop the graphs first, and on any failure put the old shapes and
    // their tiles back.
    HIPCHK(c, hipStreamSynchronize(c->stream));
    FreeGraphs(c);
    std::array<std::vector<std::array<SolveShape, 2>>, 3> saved;
    for (int cl = 0; cl < 3; ++cl)
        if (c->plan.present[cl]) saved[cl] = c->plan.cls[cl].shapes;
    auto restore = [&](int code) {
        (void)hipStreamSynchronize(c->stream);
        for (int cl = 0; cl < 3; ++cl)
            if (c->plan.present[cl]) c->plan.cls[cl].shapes = saved[cl];
        const std::string msg = c->err;
        (void)RebuildSolveTiles(c);
        c->err = msg;
        return code;
    };
    int rc = EnsureWorkspace(c, k, 1);
    if (rc) return restore(rc);
    hipStream_t s = c->stream;
    hipEvent_t ea, eb;
    HIPCHK(c, hipEventCreate(&ea));
    HIPCHK(c, hipEventCreate(&eb));
    const Geometry& g = c->plan.geo;
    double tb = 0, ta = 0;
    for (int cl = 0; cl < 3 && !rc; ++cl) {
        if (!c->plan.present[cl]) continue;
        ClassPlan& P = c->plan.cls[cl];
        ClassDev& D = c->dev[cl];
        double weight = 0;
        for (int p = 0; p < g.numPanels; ++p)
            if (g.cls(p) == cl) weight += p + 1 < g.numPanels ? 2 : 1;
        auto timeShape = [&](int d, int ph, SolveShape sh, double* t) -> int {
            P.shapes[d][ph] = sh;
            BuildSolveTiles(P);
            BuildClassQueues(c, P);  // keeps the partial slots of the fused launches
            int st;
            if ((st = ReUpload(c, &D.mv, SolveTiles(P, P.mv))) ||
                (st = ReUpload(c, &D.mvred, SolveTiles(P, P.mvred))) ||
                (st = ReUpload(c, &D.redcBuf, ReduceCounters(P))) ||
                (st = EnsureWorkspace(c, k, 1)))
                return st;
            SetSliceReduce(c, D);
            const size_t eb_ = FrontBytes(c);
            LaunchFillZero(c->Xp, (int64_t)(eb_ * P.n * k), s);
            LaunchFillZero(c->W, (int64_t)(eb_ * P.wstore * k), s);
            LaunchPanelSolve<T>(D, P, 0, c->Xp, c->W, c->P, c->G[cl], k, s, nullptr, d, ph);  // warm
            // each repetition applies another panel's factors (where the
            // class has several), so they come from HBM as in a sweep rather
            // than from the caches the previous repetition filled
            constexpr int kReps = 3;
            HIPCHK(c, hipEventRecord(ea, s));
            for (int r = 0; r < kReps; ++r)
                LaunchPanelSolve<T>(D, P, (r + 1) % D.npanels, c->Xp, c->W, c->P, c->G[cl], k, s,
                                    nullptr, d, ph);
            HIPCHK(c, hipEventRecord(eb, s));
            HIPCHK(c, hipEventSynchronize(eb));
            HIPCHK(c, hipGetLastError());
            float ms = 0;
            HIPCHK(c, hipEventElapsedTime(&ms, ea, eb));
            *t = ms * 1e-3 / kReps;
            return PSP_HIP_OK;
        };
        for (int d = 0; d <= P.maxDepth && !rc; ++d)
            for (int ph = 0; ph < 2 && !rc; ++ph) {
                const SolveShape cur = P.shapes[d][ph];
                SolveShape best = cur;
                double tcur = 0;
                if ((rc = timeShape(d, ph, cur, &tcur))) break;
                double tbest = tcur;
                std::vector<SolveShape> cands;
                for (int rows : {64, 32, 16})
                    for (int split : {1, 2, 4, 8, 16}) cands.push_back(SolveShape{rows, split, 4});
                if (k >= 8 && k < 32)  // 16-wave workgroups exist on the matrix-core path
                    for (int rows : {32, 16})
                        for (int split : {1, 2, 4}) cands.push_back(SolveShape{rows, split, 16});
                if (k >= 8 && k <= 16)  // persistent pipelined workgroups
                    for (int rows : {64, 32, 16})
                        for (int split : {1, 2, 4}) cands.push_back(SolveShape{rows, split, 4, 1});
                for (const SolveShape& cand : cands) {
                    if (rc || (cand.rows == cur.rows && cand.split == cur.split &&
                               cand.waves == cur.waves && cand.persist == cur.persist))
                        continue;
                    double t = 0;
                    if ((rc = timeShape(d, ph, cand, &t))) break;
                    if (t < 0.97 * tbest) {
                        tbest = t;
                        best = cand;
                    }
                }
                P.shapes[d][ph] = best;
                tb += weight * tcur;
                ta += weight * tbest;
            }
    }
    (void)hipEventDestroy(ea);
    (void)hipEventDestroy(eb);
    if (rc) return restore(rc);
    if ((rc = RebuildSolveTiles(c))) return restore(rc);
    if (before) *before = tb;
    if (after) *after = ta;
    return PSP_HIP_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int psp_hip_create(const psp_hip_desc* desc, psp_hip_ctx** out)
{
    if (!out) return PSP_HIP_ERR_INVALID;
    *out = nullptr;
    if (!desc) return PSP_HIP_ERR_INVALID;
    if (desc->precision != PSP_HIP_FRONTS_FP64 && desc->precision != PSP_HIP_FRONTS_FP32)
        return PSP_HIP_ERR_INVALID;
    psp_hip_ctx* c = new psp_hip_ctx();
    c->desc = *desc;
    std::string msg;
    const int st = BuildPlan(*desc, c->plan, msg);
    if (st != PSP_HIP_OK) {
        delete c;
        return st;
    }
    *out = c;
    return PSP_HIP_OK;
}

int psp_hip_finalize(psp_hip_ctx* c)
{
    if (!c) return PSP_HIP_ERR_INVALID;
    if (c->devReady) (void)hipStreamSynchronize(c->stream);
    FreeWorkspace(c);
    FreeFactorScratch(c);
    FreeFactorStorage(c);
    return PSP_HIP_OK;
}

int psp_hip_set_iterate_callback(psp_hip_ctx* c, psp_hip_iterate_fn fn, void* user)
{
    if (!c) return PSP_HIP_ERR_INVALID;
    c->iterCb = fn;
    c->iterUser = user;
    return PSP_HIP_OK;
}

int psp_hip_destroy(psp_hip_ctx* c)
{
    if (!c) return PSP_HIP_OK;
    if (c->devReady) (void)hipStreamSynchronize(c->stream);
    FreeWorkspace(c);
    FreeFactorScratch(c);
    FreeFactorStorage(c);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->own) (void)hipStreamDestroy(c->own);
    for (hipStream_t x : c->side)
        if (x) (void)hipStreamDestroy(x);
    delete c;
    return PSP_HIP_OK;
}

const char* psp_hip_last_error(const psp_hip_ctx* c)
{
    return c ? c->err.c_str() : "null context";
}

int psp_hip_set_stream(psp_hip_ctx* c, void* stream)
{
    if (!c) return PSP_HIP_ERR_INVALID;
    hipStream_t ns = stream ? (hipStream_t)stream : c->own;
    if (ns != c->stream) FreeGraphs(c);  // graphs are launched on the capture stream's device
    c->stream = ns;
    return PSP_HIP_OK;
}

int psp_hip_set_option(psp_hip_ctx* c, int option, int value)
{
    if (!c) return PSP_HIP_ERR_INVALID;
    if (option == PSP_HIP_OPT_SWEEP_GRAPH) {
        if (c->devReady) (void)hipStreamSynchronize(c->stream);
        FreeGraphs(c);
        c->useGraph = value != 0;
        return PSP_HIP_OK;
    }
    if (option == PSP_HIP_OPT_FACTOR_OVERLAP) {
        c->factorOverlap = value != 0;
        return PSP_HIP_OK;
    }
    if (option == PSP_HIP_OPT_SOLVE_XCD_REDUCE) {
        if (c->devReady) (void)hipStreamSynchronize(c->stream);
        FreeGraphs(c);  // captured launches keep their reduce launches or counters
        c->xcdReduce = value != 0;
        for (auto& D : c->dev) SetSliceReduce(c, D);
        return PSP_HIP_OK;
    }
    if (option == PSP_HIP_OPT_SOLVE_PERSIST_GRID) {
        if (value < 0) return Fail(c, PSP_HIP_ERR_INVALID, "persistent grid cap must be >= 0");
        if (c->devReady) (void)hipStreamSynchronize(c->stream);
        FreeGraphs(c);  // captured launches keep their grid sizes
        c->mvpGrid = value;
        for (auto& D : c->dev) D.mvpGrid = value;
        return PSP_HIP_OK;
    }
    if (option == PSP_HIP_OPT_SOLVE_FUSED || option == PSP_HIP_OPT_SOLVE_FUSED_QUEUES ||
        option == PSP_HIP_OPT_SOLVE_FUSED_ROWS || option == PSP_HIP_OPT_SOLVE_FUSED_CUT ||
        option == PSP_HIP_OPT_SOLVE_FUSED_WG) {
        if (option == PSP_HIP_OPT_SOLVE_FUSED) {
            c->plan.fused = value != 0;
        } else if (option == PSP_HIP_OPT_SOLVE_FUSED_QUEUES) {
            if (value < 1 || value > 64)
                return Fail(c, PSP_HIP_ERR_INVALID, "fused queues out of range [1, 64]");
            c->plan.fusedQueues = value;
        } else if (option == PSP_HIP_OPT_SOLVE_FUSED_ROWS) {
            if (value != 16 && value != 32 && value != 64)
                return Fail(c, PSP_HIP_ERR_INVALID, "fused tile rows must be 16, 32 or 64");
            c->plan.fusedRows = value;
        } else if (option == PSP_HIP_OPT_SOLVE_FUSED_WG) {
            if (value < 1 || value > 4096)
                return Fail(c, PSP_HIP_ERR_INVALID, "fused workgroups per queue out of range");
            c->plan.fusedWgPerQueue = value;
        } else {
            if (value < -1 || value > 64)
                return Fail(c, PSP_HIP_ERR_INVALID, "fused cut depth out of range [-1, 64]");
            c->plan.fusedCut = value;
        }
        return RebuildSolveTiles(c);
    }
    if (option == PSP_HIP_OPT_SOLVE_MIN_TILES || option == PSP_HIP_OPT_SOLVE_MAX_SPLIT ||
        option == PSP_HIP_OPT_SOLVE_WIDE_WG || option == PSP_HIP_OPT_SOLVE_PERSIST) {
        if (option == PSP_HIP_OPT_SOLVE_WIDE_WG) {
            c->plan.wideWg = value != 0;
        } else if (option == PSP_HIP_OPT_SOLVE_PERSIST) {
            c->plan.persist = value != 0;
        } else if (option == PSP_HIP_OPT_SOLVE_MIN_TILES) {
            if (value < 1 || value > (1 << 20))
                return Fail(c, PSP_HIP_ERR_INVALID, "solve min tiles out of range");
            c->plan.minTiles = value;
        } else {
            if (value < 1 || value > 64)
                return Fail(c, PSP_HIP_ERR_INVALID, "solve max split out of range [1, 64]");
            c->plan.maxSplit = value;
        }
        for (int cl = 0; cl < 3; ++cl)
            if (c->plan.present[cl])
                DefaultSolveShapes(c->plan.cls[cl], c->plan.minTiles, c->plan.maxSplit,
                                   c->plan.wideWg, c->plan.persist);
        return RebuildSolveTiles(c);
    }
    return Fail(c, PSP_HIP_ERR_INVALID, "unknown option");
}

int psp_hip_factor(psp_hip_ctx* c, const double* velocity, int scheme, psp_hip_factor_report* rep)
{
    if (!c) return PSP_HIP_ERR_INVALID;
    if (c->desc.omega == 0.0)
        return Fail(c, PSP_HIP_ERR_INVALID, "PML does not work at zero frequency");
    if (scheme != PSP_HIP_LDL_1D && scheme != PSP_HIP_LDL_SELINV_1D)
        return Fail(c, PSP_HIP_ERR_INVALID, "Invalid panelScheme value");
    if (!velocity) return Fail(c, PSP_HIP_ERR_INVALID, "Velocity grid is incorrect");
    const double t0 = Now();
    int st = EnsureDevice(c);
    if (st) return st;
    if ((st = EnsureFactorStorage(c))) return st;
    // a new factorization invalidates the captured sweeps' results only, not
    // their launches: the graphs keep reading the same factor arrays
    c->factored = false;
    const int64_t N = (int64_t)c->desc.nx * c->desc.ny * c->desc.nz;
    HIPCHK(c, hipMemcpyAsync(c->vel, velocity, sizeof(double) * N, hipMemcpyHostToDevice,
                             c->stream));
    // The Schur-complement (U) and inversion (T) scratch lives only while
    // factoring, so that many factored problems fit in HBM at once.
    const size_t eb = FrontBytes(c);
    for (int cl = 0; cl < 3; ++cl) {
        if (!c->plan.present[cl]) continue;
        ClassDev& D = c->dev[cl];
        if ((st = Alloc(c, &D.U, eb * (size_t)D.npanels * D.ustore, c->facAllocs)) ||
            (st = Alloc(c, &D.T, eb * (size_t)D.npanels * D.tstore, c->facAllocs))) {
            FreeFactorScratch(c);
            return st;
        }
    }
    double tf = 0, tl = 0, ti = 0;
    double tsc = 0;
    st = c->desc.precision == PSP_HIP_FRONTS_FP32 ? FactorClasses<float>(c, &tf, &tl, &ti, &tsc)
                                                  : FactorClasses<double>(c, &tf, &tl, &ti, &tsc);
    if (st) {
        (void)hipStreamSynchronize(c->stream);
        for (hipStream_t x : c->side)
            if (x) (void)hipStreamSynchronize(x);
        FreeFactorScratch(c);
        return st;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    FreeFactorScratch(c);
    c->factored = true;
    if (rep) {
        rep->seconds_total = Now() - t0;
        rep->seconds_fill = tf;
        rep->seconds_ldl = tl;
        rep->seconds_inverse = ti;
        rep->seconds_schur = tsc;
        double fl = 0, sc = 0, by = 0;
        int nf = 0;
        for (int cl = 0; cl < 3; ++cl) {
            if (!c->plan.present[cl]) continue;
            const double np = (double)c->classPanels[cl].size();
            fl += np * c->plan.cls[cl].flops;
            sc += np * c->plan.cls[cl].schurFlops;
            by += np * c->plan.cls[cl].factorEntries * FrontBytes(c);
            nf += (int)np * (int)c->plan.cls[cl].nodes.size();
        }
        rep->flops = fl;
        rep->schur_flops = sc;
        rep->bytes_factor = by;
        rep->num_panels = c->plan.geo.numPanels;
        rep->num_fronts = nf;
    }
    return PSP_HIP_OK;
}

int psp_hip_solve(psp_hip_ctx* c, void* B, int k, int m, double relTol, int maxIts, int flags,
                  psp_hip_solve_report* rep)
{
    return RunSolve(c, B, k, m, relTol, maxIts, flags, 0, nullptr, rep);
}

int psp_hip_solve_batch(psp_hip_ctx* const* ctxs, void* const* Bs, const int* ks, int n,
                        int m, double relTol, int maxIts, int flags,
                        psp_hip_solve_report* reports)
{
    if (n < 0 || (n > 0 && (!ctxs || !Bs || !ks))) return PSP_HIP_ERR_INVALID;
    for (int i = 0; i < n; ++i) {
        if (!ctxs[i]) return PSP_HIP_ERR_INVALID;
        for (int j = 0; j < i; ++j)
            if (ctxs[j] == ctxs[i])
                return Fail(ctxs[i], PSP_HIP_ERR_INVALID, "solve_batch: contexts must be distinct");
    }
    if (n == 0) return PSP_HIP_OK;
    if (!(flags & PSP_HIP_BATCH_THREADS)) {
        // one merged batch per front precision, in turn
        int first = PSP_HIP_OK;
        for (int prec : {PSP_HIP_FRONTS_FP32, PSP_HIP_FRONTS_FP64}) {
            std::vector<psp_hip_ctx*> gc;
            std::vector<void*> gb;
            std::vector<int> gk, idx;
            for (int i = 0; i < n; ++i)
                if (ctxs[i]->desc.precision == prec) {
                    gc.push_back(ctxs[i]);
                    gb.push_back(Bs[i]);
                    gk.push_back(ks[i]);
                    idx.push_back(i);
                }
            if (gc.empty()) continue;
            std::vector<psp_hip_solve_report> gr(gc.size());
            const int rc = SolveBatchMerged(gc.data(), gb.data(), gk.data(), (int)gc.size(), m,
                                            relTol, maxIts, flags, gr.data());
            if (reports)
                for (size_t j = 0; j < idx.size(); ++j) reports[idx[j]] = gr[j];
            if (rc != PSP_HIP_OK && first == PSP_HIP_OK) first = rc;
        }
        return first;
    }
    std::vector<int> st(n, PSP_HIP_OK);
    std::vector<std::thread> workers;
    workers.reserve(n);
    for (int i = 0; i < n; ++i)
        workers.emplace_back([&, i]() {
            st[i] = RunSolve(ctxs[i], Bs[i], ks[i], m, relTol, maxIts, flags, 0, nullptr,
                             reports ? &reports[i] : nullptr);
        });
    for (std::thread& t : workers) t.join();
    for (int i = 0; i < n; ++i)
        if (st[i] != PSP_HIP_OK) return st[i];
    return PSP_HIP_OK;
}

int psp_hip_solve_fixed(psp_hip_ctx* c, void* B, int k, int m, int steps, int flags,
                        double* relHist)
{
    if (steps < 1) return c ? Fail(c, PSP_HIP_ERR_INVALID, "steps must be >= 1") : 1;
    return RunSolve(c, B, k, m, 0.0, 0, flags, steps, relHist, nullptr);
}

int psp_hip_multiply(psp_hip_ctx* c, const void* x, void* y, int k)
{
    int rc = CheckSolveArgs(c, x, k, 1);
    if (rc) return rc;
    LaunchSpmv(Stencil(c), c->vel, (const c128*)x, nullptr, (c128*)y, k, 0, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSP_HIP_OK;
}

int psp_hip_precondition(psp_hip_ctx* c, void* B, int k)
{
    int rc = CheckSolveArgs(c, B, k, 1);
    if (rc) return rc;
    if ((rc = EnsureWorkspace(c, k, std::max(1, c->wsK == k ? c->wsM : 1)))) return rc;
    if ((rc = Precondition(c, (c128*)B, k))) return rc;
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSP_HIP_OK;
}

int psp_hip_precondition_noalloc(psp_hip_ctx* c, void* B, int k)
{
    int rc = CheckSolveArgs(c, B, k, 1);
    if (rc) return rc;
    if (!c->scratch) return Fail(c, PSP_HIP_ERR_STATE, "device scratch not set up");
    const SweepSizes z = SweepBufferSizes(c, k);
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t need = up(z.xp) + up(z.w) + up(z.p) + up(z.g[0]) + up(z.g[1]) + up(z.g[2]) +
                        up(z.s[0]) + up(z.s[1]) + up(z.s[2]);
    if (need > kScratchBytes)
        return Fail(c, PSP_HIP_ERR_INVALID,
                    "precondition_noalloc: the sweep needs " + std::to_string(need) +
                        " bytes of scratch for this grid and k; the maximum is " +
                        std::to_string(kScratchBytes) + " (use psp_hip_precondition)");
    char* base = (char*)c->scratch;
    // the slot copies and the vectors' padding rows must read as zero; a
    // fill kernel clears the whole region (no memset node if captured)
    LaunchFillZero(base, (int64_t)need, c->stream);
    void* keep[6] = {c->Xp, c->W, c->P, c->G[0], c->G[1], c->G[2]};
    int32_t* keepS[3] = {c->S[0], c->S[1], c->S[2]};
    size_t off = 0;
    auto take = [&](size_t b) {
        void* q = b ? base + off : nullptr;
        off += up(b);
        return q;
    };
    c->Xp = take(z.xp);
    c->W = take(z.w);
    c->P = take(z.p);
    for (int cl = 0; cl < 3; ++cl) c->G[cl] = take(z.g[cl]);
    for (int cl = 0; cl < 3; ++cl) c->S[cl] = (int32_t*)take(z.s[cl]);
    SweepEager(c, (c128*)B, k);
    c->Xp = keep[0];
    c->W = keep[1];
    c->P = keep[2];
    for (int cl = 0; cl < 3; ++cl) c->G[cl] = keep[3 + cl];
    for (int cl = 0; cl < 3; ++cl) c->S[cl] = keepS[cl];
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSP_HIP_OK;
}

int psp_hip_autotune(psp_hip_ctx* c, int k, double* before, double* after)
{
    if (!c) return PSP_HIP_ERR_INVALID;
    if (!c->factored) return Fail(c, PSP_HIP_ERR_STATE, "autotune needs Initialize first");
    if (k < 1 || k > 256) return Fail(c, PSP_HIP_ERR_INVALID, "k out of range [1, 256]");
    return c->desc.precision == PSP_HIP_FRONTS_FP32 ? Autotune<float>(c, k, before, after)
                                                    : Autotune<double>(c, k, before, after);
}

int psp_hip_profile_sweep(psp_hip_ctx* c, void* B, int k, psp_hip_kernel_stats* stats)
{
    int rc = CheckSolveArgs(c, B, k, 1);
    if (rc) return rc;
    if (!stats) return Fail(c, PSP_HIP_ERR_INVALID, "null stats");
    if ((rc = EnsureWorkspace(c, k, std::max(1, c->wsK == k ? c->wsM : 1)))) return rc;
    MvProfile prof;
    HIPCHK(c, hipEventCreate(&prof.a));
    HIPCHK(c, hipEventCreate(&prof.b));
    if (c->desc.precision == PSP_HIP_FRONTS_FP32)
        Sweep<float>(c, (c128*)B, k, &prof);
    else
        Sweep<double>(c, (c128*)B, k, &prof);
    const hipError_t e = hipStreamSynchronize(c->stream);
    (void)hipEventDestroy(prof.a);
    (void)hipEventDestroy(prof.b);
    HIPCHK(c, e);
    HIPCHK(c, hipGetLastError());
    stats->launches = prof.launches;
    stats->seconds = prof.seconds;
    stats->bytes = prof.bytes;
    stats->flops = prof.flops;
    return PSP_HIP_OK;
}

int psp_hip_solve_panel(psp_hip_ctx* c, void* B, int k, int panel)
{
    int rc = CheckSolveArgs(c, B, k, 1);
    if (rc) return rc;
    if (panel < 0 || panel >= c->plan.geo.numPanels)
        return Fail(c, PSP_HIP_ERR_INVALID, "panel index out of range");
    if ((rc = EnsureWorkspace(c, k, std::max(1, c->wsK == k ? c->wsM : 1)))) return rc;
    if (c->desc.precision == PSP_HIP_FRONTS_FP32)
        PanelForward<float>(c, (c128*)B, k, panel);
    else
        PanelForward<double>(c, (c128*)B, k, panel);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSP_HIP_OK;
}

int psp_hip_adjoint_gradient(psp_hip_ctx* c, const void* u, const void* lambda, int k,
                             double* grad, int accumulate)
{
    int rc = CheckSolveArgs(c, u, k, 1);
    if (rc) return rc;
    if (!lambda || !grad) return Fail(c, PSP_HIP_ERR_INVALID, "null adjoint buffers");
    LaunchAdjointGrad(Stencil(c), (const c128*)u, (const c128*)lambda, k, grad, accumulate,
                      c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PSP_HIP_OK;
}

int psp_hip_num_panels(const psp_hip_ctx* c) { return c ? c->plan.geo.numPanels : -1; }

int psp_hip_panel_info(const psp_hip_ctx* c, int p, int* vPanel, int* depth, int* padding,
                       int* cls)
{
    if (!c || p < 0 || p >= c->plan.geo.numPanels) return PSP_HIP_ERR_INVALID;
    const Geometry& g = c->plan.geo;
    if (vPanel) *vPanel = g.panelV(p);
    if (depth) *depth = g.depth(p);
    if (padding) *padding = g.padding(p);
    if (cls) *cls = g.cls(p);
    return PSP_HIP_OK;
}

int psp_hip_class_num_nodes(const psp_hip_ctx* c, int cls)
{
    if (!c || cls < 0 || cls > 2 || !c->plan.present[cls]) return 0;
    return (int)c->plan.cls[cls].nodes.size();
}

int psp_hip_class_nodes(const psp_hip_ctx* c, int cls, int* off, int* size, int* parent,
                        int* usize, int* structs)
{
    if (!c || cls < 0 || cls > 2 || !c->plan.present[cls]) return PSP_HIP_ERR_INVALID;
    const ClassPlan& P = c->plan.cls[cls];
    for (size_t t = 0; t < P.nodes.size(); ++t) {
        if (off) off[t] = P.nodes[t].off;
        if (size) size[t] = P.nodes[t].size;
        if (parent) parent[t] = P.nodes[t].parent;
        if (usize) usize[t] = P.nodes[t].usize;
    }
    if (structs) std::memcpy(structs, P.structs.data(), sizeof(int) * P.structs.size());
    return PSP_HIP_OK;
}

int psp_hip_reordered_index(const psp_hip_ctx* c, int x, int y, int vLocal, int vSize)
{
    if (!c) return -1;
    return NestedIndex(x, y, vLocal, c->plan.geo.nx, c->plan.geo.ny, vSize, c->plan.geo.cutoff);
}

int psp_hip_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

}  // extern "C"
