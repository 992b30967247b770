// HIP kernels for gfx950 (MI355X): front assembly, multifrontal LDL^T,
// front solves, sweep couplings, the PML 7-point operator and the GMRES
// vector primitives.  All kernels use 256-thread workgroups (4 wave64).
#include "kernels.hpp"

#include <algorithm>
#include <type_traits>

namespace psp_amd {

namespace {

constexpr int kThreads = 256;

// --------------------------------------------------------------------------
// Complex arithmetic (interleaved re, im).
// --------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ cx<T> mk(T r, T i) { return cx<T>{r, i}; }
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.r + b.r, a.i + b.i}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.r - b.r, a.i - b.i}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a) { return {-a.r, -a.i}; }
template <typename T>
__device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b)
{
    return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
template <typename T>
__device__ __forceinline__ cx<T> operator*(cx<T> a, T s) { return {a.r * s, a.i * s}; }
// acc + a*b
template <typename T>
__device__ __forceinline__ cx<T> cfma(cx<T> a, cx<T> b, cx<T> acc)
{
    acc.r = fma(a.r, b.r, acc.r);
    acc.r = fma(-a.i, b.i, acc.r);
    acc.i = fma(a.r, b.i, acc.i);
    acc.i = fma(a.i, b.r, acc.i);
    return acc;
}
// acc - a*b
template <typename T>
__device__ __forceinline__ cx<T> cfms(cx<T> a, cx<T> b, cx<T> acc)
{
    acc.r = fma(-a.r, b.r, acc.r);
    acc.r = fma(a.i, b.i, acc.r);
    acc.i = fma(-a.r, b.i, acc.i);
    acc.i = fma(-a.i, b.r, acc.i);
    return acc;
}
// Smith's complex division a / b.
template <typename T>
__device__ __forceinline__ cx<T> cdiv(cx<T> a, cx<T> b)
{
    if (fabs(b.r) >= fabs(b.i)) {
        const T rr = b.i / b.r, den = b.r + b.i * rr;
        return {(a.r + a.i * rr) / den, (a.i - a.r * rr) / den};
    }
    const T rr = b.r / b.i, den = b.i + b.r * rr;
    return {(a.r * rr + a.i) / den, (a.i * rr - a.r) / den};
}
template <typename T>
__device__ __forceinline__ cx<T> cinv(cx<T> b) { return cdiv(cx<T>{T(1), T(0)}, b); }
template <typename T>
__device__ __forceinline__ cx<T> cdivr(cx<T> a, T s) { return {a.r / s, a.i / s}; }

template <typename To, typename Ti>
__device__ __forceinline__ cx<To> cvt(cx<Ti> a) { return {To(a.r), To(a.i)}; }

// Two reals of one type (a pre-swizzled LDS operand pair).
template <typename T>
struct T2 {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> czero() { return {T(0), T(0)}; }

// --------------------------------------------------------------------------
// PML 7-point stencil: FormGlobalRow (initialize_finalize.hpp:527-587) and
// FormLowerColumnOfNode (:589-640).  out[0] centre, out[1..6] the values of
// the x-1, x+1, y-1, y+1, v-1, v+1 couplings.
// --------------------------------------------------------------------------
__device__ __forceinline__ void Stencil(c128 s1L, c128 s1M, c128 s1R, c128 s2L, c128 s2M,
                                        c128 s2R, c128 s3L, c128 s3M, c128 s3R, c128 k2,
                                        double hx, double hy, double hz, c128 out[7])
{
    const c128 xTop = s2M * s3M;
    const c128 xTempL = cdiv(xTop, s1L), xTempM = cdiv(xTop, s1M), xTempR = cdiv(xTop, s1R);
    const c128 xTermL = cdivr(xTempL + xTempM, 2 * hx * hx);
    const c128 xTermR = cdivr(xTempR + xTempM, 2 * hx * hx);
    const c128 yTop = s1M * s3M;
    const c128 yTempL = cdiv(yTop, s2L), yTempM = cdiv(yTop, s2M), yTempR = cdiv(yTop, s2R);
    const c128 yTermL = cdivr(yTempL + yTempM, 2 * hy * hy);
    const c128 yTermR = cdivr(yTempR + yTempM, 2 * hy * hy);
    const c128 zTop = s1M * s2M;
    const c128 vTempL = cdiv(zTop, s3L), vTempM = cdiv(zTop, s3M), vTempR = cdiv(zTop, s3R);
    const c128 vTermL = cdivr(vTempL + vTempM, 2 * hz * hz);
    const c128 vTermR = cdivr(vTempR + vTempM, 2 * hz * hz);
    out[0] = (xTermL + xTermR + yTermL + yTermR + vTermL + vTermR) - k2 * s1M * s2M * s3M;
    out[1] = -xTermL;
    out[2] = -xTermR;
    out[3] = -yTermL;
    out[4] = -yTermR;
    out[5] = -vTermL;
    out[6] = -vTermR;
}

// Global operator row at (x, y, v): undamped frequency, physical PML.
__device__ __forceinline__ void GlobalRow(const StencilArgs& st, double alpha, int x, int y,
                                          int v, c128 out[7])
{
    const double wa = st.omega / alpha;
    const c128 k2 = {wa * wa, 0.0};
    Stencil(st.s1[x], st.s1[x + 1], st.s1[x + 2], st.s2[y], st.s2[y + 1], st.s2[y + 2],
            st.s3[v], st.s3[v + 1], st.s3[v + 2], k2, st.hx, st.hy, st.hz, out);
}

// The same row from per-axis tables: xTermL = (s2 s3) (1/s1(x-1) + 1/s1(x))
// / (2 hx^2) etc., i.e. FormGlobalRow's terms with the divisions hoisted into
// host tables (equal up to rounding).
__device__ __forceinline__ void GlobalRowFast(const StencilArgs& st, double alpha, int x, int y,
                                              int v, c128 out[7])
{
    const c128 s1M = st.s1[x + 1], s2M = st.s2[y + 1], s3M = st.s3[v + 1];
    const c128 xTop = s2M * s3M, yTop = s1M * s3M, zTop = s1M * s2M;
    const c128 xL = xTop * st.ax[2 * x], xR = xTop * st.ax[2 * x + 1];
    const c128 yL = yTop * st.ay[2 * y], yR = yTop * st.ay[2 * y + 1];
    const c128 vL = zTop * st.av[2 * v], vR = zTop * st.av[2 * v + 1];
    const double wa = st.omega / alpha;
    out[0] = (xL + xR + yL + yR + vL + vR) - (zTop * s3M) * (wa * wa);
    out[1] = -xL;
    out[2] = -xR;
    out[3] = -yL;
    out[4] = -yR;
    out[5] = -vL;
    out[6] = -vR;
}

// --------------------------------------------------------------------------
// Fill
// --------------------------------------------------------------------------
__global__ void k_fill_zero(uint4* p, int64_t n16)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        p[i] = make_uint4(0, 0, 0, 0);
}

// rows ranges of n8 8-byte words, range r at p + r * stride8
__global__ void k_fill_zero_rows(uint2* p, int64_t n8, int64_t stride8)
{
    uint2* q = p + (int64_t)blockIdx.y * stride8;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n8; i += stride)
        q[i] = make_uint2(0, 0);
}

__global__ void k_fill_zero_bytes(unsigned char* p, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

// --------------------------------------------------------------------------
// Front assembly of the original panel operator (FillPanelFronts,
// initialize_finalize.hpp:1284-1337): one thread per padded unknown of one
// panel of the batch; writes the diagonal and the strictly-lower couplings
// of its column into its front.
// --------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void k_scatter_orig(
    int n, int vsize, const int32_t* qxy, const int32_t* qv, const int32_t* qnode,
    const int32_t* qcol, const int32_t* qrows, const int64_t* nodeL, const int32_t* nodeS,
    const int32_t* nodeU, const int32_t* panelVOff, const c128* s3art, StencilArgs st,
    const double* vel, double damping, cx<T>* L, int64_t fstore)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int b = blockIdx.y;
    const int x = qxy[q] & 0xffff, y = qxy[q] >> 16, vl = qv[q];
    const int vOffset = panelVOff[b];
    const int v = vOffset + vl;
    const int z = st.nz - 1 - v;
    const double alpha = vel[x + (int64_t)y * st.nx + (int64_t)z * st.nx * st.ny];
    const c128* s3 = s3art + (int64_t)b * (vsize + 2);
    const c128 so = {st.omega / alpha, damping / alpha};  // (omega + i damping)/alpha
    const c128 k2 = so * so;
    c128 e[7];
    Stencil(st.s1[x], st.s1[x + 1], st.s1[x + 2], st.s2[y], st.s2[y + 1], st.s2[y + 2],
            s3[vl], s3[vl + 1], s3[vl + 2], k2, st.hx, st.hy, st.hz, e);
    const int t = qnode[q], j = qcol[q];
    const int64_t fs = nodeS[t] + nodeU[t];
    cx<T>* col = L + (int64_t)b * fstore + nodeL[t] + (int64_t)j * fs;
    col[j] = cvt<T>(e[0]);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const int row = qrows[(int64_t)q * 6 + k];
        if (row >= 0) col[row] = cvt<T>(e[k + 1]);
    }
}

// --------------------------------------------------------------------------
// Extend-add of a child's Schur complement into its parent front.
// --------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kThreads) void k_extend_add(const ExtAddItem* items, int start,
                                                          const int32_t* relinds, cx<T>* L,
                                                          int64_t fstore, cx<T>* U,
                                                          int64_t ustore)
{
    const ExtAddItem it = items[start + blockIdx.x];
    const int b = blockIdx.y;
    const cx<T>* Uc = U + (int64_t)b * ustore + it.uChild;
    cx<T>* Lp = L + (int64_t)b * fstore + it.lParent;
    cx<T>* Up = U + (int64_t)b * ustore + it.uParent;
    const int32_t* rel = relinds + it.relStart;
    const int ldp = it.sp + it.up;
    for (int cb = it.col0; cb < it.col1; ++cb) {
        const int rb = rel[cb];
        for (int a = cb + threadIdx.x; a < it.uc; a += blockDim.x) {
            const cx<T> val = Uc[a + (int64_t)cb * it.uc];
            const int ra = rel[a];
            if (rb < it.sp) {
                cx<T>* p = Lp + ra + (int64_t)rb * ldp;
                *p = *p + val;
            } else {
                cx<T>* p = Up + (ra - it.sp) + (int64_t)(rb - it.sp) * it.up;
                *p = *p + val;
            }
        }
    }
}

// --------------------------------------------------------------------------
// Blocked LDL^T, step 1: unpivoted LDL^T of the diagonal block in LDS.
// --------------------------------------------------------------------------
template <typename T, int NB>
__global__ __launch_bounds__(kThreads) void k_ldl_diag(const LdlItem* items, int start, cx<T>* L,
                                                        int64_t fstore)
{
    const LdlItem it = items[start + blockIdx.x];
    cx<T>* F = L + (int64_t)blockIdx.y * fstore + it.l;
    const int ld = it.ld, j0 = it.j0, m = it.j1 - it.j0;
    static_assert(NB == 32 && kThreads == 256, "each thread owns 2 x 2 elements of the block");
    __shared__ cx<T> A[NB][NB + 1];  // A[col][row]
    // thread (tx, ty) owns rows tx, tx + 16 and columns ty, ty + 16
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int r = tx + 16 * a, c = ty + 16 * b;
            if (r < m && c <= r) A[c][r] = F[(j0 + r) + (int64_t)(j0 + c) * ld];
        }
    __syncthreads();
    // Step j updates the trailing block with column j (unscaled) and scales
    // column j - 1, which step j no longer reads: one barrier per column.
    for (int j = 0; j < m; ++j) {
        const cx<T> dinv = cinv(A[j][j]);
        if (j > 0 && ty == 0) {
            const cx<T> dp = cinv(A[j - 1][j - 1]);
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const int r = tx + 16 * a;
                if (r > j - 1 && r < m) A[j - 1][r] = A[j - 1][r] * dp;
            }
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int r = tx + 16 * a, c = ty + 16 * b;
                if (c > j && r >= c && r < m) A[c][r] = cfms(A[j][r] * dinv, A[j][c], A[c][r]);
            }
        __syncthreads();  // no global loads are in flight here
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int r = tx + 16 * a, c = ty + 16 * b;
            if (r < m && c <= r) F[(j0 + r) + (int64_t)(j0 + c) * ld] = A[c][r];
        }
}

// Blocked LDL^T, step 2: rows below the diagonal block,
// L21 = F21 L11^{-T} D1^{-1}, one thread per row.
template <typename T, int NB>
__global__ __launch_bounds__(kThreads) void k_ldl_trsm(const LdlItem* items, int start, cx<T>* L,
                                                        int64_t fstore)
{
    const LdlItem it = items[start + blockIdx.x];
    cx<T>* F = L + (int64_t)blockIdx.y * fstore + it.l;
    const int ld = it.ld, j0 = it.j0, m = it.j1 - it.j0;
    __shared__ cx<T> W[NB][NB];  // W[c][q] = d_q L11(c, q), q < c
    __shared__ cx<T> dinv[NB];
    for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
        const int c = e / m, q = e % m;
        if (q < c)
            W[c][q] = F[(j0 + q) + (int64_t)(j0 + q) * ld] * F[(j0 + c) + (int64_t)(j0 + q) * ld];
    }
    for (int c = threadIdx.x; c < m; c += blockDim.x)
        dinv[c] = cinv(F[(j0 + c) + (int64_t)(j0 + c) * ld]);
    __syncthreads();
    const int r = it.r0 + threadIdx.x;
    if (r >= it.r1) return;
    cx<T> x[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) x[c] = c < m ? F[r + (int64_t)(j0 + c) * ld] : czero<T>();
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        if (c < m) {
            cx<T> acc = x[c];
#pragma unroll
            for (int q = 0; q < c; ++q) acc = cfms(x[q], W[c][q], acc);
            x[c] = acc * dinv[c];
        }
    }
#pragma unroll
    for (int c = 0; c < NB; ++c)
        if (c < m) F[r + (int64_t)(j0 + c) * ld] = x[c];
}

// Inverse of one unit-lower diagonal block in place (strict lower part).
template <typename T, int NB>
__global__ __launch_bounds__(kThreads) void k_inv_diag(const InvItem* items, cx<T>* L,
                                                        int64_t fstore)
{
    const InvItem it = items[blockIdx.x];
    cx<T>* F = L + (int64_t)blockIdx.y * fstore + it.l;
    const int ld = it.ld, j0 = it.j0, m = it.j1 - it.j0;
    __shared__ cx<T> A[NB][NB + 1];  // A[col][row] strictly lower
    __shared__ cx<T> X[NB][NB + 1];  // X[col][row] inverse
    for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
        const int r = e % m, c = e / m;
        A[c][r] = r > c ? F[(j0 + r) + (int64_t)(j0 + c) * ld] : czero<T>();
    }
    __syncthreads();
    const int c = threadIdx.x;
    if (c < m) {
        X[c][c] = cx<T>{T(1), T(0)};
        for (int r = c + 1; r < m; ++r) {
            cx<T> acc = czero<T>();
            for (int q = c; q < r; ++q) acc = cfms(A[q][r], X[c][q], acc);
            X[c][r] = acc;
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
        const int r = e % m, cc = e / m;
        if (r > cc) F[(j0 + r) + (int64_t)(j0 + cc) * ld] = X[cc][r];
    }
}

// Workgroup barrier for LDS traffic only: waits for this wave's LDS
// operations but not for its outstanding global loads, so loads issued
// before the barrier (the next operands) stay in flight across it.  Used
// wherever a barrier only orders LDS accesses.
__device__ __forceinline__ void LdsBarrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// --------------------------------------------------------------------------
// Front solves on the padded panel vector Xp (rows x k, interleaved), with
// the factors in the solve layout (rows < s of L: Linv with a stored unit
// diagonal; rows >= s: -L_BL Linv; LT: [Linv^T D^-1 | (-L_BL Linv)^T]).
// Per node, with g = [Xp_own; 0] + the children's update vectors (read from
// per-slot copies of the node's front rows that the children write):
//   forward  F: W_top = Linv g_top; the update g_bot + (-L_BL Linv) g_top
//               goes into the parent's slot copy;
//   backward B: Xp_own = LT [W_top; x_struct].
// Each is one product over the node's factor block, split into row tiles
// (and, near the root, inner slices summed by a reduce launch).
// --------------------------------------------------------------------------
enum { PH_F = 0, PH_B = 1 };

// The children's update rows: each child writes its rows straight into a
// per-slot copy of its parent's front rows (slot 0: first child, slot 1:
// second child), at the parent rows given by relinds.  Rows a child does not
// touch stay zero from the workspace's allocation (each slot copy belongs to
// one panel class, and a class writes the same rows in every solve), so a
// parent reads its gathered input with plain loads.
template <typename T>
struct Slots {
    cx<T>* G;                // slot 0; slot 1 at G + half
    int64_t half;            // work rows of the class * k
    const int32_t* relinds;  // row in the parent's front of each update row
};

// Raw loads of a staged operand element, combined after they have arrived.
template <typename T>
struct Raw2 {
    cx<T> a, b;
};
template <typename T>
struct Raw3 {
    cx<T> a, b, c;
};
template <typename T>
struct RawRow {  // gathered value of a bottom row (two slot copies) and its parent row
    cx<T> a, b;
    int32_t prow;
};

// The value v, passed through an empty asm: the compiler must hold it in a
// register at this point, so a load that produced it is not sunk into the
// arm of a later select (where its wait would sit inside a branch, one
// round trip per element).
template <class X>
__device__ __forceinline__ X Opaque(X v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// An empty asm that reads v: every load that may still write v is waited for
// here, so its register is free of pending writes on every path that
// follows (without it, a register whose load result is used on one branch
// only can be reused on the other while the load is still in flight, and
// the reuse then waits for every outstanding load).
template <class X>
__device__ __forceinline__ void Sink(const X& v)
{
    asm volatile("" ::"v"(v));
}

// Branch-free staging of inner operands, in two steps so that a thread's
// index loads (B: the struct rows of the update columns) are all in flight
// together before the value loads that depend on them.  Every load is issued
// unconditionally from a valid address (F: col < s, the forward inner range;
// B: the struct list carries a sentinel entry, so structs[ustart] exists at a
// root).
template <int PH>
__device__ __forceinline__ int32_t StageIndex(const NodeDev& nd, const int32_t* structs, int col)
{
    return PH == PH_F ? 0 : structs[nd.ustart + max(col - nd.s, 0)];
}

template <typename T, int PH>
__device__ __forceinline__ cx<T> StageValue(const NodeDev& nd, const Slots<T>& gs, const cx<T>* Xp,
                                            const cx<T>* W, int col, int rr, int k, int32_t srow)
{
    if (PH == PH_F) {
        const int64_t g = (nd.w + col) * k + rr;
        return gs.G[g] + gs.G[gs.half + g] + Xp[(int64_t)(nd.off + col) * k + rr];
    }
    const cx<T>* p = col < nd.s ? W + (nd.w + col) * k + rr : Xp + (int64_t)srow * k + rr;
    return *p;
}

template <typename T, int PH>
__device__ __forceinline__ cx<T> StageOperand(const NodeDev& nd, const int32_t* structs,
                                              const Slots<T>& gs, const cx<T>* Xp,
                                              const cx<T>* W, int col, int rr, int k)
{
    return StageValue<T, PH>(nd, gs, Xp, W, col, rr, k, Opaque(StageIndex<PH>(nd, structs, col)));
}

// N staged elements of a thread; at(q) gives element q's (inner index,
// right-hand side) as an int2.
template <typename T, int PH, int N, class AT>
__device__ __forceinline__ void StageBatch(const NodeDev& nd, const int32_t* structs,
                                           const Slots<T>& gs, const cx<T>* Xp, const cx<T>* W,
                                           int k, const AT& at, cx<T> (&sv)[N])
{
    int32_t sx[N];
#pragma unroll
    for (int q = 0; q < N; ++q) sx[q] = StageIndex<PH>(nd, structs, at(q).x);
    if (PH == PH_B) {
#pragma unroll
        for (int q = 0; q < N; ++q) sx[q] = Opaque(sx[q]);
    }
#pragma unroll
    for (int q = 0; q < N; ++q) {
        const int2 p = at(q);
        sv[q] = StageValue<T, PH>(nd, gs, Xp, W, p.x, p.y, k, sx[q]);
    }
}

// Factor entry of phase PH at (row, col) of the node's block.
template <typename T, int PH>
__device__ __forceinline__ cx<T> FactorEntry(const cx<T>* Lf, const cx<T>* Lt, int s, int fs,
                                             int row, int col)
{
    if (PH == PH_F) return col <= row ? Lf[row + (int64_t)col * fs] : czero<T>();
    return col >= row ? Lt[row + (int64_t)col * s] : czero<T>();
}

// Finish of output row `row`, right-hand side r, product a.  F: top rows
// are y (into W); a bottom row is g + a, with g the gathered value of that
// row, written into the parent's slot copy at parent row `prow`.  B: the
// node's Xp rows.
template <typename T, int PH>
__device__ __forceinline__ void Finish(const NodeDev& nd, const Slots<T>& gs, cx<T>* Xp, cx<T>* W,
                                       int row, int r, int k, cx<T> a, cx<T> g, int prow)
{
    if (PH == PH_F) {
        if (row < nd.s)
            W[(nd.w + row) * k + r] = a;
        else if (nd.cslot >= 0)
            gs.G[nd.cslot * gs.half + (int64_t)(nd.pw + prow) * k + r] = g + a;
    } else {
        Xp[(int64_t)(nd.off + row) * k + r] = a;
    }
}

template <typename T, int PH>
__device__ __forceinline__ void FinishGather(const NodeDev& nd, const Slots<T>& gs, cx<T>* Xp,
                                             cx<T>* W, int row, int r, int k, cx<T> a)
{
    if (PH == PH_F && row >= nd.s) {
        const int64_t g = (nd.w + row) * k + r;
        Finish<T, PH>(nd, gs, Xp, W, row, r, k, a, gs.G[g] + gs.G[gs.half + g],
                      gs.relinds[nd.ustart + row - nd.s]);
    } else {
        Finish<T, PH>(nd, gs, Xp, W, row, r, k, a, czero<T>(), 0);
    }
}

constexpr int kMvThreads = 512;  // 8 waves split each row block's inner indices

// Vector-unit path (k < 8).  One workgroup per tile (a block of RT output
// rows and an inner range).  Lane l of a wave owns row l % RT and column
// slot l / RT (64 / RT columns per wave step); the inner range is cut into
// chunks of NI columns per lane.  For each chunk the lane first issues all
// NI of its factor loads, then the workgroup stages the matching vector
// rows in LDS, so the factor and vector reads of a chunk are in flight
// together.  Column slots are combined by lane shuffles and the 8 waves'
// partial sums by a fixed-order tree in LDS: every output element is summed
// in a fixed order.
template <typename T, int PH, int RT>
__global__ __launch_bounds__(kMvThreads) void k_mv(const SolveTile* tiles, int start,
                                                    const int32_t* structs,
                                                    Slots<T> gs, const cx<T>* L,
                                                    const cx<T>* LT, cx<T>* Xp, cx<T>* W,
                                                    cx<T>* P, int k)
{
    constexpr int KC = 8, NW = kMvThreads / 64, CPW = 64 / RT, STRIDE = CPW * NW;
    constexpr int NI = RT == 16 ? 16 : 32, CC = STRIDE * NI;
    const SolveTile tl = tiles[start + blockIdx.x];
    if (tl.nrows <= 0) return;  // a hole of the XCD placement
    const NodeDev nd = tl.nd;
    const int s = nd.s, fs = nd.s + nd.u;
    const cx<T>* Lf = L + nd.l;
    const cx<T>* Lt = LT + nd.l;
    __shared__ cx<T> vs[CC][KC];            // staged vector chunk
    __shared__ cx<T> red[NW / 2][RT][KC];   // wave partial sums
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rl = lane % RT, slot = lane / RT;
    const int row = tl.row0 + rl;
    const bool rowOk = rl < tl.nrows;
    const int first = slot + CPW * wave;  // this lane's first column in a chunk
    for (int r0 = 0; r0 < k; r0 += KC) {
        const int kc = min(KC, k - r0);
        cx<T> acc[KC];
#pragma unroll
        for (int r = 0; r < KC; ++r) acc[r] = czero<T>();
        for (int cb = tl.col0; cb < tl.col1; cb += CC) {
            const int ncol = min(CC, tl.col1 - cb);
            // staging loads first, then the factor loads, all in flight
            // together; the LDS stores wait only for the staging loads
            constexpr int SPT = CC * KC / kMvThreads;
            cx<T> sv[SPT];
            StageBatch<T, PH>(nd, structs, gs, Xp, W, k, [&](int q) {
                const int e = tid + kMvThreads * q, c = e / KC, r = e % KC;
                return int2{cb + min(c, ncol - 1), min(r0 + r, k - 1)};
            }, sv);
            cx<T> m[NI];
#pragma unroll
            for (int u = 0; u < NI; ++u) {
                const int c = first + STRIDE * u;
                m[u] = (rowOk && c < ncol) ? FactorEntry<T, PH>(Lf, Lt, s, fs, row, cb + c)
                                           : czero<T>();
            }
            LdsBarrier();  // the previous chunk's readers are done with vs
#pragma unroll
            for (int q = 0; q < SPT; ++q) {
                const int e = tid + kMvThreads * q, c = e / KC, r = e % KC;
                if (c < ncol) vs[c][r] = r < kc ? sv[q] : czero<T>();
            }
            LdsBarrier();
#pragma unroll
            for (int u = 0; u < NI; ++u) {
                const int c = first + STRIDE * u;
                if (c < ncol) {
#pragma unroll
                    for (int r = 0; r < KC; ++r) acc[r] = cfma(m[u], vs[c][r], acc[r]);
                }
            }
        }
        // combine the column slots of a wave (lanes l, l + RT, ...)
#pragma unroll
        for (int o = RT; o < 64; o <<= 1) {
#pragma unroll
            for (int r = 0; r < KC; ++r) {
                acc[r].r += __shfl_xor(acc[r].r, o, 64);
                acc[r].i += __shfl_xor(acc[r].i, o, 64);
            }
        }
        for (int h = NW / 2; h >= 1; h >>= 1) {
            if (wave >= h && wave < 2 * h && slot == 0) {
#pragma unroll
                for (int r = 0; r < KC; ++r) red[wave - h][rl][r] = acc[r];
            }
            LdsBarrier();
            if (wave < h && slot == 0) {
#pragma unroll
                for (int r = 0; r < KC; ++r) acc[r] = acc[r] + red[wave][rl][r];
            }
            LdsBarrier();
        }
        if (wave == 0 && slot == 0 && rowOk) {
#pragma unroll
            for (int r = 0; r < KC; ++r)
                if (r < kc) {
                    if (tl.slot >= 0)
                        P[((int64_t)tl.slot * 64 + rl) * k + r0 + r] = acc[r];
                    else
                        FinishGather<T, PH>(nd, gs, Xp, W, row, r0 + r, k, acc[r]);
                }
        }
    }
}

// --------------------------------------------------------------------------
// Front-solve products on the matrix cores (k >= 8 right-hand sides).
//
// One 16x16x4 real MFMA computes D(16x16) += A(16x4) B(4x16).  A complex
// product out = M v (M: rows x inner, v: inner x k) maps onto two of them by
// splitting the 16 output columns into the real parts (j < 8) and imaginary
// parts (j >= 8) of 8 right-hand sides:
//     D += Re(M) [Re v | Im v]  +  Im(M) [-Im v | Re v],
// which is exactly 8 real flops per complex multiply-add, all on the MFMA
// units (no 3M recombination).  Lane l of a wave holds A(l % 16, l / 16) and
// B(l / 16, l % 16); its accumulator register q holds D(row, l % 16) with
// row = 4 (l / 16) + q for the f32 instruction and l / 16 + 4 q for f64.
// --------------------------------------------------------------------------
typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
    typedef v4f acc;
    static __device__ __forceinline__ acc step(float a, float b, acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int q) { return 4 * lk + q; }
};
template <>
struct Mfma<double> {
    typedef v4d acc;
    static __device__ __forceinline__ acc step(double a, double b, acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int q) { return lk + 4 * q; }
};

// Build-time tuning of the matrix-core solve kernel: factor loads in flight
// per lane (FP32 / FP64 fronts) and the waves per SIMD the register
// allocation must allow (0: the compiler's choice).
#ifndef PSP_MVM_U32
#define PSP_MVM_U32 8
#endif
#ifndef PSP_MVM_U64
#define PSP_MVM_U64 8
#endif
// Staged elements (columns x right-hand sides) per chunk, by front precision:
// smaller chunks keep fewer staging registers, so more workgroups fit on a
// CU (each launch then runs in fewer rounds of workgroups).
#ifndef PSP_MVM_CHUNK32
#define PSP_MVM_CHUNK32 1024
#endif
#ifndef PSP_MVM_CHUNK64
#define PSP_MVM_CHUNK64 512
#endif
#ifndef PSP_MVM_WAVES_PER_EU
#define PSP_MVM_WAVES_PER_EU 0
#endif
#if PSP_MVM_WAVES_PER_EU > 0
#define PSP_MVM_OCCUPANCY __attribute__((amdgpu_waves_per_eu(PSP_MVM_WAVES_PER_EU)))
#else
#define PSP_MVM_OCCUPANCY
#endif

// --------------------------------------------------------------------------
// Hand-offs between workgroups of one XCD.  Workgroups are dealt to the 8
// XCDs round-robin (blockIdx % 8 == XCC_ID, checked at device set-up), and
// the XCDs' L2 caches are not coherent with each other, so a hand-off that
// stays on one XCD needs no L2 write-back: the producer waits for its
// stores and raises a device-scope counter, the consumer sees the counter
// and reads the data with loads that miss the L1 (L2Buf).
// --------------------------------------------------------------------------
__device__ __forceinline__ int XccId()
{
    return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20);  // HW_REG_XCC_ID, bits 3:0
}

__device__ __forceinline__ int LoadAgent(const int32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int AddAgent(int32_t* p, int v)
{
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// This wave's stores (and loads) have completed: s_waitcnt vmcnt(0).
__device__ __forceinline__ void WaitVmem() { __builtin_amdgcn_s_waitcnt(0xF70); }

// Loads of vectors that another workgroup of the XCD wrote during this
// launch: buffer loads with the device-scope cache policy (sc1), which miss
// the CU's L1 and read the L2, where the producer's stores are.  They are
// ordinary loads to the compiler (scheduled and batched like any other).
// Measured on MI355X: with sc0 loads (workgroup scope), or with an L1
// invalidate after the wait, results went wrong whenever a vector row was
// not a whole number of 128-byte lines (a line then holds rows of two nodes,
// and the L1 can keep or regain a copy read before the other node was
// written); with sc1 loads every case is exact, and so are full
// device-scope release/acquire fences, which also write back and
// invalidate the L2.  A buffer spans at most 4 GiB (the host checks).
template <typename T>
struct L2Buf {
    __amdgpu_buffer_rsrc_t r;
    __device__ explicit L2Buf(const void* base)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, -1, 0x00020000))
    {
    }
    __device__ cx<T> operator[](int64_t i) const
    {
        constexpr int kSc1 = 16;  // cache policy: SC1 set (device scope)
        const uint32_t off = (uint32_t)(i * (int64_t)sizeof(cx<T>));
        cx<T> v;
        if constexpr (sizeof(T) == 4) {
            const auto x = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, kSc1);
            __builtin_memcpy(&v, &x, 8);
        } else {
            const auto x = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, kSc1);
            __builtin_memcpy(&v, &x, 16);
        }
        return v;
    }
};

__global__ void k_xcc_probe(int32_t* out)
{
    if (threadIdx.x == 0) out[blockIdx.x] = XccId();
}

// Matrix-core path (k >= 8).  One workgroup per tile; wave w owns rows
// 16 (w / WK) .. +16 of the tile and every WK-th 4-column step of its inner
// range.  The inner range streams through LDS in chunks of CC columns (all
// KB right-hand sides of a pass); the factor loads of the next U steps are
// issued before the current ones are consumed.  Waves that share rows sum
// their partial tiles in a fixed order through LDS.  Loads that do not
// depend on the products (the first factor batch, the staged vector, and
// for single-pass blocks the gathered bottom rows of F) are issued right
// after the tile record arrives, so a workgroup waits for few memory round
// trips.
//
// Inner slices of a split row block (slot >= 0) write partial sums to their
// slots.  With redc given, the host has placed the slices of each row block
// at launch positions that are equal mod 8, so they all run on one XCD: each
// slice waits for its partial stores and counts itself in redc[red], and the
// last to arrive sums the block's slots in slot order (the sums of
// k_mv_reduce, bit for bit), finishes the rows and clears the counter.
// Without redc a reduce launch follows.  Tiles with no rows are holes in that
// placement and exit at once.
//
// MvmTile is the work of one workgroup on tile tl (nrows > 0); k_mvm runs it
// on one panel solve's tiles, k_mvm_batch on the tiles of many.
template <typename T, int PH, int RT, int KB, int NWV>
__device__ __forceinline__ void MvmTile(const SolveTile tl, const int32_t* structs,
                                        const Slots<T> gs, const cx<T>* L, const cx<T>* LT,
                                        cx<T>* Xp, cx<T>* W, cx<T>* P, int k, int32_t* redc,
                                        const SolveTile* reds);

template <typename T, int PH, int RT, int KB, int NWV>
__global__ __launch_bounds__(64 * NWV) PSP_MVM_OCCUPANCY void k_mvm(const SolveTile* tiles, int start,
                                                     const int32_t* structs,
                                                     Slots<T> gs, const cx<T>* L,
                                                     const cx<T>* LT, cx<T>* Xp, cx<T>* W,
                                                     cx<T>* P, int k, int32_t* redc,
                                                     const SolveTile* reds)
{
    const SolveTile tl = tiles[start + blockIdx.x];
    if (tl.nrows <= 0) return;
    MvmTile<T, PH, RT, KB, NWV>(tl, structs, gs, L, LT, Xp, W, P, k, redc, reds);
}

template <typename T, int PH, int RT, int KB, int NWV>
__device__ __forceinline__ void MvmTile(const SolveTile tl, const int32_t* structs,
                                        const Slots<T> gs, const cx<T>* L, const cx<T>* LT,
                                        cx<T>* Xp, cx<T>* W, cx<T>* P, int k, int32_t* redc,
                                        const SolveTile* reds)
{
    typedef typename Mfma<T>::acc acc_t;
    constexpr int NT = 64 * NWV, WR = RT / 16, WK = NWV / WR, NB = KB / 8;
    // factor loads in flight per lane; 16-wave workgroups (128 VGPRs a lane)
    // keep fewer
    constexpr int U = (sizeof(T) == 4 ? PSP_MVM_U32 : PSP_MVM_U64) / (NWV == 16 ? 4 : 1);
    // staged inner columns per chunk; 16-wave tiles reuse the chunk buffer
    // as reduction scratch and keep the large one
    constexpr int CC = (NWV == 16 ? 2048 : sizeof(T) == 4 ? PSP_MVM_CHUNK32 : PSP_MVM_CHUNK64) / KB;
    constexpr bool PRE = KB <= 16 && PH == PH_F;
    static_assert(WR * WK == NWV && CC % 4 == 0 && U >= 1, "tile shape");
    static_assert(WK == 1 || 2 * CC * KB >= WR * (WK / 2) * NB * 256,
                  "reduction scratch must fit in vs");
    const NodeDev nd = tl.nd;
    const int s = nd.s, fs = nd.s + nd.u;
    const cx<T>* Lf = L + nd.l;
    const cx<T>* Lt = LT + nd.l;
    __shared__ cx<T> vs[CC][KB];
    __shared__ int arrived;
    T* red = reinterpret_cast<T*>(&vs[0][0]);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave / WK, wk = wave % WK;
    const int li = lane & 15, lk = lane >> 4;
    // Finishing lanes: wk == 0, li < 8; element (b, q) is row
    // 16 wr + row(lk, q) of the tile and right-hand side 8 b + li of the pass.
    const bool fin = wk == 0 && li < 8 && tl.slot < 0;
    cx<T> pre[PRE ? NB : 1][4];
    int prow[4] = {0, 0, 0, 0};
    if (PRE && fin && k <= KB) {
        // branch-free: rows clamped into the node's front (the relinds list
        // carries a sentinel entry), unused values dropped at the finish
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = min(tl.row0 + 16 * wr + Mfma<T>::row(lk, q), fs - 1);
            prow[q] = gs.relinds[nd.ustart + max(row - s, 0)];
        }
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = min(tl.row0 + 16 * wr + Mfma<T>::row(lk, q), fs - 1);
                const int64_t o = (nd.w + row) * k + min(8 * b + li, k - 1);
                pre[PRE ? b : 0][q] = gs.G[o] + gs.G[gs.half + o];
            }
    }
    for (int r0 = 0; r0 < k; r0 += KB) {
        const int kc = min(KB, k - r0);
        acc_t acc[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[b] = acc_t{0, 0, 0, 0};
        for (int cb = tl.col0; cb < tl.col1; cb += CC) {
            const int ncol = min(CC, tl.col1 - cb);
            const int nsteps = (ncol + 3) >> 2;
            const int mine = nsteps > wk ? (nsteps - wk + WK - 1) / WK : 0;
            cx<T> m[U];
            // Factor entries from a scalar base and a per-lane offset, rows
            // clamped into the tile, unmasked: the unused triangles of the
            // solve storage are zero, and a wave reads at most 3 columns past
            // the chunk (inside the storage's zero padding), against zeros
            // in the staged vector.  Steps past the wave's share reload its
            // last one and are not multiplied.
            const int fa = PH == PH_F ? fs : s;
            const cx<T>* Ab = (PH == PH_F ? Lf : Lt) + tl.row0 + (int64_t)cb * fa;
            const uint32_t la = (uint32_t)(min(16 * wr + li, tl.nrows - 1) + (4 * wk + lk) * fa);
            const int64_t ustep = (int64_t)(4 * WK) * fa;
            const int ulast = max(mine - 1, 0);
            auto load = [&](cx<T>* dst, int u0) {
#pragma unroll
                for (int u = 0; u < U; ++u) dst[u] = (Ab + min(u0 + u, ulast) * ustep)[la];
            };
            // staging loads first, then the first factor batch, all in flight
            // together; the LDS stores wait only for the staging loads
            // (columns up to the next multiple of 4 are written, as zeros
            // past the range)
            constexpr int SPT = CC * KB / NT;
            const int nst = min(CC, 4 * nsteps);
            cx<T> sv[SPT];
            StageBatch<T, PH>(nd, structs, gs, Xp, W, k, [&](int q) {
                const int e = tid + NT * q, c = e / KB, r = e % KB;
                return int2{cb + min(c, ncol - 1), min(r0 + r, k - 1)};
            }, sv);
            load(m, 0);
            LdsBarrier();  // the previous chunk's (or pass's) readers are done with vs
#pragma unroll
            for (int q = 0; q < SPT; ++q) {
                const int e = tid + NT * q, c = e / KB, r = e % KB;
                if (c < nst) vs[c][r] = (r < kc && c < ncol) ? sv[q] : czero<T>();
            }
            LdsBarrier();
            for (int u0 = 0; u0 < mine; u0 += U) {
                cx<T> mn[U];
                if (u0 + U < mine) load(mn, u0 + U);
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (u0 + u < mine) {
                        const int c = 4 * (wk + WK * (u0 + u)) + lk;
#pragma unroll
                        for (int b = 0; b < NB; ++b) {
                            const cx<T> v = vs[c][8 * b + (li & 7)];
                            const T b1 = li < 8 ? v.r : v.i;
                            const T b2 = li < 8 ? -v.i : v.r;
                            acc[b] = Mfma<T>::step(m[u].r, b1, acc[b]);
                            acc[b] = Mfma<T>::step(m[u].i, b2, acc[b]);
                        }
                    }
                }
                if (u0 + U < mine) {
#pragma unroll
                    for (int u = 0; u < U; ++u) m[u] = mn[u];
                }
            }
        }
        // waves that share rows sum their partial tiles by a fixed tree
        if (WK > 1) {
            for (int h = WK / 2; h >= 1; h >>= 1) {
                LdsBarrier();
                if (wk >= h && wk < 2 * h) {
#pragma unroll
                    for (int b = 0; b < NB; ++b)
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            red[(((wr * h + wk - h) * NB + b) * 64 + lane) * 4 + q] = acc[b][q];
                }
                LdsBarrier();
                if (wk < h) {
#pragma unroll
                    for (int b = 0; b < NB; ++b)
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            acc[b][q] += red[(((wr * h + wk) * NB + b) * 64 + lane) * 4 + q];
                }
            }
        }
        if (wk == 0) {
#pragma unroll
            for (int b = 0; b < NB; ++b) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const T val = acc[b][q];
                    const T other = __shfl_xor(val, 8, 64);
                    const int rl = 16 * wr + Mfma<T>::row(lk, q);  // row within the tile
                    const int rr = 8 * b + li;                      // right-hand side within the pass
                    if (li < 8 && rl < tl.nrows && rr < kc) {
                        const cx<T> a{val, other};
                        const int row = tl.row0 + rl;
                        if (tl.slot >= 0)
                            P[((int64_t)tl.slot * 64 + rl) * k + r0 + rr] = a;
                        else if (PRE && k <= KB)
                            Finish<T, PH>(nd, gs, Xp, W, row, rr, k, a, pre[PRE ? b : 0][q],
                                          prow[q]);
                        else
                            FinishGather<T, PH>(nd, gs, Xp, W, row, r0 + rr, k, a);
                    }
                }
            }
        }
    }
    if (redc != nullptr && tl.slot >= 0) {
        const int rix = tl.pad_;
        WaitVmem();
        LdsBarrier();  // every wave's partial sums have reached the L2
        if (tid == 0) arrived = AddAgent(redc + rix, 1);
        LdsBarrier();
        const SolveTile it = reds[rix];  // row block: first slot, slot count
        if (__builtin_amdgcn_readfirstlane(arrived) == it.col1 - 1) {
            const L2Buf<T> lP(P);
            constexpr int kS = 16;  // slots loaded together (more: in groups)
            for (int e = tid; e < it.nrows * k; e += NT) {
                const int rl = e / k, r = e % k;
                cx<T> a = czero<T>();
                for (int j0 = 0; j0 < it.col1; j0 += kS) {
                    // all of a group's loads in flight at once, then summed in
                    // slot order (slots past the block add nothing)
                    cx<T> v[kS];
#pragma unroll
                    for (int j = 0; j < kS; ++j)
                        v[j] = lP[((int64_t)(it.col0 + min(j0 + j, it.col1 - 1)) * 64 + rl) * k + r];
#pragma unroll
                    for (int j = 0; j < kS; ++j)
                        if (j0 + j < it.col1) a = a + v[j];
                }
                FinishGather<T, PH>(it.nd, gs, Xp, W, it.row0 + rl, r, k, a);
            }
            if (tid == 0) redc[rix] = 0;  // no slice of this launch counts any more
        }
    }
}

// Persistent, pipelined matrix-core path (8 <= k <= KB, KB = 8 or 16: one
// pass over the right-hand sides).  gridDim.x workgroups take the launch's
// tiles round-robin (their descriptors are read into LDS once) and work
// through them as a sequence of steps of CW inner columns.  Two register
// sets alternate: while step n is multiplied, the staging and factor loads
// of step n + 1 are in flight, and in the backward phase also the struct
// rows that the gather of step n + 2 needs.  The products, the summation
// order within a tile (4-column steps in increasing order, the waves of a
// row block summed by the same fixed tree) and the finish are those of
// k_mvm; only the schedule differs.  No workgroup waits for another.
//
// The loop is written for few vector instructions per load: a tile's
// fields are made wave-uniform (so its base addresses are scalar), every
// lane's offsets within a step are fixed for the launch, and factor entries
// are used unmasked.  That relies on the solve storage: the triangles a
// phase does not use are zero (k_transpose_scale), the factor and vector
// buffers end in kSolvePad padding columns / rows of zeros, so the entries
// a partial step reads past its tile are finite and meet a zero in the
// staged vector (the staged vector is masked), and rows past a partial
// tile read the tile's last row and give outputs that are not stored.
constexpr int kMvpTiles = 32;  // tiles per workgroup at most (host sizes the grid)
#ifndef PSP_MVP_U
#define PSP_MVP_U 8  // factor loads per lane per step, 64- and 32-row tiles
#endif
#ifndef PSP_MVP_U16
#define PSP_MVP_U16 4  // factor loads per lane per step, 16-row tiles
#endif

template <typename T, int PH, int RT, int KB>
struct MvpShape {
    static constexpr int NT = 256, WR = RT / 16, WK = 4 / WR, NB = KB / 8;
    static constexpr int U = RT == 16 ? PSP_MVP_U16 : PSP_MVP_U;  // factor loads per lane per step
    static constexpr int CW = 4 * WK * U;        // inner columns per step
    static constexpr int SPT = CW * KB / NT;     // staged elements per thread per step
    static_assert(CW * KB == SPT * NT && SPT >= 1, "step shape");
    static_assert(CW <= kSolvePad, "a step reads at most kSolvePad columns past its tile");
};

__device__ __forceinline__ int Uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t Uni(int64_t v)
{
    const uint64_t u = (uint64_t)v;
    return (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32)) << 32) |
                     (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u));
}

template <typename T, int PH, int RT, int KB>
__global__ __launch_bounds__(256) void k_mvp(const SolveTile* tiles, int start, int count,
                                             const int32_t* structs, Slots<T> gs,
                                             const cx<T>* L, const cx<T>* LT, cx<T>* Xp,
                                             cx<T>* W, cx<T>* P, int k)
{
    typedef typename Mfma<T>::acc acc_t;
    typedef MvpShape<T, PH, RT, KB> S;
    constexpr int NT = S::NT, WK = S::WK, NB = S::NB, U = S::U, CW = S::CW, SPT = S::SPT;
    constexpr int NRED = WK > 1 ? S::WR * (WK / 2) * NB * 256 : 1;
    typedef typename std::conditional<PH == PH_F, Raw3<T>, cx<T>>::type Staged;
    struct Pre {       // forward finish of a tile: gathered rows and parent rows
        Raw2<T> g[PH == PH_F ? NB : 1][4];
        int32_t prow[4];
    };
    struct Data {      // one step's loads
        cx<T> m[U];    // factor entries
        Staged v[SPT]; // staged inner operand (raw)
        Pre pre;       // forward, last step of a tile: its finish data
    };
    struct Idx {       // backward: panel rows of one step's update columns
        int32_t x[PH == PH_B ? SPT : 1];
    };
    struct Tile {      // wave-uniform fields of a tile
        int64_t l, w;
        int off, s, u, ustart, pw, cslot, row0, nrows, col0, col1, slot;
    };
    __shared__ SolveTile tq[kMvpTiles];
    // staged vector in the two forms the B operand takes: lanes l % 16 < 8
    // read (re, -im), the others (im, re)
    __shared__ T2<T> vs[2][2][CW][KB];
    __shared__ T red[NRED];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave / WK, wk = wave % WK;
    const int li = lane & 15, lk = lane >> 4;
    const int G = gridDim.x;
    const int nmine = (count - (int)blockIdx.x + G - 1) / G;
    {
        const int4* src = reinterpret_cast<const int4*>(tiles + start);
        int4* dst = reinterpret_cast<int4*>(tq);
        for (int e = tid; e < nmine * 4; e += NT)
            dst[e] = src[((int64_t)blockIdx.x + (int64_t)(e >> 2) * G) * 4 + (e & 3)];
    }
    LdsBarrier();
    if (nmine <= 0) return;
    auto tile = [&](int i) {
        const SolveTile& t = tq[i];
        Tile r;
        r.l = Uni(t.nd.l);
        r.w = Uni(t.nd.w);
        r.off = Uni(t.nd.off);
        r.s = Uni(t.nd.s);
        r.u = Uni(t.nd.u);
        r.ustart = Uni(t.nd.ustart);
        r.pw = Uni(t.nd.pw);
        r.cslot = Uni(t.nd.cslot);
        r.row0 = Uni(t.row0);
        r.nrows = Uni(t.nrows);
        r.col0 = Uni(t.col0);
        r.col1 = Uni(t.col1);
        r.slot = Uni(t.slot);
        return r;
    };
    // lane constants of the launch
    const int lrow = 16 * wr + li;        // factor row within the tile
    const int lcol = 4 * wk + lk;         // factor column within a 4 WK-column group
    int sc[SPT], sr[SPT];                 // staged element: column within the step, rhs
#pragma unroll
    for (int q = 0; q < SPT; ++q) sc[q] = (tid + NT * q) / KB, sr[q] = (tid + NT * q) % KB;

    struct Cur {
        int i, c0;  // tile (of this workgroup) and first inner column of the step
    };
    auto ncolsOf = [&](int i) { return max(Uni(tq[i].col1) - Uni(tq[i].col0), 1); };
    auto advance = [&](Cur c) {
        if (c.i >= nmine) return c;  // past the end stays there
        c.c0 += CW;
        if (c.c0 >= ncolsOf(c.i)) {
            ++c.i;
            c.c0 = 0;
        }
        return c;
    };
    auto issueIdx = [&](Cur c, Idx& ix) {
        if constexpr (PH == PH_B) {
            const int ustart = Uni(tq[c.i].nd.ustart), s = Uni(tq[c.i].nd.s);
            const int cb = Uni(tq[c.i].col0) + c.c0;
            const int32_t* sb = structs + ustart;  // read up to kSolvePad past the list
#pragma unroll
            for (int q = 0; q < SPT; ++q) ix.x[q] = sb[max(cb + sc[q] - s, 0)];
        }
    };
    auto issueData = [&](Cur c, Data& d, Idx& ix) {
        const Tile t = tile(c.i);
        const int cb = t.col0 + c.c0;
        const int fa = PH == PH_F ? t.s + t.u : t.s;  // column stride of the factor block
        const cx<T>* Ab = (PH == PH_F ? L : LT) + t.l + t.row0 + (int64_t)cb * fa;
        const uint32_t la = (uint32_t)(min(lrow, t.nrows - 1) + lcol * fa);
        const int64_t ustep = (int64_t)(4 * WK) * fa;
        if constexpr (PH == PH_B) {
#pragma unroll
            for (int q = 0; q < SPT; ++q) ix.x[q] = Opaque(ix.x[q]);
        }
        if constexpr (PH == PH_F) {
            const cx<T>* g0 = gs.G + (t.w + cb) * k;
            const cx<T>* g1 = g0 + gs.half;
            const cx<T>* xb = Xp + (int64_t)(t.off + cb) * k;
#pragma unroll
            for (int q = 0; q < SPT; ++q) {
                const uint32_t o = (uint32_t)(sc[q] * k + sr[q]);
                d.v[q] = Raw3<T>{g0[o], g1[o], xb[o]};
            }
        } else {
#pragma unroll
            for (int q = 0; q < SPT; ++q) {
                const int col = cb + sc[q];
                const int64_t row = col < t.s ? t.w + col : (int64_t)ix.x[q];
                d.v[q] = (col < t.s ? W : Xp)[row * k + sr[q]];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) d.m[u] = (Ab + u * ustep)[la];
    };
    auto issuePre = [&](int i, Pre& pr) {
        if constexpr (PH == PH_F) {
            const Tile t = tile(i);
            const int fs = t.s + t.u;
            const int32_t* rb = gs.relinds + t.ustart;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = min(t.row0 + 16 * wr + Mfma<T>::row(lk, q), fs - 1);
                pr.prow[q] = rb[max(row - t.s, 0)];
            }
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int row = min(t.row0 + 16 * wr + Mfma<T>::row(lk, q), fs - 1);
                    const int64_t o = (t.w + row) * k + min(8 * b + li, k - 1);
                    pr.g[b][q] = Raw2<T>{gs.G[o], gs.G[gs.half + o]};
                }
        }
    };
    acc_t acc[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[b] = acc_t{0, 0, 0, 0};
    auto finishTile = [&](int i, const Pre& pr) {
        if (WK > 1) {  // waves that share rows sum their partial tiles by a fixed tree
            for (int h = WK / 2; h >= 1; h >>= 1) {
                LdsBarrier();
                if (wk >= h && wk < 2 * h) {
#pragma unroll
                    for (int b = 0; b < NB; ++b)
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            red[(((wr * h + wk - h) * NB + b) * 64 + lane) * 4 + q] = acc[b][q];
                }
                LdsBarrier();
                if (wk < h) {
#pragma unroll
                    for (int b = 0; b < NB; ++b)
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            acc[b][q] += red[(((wr * h + wk) * NB + b) * 64 + lane) * 4 + q];
                }
            }
        }
        if (wk == 0) {
            const Tile t = tile(i);
            NodeDev nd;
            nd.l = t.l, nd.w = t.w, nd.off = t.off, nd.s = t.s, nd.u = t.u;
            nd.ustart = t.ustart, nd.pw = t.pw, nd.cslot = t.cslot;
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const T val = acc[b][q];
                    const T other = __shfl_xor(val, 8, 64);
                    const int rl = 16 * wr + Mfma<T>::row(lk, q), rr = 8 * b + li;
                    if (li < 8 && rl < t.nrows && rr < k) {
                        const cx<T> a{val, other};
                        if (t.slot >= 0)
                            P[((int64_t)t.slot * 64 + rl) * k + rr] = a;
                        else if constexpr (PH == PH_F)
                            Finish<T, PH>(nd, gs, Xp, W, t.row0 + rl, rr, k, a,
                                          pr.g[b][q].a + pr.g[b][q].b, pr.prow[q]);
                        else
                            Finish<T, PH>(nd, gs, Xp, W, t.row0 + rl, rr, k, a, czero<T>(), 0);
                    }
                }
        }
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[b] = acc_t{0, 0, 0, 0};
    };
    // B operand source of this lane (fixed for the launch)
    const int form = li < 8 ? 0 : 1, bj = li & 7;
    auto consume = [&](Cur c, const Data& d, int buf) {
        const int nc = Uni(tq[c.i].col1) - (Uni(tq[c.i].col0) + c.c0);
#pragma unroll
        for (int q = 0; q < SPT; ++q) {
            cx<T> v;
            if constexpr (PH == PH_F)
                v = d.v[q].a + d.v[q].b + d.v[q].c;
            else
                v = d.v[q];
            const bool ok = sc[q] < nc && sr[q] < k;
            const T re = ok ? v.r : T(0), im = ok ? v.i : T(0);
            vs[buf][0][sc[q]][sr[q]] = T2<T>{re, -im};
            vs[buf][1][sc[q]][sr[q]] = T2<T>{im, re};
        }
        LdsBarrier();
        const T2<T>(*vb)[KB] = vs[buf][form];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int cl = lcol + 4 * WK * u;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const T2<T> w = vb[cl][8 * b + bj];
                acc[b] = Mfma<T>::step(d.m[u].r, w.x, acc[b]);
                acc[b] = Mfma<T>::step(d.m[u].i, w.y, acc[b]);
            }
        }
        if (c.c0 + CW >= ncolsOf(c.i)) finishTile(c.i, d.pre);
        if constexpr (PH == PH_F) {  // the finish data was issued one step ago
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    Sink(d.pre.g[b][q].a.r), Sink(d.pre.g[b][q].a.i);
                    Sink(d.pre.g[b][q].b.r), Sink(d.pre.g[b][q].b.i);
                }
#pragma unroll
            for (int q = 0; q < 4; ++q) Sink(d.pre.prow[q]);
        }
    };

    Data d0, d1;
    Idx i0, i1;
    // steps n, n + 1, n + 2: data of n sits in d[n % 2], its indices were in
    // i[n % 2]; a cursor is past the end when its tile index reaches nmine.
    // The index and data loads are issued unconditionally (a cursor past the
    // end reloads the current step, unused), so the wait counts before a
    // step's use never have to allow for a path without the loads issued
    // after it.  The finish data of a tile travels with its last step in the
    // same register set; it is issued after the multiply of the previous
    // step, one step before it is used.
    Cur cn{0, 0};
    Cur cn1 = advance(cn);
    Cur cn2 = advance(cn1);
    auto lastStep = [&](Cur c) { return c.i < nmine && c.c0 + CW >= ncolsOf(c.i); };
    issueIdx(cn, i0);
    issueIdx(cn1.i < nmine ? cn1 : cn, i1);
    issueData(cn, d0, i0);
    if (PH == PH_F && lastStep(cn)) issuePre(cn.i, d0.pre);
    auto half = [&](Idx& iN, Idx& iN1, Data& dN, Data& dN1, int buf) {
        issueIdx(cn2.i < nmine ? cn2 : cn, iN);
        issueData(cn1.i < nmine ? cn1 : cn, dN1, iN1);
        consume(cn, dN, buf);
        if (PH == PH_F && lastStep(cn1)) issuePre(cn1.i, dN1.pre);
        cn = cn1;
        cn1 = cn2;
        cn2 = advance(cn2);
    };
    // On the way out, every load still in flight (the reloads past the end)
    // is waited for: the compiler merges the two exits with the path back
    // to the loop head, and a load pending on any of them would otherwise
    // be drained at the head of every iteration.
    auto drain = [&](const Data& d, const Idx& ix) {
#pragma unroll
        for (int u = 0; u < U; ++u) Sink(d.m[u].r), Sink(d.m[u].i);
#pragma unroll
        for (int q = 0; q < SPT; ++q) {
            if constexpr (PH == PH_F)
                Sink(d.v[q].a.r), Sink(d.v[q].a.i), Sink(d.v[q].b.r), Sink(d.v[q].b.i),
                    Sink(d.v[q].c.r), Sink(d.v[q].c.i);
            else
                Sink(d.v[q].r), Sink(d.v[q].i), Sink(ix.x[q]);
        }
    };
    for (;;) {
        half(i0, i1, d0, d1, 0);
        if (cn.i >= nmine) {
            drain(d1, i0);
            break;
        }
        half(i1, i0, d1, d0, 1);
        if (cn.i >= nmine) {
            drain(d0, i1);
            break;
        }
    }
}

// --------------------------------------------------------------------------
// Fused subtree solves (8 <= k <= KB).  One persistent launch runs every
// level of a phase from the class's cut depth down (forward: deepest level
// first; backward: the cut level first).  The subtrees under the cut nodes
// are split over nqueues queues, one per XCD: a workgroup reads the XCD it
// runs on (hardware register XCC_ID), serves queue XCC_ID (if there is
// one; no two XCDs ever serve one queue), takes its tiles in
// queue order (an atomic head), and before reading a tile's inputs waits
// until the nodes the tile depends on have all their row blocks done
// (per-node counters; forward: the node's children, backward: its nearest
// ancestor with pivots).  Producer and consumer of every hand-off then
// share one XCD and its L2, and the hand-off is the cross-CU one of the
// memory model, with no L2 write-back:
//   producer: its stores; s_waitcnt vmcnt(0) (the stores have reached the
//             L2); barrier; one lane increments the node's counter;
//   consumer: polls the counter; barrier; reads what the launch wrote with
//             loads that miss the L1 and read the L2 (sc1, L2Buf).
// Counters and heads are device-scope atomics.  Near the cut a row block
// is cut into inner slices: slices write partial sums to their slots, and
// the slice that arrives last (a counter per row block) sums the slots in
// slot order and finishes the rows, so sums have a fixed order.
//
// No workgroup waits forever: a tile's dependencies come earlier in its
// queue, and the earliest tile not yet done is held by a running workgroup
// whose own earlier tiles are done, so it proceeds.  A queue that no
// workgroup serves (fewer XCDs than queues) is run, tile by tile, by the
// last workgroup to leave (an exit counter).  The per-tile body is that of
// k_mvm: waves own 16-row blocks and every WK-th 4-column step, the inner
// range streams through LDS, factor loads run U steps ahead; the first
// factor batch of a tile is issued before its dependency wait.  The launch
// state (counters, heads) is cleared by the copy-in or zero-set launch
// that precedes each panel solve.
// --------------------------------------------------------------------------
// The fused launch's buffers read through the L2: slot copies, work vectors,
// panel vector, partial sums.
template <typename T>
struct L2Bufs {
    L2Buf<T> G, W, X, P;
};

// StageBatch of the fused launch: the same operand (the same sums in the
// same order) with the vectors written in this launch read from the L2.
// Backward: both the work-vector and the panel-vector element are loaded
// (a buffer per vector, so no per-lane choice of descriptor) and one kept.
template <typename T, int PH, int N, class AT>
__device__ __forceinline__ void StageBatchL2(const NodeDev& nd, const int32_t* structs,
                                             const Slots<T>& gs, const L2Bufs<T>& lb,
                                             const cx<T>* Xp, int k, const AT& at,
                                             cx<T> (&sv)[N])
{
    int32_t sx[N];
#pragma unroll
    for (int q = 0; q < N; ++q) sx[q] = StageIndex<PH>(nd, structs, at(q).x);
    if (PH == PH_B) {
#pragma unroll
        for (int q = 0; q < N; ++q) sx[q] = Opaque(sx[q]);
    }
    // all loads first, combined once they have arrived
    cx<T> ra[N], rb[N], rc[N];
#pragma unroll
    for (int q = 0; q < N; ++q) {
        const int2 p = at(q);
        const int col = p.x, rr = p.y;
        if (PH == PH_F) {
            const int64_t g = (nd.w + col) * k + rr;
            ra[q] = (lb.G[g]);
            rb[q] = (lb.G[gs.half + g]);
            rc[q] = (Xp[(int64_t)(nd.off + col) * k + rr]);
        } else {
            ra[q] = (lb.W[(nd.w + min(col, nd.s - 1)) * k + rr]);
            rb[q] = (lb.X[(int64_t)sx[q] * k + rr]);
        }
    }
#pragma unroll
    for (int q = 0; q < N; ++q) {
        if (PH == PH_F)
            sv[q] = ra[q] + rb[q] + rc[q];
        else
            sv[q] = at(q).x < nd.s ? ra[q] : rb[q];
    }
}

template <typename T, int PH, int RT, int KB>
__global__ __launch_bounds__(256) void k_mvq(const QTile* tiles, const int32_t* qstart,
                                             const QRed* reds, const int4* dep,
                                             const NodeDev* nodes, int32_t* sync, int nnodes,
                                             int nqueues, const int32_t* structs, Slots<T> gs,
                                             const cx<T>* L, const cx<T>* LT, cx<T>* Xp,
                                             cx<T>* W, cx<T>* P, int k)
{
    typedef typename Mfma<T>::acc acc_t;
    constexpr int NWV = 4, NT = 256, WR = RT / 16, WK = NWV / WR, NB = KB / 8;
    constexpr int U = sizeof(T) == 4 ? PSP_MVM_U32 : PSP_MVM_U64;
    constexpr int CC = (sizeof(T) == 4 ? PSP_MVM_CHUNK32 : PSP_MVM_CHUNK64) / KB;
    constexpr int SPT = CC * KB / NT;
    constexpr bool PRE = PH == PH_F;
    static_assert(WR * WK == NWV && CC % 4 == 0 && SPT >= 1, "tile shape");
    static_assert(WK == 1 || 2 * CC * KB >= WR * (WK / 2) * NB * 256,
                  "reduction scratch must fit in vs");
    int32_t* done = sync;
    int32_t* head = sync + nnodes;
    int32_t* exitc = head + nqueues;
    int32_t* redc = exitc + 1;
    const L2Bufs<T> lb{L2Buf<T>(gs.G), L2Buf<T>(W), L2Buf<T>(Xp), L2Buf<T>(P)};
    __shared__ cx<T> vs[CC][KB];
    __shared__ int bc[4];
    T* red = reinterpret_cast<T*>(&vs[0][0]);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave / WK, wk = wave % WK;
    const int li = lane & 15, lk = lane >> 4;

    // One queue tile: wait for its dependencies, the products, the finish
    // (or the partial sums and, for the last slice, the row block's sum and
    // finish), then the node's counter.
    auto runTile = [&](int idx) {
        const QTile qt = tiles[idx];
        const NodeDev nd = nodes[qt.node];
        const int s = nd.s, fs = nd.s + nd.u;
        const int fa = PH == PH_F ? fs : s;
        const cx<T>* Ab0 = (PH == PH_F ? L : LT) + nd.l + qt.row0;
        const uint32_t la = (uint32_t)(min(16 * wr + li, qt.nrows - 1) + (4 * wk + lk) * fa);
        const int64_t ustep = (int64_t)(4 * WK) * fa;
        // the first factor batch (independent of the dependencies)
        const int ncol0 = min(CC, qt.col1 - qt.col0);
        const int nsteps0 = (max(ncol0, 0) + 3) >> 2;
        const int mine0 = nsteps0 > wk ? (nsteps0 - wk + WK - 1) / WK : 0;
        cx<T> m[U];
        {
            const cx<T>* Ab = Ab0 + (int64_t)qt.col0 * fa;
            const int ulast = max(mine0 - 1, 0);
#pragma unroll
            for (int u = 0; u < U; ++u) m[u] = (Ab + min(u, ulast) * ustep)[la];
        }
        int prow[4] = {0, 0, 0, 0};
        const bool fin = wk == 0 && li < 8 && qt.red < 0;
        if (PRE && fin) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = min(qt.row0 + 16 * wr + Mfma<T>::row(lk, q), fs - 1);
                prow[q] = gs.relinds[nd.ustart + max(row - s, 0)];
            }
        }
        // dependencies
        {
            const int4 dp = dep[qt.node];
            if (tid == 0) {
                if (dp.x >= 0)
                    while (LoadAgent(done + dp.x) < dp.z) __builtin_amdgcn_s_sleep(1);
                if (dp.y >= 0)
                    while (LoadAgent(done + dp.y) < dp.w) __builtin_amdgcn_s_sleep(1);
            }
            LdsBarrier();
        }
        cx<T> pre[PRE ? NB : 1][4], pre1[PRE ? NB : 1][4];
        if (PRE && fin) {
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int row = min(qt.row0 + 16 * wr + Mfma<T>::row(lk, q), fs - 1);
                    const int64_t o = (nd.w + row) * k + min(8 * b + li, k - 1);
                    pre[PRE ? b : 0][q] = (lb.G[o]);
                    pre1[PRE ? b : 0][q] = (lb.G[gs.half + o]);
                }
        }
        acc_t acc[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[b] = acc_t{0, 0, 0, 0};
        for (int cb = qt.col0; cb < qt.col1; cb += CC) {
            const int ncol = min(CC, qt.col1 - cb);
            const int nsteps = (ncol + 3) >> 2;
            const int mine = nsteps > wk ? (nsteps - wk + WK - 1) / WK : 0;
            const cx<T>* Ab = Ab0 + (int64_t)cb * fa;
            const int ulast = max(mine - 1, 0);
            auto load = [&](cx<T>* dst, int u0) {
#pragma unroll
                for (int u = 0; u < U; ++u) dst[u] = (Ab + min(u0 + u, ulast) * ustep)[la];
            };
            const int nst = min(CC, 4 * nsteps);
            cx<T> sv[SPT];
            StageBatchL2<T, PH>(nd, structs, gs, lb, Xp, k, [&](int q) {
                const int e = tid + NT * q, c = e / KB, r = e % KB;
                return int2{cb + min(c, ncol - 1), min(r, k - 1)};
            }, sv);
            if (cb != qt.col0) load(m, 0);
            LdsBarrier();  // the previous chunk's readers are done with vs
#pragma unroll
            for (int q = 0; q < SPT; ++q) {
                const int e = tid + NT * q, c = e / KB, r = e % KB;
                if (c < nst) vs[c][r] = (r < k && c < ncol) ? sv[q] : czero<T>();
            }
            LdsBarrier();
            for (int u0 = 0; u0 < mine; u0 += U) {
                cx<T> mn[U];
                if (u0 + U < mine) load(mn, u0 + U);
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (u0 + u < mine) {
                        const int c = 4 * (wk + WK * (u0 + u)) + lk;
#pragma unroll
                        for (int b = 0; b < NB; ++b) {
                            const cx<T> v = vs[c][8 * b + (li & 7)];
                            const T b1 = li < 8 ? v.r : v.i;
                            const T b2 = li < 8 ? -v.i : v.r;
                            acc[b] = Mfma<T>::step(m[u].r, b1, acc[b]);
                            acc[b] = Mfma<T>::step(m[u].i, b2, acc[b]);
                        }
                    }
                }
                if (u0 + U < mine) {
#pragma unroll
                    for (int u = 0; u < U; ++u) m[u] = mn[u];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) Sink(m[u].r), Sink(m[u].i);
        if (WK > 1) {  // waves that share rows sum their partial tiles by a fixed tree
            for (int h = WK / 2; h >= 1; h >>= 1) {
                LdsBarrier();
                if (wk >= h && wk < 2 * h) {
#pragma unroll
                    for (int b = 0; b < NB; ++b)
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            red[(((wr * h + wk - h) * NB + b) * 64 + lane) * 4 + q] = acc[b][q];
                }
                LdsBarrier();
                if (wk < h) {
#pragma unroll
                    for (int b = 0; b < NB; ++b)
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            acc[b][q] += red[(((wr * h + wk) * NB + b) * 64 + lane) * 4 + q];
                }
            }
        }
        if (wk == 0) {
#pragma unroll
            for (int b = 0; b < NB; ++b) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const T val = acc[b][q];
                    const T other = __shfl_xor(val, 8, 64);
                    const int rl = 16 * wr + Mfma<T>::row(lk, q);
                    const int rr = 8 * b + li;
                    if (li < 8 && rl < qt.nrows && rr < k) {
                        const cx<T> a{val, other};
                        if (qt.red >= 0)
                            P[((int64_t)qt.slot * 64 + rl) * k + rr] = a;
                        else if (PRE)
                            Finish<T, PH>(nd, gs, Xp, W, qt.row0 + rl, rr, k, a,
                                          pre[PRE ? b : 0][q] + pre1[PRE ? b : 0][q], prow[q]);
                        else
                            Finish<T, PH>(nd, gs, Xp, W, qt.row0 + rl, rr, k, a, czero<T>(), 0);
                    }
                }
            }
        }
        WaitVmem();
        LdsBarrier();
        bool rowsDone = qt.red < 0;
        if (!rowsDone) {
            const QRed rd = reds[qt.red];
            if (tid == 0) bc[2] = AddAgent(redc + qt.red, 1);
            LdsBarrier();
            if (__builtin_amdgcn_readfirstlane(bc[2]) == rd.nslots - 1) {
                // the last slice: every other slice's partials are in L2
                for (int e = tid; e < rd.nrows * k; e += NT) {
                    const int rl = e / k, r = e % k;
                    cx<T> a = czero<T>();
                    for (int j = 0; j < rd.nslots; ++j)
                        a = a + lb.P[((int64_t)(rd.slot0 + j) * 64 + rl) * k + r];
                    if (PH == PH_F && rd.row0 + rl >= nd.s) {
                        const int row = rd.row0 + rl;
                        const int64_t g = (nd.w + row) * k + r;
                        Finish<T, PH>(nd, gs, Xp, W, row, r, k, a, lb.G[g] + lb.G[gs.half + g],
                                      gs.relinds[nd.ustart + row - nd.s]);
                    } else {
                        Finish<T, PH>(nd, gs, Xp, W, rd.row0 + rl, r, k, a, czero<T>(), 0);
                    }
                }
                WaitVmem();
                LdsBarrier();
                rowsDone = true;
            }
        }
        if (rowsDone && tid == 0) AddAgent(done + qt.node, 1);
    };

    // queue x is served by the workgroups of XCD x only (none of another XCD
    // ever shares it); XCDs without a queue go straight to the exit
    const int q = __builtin_amdgcn_readfirstlane(XccId());
    int cur = 0, qs = 0, qn = 0;
    if (q < nqueues) {
        qs = qstart[q];
        qn = qstart[q + 1] - qs;
        if (tid == 0) bc[0] = AddAgent(head + q, 1);
        LdsBarrier();
        cur = __builtin_amdgcn_readfirstlane(bc[0]);
    }
    while (cur < qn) {
        int nxt = 0;
        if (tid == 0) nxt = AddAgent(head + q, 1);  // the tile after this one
        runTile(qs + cur);
        if (tid == 0) bc[1] = nxt;
        LdsBarrier();
        cur = __builtin_amdgcn_readfirstlane(bc[1]);
    }
    // the last workgroup out runs the queues no workgroup served
    if (tid == 0) bc[3] = AddAgent(exitc, 1);
    LdsBarrier();
    if (__builtin_amdgcn_readfirstlane(bc[3]) != (int)gridDim.x - 1) return;
    for (int qq = 0; qq < nqueues; ++qq) {
        if (tid == 0) bc[0] = LoadAgent(head + qq);
        LdsBarrier();
        const int served = __builtin_amdgcn_readfirstlane(bc[0]);
        LdsBarrier();
        if (served != 0) continue;
        for (int i = qstart[qq]; i < qstart[qq + 1]; ++i) runTile(i);
    }
}


// --------------------------------------------------------------------------
// Complex tile engine on the matrix cores: one 256-thread workgroup computes
// an RT x 64 complex tile C = A B over an inner range, K staged through LDS
// 16 columns at a time with two buffers (the global loads of stage s + 1
// are in flight while stage s is multiplied).  Waves split the tile into
// WRW x WCW parts; each wave holds RB x CB accumulator blocks of 16 rows x 8
// complex columns (16 real MFMA columns: real parts, then imaginary parts).
// A and B are read through policy objects: a(i, k) and b(k, j) return the
// complex entry or zero outside the operand.
// --------------------------------------------------------------------------
template <int RT, int CT = 64>
struct TileShape {
    static constexpr int WRW = RT >= 64 ? 2 : 1;       // waves along rows
    static constexpr int WCW = 4 / WRW;                // waves along columns
    static constexpr int RB = RT / 16 / WRW;           // 16-row blocks per wave
    static constexpr int CB = CT / 8 / WCW;            // 8-column blocks per wave
};

constexpr int kBK = 16;  // inner columns per LDS stage (solve tiles)

// la(i, k) / lb(k, j) return the raw loads of one A / B element and must be
// safe to call for every i < RT, j < CT and k in [kbeg, kend) (loaders clamp
// their own indices, so a stage's loads are issued without branches);
// fa(i, k, raw) / fb(k, j, raw) turn them into the operand value, masks and
// scaling included.  The transforms run when the next stage is written to
// LDS, after this stage's matrix-core steps, so the loads of stage s + 1 are
// in flight while stage s computes.
// skipAbove: (lower-triangular results) a 16-row x 8-column block whose
// rows all lie above its columns, i + skipAbove < j for tile coordinates
// (i, j), is not multiplied; its accumulators stay zero.  1 << 30: none.
template <typename T, int RT, bool BJFAST, int BK = kBK, int CT = 64, class LA, class FA, class LB,
          class FB>
__device__ __forceinline__ void MmaTile(const LA& la, const FA& fa, const LB& lb, const FB& fb,
                                        int kbeg, int kend, cx<T> (*As)[BK][RT + 1],
                                        cx<T> (*Bs)[BK][CT + 1],
                                        typename Mfma<T>::acc (&acc)[TileShape<RT, CT>::RB]
                                                                   [TileShape<RT, CT>::CB],
                                        int skipAbove = 1 << 30)
{
    typedef TileShape<RT, CT> S;
    constexpr int kBK = BK;
    constexpr int NA = RT * kBK / 256, NBL = CT * kBK / 256;  // loads per thread per stage
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave / S::WCW, wc = wave % S::WCW;
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int a = 0; a < S::RB; ++a)
#pragma unroll
        for (int b = 0; b < S::CB; ++b) acc[a][b] = typename Mfma<T>::acc{0, 0, 0, 0};
    if (kend <= kbeg) return;
    decltype(la(0, kbeg)) ra[NA];
    decltype(lb(kbeg, 0)) rb[NBL];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < NA; ++q) {
            const int e = tid + 256 * q, i = e % RT, kk = e / RT;
            ra[q] = la(i, min(k0 + kk, kend - 1));
        }
#pragma unroll
        for (int q = 0; q < NBL; ++q) {
            const int e = tid + 256 * q;
            const int j = BJFAST ? e % CT : e / kBK, kk = BJFAST ? e / CT : e % kBK;
            rb[q] = lb(min(k0 + kk, kend - 1), j);
        }
    };
    auto put = [&](int buf, int k0) {
#pragma unroll
        for (int q = 0; q < NA; ++q) {
            const int e = tid + 256 * q, i = e % RT, kk = e / RT;
            As[buf][kk][i] = k0 + kk < kend ? fa(i, k0 + kk, ra[q]) : czero<T>();
        }
#pragma unroll
        for (int q = 0; q < NBL; ++q) {
            const int e = tid + 256 * q;
            const int j = BJFAST ? e % CT : e / kBK, kk = BJFAST ? e / CT : e % kBK;
            Bs[buf][kk][j] = k0 + kk < kend ? fb(k0 + kk, j, rb[q]) : czero<T>();
        }
    };
    const int nst = (kend - kbeg + kBK - 1) / kBK;
    fetch(kbeg);
    LdsBarrier();  // the caller's previous use of As / Bs is over
    put(0, kbeg);
    LdsBarrier();
    for (int st = 0; st < nst; ++st) {
        const int buf = st & 1;
        if (st + 1 < nst) fetch(kbeg + (st + 1) * kBK);
#pragma unroll
        for (int k4 = 0; k4 < kBK / 4; ++k4) {
            cx<T> av[S::RB];
#pragma unroll
            for (int a = 0; a < S::RB; ++a) av[a] = As[buf][4 * k4 + lk][16 * (wr * S::RB + a) + li];
#pragma unroll
            for (int b = 0; b < S::CB; ++b) {
                const int c0 = 8 * (wc * S::CB + b);  // the block's first column (wave-uniform)
                if (16 * (wr * S::RB + S::RB - 1) + 15 + skipAbove < c0) continue;
                const cx<T> v = Bs[buf][4 * k4 + lk][c0 + (li & 7)];
                const T b1 = li < 8 ? v.r : v.i;
                const T b2 = li < 8 ? -v.i : v.r;
#pragma unroll
                for (int a = 0; a < S::RB; ++a) {
                    if (16 * (wr * S::RB + a) + 15 + skipAbove < c0) continue;
                    acc[a][b] = Mfma<T>::step(av[a].r, b1, acc[a][b]);
                    acc[a][b] = Mfma<T>::step(av[a].i, b2, acc[a][b]);
                }
            }
        }
        if (st + 1 < nst) put(buf ^ 1, kbeg + (st + 1) * kBK);
        LdsBarrier();
    }
}

// Epilogue over the tile elements this lane finishes, one 16-row block at a
// time: ld(i, j) is called for every element of the block first (it must be
// safe for any i < RT, j < CT: the loads of the block are all in flight
// together), then st(i, j, value, loaded) for the elements of lanes with
// l % 16 < 8, after pairing the real and imaginary halves.
template <typename T, int RT, int CT = 64, class LD, class ST>
__device__ __forceinline__ void MmaEpilogue(const typename Mfma<T>::acc (&acc)[TileShape<RT, CT>::RB]
                                                                             [TileShape<RT, CT>::CB],
                                            const LD& ld, const ST& st)
{
    typedef TileShape<RT, CT> S;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave / S::WCW, wc = wave % S::WCW;
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int a = 0; a < S::RB; ++a) {
        decltype(ld(0, 0)) old[S::CB][4];
#pragma unroll
        for (int b = 0; b < S::CB; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                old[b][q] = ld(16 * (wr * S::RB + a) + Mfma<T>::row(lk, q), 8 * (wc * S::CB + b) + (li & 7));
#pragma unroll
        for (int b = 0; b < S::CB; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const T val = acc[a][b][q];
                const T other = __shfl_xor(val, 8, 64);
                if (li < 8)
                    st(16 * (wr * S::RB + a) + Mfma<T>::row(lk, q), 8 * (wc * S::CB + b) + li,
                       cx<T>{val, other}, old[b][q]);
            }
    }
}

// Calls f(i, j, value) for every tile element this lane finishes (lanes with
// l % 16 < 8 after pairing the real and imaginary halves).
template <typename T, int RT, int CT = 64, class F>
__device__ __forceinline__ void MmaForEach(const typename Mfma<T>::acc (&acc)[TileShape<RT, CT>::RB]
                                                                            [TileShape<RT, CT>::CB],
                                           const F& f)
{
    typedef TileShape<RT, CT> S;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave / S::WCW, wc = wave % S::WCW;
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int a = 0; a < S::RB; ++a)
#pragma unroll
        for (int b = 0; b < S::CB; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const T val = acc[a][b][q];
                const T other = __shfl_xor(val, 8, 64);
                if (li < 8)
                    f(16 * (wr * S::RB + a) + Mfma<T>::row(lk, q), 8 * (wc * S::CB + b) + li,
                      cx<T>{val, other});
            }
}

// Waves per SIMD the wide-block kernel is compiled for (its register budget:
// 512 / waves), per precision, phase and tile rows; 1 leaves the compiler
// free.  Each value is the most that compiles without spills.  Unbounded,
// FP32 64-row forward tiles took 293 registers (one workgroup per CU) and
// 16-row ones 140 (three).
template <typename T, int PH, int RT>
constexpr int MmwWavesPerEu()
{
    if (sizeof(T) == 4) return RT == 16 ? 4 : 2;
    if (RT == 16) return PH == PH_F ? 3 : 1;
    return RT == 64 && PH == PH_B ? 2 : 1;
}

// Wide-block front solves (k >= 32): one workgroup per solve tile and pass
// of 64 right-hand sides, through the tile engine.
template <typename T, int PH, int RT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MmwWavesPerEu<T, PH, RT>())))
void k_mmw(const SolveTile* tiles, int start,
                                              const int32_t* structs, Slots<T> gs,
                                              const cx<T>* L, const cx<T>* LT, cx<T>* Xp,
                                              cx<T>* W, cx<T>* P, int k)
{
    typedef TileShape<RT> S;
    constexpr int IX = 2048;  // inner indices whose gather addresses are cached in LDS
    __shared__ cx<T> As[2][kBK][RT + 1];
    __shared__ cx<T> Bs[2][kBK][65];
    __shared__ int32_t ix[PH == PH_F ? 1 : IX];
    const SolveTile tl = tiles[start + blockIdx.x];
    if (tl.nrows <= 0) return;  // a hole of the XCD placement
    const NodeDev nd = tl.nd;
    const int s = nd.s, fs = nd.s + nd.u;
    const cx<T>* Lf = L + nd.l;
    const cx<T>* Lt = LT + nd.l;
    // B: the panel rows of the update columns are read once into LDS, so each
    // K stage waits for one memory round trip instead of two (F needs no
    // indices: its gathered input sits in the slot copies).
    const int ncols = tl.col1 - tl.col0;
    const bool cached = PH == PH_B && ncols <= IX;
    if (cached) {
        for (int c = threadIdx.x; c < ncols; c += 256) {
            const int col = tl.col0 + c;
            ix[c] = col >= s ? structs[nd.ustart + col - s] : col;
        }
    }
    LdsBarrier();  // the cache is read by every thread's stage loads
    const int ilast = tl.nrows - 1;
    // A: the factor block, rows clamped into the tile, triangle masked later.
    auto la = [&](int i, int col) {
        const int row = tl.row0 + min(i, ilast);
        return PH == PH_F ? Lf[row + (int64_t)col * fs] : Lt[row + (int64_t)col * s];
    };
    auto fa = [&](int i, int col, cx<T> v) {
        const int row = tl.row0 + i;
        const bool in = i < tl.nrows && (PH == PH_F ? col <= row : col >= row);
        return in ? v : czero<T>();
    };
    for (int r0 = 0; r0 < k; r0 += 64) {
        const int kc = min(64, k - r0);
        typename Mfma<T>::acc acc[S::RB][S::CB];
        if constexpr (PH == PH_F) {
            // B: the gathered forward input (slot copies plus the panel rows).
            auto lb = [&](int col, int j) {
                const int rr = r0 + min(j, kc - 1);
                const int64_t g = (nd.w + col) * k + rr;
                return Raw3<T>{gs.G[g], gs.G[gs.half + g], Xp[(int64_t)(nd.off + col) * k + rr]};
            };
            auto fb = [&](int, int j, const Raw3<T>& v) { return j < kc ? v.a + v.b + v.c : czero<T>(); };
            MmaTile<T, RT, true>(la, fa, lb, fb, tl.col0, tl.col1, As, Bs, acc);
        } else {
            auto fb = [&](int, int j, cx<T> v) { return j < kc ? v : czero<T>(); };
            if (cached) {
                auto lb = [&](int col, int j) {
                    const int rr = r0 + min(j, kc - 1);
                    // ix holds the W row (top columns, plus nd.w) or the
                    // panel row: the read feeds both arms, so no branch
                    const int64_t row = (int64_t)ix[col - tl.col0] + (col < s ? nd.w : 0);
                    return (col < s ? W : Xp)[row * k + rr];
                };
                MmaTile<T, RT, true>(la, fa, lb, fb, tl.col0, tl.col1, As, Bs, acc);
            } else {
                auto lb = [&](int col, int j) {
                    return StageOperand<T, PH>(nd, structs, gs, Xp, W, col, r0 + min(j, kc - 1), k);
                };
                MmaTile<T, RT, true>(la, fa, lb, fb, tl.col0, tl.col1, As, Bs, acc);
            }
        }
        if (tl.slot >= 0) {
            MmaForEach<T, RT>(acc, [&](int i, int j, cx<T> a) {
                if (i < tl.nrows && j < kc) P[((int64_t)tl.slot * 64 + i) * k + r0 + j] = a;
            });
        } else if constexpr (PH == PH_F) {
            // Bottom rows add their gathered value; its loads (and the parent
            // row of each) are issued for a whole row block before any store.
            auto ld = [&](int i, int j) {
                const int row = tl.row0 + min(i, ilast);
                const int64_t g = (nd.w + row) * k + r0 + min(j, kc - 1);
                return RawRow<T>{gs.G[g], gs.G[gs.half + g],
                                 gs.relinds[nd.ustart + max(row - s, 0)]};
            };
            MmaEpilogue<T, RT>(acc, ld, [&](int i, int j, cx<T> a, const RawRow<T>& g) {
                if (i < tl.nrows && j < kc)
                    Finish<T, PH>(nd, gs, Xp, W, tl.row0 + i, r0 + j, k, a, g.a + g.b, g.prow);
            });
        } else {
            MmaForEach<T, RT>(acc, [&](int i, int j, cx<T> a) {
                if (i < tl.nrows && j < kc) Finish<T, PH>(nd, gs, Xp, W, tl.row0 + i, r0 + j, k, a, czero<T>(), 0);
            });
        }
    }
}

// Batched tile GEMM on the matrix cores: C (+)= +/- opA(A) diag(D) opB(B)
// over a list of 64x64 tiles (see GemmFlags), through the tile engine.
// 8-column LDS stages keep the registers low enough for 5 waves per SIMD in
// FP32 and 3 in FP64 without spills; on the factorization's many short-K
// Schur updates that occupancy is worth more than deeper stages.
template <typename T>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? 5 : 3) void k_gemm_mma(const GemmDesc* descs, const GemmTile* tiles,
                                                  int tileStart, cx<T>* b0, int64_t st0,
                                                  cx<T>* b1, int64_t st1, cx<T>* b2,
                                                  int64_t st2, cx<T>* b3)
{
    constexpr int BK = 8;  // inner columns per stage
    constexpr int TS = 64;   // 128x128 tiles measured slower on the Schur updates
    __shared__ cx<T> As[2][BK][TS + 1];
    __shared__ cx<T> Bs[2][BK][TS + 1];
    const GemmTile tl = tiles[tileStart + blockIdx.x];
    const GemmDesc g = descs[tl.desc];
    const int bt = blockIdx.y;
    cx<T>* bases[4] = {b0 + (int64_t)bt * st0, b1 + (int64_t)bt * st1, b2 + (int64_t)bt * st2,
                       b3 + (int64_t)bt * st0};  // b3 (LT) has L's stride
    const cx<T>* A = bases[g.abuf] + g.a;
    const cx<T>* B = bases[g.bbuf] + g.b;
    cx<T>* Cm = bases[g.cbuf] + g.c;
    const cx<T>* D = bases[g.dbuf] + g.d;
    const int tm = tl.tm, tn = tl.tn, f = g.flags;
    int kbeg = 0, kend = g.K;
    if (f & G_BUNIT) kbeg = tn;
    if (f & G_AUNIT) kend = min(kend, tm + TS);
    // Every loader reads from a clamped index inside the operand's rectangle
    // (always allocated), so a stage's loads go out without branches; the
    // masks, the unit diagonals and the diag(D) scaling are applied when the
    // stage is written to LDS.  Without G_HASD the D load reads A's first
    // element (a broadcast) and is ignored.
    const bool hasd = f & G_HASD;
    const cx<T>* Dv = hasd ? D : A;
    const int64_t dinc = hasd ? g.dinc : 0;
    auto la = [&](int i, int k) {
        return Raw2<T>{A[min(tm + i, g.M - 1) + (int64_t)k * g.lda], Dv[(int64_t)k * dinc]};
    };
    auto fa = [&](int i, int k, const Raw2<T>& r) {
        const int gi = tm + i;
        cx<T> v = r.a;
        if (f & G_AUNIT) v = gi > k ? v : (gi == k ? cx<T>{T(1), T(0)} : czero<T>());
        if (hasd) v = v * r.b;
        return gi < g.M ? v : czero<T>();
    };
    typename Mfma<T>::acc acc[TileShape<TS, TS>::RB][TileShape<TS, TS>::CB];
    // lower-triangular results: blocks above the diagonal are not computed
    // (only tiles that cross the diagonal have any)
    const int skip = (f & G_CLOWER) ? tm - tn : 1 << 30;
    if (f & G_BTRANS) {
        auto lb = [&](int k, int j) { return B[min(tn + j, g.N - 1) + (int64_t)k * g.ldb]; };
        auto fb = [&](int, int j, cx<T> v) { return tn + j < g.N ? v : czero<T>(); };
        MmaTile<T, TS, true, BK, TS>(la, fa, lb, fb, kbeg, kend, As, Bs, acc, skip);
    } else {
        auto lb = [&](int k, int j) { return B[k + (int64_t)min(tn + j, g.N - 1) * g.ldb]; };
        auto fb = [&](int k, int j, cx<T> v) {
            const int gj = tn + j;
            if (f & G_BUNIT) v = k > gj ? v : (k == gj ? cx<T>{T(1), T(0)} : czero<T>());
            return gj < g.N ? v : czero<T>();
        };
        MmaTile<T, TS, false, BK, TS>(la, fa, lb, fb, kbeg, kend, As, Bs, acc, skip);
    }
    auto offset = [&](int gi, int gj) {
        return (f & G_CTRANS) ? gj + (int64_t)gi * g.ldc : gi + (int64_t)gj * g.ldc;
    };
    auto keep = [&](int gi, int gj) {
        return gi < g.M && gj < g.N && !((f & G_CLOWER) && gi < gj);
    };
    if (f & G_BETA0) {
        MmaForEach<T, TS, TS>(acc, [&](int i, int j, cx<T> v) {
            const int gi = tm + i, gj = tn + j;
            if (keep(gi, gj)) Cm[offset(gi, gj)] = (f & G_NEG) ? -v : v;
        });
    } else {
        // C is read for a whole row block before the first store goes out.
        auto ld = [&](int i, int j) { return Cm[offset(min(tm + i, g.M - 1), min(tn + j, g.N - 1))]; };
        MmaEpilogue<T, TS, TS>(acc, ld, [&](int i, int j, cx<T> v, cx<T> c) {
            const int gi = tm + i, gj = tn + j;
            if (keep(gi, gj)) Cm[offset(gi, gj)] = c + ((f & G_NEG) ? -v : v);
        });
    }
}

// Sum of the inner slices of split row blocks, in slot order, then the
// phase's finish.  One workgroup per row block.
template <typename T, int PH>
__device__ __forceinline__ void MvReduceItem(const SolveTile& it, const Slots<T>& gs, cx<T>* Xp,
                                             cx<T>* W, const cx<T>* P, int k)
{
    const NodeDev nd = it.nd;
    for (int e = threadIdx.x; e < it.nrows * k; e += kThreads) {
        const int rl = e / k, r = e % k;
        cx<T> a = czero<T>();
        for (int j = 0; j < it.col1; ++j) a = a + P[((int64_t)(it.col0 + j) * 64 + rl) * k + r];
        FinishGather<T, PH>(nd, gs, Xp, W, it.row0 + rl, r, k, a);
    }
}

template <typename T, int PH>
__global__ __launch_bounds__(kThreads) void k_mv_reduce(const SolveTile* items, int start,
                                                        Slots<T> gs, cx<T>* Xp,
                                                        cx<T>* W, const cx<T>* P, int k)
{
    MvReduceItem<T, PH>(items[start + blockIdx.x], gs, Xp, W, P, k);
}

// --------------------------------------------------------------------------
// Batched panel solves: one launch runs the same step (a tree level's
// products or reduce, a panel copy) of the sweeps of many independent
// problems.  Every workgroup reads its BatchBlock: the tile or reduce item it
// works on (tiles of the problems' own lists, unchanged), the segment (one
// problem's panel solve: buffers, factors of that panel, k) and the problem,
// whose flag in `active` is zero once its solve has finished (its
// workgroups then exit at once).  The work of a workgroup is that of the
// single-problem kernels, with the same summation order.
// --------------------------------------------------------------------------
template <typename T, int PH, int RT, int KB, int NWV>
__global__ __launch_bounds__(64 * NWV) PSP_MVM_OCCUPANCY void k_mvm_batch(
    const BatchBlock* blocks, const BatchSolveSeg* segs, const int32_t* active)
{
    const BatchBlock bb = blocks[blockIdx.x];
    const SolveTile tl = *bb.item;
    const BatchSolveSeg sg = segs[bb.seg];
    if (!active[bb.problem] || tl.nrows <= 0) return;
    const Slots<T> gs{(cx<T>*)sg.G, sg.half, sg.relinds};
    MvmTile<T, PH, RT, KB, NWV>(tl, sg.structs, gs, (const cx<T>*)sg.L, (const cx<T>*)sg.LT,
                                (cx<T>*)sg.Xp, (cx<T>*)sg.W, (cx<T>*)sg.P, sg.k, nullptr,
                                nullptr);
}

template <typename T, int PH>
__global__ __launch_bounds__(kThreads) void k_mv_reduce_batch(const BatchBlock* blocks,
                                                              const BatchSolveSeg* segs,
                                                              const int32_t* active)
{
    const BatchBlock bb = blocks[blockIdx.x];
    const SolveTile it = *bb.item;
    const BatchSolveSeg sg = segs[bb.seg];
    if (!active[bb.problem]) return;
    const Slots<T> gs{(cx<T>*)sg.G, sg.half, sg.relinds};
    MvReduceItem<T, PH>(it, gs, (cx<T>*)sg.Xp, (cx<T>*)sg.W, (const cx<T>*)sg.P, sg.k);
}

// LT[:, :s] = Linv^T D^-1 per front (upper triangle; D^-1 on the diagonal),
// through a 32x32 LDS tile of L_TL (strictly lower part: Linv; diagonal: D).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_transpose_scale(const TransItem* items,
                                                               const NodeDev* nodes,
                                                               cx<T>* L, cx<T>* LT,
                                                               int64_t fstore)
{
    const TransItem it = items[blockIdx.x];
    const NodeDev nd = nodes[it.node];
    const int s = nd.s, fs = nd.s + nd.u;
    cx<T>* Lf = L + (int64_t)blockIdx.y * fstore + nd.l;
    cx<T>* Lt = LT + (int64_t)blockIdx.y * fstore + nd.l;
    __shared__ cx<T> tile[32][33];
    __shared__ cx<T> dinv[32];
    const int t32 = threadIdx.x & 31, q = threadIdx.x >> 5;
    if (threadIdx.x < 32) {
        const int i = it.i0 + t32;
        dinv[t32] = i < s ? cinv(Lf[i + (int64_t)i * fs]) : czero<T>();
    }
    for (int jj = q; jj < 32; jj += 8) {
        const int i = it.i0 + t32, j = it.j0 + jj;
        if (i < s && j < s) tile[jj][t32] = Lf[i + (int64_t)j * fs];
    }
    __syncthreads();
    // LT's block, with zeros below its diagonal; L's block is read by now,
    // and its part above the diagonal (which no solve uses) becomes zero
    for (int il = q; il < 32; il += 8) {
        const int j = it.j0 + t32, i = it.i0 + il;
        if (i < s && j < s)
            Lt[j + (int64_t)i * s] =
                i > j ? tile[t32][il] * dinv[il] : (i == j ? dinv[il] : czero<T>());
    }
    for (int jj = q; jj < 32; jj += 8) {
        const int i = it.i0 + t32, j = it.j0 + jj;
        if (i < s && j < s && j > i) Lf[i + (int64_t)j * fs] = czero<T>();
    }
}

// L_BL = (LT[:, s:])^T, i.e. -L_BL Linv back in column-major L, and the unit
// diagonal of Linv stored in L (items with i0 == 0 of each column block).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_fold_back(const TransItem* items,
                                                         const NodeDev* nodes, cx<T>* L,
                                                         const cx<T>* LT, int64_t fstore)
{
    const TransItem it = items[blockIdx.x];
    const NodeDev nd = nodes[it.node];
    const int s = nd.s, u = nd.u, fs = nd.s + nd.u;
    cx<T>* Lf = L + (int64_t)blockIdx.y * fstore + nd.l;
    const cx<T>* Lt = LT + (int64_t)blockIdx.y * fstore + nd.l;
    __shared__ cx<T> tile[32][33];
    const int t32 = threadIdx.x & 31, q = threadIdx.x >> 5;
    for (int il = q; il < 32; il += 8) {
        const int i = it.i0 + il, j = it.j0 + t32;  // LT(j, s + i)
        if (i < u && j < s) tile[il][t32] = Lt[j + (int64_t)(s + i) * s];
    }
    __syncthreads();
    for (int jj = q; jj < 32; jj += 8) {
        const int i = it.i0 + t32, j = it.j0 + jj;  // L(s + i, j)
        if (i < u && j < s) Lf[s + i + (int64_t)j * fs] = tile[t32][jj];
    }
    if (it.i0 == 0 && threadIdx.x < 32) {
        const int j = it.j0 + threadIdx.x;
        if (j < s) Lf[j + (int64_t)j * fs] = cx<T>{T(1), T(0)};
    }
}

// --------------------------------------------------------------------------
// Natural-order vector <-> padded panel vector (SolvePanel's copies,
// solve.hpp:756-778 and :825-847).
// --------------------------------------------------------------------------
// Panel copies, element e = (panel row q, right-hand side r) of a panel
// vector; the kernels below run them on one panel (k_copy_in, k_copy_out,
// k_zero_set_plane) or on panels of many problems (k_copy_batch).
template <typename T>
__device__ __forceinline__ void CopyInElem(int64_t e, int n, int k, const int32_t* qxy,
                                           const int32_t* qv, const c128* B, cx<T>* Xp, int nx,
                                           int ny, int nz, int vOffset, int padding)
{
    if (e >= (int64_t)n * k) return;
    const int q = (int)(e / k), r = (int)(e - (int64_t)q * k);
    const int vl = qv[q];
    if (vl < padding) {
        Xp[e] = czero<T>();
        return;
    }
    const int x = qxy[q] & 0xffff, y = qxy[q] >> 16;
    const int z = nz - 1 - (vOffset + vl);
    const int64_t nat = x + (int64_t)y * nx + (int64_t)z * nx * ny;
    Xp[e] = cvt<T>(B[nat * k + r]);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_copy_in(int n, int k, const int32_t* qxy,
                                                       const int32_t* qv, const c128* B,
                                                       cx<T>* Xp, int nx, int ny, int nz,
                                                       int vOffset, int padding,
                                                       int32_t* sync, int nsync)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // the fused-launch state of the panel solve that follows starts at zero
    for (int64_t j = e; j < nsync; j += (int64_t)gridDim.x * blockDim.x) sync[j] = 0;
    CopyInElem<T>(e, n, k, qxy, qv, B, Xp, nx, ny, nz, vOffset, padding);
}

template <typename T>
__device__ __forceinline__ void CopyOutElem(int64_t e, int n, int k, const int32_t* qxy,
                                            const int32_t* qv, c128* B, const cx<T>* Xp, int nx,
                                            int ny, int nz, int vOffset, int padding, int mode,
                                            const StencilArgs& st, const double* vel, int v1);

template <typename T>
__global__ __launch_bounds__(kThreads) void k_copy_out(int n, int k, const int32_t* qxy,
                                                        const int32_t* qv, c128* B,
                                                        const cx<T>* Xp, int nx, int ny, int nz,
                                                        int vOffset, int padding, int mode,
                                                        StencilArgs st, const double* vel, int v1)
{
    CopyOutElem<T>((int64_t)blockIdx.x * blockDim.x + threadIdx.x, n, k, qxy, qv, B, Xp, nx, ny,
                   nz, vOffset, padding, mode, st, vel, v1);
}

template <typename T>
__device__ __forceinline__ void CopyOutElem(int64_t e, int n, int k, const int32_t* qxy,
                                            const int32_t* qv, c128* B, const cx<T>* Xp, int nx,
                                            int ny, int nz, int vOffset, int padding, int mode,
                                            const StencilArgs& st, const double* vel, int v1)
{
    if (e >= (int64_t)n * k) return;
    const int q = (int)(e / k), r = (int)(e - (int64_t)q * k);
    const int vl = qv[q];
    if (vl < padding) return;
    const int x = qxy[q] & 0xffff, y = qxy[q] >> 16;
    const int z = nz - 1 - (vOffset + vl);
    const int64_t nat = x + (int64_t)y * nx + (int64_t)z * nx * ny;
    const c128 v = cvt<double>(Xp[e]);
    if (mode == 0)
        B[nat * k + r] = v;
    else
        B[nat * k + r] = B[nat * k + r] + v;
    // forward sweep: the subdiagonal coupling into the next panel's first
    // plane v1 (k_subdiag's update), applied by the thread that wrote the
    // value it reads: row (x,y,v1) couples to (x,y,v1-1) with value c[5]
    if (v1 >= 0 && vOffset + vl == v1 - 1) {
        const int64_t natN = nat - (int64_t)nx * ny;
        c128 c[7];
        GlobalRowFast(st, vel[natN], x, y, v1, c);
        B[natN * k + r] = B[natN * k + r] - c[5] * v;
    }
}

// --------------------------------------------------------------------------
// Sweep couplings between consecutive panels (solve.hpp:884-1047).
// --------------------------------------------------------------------------
// Backward sweep (ExtractPanel + MultiplySuperdiagonal): the panel vector is
// zero except on its last local plane vlLast (global plane v2), which holds
// -A(v2, v2+1) B(plane v2+1), i.e. row (x,y,v2) times its coupling c[6] to
// (x,y,v2+1) = natural z-1; one pass over the panel vector.  The forward
// sweep's subdiagonal coupling is applied by k_copy_out.
template <typename T>
__device__ __forceinline__ void ZeroSetPlaneElem(int64_t e, int n, int k, const int32_t* qxy,
                                                 const int32_t* qv, const StencilArgs& st,
                                                 const double* vel, const c128* B, cx<T>* Xp,
                                                 int v2, int vlLast);

template <typename T>
__global__ __launch_bounds__(kThreads) void k_zero_set_plane(int n, int k, const int32_t* qxy,
                                                              const int32_t* qv, StencilArgs st,
                                                              const double* vel, const c128* B,
                                                              cx<T>* Xp, int v2, int vlLast,
                                                              int32_t* sync, int nsync)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t j = e; j < nsync; j += (int64_t)gridDim.x * blockDim.x) sync[j] = 0;
    ZeroSetPlaneElem<T>(e, n, k, qxy, qv, st, vel, B, Xp, v2, vlLast);
}

// One batched panel copy: op BATCH_COPY_IN, BATCH_COPY_OUT (mode and
// coupling plane v1 per segment) or BATCH_ZERO_SET_PLANE, with the segment's
// elements split into whole workgroups (BatchBlock.item unused, blk: the
// workgroup's index inside its segment).
template <typename T>
__global__ __launch_bounds__(kBatchCopyThreads) void k_copy_batch(int op, const BatchBlock* blocks,
                                                                  const BatchCopySeg* segs,
                                                                  const int32_t* active)
{
    const BatchBlock bb = blocks[blockIdx.x];
    const BatchCopySeg sg = segs[bb.seg];
    if (!active[bb.problem]) return;
    const int64_t e = (int64_t)bb.blk * kBatchCopyThreads + threadIdx.x;
    if (op == BATCH_COPY_IN)
        CopyInElem<T>(e, sg.n, sg.k, sg.qxy, sg.qv, sg.B, (cx<T>*)sg.Xp, sg.st.nx, sg.st.ny,
                      sg.st.nz, sg.vOffset, sg.padding);
    else if (op == BATCH_COPY_OUT)
        CopyOutElem<T>(e, sg.n, sg.k, sg.qxy, sg.qv, sg.B, (const cx<T>*)sg.Xp, sg.st.nx,
                       sg.st.ny, sg.st.nz, sg.vOffset, sg.padding, sg.mode, sg.st, sg.vel, sg.v1);
    else
        ZeroSetPlaneElem<T>(e, sg.n, sg.k, sg.qxy, sg.qv, sg.st, sg.vel, sg.B, (cx<T>*)sg.Xp,
                            sg.v2, sg.vlLast);
}

template <typename T>
__device__ __forceinline__ void ZeroSetPlaneElem(int64_t e, int n, int k, const int32_t* qxy,
                                                 const int32_t* qv, const StencilArgs& st,
                                                 const double* vel, const c128* B, cx<T>* Xp,
                                                 int v2, int vlLast)
{
    if (e >= (int64_t)n * k) return;
    const int q = (int)(e / k), r = (int)(e - (int64_t)q * k);
    cx<T> out = czero<T>();
    if (qv[q] == vlLast) {
        const int x = qxy[q] & 0xffff, y = qxy[q] >> 16;
        const int64_t plane = (int64_t)st.nx * st.ny;
        const int64_t nat = x + (int64_t)y * st.nx + (int64_t)(st.nz - 1 - v2) * plane;
        c128 c[7];
        GlobalRowFast(st, vel[nat], x, y, v2, c);
        out = cvt<T>(-(c[6] * B[(nat - plane) * k + r]));
    }
    Xp[e] = out;
}

// --------------------------------------------------------------------------
// Multiply (solve.hpp:613-679) in natural order: y = A x or y = b - A x.
// --------------------------------------------------------------------------
// I: index type, uint32_t when every element index (and the grid's last
// thread) fits, which keeps the index divisions 32-bit; int64_t otherwise.
template <typename I>
__global__ __launch_bounds__(kThreads) void k_spmv(StencilArgs st, const double* vel,
                                                    const c128* X, const c128* Bv, c128* Y, int k,
                                                    int mode)
{
    const I e = (I)blockIdx.x * (I)blockDim.x + (I)threadIdx.x;
    const I nx = (I)st.nx, ny = (I)st.ny, kk = (I)k;
    const I plane = nx * ny;
    if (e >= plane * (I)st.nz * kk) return;
    const I p = e / kk;
    const I py = p / nx;
    const int x = (int)(p - py * nx);
    const int y = (int)(py % ny);
    const int z = (int)(py / ny);
    const int v = st.nz - 1 - z;
    c128 c[7];
    GlobalRowFast(st, vel[p], x, y, v, c);
    c128 acc = X[e] * c[0];
    if (x > 0) acc = cfma(c[1], X[e - kk], acc);
    if (x + 1 < st.nx) acc = cfma(c[2], X[e + kk], acc);
    if (y > 0) acc = cfma(c[3], X[e - nx * kk], acc);
    if (y + 1 < st.ny) acc = cfma(c[4], X[e + nx * kk], acc);
    if (v > 0) acc = cfma(c[5], X[e + plane * kk], acc);
    if (v + 1 < st.nz) acc = cfma(c[6], X[e - plane * kk], acc);
    Y[e] = mode == 0 ? acc : (Bv[e] - acc);
}

// dL/dm_i = sum_s Re(omega^2 s1 s2 s3 lambda_s(i) u_s(i)).
__global__ __launch_bounds__(kThreads) void k_adjoint_grad(StencilArgs st, const c128* U,
                                                            const c128* Lam, int k, double* grad,
                                                            int accumulate)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t plane = (int64_t)st.nx * st.ny;
    if (p >= plane * st.nz) return;
    const int x = (int)(p % st.nx);
    const int y = (int)((p / st.nx) % st.ny);
    const int z = (int)(p / plane);
    const int v = st.nz - 1 - z;
    const c128 s = st.s1[x + 1] * st.s2[y + 1] * st.s3[v + 1];
    double g = 0.0;
    for (int r = 0; r < k; ++r) {
        const c128 t = Lam[p * k + r] * U[p * k + r];
        const c128 w = s * t;
        g += w.r;
    }
    g *= st.omega * st.omega;
    grad[p] = accumulate ? grad[p] + g : g;
}

// --------------------------------------------------------------------------
// Column reductions (Norms / InnerProducts, environment.hpp:177-211).
// Each workgroup reduces a contiguous point range; thread t < lanes owns
// column t % k; partials are summed in a fixed order (deterministic).
// --------------------------------------------------------------------------
constexpr int kReduceBlocks = 512;

__global__ __launch_bounds__(kThreads) void k_colnorm_part(const c128* X, int64_t np, int k,
                                                            double* part)
{
    __shared__ double sh[kThreads];
    const int lanes = (kThreads / k) * k, per = kThreads / k;
    const int t = threadIdx.x;
    const int64_t chunk = (np + gridDim.x - 1) / gridDim.x;
    const int64_t p0 = (int64_t)blockIdx.x * chunk, p1 = min(np, p0 + chunk);
    double acc = 0.0;
    if (t < lanes) {
        const int r = t % k;
        for (int64_t p = p0 + t / k; p < p1; p += per) {
            const c128 v = X[p * k + r];
            acc = fma(v.r, v.r, acc);
            acc = fma(v.i, v.i, acc);
        }
    }
    sh[t] = acc;
    __syncthreads();
    if (t < k) {
        double s = 0.0;
        for (int i = 0; i < per; ++i) s += sh[t + i * k];
        part[(int64_t)blockIdx.x * k + t] = s;
    }
}

// One workgroup per column: fixed-order tree sum of the block partials.
__global__ __launch_bounds__(kThreads) void k_colnorm_final(const double* part, int nblk, int k,
                                                             double* out)
{
    __shared__ double sh[kThreads];
    const int r = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kThreads) s += part[(int64_t)b * k + r];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[r] = sqrt(sh[0]);
}

__global__ __launch_bounds__(kThreads) void k_coldot_part(const c128* __restrict__ V,
                                                           const c128* __restrict__ W,
                                                           int64_t np, int k,
                                                           c128* __restrict__ part)
{
    __shared__ c128 sh[kThreads];
    const int lanes = (kThreads / k) * k, per = kThreads / k;
    const int t = threadIdx.x;
    const int64_t chunk = (np + gridDim.x - 1) / gridDim.x;
    const int64_t p0 = (int64_t)blockIdx.x * chunk, p1 = min(np, p0 + chunk);
    c128 acc = {0.0, 0.0};
    if (t < lanes) {
        const int r = t % k;
#pragma unroll 4
        for (int64_t p = p0 + t / k; p < p1; p += per) {
            const c128 v = V[p * k + r], w = W[p * k + r];
            // conj(v) * w
            acc.r = fma(v.r, w.r, acc.r);
            acc.r = fma(v.i, w.i, acc.r);
            acc.i = fma(v.r, w.i, acc.i);
            acc.i = fma(-v.i, w.r, acc.i);
        }
    }
    sh[t] = acc;
    __syncthreads();
    if (t < k) {
        c128 s = {0.0, 0.0};
        for (int i = 0; i < per; ++i) s = s + sh[t + i * k];
        part[(int64_t)blockIdx.x * k + t] = s;
    }
}

// Modified Gram-Schmidt, two steps in one pass: W -= alpha V (step i's
// update) and the partial sums of conj(Vn) W for step i + 1, with the
// block partition and summation order of k_coldot_part.
__global__ __launch_bounds__(kThreads) void k_axpy_dot_part(const c128* __restrict__ alpha,
                                                             const c128* __restrict__ V,
                                                             const c128* __restrict__ Vn,
                                                             c128* __restrict__ W, int64_t np,
                                                             int k, c128* __restrict__ part)
{
    __shared__ c128 sh[kThreads];
    const int lanes = (kThreads / k) * k, per = kThreads / k;
    const int t = threadIdx.x;
    const int64_t chunk = (np + gridDim.x - 1) / gridDim.x;
    const int64_t p0 = (int64_t)blockIdx.x * chunk, p1 = min(np, p0 + chunk);
    c128 acc = {0.0, 0.0};
    if (t < lanes) {
        const int r = t % k;
        const c128 a = alpha[r];
        // unrolled so a thread has several elements' loads in flight (the
        // sums still run in element order)
#pragma unroll 4
        for (int64_t p = p0 + t / k; p < p1; p += per) {
            const int64_t e = p * k + r;
            const c128 w = cfms(a, V[e], W[e]), v = Vn[e];
            W[e] = w;
            acc.r = fma(v.r, w.r, acc.r);
            acc.r = fma(v.i, w.i, acc.r);
            acc.i = fma(v.r, w.i, acc.i);
            acc.i = fma(-v.i, w.r, acc.i);
        }
    }
    sh[t] = acc;
    __syncthreads();
    if (t < k) {
        c128 s = {0.0, 0.0};
        for (int i = 0; i < per; ++i) s = s + sh[t + i * k];
        part[(int64_t)blockIdx.x * k + t] = s;
    }
}

__global__ __launch_bounds__(kThreads) void k_coldot_final(const c128* part, int nblk, int k,
                                                            c128* out)
{
    __shared__ c128 sh[kThreads];
    const int r = blockIdx.x;
    c128 s = {0.0, 0.0};
    for (int b = threadIdx.x; b < nblk; b += kThreads) s = s + part[(int64_t)b * k + r];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[r] = sh[0];
}

__global__ __launch_bounds__(kThreads) void k_axpy_cols(const c128* alpha, int sign,
                                                         const c128* V, c128* W, int64_t n, int k)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int r = (int)(e % k);
    const c128 a = alpha[r];
    W[e] = sign > 0 ? cfma(a, V[e], W[e]) : cfms(a, V[e], W[e]);
}

__global__ __launch_bounds__(kThreads) void k_scale_cols(const double* delta, const c128* W,
                                                          c128* V, int64_t n, int k)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int r = (int)(e % k);
    const double inv = 1.0 / delta[r];
    V[e] = W[e] * inv;
}

__global__ __launch_bounds__(kThreads) void k_update_x(const c128* X0, const c128* V,
                                                        const c128* Y, int j, int64_t n, int k,
                                                        int64_t vstride, c128* X)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int r = (int)(e % k);
    c128 acc = X0[e];
    for (int i = 0; i <= j; ++i) acc = cfma(Y[i * k + r], V[i * vstride + e], acc);
    X[e] = acc;
}

__global__ __launch_bounds__(kThreads) void k_neg(const c128* X, c128* W, int64_t n)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) W[e] = -X[e];
}

inline unsigned Blocks(int64_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

}  // namespace

// ==========================================================================
// Launch wrappers
// ==========================================================================
bool XcdRoundRobin(hipStream_t s)
{
    constexpr int kBlocks = 64;
    int32_t* d = nullptr;
    if (hipMalloc(&d, sizeof(int32_t) * kBlocks) != hipSuccess) return false;
    int32_t h[kBlocks] = {};
    bool ok = hipMemsetAsync(d, 0xff, sizeof(h), s) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(k_xcc_probe, dim3(kBlocks), dim3(64), 0, s, d);
        ok = hipGetLastError() == hipSuccess &&
             hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
    }
    (void)hipFree(d);
    for (int b = 0; ok && b < kBlocks; ++b) ok = h[b] == b % 8;
    return ok;
}

void LaunchFillZeroRows(void* p, int64_t bytes, int rows, int64_t strideBytes, hipStream_t s)
{
    if (bytes <= 0 || rows <= 0) return;
    const int64_t n8 = bytes / 8;
    const unsigned g = (unsigned)std::min<int64_t>(Blocks(n8), 4096);
    hipLaunchKernelGGL(k_fill_zero_rows, dim3(g, rows), dim3(kThreads), 0, s, (uint2*)p, n8,
                       strideBytes / 8);
}

void LaunchFillZero(void* p, int64_t bytes, hipStream_t s)
{
    if (bytes <= 0) return;
    const int64_t n16 = bytes / 16;
    if (n16 > 0) {
        const unsigned g = (unsigned)std::min<int64_t>(Blocks(n16), 65535 * 4);
        hipLaunchKernelGGL(k_fill_zero, dim3(g), dim3(kThreads), 0, s, (uint4*)p, n16);
    }
    const int64_t rem = bytes - n16 * 16;
    if (rem > 0)
        hipLaunchKernelGGL(k_fill_zero_bytes, dim3(1), dim3(kThreads), 0, s,
                           (unsigned char*)p + n16 * 16, rem);
}

template <typename T>
void LaunchScatterOrig(const ClassDev& c, const StencilArgs& st, const double* vel,
                       double damping, hipStream_t s)
{
    hipLaunchKernelGGL((k_scatter_orig<T>), dim3(Blocks(c.n), c.npanels), dim3(kThreads), 0, s,
                       c.n, c.vsize, c.qxy, c.qv, c.qnode, c.qcol, c.qrows, c.nodeL, c.nodeS,
                       c.nodeU, c.panelVOff, c.s3art, st, vel, damping, (cx<T>*)c.L, c.fstore);
}

template <typename T>
void LaunchExtendAdd(const ClassDev& c, int start, int count, hipStream_t s)
{
    if (count <= 0) return;
    hipLaunchKernelGGL((k_extend_add<T>), dim3(count, c.npanels), dim3(kThreads), 0, s,
                       c.extadd, start, c.relinds, (cx<T>*)c.L, c.fstore, (cx<T>*)c.U, c.ustore);
}

template <typename T>
void LaunchLdlDiag(const ClassDev& c, int start, int count, int nb, hipStream_t s)
{
    if (count <= 0) return;
    (void)nb;
    hipLaunchKernelGGL((k_ldl_diag<T, 32>), dim3(count, c.npanels), dim3(kThreads), 0, s, c.ldl,
                       start, (cx<T>*)c.L, c.fstore);
}

template <typename T>
void LaunchLdlTrsm(const ClassDev& c, int start, int count, int nb, hipStream_t s)
{
    if (count <= 0) return;
    (void)nb;
    hipLaunchKernelGGL((k_ldl_trsm<T, 32>), dim3(count, c.npanels), dim3(kThreads), 0, s, c.ldl,
                       start, (cx<T>*)c.L, c.fstore);
}

template <typename T>
void LaunchGemm(const ClassDev& c, int tileStart, int tileCount, hipStream_t s)
{
    if (tileCount <= 0) return;
    hipLaunchKernelGGL((k_gemm_mma<T>), dim3(tileCount, c.npanels), dim3(256), 0, s, c.gemms,
                       c.tiles, tileStart, (cx<T>*)c.L, c.fstore, (cx<T>*)c.U, c.ustore,
                       (cx<T>*)c.T, c.tstore, (cx<T>*)c.LT);
}

template <typename T>
void LaunchInvDiag(const ClassDev& c, int nb, hipStream_t s)
{
    if (c.ninv <= 0) return;
    (void)nb;
    hipLaunchKernelGGL((k_inv_diag<T, 32>), dim3(c.ninv, c.npanels), dim3(kThreads), 0, s, c.inv,
                       (cx<T>*)c.L, c.fstore);
}

// Workgroups of a persistent solve launch: as many as fit on the device at
// once, more only when a workgroup would otherwise get over kMvpTiles tiles.
template <typename T, int PH, int RT, int KB>
int MvpGrid(int cus, int count, int cap0 = 0)
{
    static const int perCU = [] {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_mvp<T, PH, RT, KB>, 256, 0) !=
                hipSuccess ||
            b < 1)
            b = 1;
        return b;
    }();
    const int fit = cap0 > 0 ? cap0 : perCU * std::max(cus, 1);
    const int cap = std::max(fit, (count + kMvpTiles - 1) / kMvpTiles);
    return std::max(1, std::min(count, cap));
}

// Workgroups of a fused launch: as many as fit on the device at once.
template <typename T, int PH, int RT, int KB>
int MvqGrid(int cus)
{
    static const int perCU = [] {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_mvq<T, PH, RT, KB>, 256, 0) !=
                hipSuccess ||
            b < 1)
            b = 1;
        return b;
    }();
    return perCU * std::max(cus, 1);
}

void PrimeSolveOccupancy()
{
    auto prime = [](auto t) {
        typedef decltype(t) T;
        (void)MvpGrid<T, PH_F, 64, 8>(1, 1), (void)MvpGrid<T, PH_F, 32, 8>(1, 1);
        (void)MvpGrid<T, PH_F, 16, 8>(1, 1), (void)MvpGrid<T, PH_F, 64, 16>(1, 1);
        (void)MvpGrid<T, PH_F, 32, 16>(1, 1), (void)MvpGrid<T, PH_F, 16, 16>(1, 1);
        (void)MvpGrid<T, PH_B, 64, 8>(1, 1), (void)MvpGrid<T, PH_B, 32, 8>(1, 1);
        (void)MvpGrid<T, PH_B, 16, 8>(1, 1), (void)MvpGrid<T, PH_B, 64, 16>(1, 1);
        (void)MvpGrid<T, PH_B, 32, 16>(1, 1), (void)MvpGrid<T, PH_B, 16, 16>(1, 1);
        (void)MvqGrid<T, PH_F, 64, 8>(1), (void)MvqGrid<T, PH_F, 32, 8>(1);
        (void)MvqGrid<T, PH_F, 16, 8>(1), (void)MvqGrid<T, PH_F, 64, 16>(1);
        (void)MvqGrid<T, PH_F, 32, 16>(1), (void)MvqGrid<T, PH_F, 16, 16>(1);
        (void)MvqGrid<T, PH_B, 64, 8>(1), (void)MvqGrid<T, PH_B, 32, 8>(1);
        (void)MvqGrid<T, PH_B, 16, 8>(1), (void)MvqGrid<T, PH_B, 64, 16>(1);
        (void)MvqGrid<T, PH_B, 32, 16>(1), (void)MvqGrid<T, PH_B, 16, 16>(1);
    };
    prime(float());
    prime(double());
}

template <typename T>
void LaunchPanelSolve(const ClassDev& c, const ClassPlan& plan, int b, void* Xpv, void* Wv, void* Pv,
                      void* Gv, int k, hipStream_t s, MvProfile* prof, int onlyLevel,
                      int onlyPhase, int32_t* sync)
{
    const cx<T>* L = (const cx<T>*)c.L + (int64_t)b * c.fstore;
    const cx<T>* LT = (const cx<T>*)c.LT + (int64_t)b * c.fstore;
    cx<T>* Xp = (cx<T>*)Xpv;
    cx<T>* W = (cx<T>*)Wv;
    cx<T>* P = (cx<T>*)Pv;
    const Slots<T> gs{(cx<T>*)Gv, c.wstore * k, c.relinds};
    auto mv = [&](auto ph, const ClassPlan::SolveLevel& lv, const ClassPlan::Range& r) {
        if (r.count <= 0) return;
        constexpr int PH = decltype(ph)::value;
        bool reduced = false;  // split row blocks already summed by the product launch
        if (prof) (void)hipEventRecord(prof->a, s);
        if (k >= 32) {
            if (r.rows == 64)
                hipLaunchKernelGGL((k_mmw<T, PH, 64>), dim3(r.count), dim3(256), 0, s, c.mv,
                                   r.start, c.structs, gs, L, LT, Xp, W, P, k);
            else if (r.rows == 32)
                hipLaunchKernelGGL((k_mmw<T, PH, 32>), dim3(r.count), dim3(256), 0, s, c.mv,
                                   r.start, c.structs, gs, L, LT, Xp, W, P, k);
            else
                hipLaunchKernelGGL((k_mmw<T, PH, 16>), dim3(r.count), dim3(256), 0, s, c.mv,
                                   r.start, c.structs, gs, L, LT, Xp, W, P, k);
        } else if (k >= 8 && k <= 16 && r.persist && r.waves == 4) {
            auto launchP = [&](auto rt, auto kb) {
                constexpr int RT = decltype(rt)::value, KB = decltype(kb)::value;
                const int G = MvpGrid<T, PH, RT, KB>(c.cus, r.count, c.mvpGrid);
                hipLaunchKernelGGL((k_mvp<T, PH, RT, KB>), dim3(G), dim3(256), 0, s, c.mv,
                                   r.start, r.count, c.structs, gs, L, LT, Xp, W, P, k);
            };
            auto byRowsP = [&](auto kb) {
                if (r.rows == 64)
                    launchP(std::integral_constant<int, 64>(), kb);
                else if (r.rows == 32)
                    launchP(std::integral_constant<int, 32>(), kb);
                else
                    launchP(std::integral_constant<int, 16>(), kb);
            };
            if (k <= 8)
                byRowsP(std::integral_constant<int, 8>());
            else
                byRowsP(std::integral_constant<int, 16>());
        } else if (k >= 8) {
            // split row blocks summed in the launch (slices placed per XCD;
            // the partial sums are read through a buffer of at most 4 GiB)
            const bool pfits = (int64_t)sizeof(cx<T>) * k * 64 * plan.maxSlots < ((int64_t)1 << 32);
            int32_t* rc = r.xcd && c.redc && pfits ? c.redc : nullptr;
            reduced = rc != nullptr;
            auto launch = [&](auto rt, auto kb) {
                constexpr int RT = decltype(rt)::value, KB = decltype(kb)::value;
                if (r.waves == 16)
                    hipLaunchKernelGGL((k_mvm<T, PH, RT, KB, 16>), dim3(r.count), dim3(1024), 0,
                                       s, c.mv, r.start, c.structs, gs, L, LT, Xp, W, P, k, rc,
                                       c.mvred);
                else
                    hipLaunchKernelGGL((k_mvm<T, PH, RT, KB, 4>), dim3(r.count), dim3(256), 0, s,
                                       c.mv, r.start, c.structs, gs, L, LT, Xp, W, P, k, rc,
                                       c.mvred);
            };
            auto byRows = [&](auto kb) {
                if (r.rows == 64)
                    launch(std::integral_constant<int, 64>(), kb);
                else if (r.rows == 32)
                    launch(std::integral_constant<int, 32>(), kb);
                else
                    launch(std::integral_constant<int, 16>(), kb);
            };
            if (k <= 8)
                byRows(std::integral_constant<int, 8>());
            else
                byRows(std::integral_constant<int, 16>());
        } else if (r.rows == 64)
            hipLaunchKernelGGL((k_mv<T, PH, 64>), dim3(r.count), dim3(kMvThreads), 0, s, c.mv,
                               r.start, c.structs, gs, L, LT, Xp, W, P, k);
        else if (r.rows == 32)
            hipLaunchKernelGGL((k_mv<T, PH, 32>), dim3(r.count), dim3(kMvThreads), 0, s, c.mv,
                               r.start, c.structs, gs, L, LT, Xp, W, P, k);
        else
            hipLaunchKernelGGL((k_mv<T, PH, 16>), dim3(r.count), dim3(kMvThreads), 0, s, c.mv,
                               r.start, c.structs, gs, L, LT, Xp, W, P, k);
        if (prof) {
            (void)hipEventRecord(prof->b, s);
            (void)hipEventSynchronize(prof->b);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, prof->a, prof->b);
            prof->launches += 1;
            prof->seconds += ms * 1e-3;
            const double eb = sizeof(cx<T>);
            prof->bytes += eb * (lv.entries[PH] + lv.vecRows[PH] * k);
            prof->flops += 8.0 * lv.entries[PH] * k;
        }
        if (r.rcount > 0 && !reduced)
            hipLaunchKernelGGL((k_mv_reduce<T, PH>), dim3(r.rcount), dim3(kThreads), 0, s,
                               c.mvred, r.rstart, gs, Xp, W, P, k);
    };
    if (onlyLevel >= 0) {  // one (level, phase) launch, for tile-shape tuning
        const ClassPlan::SolveLevel& lv = plan.solveLevels[onlyLevel];
        if (onlyPhase == PH_F)
            mv(std::integral_constant<int, PH_F>(), lv, lv.f);
        else
            mv(std::integral_constant<int, PH_B>(), lv, lv.b);
        return;
    }
    // Fused subtree launches: the levels from the cut down as one launch
    // per phase (8 <= k <= 16, with the class's launch state to count on).
    // (the fused kernel reads its vectors through buffer descriptors: each
    // buffer must span less than 4 GiB)
    const int64_t eb_ = sizeof(cx<T>), span = (int64_t)1 << 32;
    const bool fits = eb_ * k * (2 * c.wstore + kSolvePad + 2) < span &&
                      eb_ * k * (int64_t)(c.n + kSolvePad + 2) < span &&
                      eb_ * k * 64 * (int64_t)plan.maxSlots < span;
    const bool fused = sync && c.nqueues > 0 && plan.qcut >= 0 && k >= 8 && k <= 16 && fits;
    auto fusedLaunch = [&](auto ph) {
        constexpr int PH = decltype(ph)::value;
        const int nt = c.nqtiles[PH];
        if (nt <= 0) return;
        int32_t* st = sync + (PH == PH_F ? 0 : QueueSyncInts(c.nnodes, c.nqueues, c.nqreds[0]));
        if (prof) (void)hipEventRecord(prof->a, s);
        auto go = [&](auto rt, auto kb) {
            constexpr int RT = decltype(rt)::value, KB = decltype(kb)::value;
            const int G = std::max(1, std::min(nt, MvqGrid<T, PH, RT, KB>(c.cus)));
            hipLaunchKernelGGL((k_mvq<T, PH, RT, KB>), dim3(G), dim3(256), 0, s, c.qtiles[PH],
                               c.qstart[PH], c.qreds[PH], c.qdep[PH], c.nodes, st, c.nnodes,
                               c.nqueues, c.structs, gs, L, LT, Xp, W, P, k);
        };
        auto byRows = [&](auto kb) {
            if (c.queueRows == 64)
                go(std::integral_constant<int, 64>(), kb);
            else if (c.queueRows == 16)
                go(std::integral_constant<int, 16>(), kb);
            else
                go(std::integral_constant<int, 32>(), kb);
        };
        if (k <= 8)
            byRows(std::integral_constant<int, 8>());
        else
            byRows(std::integral_constant<int, 16>());
        if (prof) {
            (void)hipEventRecord(prof->b, s);
            (void)hipEventSynchronize(prof->b);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, prof->a, prof->b);
            prof->launches += 1;
            prof->seconds += ms * 1e-3;
            const double eb = sizeof(cx<T>);
            const ClassPlan::QueuePhase& Q = plan.queue[PH];
            prof->bytes += eb * (Q.entries + Q.vecRows * k);
            prof->flops += 8.0 * Q.entries * k;
        }
    };
    const int lowest = fused ? plan.qcut : plan.maxDepth + 1;  // first level of the fused launch
    if (fused) fusedLaunch(std::integral_constant<int, PH_F>());
    for (int d = std::min(plan.maxDepth, lowest - 1); d >= 0; --d) {
        const ClassPlan::SolveLevel& lv = plan.solveLevels[d];
        mv(std::integral_constant<int, PH_F>(), lv, lv.f);
    }
    for (int d = 0; d < std::min(plan.maxDepth + 1, lowest); ++d) {
        const ClassPlan::SolveLevel& lv = plan.solveLevels[d];
        mv(std::integral_constant<int, PH_B>(), lv, lv.b);
    }
    if (fused) fusedLaunch(std::integral_constant<int, PH_B>());
}

template <typename T>
void LaunchTransposeScale(const ClassDev& c, hipStream_t s)
{
    if (c.ntrans <= 0) return;
    hipLaunchKernelGGL((k_transpose_scale<T>), dim3(c.ntrans, c.npanels), dim3(kThreads), 0, s,
                       c.trans, c.nodes, (cx<T>*)c.L, (cx<T>*)c.LT, c.fstore);
}

template <typename T>
void LaunchFoldBack(const ClassDev& c, hipStream_t s)
{
    if (c.nfold <= 0) return;
    hipLaunchKernelGGL((k_fold_back<T>), dim3(c.nfold, c.npanels), dim3(kThreads), 0, s,
                       c.foldBack, c.nodes, (cx<T>*)c.L, (const cx<T>*)c.LT, c.fstore);
}

template <typename T>
void LaunchCopyIn(const ClassDev& c, const c128* B, void* Xp, int k, int nx, int ny, int nz,
                  int vOffset, int padding, hipStream_t s, int32_t* sync)
{
    hipLaunchKernelGGL((k_copy_in<T>), dim3(Blocks((int64_t)c.n * k)), dim3(kThreads), 0, s, c.n,
                       k, c.qxy, c.qv, B, (cx<T>*)Xp, nx, ny, nz, vOffset, padding, sync,
                       sync ? c.nsync : 0);
}

template <typename T>
void LaunchCopyOut(const ClassDev& c, c128* B, const void* Xp, int k, int nx, int ny, int nz,
                   int vOffset, int padding, int mode, hipStream_t s, const StencilArgs* st,
                   const double* vel, int v1)
{
    const StencilArgs none{};
    hipLaunchKernelGGL((k_copy_out<T>), dim3(Blocks((int64_t)c.n * k)), dim3(kThreads), 0, s, c.n,
                       k, c.qxy, c.qv, B, (const cx<T>*)Xp, nx, ny, nz, vOffset, padding, mode,
                       st ? *st : none, vel, st ? v1 : -1);
}

template <typename T>
void LaunchZeroSetPlane(const ClassDev& c, const StencilArgs& st, const double* vel,
                        const c128* B, void* Xp, int k, int v2, int vlLast, hipStream_t s,
                        int32_t* sync)
{
    hipLaunchKernelGGL((k_zero_set_plane<T>), dim3(Blocks((int64_t)c.n * k)), dim3(kThreads), 0, s,
                       c.n, k, c.qxy, c.qv, st, vel, B, (cx<T>*)Xp, v2, vlLast, sync,
                       sync ? c.nsync : 0);
}

template <typename T>
void LaunchSolveBatch(int ph, int rows, int waves, int kb, const BatchBlock* blocks, int nblocks,
                      const BatchSolveSeg* segs, const int32_t* active, hipStream_t s)
{
    if (nblocks <= 0) return;
    auto go = [&](auto phc, auto rt) {
        constexpr int PH = decltype(phc)::value, RT = decltype(rt)::value;
        auto byKb = [&](auto kbc) {
            constexpr int KB = decltype(kbc)::value;
            if (waves == 16)
                hipLaunchKernelGGL((k_mvm_batch<T, PH, RT, KB, 16>), dim3(nblocks), dim3(1024), 0,
                                   s, blocks, segs, active);
            else
                hipLaunchKernelGGL((k_mvm_batch<T, PH, RT, KB, 4>), dim3(nblocks), dim3(256), 0, s,
                                   blocks, segs, active);
        };
        if (kb == 8)
            byKb(std::integral_constant<int, 8>());
        else
            byKb(std::integral_constant<int, 16>());
    };
    auto byRows = [&](auto phc) {
        if (rows == 64)
            go(phc, std::integral_constant<int, 64>());
        else if (rows == 32)
            go(phc, std::integral_constant<int, 32>());
        else
            go(phc, std::integral_constant<int, 16>());
    };
    if (ph == PH_F)
        byRows(std::integral_constant<int, PH_F>());
    else
        byRows(std::integral_constant<int, PH_B>());
}

template <typename T>
void LaunchReduceBatch(int ph, const BatchBlock* blocks, int nblocks, const BatchSolveSeg* segs,
                       const int32_t* active, hipStream_t s)
{
    if (nblocks <= 0) return;
    if (ph == PH_F)
        hipLaunchKernelGGL((k_mv_reduce_batch<T, PH_F>), dim3(nblocks), dim3(kThreads), 0, s,
                           blocks, segs, active);
    else
        hipLaunchKernelGGL((k_mv_reduce_batch<T, PH_B>), dim3(nblocks), dim3(kThreads), 0, s,
                           blocks, segs, active);
}

template <typename T>
void LaunchCopyBatch(int op, const BatchBlock* blocks, int nblocks, const BatchCopySeg* segs,
                     const int32_t* active, hipStream_t s)
{
    if (nblocks <= 0) return;
    hipLaunchKernelGGL((k_copy_batch<T>), dim3(nblocks), dim3(kBatchCopyThreads), 0, s, op,
                       blocks, segs, active);
}

void LaunchSpmv(const StencilArgs& st, const double* vel, const c128* x, const c128* b, c128* y,
                int k, int mode, hipStream_t s)
{
    const int64_t n = (int64_t)st.nx * st.ny * st.nz * k;
    if ((int64_t)Blocks(n) * kThreads <= (int64_t)UINT32_MAX)
        hipLaunchKernelGGL((k_spmv<uint32_t>), dim3(Blocks(n)), dim3(kThreads), 0, s, st, vel, x,
                           b, y, k, mode);
    else
        hipLaunchKernelGGL((k_spmv<int64_t>), dim3(Blocks(n)), dim3(kThreads), 0, s, st, vel, x, b,
                           y, k, mode);
}

int ReducePartials(int64_t npoints)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>(kReduceBlocks, (npoints + 255) / 256));
}

void LaunchColNorm2(const c128* x, int64_t np, int k, double* part, double* out, hipStream_t s)
{
    const int g = ReducePartials(np);
    hipLaunchKernelGGL(k_colnorm_part, dim3(g), dim3(kThreads), 0, s, x, np, k, part);
    hipLaunchKernelGGL(k_colnorm_final, dim3(k), dim3(kThreads), 0, s, part, g, k, out);
}

void LaunchColDot(const c128* v, const c128* w, int64_t np, int k, c128* part, c128* out,
                  hipStream_t s)
{
    const int g = ReducePartials(np);
    hipLaunchKernelGGL(k_coldot_part, dim3(g), dim3(kThreads), 0, s, v, w, np, k, part);
    hipLaunchKernelGGL(k_coldot_final, dim3(k), dim3(kThreads), 0, s, part, g, k, out);
}

void LaunchAxpyDot(const c128* alpha, const c128* v, const c128* vn, c128* w, int64_t np, int k,
                   c128* part, c128* out, hipStream_t s)
{
    const int g = ReducePartials(np);
    hipLaunchKernelGGL(k_axpy_dot_part, dim3(g), dim3(kThreads), 0, s, alpha, v, vn, w, np, k, part);
    hipLaunchKernelGGL(k_coldot_final, dim3(k), dim3(kThreads), 0, s, part, g, k, out);
}

void LaunchAxpyCols(const c128* alpha, int sign, const c128* v, c128* w, int64_t np, int k,
                    hipStream_t s)
{
    const int64_t n = np * k;
    hipLaunchKernelGGL(k_axpy_cols, dim3(Blocks(n)), dim3(kThreads), 0, s, alpha, sign, v, w, n,
                       k);
}

void LaunchScaleCols(const double* delta, const c128* w, c128* v, int64_t np, int k,
                     hipStream_t s)
{
    const int64_t n = np * k;
    hipLaunchKernelGGL(k_scale_cols, dim3(Blocks(n)), dim3(kThreads), 0, s, delta, w, v, n, k);
}

void LaunchUpdateX(const c128* x0, const c128* V, const c128* y, int j, int64_t np, int k,
                   c128* x, hipStream_t s)
{
    const int64_t n = np * k;
    hipLaunchKernelGGL(k_update_x, dim3(Blocks(n)), dim3(kThreads), 0, s, x0, V, y, j, n, k, n,
                       x);
}

void LaunchNeg(const c128* x, c128* w, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_neg, dim3(Blocks(n)), dim3(kThreads), 0, s, x, w, n);
}

void LaunchAdjointGrad(const StencilArgs& st, const c128* u, const c128* lam, int k, double* grad,
                       int accumulate, hipStream_t s)
{
    const int64_t n = (int64_t)st.nx * st.ny * st.nz;
    hipLaunchKernelGGL(k_adjoint_grad, dim3(Blocks(n)), dim3(kThreads), 0, s, st, u, lam, k, grad,
                       accumulate);
}

// Explicit instantiations for complex double and complex float fronts.
#define PSP_INST(T)                                                                               \
    template void LaunchScatterOrig<T>(const ClassDev&, const StencilArgs&, const double*, double, \
                                       hipStream_t);                                              \
    template void LaunchExtendAdd<T>(const ClassDev&, int, int, hipStream_t);                     \
    template void LaunchLdlDiag<T>(const ClassDev&, int, int, int, hipStream_t);                  \
    template void LaunchLdlTrsm<T>(const ClassDev&, int, int, int, hipStream_t);                  \
    template void LaunchGemm<T>(const ClassDev&, int, int, hipStream_t);                          \
    template void LaunchInvDiag<T>(const ClassDev&, int, hipStream_t);                            \
    template void LaunchPanelSolve<T>(const ClassDev&, const ClassPlan&, int, void*, void*, void*,   \
                                      void*, int, hipStream_t, MvProfile*, int, int, int32_t*);   \
    template void LaunchTransposeScale<T>(const ClassDev&, hipStream_t);                          \
    template void LaunchFoldBack<T>(const ClassDev&, hipStream_t);                                \
    template void LaunchCopyIn<T>(const ClassDev&, const c128*, void*, int, int, int, int, int,   \
                                  int, hipStream_t, int32_t*);                                    \
    template void LaunchCopyOut<T>(const ClassDev&, c128*, const void*, int, int, int, int, int,  \
                                   int, int, hipStream_t, const StencilArgs*, const double*,      \
                                   int);                                                          \
    template void LaunchZeroSetPlane<T>(const ClassDev&, const StencilArgs&, const double*,       \
                                        const c128*, void*, int, int, int, hipStream_t, int32_t*); \
    template void LaunchSolveBatch<T>(int, int, int, int, const BatchBlock*, int,                 \
                                      const BatchSolveSeg*, const int32_t*, hipStream_t);         \
    template void LaunchReduceBatch<T>(int, const BatchBlock*, int, const BatchSolveSeg*,         \
                                       const int32_t*, hipStream_t);                              \
    template void LaunchCopyBatch<T>(int, const BatchBlock*, int, const BatchCopySeg*,            \
                                     const int32_t*, hipStream_t);
PSP_INST(double)
PSP_INST(float)
#undef PSP_INST

}  // namespace psp_amd
