/*
 * psp_oracle.c -- CPU restatement of PSP's sweeping-preconditioned Helmholtz
 * solve in complex double.  TEST INFRASTRUCTURE ONLY: the parity checker for
 * the HIP product path and the CPU baseline.  The product never calls it.
 *
 * Paths in citations are relative to the PSP source tree.  Clique (the
 * multifrontal solver PSP calls) is absent from that tree; its LDL^T and
 * triangular solves are restated here as the textbook unpivoted multifrontal
 * method over PSP's own elimination tree (main.hpp:1479-1900).
 *
 * Layout follows the reference: the global vector is in PSP's nested-
 * dissection (ND) order, column-major N x k (solve.hpp:139); fronts are
 * dense column-major (s+u) x s with D on the diagonal and unit-lower L below
 * (Clique's frontL, initialize_finalize.hpp:1316).
 */
#include "psp_oracle.h"

#include <complex.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

typedef double complex zc;

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
    int off, size, parent, nchild, child[2];
    long ustart;
    int usize;
    int depth;
} node_t;

typedef struct {
    int vsize;      /* padded panel depth                                  */
    int n;          /* nx*ny*vsize                                          */
    int nnodes;
    node_t* nodes;  /* post-order, root last                                */
    int* structs;   /* concatenated sorted lower structures                 */
    long nstruct;
    int* relinds;   /* per struct entry: its row in the parent's front      */
    int* nd2xyv;    /* padded ND index -> x + y*nx + vLocal*nx*ny           */
    long* loff;     /* per node offset into the panel's factor storage      */
    long fstore;    /* sum over nodes of (s+u)*s                            */
    int present;
} class_t;

struct ora_ctx {
    ora_disc d;
    double hx, hy, hz;
    int bottomDepth, topOrigDepth, innerDepth, leftoverDepth;
    int haveLeftover, numFullInner, numPanels;
    long N;
    class_t cls[3];
    int* l2n;        /* global ND -> natural                                */
    int* n2l;        /* natural -> global ND                                */
    int* rowptr;     /* global operator, CSR in PSP's fixed entry order     */
    int* cols;
    zc* vals;
    long nnz;
    int** pmap;      /* per panel: padded ND q -> global ND row or -1       */
    zc** L;          /* per panel factor storage                            */
    int factored;
    double tfactor;
};

/* ------------------------------------------------------------------ */
/* Panel geometry: dist_helmholtz/main.hpp:35-57 and :728-821          */
/* ------------------------------------------------------------------ */

static int panel_padding(const ora_ctx* X, int p) /* main.hpp:728-735 */
{
    return p == 0 ? 0 : X->d.bz;
}

static int panel_depth(const ora_ctx* X, int p) /* main.hpp:737-749 */
{
    if (p == 0) return X->bottomDepth;
    if (p < 1 + X->numFullInner) return X->d.planesPerPanel;
    if (X->haveLeftover && p == 1 + X->numFullInner) return X->leftoverDepth;
    return X->topOrigDepth;
}

static int panel_v(const ora_ctx* X, int p) /* main.hpp:775-789 */
{
    const int P = X->d.planesPerPanel;
    if (p == 0) return 0;
    if (p < X->numFullInner + 1) return X->bottomDepth + P * (p - 1);
    if (X->haveLeftover && p == X->numFullInner + 1)
        return X->bottomDepth + P * X->numFullInner;
    return X->bottomDepth + P * X->numFullInner + X->leftoverDepth;
}

/* Which factorization class a panel uses (main.hpp:704-726): bottom and
 * full inner panels share one analysis. */
static int panel_class(const ora_ctx* X, int p)
{
    if (p < 1 + X->numFullInner) return 0;
    if (X->haveLeftover && p == 1 + X->numFullInner) return 1;
    return 2;
}

/* ------------------------------------------------------------------ */
/* Nested dissection: ReorderedIndexRecursion, main.hpp:1425-1477,     */
/* at depthTilSerial == 0 (one process).                               */
/* ------------------------------------------------------------------ */
static int nd_index(int x, int y, int vl, int xs, int ys, int vs, int cutoff)
{
    int offset = 0;
    for (;;) {
        if (xs * ys <= cutoff) return offset + (x + y * xs + vl * xs * ys);
        if (xs >= ys) {
            const int mid = (xs - 1) / 2;
            if (x < mid) {
                xs = mid;
            } else if (x == mid) {
                return offset + (xs - 1 > 0 ? xs - 1 : 0) * ys * vs +
                       (y + vl * ys);
            } else {
                x -= mid + 1;
                offset += mid * ys * vs;
                xs = xs - mid - 1 > 0 ? xs - mid - 1 : 0;
            }
        } else {
            const int mid = (ys - 1) / 2;
            if (y < mid) {
                ys = mid;
            } else if (y == mid) {
                return offset + xs * (ys - 1 > 0 ? ys - 1 : 0) * vs +
                       (x + vl * xs);
            } else {
                y -= mid + 1;
                offset += xs * mid * vs;
                ys = ys - mid - 1 > 0 ? ys - mid - 1 : 0;
            }
        }
    }
}

int ora_reordered_index(const ora_ctx* X, int x, int y, int vLocal, int vSize)
{
    return nd_index(x, y, vLocal, X->d.nx, X->d.ny, vSize, X->d.cutoff);
}

/* ------------------------------------------------------------------ */
/* Elimination tree: FillPanelLocalElimTree, main.hpp:1709-1900.       */
/* Nodes are numbered in post-order (left, right, separator), which is */
/* what the reference's stack construction produces (root last).       */
/* ------------------------------------------------------------------ */
typedef struct {
    int* data;
    int len;
} ivec;

static int count_nodes(int bnx, int bny, int cutoff) /* main.hpp:600-690 */
{
    if (bnx * bny <= cutoff) return 1;
    if (bnx >= bny) {
        const int mid = (bnx - 1) / 2;
        return 1 + count_nodes(mid, bny, cutoff) +
               count_nodes(bnx - mid - 1 > 0 ? bnx - mid - 1 : 0, bny, cutoff);
    }
    const int mid = (bny - 1) / 2;
    return 1 + count_nodes(bnx, mid, cutoff) +
           count_nodes(bnx, bny - mid - 1 > 0 ? bny - mid - 1 : 0, cutoff);
}

static int cmp_int(const void* a, const void* b)
{
    const int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

static int build_node(const ora_ctx* X, class_t* c, ivec* orig, int* counter,
                      int xo, int yo, int bnx, int bny, int depth)
{
    const int nx = X->d.nx, ny = X->d.ny, vs = c->vsize, cut = X->d.cutoff;
    node_t nd;
    memset(&nd, 0, sizeof(nd));
    nd.depth = depth;
    nd.parent = -1;
    int* os = NULL;
    int olen = 0;
    if (bnx * bny <= cut) {
        /* leaf: the whole box; struct = the box's ring (main.hpp:1762-1797) */
        nd.size = bnx * bny * vs;
        nd.off = nd_index(xo, yo, 0, nx, ny, vs, cut);
        int join = 0;
        if (xo > 0) join += bny;
        if (xo + bnx < nx) join += bny;
        if (yo > 0) join += bnx;
        if (yo + bny < ny) join += bnx;
        os = (int*)malloc(sizeof(int) * (size_t)(join * vs + 1));
        if (xo > 0)
            for (int i = 0; i < vs; ++i)
                for (int j = 0; j < bny; ++j)
                    os[olen++] = nd_index(xo - 1, yo + j, i, nx, ny, vs, cut);
        if (xo + bnx < nx)
            for (int i = 0; i < vs; ++i)
                for (int j = 0; j < bny; ++j)
                    os[olen++] = nd_index(xo + bnx, yo + j, i, nx, ny, vs, cut);
        if (yo > 0)
            for (int i = 0; i < vs; ++i)
                for (int j = 0; j < bnx; ++j)
                    os[olen++] = nd_index(xo + j, yo - 1, i, nx, ny, vs, cut);
        if (yo + bny < ny)
            for (int i = 0; i < vs; ++i)
                for (int j = 0; j < bnx; ++j)
                    os[olen++] = nd_index(xo + j, yo + bny, i, nx, ny, vs, cut);
    } else if (bnx >= bny) {
        /* x separator (main.hpp:1801-1848) */
        const int mid = (bnx - 1) / 2;
        const int l = build_node(X, c, orig, counter, xo, yo, mid, bny, depth + 1);
        const int r = build_node(X, c, orig, counter, xo + mid + 1, yo,
                                 bnx - mid - 1 > 0 ? bnx - mid - 1 : 0, bny,
                                 depth + 1);
        nd.nchild = 2;
        nd.child[0] = l;
        nd.child[1] = r;
        nd.size = bny * vs;
        nd.off = nd_index(xo + mid, yo, 0, nx, ny, vs, cut);
        os = (int*)malloc(sizeof(int) * (size_t)(2 * vs + 1));
        if (yo > 0)
            for (int i = 0; i < vs; ++i)
                os[olen++] = nd_index(xo + mid, yo - 1, i, nx, ny, vs, cut);
        if (yo + bny < ny)
            for (int i = 0; i < vs; ++i)
                os[olen++] = nd_index(xo + mid, yo + bny, i, nx, ny, vs, cut);
    } else {
        /* y separator (main.hpp:1850-1896) */
        const int mid = (bny - 1) / 2;
        const int l = build_node(X, c, orig, counter, xo, yo, bnx, mid, depth + 1);
        const int r = build_node(X, c, orig, counter, xo, yo + mid + 1, bnx,
                                 bny - mid - 1 > 0 ? bny - mid - 1 : 0,
                                 depth + 1);
        nd.nchild = 2;
        nd.child[0] = l;
        nd.child[1] = r;
        nd.size = bnx * vs;
        nd.off = nd_index(xo, yo + mid, 0, nx, ny, vs, cut);
        os = (int*)malloc(sizeof(int) * (size_t)(2 * vs + 1));
        if (xo > 0)
            for (int i = 0; i < vs; ++i)
                os[olen++] = nd_index(xo - 1, yo + mid, i, nx, ny, vs, cut);
        if (xo + bnx < nx)
            for (int i = 0; i < vs; ++i)
                os[olen++] = nd_index(xo + bnx, yo + mid, i, nx, ny, vs, cut);
    }
    if (nd.size == 0) nd.off = 0;
    qsort(os, (size_t)olen, sizeof(int), cmp_int);
    const int idx = (*counter)++;
    c->nodes[idx] = nd;
    orig[idx].data = os;
    orig[idx].len = olen;
    for (int k = 0; k < nd.nchild; ++k) c->nodes[nd.child[k]].parent = idx;
    return idx;
}

/* Symbolic analysis (the part of cliq::SymmetricAnalysis PSP relies on,
 * call site main.hpp:432-435): each node's lower structure is the sorted
 * union of its original structure and its children's structures, minus the
 * node's own indices; relinds map a child's structure into the parent front. */
static void analyse_class(const ora_ctx* X, class_t* c, int vsize)
{
    const int nx = X->d.nx, ny = X->d.ny, cut = X->d.cutoff;
    c->vsize = vsize;
    c->n = nx * ny * vsize;
    c->nnodes = count_nodes(nx, ny, cut);
    c->nodes = (node_t*)calloc((size_t)c->nnodes, sizeof(node_t));
    ivec* orig = (ivec*)calloc((size_t)c->nnodes, sizeof(ivec));
    int counter = 0;
    build_node(X, c, orig, &counter, 0, 0, nx, ny, 0);

    ivec* st = (ivec*)calloc((size_t)c->nnodes, sizeof(ivec));
    long total = 0;
    for (int t = 0; t < c->nnodes; ++t) {
        node_t* nd = &c->nodes[t];
        int cap = orig[t].len;
        for (int k = 0; k < nd->nchild; ++k) cap += st[nd->child[k]].len;
        int* tmp = (int*)malloc(sizeof(int) * (size_t)(cap + 1));
        int len = 0;
        for (int i = 0; i < orig[t].len; ++i) tmp[len++] = orig[t].data[i];
        for (int k = 0; k < nd->nchild; ++k) {
            const ivec* cs = &st[nd->child[k]];
            for (int i = 0; i < cs->len; ++i) tmp[len++] = cs->data[i];
        }
        qsort(tmp, (size_t)len, sizeof(int), cmp_int);
        int out = 0;
        for (int i = 0; i < len; ++i) {
            const int g = tmp[i];
            if (g >= nd->off && g < nd->off + nd->size) continue;
            if (out > 0 && tmp[out - 1] == g) continue;
            tmp[out++] = g;
        }
        st[t].data = tmp;
        st[t].len = out;
        nd->usize = out;
        nd->ustart = total;
        total += out;
    }
    c->nstruct = total;
    c->structs = (int*)malloc(sizeof(int) * (size_t)(total + 1));
    c->relinds = (int*)malloc(sizeof(int) * (size_t)(total + 1));
    c->loff = (long*)malloc(sizeof(long) * (size_t)(c->nnodes + 1));
    long fs = 0;
    for (int t = 0; t < c->nnodes; ++t) {
        node_t* nd = &c->nodes[t];
        memcpy(c->structs + nd->ustart, st[t].data, sizeof(int) * (size_t)nd->usize);
        c->loff[t] = fs;
        fs += (long)(nd->size + nd->usize) * nd->size;
    }
    c->loff[c->nnodes] = fs;
    c->fstore = fs;
    for (int t = 0; t < c->nnodes; ++t) {
        const node_t* nd = &c->nodes[t];
        if (nd->parent < 0) continue;
        const node_t* pa = &c->nodes[nd->parent];
        const int* ps = c->structs + pa->ustart;
        for (int a = 0; a < nd->usize; ++a) {
            const int g = c->structs[nd->ustart + a];
            int rel;
            if (g >= pa->off && g < pa->off + pa->size) {
                rel = g - pa->off;
            } else {
                const int* f = (const int*)bsearch(&g, ps, (size_t)pa->usize,
                                                   sizeof(int), cmp_int);
                rel = f ? pa->size + (int)(f - ps) : -1;
            }
            c->relinds[nd->ustart + a] = rel;
        }
    }
    /* padded ND index -> panel-natural (x + y*nx + vLocal*nx*ny) */
    c->nd2xyv = (int*)malloc(sizeof(int) * (size_t)c->n);
    for (int v = 0; v < vsize; ++v)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
                c->nd2xyv[nd_index(x, y, v, nx, ny, vsize, cut)] =
                    x + y * nx + v * nx * ny;
    for (int t = 0; t < c->nnodes; ++t) {
        free(orig[t].data);
        free(st[t].data);
    }
    free(orig);
    free(st);
    c->present = 1;
}

/* ------------------------------------------------------------------ */
/* PML profiles: initialize_finalize.hpp:415-525                       */
/* ------------------------------------------------------------------ */
/* Planted mistakes for the pinning tests' mutant builds (-DORA_PLANT=n,
 * compiled into a separate library by tests/test_oracle_pml.py); the
 * oracle proper is built with ORA_PLANT = 0 and none of them.
 *   1: back face of the x profile shifted by one cell
 *   2: imaginary part of every profile negated
 *   3: profile linear instead of quadratic in the depth
 *   4: artificial PML of a panel one plane too deep
 *   5: damping applied as omega - i*damping                               */
#ifndef ORA_PLANT
#define ORA_PLANT 0
#endif

/* 1 + i (sigma/eta) (delta/b)^2 (2 pi/omega), eta = b h: the profile value
 * at depth delta cells into a layer of b cells (initialize_finalize.hpp:
 * 420-426 and the seven copies after it). */
static zc pml_value(double sigma, int b, double h, double omega, double delta)
{
    const double eta = b * h;
    const double q = (sigma / eta) * (delta / b);
    const double im = (ORA_PLANT == 3 ? q : q * (delta / b)) * (2 * M_PI / omega);
    return 1.0 + (ORA_PLANT == 2 ? -im : im) * I;
}

static zc s1inv(const ora_ctx* X, int x)
{
    const int bx = X->d.bx;
    const int back = X->d.nx - bx - (ORA_PLANT == 1);
    if (x < (bx - 1) && X->d.frontBC == ORA_PML)
        return pml_value(X->d.sigmax, bx, X->hx, X->d.omega, (bx - 1) - x);
    else if (x > back && X->d.backBC == ORA_PML)
        return pml_value(X->d.sigmax, bx, X->hx, X->d.omega, x - back);
    return 1.0;
}

static zc s2inv(const ora_ctx* X, int y)
{
    const int by = X->d.by;
    if (y < (by - 1) && X->d.leftBC == ORA_PML)
        return pml_value(X->d.sigmay, by, X->hy, X->d.omega, (by - 1) - y);
    else if (y > (X->d.ny - by) && X->d.rightBC == ORA_PML)
        return pml_value(X->d.sigmay, by, X->hy, X->d.omega, y - (X->d.ny - by));
    return 1.0;
}

static zc s3inv(const ora_ctx* X, int v)
{
    const int bz = X->d.bz;
    if (v < bz - 1)
        return pml_value(X->d.sigmaz, bz, X->hz, X->d.omega, (bz - 1) - v);
    else if (v > (X->d.nz - bz) && X->d.topBC == ORA_PML)
        return pml_value(X->d.sigmaz, bz, X->hz, X->d.omega, v - (X->d.nz - bz));
    return 1.0;
}

/* s3InvArtificial (initialize_finalize.hpp:500-525): the panel's own PML
 * over the bz planes below vOffset + bz - 1. */
static zc s3inv_art(const ora_ctx* X, int v, int vOffset)
{
    const int bz = X->d.bz;
    const int lim = vOffset + bz - 1 + (ORA_PLANT == 4);
    if (v < lim)
        return pml_value(X->d.sigmaz, bz, X->hz, X->d.omega, lim - v);
    else if (v > (X->d.nz - bz) && X->d.topBC == ORA_PML)
        return pml_value(X->d.sigmaz, bz, X->hz, X->d.omega, v - (X->d.nz - bz));
    return 1.0;
}

/* Stencil terms shared by FormGlobalRow (initialize_finalize.hpp:527-587)
 * and FormLowerColumnOfNode (:589-806).  out[0] = centre, out[1..6] = the
 * off-diagonal values for x-1, x+1, y-1, y+1, v-1, v+1. */
static void stencil(const ora_ctx* X, zc s1L, zc s1M, zc s1R, zc s2L, zc s2M,
                    zc s2R, zc s3L, zc s3M, zc s3R, zc k2, zc out[7])
{
    const double hx = X->hx, hy = X->hy, hz = X->hz;
    const zc xTop = s2M * s3M;
    const zc xTempL = xTop / s1L, xTempM = xTop / s1M, xTempR = xTop / s1R;
    const zc xTermL = (xTempL + xTempM) / (2 * hx * hx);
    const zc xTermR = (xTempR + xTempM) / (2 * hx * hx);
    const zc yTop = s1M * s3M;
    const zc yTempL = yTop / s2L, yTempM = yTop / s2M, yTempR = yTop / s2R;
    const zc yTermL = (yTempL + yTempM) / (2 * hy * hy);
    const zc yTermR = (yTempR + yTempM) / (2 * hy * hy);
    const zc zTop = s1M * s2M;
    const zc vTempL = zTop / s3L, vTempM = zTop / s3M, vTempR = zTop / s3R;
    const zc vTermL = (vTempL + vTempM) / (2 * hz * hz);
    const zc vTermR = (vTempR + vTempM) / (2 * hz * hz);
    out[0] = (xTermL + xTermR + yTermL + yTermR + vTermL + vTermR) -
             k2 * s1M * s2M * s3M;
    out[1] = -xTermL;
    out[2] = -xTermR;
    out[3] = -yTermL;
    out[4] = -yTermR;
    out[5] = -vTermL;
    out[6] = -vTermR;
}

/* FormGlobalRow (initialize_finalize.hpp:527-587): undamped omega,
 * physical PML only. */
static void global_row(const ora_ctx* X, double alpha, int x, int y, int v,
                       zc out[7])
{
    const double wa = X->d.omega / alpha;
    stencil(X, s1inv(X, x - 1), s1inv(X, x), s1inv(X, x + 1), s2inv(X, y - 1),
            s2inv(X, y), s2inv(X, y + 1), s3inv(X, v - 1), s3inv(X, v),
            s3inv(X, v + 1), wa * wa, out);
}

/* The column values of FormLowerColumnOfNode (initialize_finalize.hpp:
 * 589-640): shifted frequency omega + i*damping, artificial PML below the
 * panel. */
static void panel_row(const ora_ctx* X, double alpha, int x, int y, int v,
                      int vOffset, zc out[7])
{
    const zc so = X->d.omega + (ORA_PLANT == 5 ? -X->d.damping : X->d.damping) * I;
    const zc sa = so / alpha;
    stencil(X, s1inv(X, x - 1), s1inv(X, x), s1inv(X, x + 1), s2inv(X, y - 1),
            s2inv(X, y), s2inv(X, y + 1), s3inv_art(X, v - 1, vOffset),
            s3inv_art(X, v, vOffset), s3inv_art(X, v + 1, vOffset), sa * sa,
            out);
}

/* ------------------------------------------------------------------ */
/* Construction: DistHelmholtz ctor, main.hpp:12-436 (one process).    */
/* ------------------------------------------------------------------ */
int ora_create(const ora_disc* disc, ora_ctx** out)
{
    *out = NULL;
    if (!disc || disc->nx < 1 || disc->ny < 1 || disc->nz < 1 ||
        disc->planesPerPanel < 1 || disc->cutoff < 1 || disc->bz < 1)
        return ORA_ERR_INVALID;
    ora_ctx* X = (ora_ctx*)calloc(1, sizeof(ora_ctx));
    X->d = *disc;
    X->hx = disc->wx / (disc->nx + 1);
    X->hy = disc->wy / (disc->ny + 1);
    X->hz = disc->wz / (disc->nz + 1);
    const int P = disc->planesPerPanel, bz = disc->bz, nz = disc->nz;
    X->bottomDepth = bz + P;
    X->topOrigDepth = disc->topBC == ORA_PML ? bz + P : P;
    if (nz <= X->bottomDepth + X->topOrigDepth) {
        free(X);
        return ORA_ERR_SHALLOW; /* main.hpp:38-41 */
    }
    X->innerDepth = nz - (X->bottomDepth + X->topOrigDepth);
    X->leftoverDepth = X->innerDepth % P;
    X->haveLeftover = X->leftoverDepth != 0;
    X->numFullInner = X->innerDepth / P;
    X->numPanels = 1 + X->numFullInner + X->haveLeftover + 1;
    const int nx = disc->nx, ny = disc->ny;
    X->N = (long)nx * ny * nz;

    analyse_class(X, &X->cls[0], X->bottomDepth);
    if (X->haveLeftover) analyse_class(X, &X->cls[1], X->leftoverDepth + bz);
    analyse_class(X, &X->cls[2], X->topOrigDepth + bz);

    /* Global row order: MapLocalPanelIndices (main.hpp:823-1059). */
    X->l2n = (int*)malloc(sizeof(int) * (size_t)X->N);
    X->n2l = (int*)malloc(sizeof(int) * (size_t)X->N);
    for (int p = 0; p < X->numPanels; ++p) {
        const int vP = panel_v(X, p), dep = panel_depth(X, p);
        const long base = (long)nx * ny * vP;
        for (int vl = 0; vl < dep; ++vl)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    const int v = vP + vl, z = nz - 1 - v;
                    const long g = base + nd_index(x, y, vl, nx, ny, dep, disc->cutoff);
                    const int nat = x + y * nx + z * nx * ny;
                    X->l2n[g] = nat;
                    X->n2l[nat] = (int)g;
                }
    }
    /* Connection structure: MapLocalConnectionIndices (main.hpp:1061-1273),
     * fixed order [self, x-1, x+1, y-1, y+1, v-1 (z+1), v+1 (z-1)]. */
    X->rowptr = (int*)malloc(sizeof(int) * (size_t)(X->N + 1));
    X->rowptr[0] = 0;
    for (long g = 0; g < X->N; ++g) {
        const int nat = X->l2n[g];
        const int x = nat % nx, y = (nat / nx) % ny, z = nat / (nx * ny);
        const int v = nz - 1 - z;
        int cnt = 1 + (x > 0) + (x + 1 < nx) + (y > 0) + (y + 1 < ny) +
                  (v > 0) + (v + 1 < nz);
        X->rowptr[g + 1] = X->rowptr[g] + cnt;
    }
    X->nnz = X->rowptr[X->N];
    X->cols = (int*)malloc(sizeof(int) * (size_t)X->nnz);
    X->vals = (zc*)calloc((size_t)X->nnz, sizeof(zc));
    for (long g = 0; g < X->N; ++g) {
        const int nat = X->l2n[g];
        const int x = nat % nx, y = (nat / nx) % ny, z = nat / (nx * ny);
        const int v = nz - 1 - z;
        int o = X->rowptr[g];
        X->cols[o++] = (int)g;
        if (x > 0) X->cols[o++] = X->n2l[nat - 1];
        if (x + 1 < nx) X->cols[o++] = X->n2l[nat + 1];
        if (y > 0) X->cols[o++] = X->n2l[nat - nx];
        if (y + 1 < ny) X->cols[o++] = X->n2l[nat + nx];
        if (v > 0) X->cols[o++] = X->n2l[nat + nx * ny];
        if (v + 1 < nz) X->cols[o++] = X->n2l[nat - nx * ny];
    }
    /* Per panel: padded ND index -> global row (SolvePanel's memcpy plan,
     * solve.hpp:756-778), -1 on the padding planes. */
    X->pmap = (int**)calloc((size_t)X->numPanels, sizeof(int*));
    for (int p = 0; p < X->numPanels; ++p) {
        const class_t* c = &X->cls[panel_class(X, p)];
        const int pad = panel_padding(X, p), dep = panel_depth(X, p);
        const long base = (long)nx * ny * panel_v(X, p);
        int* m = (int*)malloc(sizeof(int) * (size_t)c->n);
        for (int q = 0; q < c->n; ++q) {
            const int w = c->nd2xyv[q];
            const int x = w % nx, y = (w / nx) % ny, vl = w / (nx * ny);
            m[q] = vl < pad ? -1
                            : (int)(base + nd_index(x, y, vl - pad, nx, ny, dep,
                                                    disc->cutoff));
        }
        X->pmap[p] = m;
    }
    X->L = (zc**)calloc((size_t)X->numPanels, sizeof(zc*));
    *out = X;
    return ORA_OK;
}

static void free_class(class_t* c)
{
    if (!c->present) return;
    free(c->nodes);
    free(c->structs);
    free(c->relinds);
    free(c->nd2xyv);
    free(c->loff);
}

void ora_destroy(ora_ctx* X)
{
    if (!X) return;
    for (int i = 0; i < 3; ++i) free_class(&X->cls[i]);
    for (int p = 0; p < X->numPanels; ++p) {
        free(X->pmap[p]);
        free(X->L[p]);
    }
    free(X->pmap);
    free(X->L);
    free(X->l2n);
    free(X->n2l);
    free(X->rowptr);
    free(X->cols);
    free(X->vals);
    free(X);
}

int ora_num_panels(const ora_ctx* X) { return X->numPanels; }
int ora_num_threads(void) { return omp_get_max_threads(); }
long ora_global_size(const ora_ctx* X) { return X->N; }

int ora_panel_info(const ora_ctx* X, int p, int* vPanel, int* depth,
                   int* padding, int* cls)
{
    if (p < 0 || p >= X->numPanels) return ORA_ERR_INVALID;
    if (vPanel) *vPanel = panel_v(X, p);
    if (depth) *depth = panel_depth(X, p);
    if (padding) *padding = panel_padding(X, p);
    if (cls) *cls = panel_class(X, p);
    return ORA_OK;
}

void ora_local_to_natural(const ora_ctx* X, int* map)
{
    memcpy(map, X->l2n, sizeof(int) * (size_t)X->N);
}

int ora_class_vsize(const ora_ctx* X, int cls)
{
    return X->cls[cls].present ? X->cls[cls].vsize : 0;
}
int ora_class_num_nodes(const ora_ctx* X, int cls)
{
    return X->cls[cls].present ? X->cls[cls].nnodes : 0;
}
void ora_class_nodes(const ora_ctx* X, int cls, int* off, int* size,
                     int* parent, int* usize)
{
    const class_t* c = &X->cls[cls];
    for (int t = 0; t < c->nnodes; ++t) {
        if (off) off[t] = c->nodes[t].off;
        if (size) size[t] = c->nodes[t].size;
        if (parent) parent[t] = c->nodes[t].parent;
        if (usize) usize[t] = c->nodes[t].usize;
    }
}
void ora_class_structs(const ora_ctx* X, int cls, int* structs)
{
    const class_t* c = &X->cls[cls];
    memcpy(structs, c->structs, sizeof(int) * (size_t)c->nstruct);
}

/* ------------------------------------------------------------------ */
/* Global operator values: the global sparse matrix block of           */
/* Initialize, initialize_finalize.hpp:340-376.                        */
/* ------------------------------------------------------------------ */
static void fill_global(ora_ctx* X, const double* vel)
{
    const int nx = X->d.nx, ny = X->d.ny, nz = X->d.nz;
#pragma omp parallel for schedule(static)
    for (long g = 0; g < X->N; ++g) {
        const int nat = X->l2n[g];
        const int x = nat % nx, y = (nat / nx) % ny, z = nat / (nx * ny);
        const int v = nz - 1 - z;
        zc e[7];
        global_row(X, vel[nat], x, y, v, e);
        int o = X->rowptr[g];
        X->vals[o++] = e[0];
        if (x > 0) X->vals[o++] = e[1];
        if (x + 1 < nx) X->vals[o++] = e[2];
        if (y > 0) X->vals[o++] = e[3];
        if (y + 1 < ny) X->vals[o++] = e[4];
        if (v > 0) X->vals[o++] = e[5];
        if (v + 1 < nz) X->vals[o++] = e[6];
    }
}

long ora_global_nnz(const ora_ctx* X) { return X->nnz; }

int ora_global_csr(const ora_ctx* X, int* rowptr, int* cols, double* vals)
{
    memcpy(rowptr, X->rowptr, sizeof(int) * (size_t)(X->N + 1));
    memcpy(cols, X->cols, sizeof(int) * (size_t)X->nnz);
    memcpy(vals, X->vals, sizeof(zc) * (size_t)X->nnz);
    return ORA_OK;
}

/* ------------------------------------------------------------------ */
/* Front filling: FillPanelFronts + FormLowerColumnOfNode              */
/* (initialize_finalize.hpp:1284-1337, 589-806).                       */
/* Calls emit(row, col, value) for the strictly-lower and diagonal     */
/* entries of node t's columns, rows in front coordinates.             */
/* ------------------------------------------------------------------ */
typedef void (*emit_fn)(void* ud, int row, int col, zc val);

static int front_row(const class_t* c, const node_t* nd, int nested)
{
    if (nested < nd->off + nd->size) return nested - nd->off;
    const int* s = c->structs + nd->ustart;
    const int* f = (const int*)bsearch(&nested, s, (size_t)nd->usize,
                                       sizeof(int), cmp_int);
    return f ? nd->size + (int)(f - s) : -1;
}

static void fill_node_columns(const ora_ctx* X, const class_t* c, int t,
                              int vOffset, const double* vel, emit_fn emit,
                              void* ud)
{
    const int nx = X->d.nx, ny = X->d.ny, nz = X->d.nz, vs = c->vsize;
    const int cut = X->d.cutoff;
    const node_t* nd = &c->nodes[t];
    for (int j = 0; j < nd->size; ++j) {
        const int w = c->nd2xyv[nd->off + j];
        const int x = w % nx, y = (w / nx) % ny, vl = w / (nx * ny);
        const int v = vOffset + vl, z = nz - 1 - v;
        zc e[7];
        panel_row(X, vel[x + y * nx + z * nx * ny], x, y, v, vOffset, e);
        emit(ud, j, j, e[0]);
        const int me = nd->off + j;
        int nb[6];
        int ok[6];
        ok[0] = x > 0;
        ok[1] = x + 1 < nx;
        ok[2] = y > 0;
        ok[3] = y + 1 < ny;
        ok[4] = vl > 0;
        ok[5] = vl + 1 < vs;
        nb[0] = ok[0] ? nd_index(x - 1, y, vl, nx, ny, vs, cut) : -1;
        nb[1] = ok[1] ? nd_index(x + 1, y, vl, nx, ny, vs, cut) : -1;
        nb[2] = ok[2] ? nd_index(x, y - 1, vl, nx, ny, vs, cut) : -1;
        nb[3] = ok[3] ? nd_index(x, y + 1, vl, nx, ny, vs, cut) : -1;
        nb[4] = ok[4] ? nd_index(x, y, vl - 1, nx, ny, vs, cut) : -1;
        nb[5] = ok[5] ? nd_index(x, y, vl + 1, nx, ny, vs, cut) : -1;
        for (int k = 0; k < 6; ++k) {
            if (!ok[k] || nb[k] <= me) continue;
            const int row = front_row(c, nd, nb[k]);
            if (row >= 0) emit(ud, row, j, e[k + 1]);
        }
    }
}

typedef struct {
    zc* F;
    int ld;
} dense_ud;

static void emit_dense(void* ud, int row, int col, zc val)
{
    dense_ud* d = (dense_ud*)ud;
    d->F[row + (long)col * d->ld] = val;
}

typedef struct {
    int* rows;
    int* cols;
    double* vals;
    long n;
    int off;
    const class_t* c;
    const node_t* nd;
} coo_ud;

static void emit_coo(void* ud, int row, int col, zc val)
{
    coo_ud* d = (coo_ud*)ud;
    const int gr = row < d->nd->size ? d->nd->off + row
                                     : d->c->structs[d->nd->ustart + row - d->nd->size];
    if (d->rows) {
        d->rows[d->n] = gr;
        d->cols[d->n] = d->nd->off + col;
        d->vals[2 * d->n] = creal(val);
        d->vals[2 * d->n + 1] = cimag(val);
    }
    ++d->n;
}

long ora_panel_lower_coo(ora_ctx* X, int panel, const double* vel, int* rows,
                         int* cols, double* vals)
{
    if (panel < 0 || panel >= X->numPanels) return -1;
    const class_t* c = &X->cls[panel_class(X, panel)];
    const int vOffset = panel_v(X, panel) - panel_padding(X, panel);
    coo_ud ud;
    ud.rows = rows;
    ud.cols = cols;
    ud.vals = vals;
    ud.n = 0;
    ud.c = c;
    for (int t = 0; t < c->nnodes; ++t) {
        ud.nd = &c->nodes[t];
        fill_node_columns(X, c, t, vOffset, vel, emit_coo, &ud);
    }
    return ud.n;
}

/* ------------------------------------------------------------------ */
/* Dense partial LDL^T of one front (the per-front work of cliq::LDL,  */
/* call sites initialize_finalize.hpp:145-150): eliminates the first s */
/* of fs columns of a column-major fs x fs lower-stored front; leaves  */
/* D on the diagonal, unit-lower L below it, and the Schur complement  */
/* in the trailing (fs-s) x (fs-s) lower block.  Unpivoted.            */
/* ------------------------------------------------------------------ */
static void partial_ldl(zc* F, int fs, int s)
{
    const int nb = 48;
    const long ld = fs;
    zc* w = (zc*)malloc(sizeof(zc) * (size_t)(fs > 0 ? fs : 1) * nb);
    for (int jb = 0; jb < s; jb += nb) {
        const int je = jb + nb < s ? jb + nb : s;
        for (int j = jb; j < je; ++j) {
            const zc d = F[j + j * ld];
            const zc dinv = 1.0 / d;
            for (int c = j + 1; c < je; ++c) {
                const zc lc = F[c + j * ld] * dinv;
                zc* Fc = F + c * ld;
                const zc* Fj = F + j * ld;
                for (int i = c; i < fs; ++i) Fc[i] -= Fj[i] * lc;
            }
            for (int i = j + 1; i < fs; ++i) F[i + j * ld] *= dinv;
        }
        /* w(c, j) = d_j * l_c,j for the trailing columns */
        const int nt = fs - je;
        if (nt <= 0) continue;
        for (int j = jb; j < je; ++j) {
            const zc d = F[j + j * ld];
            for (int c = je; c < fs; ++c) w[(c - je) + (long)(j - jb) * nt] = d * F[c + j * ld];
        }
#pragma omp parallel for schedule(dynamic, 8) if (!omp_in_parallel() && nt > 256)
        for (int c = je; c < fs; ++c) {
            zc* Fc = F + c * ld;
            for (int j = jb; j < je; ++j) {
                const zc wc = w[(c - je) + (long)(j - jb) * nt];
                const zc* Fj = F + j * ld;
                for (int i = c; i < fs; ++i) Fc[i] -= Fj[i] * wc;
            }
        }
    }
    free(w);
}

int ora_partial_ldl(double* F, int fs, int s)
{
    if (fs < 0 || s < 0 || s > fs) return ORA_ERR_INVALID;
    partial_ldl((zc*)F, fs, s);
    return ORA_OK;
}

/* Multifrontal factorization of one panel: assemble each front from the
 * original columns (FillPanelFronts) plus the extend-add of its children's
 * Schur complements, eliminate, keep L, pass the Schur complement up. */
static void factor_panel(ora_ctx* X, int panel, const double* vel)
{
    const class_t* c = &X->cls[panel_class(X, panel)];
    const int vOffset = panel_v(X, panel) - panel_padding(X, panel);
    if (!X->L[panel]) X->L[panel] = (zc*)malloc(sizeof(zc) * (size_t)(c->fstore + 1));
    zc* Lst = X->L[panel];
    zc** U = (zc**)calloc((size_t)c->nnodes, sizeof(zc*));
    for (int t = 0; t < c->nnodes; ++t) {
        const node_t* nd = &c->nodes[t];
        const int s = nd->size, u = nd->usize, fs = s + u;
        zc* F = (zc*)calloc((size_t)fs * fs + 1, sizeof(zc));
        dense_ud ud = {F, fs};
        fill_node_columns(X, c, t, vOffset, vel, emit_dense, &ud);
        for (int k = 0; k < nd->nchild; ++k) {
            const int ch = nd->child[k];
            const node_t* cn = &c->nodes[ch];
            const int cu = cn->usize;
            const int* rel = c->relinds + cn->ustart;
            const zc* Uc = U[ch];
            for (int b = 0; b < cu; ++b)
                for (int a = b; a < cu; ++a)
                    F[rel[a] + (long)rel[b] * fs] += Uc[a + (long)b * cu];
            free(U[ch]);
            U[ch] = NULL;
        }
        partial_ldl(F, fs, s);
        memcpy(Lst + c->loff[t], F, sizeof(zc) * (size_t)fs * s);
        if (nd->parent >= 0 && u > 0) {
            zc* Ut = (zc*)malloc(sizeof(zc) * (size_t)u * u);
            for (int b = 0; b < u; ++b)
                memcpy(Ut + (long)b * u, F + s + (long)(s + b) * fs, sizeof(zc) * (size_t)u);
            U[t] = Ut;
        }
        free(F);
    }
    for (int t = 0; t < c->nnodes; ++t) free(U[t]);
    free(U);
}

int ora_factor_sample(ora_ctx* X, const double* vel, int first, int count,
                      double* seconds)
{
    if (first < 0 || first + count > X->numPanels) return ORA_ERR_INVALID;
    const double t0 = omp_get_wtime();
    /* one panel at a time so the dense front kernels use every thread */
    for (int p = first; p < first + count; ++p) factor_panel(X, p, vel);
    if (seconds) *seconds = omp_get_wtime() - t0;
    return ORA_OK;
}

/* Initialize (initialize_finalize.hpp:12-377). */
int ora_factor(ora_ctx* X, const double* vel)
{
    if (X->d.omega == 0.0) return ORA_ERR_INVALID; /* :17-18 */
    const double t0 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < X->numPanels; ++p) factor_panel(X, p, vel);
    fill_global(X, vel);
    X->tfactor = omp_get_wtime() - t0;
    X->factored = 1;
    return ORA_OK;
}

double ora_factor_seconds(const ora_ctx* X) { return X->tfactor; }

int ora_fill_global(ora_ctx* X, const double* vel)
{
    fill_global(X, vel);
    X->factored = 1;
    return ORA_OK;
}

/* ------------------------------------------------------------------ */
/* Multifrontal triangular solves (cliq::Solve, call solve.hpp:822):   */
/* forward L, diagonal D, backward L^T over one padded panel vector.   */
/* ------------------------------------------------------------------ */
static void panel_ldl_solve(const ora_ctx* X, int panel, zc* xr)
{
    const class_t* c = &X->cls[panel_class(X, panel)];
    const zc* Lst = X->L[panel];
    for (int t = 0; t < c->nnodes; ++t) {
        const node_t* nd = &c->nodes[t];
        const int s = nd->size, u = nd->usize, fs = s + u;
        const zc* L = Lst + c->loff[t];
        zc* own = xr + nd->off;
        const int* st = c->structs + nd->ustart;
        for (int j = 0; j < s; ++j) {
            const zc yj = own[j];
            const zc* Lj = L + (long)j * fs;
            for (int i = j + 1; i < s; ++i) own[i] -= Lj[i] * yj;
            for (int a = 0; a < u; ++a) xr[st[a]] -= Lj[s + a] * yj;
        }
    }
    for (int t = 0; t < c->nnodes; ++t) {
        const node_t* nd = &c->nodes[t];
        const int s = nd->size, fs = s + nd->usize;
        const zc* L = Lst + c->loff[t];
        zc* own = xr + nd->off;
        for (int j = 0; j < s; ++j) own[j] /= L[j + (long)j * fs];
    }
    for (int t = c->nnodes - 1; t >= 0; --t) {
        const node_t* nd = &c->nodes[t];
        const int s = nd->size, u = nd->usize, fs = s + u;
        const zc* L = Lst + c->loff[t];
        zc* own = xr + nd->off;
        const int* st = c->structs + nd->ustart;
        for (int j = s - 1; j >= 0; --j) {
            const zc* Lj = L + (long)j * fs;
            zc acc = own[j];
            for (int a = 0; a < u; ++a) acc -= Lj[s + a] * xr[st[a]];
            for (int i = j + 1; i < s; ++i) acc -= Lj[i] * own[i];
            own[j] = acc;
        }
    }
}

/* SolvePanel (solve.hpp:736-882): zero the padding, copy the panel's rows
 * in, solve, copy them back out. */
static void solve_panel(const ora_ctx* X, zc* B, int k, int panel)
{
    const class_t* c = &X->cls[panel_class(X, panel)];
    const int* m = X->pmap[panel];
    const long N = X->N;
    const int n = c->n;
#pragma omp parallel for schedule(static) if (k > 1)
    for (int r = 0; r < k; ++r) {
        zc* xp = (zc*)calloc((size_t)n, sizeof(zc));
        zc* Br = B + r * N;
        for (int q = 0; q < n; ++q)
            if (m[q] >= 0) xp[q] = Br[m[q]];
        panel_ldl_solve(X, panel, xp);
        for (int q = 0; q < n; ++q)
            if (m[q] >= 0) Br[m[q]] = xp[q];
        free(xp);
    }
}

int ora_solve_panel(const ora_ctx* X, double* B, int k, int panel)
{
    if (!X->factored) return ORA_ERR_STATE;
    if (panel < 0 || panel >= X->numPanels) return ORA_ERR_INVALID;
    solve_panel(X, (zc*)B, k, panel);
    return ORA_OK;
}

/* Index of the v-1 (dv = -1) or v+1 (dv = +1) entry of a global row. */
static int vert_entry(const ora_ctx* X, long g, int dv)
{
    const int nx = X->d.nx, ny = X->d.ny;
    const int nat = X->l2n[g];
    const int target = X->n2l[nat + (dv < 0 ? nx * ny : -nx * ny)];
    for (int o = X->rowptr[g] + 1; o < X->rowptr[g + 1]; ++o)
        if (X->cols[o] == target) return o;
    return -1;
}

/* B_{i+1} := B_{i+1} - A_{i+1,i} B_i  (SubdiagonalUpdate, solve.hpp:884-948) */
static void subdiag_update(const ora_ctx* X, zc* B, int k, int i)
{
    const int nx = X->d.nx, ny = X->d.ny, nz = X->d.nz;
    const int v = panel_v(X, i + 1), z = nz - 1 - v;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const int nat = x + y * nx + z * nx * ny;
            const long g = X->n2l[nat];
            const int o = vert_entry(X, g, -1);
            const zc a = X->vals[o];
            const long col = X->cols[o];
            for (int r = 0; r < k; ++r) B[g + r * X->N] -= a * B[col + r * X->N];
        }
}

/* B_i := -A_{i,i+1} B_{i+1} on the last plane of panel i
 * (MultiplySuperdiagonal, solve.hpp:998-1047). */
static void supdiag_set(const ora_ctx* X, zc* B, int k, int i)
{
    const int nx = X->d.nx, ny = X->d.ny, nz = X->d.nz;
    const int v = panel_v(X, i) + panel_depth(X, i) - 1, z = nz - 1 - v;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const int nat = x + y * nx + z * nx * ny;
            const long g = X->n2l[nat];
            const int o = vert_entry(X, g, +1);
            const zc a = X->vals[o];
            const long col = X->cols[o];
            for (int r = 0; r < k; ++r) B[g + r * X->N] = -a * B[col + r * X->N];
        }
}

/* Precondition (solve.hpp:681-734). */
static void precondition(const ora_ctx* X, zc* B, int k)
{
    const long N = X->N;
    const int np = X->numPanels, nx = X->d.nx, ny = X->d.ny;
    for (int i = 0; i < np - 1; ++i) {
        solve_panel(X, B, k, i);
        subdiag_update(X, B, k, i);
    }
    solve_panel(X, B, k, np - 1);
    for (int i = np - 2; i >= 0; --i) {
        /* ExtractPanel (solve.hpp:950-966) */
        const long off = (long)nx * ny * panel_v(X, i);
        const long len = (long)nx * ny * panel_depth(X, i);
        zc* Z = (zc*)malloc(sizeof(zc) * (size_t)(len * k + 1));
        for (int r = 0; r < k; ++r) {
            memcpy(Z + r * len, B + off + r * N, sizeof(zc) * (size_t)len);
            memset(B + off + r * N, 0, sizeof(zc) * (size_t)len);
        }
        supdiag_set(X, B, k, i);
        solve_panel(X, B, k, i);
        /* UpdatePanel (solve.hpp:1049-1061) */
        for (int r = 0; r < k; ++r)
            for (long s = 0; s < len; ++s) B[off + s + r * N] += Z[r * len + s];
        free(Z);
    }
}

int ora_precondition(const ora_ctx* X, double* B, int k)
{
    if (!X->factored) return ORA_ERR_STATE;
    precondition(X, (zc*)B, k);
    return ORA_OK;
}

/* Multiply (solve.hpp:613-679): B := A B. */
static void multiply(const ora_ctx* X, zc* B, int k, zc* tmp)
{
    const long N = X->N;
#pragma omp parallel for schedule(static)
    for (long g = 0; g < N; ++g) {
        const int o0 = X->rowptr[g], o1 = X->rowptr[g + 1];
        for (int r = 0; r < k; ++r) {
            zc acc = B[g + r * N] * X->vals[o0];
            for (int o = o0 + 1; o < o1; ++o) acc += X->vals[o] * B[X->cols[o] + r * N];
            tmp[g + r * N] = acc;
        }
    }
    memcpy(B, tmp, sizeof(zc) * (size_t)(N * k));
}

int ora_multiply(const ora_ctx* X, double* B, int k)
{
    if (!X->factored) return ORA_ERR_STATE;
    zc* tmp = (zc*)malloc(sizeof(zc) * (size_t)(X->N * k));
    multiply(X, (zc*)B, k, tmp);
    free(tmp);
    return ORA_OK;
}

/* ------------------------------------------------------------------ */
/* BLAS-1 helpers: Norms / InnerProducts / DivideColumns /             */
/* SubtractScaledColumns (environment.hpp:177-257).                    */
/* ------------------------------------------------------------------ */
static double nrm2(const zc* x, long n)
{
    double s = 0.0;
    for (long i = 0; i < n; ++i) s += creal(x[i]) * creal(x[i]) + cimag(x[i]) * cimag(x[i]);
    return sqrt(s);
}

static zc dotc(const zc* x, const zc* y, long n)
{
    zc s = 0.0;
    for (long i = 0; i < n; ++i) s += conj(x[i]) * y[i];
    return s;
}

/* lapack::ComputeGivens (solve.hpp:486): LAPACK zlartg semantics,
 * [c s; -conj(s) c] [f; g] = [r; 0] with c real. */
static void givens(zc f, zc g, double* c, zc* s, zc* r)
{
    if (g == 0.0) {
        *c = 1.0;
        *s = 0.0;
        *r = f;
        return;
    }
    if (f == 0.0) {
        const double ag = cabs(g);
        *c = 0.0;
        *s = conj(g) / ag;
        *r = ag;
        return;
    }
    const double af = cabs(f), ag = cabs(g);
    const double d = hypot(af, ag);
    const zc ph = f / af;
    *c = af / d;
    *s = ph * conj(g) / d;
    *r = ph * d;
}

static int has_nan(const double* a, int n)
{
    for (int i = 0; i < n; ++i)
        if (a[i] != a[i]) return 1;
    return 0;
}

/* InternalSolveWithGMRES (solve.hpp:241-611), one restarted GMRES(m) per
 * right-hand side sharing the operator applications.  `steps` > 0 runs a
 * fixed number of inner steps and records the residual history. */
static int gmres(const ora_ctx* X, zc* b, int k, int m, double relTol,
                 int maxIts, int steps, double* relHist, ora_report* rep)
{
    const long N = X->N;
    zc* V = (zc*)malloc(sizeof(zc) * (size_t)(N * k * m));
    zc* x0 = (zc*)malloc(sizeof(zc) * (size_t)(N * k));
    zc* x = (zc*)calloc((size_t)(N * k), sizeof(zc));
    zc* w = (zc*)malloc(sizeof(zc) * (size_t)(N * k));
    zc* tmp = (zc*)malloc(sizeof(zc) * (size_t)(N * k));
    zc* z = (zc*)malloc(sizeof(zc) * (size_t)((m + 1) * k));
    zc* H = (zc*)calloc((size_t)(m * m * k), sizeof(zc));
    double* cs = (double*)malloc(sizeof(double) * (size_t)(m * k));
    zc* sn = (zc*)malloc(sizeof(zc) * (size_t)(m * k));
    double* orig = (double*)malloc(sizeof(double) * (size_t)k);
    double* beta = (double*)malloc(sizeof(double) * (size_t)k);
    double* delta = (double*)malloc(sizeof(double) * (size_t)k);
    zc* alpha = (zc*)malloc(sizeof(zc) * (size_t)k);
    zc* y = (zc*)malloc(sizeof(zc) * (size_t)(m + 1));
    int status = ORA_OK;
    double tP = 0, tM = 0, tA = 0, t0;
    int nP = 0, nM = 0;
    int it = 0, restarts = 0, converged = 0;
    double maxRel = 0.0;

    memcpy(w, b, sizeof(zc) * (size_t)(N * k));
    for (int r = 0; r < k; ++r) orig[r] = nrm2(w + r * N, N);
    if (has_nan(orig, k)) {
        status = ORA_ERR_NAN; /* solve.hpp:280-282 */
        goto done;
    }
    while (!converged) {
        memcpy(x0, x, sizeof(zc) * (size_t)(N * k));
        t0 = omp_get_wtime();
        precondition(X, w, k);
        tP += omp_get_wtime() - t0;
        ++nP;
        for (int r = 0; r < k; ++r) beta[r] = nrm2(w + r * N, N);
        if (has_nan(beta, k)) {
            status = ORA_ERR_NAN;
            goto done;
        }
        for (int r = 0; r < k; ++r) {
            const double inv = 1.0 / beta[r];
            zc* v0 = V + (0 * k + r) * N;
            for (long i = 0; i < N; ++i) v0[i] = w[r * N + i] * inv;
        }
        memset(z, 0, sizeof(zc) * (size_t)((m + 1) * k));
        for (int r = 0; r < k; ++r) z[0 + r * (m + 1)] = beta[r];
        for (int j = 0; j < m; ++j) {
            memcpy(w, V + (long)j * k * N, sizeof(zc) * (size_t)(N * k));
            t0 = omp_get_wtime();
            multiply(X, w, k, tmp);
            tM += omp_get_wtime() - t0;
            ++nM;
            t0 = omp_get_wtime();
            precondition(X, w, k);
            tP += omp_get_wtime() - t0;
            ++nP;
            t0 = omp_get_wtime();
            /* Modified Gram-Schmidt (solve.hpp:382-399) */
            for (int i = 0; i <= j; ++i) {
#pragma omp parallel for schedule(static) if (k > 1)
                for (int r = 0; r < k; ++r) {
                    const zc* vi = V + ((long)i * k + r) * N;
                    zc* wr = w + r * N;
                    const zc a = dotc(vi, wr, N);
                    alpha[r] = a;
                    for (long q = 0; q < N; ++q) wr[q] -= a * vi[q];
                }
                for (int r = 0; r < k; ++r) H[i + (j + r * m) * m] = alpha[r];
            }
            for (int r = 0; r < k; ++r) delta[r] = nrm2(w + r * N, N);
            tA += omp_get_wtime() - t0;
            if (has_nan(delta, k)) {
                status = ORA_ERR_NAN; /* :401-403 */
                goto done;
            }
            for (int r = 0; r < k; ++r)
                if (delta[r] == 0.0) {
                    status = ORA_ERR_BREAKDOWN; /* :404-414 */
                    goto done;
                }
            if (j + 1 != m)
                for (int r = 0; r < k; ++r) {
                    const double inv = 1.0 / delta[r];
                    zc* vn = V + ((long)(j + 1) * k + r) * N;
                    for (long q = 0; q < N; ++q) vn[q] = w[r * N + q] * inv;
                }
            for (int r = 0; r < k; ++r) {
                zc* Hr = H + (long)r * m * m; /* H(i,j) = Hr[i + j*m] */
                /* previous rotations (:442-456) */
                for (int i = 0; i < j; ++i) {
                    const double c = cs[i + r * m];
                    const zc s = sn[i + r * m];
                    const zc e0 = Hr[i + j * m], e1 = Hr[i + 1 + j * m];
                    Hr[i + j * m] = c * e0 + s * e1;
                    Hr[i + 1 + j * m] = -conj(s) * e0 + c * e1;
                }
                /* new rotation (:472-503) */
                double c;
                zc s, rho;
                const zc ejj = Hr[j + j * m];
                const zc ejp = delta[r];
                if (ejj != ejj) {
                    status = ORA_ERR_NAN;
                    goto done;
                }
                givens(ejj, ejp, &c, &s, &rho);
                if (c != c || s != s || rho != rho) {
                    status = ORA_ERR_NAN;
                    goto done;
                }
                Hr[j + j * m] = rho;
                cs[j + r * m] = c;
                sn[j + r * m] = s;
                zc* zr = z + (long)r * (m + 1);
                const zc zj = zr[j], zj1 = zr[j + 1];
                zr[j] = c * zj + s * zj1;
                zr[j + 1] = -conj(s) * zj + c * zj1;
                /* y := H \ z (upper triangular, :506-509) */
                for (int i = 0; i <= j; ++i) y[i] = zr[i];
                for (int i = j; i >= 0; --i) {
                    zc acc = y[i];
                    for (int l = i + 1; l <= j; ++l) acc -= Hr[i + l * m] * y[l];
                    y[i] = acc / Hr[i + i * m];
                }
                /* x := x0 + V y (:512-521) */
                zc* xr = x + r * N;
                const zc* x0r = x0 + r * N;
                memcpy(xr, x0r, sizeof(zc) * (size_t)N);
                for (int i = 0; i <= j; ++i) {
                    const zc* vi = V + ((long)i * k + r) * N;
                    const zc eta = y[i];
                    for (long q = 0; q < N; ++q) xr[q] += eta * vi[q];
                }
            }
            /* w := b - A x (:529-546) */
            for (long q = 0; q < N * k; ++q) w[q] = -x[q];
            t0 = omp_get_wtime();
            multiply(X, w, k, tmp);
            tM += omp_get_wtime() - t0;
            ++nM;
            for (long q = 0; q < N * k; ++q) w[q] += b[q];
            maxRel = 0.0;
            for (int r = 0; r < k; ++r) {
                const double rn = nrm2(w + r * N, N);
                if (rn != rn) {
                    status = ORA_ERR_NAN;
                    goto done;
                }
                const double rel = rn / orig[r];
                if (rel > maxRel) maxRel = rel;
            }
            if (relHist && it < steps) relHist[it] = maxRel;
            ++it;
            if (steps > 0) {
                if (it >= steps) {
                    converged = 1;
                    break;
                }
            } else if (maxRel < relTol) {
                converged = 1;
                break;
            }
            if (maxIts > 0 && it >= maxIts) {
                status = ORA_ERR_NOCONV;
                goto done;
            }
        }
        ++restarts;
    }
done:
    memcpy(b, x, sizeof(zc) * (size_t)(N * k));
    if (rep) {
        rep->iterations = it;
        rep->restarts = restarts;
        rep->final_max_rel = maxRel;
        rep->t_precond = tP;
        rep->t_multiply = tM;
        rep->t_arnoldi = tA;
        rep->n_precond = nP;
        rep->n_multiply = nM;
    }
    free(V);
    free(x0);
    free(x);
    free(w);
    free(tmp);
    free(z);
    free(H);
    free(cs);
    free(sn);
    free(orig);
    free(beta);
    free(delta);
    free(alpha);
    free(y);
    return status;
}

double ora_time_arnoldi(const ora_ctx* X, int k, int m)
{
    const long N = X->N;
    zc* V = (zc*)malloc(sizeof(zc) * (size_t)(N * k * m));
    zc* w = (zc*)malloc(sizeof(zc) * (size_t)(N * k));
    zc* x = (zc*)malloc(sizeof(zc) * (size_t)(N * k));
    zc* alpha = (zc*)malloc(sizeof(zc) * (size_t)k);
#pragma omp parallel for schedule(static)
    for (long q = 0; q < N * k * m; ++q) V[q] = (double)((q * 2654435761u) % 1000) * 1e-3;
    for (long q = 0; q < N * k; ++q) w[q] = 1.0;
    const double t0 = omp_get_wtime();
    for (int j = 0; j < m; ++j) {
        for (int i = 0; i <= j; ++i) {
#pragma omp parallel for schedule(static) if (k > 1)
            for (int r = 0; r < k; ++r) {
                const zc* vi = V + ((long)i * k + r) * N;
                zc* wr = w + r * N;
                const zc a = dotc(vi, wr, N) * 1e-6;
                alpha[r] = a;
                for (long q = 0; q < N; ++q) wr[q] -= a * vi[q];
            }
        }
        for (int r = 0; r < k; ++r) alpha[r] = nrm2(w + r * N, N);
#pragma omp parallel for schedule(static) if (k > 1)
        for (int r = 0; r < k; ++r) {
            zc* xr = x + r * N;
            for (long q = 0; q < N; ++q) xr[q] = w[r * N + q];
            for (int i = 0; i <= j; ++i) {
                const zc* vi = V + ((long)i * k + r) * N;
                for (long q = 0; q < N; ++q) xr[q] += 1e-3 * vi[q];
            }
        }
    }
    const double t = omp_get_wtime() - t0;
    free(V);
    free(w);
    free(x);
    free(alpha);
    return t;
}

/* PullRightHandSides / PushRightHandSides (solve.hpp:58-239). */
static void pull(const ora_ctx* X, const zc* grid, zc* B, int k)
{
    for (long g = 0; g < X->N; ++g) {
        const long nat = X->l2n[g];
        for (int r = 0; r < k; ++r) B[g + r * X->N] = grid[nat * k + r];
    }
}

static void push(const ora_ctx* X, zc* grid, const zc* B, int k)
{
    for (long g = 0; g < X->N; ++g) {
        const long nat = X->l2n[g];
        for (int r = 0; r < k; ++r) grid[nat * k + r] = B[g + r * X->N];
    }
}

int ora_solve(const ora_ctx* X, double* Bgrid, int k, int m, double relTol,
              int maxIts, ora_report* rep)
{
    if (!X->factored) return ORA_ERR_STATE;
    if (k < 1 || m < 1) return ORA_ERR_INVALID;
    const double t0 = omp_get_wtime();
    zc* B = (zc*)malloc(sizeof(zc) * (size_t)(X->N * k));
    pull(X, (const zc*)Bgrid, B, k);
    const int st = gmres(X, B, k, m, relTol, maxIts, 0, NULL, rep);
    push(X, (zc*)Bgrid, B, k);
    free(B);
    if (rep) rep->t_total = omp_get_wtime() - t0;
    return st;
}

int ora_solve_fixed(const ora_ctx* X, double* Bgrid, int k, int m, int steps,
                    double* relHist)
{
    if (!X->factored) return ORA_ERR_STATE;
    if (k < 1 || m < 1 || steps < 1) return ORA_ERR_INVALID;
    zc* B = (zc*)malloc(sizeof(zc) * (size_t)(X->N * k));
    pull(X, (const zc*)Bgrid, B, k);
    const int st = gmres(X, B, k, m, 0.0, 0, steps, relHist, NULL);
    push(X, (zc*)Bgrid, B, k);
    free(B);
    return st;
}
