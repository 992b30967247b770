/* Drives the CPU oracle (oracle/psp_oracle.c) through its whole path for
 * one small grid, built with AddressSanitizer and UndefinedBehaviorSanitizer
 * by tests/test_sanitizers.py: construction and symbolic analysis, every
 * query, the global CSR and panel COO assembly, the factorization, Multiply,
 * Precondition, single-panel solves, fixed-step and converged GMRES, and
 * the dense partial LDL^T.  Any invalid access or undefined operation aborts
 * the run; the exit status is nonzero when a step fails.
 *   sanitize_oracle nx ny nz k */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "psp_oracle.h"

#define CHECK(cond, what)                                    \
    do {                                                     \
        if (!(cond)) {                                       \
            fprintf(stderr, "sanitize_oracle: %s\n", what);  \
            return 1;                                        \
        }                                                    \
    } while (0)

static double urand(unsigned long long* s)
{
    *s = *s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(*s >> 11) / 9007199254740992.0;
}

int main(int argc, char** argv)
{
    const int nx = argc > 1 ? atoi(argv[1]) : 12, ny = argc > 2 ? atoi(argv[2]) : 10;
    const int nz = argc > 3 ? atoi(argv[3]) : 24, k = argc > 4 ? atoi(argv[4]) : 3;
    const int nmax = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
    const double h = 1.0 / (nmax + 1);
    ora_disc d;
    memset(&d, 0, sizeof d);
    d.nx = nx, d.ny = ny, d.nz = nz;
    d.wx = (nx + 1) * h, d.wy = (ny + 1) * h, d.wz = (nz + 1) * h;
    d.omega = 2 * M_PI * 1.5 * nmax / (10.0 * (nmax + 1) * h);
    d.frontBC = d.rightBC = d.backBC = d.leftBC = d.topBC = ORA_PML;
    d.bx = d.by = d.bz = 5;
    d.sigmax = d.sigmay = d.sigmaz = 1.5;
    d.damping = 7.0;
    d.planesPerPanel = 4;
    d.cutoff = 12;

    ora_ctx* ctx = NULL;
    CHECK(ora_create(&d, &ctx) == ORA_OK, "create");
    const long n = ora_global_size(ctx);
    CHECK(n == (long)nx * ny * nz, "global size");

    /* queries */
    int* map = malloc(sizeof(int) * n);
    ora_local_to_natural(ctx, map);
    for (long i = 0; i < n; ++i) CHECK(map[i] >= 0 && map[i] < n, "local-to-natural map");
    free(map);
    for (int p = 0; p < ora_num_panels(ctx); ++p) {
        int vp = 0, depth = 0, pad = 0, cls = 0;
        CHECK(ora_panel_info(ctx, p, &vp, &depth, &pad, &cls) == ORA_OK, "panel info");
    }
    for (int c = 0; c < 3; ++c) {
        const int nn = ora_class_num_nodes(ctx, c);
        if (nn <= 0) continue;
        int *off = malloc(sizeof(int) * nn), *size = malloc(sizeof(int) * nn);
        int *parent = malloc(sizeof(int) * nn), *usize = malloc(sizeof(int) * nn);
        ora_class_nodes(ctx, c, off, size, parent, usize);
        long nst = 0;
        for (int t = 0; t < nn; ++t) nst += usize[t];
        int* structs = malloc(sizeof(int) * (nst + 1));
        ora_class_structs(ctx, c, structs);
        free(off), free(size), free(parent), free(usize), free(structs);
    }

    /* a smooth velocity with a few bumps */
    double* vel = malloc(sizeof(double) * n);
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x)
                vel[x + (long)y * nx + (long)z * nx * ny] =
                    2.5 + 0.8 * sin(0.3 * x) * cos(0.2 * y) + 0.5 * sin(0.15 * z);

    /* factor, apply, sweep, solve */
    CHECK(ora_factor(ctx, vel) == ORA_OK, "factor");
    /* assembly (the global operator exists once factored) */
    const long nnz = ora_global_nnz(ctx);
    CHECK(nnz > 0, "global nnz");
    CHECK(ora_fill_global(ctx, vel) == ORA_OK, "fill global");
    int* rowptr = malloc(sizeof(int) * (n + 1));
    int* cols = malloc(sizeof(int) * nnz);
    double* vals = malloc(sizeof(double) * 2 * nnz);
    CHECK(ora_global_csr(ctx, rowptr, cols, vals) == ORA_OK, "global csr");
    CHECK(rowptr[n] == nnz, "csr row pointers");
    free(rowptr), free(cols), free(vals);
    {
        const long m = ora_panel_lower_coo(ctx, 0, vel, NULL, NULL, NULL);
        CHECK(m > 0, "panel coo size");
        int *ri = malloc(sizeof(int) * m), *ci = malloc(sizeof(int) * m);
        double* vv = malloc(sizeof(double) * 2 * m);
        CHECK(ora_panel_lower_coo(ctx, 0, vel, ri, ci, vv) == m, "panel coo");
        free(ri), free(ci), free(vv);
    }

    unsigned long long seed = 12345;
    double* B = malloc(sizeof(double) * 2 * n * k);
    double* C = malloc(sizeof(double) * 2 * n * k);
    for (long i = 0; i < 2 * n * k; ++i) B[i] = urand(&seed) - 0.5;
    memcpy(C, B, sizeof(double) * 2 * n * k);
    CHECK(ora_multiply(ctx, C, k) == ORA_OK, "multiply");
    memcpy(C, B, sizeof(double) * 2 * n * k);
    CHECK(ora_precondition(ctx, C, k) == ORA_OK, "precondition");
    for (long i = 0; i < 2 * n * k; ++i) CHECK(isfinite(C[i]), "finite sweep");
    for (int p = 0; p < ora_num_panels(ctx); ++p) {
        memcpy(C, B, sizeof(double) * 2 * n * k);
        CHECK(ora_solve_panel(ctx, C, k, p) == ORA_OK, "panel solve");
    }
    double hist[3];
    memcpy(C, B, sizeof(double) * 2 * n * k);
    CHECK(ora_solve_fixed(ctx, C, k, 4, 3, hist) == ORA_OK, "fixed-step GMRES");
    ora_report rep;
    memcpy(C, B, sizeof(double) * 2 * n * k);
    CHECK(ora_solve(ctx, C, k, 20, 1e-8, 300, &rep) == ORA_OK, "GMRES solve");
    CHECK(rep.final_max_rel < 1e-8, "GMRES residual");
    CHECK(ora_time_arnoldi(ctx, k, 4) >= 0.0, "arnoldi timing");

    /* dense partial LDL^T of a diagonally dominant complex symmetric block */
    {
        const int fs = 9, s = 5;
        double* F = calloc(2 * fs * fs, sizeof(double));
        for (int j = 0; j < fs; ++j)
            for (int i = j; i < fs; ++i) {
                F[2 * (i + j * fs)] = i == j ? 10.0 + i : urand(&seed) - 0.5;
                F[2 * (i + j * fs) + 1] = urand(&seed) - 0.5;
            }
        CHECK(ora_partial_ldl(F, fs, s) == ORA_OK, "partial ldl");
        free(F);
    }

    printf("sanitize_oracle %dx%dx%d k=%d: %d iterations, rel %.2e\n", nx, ny, nz, k,
           rep.iterations, rep.final_max_rel);
    free(B), free(C), free(vel);
    ora_destroy(ctx);
    return 0;
}
