/*
 * psp_oracle.h -- CPU restatement of PSP's sweeping-preconditioned Helmholtz
 * solve (FP64, complex double), written from scratch in C.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product path and the CPU baseline timed beside it.  Nothing in the product
 * (psp_amd/) links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the PSP source tree).  Clique's multifrontal LDL^T, which PSP calls but does
 * not ship, is restated as a textbook unpivoted multifrontal LDL^T
 * (Duff & Reid 1983; Liu 1992) over PSP's own elimination tree.
 */
#ifndef PSP_ORACLE_H
#define PSP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary kinds, include/psp/discretization.hpp:15 */
enum { ORA_PML = 0, ORA_DIRICHLET = 1 };

/* Status codes (mirrors the product's C ABI codes). */
enum {
    ORA_OK = 0,
    ORA_ERR_INVALID = 1,   /* LogicError: bad argument                  */
    ORA_ERR_SHALLOW = 2,   /* LogicError: domain too shallow to sweep   */
    ORA_ERR_NAN = 3,       /* RuntimeError: NaN in a norm / rotation    */
    ORA_ERR_BREAKDOWN = 4, /* exact Arnoldi breakdown (delta == 0)      */
    ORA_ERR_NOCONV = 5,    /* iteration cap reached                     */
    ORA_ERR_STATE = 6      /* not initialized                           */
};

/* Discretization<R> (include/psp/discretization.hpp:36-110) plus the
 * DistHelmholtz constructor knobs (include/psp/dist_helmholtz.hpp:27-29). */
typedef struct {
    double omega;
    int nx, ny, nz;
    double wx, wy, wz;
    int frontBC, rightBC, backBC, leftBC, topBC;
    int bx, by, bz;
    double sigmax, sigmay, sigmaz;
    double damping;
    int planesPerPanel;
    int cutoff;
} ora_disc;

typedef struct {
    int iterations;          /* GMRES iterations (inner steps, summed)     */
    int restarts;
    double final_max_rel;    /* max_r ||b_r - A x_r|| / ||b_r||           */
    double t_total;          /* seconds: pull + GMRES + push               */
    double t_precond;        /* seconds spent in Precondition              */
    double t_multiply;       /* seconds spent in Multiply                  */
    double t_arnoldi;        /* seconds spent in MGS + norms               */
    int n_precond, n_multiply;
} ora_report;

typedef struct ora_ctx ora_ctx;

int  ora_create(const ora_disc* disc, ora_ctx** out);
int  ora_num_threads(void);  /* OpenMP threads the oracle runs on */
void ora_destroy(ora_ctx* ctx);

/* Geometry queries (main.hpp:35-57, 728-821). */
int  ora_num_panels(const ora_ctx* ctx);
int  ora_panel_info(const ora_ctx* ctx, int panel, int* vPanel, int* depth,
                    int* padding, int* cls);
long ora_global_size(const ora_ctx* ctx);

/* Global ND order -> natural index x + y*nx + z*nx*ny (main.hpp:823-1059). */
void ora_local_to_natural(const ora_ctx* ctx, int* map);

/* In-panel nested-dissection index (main.hpp:1425-1477). */
int  ora_reordered_index(const ora_ctx* ctx, int x, int y, int vLocal,
                         int vSize);

/* Class symbolic data.  cls: 0 = bottom/full inner, 1 = leftover, 2 = top. */
int  ora_class_vsize(const ora_ctx* ctx, int cls);
int  ora_class_num_nodes(const ora_ctx* ctx, int cls);
/* per node: off, size, parent, usize, depth */
void ora_class_nodes(const ora_ctx* ctx, int cls, int* off, int* size,
                     int* parent, int* usize);
/* concatenated struct indices of all nodes (node order) */
void ora_class_structs(const ora_ctx* ctx, int cls, int* structs);

/* Global operator in CSR (ND order, row entries in PSP's fixed order
 * [self, x-1, x+1, y-1, y+1, v-1, v+1]).  rowptr has N+1 entries; cols are
 * global ND indices; vals interleaved (re, im).  Needs ora_factor first. */
long ora_global_nnz(const ora_ctx* ctx);
int  ora_global_csr(const ora_ctx* ctx, int* rowptr, int* cols, double* vals);

/* Lower-triangular panel operator (padded ND order) as COO, from
 * FormLowerColumnOfNode.  Returns nnz when rows==NULL.  Needs a velocity. */
long ora_panel_lower_coo(ora_ctx* ctx, int panel, const double* velocity,
                         int* rows, int* cols, double* vals);

/* Initialize: factor all panels and build the global operator
 * (initialize_finalize.hpp:12-377).  velocity in natural order. */
int  ora_factor(ora_ctx* ctx, const double* velocity);
double ora_factor_seconds(const ora_ctx* ctx);
/* Factor only `count` panels starting at `first` (CPU baseline sampling). */
int  ora_factor_sample(ora_ctx* ctx, const double* velocity, int first,
                       int count, double* seconds);

/* Build only the global operator (the last block of Initialize), for CPU
 * baseline sampling. */
int  ora_fill_global(ora_ctx* ctx, const double* velocity);

/* Seconds for one restart cycle of the GMRES(m) Arnoldi work on k columns:
 * the modified Gram-Schmidt steps j = 0..m-1 plus the x := x0 + V y updates
 * (solve.hpp:382-399, 512-521), on random data (CPU baseline sampling). */
double ora_time_arnoldi(const ora_ctx* ctx, int k, int m);

/* Operators on a global ND-ordered, column-major N x k block (ld = N).
 * B interleaved complex (re, im). */
int  ora_multiply(const ora_ctx* ctx, double* B, int k);
int  ora_precondition(const ora_ctx* ctx, double* B, int k);
int  ora_solve_panel(const ora_ctx* ctx, double* B, int k, int panel);

/* Solve (solve.hpp:12-56): Bgrid in grid order, (x+y*nx+z*nx*ny)*k + r,
 * overwritten with the solution.  maxIts <= 0 means no cap. */
int  ora_solve(const ora_ctx* ctx, double* Bgrid, int k, int m, double relTol,
               int maxIts, ora_report* rep);

/* GMRES with a fixed number of inner steps (no convergence stop), for
 * parity at a fixed iteration count.  Returns the residual history. */
int  ora_solve_fixed(const ora_ctx* ctx, double* Bgrid, int k, int m,
                     int steps, double* relHist);

/* Dense-front factorization self-check helpers (tests only). */
int  ora_partial_ldl(double* F, int fs, int s);

#ifdef __cplusplus
}
#endif
#endif
