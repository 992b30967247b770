/*
 * psp_hip.h -- C ABI of the MI355X-native sweeping-preconditioned Helmholtz
 * solver.  Plain pointers and sizes only; no C++ or torch types cross it.
 *
 * The reference (PSP) exposes a header-only C++ template API and no C ABI.
 * Each entry point below names the reference method it replaces
 * (paths relative to the PSP source tree).  The C++ mirror of that API,
 * psp::DistHelmholtz<double> (psp_amd/include/psp/dist_helmholtz.hpp), is a
 * thin layer over these functions.
 *
 * Conventions
 *  - Complex data is interleaved (re, im) doubles, layout-compatible with
 *    std::complex<double>.
 *  - Right-hand-side blocks use the DistUniformGrid layout at one process
 *    (include/psp/dist_uniform_grid.hpp:322-341): entry (x, y, z, r) at
 *    ((x + y*nx + z*nx*ny) * k + r).
 *  - Functions never throw; they return a status code (PSP_HIP_OK == 0) and
 *    leave a message readable through psp_hip_last_error().
 *  - The context owns all device memory.  Calls on one context are
 *    serialized on the context's stream; use one context per host thread.
 */
#ifndef PSP_HIP_H
#define PSP_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
enum {
    PSP_HIP_OK = 0,
    PSP_HIP_ERR_INVALID = 1,   /* LogicError: invalid argument            */
    PSP_HIP_ERR_SHALLOW = 2,   /* LogicError: domain too shallow (main.hpp:38-41) */
    PSP_HIP_ERR_NAN = 3,       /* RuntimeError: NaN in norms (solve.hpp:280-282, 401-403) */
    PSP_HIP_ERR_BREAKDOWN = 4, /* exact Arnoldi breakdown (solve.hpp:404-414) */
    PSP_HIP_ERR_NOCONV = 5,    /* iteration cap reached (the reference has none) */
    PSP_HIP_ERR_STATE = 6,     /* called before psp_hip_factor           */
    PSP_HIP_ERR_HIP = 7,       /* HIP runtime error / no device          */
    PSP_HIP_ERR_MEMORY = 8     /* device allocation failed               */
};

/* Boundary (include/psp/discretization.hpp:15). */
enum { PSP_HIP_PML = 0, PSP_HIP_DIRICHLET = 1 };

/* Front precision: FP64 fronts reproduce PSP's method; FP32 fronts are
 * factored and applied in complex float inside the FP64 GMRES. */
enum { PSP_HIP_FRONTS_FP64 = 0, PSP_HIP_FRONTS_FP32 = 1 };

/* PanelScheme (include/psp/dist_helmholtz.hpp:15-19). */
enum { PSP_HIP_LDL_1D = 0, PSP_HIP_LDL_SELINV_1D = 1 };

/* Options for psp_hip_set_option. */
enum {
    PSP_HIP_OPT_SWEEP_GRAPH = 1,     /* 1 (default): replay the sweep from a captured
                                        hipGraph; 0: launch it kernel by kernel */
    PSP_HIP_OPT_SOLVE_MIN_TILES = 2, /* front-solve workgroups wanted per launch when
                                        choosing default tile shapes (default 256) */
    PSP_HIP_OPT_SOLVE_MAX_SPLIT = 3, /* most inner slices per row block in the
                                        default tile shapes (default 8) */
    PSP_HIP_OPT_SOLVE_WIDE_WG = 4,   /* 1: default shapes give levels with few row
                                        blocks 16-wave workgroups instead of inner
                                        slices (8 <= k < 32); 0 (default): slices */
    PSP_HIP_OPT_SOLVE_PERSIST = 5,   /* 1 (default): default shapes run the front
                                        solves with persistent, pipelined
                                        workgroups (8 <= k <= 16); 0: one
                                        workgroup per tile */
    PSP_HIP_OPT_SOLVE_PERSIST_GRID = 6, /* persistent launches use at most this many
                                          workgroups (more only when one would
                                          get over 32 tiles); 0 (default): as many
                                          as fit on the device at once */
    PSP_HIP_OPT_SOLVE_FUSED = 7,       /* 1: for 8 <= k <= 16, the tree levels from
                                          the cut depth down run as one persistent
                                          launch per direction, a subtree queue per
                                          XCD; 0 (default): one launch per level */
    PSP_HIP_OPT_SOLVE_FUSED_QUEUES = 8, /* queues of the fused launches (default 8,
                                          one per XCD; more leaves queues that the
                                          last workgroup runs alone) */
    PSP_HIP_OPT_SOLVE_FUSED_ROWS = 9,  /* output rows per fused tile: 16, 32 (default)
                                          or 64 */
    PSP_HIP_OPT_SOLVE_FUSED_CUT = 10,  /* cut depth of the fused launches; -1 (default):
                                          the first depth with 8 nodes */
    PSP_HIP_OPT_SOLVE_FUSED_WG = 11,   /* workgroups a fused queue expects (default 64):
                                          near the cut, row blocks are cut into inner
                                          slices until a level has about that many
                                          tiles per queue; 1: no slices */
    PSP_HIP_OPT_SOLVE_XCD_REDUCE = 12, /* 1 (default): the inner slices of a row block
                                          run on one XCD and the last to finish sums
                                          them in the product launch (one-workgroup-
                                          per-tile levels, 8 <= k < 32), when the
                                          device deals workgroups to its XCDs
                                          round-robin (checked at set-up); 0: a
                                          reduce launch follows the product launch */
    PSP_HIP_OPT_FACTOR_OVERLAP = 13    /* 1 (default): Initialize factors the leftover
                                          and top panels on their own streams beside
                                          the batch of full panels; 0: one class after
                                          another on the context's stream (the Schur
                                          timings of the report then time the GEMM
                                          launches alone) */
};

/* Flags for psp_hip_solve. */
enum {
    PSP_HIP_DEVICE_POINTER = 1, /* B is a device pointer (HBM-resident)   */
    PSP_HIP_BATCH_THREADS = 2   /* psp_hip_solve_batch: one host thread and
                                   stream per problem instead of the merged
                                   sweep launches */
};

/* Discretization<double> (include/psp/discretization.hpp:36-110) plus the
 * DistHelmholtz constructor knobs (include/psp/dist_helmholtz.hpp:27-29). */
typedef struct {
    double omega;
    int nx, ny, nz;
    double wx, wy, wz;
    int frontBC, rightBC, backBC, leftBC, topBC;
    int bx, by, bz;
    double sigmax, sigmay, sigmaz;
    double damping;       /* default 7.0 */
    int planesPerPanel;   /* default 4   */
    int cutoff;           /* default 12  */
    int precision;        /* PSP_HIP_FRONTS_FP64 or PSP_HIP_FRONTS_FP32 */
} psp_hip_desc;

/* Timings mirroring the per-panel gather/fill/ldl report of
 * initialize_finalize.hpp:153-165. */
typedef struct {
    double seconds_total;    /* wall time of psp_hip_factor               */
    double seconds_fill;     /* front assembly (FillPanelFronts)          */
    double seconds_ldl;      /* multifrontal LDL^T                        */
    double seconds_inverse;  /* diagonal-block inversion (SELINV)         */
    double flops;            /* algorithmic real flops, 8 per complex MAC */
    double schur_flops;      /* subset spent in Schur-complement GEMMs    */
    double bytes_factor;     /* bytes of stored factors                   */
    int num_panels;
    int num_fronts;          /* fronts over all panels                    */
    double seconds_schur;    /* GPU time of the Schur-complement GEMMs    */
} psp_hip_factor_report;

typedef struct {
    int status;
    int iterations;          /* GMRES inner steps over all restarts       */
    int restarts;
    double final_max_rel;    /* max_r ||b_r - A x_r|| / ||b_r||           */
    double seconds_total;
    double seconds_precond;
    double seconds_multiply;
    double seconds_arnoldi;
    int num_precond, num_multiply;
} psp_hip_solve_report;

/* Front-solve product kernels of one sweep, timed launch by launch. */
typedef struct {
    int launches;        /* kernel launches timed                          */
    double seconds;      /* summed HIP-event durations of those launches   */
    double bytes;        /* algorithmic bytes: factor entries + vectors    */
    double flops;        /* algorithmic real flops (8 per complex MAC)     */
} psp_hip_kernel_stats;

typedef struct psp_hip_ctx psp_hip_ctx;

/* DistHelmholtz ctor (dist_helmholtz/main.hpp:12-436): panel split, nested
 * dissection and symbolic analysis.  Host work only; no device memory is
 * touched until psp_hip_factor. */
int psp_hip_create(const psp_hip_desc* desc, psp_hip_ctx** out);

/* DistHelmholtz::Finalize (initialize_finalize.hpp:379-396): frees the
 * factors and the solve workspace; the plan stays, so a later
 * psp_hip_factor (Initialize) rebuilds them.  Solve in between is a
 * PSP_HIP_ERR_STATE error. */
int psp_hip_finalize(psp_hip_ctx* ctx);

/* ~DistHelmholtz: frees everything, including the context. */
int psp_hip_destroy(psp_hip_ctx* ctx);

/* viewIterates (solve.hpp:579-585): with a callback set, every GMRES step
 * that has not converged hands the current iterate x (host copy, k
 * right-hand sides in grid layout) and the step index to fn before the next
 * step; NULL removes it. */
typedef void (*psp_hip_iterate_fn)(void* user, int iteration, const void* x, int k);
int psp_hip_set_iterate_callback(psp_hip_ctx* ctx, psp_hip_iterate_fn fn, void* user);

/* Message of the last failing call on this context (never NULL). */
const char* psp_hip_last_error(const psp_hip_ctx* ctx);

/* Use `stream` (a hipStream_t, or NULL for the context's own stream) for all
 * later device work of this context. */
int psp_hip_set_stream(psp_hip_ctx* ctx, void* stream);

/* Set a context option (PSP_HIP_OPT_*); no reference counterpart. */
int psp_hip_set_option(psp_hip_ctx* ctx, int option, int value);

/* DistHelmholtz::Initialize (initialize_finalize.hpp:12-377): velocity is a
 * host array of nx*ny*nz doubles in natural order x + y*nx + z*nx*ny.
 * panelScheme: PSP_HIP_LDL_1D or PSP_HIP_LDL_SELINV_1D. */
int psp_hip_factor(psp_hip_ctx* ctx, const double* velocity, int panelScheme,
                   psp_hip_factor_report* report);

/* DistHelmholtz::Solve (solve.hpp:12-56) = Pull, InternalSolveWithGMRES
 * (solve.hpp:241-611), Push.  B holds k right-hand sides in grid layout and
 * is overwritten with the solutions.  flags & PSP_HIP_DEVICE_POINTER: B is
 * in device memory.  maxIts <= 0: no cap (as the reference). */
int psp_hip_solve(psp_hip_ctx* ctx, void* B, int k, int m, double relTol,
                  int maxIts, int flags, psp_hip_solve_report* report);

/* DistHelmholtz::Multiply (solve.hpp:613-679) on device blocks in grid
 * layout: y := A x. */
int psp_hip_multiply(psp_hip_ctx* ctx, const void* x, void* y, int k);

/* DistHelmholtz::Precondition (solve.hpp:681-734) in place on a device
 * block in grid layout. */
int psp_hip_precondition(psp_hip_ctx* ctx, void* B, int k);

/* Precondition without allocating (no reference counterpart: for callers
 * that cannot allocate device memory, such as code under HIP graph capture
 * or serving loops that forbid it).  Same result as psp_hip_precondition,
 * with the sweep's buffers in a fixed 64 MiB device array of the library
 * (one per device, held until the process ends) instead of the context's
 * workspace; calls no hipMalloc or hipFree and replays no graph.  A grid and
 * k whose sweep needs more than 64 MiB is PSP_HIP_ERR_INVALID.  Calls that
 * overlap on one device share the array and race: serialize them. */
int psp_hip_precondition_noalloc(psp_hip_ctx* ctx, void* B, int k);

/* DistHelmholtz::SolvePanel (solve.hpp:736-882) in place on a device block
 * in grid layout (only panel `panel`'s planes are read and written). */
int psp_hip_solve_panel(psp_hip_ctx* ctx, void* B, int k, int panel);

/* One Precondition on a device block, launched kernel by kernel with HIP
 * events around every front-solve product launch (no reference
 * counterpart; measurement only). */
int psp_hip_profile_sweep(psp_hip_ctx* ctx, void* B, int k,
                          psp_hip_kernel_stats* stats);

/* Tile-shape tuner for the front solves (no reference counterpart): for
 * each panel class, tree level and direction, times every candidate shape
 * (rows per tile 64/32/16, inner slices 1/2/4/8) with k right-hand sides on
 * the class's first panel and keeps the fastest; the chosen shapes stay for
 * later solves of this context.  before/after (may be NULL) receive the
 * estimated front-solve time of one sweep with the old and new shapes.
 * Needs a factored context. */
int psp_hip_autotune(psp_hip_ctx* ctx, int k, double* before, double* after);

/* Many independent solves at once (no reference counterpart; for callers
 * with batches of uneven problems): solve i gives the result of
 * psp_hip_solve(ctxs[i], Bs[i], ks[i], m, relTol, maxIts, flags,
 * &reports[i]), each problem's GMRES (solve.hpp:241-611) stepping as it
 * would alone.  Default: the GMRES steps of all problems in lockstep on the
 * first context's stream, and the sweeps of all unfinished problems merged:
 * the problems form two groups of about equal work, and one launch per
 * sweep step (a panel copy, or one tree level's products or reduce) covers
 * that step of every problem of a group; the two groups' launch chains are
 * parallel branches of one hipGraph, replayed for every sweep; one host
 * wait per GMRES step for all problems.  The front products of a
 * problem then run on the one-workgroup-per-tile matrix-core kernel with
 * its tuned tile rows, slices and waves (persistent tiles included), so a
 * problem with k < 8 or k >= 32 sums in another order than alone (results
 * agree to rounding; they are bitwise equal for
 * 8 <= k < 32).  Problems with FP32 and FP64 fronts form two such batches,
 * run in turn.  flags & PSP_HIP_BATCH_THREADS: solve i runs
 * psp_hip_solve on its own host thread and its context's stream instead.
 * Contexts must be distinct.  reports may be NULL (batch-wide phase times
 * and total time in each report with the merged launches).  Returns the
 * first failing status, or OK. */
int psp_hip_solve_batch(psp_hip_ctx* const* ctxs, void* const* Bs, const int* ks, int n,
                        int m, double relTol, int maxIts, int flags,
                        psp_hip_solve_report* reports);

/* GMRES with a fixed number of inner steps, for parity at a fixed
 * iteration count; relHist (host, steps doubles) gets max_r rel. residual. */
int psp_hip_solve_fixed(psp_hip_ctx* ctx, void* B, int k, int m, int steps,
                        int flags, double* relHist);

/* Adjoint-state gradient of L = 1/2 sum_s sum_r |u_s(x_r) - d_s(x_r)|^2 with
 * respect to squared slowness m = 1/c^2 (no reference counterpart; the
 * extension uses the diagonal term of FormGlobalRow,
 * initialize_finalize.hpp:567-568).  u: device k source solutions (grid
 * layout); lambda: device adjoint solutions; grad (device, nx*ny*nz doubles)
 * receives sum_s Re(omega^2 s1 s2 s3 lambda_s u_s), accumulated if
 * accumulate != 0. */
int psp_hip_adjoint_gradient(psp_hip_ctx* ctx, const void* u,
                             const void* lambda, int k, double* grad,
                             int accumulate);

/* Geometry queries (main.hpp:35-57, 728-821); host only. */
int psp_hip_num_panels(const psp_hip_ctx* ctx);
int psp_hip_panel_info(const psp_hip_ctx* ctx, int panel, int* vPanel,
                       int* depth, int* padding, int* cls);
/* Symbolic analysis of panel class cls (0 bottom/inner, 1 leftover, 2 top):
 * per node offset, size, parent, struct size (arrays of
 * psp_hip_class_num_nodes entries, any may be NULL) and the concatenated
 * sorted structures. */
int psp_hip_class_num_nodes(const psp_hip_ctx* ctx, int cls);
int psp_hip_class_nodes(const psp_hip_ctx* ctx, int cls, int* off, int* size,
                        int* parent, int* usize, int* structs);
/* In-panel nested-dissection index (main.hpp:1425-1477). */
int psp_hip_reordered_index(const psp_hip_ctx* ctx, int x, int y, int vLocal,
                            int vSize);

/* Device inventory string (host). */
int psp_hip_device_count(void);

/* ---- Grid container I/O (host only, no device needed) -------------------
 * DistUniformGrid<F> at one process stores numScalars values per point,
 * points in the grid's GridDataOrder (the fastest axis first), so point
 * (x, y, z) starts at element psp_hip_grid_index(...) * numScalars.
 * Scalars are doubles (PSP_HIP_REAL64) or interleaved complex doubles
 * (PSP_HIP_COMPLEX128).  Errors: PSP_HIP_ERR_INVALID, message in
 * psp_hip_grid_last_error() (per host thread). */

/* GridDataOrder (include/psp/dist_uniform_grid.hpp:15-22). */
enum { PSP_HIP_XYZ = 0, PSP_HIP_XZY = 1, PSP_HIP_YXZ = 2, PSP_HIP_YZX = 3, PSP_HIP_ZXY = 4,
       PSP_HIP_ZYX = 5 };
/* PlaneType (include/psp/dist_uniform_grid.hpp:24-28). */
enum { PSP_HIP_PLANE_XY = 0, PSP_HIP_PLANE_XZ = 1, PSP_HIP_PLANE_YZ = 2 };
/* Scalar type of a grid. */
enum { PSP_HIP_REAL64 = 0, PSP_HIP_COMPLEX128 = 1 };

const char* psp_hip_grid_last_error(void);

/* DistUniformGrid::LocalIndex / numScalars (dist_uniform_grid.hpp:302-342)
 * at one process; -1 for an out-of-range point or order. */
long long psp_hip_grid_index(int nx, int ny, int nz, int order, int x, int y, int z);

/* DistUniformGrid::SequentialLoad (dist_uniform_grid.hpp:349-482): fills
 * data from a raw binary file holding the grid's points in its order,
 * numScalars native-endian scalars each.  A missing or short file is an
 * error ("Could not open <file>"). */
int psp_hip_grid_load(const char* filename, int nx, int ny, int nz, int numScalars, int type,
                      int order, void* data);

/* Copy between two storage orders of one grid (no reference counterpart:
 * the C++ and Python grids use it to hand natural-order XYZ data to the
 * solver). */
int psp_hip_grid_reorder(const void* in, int inOrder, void* out, int outOrder, int nx, int ny,
                         int nz, int numScalars, int type);

/* DistUniformGrid::InterpolateTo (dist_uniform_grid.hpp:484-714): trilinear
 * resampling of an nx0 x ny0 x nz0 grid onto nx x ny x nz points spanning
 * the same box (both in `order`). */
int psp_hip_grid_interpolate(const void* in, int nx0, int ny0, int nz0, void* out, int nx, int ny,
                             int nz, int numScalars, int type, int order);

/* DistUniformGrid::WritePlane (dist_uniform_grid.hpp:716-1629): ascii VTK
 * ImageData files basename_<k>.vti (real) or basename_<k>_real.vti and
 * basename_<k>_imag.vti (complex), one per scalar k. */
int psp_hip_grid_write_plane(const void* data, int nx, int ny, int nz, int numScalars, int type,
                             int order, int planeType, int whichPlane, const char* basename);

/* DistUniformGrid::WriteVolume (dist_uniform_grid.hpp:1786-2198) at one
 * process: basename_<k>[_real|_imag].pvti plus the piece
 * basename_<k>[_real|_imag]/0.vti. */
int psp_hip_grid_write_volume(const void* data, int nx, int ny, int nz, int numScalars, int type,
                              int order, const char* basename);

#ifdef __cplusplus
}
#endif
#endif /* PSP_HIP_H */
