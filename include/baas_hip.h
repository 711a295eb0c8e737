/*
 * baas_hip.h -- C-ABI of the MI355X (gfx950) sparse triangular pipeline.
 *
 * Plain pointers and sizes only.  Every numeric pointer is a DEVICE pointer
 * unless the comment says host.  Indices are 0-based int32; rows are sorted
 * ascending by column; a lower-triangular CSR keeps its diagonal last in each
 * row, an upper-triangular CSR keeps it first (the conventions of the
 * reference's SPM::CSR, spm/include/CSR.hpp:16-27).
 *
 * Ownership: the caller owns every input and output buffer; the library owns
 * handles, schedules, plans, ready flags and scratch inside the opaque
 * objects and frees them in the matching destroy call.
 * Threading: one handle = one HIP stream; a handle is not thread-safe, many
 * handles may run concurrently.  Numeric calls are asynchronous on the
 * handle's stream; baas_check() synchronises and returns the first error a
 * kernel raised (zero pivot, watchdog timeout) since the previous check.
 * Ready flags are reset by epoch per call, so back-to-back calls and graph
 * replays need no memset.
 */
#ifndef BAAS_HIP_H
#define BAAS_HIP_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct baas_handle baas_handle;
typedef struct baas_schedule baas_schedule;
typedef struct baas_pcg baas_pcg;
typedef struct baas_batch baas_batch;

enum baas_status {
    BAAS_OK = 0,
    BAAS_INVALID_ARG = 1,      /* bad size, unsorted row, missing diagonal, wrong kind */
    BAAS_ZERO_PIVOT = 2,       /* ILU0 pivot |u_jj| <= 1e-16 (the reference asserts) */
    BAAS_NONPOSITIVE_PIVOT = 3,/* IC0 A_ii - sum L_ik^2 <= 0; diagonal kept as the reference does */
    BAAS_HIP_ERROR = 4,
    BAAS_TIMEOUT = 5,          /* a ready-flag wait exceeded the watchdog spin limit */
    BAAS_NOT_CONVERGED = 6,
    BAAS_ALLOC = 7,
    BAAS_UNSUPPORTED = 8,
    BAAS_TOO_LARGE = 9         /* beyond the fixed scratch of baas_sptrsv_f64_noalloc */
};

/* Which numeric kernel a schedule is analysed for; selects the threads-per-
 * level rule (merge/LevelMerge.cpp:118-161) and, for IC0/ILU0, extraction of
 * the lower pattern from a full matrix (test/ICRuntime.hpp:416-435). */
enum baas_kind { BAAS_KIND_TRSV = 0, BAAS_KIND_IC0 = 1, BAAS_KIND_ILU0 = 2 };

/* Arrays a schedule exports for bit-exact comparison with the reference. */
enum baas_export {
    BAAS_EXPORT_GROUP_PTR = 0,     /* int[ngroups+1]  Coarsen groups (merge/Coarsen.cpp) */
    BAAS_EXPORT_GROUP_SET = 1,     /* int[n]          nodes, ascending inside a group */
    BAAS_EXPORT_LEVEL_PTR = 2,     /* int[nlevels+1]  group level set (spm/DAG.cpp:572-629) */
    BAAS_EXPORT_PERM_TO_ORIG = 3,  /* int[ngroups]    groups ordered by (level, id) */
    BAAS_EXPORT_NODE_TO_LEVEL = 4, /* int[ngroups]    1-based level of each group */
    BAAS_EXPORT_NTHREADS = 5,      /* int[nlevels]    threads per level (LevelMerge.cpp) */
    BAAS_EXPORT_MERGED_PTR = 6,    /* int[nmerged+1]  1-thread-run merge of level_ptr */
    BAAS_EXPORT_ROW_LEVEL_PTR = 7, /* int[nrowlevels+1] row-level level set of L */
    BAAS_EXPORT_ROW_PERM = 8,      /* int[n]          rows ordered by (row level, id) */
    BAAS_EXPORT_ROW_LEVEL = 9,     /* int[n]          1-based row level */
    BAAS_EXPORT_REDUCED_PTR = 10,  /* int[n+1]        two-hop reduced CSC DAG (spm/DAG.cpp:295-344) */
    BAAS_EXPORT_REDUCED_SET = 11,  /* int[edges] */
    BAAS_EXPORT_UNIT_PTR = 12,     /* int[nunits+1]   execution plan: rows per work unit */
    BAAS_EXPORT_UNIT_ROWS = 13,    /* int[n]          execution plan: row order */
    BAAS_EXPORT_GROUP_COST = 14    /* int[ngroups]    nonzeros of each group's rows in the lower
                                    *                 CSR (costComputation, merge/Coarsen.cpp:854-891;
                                    *                 the reference keeps them as doubles) */
};

/* Execution-plan choice (the autotuner fills these when mode = AUTO). */
/* LEVELS: one persistent launch walks the row level sets, all workgroups
 * sharing each level, with a device-wide barrier between levels: the GPU
 * form of the reference's level-set solve (merge/SpTRSV.cpp:41-80), best
 * when levels are few and wide. */
/* TILES: work units are tiles of the (row level, row id) plane -- a band of
 * consecutive levels by a range of consecutive row ids, band by band -- so
 * on matrices numbered along their geometry (natural stencil orderings,
 * banded and RCM-ordered matrices) a unit is a compact block of the domain
 * and a dependency chain crosses few units.  rows_per_unit is the row range;
 * the band height is set with baas_replan_tiles (default 16 levels).
 * BOXES: for a matrix whose lower triangle has at most three distinct
 * offsets (row - column) 1 < sy < sz with sy dividing sz -- a 5- or
 * 7-point stencil grid in natural order, recognised from the pattern alone
 * -- units are boxes of the grid of about rows_per_unit rows (default 512);
 * other matrices get BAAS_UNSUPPORTED.
 * PIPE: boxes of such a grid shaped for one wavefront each -- 8 x 8 rows
 * across (x, y) by rows_per_unit / 64 planes along z on a 3D grid, 64 rows
 * along x by rows_per_unit / 64 rows along y on a 2D grid -- which the
 * wavefront sweeps as a diagonal pipeline (every lane solves one row per
 * step and hands its value to its neighbours' lanes); also requires every
 * parent to lie in the grid direction its offset names (no wrap-around).
 * Single right-hand-side sweeps with dataflow 5 run that kernel; other
 * sweeps and factorisations run these boxes on the other kernels.
 * For LEVELS plans rows_per_unit caps the workgroups (0: as many as the
 * widest level needs); with one workgroup the levels are separated by the
 * workgroup's own barrier instead of a device-wide one. */
enum baas_plan_mode {
    BAAS_PLAN_AUTO = 0,
    BAAS_PLAN_RANGE = 1,
    BAAS_PLAN_SCHEDULE = 2,
    BAAS_PLAN_LEVELS = 3,
    BAAS_PLAN_TILES = 4,
    BAAS_PLAN_BOXES = 5,
    BAAS_PLAN_PIPE = 6
};

typedef struct baas_opts {
    int ncore;            /* thread count the reference rule is evaluated at (default 8) */
    int plan_mode;        /* baas_plan_mode */
    int rows_per_unit;    /* rows per work unit (0 = default for the mode) */
    int block_threads;    /* threads per workgroup (0 = default 256) */
    int merge_threshold;  /* schedule mode: merged 1-thread levels up to this many rows
                             stay in one unit (0 = default) */
} baas_opts;

/* ---- handles -------------------------------------------------------------- */
/* stream: a hipStream_t (NULL = the default stream). */
int baas_create(baas_handle **h, void *stream);
int baas_destroy(baas_handle *h);
int baas_set_stream(baas_handle *h, void *stream);
/* Synchronise the handle's stream, return and clear the first kernel error. */
int baas_check(baas_handle *h);
const char *baas_status_string(int status);
/* Watchdog spin limit for ready-flag waits (default 1<<22 polls). */
int baas_set_spin_limit(baas_handle *h, long long polls);

/* ---- analysis (replaces the runtimes' buildset(), e.g.
 * test/SpTRSVRuntime.hpp:675-775, test/ICRuntime.hpp:411-531,
 * test/ILURuntime.hpp:265-385, test/Schedule.hpp:181-296) ------------------- */
/* rowptr/colidx: device arrays of an n x n CSR.  kind TRSV: the lower-
 * triangular L itself.  kind IC0/ILU0: the full A (its lower part is
 * extracted; the threads-per-level rule counts rows of A as the reference
 * runtimes do).  Synchronous.  The schedule keeps copies of what it needs. */
int baas_analyze(baas_handle *h, int n, const int *rowptr, const int *colidx, int kind,
                 const baas_opts *opts, baas_schedule **s);
int baas_schedule_destroy(baas_schedule *s);
/* counts[0..7] = ngroups, nlevels, nmerged, reduced_edges, nrowlevels, nunits,
 * nnz_lower, plan_mode.  Host pointer. */
int baas_schedule_info(const baas_schedule *s, int *counts);
/* Copy one exported array to host memory; out == NULL only queries *len. */
int baas_schedule_export(const baas_schedule *s, int which, int *out, long long *len);
/* Seconds the analysis took, split into GPU phase and host phase (host). */
int baas_schedule_times(const baas_schedule *s, double *gpu_s, double *host_s);
/* Persistence.  The reference keeps its inputs across runs as binary
 * matrices (CSR::storeBin/loadBin, spm/CSR.cpp:332-435) and redoes its
 * buildset() analysis every run; a schedule can be kept too.  save writes the
 * analysis (groups, levels, threads per level, merged levels, reduced DAG,
 * group costs), the plan in use (mode, rows per unit, threads, dataflow
 * kernel, TILES band: an autotuned choice included) and a digest of the
 * pattern into buf (host); buf == NULL only queries *len.  load rebuilds the
 * schedule for the same pattern and kind (device rowptr/colidx as for
 * baas_analyze) without re-running the coarsening analysis or the
 * autotuner: the exports, the plan and every result are the original's.
 * Bytes from another pattern, kind or format version: BAAS_INVALID_ARG. */
int baas_schedule_save(const baas_schedule *s, void *buf, size_t *len);
int baas_schedule_load(baas_handle *h, int n, const int *rowptr, const int *colidx, int kind,
                       const void *buf, size_t len, baas_schedule **out);
/* Device pointers of U = L^T's CSR pattern (rows with the diagonal first). */
int baas_schedule_upper_pattern(const baas_schedule *s, const int **rowptr, const int **colidx);

/* ---- triangular solves (sptrsv_p2p_csr_group_merge_no_perm_alloc,
 * merge/SpTRSV.cpp:421-478; backward form :601-660) ------------------------ */
/* uplo 0: L x = b with vals in L's CSR layout (pattern given at analysis).
 * uplo 1: U x = b with vals in U = L^T's CSR layout; the sum runs in
 *         descending column order as in the backward P2P kernel.
 * nrhs >= 1: b and x are row-major n x nrhs.  x must not alias b. */
int baas_sptrsv_f64(baas_handle *h, baas_schedule *s, int uplo, const double *vals,
                    const double *b, double *x, int nrhs);
int baas_sptrsv_f32(baas_handle *h, baas_schedule *s, int uplo, const float *vals,
                    const float *b, float *x, int nrhs);
/* No counterpart in the reference.  For callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the solve of baas_sptrsv_f64, same arguments and the same
 * bits, but it calls no hipMalloc/hipFree.  baas_sptrsv_f64 allocates a
 * plan's publication area (2 x n x nrhs values) on its first dataflow call;
 * this entry point uses a fixed device array of BAAS_NOALLOC_MAX_VALUES
 * doubles (32 MiB, held by each GPU the library runs on until the process
 * ends) that every call on that GPU shares, re-armed at the start of each
 * call by a fill kernel.  Solves with 2 x n x nrhs > BAAS_NOALLOC_MAX_VALUES
 * return BAAS_TOO_LARGE and launch nothing.  Calls that overlap on one GPU
 * (several streams or threads) race on the array: serialise them.  The
 * schedule (baas_analyze) must exist before; it is not allocated here. */
#define BAAS_NOALLOC_MAX_VALUES 4194304
int baas_sptrsv_f64_noalloc(baas_handle *h, baas_schedule *s, int uplo, const double *vals,
                            const double *b, double *x, int nrhs);
/* Level-synchronous solve, one launch per row level (sptrsv_level_csr_no_perm,
 * merge/SpTRSV.cpp:41-80).  Same bits as baas_sptrsv_f64 with uplo 0. */
int baas_sptrsv_levelsync_f64(baas_handle *h, baas_schedule *s, const double *vals,
                              const double *b, double *x);
/* Gather L's values into U = L^T's CSR layout (and FP32 casts of either). */
int baas_transpose_values_f64(baas_handle *h, baas_schedule *s, const double *lvals, double *uvals);
int baas_cast_f64_f32(baas_handle *h, int n, const double *in, float *out);
int baas_cast_f32_f64(baas_handle *h, int n, const float *in, double *out);

/* Adjoint of x = L^-1 b: given lambda = L^-T g (solve with uplo 1 on the
 * transposed values), dL_ij = -lambda_i * x_j on L's pattern (row-major over
 * L's nonzeros).  nrhs columns are summed. */
int baas_sptrsv_grad_vals_f64(baas_handle *h, baas_schedule *s, const double *lambda,
                              const double *x, double *dvals, int nrhs);

/* Adjoint of x = A^-1 b for any CSR A: dA_ij = -sum_r lambda_{i,r} x_{j,r}
 * with lambda = A^-T g (for ICCG: a second PCG solve on the symmetric A). */
int baas_csr_outer_grad_f64(baas_handle *h, int m, const int *rowptr, const int *colidx,
                            const double *lambda, const double *x, double *dvals, int nrhs);

/* ---- factorizations --------------------------------------------------------
 * IC0  (spic0_csr_uL_p2p_group_merge_no_perm, merge/IC.cpp:296-393)
 * ILU0 (ilu0_p2p_csr_group_merge_no_perm_alloc, merge/ILU.cpp:638-720)
 * a_vals: values of the full A analysed with kind IC0/ILU0.  lu: nnz(A)
 * output, starts as a copy of a_vals (may alias a_vals).  IC0 writes the
 * lower part including the diagonal; ILU0 writes every entry. */
int baas_ic0_f64(baas_handle *h, baas_schedule *s, const double *a_vals, double *lu);
int baas_ilu0_f64(baas_handle *h, baas_schedule *s, const double *a_vals, double *lu);
/* Split an IC0 result into L values (lower CSR) and U = L^T values (upper
 * CSR of the schedule), test/iccg_level_merge.cpp:809-891; FP64 or FP32. */
int baas_ic0_split_f64(baas_handle *h, baas_schedule *s, const double *lu, double *lvals, double *uvals);
int baas_ic0_split_f32(baas_handle *h, baas_schedule *s, const double *lu, float *lvals, float *uvals);

/* ---- SpMV / BLAS-1 (blas/spmv.cpp, blas/blasone.cpp) ---------------------- */
/* y = alpha * (A x) + beta * y  (beta == 0: y is not read). */
int baas_spmv_f64(baas_handle *h, int m, const int *rowptr, const int *colidx, const double *vals,
                  const double *x, double *y, double alpha, double beta);
/* *result (DEVICE double) = x . y */
int baas_dot_f64(baas_handle *h, int n, const double *x, const double *y, double *result);
int baas_nrm2_f64(baas_handle *h, int n, const double *x, double *result);
/* w = alpha x + beta y (w may alias x or y) */
int baas_axpby_f64(baas_handle *h, int n, double *w, double alpha, const double *x, double beta,
                   const double *y);

/* ---- ICCG (pcg_opt, test/iccg_level_merge.cpp:380-533) --------------------
 * A: full SPD matrix (device CSR) analysed with kind IC0.  precision 0: all
 * FP64; 1: IC0 factor stored and both triangular solves run in FP32, CG
 * (residual, dots, updates) in FP64.  The iteration body is captured in a
 * hipGraph of `chunk` iterations; convergence (||r||/||r0|| < tol) is
 * checked on device and stops further work inside the chunk. */
int baas_pcg_create(baas_handle *h, baas_schedule *s, const double *a_vals, int precision,
                    baas_pcg **p);
int baas_pcg_destroy(baas_pcg *p);
/* x: initial guess in, solution out.  iters/relres: host outputs (first k
 * with relres < tol, or maxiter).  Re-runs IC0 when refactor != 0. */
int baas_pcg_solve(baas_pcg *p, const double *b, double *x, double tol, int maxiter,
                   int refactor, int *iters, double *relres);
/* Seconds of the last solve split into IC0, solve loop (host). */
int baas_pcg_times(const baas_pcg *p, double *ic_s, double *iter_s);
/* Factor if needed, then autotune the schedule's plans for this PCG's
 * triangular sweeps in its precision (no reference counterpart: the
 * reference fixes its schedule at analysis).  chosen as baas_autotune. */
int baas_pcg_autotune(baas_pcg *p, int *chosen);

/* ---- batched solves -------------------------------------------------------
 * nsys independent SPD systems packed back to back (block-diagonal CSR of
 * total size N, system i occupying rows sys_ptr[i]..sys_ptr[i+1]-1, host
 * array), each with nrhs right-hand sides (row-major N x nrhs).  One
 * schedule over the block-diagonal matrix; one persistent launch per sweep. */
int baas_batch_create(baas_handle *h, int nsys, const int *sys_ptr_host, int N,
                      const int *rowptr, const int *colidx, const double *a_vals, int nrhs,
                      const baas_opts *opts, baas_batch **b);
int baas_batch_destroy(baas_batch *b);
/* IC0 (FP64) then PCG with an FP32 factor and FP64 CG on all systems and
 * right-hand sides; iters (host int[nsys*nrhs]) receives per-column counts. */
int baas_batch_pcg(baas_batch *b, const double *B, double *X, double tol, int maxiter, int *iters);
/* times[0..2] (host) = IC0 + setup, CG loop, autotuning seconds of the last
 * call; rpu[0..2] (host) = tuned rows per unit for systems of <= 4096,
 * <= 16384 and more rows, {0, 0, workgroup size} when the level-set plan
 * (all systems' rows of one level at a time) won, or the box size three
 * times when boxes of each system's stencil grid on the multi-RHS lane
 * kernel won (every system's pattern a 5- or 7-point grid in natural
 * order; the boxes of all systems are claimed in one wavefront order). */
int baas_batch_info(const baas_batch *b, double *times, int *rpu);

/* Re-pack a schedule's execution plan (baas_plan_mode RANGE, SCHEDULE,
 * LEVELS or TILES; rows per unit; threads per workgroup: a multiple of 64 up
 * to 1024). */
int baas_replan(baas_schedule *s, int mode, int rows_per_unit, int block_threads);
/* Re-pack into TILES units of row_range consecutive row ids by band_levels
 * consecutive row levels (no reference counterpart: a packing of the
 * reference's level sets for GPU work units). */
int baas_replan_tiles(baas_schedule *s, int row_range, int band_levels, int block_threads);
/* The plan in use (host int[5]): mode, rows per unit (TILES: row range),
 * threads per workgroup, dataflow, band levels (TILES, else 0). */
int baas_schedule_plan(const baas_schedule *s, int *plan);
/* Where the PIPE plan in use applies (host int[2]): [0] = 1 when the
 * forward plan is a PIPE plan (stencil boxes), [1] = 1 when IC0 / ILU0 of
 * this schedule run pipelined on it (A is the full stencil).  No reference
 * counterpart: a diagnostic of this library's plan kinds. */
int baas_schedule_pipe(const baas_schedule *s, int *info);
/* Sweeps of this schedule use a dataflow kernel when the unit's data fits
 * in LDS: enable 1 (single right-hand side) = rows in rounds of one per lane with prefetch
 * and exchange wavefronts for parents of other units; enable 2 = the unit's
 * rows staged in LDS, every lane advancing through its own rows (one row and
 * right-hand side per item, any nrhs) on its own and publishing them itself
 * (workgroups of at least 128 threads); 0 = off.  Enable 2 also runs IC0 and
 * ILU0 on this schedule's units with the same design (rows with at most 255
 * lower entries; others keep the unit kernels).  Enable 3 and 4 = the
 * multi-right-hand-side lane kernel: (row, right-hand side) items over every
 * lane of the workgroup, parents in other units read straight from the
 * publication array by the lane that needs them (no exchange wavefront),
 * x written as each item is solved, so LDS holds only the unit's own
 * values; 3 also stages the unit's b in LDS where it fits, 4 reads b from
 * memory.  Sweeps only (factorisations keep the unit kernels).  Enable 5
 * (PIPE plans) = the wave-pipelined grid sweep for one right-hand side (a
 * launch per column for several); factorisations then take the lane
 * kernel's path. */
int baas_plan_dataflow(baas_schedule *s, int enable);
/* Autotune the execution plan for sweeps with `uplo` and `nrhs` on these
 * values: times RANGE, SCHEDULE, LEVELS and TILES packings over their unit
 * sizes, workgroup size and the dataflow kernel, keeps the fastest.  chosen
 * (host int[5], may be NULL) = mode, rows per unit, threads per workgroup,
 * dataflow, microseconds (the TILES band: baas_schedule_plan). */
int baas_autotune(baas_handle *h, baas_schedule *s, const double *vals, int uplo, int nrhs, int *chosen);
/* Time the numeric factorisation of an IC0/ILU0 schedule over unit,
 * schedule and level-set plans on these values and keep the fastest (no
 * reference counterpart: the reference fixes its schedule at analysis). */
int baas_autotune_factor(baas_handle *h, baas_schedule *s, const double *a_vals, int *chosen);
/* The same for FP32 sweeps (whose dataflow units may hold twice the rows). */
int baas_autotune_f32(baas_handle *h, baas_schedule *s, const float *vals, int uplo, int nrhs, int *chosen);

#ifdef __cplusplus
}
#endif
#endif
