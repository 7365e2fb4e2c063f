/*
 * brdc.h -- C ABI of the MI355X boundary-row (BR) values-only
 * divide-and-conquer eigensolver for symmetric tridiagonal matrices.
 *
 * Arithmetic is FP64 throughout.  D holds the diagonal (n entries), E the
 * off-diagonal (n-1 entries); on success D is overwritten by the eigenvalues
 * in ascending order.  Indices are 0-based; sizes are element counts unless
 * named *_bytes.  Plain pointers and sizes only: any FFI (ctypes, cgo, JNI,
 * Fortran ISO_C_BINDING) binds these symbols directly.
 */
#ifndef BRDC_H
#define BRDC_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status of the device entry points (the launch itself; numerical failures
 * are reported through the device-side info word).  Mirrors the reference's
 * evd_status (gpu/include/evd/status.h:8-18) for the codes this path uses. */
typedef enum brdc_status {
    BRDC_STATUS_SUCCESS = 0,
    BRDC_STATUS_INVALID_VALUE = 3,
    BRDC_STATUS_INVALID_POINTER = 2,
    BRDC_STATUS_INVALID_WORKSPACE = 4,
    BRDC_STATUS_ALLOC_FAILED = 5,
    BRDC_STATUS_HIP_ERROR = 6,
    BRDC_STATUS_INTERNAL_ERROR = 7
} brdc_status;

/* Tuning knobs of the device path; zero-initialise for the defaults. */
typedef struct brdc_options {
    int leaf_max;           /* largest leaf (rows) of the tree, solved by implicit
                               QL, one lane per leaf, 2..64; 0 selects the
                               default: 8 when the subtree kernel (below) takes
                               the bottom of the tree, else 16 */
    int stream_min_k;       /* merges with at least this many non-deflated poles
                               go to the streamed secular kernel, 256..2^28;
                               0 selects the default (2048) */
    int secular_lanes_log2; /* log2 of the secular lanes a level aims to keep in
                               flight (lanes per root follow from it), 12..24;
                               0 selects the default (18) */
    int wide_max_waves;     /* widest group (waves) of the product-form and
                               boundary-row kernels, 1, 2, 4 or 8; 0 selects
                               the default (8) */
    int subtree_max_log2;   /* the bottom of the tree (nodes of at most 128 rows
                               with all their leaves and merges) runs in one
                               launch, a block per node, when there are at most
                               2^subtree_max_log2 such nodes (a batch: over all
                               its matrices), 1..30; -1 never; 0 selects the
                               default (10) */
    int reserved[3];        /* must be zero */
} brdc_options;
/* The launch-shape knobs after leaf_max change which kernel solves a merge
 * and how sums are split over lanes, so results move within rounding (not bit
 * for bit) when they change; brdc.autotune picks them by timing. */

/*
 * Host-array entry.  Replaces DSTEDC_DC_BR (cpu/src/dstedc_dc_br.f:7-8), the
 * reference's values-only BR driver, with the same argument meaning:
 *   n       order of the matrix (n >= 0)
 *   d[n]    in: diagonal; out: eigenvalues ascending
 *   e[n-1]  in: off-diagonal; contents are not used afterwards
 *   work, lwork, iwork, liwork  workspace as in the reference; lwork == -1 or
 *           liwork == -1 is a query that stores the minimum sizes in work[0]
 *           and iwork[0] ((16 + 2*NTHR)*n and 7*n for n > 25, else 1 and 1,
 *           with NTHR = 1: the solve itself runs on the GPU)
 * Returns INFO: 0 success; -1 bad n; -5 lwork too small; -7 liwork too small;
 * > 0 numerical failure of a tree node (the reference's encoding,
 * dlaed0_br.f:302-303 remapped at dstedc_dc_br.f:123-124: the failing node's
 * first row r and size m give r*(n+1) + r + m - 1, r 1-based) or non-finite
 * input (n + 1); below -1000, a HIP runtime failure (-1000 - brdc_status).
 * Blocks the calling thread; device memory is allocated for the call and
 * released before it returns.
 */
int brdc_dstebr(int n, double *d, double *e, double *work, int lwork, int *iwork,
                int liwork);

/* Fortran-callable alias with all arguments by reference (LP64), the calling
 * convention of dstedc_dc_br_ (scripts/cpu/smoke_br.f90:24-27). */
void brdc_dstebr_(const int *n, double *d, double *e, double *work, const int *lwork,
                  int *iwork, const int *liwork, int *info);

/* The same entry under the reference's own symbol (SUBROUTINE DSTEDC_DC_BR,
 * cpu/src/dstedc_dc_br.f:7-8, as gfortran and amdflang name it), so an
 * unchanged Fortran caller links against this library in place of the
 * reference's objects.  Weak: a strong definition in the program wins. */
void dstedc_dc_br_(const int *n, double *d, double *e, double *work, const int *lwork,
                   int *iwork, const int *liwork, int *info);

/*
 * Device workspace query.  Replaces the reference's
 * evd::internal::tridiag_solver_workspace_size (gpu/src/tridiag_solver.cu:2373).
 * Writes the bytes the device entry needs for order n (O(n): about 40 doubles
 * per row; no O(n^2) slab).  Returns a brdc_status.
 */
int brdc_dstebr_device_workspace(int n, size_t *bytes);
int brdc_dstebr_device_workspace_ex(int n, const brdc_options *opts, size_t *bytes);

/*
 * Device entry.  Replaces evd::internal::tridiag_solver_run_from_device_tridiagonal
 * (gpu/src/tridiag_solver.cu:2541-2583) on the values-only path:
 *   d_dev[n]     device; in: diagonal, out: eigenvalues ascending
 *   e_dev[n-1]   device; off-diagonal, not modified
 *   workspace    device, caller-owned, >= brdc_dstebr_device_workspace bytes,
 *                256-byte aligned
 *   stream       hipStream_t the work is enqueued on (NULL = default stream)
 *   info_dev     device int receiving INFO as brdc_dstebr defines it (written
 *                by the last kernel of the solve; may be NULL)
 * Asynchronous: returns after enqueueing, with no host synchronisation and no
 * device allocation, so it can be captured into a hipGraph.  Returns a
 * brdc_status for the launch.
 */
int brdc_dstebr_device(int n, double *d_dev, const double *e_dev, void *workspace,
                       size_t workspace_bytes, void *stream, int *info_dev);
int brdc_dstebr_device_ex(int n, double *d_dev, const double *e_dev, void *workspace,
                          size_t workspace_bytes, void *stream, int *info_dev,
                          const brdc_options *opts);

/*
 * Allocation-free device entry.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory or bring a workspace (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as brdc_dstebr_device_ex without
 * the workspace, which comes from a static device array of
 * BRDC_NOALLOC_BYTES per GPU (held by each GPU the library runs on until the
 * process ends).  Orders up to BRDC_NOALLOC_MAX_N; a larger n returns
 * BRDC_STATUS_INVALID_VALUE.  Calls no allocating routine.  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (two streams,
 * two host threads) race and must be serialised by the caller.
 */
#define BRDC_NOALLOC_MAX_N 65536
#define BRDC_NOALLOC_BYTES (20u << 20)
int brdc_dstebr_device_noalloc(int n, double *d_dev, const double *e_dev, void *stream,
                               int *info_dev, const brdc_options *opts);

/*
 * Per-merge diagnostics of the last solve run in a workspace (the role of
 * TridiagMergeLevelDiagnostics, gpu/src/handle_impl.hpp:16-85, and of the
 * per-merge (N, K) record of the instrumented CPU copy).  Copies up to
 * max_merges records of BRDC_MERGE_STAT_FIELDS ints -- level (0 = root
 * merge), first row, size, K (non-deflated poles), G (close-pole Givens
 * deflations: close_deflated_total), secular function evaluations summed over
 * the merge's roots (secular_iter_sum), T (tiny-z deflations:
 * zero_deflated_total), the most evaluations of one root (secular_iter_max),
 * roots that did not converge (secular_iter_failed) -- into out_host, ordered
 * by level from the root down, and returns the number of merges (or a
 * negative brdc_status).  Synchronises the stream.
 */
#define BRDC_MERGE_STAT_FIELDS 9
int brdc_merge_stats(int n, const void *workspace, size_t workspace_bytes, void *stream,
                     const brdc_options *opts, int *out_host, int max_merges);

/*
 * Kernel-class timing of the solves enqueued from the calling host thread
 * between begin and end (diagnostics only; the role of the reference's
 * MergeLevelTimer, gpu/src/tridiag_solver.cu:158-359).  Each launch is
 * bracketed by two HIP events on its own stream.  end synchronises those
 * events and writes, for each of the first nclasses classes in the order
 * norm, prep_e, leaf, prepare, deflate, secular (merges below 2048 poles),
 * wprod and rows (the product form and boundary rows of merges of more than
 * 128 poles at levels of nodes of more than 256 rows), finalize (which
 * computes them for the other merges), secular_stream (merges of 2048
 * poles or more, every evaluation after the first), secular_mid (their
 * first, midpoint evaluation), subtree (the leaves and every merge of nodes
 * of at most 128 rows, in one launch) (BRDC_KERNEL_CLASSES of them), the
 * summed milliseconds and launch counts.
 */
#define BRDC_KERNEL_CLASSES 12
int brdc_profile_begin(void);
int brdc_profile_end(double *ms_out, int *launches_out, int nclasses);

/* As brdc_profile_end, split by merge level (the per-level values_*_ms of
 * TridiagMergeLevelDiagnostics): ms_out and launches_out hold (nlevels + 1)
 * rows of nclasses entries; row 0 gathers the launches before the merges
 * (norm, prep_e, leaf, subtree: the subtree launch holds the levels of nodes
 * of at most 128 rows, which then have no row of their own), row 1 + l the
 * merges of level l (tree depth, 0 = the root merge; for a batch, forest
 * height l, 1 = one level above the leaves).  Levels beyond nlevels are
 * dropped. */
int brdc_profile_end_levels(double *ms_out, int *launches_out, int nclasses, int nlevels);

/*
 * Batched solves: many independent tridiagonals of uneven sizes as one
 * flattened merge forest.  No counterpart in the reference (which solves one
 * matrix per call); the entry points follow the device entry's conventions.
 *
 * The matrices' rows are concatenated: matrix b occupies rows
 * [off_b, off_b + n_b) with off_b = n_0 + ... + n_{b-1}; d holds the
 * diagonals (rows entries) and e the off-diagonals in the same rows, e[off_b +
 * i] for i < n_b - 1 (the slot of each matrix's last row is not read).
 *
 * A plan is built once per size list on the host (every matrix's tree and the
 * per-height launch shapes), its tables are copied into a workspace once
 * (brdc_batch_upload, synchronous), and brdc_batch_solve then enqueues a
 * host-sync-free, allocation-free launch sequence -- one sweep per height of
 * the forest -- that can be captured into a hipGraph and replayed.
 */
typedef struct brdc_batch_plan brdc_batch_plan;

/* Builds the plan of nmat matrices of orders sizes[0..nmat) (each >= 0; all
 * rows together below 2^29) with the options' leaf bound.  Returns a status;
 * *plan is NULL on failure. */
int brdc_batch_plan_create(int nmat, const int *sizes, const brdc_options *opts,
                           brdc_batch_plan **plan);
void brdc_batch_plan_destroy(brdc_batch_plan *plan);

/* Device workspace bytes for the plan (O(rows) plus the plan tables). */
int brdc_batch_workspace(const brdc_batch_plan *plan, size_t *bytes);

/* Copies the plan tables into the workspace on stream and waits for the copy.
 * Must precede brdc_batch_solve on that workspace (once per workspace). */
int brdc_batch_upload(const brdc_batch_plan *plan, void *workspace, size_t workspace_bytes,
                      void *stream);

/* Enqueues the batched solve: d_dev[rows] (in: diagonals, out: each matrix's
 * eigenvalues ascending in its own rows), e_dev[rows] (not modified),
 * info_dev[nmat] (each matrix's INFO as brdc_dstebr defines it, rows numbered
 * within the matrix).  Asynchronous; returns a status for the launch. */
int brdc_batch_solve(const brdc_batch_plan *plan, double *d_dev, const double *e_dev,
                     void *workspace, size_t workspace_bytes, void *stream, int *info_dev);

/* Per-merge diagnostics of the last batched solve in a workspace, as
 * brdc_merge_stats: records of BRDC_MERGE_STAT_FIELDS ints whose level is the
 * forest height (1 = one level above the leaves) and whose first row counts in
 * the concatenated rows; ordered by height, then by matrix.  Returns the
 * number of merges (or a negative brdc_status); synchronises the stream. */
int brdc_batch_merge_stats(const brdc_batch_plan *plan, const void *workspace,
                           size_t workspace_bytes, void *stream, int *out_host, int max_merges);

/* Plan summary: up to nout of (rows, matrices, leaves, forest heights, merges,
 * largest leaf, leaf bound).  Returns the number written or a negative status. */
int brdc_batch_plan_stats(const brdc_batch_plan *plan, int *out, int nout);

/*
 * Backward of the eigenvalues (no counterpart in the reference, whose solver
 * is values-only).  Given D, E, the ascending eigenvalues W from the solve and
 * an upstream gradient G (all device, FP64), writes
 *   DD[j] = sum_i G[i] v_ij^2             (dL/dD_j)
 *   DE[j] = 2 sum_i G[i] v_ij v_i,j+1     (dL/dE_j; DE[n-1] = 0)
 * with each unit eigenvector v_i recomputed at W[i] by twisted factorization
 * (nothing is stored by the forward pass).  The scratch buffer (caller-owned,
 * 256-byte aligned) holds per-wave pivot arrays and per-task partial sums;
 * the sums over tasks are added in a fixed order, so results are
 * reproducible.  Asynchronous on stream; returns a status for the launch.
 * For n >= 8192 the eigenvectors are crossed segment by segment with stored
 * transfers; a transfer whose partial products exceed 2^400 is replaced by a
 * walk of the rows.  The scratch grows as about n^2 / 8 doubles (1.3 GB at
 * n = 16384, 18.9 GB at n = 65536); above 2^37 bytes (n of about 170000) the
 * workspace queries and the entries return BRDC_STATUS_INVALID_VALUE.
 * Eigenvalues within 2^-44 ||T|| of a neighbour (numerically repeated: for
 * instance a matrix split into identical blocks) get their vectors from
 * inverse iteration with Gram-Schmidt inside the cluster, so the cluster's
 * vectors stay orthonormal.  The environment variable BRDC_GRAD_SCAN_LIMIT (read at
 * each call, a value in [0, 2^400]) lowers that bound; 0 walks every row.
 */
int brdc_dstebr_grad_workspace(int n, size_t *bytes);
int brdc_dstebr_grad_device(int n, const double *d_dev, const double *e_dev,
                            const double *w_dev, const double *g_dev, double *dd_dev,
                            double *de_dev, void *scratch, size_t scratch_bytes, void *stream);

/* Batched backward over a plan's concatenated rows (same layout as
 * brdc_batch_solve; DE is 0 at each matrix's last row).  workspace: a
 * workspace the plan was uploaded to (its tables are read). */
int brdc_batch_grad_workspace(const brdc_batch_plan *plan, size_t *bytes);
int brdc_batch_grad(const brdc_batch_plan *plan, const double *d_dev, const double *e_dev,
                    const double *w_dev, const double *g_dev, double *dd_dev, double *de_dev,
                    const void *workspace, size_t workspace_bytes, void *scratch,
                    size_t scratch_bytes, void *stream);

/* Human-readable name of a brdc_status. */
const char *brdc_status_string(int status);

#ifdef __cplusplus
}
#endif
#endif /* BRDC_H */
