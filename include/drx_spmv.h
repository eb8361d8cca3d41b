/*
 * drx_spmv.h -- C ABI of the MI355X (gfx950) sparse gemv library.
 *
 * This is the drop-in boundary for Distributed Ranges' sparse gemv path: the
 * per-share / per-tile kernel seam that dr::mp::gemv and dr::sp::gemv call.
 * All pointers named rowptr/colind/vals/x/y/... are DEVICE pointers owned by
 * the caller; the library owns only the opaque plan / transpose / batch
 * objects. Every function returns an int status (DRX_OK == 0, negative on
 * error), never throws and never aborts; drx_last_error() gives the message
 * of the calling thread's last failure. Work is enqueued on `stream` (a
 * hipStream_t; NULL = the null stream) and is asynchronous unless stated.
 *
 * Index semantics shared by every CSR entry point. A CSR share is
 *   rowptr[0 .. nrows)        absolute offsets into the full matrix's nnz
 *   rowptr[nrows]             read only when has_sentinel != 0
 *   colind[0 .. nnz), vals[0 .. nnz)   the share's entries
 *   rowptr_base               the absolute offset of colind[0] / vals[0]
 * Row r covers share entries [clamp(rowptr[r] - base), clamp(end_r - base))
 * with end_r = rowptr[r+1] (r + 1 < nrows or a sentinel) else base + nnz,
 * and clamp() into [0, nnz]. This single rule is
 *   - a full CSR matrix / dr::sp tile: base 0, sentinel present
 *     (include/dr/views/csr_matrix_view.hpp:125-218,
 *      include/dr/sp/containers/sparse_matrix.hpp:193-207);
 *   - a dr::mp csr_row share: base = rowptr[first row], no sentinel
 *     (include/dr/mp/containers/matrix_formats/csr_row_distribution.hpp:92-99);
 *   - a dr::mp csr_eq share: rows cut at both ends of the nnz range, no
 *     sentinel (include/dr/mp/containers/matrix_formats/csr_eq_distribution.hpp:
 *     62-111), so the first and last outputs are partial sums.
 * Dense operands with k right-hand sides are column-major like
 * dr::mp::broadcasted_slim_matrix (include/dr/mp/containers/
 * broadcasted_slim_matrix.hpp:123): element (i, j) of x is x[i + j*ldx].
 */
#ifndef DRX_SPMV_H
#define DRX_SPMV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define DRX_OK 0
#define DRX_ERR_INVALID (-1)     /* bad argument (null, size, dtype combo) */
#define DRX_ERR_UNSUPPORTED (-2) /* valid but not built (dtype/index combo) */
#define DRX_ERR_HIP (-3)         /* a HIP runtime call failed */
#define DRX_ERR_NOMEM (-4)       /* device or host allocation failed */

/* Element types (storage). Accumulation is f64 for DRX_F64 and f32 for the
 * others; DRX_BF16 / DRX_F16 are storage-only types for vals and x. */
#define DRX_F64 0
#define DRX_F32 1
#define DRX_BF16 2
#define DRX_F16 3

/* Index types of rowptr and colind (both the same type). */
#define DRX_I32 0
#define DRX_I64 1

typedef void *drx_stream_t; /* a hipStream_t */
typedef struct drx_plan_s *drx_plan;
typedef struct drx_transpose_s *drx_transpose;
typedef struct drx_batch_s *drx_batch;

/* ---- library ----------------------------------------------------------- */
int drx_version(void);
const char *drx_last_error(void);
/* Process-wide tuning knobs. No counterpart in the reference.
 *   "block_nnz"   default entries per row block of new plans (0 restores the
 *                 built-in default; env DRX_BLOCK_NNZ seeds it).
 *   "mrhs"        k >= 2: 2 copies x into row-major chunks on the call's
 *                 stream (scratch from a per-device pool the library keeps)
 *                 and gathers one contiguous vector per entry; 0 gathers the
 *                 k columns of x directly (no scratch); 1 (default) chunks
 *                 except for k <= 4 over short rows (<= 8 entries on
 *                 average) with local columns (stencils; locality detected
 *                 once per colind pointer with one small kernel and a wait
 *                 on the stream).
 *   "mrhs_gshift" lane-group shift of the k >= 2 kernel, 0..3 (default 1).
 *   "mrhs_kp_max" widest row-major chunk for 4-byte x: 16, 8, or 0 (default:
 *                 8 for k > 8 over loosely local columns -- stencils,
 *                 bands -- else 16; batches use 16). Fixed per call, and per
 *                 batch when it is created.
 *   "vec"         consecutive entries per lane and load (32-bit indices,
 *                 one right-hand side or column gathers): 0 (default) groups
 *                 of four for values of at most 4 bytes, one for FP64; 1
 *                 one; 2 pairs; 4 groups of four (values of at most 4
 *                 bytes) -- wherever eligible (vals and colind aligned to a
 *                 group), else one. Same bits for every choice.
 *   "wide"        1 (default): the windows of 4096 entries (16 per lane)
 *                 that "fwd_cap" 0 picks, for plans whose block_nnz is the
 *                 default 1024; 0: the plan's own blocks.
 *   "grid_cap"    most workgroups of a per-share launch (0 = one per row
 *                 block); capped launches loop over the blocks.
 *   "xstage"      x windows in LDS: 2 makes each row block of the forward
 *                 product (one right-hand side) and of the fused residual
 *                 load the x window its columns span into LDS once and read
 *                 x from there (a window wider than the LDS buffer reads the
 *                 rest from memory); 0 and 1 (default) gather x per entry,
 *                 which measured faster on bands. Results are the same bit
 *                 for bit.
 *   "auto_panels" 1 (default): drx_plan_auto_panels builds the panel layout
 *                 when it pays; 0: never.
 *   "fill"        1 (default): plans that measure their longest row L
 *                 (drx_plan_create, transposes, panels, batches) start a
 *                 block every W - L entries for a window of W entries, so
 *                 blocks fill their windows (when L <= W / 2 and rows
 *                 average at least 2.5 entries); 0: every W / 2 entries.
 *   "fwd_cap"     forward window of plain CSR with one right-hand side: 0
 *                 (default) 4096 for FP32 values, and for 16-bit values
 *                 streamed in groups of four, with 32-bit indices (see
 *                 "wide") unless "rowgather" applies, else the plan's own
 *                 (2 block_nnz); 2048 or 4096 force one where the plan
 *                 holds those records.
 *   "rowgather"   row-ordered x gathers for plain CSR with one right-hand
 *                 side: each block's values and columns are staged in LDS
 *                 and each lane sums one row, gathering its x entries in
 *                 row order (stencils: a wave's gather reads 64 consecutive
 *                 rows' neighbours; lane groups of 8 or 16 per row in
 *                 blocks of longer rows). 1 (default) for plans whose
 *                 longest row has at most 128 entries and whose rows
 *                 average at least 4 (2048-entry windows, 4096 for 16-bit
 *                 values), copying the window straight
 *                 into LDS where vals and colind are aligned to 16 bytes (4
 *                 for 16-bit values), else through registers; 2 always,
 *                 through registers; 3 always, straight copies where
 *                 aligned; 0 never. The adjoint through the permutation
 *                 (drx_spmv_adjoint without stored CSC values), drx_csr_dvals
 *                 (one right-hand side) and drx_spmv_fused_residual follow
 *                 the same choice with row-ordered kernels of their own (0
 *                 and 2: their entry-ordered kernels). Sums in entry order, so results may
 *                 differ from 0 in the last bits (all within the oracle
 *                 tolerance).
 *   "gemm_tile"   drx_gemm workgroup tile: 1 128 x 128; 2 256 x 128; 3 and
 *                 4 the same with stages twice as deep in k; 5 256 x 256
 *                 (FP32; FP64 takes 4); 6 128 x 128 loading stages straight
 *                 into LDS (FP32 and FP64); 7 256 x 128 the same way (FP32;
 *                 FP64 takes 4); 0 (default) by type and size: FP64 6; FP32
 *                 6 once C holds 4096 tiles of 128 x 128 (about 8192^2), 5
 *                 while it holds at least 256 tiles of 256 x 256 (about
 *                 4096^2), else 1. */
int drx_tune_set(const char *key, int64_t value);
int drx_tune_get(const char *key, int64_t *value);

/* ---- plan: analysis of one CSR share ------------------------------------
 * Splits the share's rows into blocks for windows of 2 * block_nnz entries
 * (block c: the rows whose first entry falls in [c * pitch, (c + 1) *
 * pitch); the pitch is the window minus the share's longest row when that
 * row fits half a window, else half a window, and a last row that would
 * overflow the window is reduced by the whole workgroup). block_nnz <= 0
 * picks the tuned default. Keeps a reference to `rowptr` (the caller keeps it
 * alive and unchanged). Synchronizes `stream` once to validate the share's
 * bounds and measure its longest row.
 * Replaces the per-share setup of csr_row_distribution::init
 * (csr_row_distribution.hpp:205-318), csr_eq_distribution::init
 * (csr_eq_distribution.hpp:220-333) and oneMKL's set_csr_data for the sp
 * tile (include/dr/sp/algorithms/matrix/local_gemv.hpp:63-72). */
int drx_plan_create(drx_plan *plan, int itype, const void *rowptr,
                    int64_t rowptr_base, int64_t nrows, int64_t ncols,
                    int64_t nnz, int has_sentinel, int64_t block_nnz,
                    drx_stream_t stream);
int drx_plan_destroy(drx_plan plan);
/* Number of row blocks and the block size the plan uses. */
int drx_plan_info(drx_plan plan, int64_t *nblocks, int64_t *block_nnz);

/* ---- column-panel layout (optional) --------------------------------------
 * For matrices whose x exceeds one XCD's L2 and whose columns are irregular
 * (random 4-byte gathers beyond L2 run ~4x slower than within it): keeps a
 * panel-major copy of (colind, vals) inside the plan -- columns cut into
 * panels of panel_cols (0: ~2 MB of x per panel), rows compressed per panel
 * -- and drx_spmv then runs one L2-resident launch per panel, accumulating
 * into y in panel order (deterministic). drx_spmv on such a plan must be
 * given the same colind/vals pointers; drx_plan_update_values re-packs new
 * values through the kept permutation. Synchronizes the stream. No
 * counterpart in the reference (an optimization of its local_gemv). */
int drx_plan_set_panels(drx_plan plan, int vtype, const void *colind,
                        const void *vals, int64_t panel_cols,
                        drx_stream_t stream);
int drx_plan_update_values(drx_plan plan, const void *vals,
                           drx_stream_t stream);
int drx_plan_panels(drx_plan plan, int64_t *npanels, int64_t *panel_cols,
                    int64_t *segments);
/* Builds the panel layout (as drx_plan_set_panels with panel_cols = 0) when
 * it pays, and reports it in *built (1, else 0; may be NULL). It pays when
 * x of vtype exceeds 4 MiB (one XCD's L2), the share averages at least 4
 * entries per row, its columns scatter (fewer than half of the entries
 * sampled from up to 2048 row blocks lie within 1 MiB of x of their block's
 * middle column; one small kernel), the share holds at most 2^31 - 1
 * entries and the device has memory for the copy. Synchronizes the stream;
 * refuses a stream under capture. Knob "auto_panels" 0 turns it off. Called
 * by the host interface's share and tile setup (the seams of
 * csr_row_distribution::init, csr_row_distribution.hpp:205-318,
 * csr_eq_distribution::init, csr_eq_distribution.hpp:220-333, and the
 * sp tiles, sparse_matrix.hpp:193-207), so that their products reach the
 * panel kernels. */
int drx_plan_auto_panels(drx_plan plan, int vtype, const void *colind,
                         const void *vals, drx_stream_t stream, int *built);

/* ---- forward ------------------------------------------------------------
 * y[:, j] = alpha * A x[:, j] + beta * y[:, j]   for j in [0, k)
 * (beta == 0: y is not read). Deterministic (no atomics).
 * Supported (vtype, xtype, ytype): (F64,F64,F64), (F32,F32,F32),
 * (BF16,BF16,F32), (F16,F16,F32), (BF16,F32,F32), (F16,F32,F32).
 * Replaces the kernels of csr_row_distribution::local_gemv
 * (csr_row_distribution.hpp:63-134; beta = 0 into the zeroed partial),
 * csr_eq_distribution::local_gemv (csr_eq_distribution.hpp:54-131) and
 * dr::sp::__detail::local_gemv (local_gemv.hpp:83-98; alpha = beta = 1, as
 * oneMKL sparse::gemv is called there). */
int drx_spmv(drx_plan plan, int vtype, const void *colind, const void *vals,
             int xtype, const void *x, int64_t ldx, int ytype, void *y,
             int64_t ldy, int64_t k, double alpha, double beta,
             drx_stream_t stream);

/* One-shot form without a kept plan (plans on `stream`, runs, frees the plan
 * stream-ordered; no host synchronization). Replaces
 * dr::sp::__detail::mkl_gemv (local_gemv.hpp:59-78), which sets up the CSR
 * handle on every call. */
int drx_spmv_csr(int itype, const void *rowptr, int64_t rowptr_base,
                 int64_t nrows, int64_t ncols, int64_t nnz, int has_sentinel,
                 const void *colind, int vtype, const void *vals, int xtype,
                 const void *x, int64_t ldx, int ytype, void *y, int64_t ldy,
                 int64_t k, double alpha, double beta, drx_stream_t stream);

/* The one-shot form above without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). No counterpart in the reference. The
 * call's plan records live in a library-owned device array of 8 MiB per GPU
 * (held until the process ends); the caller passes inputs and outputs only.
 * Shares of more than drx_spmv_csr_noalloc_max_nnz() entries (about 178M)
 * are rejected with DRX_ERR_INVALID. k >= 2 gathers the k columns of x
 * directly (no row-major copy of x). Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (different streams or threads without
 * ordering) race and give wrong results; calls ordered on one stream are
 * safe. Nothing is cleared between calls (each call writes every record it
 * reads). Same arguments, semantics and status codes as drx_spmv_csr. */
int drx_spmv_csr_noalloc(int itype, const void *rowptr, int64_t rowptr_base,
                         int64_t nrows, int64_t ncols, int64_t nnz,
                         int has_sentinel, const void *colind, int vtype,
                         const void *vals, int xtype, const void *x,
                         int64_t ldx, int ytype, void *y, int64_t ldy,
                         int64_t k, double alpha, double beta,
                         drx_stream_t stream);
/* Largest nnz drx_spmv_csr_noalloc accepts (host arithmetic, no device work). */
int64_t drx_spmv_csr_noalloc_max_nnz(void);

/* ---- adjoint ------------------------------------------------------------
 * No counterpart in the reference (it has no gradient path); these serve the
 * differentiation of the gemv above.
 * drx_transpose_create builds the transposed structure (CSC of the share) on
 * the GPU by a stable sort of the entries by column: colptr[ncols+1],
 * rowind[nnz] (row ids local to the share), perm[nnz] (the share position of
 * each transposed entry) -- all in the plan's index type -- and a plan over
 * it. vals are read through perm, so new vals need no rebuild. */
int drx_transpose_create(drx_transpose *t, drx_plan plan, const void *colind,
                         drx_stream_t stream);
int drx_transpose_destroy(drx_transpose t);
int drx_transpose_arrays(drx_transpose t, const void **colptr,
                         const void **rowind, const void **perm);

/* Store the values in transposed (CSC) order inside t: one gather through
 * perm. Adjoint calls given vals == NULL then stream these like a forward
 * product instead of gathering vals[perm[e]] at random; call again when the
 * values change. */
int drx_transpose_set_values(drx_transpose t, int vtype, const void *vals,
                             drx_stream_t stream);

/* dx[:, j] = alpha * A^T dy[:, j] + beta * dx[:, j]. Same dtype combos as
 * drx_spmv with (vtype, dytype, dxtype). vals: the CSR-order values (read
 * through the permutation), or NULL to use the values stored by
 * drx_transpose_set_values. Deterministic. */
int drx_spmv_adjoint(drx_transpose t, int vtype, const void *vals, int dytype,
                     const void *dy, int64_t lddy, int dxtype, void *dx,
                     int64_t lddx, int64_t k, double alpha, double beta,
                     drx_stream_t stream);

/* dvals[e] = alpha * sum_j dy[row(e) + j*lddy] * x[col(e) + j*ldx]
 *            + beta * dvals[e]   (the gradient of y = A x w.r.t. vals).
 * dvtype is DRX_F64 (with F64 x, dy) or DRX_F32. */
int drx_csr_dvals(drx_plan plan, const void *colind, int xtype, const void *x,
                  int64_t ldx, int dytype, const void *dy, int64_t lddy,
                  int64_t k, int dvtype, void *dvals, double alpha,
                  double beta, drx_stream_t stream);

/* Fused forward + residual + dvals for the loss 0.5 * ||A x - t||^2
 * (k = 1): y = A x, r = y - t, dvals[e] = r[row(e)] * x[col(e)], in one
 * pass over the matrix (vals/colind/x gathered once). r may alias y.
 * Types as drx_spmv with ytype == t's type == r's type == dvals' type. */
int drx_spmv_fused_residual(drx_plan plan, int vtype, const void *colind,
                            const void *vals, int xtype, const void *x,
                            int ytype, const void *t, void *y, void *r,
                            void *dvals, drx_stream_t stream);

/* ---- batched products ---------------------------------------------------
 * Many independent products -- forward items y = alpha*A x + beta*y and
 * adjoint items dx = alpha*A^T dy + beta*dx -- in one persistent launch.
 * Items re-plan their matrices inside the batch with the batch's tuning
 * (block_nnz, gshift, wg_per_cu); all items share k, alpha, beta and the
 * dtype combination (as drx_spmv) and an index type. A run allocates and
 * clears nothing, so drx_batch_graph can capture it into a hipGraph
 * (instantiated once; drx_batch_graph_launch replays it on any stream).
 * drx_batch_autotune times every tuning combination on the batch and keeps
 * the fastest (invalidates a captured graph); its JSON describes the choice.
 * It runs the items about a hundred times with beta taken as 0, so on return
 * every output holds alpha * (product), whatever it held before: with
 * beta != 0, restore the outputs and run again. A default block size from
 * drx_tune_set("block_nnz") above 2048 is clamped to 2048 for batches.
 * No counterpart in the reference: it batches the per-share gemv seam. */
typedef struct drx_batch_item {
  drx_plan plan;           /* forward item: the matrix's plan (else NULL) */
  drx_transpose transpose; /* adjoint item: its transposed structure (else NULL) */
  const void *colind;      /* forward: the matrix's colind (adjoint: unused) */
  const void *vals;        /* the matrix's values, CSR order; NULL for an
                              adjoint item whose transpose holds its values */
  const void *x;           /* forward: x; adjoint: dy */
  int64_t ldx;             /* <= 0: the operand's length */
  void *y;                 /* forward: y; adjoint: dx */
  int64_t ldy;             /* <= 0: the result's length */
} drx_batch_item;

int drx_batch_create(drx_batch *batch, int64_t nitems,
                     const drx_batch_item *items, int vtype, int xtype,
                     int ytype, int64_t k, double alpha, double beta,
                     drx_stream_t stream);
int drx_batch_destroy(drx_batch batch);
int drx_batch_run(drx_batch batch, drx_stream_t stream);
int drx_batch_graph(drx_batch batch);
int drx_batch_graph_launch(drx_batch batch, drx_stream_t stream);
int drx_batch_set_tuning(drx_batch batch, int64_t block_nnz, int gshift,
                         int wg_per_cu, drx_stream_t stream);
int drx_batch_autotune(drx_batch batch, int iters, char *json,
                       int64_t json_cap, drx_stream_t stream);
int drx_batch_info(drx_batch batch, int64_t *work_units, int64_t *block_nnz,
                   int *gshift, int *wg_per_cu);

/* ---- synthetic inputs (device-side generators) --------------------------
 * Counter-based (splitmix64) so any box regenerates identical data.
 * The reference generates its inputs on the host with
 * dr::generate_random_csr / dr::generate_band_csr
 * (include/dr/detail/generate_csr.hpp:40-153); these are device-side
 * generators for the large inputs. */
/* Fill n elements with U[lo, hi) (dtype F64/F32/BF16/F16). */
int drx_gen_uniform(int dtype, void *out, int64_t n, double lo, double hi,
                    uint64_t seed, drx_stream_t stream);
/* Fill n elements with N(0, 1) (Box-Muller). */
int drx_gen_normal(int dtype, void *out, int64_t n, uint64_t seed,
                   drx_stream_t stream);
/* Row lengths ~ Zipf(s) clipped to [1, lmax] into rowptr[1..nrows] and
 * rowptr[0] = 0, then an inclusive scan: rowptr becomes a CSR row pointer.
 * *nnz_out (host) receives rowptr[nrows]; fails with DRX_ERR_INVALID if it
 * exceeds the index type. Synchronizes the stream. */
int drx_gen_zipf_rowptr(int itype, void *rowptr, int64_t nrows, double s,
                        int64_t lmax, uint64_t seed, int64_t *nnz_out,
                        drx_stream_t stream);
/* Row lengths uniform in [lmin, lmax] (inclusive), scanned into rowptr as
 * drx_gen_zipf_rowptr does. */
int drx_gen_uniform_rowptr(int itype, void *rowptr, int64_t nrows,
                           int64_t lmin, int64_t lmax, uint64_t seed,
                           int64_t *nnz_out, drx_stream_t stream);
/* Given rowptr, fill colind with sorted, distinct columns in [0, ncols): the
 * row's L entries take one uniform column in each of L equal strata of
 * [0, ncols). Requires every row length <= ncols. */
int drx_gen_stratified_colind(int itype, const void *rowptr, int64_t nrows,
                              int64_t ncols, void *colind, uint64_t seed,
                              drx_stream_t stream);
/* 2-D (dim 2: grid g x g, 5-point) or 3-D (dim 3: g x g x g, 7-point)
 * Laplacian: diagonal 2*dim, off-diagonals -1, rows in lexicographic order,
 * columns sorted. nnz_of returns the entry count (host, no device work). */
int64_t drx_gen_laplacian_nnz(int dim, int64_t g);
int drx_gen_laplacian(int dim, int64_t g, int itype, void *rowptr,
                      void *colind, int vtype, void *vals,
                      drx_stream_t stream);
/* Band matrix: row i has columns [max(0, i-lo), min(n-1, i+hi)], vals
 * N(0,1) from `seed` (vals may be NULL to skip). */
int64_t drx_gen_band_nnz(int64_t n, int64_t lo, int64_t hi);
int drx_gen_band(int64_t n, int64_t lo, int64_t hi, int itype, void *rowptr,
                 void *colind, int vtype, void *vals, uint64_t seed,
                 drx_stream_t stream);
/* Copy bytes between device (or host) buffers and wait for the copy; for
 * bindings without a device runtime of their own (e.g. reading the arrays
 * drx_transpose_arrays returns). */
int drx_copy_device(void *dst, const void *src, int64_t bytes,
                    drx_stream_t stream);
/* out[i] = x[idx[i]] for i < n (0 <= idx < nx, x below 4 GiB; elements of
 * 2, 4 or 8 bytes by xtype). No counterpart in the reference: the
 * benchmark times it on a product's own column indices to measure the
 * random-gather rate that bounds products with scattered columns. */
int drx_gather(const int32_t *idx, int64_t n, int xtype, const void *x,
               int64_t nx, void *out, drx_stream_t stream);
/* Reads `bytes` (a multiple of 16, src 16-byte aligned) from device memory
 * with streaming 16-byte loads and writes one 4-byte digest per workgroup
 * into digest[0 .. min(digest_len, workgroups)). No counterpart in the
 * reference: the benchmark times it over a buffer much larger than the
 * caches to measure the HBM read ceiling that bounds the products. */
int drx_stream_read(const void *src, int64_t bytes, void *digest,
                    int64_t digest_len, drx_stream_t stream);
/* Convert between storage dtypes (round to nearest even). */
int drx_convert(int src_dtype, const void *src, int dst_dtype, void *dst,
                int64_t n, drx_stream_t stream);

/* ---- Matrix Market input (host) ------------------------------------------
 * Replaces dr::__detail::read_csr_matrix_base + convert_csr_base_to_csr and
 * dr::read_csr (include/dr/detail/matrix_io.hpp:119-252), the reader behind
 * dr::sp::mmread (include/dr/sp/util/matrix_io.hpp:68-84) and the mp
 * examples' read_csr<double, long> (examples/mp/sparse_benchmark.cpp:58).
 * Reads a coordinate file (field pattern -> value 1, any other field parsed
 * as a number; symmetry general, or symmetric: an off-diagonal (i, j) is
 * stored at (i, j) and (j, i)) with `nthreads` host threads (0 = all) into
 * host CSR arrays that the library allocates and drx_host_free releases:
 * rowptr (nrows + 1, 0-based), colind, vals (itype I32/I64; vtype F64/F32,
 * values parsed directly in that type). Rows are sorted by column; entries
 * with equal (row, column) are kept, in file order. Errors (status
 * DRX_ERR_INVALID): unreadable file, not "matrix coordinate", unsupported
 * symmetry, index out of bounds, more entry lines than the size line
 * announces, malformed entry line. No device work. */
int drx_mm_read_csr(const char *path, int one_indexed, int itype, int vtype,
                    int nthreads, int64_t *nrows, int64_t *ncols, int64_t *nnz,
                    void **rowptr, void **colind, void **vals);
void drx_host_free(void *p);

/* ---- Dense tile product ---------------------------------------------------
 * C = alpha * A * B + beta * C for row-major A (m x k, leading dimension
 * lda), B (k x n, ldb) and C (m x n, ldc), vtype DRX_F32 or DRX_F64, on the
 * matrix cores. Replaces dr::sp::__detail::local_gemm
 * (include/dr/sp/algorithms/matrix/local_gemm.hpp:45-60, oneMKL's row-major
 * gemm called with alpha = beta = 1; :15-41 custom_gemm, c += a * b), the
 * per-tile step of dr::sp::gemm (include/dr/sp/algorithms/matrix/
 * gemm.hpp:12-146). beta == 0 does not read C. C must not overlap A or B.
 * Errors: DRX_ERR_INVALID for negative sizes, a leading dimension below the
 * row length, null operands, other vtypes. Asynchronous on `stream`. */
int drx_gemm(int vtype, int64_t m, int64_t n, int64_t k, double alpha,
             const void *a, int64_t lda, const void *b, int64_t ldb,
             double beta, void *c, int64_t ldc, drx_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* DRX_SPMV_H */
