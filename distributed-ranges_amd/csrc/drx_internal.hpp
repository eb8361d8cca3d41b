// drx_internal.hpp -- objects shared between the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>

namespace drx {

// One row block of a share: rows [r0, next.r0) whose first entry lies in
// [c*block_nnz, (c+1)*block_nnz). Entries [e0, es) are reduced from the LDS
// window; if es < next.e0 the block's last row is longer than the window and
// entries [es, next.e0) are reduced by the whole workgroup. Record nblocks is
// the end sentinel {nrows, end, end}.
struct alignas(32) BlockRec {
  int64_t r0, e0, es, pad;
};

struct PanelSet;

// Analysis of one CSR share.
struct Plan {
  int itype = 0;
  const void *rowptr = nullptr;
  int64_t base = 0, nrows = 0, ncols = 0, nnz = 0;
  int has_sentinel = 0;
  int64_t block_nnz = 1024;
  int64_t cap = 2048;  // LDS window (entries) = 2 * block_nnz
  int64_t nblocks = 0;
  BlockRec *rec = nullptr;  // device, nblocks + 1
  // Blocks of 2048 entries (window 4096) for plans of 1024-entry blocks:
  // the forward kernel's layout for 4-byte values (16 entries per lane).
  BlockRec *rec_wide = nullptr;
  int64_t nblocks_wide = 0;
  // Longest row of the share (-1: not measured); sets the block pitch.
  int64_t lmax = -1;
  bool owns_async = false;
  int device = 0;
  PanelSet *panels = nullptr;  // column-panel layout (drx_plan_set_panels)
  // Cached column locality of the share for one colind pointer (k >= 2
  // path choice): -1 unknown, else bit 0 = tight (row blocks span about as
  // many columns as entries), bit 1 = loose (within ncols / 16).
  // Atomics: calls sharing the plan from several threads read and write the
  // cache without a lock. A torn read only picks the other, equally correct,
  // k >= 2 path once.
  mutable std::atomic<const void *> loc_colind{nullptr};
  mutable std::atomic<int> loc_local{-1};
  // Records in a caller's buffer (init_plan_in): products take no scratch
  // (k >= 2 gathers the k columns of x) and probe nothing.
  bool no_scratch = false;
};

// Launch shape over a plan's blocks: grid workgroups walk per_xcd * 8
// logical blocks (see map_block).
struct Grid {
  int64_t per_xcd, grid;
};
Grid grid_for(int64_t nblocks);

inline size_t dtype_size(int t) { return t == 0 ? 8 : t == 1 ? 4 : 2; }

// Transposed structure (CSC) of a share and a plan over it.
struct Transpose {
  Plan *plan = nullptr;    // over the CSC (rows of A^T = columns of A)
  void *colptr = nullptr;  // ncols + 1
  void *rowind = nullptr;  // nnz, row ids local to the share
  void *perm = nullptr;    // nnz, share position of each CSC entry
  void *tvals = nullptr;   // nnz, values in CSC order (drx_transpose_set_values)
  int tvtype = -1;         // dtype of tvals
  int itype = 0;
};

int64_t tuned_block_nnz();
int64_t default_block_nnz(int itype);
// k >= 2: 1 = gather from row-major chunks of x (default), 0 = gather from
// the k columns of x directly; lane-group shift of the row-major kernel.
int tuned_mrhs();
int tuned_mrhs_gshift();
int tuned_mrhs_kp_max();  // widest chunk for 4-byte x: 0 by locality, 8, 16
// Entries per lane of the streaming loads: 0 = by type and mode, 1, or 2
// where eligible;
// the most workgroups of a per-share launch (0 = one per row block, up to
// 2^22), which then loop over blocks.
int tuned_vec();
// 1 (default): 4-byte values with 32-bit indices use the plan's 2048-entry
// blocks (window 4096, 16 entries per lane); 0: the plan's own blocks.
int tuned_wide();
int64_t tuned_grid_cap();
// Dense product workgroup tile: 0 = by size, 1 = 128 x 128, 2 = 256 x 128.
int tuned_gemm_tile();
// x staging in LDS for k = 1: 0 never, 1 over locally clustered columns, 2 always.
int tuned_xstage();
// drx_plan_auto_panels: 1 builds panels when they pay, 0 never.
int tuned_auto_panels();
// Block pitch: 1 (default) fills windows up to the longest row, 0 half windows.
int tuned_fill();
// Forward window for plain CSR: 0 by value type, else 2048 (the plan's own
// records) or 4096 (the wide records, where the plan holds them).
int tuned_fwd_cap();
// Row-ordered x gathers for plain CSR, k = 1: 0 never, 1 by the plan's row
// lengths and column locality, 2 always.
int tuned_rowgather();

// Stream-ordered scratch from the library's per-device pool.
void *scratch_alloc(size_t bytes, hipStream_t s);
void scratch_free(void *p, hipStream_t s);

// Row-major chunks of x for k >= 2 right-hand sides: chunk width KP and the
// bytes of all chunks; the copy kernel.
// Chunk width for k right-hand sides; `loose_local` allows 8-wide chunks
// for k > 8 (fewer running sums, more waves: faster when the gathers of
// neighbouring rows coalesce).
int mrhs_kp(int64_t k, int xtype, bool loose_local = false);
// Column locality of the share: bit 0 when most row blocks' windows span few
// more columns than entries (banded, 2-D stencils), bit 1 when they span at
// most ncols / 16 (3-D stencils too). Measured once per colind pointer (one
// small kernel and a wait), cached in the plan; 0 while the stream is being
// captured.
int column_locality(const Plan &p, const void *colind, hipStream_t s);
size_t mrhs_xr_bytes(int64_t ncols, int64_t k, int kp, int xtype);
void launch_x_rowmajor(int xtype, const void *x, int64_t ldx, int64_t ncols,
                       int64_t k, int kp, void *xr, hipStream_t s);

// fill: measure the longest row (waits for the stream) so that blocks fill
// their windows (block_pitch); otherwise blocks start every half window.
Plan *make_plan(int itype, const void *rowptr, int64_t base, int64_t nrows,
                int64_t ncols, int64_t nnz, int sent, int64_t block_nnz,
                hipStream_t s, bool validate, bool async_alloc, bool fill = false);
// Entries between the starts of consecutive blocks of windows of cap entries.
int64_t block_pitch(const Plan &p, int64_t cap);
void free_plan(Plan *p, hipStream_t s);
// Records a share of nnz entries needs with block_nnz-entry blocks (plus the
// 2048-entry set of 1024-entry plans), sentinels included.
int64_t records_needed(int64_t nnz, int64_t block_nnz);
// A plan of 1024-entry blocks whose records live in `buf` (buf_records
// records, device memory the caller owns); allocates nothing, launches the
// record kernels on s. Throws if the buffer is too small.
void init_plan_in(Plan &p, int itype, const void *rowptr, int64_t base, int64_t nrows,
                  int64_t ncols, int64_t nnz, int sent, BlockRec *buf,
                  int64_t buf_records, hipStream_t s);

void run_spmv(const Plan &p, int vtype, const void *colind, const void *vals,
              int xtype, const void *x, int64_t ldx, int ytype, void *y,
              int64_t ldy, int64_t k, double alpha, double beta,
              hipStream_t s);

void run_spmv_panels(const Plan &p, int vtype, const void *colind,
                     const void *vals, int xtype, const void *x, int64_t ldx,
                     int ytype, void *y, int64_t ldy, int64_t k, double alpha,
                     double beta, hipStream_t s);
// One column panel: its segments add alpha * sum into y[row]; a row's first
// segment (row-map top bit) writes beta * y + alpha * sum instead.
void run_spmv_rowmap(const Plan &p, int vtype, const void *colind,
                     const void *vals, const void *rowmap, int xtype,
                     const void *x, int64_t ldx, int ytype, void *y,
                     int64_t ldy, int64_t k, double alpha, double beta,
                     hipStream_t s);
void free_panels(PanelSet *ps);
// A plan with panels is called with the colind/vals it was built from (or
// last updated to); throws otherwise.
void check_panel_operands(const Plan &p, int vtype, const void *colind,
                          const void *vals);

// Forward kernel reading vals through a permutation (vals[perm[e]]), used for
// the transposed structure.
void run_spmv_perm(const Plan &p, int vtype, const void *colind,
                   const void *vals, const void *perm, int xtype,
                   const void *x, int64_t ldx, int ytype, void *y,
                   int64_t ldy, int64_t k, double alpha, double beta,
                   hipStream_t s);

}  // namespace drx
