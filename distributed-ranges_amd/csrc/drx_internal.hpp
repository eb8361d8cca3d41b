// drx_internal.hpp -- objects shared between the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

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
  bool owns_async = false;
  int device = 0;
  PanelSet *panels = nullptr;  // column-panel layout (drx_plan_set_panels)
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

Plan *make_plan(int itype, const void *rowptr, int64_t base, int64_t nrows,
                int64_t ncols, int64_t nnz, int sent, int64_t block_nnz,
                hipStream_t s, bool validate, bool async_alloc);
void free_plan(Plan *p, hipStream_t s);

void run_spmv(const Plan &p, int vtype, const void *colind, const void *vals,
              int xtype, const void *x, int64_t ldx, int ytype, void *y,
              int64_t ldy, int64_t k, double alpha, double beta,
              hipStream_t s);

void run_spmv_panels(const Plan &p, int vtype, const void *colind,
                     const void *vals, int xtype, const void *x, int64_t ldx,
                     int ytype, void *y, int64_t ldy, int64_t k, double alpha,
                     double beta, hipStream_t s);
void run_spmv_rowmap(const Plan &p, int vtype, const void *colind,
                     const void *vals, const void *rowmap, int xtype,
                     const void *x, int64_t ldx, int ytype, void *y,
                     int64_t ldy, int64_t k, double alpha, hipStream_t s);
void free_panels(PanelSet *ps);

// Forward kernel reading vals through a permutation (vals[perm[e]]), used for
// the transposed structure.
void run_spmv_perm(const Plan &p, int vtype, const void *colind,
                   const void *vals, const void *perm, int xtype,
                   const void *x, int64_t ldx, int ytype, void *y,
                   int64_t ldy, int64_t k, double alpha, double beta,
                   hipStream_t s);

}  // namespace drx
