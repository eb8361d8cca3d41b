// drx_internal.hpp -- objects shared between the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace drx {

// Analysis of one CSR share: block_row[c] .. block_row[c+1] are the rows
// whose first entry lies in [c*block_nnz, (c+1)*block_nnz) of the share.
struct Plan {
  int itype = 0;
  const void *rowptr = nullptr;
  int64_t base = 0, nrows = 0, ncols = 0, nnz = 0;
  int has_sentinel = 0;
  int64_t block_nnz = 1024;
  int64_t nblocks = 0;
  int64_t *block_row = nullptr;  // device, nblocks + 1
  bool owns_async = false;
  int device = 0;
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

// Forward kernel reading vals through a permutation (vals[perm[e]]), used for
// the transposed structure.
void run_spmv_perm(const Plan &p, int vtype, const void *colind,
                   const void *vals, const void *perm, int xtype,
                   const void *x, int64_t ldx, int ytype, void *y,
                   int64_t ldy, int64_t k, double alpha, double beta,
                   hipStream_t s);

}  // namespace drx
