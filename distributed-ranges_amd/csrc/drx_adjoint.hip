// drx_adjoint.hip -- transposed structure (CSC of a share) and the adjoint
// dx = A^T dy, built for gfx950.
//
// The transpose is a stable radix sort of the share's entries by column
// (hipCUB, keys = column, values = entry position), so within a column the
// entries keep their row order and the structure is deterministic. The
// adjoint then runs the forward kernel over the transposed structure, reading
// vals through the permutation (vals[perm[p]]), so new vals (a training step)
// need no rebuild.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "drx_common.hpp"
#include "drx_internal.hpp"

namespace drx {

template <typename TI>
__global__ void k_row_ids(const TI *rowptr, int64_t base, int64_t nrows,
                          int64_t nnz, int sent, TI *rowid) {
  int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  // last row r with row_begin(r) <= e
  int64_t lo = 0, hi = nrows - 1;
  while (lo < hi) {
    int64_t mid = lo + (hi - lo + 1) / 2;
    if (row_begin(rowptr, base, nrows, nnz, sent, mid) <= e)
      lo = mid;
    else
      hi = mid - 1;
  }
  rowid[e] = static_cast<TI>(lo);
}

template <typename TI>
__global__ void k_iota(TI *p, int64_t n) {
  int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = static_cast<TI>(i);
}

template <typename TI>
__global__ void k_colptr(const TI *sorted_cols, int64_t nnz, int64_t ncols,
                         TI *colptr) {
  int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c > ncols) return;
  int64_t lo = 0, hi = nnz;  // first p with sorted_cols[p] >= c
  while (lo < hi) {
    int64_t mid = lo + (hi - lo) / 2;
    if (static_cast<int64_t>(sorted_cols[mid]) >= c)
      hi = mid;
    else
      lo = mid + 1;
  }
  colptr[c] = static_cast<TI>(lo);
}

template <typename TI>
__global__ void k_gather_rows(const TI *rowid, const TI *perm, int64_t nnz,
                              TI *rowind) {
  int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p < nnz) rowind[p] = rowid[static_cast<int64_t>(perm[p])];
}

template <typename TI, typename TV>
__global__ void k_permute_vals(const TI *perm, const TV *vals, int64_t nnz,
                               TV *tvals) {
  for (int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       p < nnz; p += static_cast<int64_t>(gridDim.x) * blockDim.x)
    tvals[p] = vals[static_cast<int64_t>(perm[p])];
}

static inline unsigned blocks_for(int64_t n, int tpb) {
  return static_cast<unsigned>((n + tpb - 1) / tpb);
}

template <typename TI>
static void build_transpose(Transpose &t, const Plan &p, const TI *colind,
                            hipStream_t s) {
  const int64_t nnz = p.nnz, ncols = p.ncols;
  const int tpb = 256;
  const size_t isz = sizeof(TI);
  DRX_HIP(hipMalloc(&t.colptr, (ncols + 1) * isz));
  DRX_HIP(hipMalloc(&t.rowind, std::max<int64_t>(nnz, 1) * isz));
  DRX_HIP(hipMalloc(&t.perm, std::max<int64_t>(nnz, 1) * isz));
  if (nnz > 0) {
    TI *rowid = nullptr, *keys_out = nullptr, *iota = nullptr;
    DRX_HIP(hipMalloc(&rowid, nnz * isz));
    DRX_HIP(hipMalloc(&keys_out, nnz * isz));
    DRX_HIP(hipMalloc(&iota, nnz * isz));
    hipLaunchKernelGGL(k_row_ids<TI>, dim3(blocks_for(nnz, tpb)), dim3(tpb), 0,
                       s, static_cast<const TI *>(p.rowptr), p.base, p.nrows,
                       nnz, p.has_sentinel, rowid);
    hipLaunchKernelGGL(k_iota<TI>, dim3(blocks_for(nnz, tpb)), dim3(tpb), 0, s,
                       iota, nnz);
    int end_bit = 1;
    while (end_bit < static_cast<int>(8 * isz) &&
           (int64_t(1) << end_bit) < std::max<int64_t>(ncols, 2))
      end_bit++;
    size_t temp_bytes = 0;
    DRX_HIP(hipcub::DeviceRadixSort::SortPairs(
        nullptr, temp_bytes, colind, keys_out, iota, static_cast<TI *>(t.perm),
        static_cast<int>(nnz), 0, end_bit, s));
    void *temp = nullptr;
    DRX_HIP(hipMalloc(&temp, std::max<size_t>(temp_bytes, 1)));
    DRX_HIP(hipcub::DeviceRadixSort::SortPairs(
        temp, temp_bytes, colind, keys_out, iota, static_cast<TI *>(t.perm),
        static_cast<int>(nnz), 0, end_bit, s));
    hipLaunchKernelGGL(k_colptr<TI>, dim3(blocks_for(ncols + 1, tpb)),
                       dim3(tpb), 0, s, keys_out, nnz, ncols,
                       static_cast<TI *>(t.colptr));
    hipLaunchKernelGGL(k_gather_rows<TI>, dim3(blocks_for(nnz, tpb)), dim3(tpb),
                       0, s, rowid, static_cast<const TI *>(t.perm), nnz,
                       static_cast<TI *>(t.rowind));
    DRX_HIP(hipGetLastError());
    DRX_HIP(hipStreamSynchronize(s));
    (void)hipFree(temp);
    (void)hipFree(rowid);
    (void)hipFree(keys_out);
    (void)hipFree(iota);
  } else {
    DRX_HIP(hipMemsetAsync(t.colptr, 0, (ncols + 1) * isz, s));
    DRX_HIP(hipStreamSynchronize(s));
  }
  t.plan = make_plan(p.itype, t.colptr, 0, ncols, p.nrows, nnz, 1, 0, s,
                     false, false, true);
  DRX_HIP(hipStreamSynchronize(s));
}

static void free_transpose(Transpose *t) {
  if (!t) return;
  if (t->plan) free_plan(t->plan, nullptr);
  if (t->colptr) (void)hipFree(t->colptr);
  if (t->rowind) (void)hipFree(t->rowind);
  if (t->perm) (void)hipFree(t->perm);
  if (t->tvals) (void)hipFree(t->tvals);
  delete t;
}

}  // namespace drx

using namespace drx;

extern "C" {

int drx_transpose_create(drx_transpose *out, drx_plan plan,
                         const void *colind, drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(out != nullptr && plan != nullptr, "null argument");
    *out = nullptr;
    const Plan &p = *reinterpret_cast<Plan *>(plan);
    DRX_REQUIRE(p.nnz == 0 || colind != nullptr, "colind is null");
    DRX_REQUIRE(p.nnz <= INT32_MAX, "radix sort is limited to 2^31-1 entries");
    hipStream_t s = static_cast<hipStream_t>(stream);
    Transpose *t = new Transpose();
    t->itype = p.itype;
    try {
      if (p.itype == DRX_I32)
        build_transpose<int32_t>(*t, p, static_cast<const int32_t *>(colind), s);
      else
        build_transpose<int64_t>(*t, p, static_cast<const int64_t *>(colind), s);
    } catch (...) {
      free_transpose(t);
      throw;
    }
    *out = reinterpret_cast<drx_transpose>(t);
  });
}

int drx_transpose_destroy(drx_transpose t) {
  return guarded([&] { free_transpose(reinterpret_cast<Transpose *>(t)); });
}

int drx_transpose_arrays(drx_transpose t, const void **colptr,
                         const void **rowind, const void **perm) {
  return guarded([&] {
    DRX_REQUIRE(t != nullptr, "transpose is null");
    Transpose *tt = reinterpret_cast<Transpose *>(t);
    if (colptr) *colptr = tt->colptr;
    if (rowind) *rowind = tt->rowind;
    if (perm) *perm = tt->perm;
  });
}

int drx_transpose_set_values(drx_transpose t, int vtype, const void *vals,
                             drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(t != nullptr, "transpose is null");
    Transpose *tt = reinterpret_cast<Transpose *>(t);
    const int64_t nnz = tt->plan->nnz;
    DRX_REQUIRE(nnz == 0 || vals != nullptr, "vals is null");
    DRX_REQUIRE(vtype >= DRX_F64 && vtype <= DRX_F16, "bad vtype");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (tt->tvals == nullptr || tt->tvtype != vtype) {
      if (tt->tvals) (void)hipFree(tt->tvals);
      tt->tvals = nullptr;
      DRX_HIP(hipMalloc(&tt->tvals, std::max<int64_t>(nnz, 1) * dtype_size(vtype)));
      tt->tvtype = vtype;
    }
    if (nnz == 0) return;
    const dim3 g(std::min<unsigned>(blocks_for(nnz, 256), 1u << 20));
    const bool i32 = tt->itype == DRX_I32;
#define DRX_PV(TV)                                                             \
  do {                                                                         \
    if (i32)                                                                   \
      hipLaunchKernelGGL((k_permute_vals<int32_t, TV>), g, dim3(256), 0, s,    \
                         static_cast<const int32_t *>(tt->perm),               \
                         static_cast<const TV *>(vals), nnz,                   \
                         static_cast<TV *>(tt->tvals));                        \
    else                                                                       \
      hipLaunchKernelGGL((k_permute_vals<int64_t, TV>), g, dim3(256), 0, s,    \
                         static_cast<const int64_t *>(tt->perm),               \
                         static_cast<const TV *>(vals), nnz,                   \
                         static_cast<TV *>(tt->tvals));                        \
  } while (0)
    if (vtype == DRX_F64)
      DRX_PV(double);
    else if (vtype == DRX_F32)
      DRX_PV(float);
    else
      DRX_PV(uint16_t);
#undef DRX_PV
    DRX_HIP(hipGetLastError());
  });
}

int drx_spmv_adjoint(drx_transpose t, int vtype, const void *vals, int dytype,
                     const void *dy, int64_t lddy, int dxtype, void *dx,
                     int64_t lddx, int64_t k, double alpha, double beta,
                     drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(t != nullptr, "transpose is null");
    Transpose *tt = reinterpret_cast<Transpose *>(t);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (vals == nullptr) {
      DRX_REQUIRE(tt->tvals != nullptr && tt->tvtype == vtype,
                  "vals is null and no values of this dtype were set on the transpose");
      run_spmv(*tt->plan, vtype, tt->rowind, tt->tvals, dytype, dy, lddy, dxtype,
               dx, lddx, k, alpha, beta, s);
    } else {
      run_spmv_perm(*tt->plan, vtype, tt->rowind, vals, tt->perm, dytype, dy,
                    lddy, dxtype, dx, lddx, k, alpha, beta, s);
    }
  });
}

}  // extern "C"
