// drx_spmv.hip -- CSR sparse gemv for gfx950 (MI355X): plans, forward kernel,
// dvals, fused residual, and their C entry points.
//
// Kernel design (one share, y = alpha*A*x + beta*y):
//   * A plan cuts the share's rows into blocks: block c holds the rows whose
//     first entry lies in [c*NB, (c+1)*NB) of the share (NB = block_nnz). The
//     boundaries are one binary search per block on rowptr, built on the GPU.
//   * One 256-lane workgroup per block. Rows that fit the LDS window (all but
//     possibly a long last row) are streamed: lanes load vals/colind with
//     coalesced, non-temporal loads (entry e0 + lane + 256*u), gather x, and
//     store the products to LDS. The rows are then reduced from LDS by lane
//     groups of G = 1..64 lanes chosen from the block's mean row length, with
//     DPP/shuffle sums inside each group -- short rows get one lane, long
//     rows a whole wavefront.
//   * A last row longer than the window is reduced by the whole workgroup
//     (strided coalesced loads, DPP wave sums, LDS across the 4 waves).
//   * Workgroup i takes block (i % 8) * ceil(nb / 8) + i / 8, so each XCD
//     (which receives every 8th workgroup) walks a contiguous range of rows
//     and its L2 keeps the x lines neighbouring rows share.
//   * No atomics: every output is written once; results are deterministic.
//
// Reference kernels replaced: csr_row_distribution.hpp:84-115,
// csr_eq_distribution.hpp:62-111, sp/algorithms/matrix/local_gemv.hpp:22-78.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <mutex>

#include "drx_common.hpp"
#include "drx_internal.hpp"

namespace drx {

static thread_local std::string g_last_error;
void set_last_error(const char *msg) { g_last_error = msg ? msg : ""; }

// ---------------------------------------------------------------------------
// Plan
// ---------------------------------------------------------------------------
template <typename TI>
__global__ void k_plan_blocks(const TI *rowptr, int64_t base, int64_t nrows,
                              int64_t nnz, int sent, int64_t nb_nnz,
                              int64_t nblocks, int64_t *block_row) {
  int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c > nblocks) return;
  if (c == 0) {
    block_row[0] = 0;
    return;
  }
  if (c == nblocks) {
    block_row[nblocks] = nrows;
    return;
  }
  const int64_t target = c * nb_nnz;
  int64_t lo = 0, hi = nrows;  // first r with row_begin(r) >= target
  while (lo < hi) {
    int64_t mid = lo + (hi - lo) / 2;
    if (row_begin(rowptr, base, nrows, nnz, sent, mid) >= target)
      hi = mid;
    else
      lo = mid + 1;
  }
  block_row[c] = lo;
}

int64_t default_block_nnz(int itype) {
  int64_t v = tuned_block_nnz();
  if (v > 0) return v;
  (void)itype;
  return 1024;
}

static void build_plan_blocks(Plan &p, hipStream_t s, bool async_alloc) {
  p.nblocks = p.nnz / p.block_nnz + 1;
  size_t bytes = static_cast<size_t>(p.nblocks + 1) * sizeof(int64_t);
  if (async_alloc) {
    DRX_HIP(hipMallocAsync(reinterpret_cast<void **>(&p.block_row), bytes, s));
  } else {
    DRX_HIP(hipMalloc(reinterpret_cast<void **>(&p.block_row), bytes));
  }
  p.owns_async = async_alloc;
  const int tpb = 256;
  int64_t n = p.nblocks + 1;
  unsigned grid = static_cast<unsigned>((n + tpb - 1) / tpb);
  if (p.itype == DRX_I32)
    hipLaunchKernelGGL(k_plan_blocks<int32_t>, dim3(grid), dim3(tpb), 0, s,
                       static_cast<const int32_t *>(p.rowptr), p.base, p.nrows,
                       p.nnz, p.has_sentinel, p.block_nnz, p.nblocks,
                       p.block_row);
  else
    hipLaunchKernelGGL(k_plan_blocks<int64_t>, dim3(grid), dim3(tpb), 0, s,
                       static_cast<const int64_t *>(p.rowptr), p.base, p.nrows,
                       p.nnz, p.has_sentinel, p.block_nnz, p.nblocks,
                       p.block_row);
  DRX_HIP(hipGetLastError());
}

static void validate_plan_args(int itype, const void *rowptr, int64_t nrows,
                               int64_t ncols, int64_t nnz,
                               int64_t block_nnz) {
  DRX_REQUIRE(itype == DRX_I32 || itype == DRX_I64, "itype must be I32 or I64");
  DRX_REQUIRE(nrows >= 0 && ncols >= 0 && nnz >= 0, "negative size");
  DRX_REQUIRE(nrows == 0 || rowptr != nullptr, "rowptr is null");
  DRX_REQUIRE(block_nnz <= 0 || (block_nnz >= 64 && block_nnz <= 8192 &&
                                 (block_nnz & (block_nnz - 1)) == 0),
              "block_nnz must be a power of two in [64, 8192]");
  if (itype == DRX_I32) {
    DRX_REQUIRE(nnz <= INT32_MAX && ncols <= static_cast<int64_t>(INT32_MAX) + 1,
                "sizes exceed the 32-bit index type");
  }
}

Plan *make_plan(int itype, const void *rowptr, int64_t base, int64_t nrows,
                int64_t ncols, int64_t nnz, int sent, int64_t block_nnz,
                hipStream_t s, bool validate, bool async_alloc) {
  validate_plan_args(itype, rowptr, nrows, ncols, nnz, block_nnz);
  Plan *p = new Plan();
  p->itype = itype;
  p->rowptr = rowptr;
  p->base = base;
  p->nrows = nrows;
  p->ncols = ncols;
  p->nnz = nnz;
  p->has_sentinel = sent ? 1 : 0;
  p->block_nnz = block_nnz > 0 ? block_nnz : default_block_nnz(itype);
  DRX_HIP(hipGetDevice(&p->device));
  try {
    if (validate && nrows > 0) {
      // First and last row offsets must bracket the share.
      int64_t first = 0, lastv = 0;
      size_t isz = itype == DRX_I32 ? 4 : 8;
      int64_t lastidx = sent ? nrows : nrows - 1;
      unsigned char buf[16] = {0};
      DRX_HIP(hipMemcpyAsync(buf, rowptr, isz, hipMemcpyDeviceToHost, s));
      DRX_HIP(hipMemcpyAsync(buf + 8,
                             static_cast<const char *>(rowptr) + lastidx * isz,
                             isz, hipMemcpyDeviceToHost, s));
      DRX_HIP(hipStreamSynchronize(s));
      if (isz == 4) {
        int32_t a, b;
        std::memcpy(&a, buf, 4);
        std::memcpy(&b, buf + 8, 4);
        first = a;
        lastv = b;
      } else {
        std::memcpy(&first, buf, 8);
        std::memcpy(&lastv, buf + 8, 8);
      }
      DRX_REQUIRE(first - base <= nnz && lastv >= first,
                  "rowptr does not bracket the share (check rowptr_base/nnz)");
      if (sent)
        DRX_REQUIRE(lastv - base <= nnz,
                    "rowptr[nrows] - rowptr_base exceeds nnz");
    }
    build_plan_blocks(*p, s, async_alloc);
  } catch (...) {
    delete p;
    throw;
  }
  return p;
}

void free_plan(Plan *p, hipStream_t s) {
  if (!p) return;
  if (p->block_row) {
    if (p->owns_async)
      (void)hipFreeAsync(p->block_row, s);
    else
      (void)hipFree(p->block_row);
  }
  delete p;
}

// ---------------------------------------------------------------------------
// Forward kernel
// ---------------------------------------------------------------------------
struct FwdArgs {
  const void *rowptr;
  int64_t base, nrows, nnz;
  int sent;
  const int64_t *block_row;
  int64_t nblocks, per_xcd, grid;
  const void *colind, *vals, *x;
  const void *perm;  // PERM kernels: vals are read as vals[perm[e]]
  int64_t ldx;
  void *y;
  int64_t ldy, k;
  double alpha, beta;
};

constexpr int kWG = 256;

template <bool PERM, typename TV, typename TI>
__device__ __forceinline__ TV load_val(const TV *vals, const TI *perm,
                                       int64_t q) {
  if constexpr (PERM)
    return vals[static_cast<int64_t>(ld_stream(perm + q))];
  else
    return ld_stream(vals + q);
}

// Row phase: rows [ra, rb) of the block, products of the block's entries at
// prod[0 .. ) relative to entry e0; lane groups of G.
template <int G, typename TA, typename TY, typename TI>
__device__ __forceinline__ void reduce_rows(const TA *prod, const TI *rowptr,
                                            int64_t base, int64_t nrows,
                                            int64_t nnz, int sent, int64_t ra,
                                            int64_t rb, int64_t e0, TY *yj,
                                            TA alpha, TA beta) {
  const int tid = threadIdx.x;
  const int lane = tid & (G - 1);
  constexpr int NG = kWG / G;
  for (int64_t r = ra + tid / G; r < rb; r += NG) {
    const int64_t s = row_begin(rowptr, base, nrows, nnz, sent, r) - e0;
    const int64_t e = row_begin(rowptr, base, nrows, nnz, sent, r + 1) - e0;
    TA acc = 0;
    for (int64_t q = s + lane; q < e; q += G) acc += prod[q];
    acc = group_sum<G>(acc);
    if (lane == 0) {
      TA out = alpha * acc;
      if (beta != TA(0)) out += beta * to_acc<TA>(yj[r]);
      yj[r] = from_acc<TY>(out);
    }
  }
}

template <typename TA, typename TY, typename TI>
__device__ __forceinline__ void reduce_rows_dispatch(
    int64_t n_s, const TA *prod, const TI *rowptr, int64_t base,
    int64_t nrows, int64_t nnz, int sent, int64_t ra, int64_t rb, int64_t e0,
    TY *yj, TA alpha, TA beta) {
  const int64_t nr = rb - ra;
  if (nr <= 0) return;
  const int64_t avg = n_s / nr;
  if (avg <= 3)
    reduce_rows<1>(prod, rowptr, base, nrows, nnz, sent, ra, rb, e0, yj, alpha, beta);
  else if (avg <= 6)
    reduce_rows<2>(prod, rowptr, base, nrows, nnz, sent, ra, rb, e0, yj, alpha, beta);
  else if (avg <= 12)
    reduce_rows<4>(prod, rowptr, base, nrows, nnz, sent, ra, rb, e0, yj, alpha, beta);
  else if (avg <= 24)
    reduce_rows<8>(prod, rowptr, base, nrows, nnz, sent, ra, rb, e0, yj, alpha, beta);
  else if (avg <= 48)
    reduce_rows<16>(prod, rowptr, base, nrows, nnz, sent, ra, rb, e0, yj, alpha, beta);
  else if (avg <= 96)
    reduce_rows<32>(prod, rowptr, base, nrows, nnz, sent, ra, rb, e0, yj, alpha, beta);
  else
    reduce_rows<64>(prod, rowptr, base, nrows, nnz, sent, ra, rb, e0, yj, alpha, beta);
}

template <typename TA>
__device__ __forceinline__ TA block_sum(TA v, TA *red) {
  v = group_sum<64>(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  TA s = 0;
#pragma unroll
  for (int i = 0; i < kWG / 64; i++) s += red[i];
  return s;
}

__device__ __forceinline__ int64_t map_block(int64_t i, int64_t per_xcd) {
  return (i % kXcds) * per_xcd + i / kXcds;
}

template <int CAP, bool PERM, typename TV, typename TX, typename TY,
          typename TA, typename TI>
__global__ __launch_bounds__(kWG) void k_spmv_fwd(FwdArgs a) {
  constexpr int T = CAP / kWG;
  __shared__ TA prod[CAP];
  __shared__ TA red[kWG / 64];
  const int tid = threadIdx.x;
  const TI *rowptr = static_cast<const TI *>(a.rowptr);
  const TI *colind = static_cast<const TI *>(a.colind);
  const TV *vals = static_cast<const TV *>(a.vals);
  const TI *perm = static_cast<const TI *>(a.perm);
  const TX *x = static_cast<const TX *>(a.x);
  TY *y = static_cast<TY *>(a.y);
  const TA alpha = static_cast<TA>(a.alpha), beta = static_cast<TA>(a.beta);

  for (int64_t wi = blockIdx.x; wi < a.per_xcd * kXcds; wi += a.grid) {
    const int64_t blk = map_block(wi, a.per_xcd);
    if (blk >= a.nblocks) continue;
    const int64_t r0 = a.block_row[blk], r1 = a.block_row[blk + 1];
    if (r0 >= r1) continue;
    const int64_t e0 = row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r0);
    const int64_t eN = row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r1);
    const int64_t last = r1 - 1;
    const int64_t eL = row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, last);
    const bool all_in = (eN - e0) <= CAP;
    const int64_t rs_end = all_in ? r1 : last;
    const int n_s = static_cast<int>((all_in ? eN : eL) - e0);

    // Stream the window's entries into registers (coalesced, read once).
    TA v[T];
    TI c[T];
#pragma unroll
    for (int u = 0; u < T; u++) {
      const int idx = tid + u * kWG;
      if (idx < n_s) {
        v[u] = to_acc<TA>(load_val<PERM>(vals, perm, e0 + idx));
        c[u] = ld_stream(colind + e0 + idx);
      }
    }
    for (int64_t j = 0; j < a.k; j++) {
      const TX *xj = x + j * a.ldx;
#pragma unroll
      for (int u = 0; u < T; u++) {
        const int idx = tid + u * kWG;
        if (idx < n_s) prod[idx] = v[u] * to_acc<TA>(xj[static_cast<int64_t>(c[u])]);
      }
      __syncthreads();
      reduce_rows_dispatch<TA, TY, TI>(n_s, prod, rowptr, a.base, a.nrows,
                                       a.nnz, a.sent, r0, rs_end, e0,
                                       y + j * a.ldy, alpha, beta);
      __syncthreads();
    }

    if (!all_in) {
      // Long last row: the whole workgroup.
      for (int64_t j = 0; j < a.k; j++) {
        const TX *xj = x + j * a.ldx;
        TA acc0 = 0, acc1 = 0;
        int64_t q = eL + tid;
        for (; q + 3 * kWG < eN; q += 4 * kWG) {
          TV w0 = load_val<PERM>(vals, perm, q),
             w1 = load_val<PERM>(vals, perm, q + kWG),
             w2 = load_val<PERM>(vals, perm, q + 2 * kWG),
             w3 = load_val<PERM>(vals, perm, q + 3 * kWG);
          TI c0 = ld_stream(colind + q), c1 = ld_stream(colind + q + kWG),
             c2 = ld_stream(colind + q + 2 * kWG), c3 = ld_stream(colind + q + 3 * kWG);
          TA x0 = to_acc<TA>(xj[static_cast<int64_t>(c0)]);
          TA x1 = to_acc<TA>(xj[static_cast<int64_t>(c1)]);
          TA x2 = to_acc<TA>(xj[static_cast<int64_t>(c2)]);
          TA x3 = to_acc<TA>(xj[static_cast<int64_t>(c3)]);
          acc0 += to_acc<TA>(w0) * x0;
          acc1 += to_acc<TA>(w1) * x1;
          acc0 += to_acc<TA>(w2) * x2;
          acc1 += to_acc<TA>(w3) * x3;
        }
        for (; q < eN; q += kWG)
          acc0 += to_acc<TA>(load_val<PERM>(vals, perm, q)) *
                  to_acc<TA>(xj[static_cast<int64_t>(ld_stream(colind + q))]);
        const TA s = block_sum(acc0 + acc1, red);
        if (tid == 0) {
          TA out = alpha * s;
          if (beta != TA(0)) out += beta * to_acc<TA>(y[last + j * a.ldy]);
          y[last + j * a.ldy] = from_acc<TY>(out);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// dvals kernel: dvals[e] = alpha * sum_j dy[row(e)] x[col(e)] + beta dvals[e]
// ---------------------------------------------------------------------------
struct DvalsArgs {
  const void *rowptr;
  int64_t base, nrows, nnz;
  int sent;
  const int64_t *block_row;
  int64_t nblocks, per_xcd, grid;
  const void *colind, *x;
  int64_t ldx;
  const void *dy;
  int64_t lddy, k;
  void *dvals;
  double alpha, beta;
};

template <int G, typename TX, typename TD, typename TA, typename TI>
__device__ __forceinline__ void dvals_rows(const DvalsArgs &a, int64_t ra,
                                           int64_t rb) {
  const TI *rowptr = static_cast<const TI *>(a.rowptr);
  const TI *colind = static_cast<const TI *>(a.colind);
  const TX *x = static_cast<const TX *>(a.x);
  const TX *dy = static_cast<const TX *>(a.dy);
  TD *dv = static_cast<TD *>(a.dvals);
  const TA alpha = static_cast<TA>(a.alpha), beta = static_cast<TA>(a.beta);
  const int tid = threadIdx.x;
  const int lane = tid & (G - 1);
  constexpr int NG = kWG / G;
  for (int64_t r = ra + tid / G; r < rb; r += NG) {
    const int64_t s = row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r);
    const int64_t e = row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r + 1);
    for (int64_t q = s + lane; q < e; q += G) {
      const int64_t col = static_cast<int64_t>(ld_stream(colind + q));
      TA acc = 0;
      for (int64_t j = 0; j < a.k; j++)
        acc += to_acc<TA>(dy[r + j * a.lddy]) * to_acc<TA>(x[col + j * a.ldx]);
      TA out = alpha * acc;
      if (beta != TA(0)) out += beta * static_cast<TA>(dv[q]);
      dv[q] = static_cast<TD>(out);
    }
  }
}

template <typename TX, typename TD, typename TA, typename TI>
__global__ __launch_bounds__(kWG) void k_dvals(DvalsArgs a) {
  const TI *rowptr = static_cast<const TI *>(a.rowptr);
  for (int64_t wi = blockIdx.x; wi < a.per_xcd * kXcds; wi += a.grid) {
    const int64_t blk = map_block(wi, a.per_xcd);
    if (blk >= a.nblocks) continue;
    const int64_t r0 = a.block_row[blk], r1 = a.block_row[blk + 1];
    if (r0 >= r1) continue;
    const int64_t e0 = row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r0);
    const int64_t eN = row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r1);
    const int64_t avg = (eN - e0) / (r1 - r0);
    if (avg <= 3)
      dvals_rows<1, TX, TD, TA, TI>(a, r0, r1);
    else if (avg <= 12)
      dvals_rows<4, TX, TD, TA, TI>(a, r0, r1);
    else if (avg <= 48)
      dvals_rows<16, TX, TD, TA, TI>(a, r0, r1);
    else
      dvals_rows<64, TX, TD, TA, TI>(a, r0, r1);
  }
}

// ---------------------------------------------------------------------------
// Fused forward + residual + dvals (k = 1)
// ---------------------------------------------------------------------------
struct FusedArgs {
  const void *rowptr;
  int64_t base, nrows, nnz;
  int sent;
  const int64_t *block_row;
  int64_t nblocks, per_xcd, grid;
  const void *colind, *vals, *x, *t;
  void *y, *r, *dvals;
};

template <int G, typename TA, typename TY, typename TI>
__device__ __forceinline__ void fused_rows(const TA *prod, const TA *xg,
                                           const FusedArgs &a, int64_t ra,
                                           int64_t rb, int64_t e0) {
  const TI *rowptr = static_cast<const TI *>(a.rowptr);
  const TY *t = static_cast<const TY *>(a.t);
  TY *y = static_cast<TY *>(a.y);
  TY *res = static_cast<TY *>(a.r);
  TY *dv = static_cast<TY *>(a.dvals);
  const int tid = threadIdx.x;
  const int lane = tid & (G - 1);
  constexpr int NG = kWG / G;
  for (int64_t r = ra + tid / G; r < rb; r += NG) {
    const int64_t s = row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r) - e0;
    const int64_t e = row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r + 1) - e0;
    TA acc = 0;
    for (int64_t q = s + lane; q < e; q += G) acc += prod[q];
    acc = group_sum<G>(acc);
    const TA rr = acc - to_acc<TA>(t[r]);
    if (lane == 0) {
      y[r] = from_acc<TY>(acc);
      res[r] = from_acc<TY>(rr);
    }
    for (int64_t q = s + lane; q < e; q += G) dv[e0 + q] = from_acc<TY>(rr * xg[q]);
  }
}

template <int CAP, typename TV, typename TX, typename TY, typename TA,
          typename TI>
__global__ __launch_bounds__(kWG) void k_fused_residual(FusedArgs a) {
  constexpr int T = CAP / kWG;
  __shared__ TA prod[CAP];
  __shared__ TA xg[CAP];
  __shared__ TA red[kWG / 64];
  const int tid = threadIdx.x;
  const TI *rowptr = static_cast<const TI *>(a.rowptr);
  const TI *colind = static_cast<const TI *>(a.colind);
  const TV *vals = static_cast<const TV *>(a.vals);
  const TX *x = static_cast<const TX *>(a.x);
  for (int64_t wi = blockIdx.x; wi < a.per_xcd * kXcds; wi += a.grid) {
    const int64_t blk = map_block(wi, a.per_xcd);
    if (blk >= a.nblocks) continue;
    const int64_t r0 = a.block_row[blk], r1 = a.block_row[blk + 1];
    if (r0 >= r1) continue;
    const int64_t e0 = row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r0);
    const int64_t eN = row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r1);
    const int64_t last = r1 - 1;
    const int64_t eL = row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, last);
    const bool all_in = (eN - e0) <= CAP;
    const int64_t rs_end = all_in ? r1 : last;
    const int n_s = static_cast<int>((all_in ? eN : eL) - e0);
#pragma unroll
    for (int u = 0; u < T; u++) {
      const int idx = tid + u * kWG;
      if (idx < n_s) {
        const TA w = to_acc<TA>(ld_stream(vals + e0 + idx));
        const TA xv = to_acc<TA>(x[static_cast<int64_t>(ld_stream(colind + e0 + idx))]);
        xg[idx] = xv;
        prod[idx] = w * xv;
      }
    }
    __syncthreads();
    const int64_t nr = rs_end - r0;
    if (nr > 0) {
      const int64_t avg = n_s / nr;
      if (avg <= 3)
        fused_rows<1, TA, TY, TI>(prod, xg, a, r0, rs_end, e0);
      else if (avg <= 12)
        fused_rows<4, TA, TY, TI>(prod, xg, a, r0, rs_end, e0);
      else if (avg <= 48)
        fused_rows<16, TA, TY, TI>(prod, xg, a, r0, rs_end, e0);
      else
        fused_rows<64, TA, TY, TI>(prod, xg, a, r0, rs_end, e0);
    }
    __syncthreads();
    if (!all_in) {
      TA acc = 0;
      for (int64_t q = eL + tid; q < eN; q += kWG)
        acc += to_acc<TA>(ld_stream(vals + q)) *
               to_acc<TA>(x[static_cast<int64_t>(ld_stream(colind + q))]);
      const TA s = block_sum(acc, red);
      const TA rr = s - to_acc<TA>(static_cast<const TY *>(a.t)[last]);
      if (tid == 0) {
        static_cast<TY *>(a.y)[last] = from_acc<TY>(s);
        static_cast<TY *>(a.r)[last] = from_acc<TY>(rr);
      }
      TY *dv = static_cast<TY *>(a.dvals);
      for (int64_t q = eL + tid; q < eN; q += kWG)
        dv[q] = from_acc<TY>(rr * to_acc<TA>(x[static_cast<int64_t>(colind[q])]));
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// Launch helpers and type dispatch
// ---------------------------------------------------------------------------
struct Grid {
  int64_t per_xcd, grid;
};
static Grid grid_for(int64_t nblocks) {
  Grid g;
  g.per_xcd = (nblocks + kXcds - 1) / kXcds;
  int64_t want = g.per_xcd * kXcds;
  const int64_t cap = int64_t(1) << 22;  // multiple of 8; loop beyond it
  g.grid = std::max<int64_t>(kXcds, std::min(want, cap));
  return g;
}

template <bool PERM, typename TV, typename TX, typename TY, typename TA,
          typename TI>
static void launch_fwd(const Plan &p, const FwdArgs &a, hipStream_t s) {
  const int64_t cap = 2 * p.block_nnz;
  dim3 grid(static_cast<unsigned>(a.grid)), block(kWG);
  if (cap <= 1024)
    hipLaunchKernelGGL((k_spmv_fwd<1024, PERM, TV, TX, TY, TA, TI>), grid, block, 0, s, a);
  else if (cap <= 2048)
    hipLaunchKernelGGL((k_spmv_fwd<2048, PERM, TV, TX, TY, TA, TI>), grid, block, 0, s, a);
  else if (cap <= 4096)
    hipLaunchKernelGGL((k_spmv_fwd<4096, PERM, TV, TX, TY, TA, TI>), grid, block, 0, s, a);
  else
    hipLaunchKernelGGL((k_spmv_fwd<8192, PERM, TV, TX, TY, TA, TI>), grid, block, 0, s, a);
}

template <bool PERM, typename TV, typename TX, typename TY, typename TA>
static void launch_fwd_i(const Plan &p, const FwdArgs &a, hipStream_t s) {
  if (p.itype == DRX_I32)
    launch_fwd<PERM, TV, TX, TY, TA, int32_t>(p, a, s);
  else
    launch_fwd<PERM, TV, TX, TY, TA, int64_t>(p, a, s);
}

template <bool PERM>
static void run_spmv_impl(const Plan &p, int vtype, const void *colind,
                          const void *vals, const void *perm, int xtype,
                          const void *x, int64_t ldx, int ytype, void *y,
                          int64_t ldy, int64_t k, double alpha, double beta,
                          hipStream_t s) {
  DRX_REQUIRE(k >= 1, "k must be >= 1");
  DRX_REQUIRE(p.nrows == 0 || y != nullptr, "y is null");
  DRX_REQUIRE(p.nnz == 0 || (colind && vals && x), "colind/vals/x is null");
  DRX_REQUIRE(!PERM || p.nnz == 0 || perm != nullptr, "perm is null");
  DRX_REQUIRE(k == 1 || (ldx >= p.ncols && ldy >= p.nrows),
              "ldx/ldy smaller than the operand");
  if (p.nrows == 0) return;
  FwdArgs a{};
  a.rowptr = p.rowptr;
  a.base = p.base;
  a.nrows = p.nrows;
  a.nnz = p.nnz;
  a.sent = p.has_sentinel;
  a.block_row = p.block_row;
  a.nblocks = p.nblocks;
  Grid g = grid_for(p.nblocks);
  a.per_xcd = g.per_xcd;
  a.grid = g.grid;
  a.colind = colind;
  a.vals = vals;
  a.perm = perm;
  a.x = x;
  a.ldx = ldx;
  a.y = y;
  a.ldy = ldy;
  a.k = k;
  a.alpha = alpha;
  a.beta = beta;
  if (vtype == DRX_F64 && xtype == DRX_F64 && ytype == DRX_F64)
    launch_fwd_i<PERM, double, double, double, double>(p, a, s);
  else if (vtype == DRX_F32 && xtype == DRX_F32 && ytype == DRX_F32)
    launch_fwd_i<PERM, float, float, float, float>(p, a, s);
  else if (vtype == DRX_BF16 && xtype == DRX_BF16 && ytype == DRX_F32)
    launch_fwd_i<PERM, bf16, bf16, float, float>(p, a, s);
  else if (vtype == DRX_F16 && xtype == DRX_F16 && ytype == DRX_F32)
    launch_fwd_i<PERM, f16, f16, float, float>(p, a, s);
  else if (vtype == DRX_BF16 && xtype == DRX_F32 && ytype == DRX_F32)
    launch_fwd_i<PERM, bf16, float, float, float>(p, a, s);
  else if (vtype == DRX_F16 && xtype == DRX_F32 && ytype == DRX_F32)
    launch_fwd_i<PERM, f16, float, float, float>(p, a, s);
  else
    throw Error(DRX_ERR_UNSUPPORTED, "unsupported (vtype, xtype, ytype)");
  DRX_HIP(hipGetLastError());
}

void run_spmv(const Plan &p, int vtype, const void *colind, const void *vals,
              int xtype, const void *x, int64_t ldx, int ytype, void *y,
              int64_t ldy, int64_t k, double alpha, double beta,
              hipStream_t s) {
  run_spmv_impl<false>(p, vtype, colind, vals, nullptr, xtype, x, ldx, ytype,
                       y, ldy, k, alpha, beta, s);
}

void run_spmv_perm(const Plan &p, int vtype, const void *colind,
                   const void *vals, const void *perm, int xtype,
                   const void *x, int64_t ldx, int ytype, void *y,
                   int64_t ldy, int64_t k, double alpha, double beta,
                   hipStream_t s) {
  run_spmv_impl<true>(p, vtype, colind, vals, perm, xtype, x, ldx, ytype, y,
                      ldy, k, alpha, beta, s);
}

void run_dvals(const Plan &p, const void *colind, int xtype, const void *x,
               int64_t ldx, int dytype, const void *dy, int64_t lddy,
               int64_t k, int dvtype, void *dvals, double alpha, double beta,
               hipStream_t s) {
  DRX_REQUIRE(k >= 1, "k must be >= 1");
  DRX_REQUIRE(p.nnz == 0 || (colind && x && dy && dvals), "null operand");
  DRX_REQUIRE(xtype == dytype, "x and dy must share a dtype");
  if (p.nrows == 0 || p.nnz == 0) return;
  DvalsArgs a{};
  a.rowptr = p.rowptr;
  a.base = p.base;
  a.nrows = p.nrows;
  a.nnz = p.nnz;
  a.sent = p.has_sentinel;
  a.block_row = p.block_row;
  a.nblocks = p.nblocks;
  Grid g = grid_for(p.nblocks);
  a.per_xcd = g.per_xcd;
  a.grid = g.grid;
  a.colind = colind;
  a.x = x;
  a.ldx = ldx;
  a.dy = dy;
  a.lddy = lddy;
  a.k = k;
  a.dvals = dvals;
  a.alpha = alpha;
  a.beta = beta;
  dim3 grid(static_cast<unsigned>(a.grid)), block(kWG);
  const bool i32 = p.itype == DRX_I32;
#define DRX_DV(TX, TD, TA)                                                     \
  do {                                                                         \
    if (i32)                                                                   \
      hipLaunchKernelGGL((k_dvals<TX, TD, TA, int32_t>), grid, block, 0, s, a); \
    else                                                                       \
      hipLaunchKernelGGL((k_dvals<TX, TD, TA, int64_t>), grid, block, 0, s, a); \
  } while (0)
  if (xtype == DRX_F64 && dvtype == DRX_F64)
    DRX_DV(double, double, double);
  else if (xtype == DRX_F32 && dvtype == DRX_F32)
    DRX_DV(float, float, float);
  else if (xtype == DRX_BF16 && dvtype == DRX_F32)
    DRX_DV(bf16, float, float);
  else if (xtype == DRX_F16 && dvtype == DRX_F32)
    DRX_DV(f16, float, float);
  else
    throw Error(DRX_ERR_UNSUPPORTED, "unsupported (xtype, dvtype)");
#undef DRX_DV
  DRX_HIP(hipGetLastError());
}

void run_fused(const Plan &p, int vtype, const void *colind, const void *vals,
               int xtype, const void *x, int ytype, const void *t, void *y,
               void *r, void *dvals, hipStream_t s) {
  DRX_REQUIRE(p.nrows == 0 || (y && r && t), "null output");
  DRX_REQUIRE(p.nnz == 0 || (colind && vals && x && dvals), "null operand");
  if (p.nrows == 0) return;
  FusedArgs a{};
  a.rowptr = p.rowptr;
  a.base = p.base;
  a.nrows = p.nrows;
  a.nnz = p.nnz;
  a.sent = p.has_sentinel;
  a.block_row = p.block_row;
  a.nblocks = p.nblocks;
  Grid g = grid_for(p.nblocks);
  a.per_xcd = g.per_xcd;
  a.grid = g.grid;
  a.colind = colind;
  a.vals = vals;
  a.x = x;
  a.t = t;
  a.y = y;
  a.r = r;
  a.dvals = dvals;
  dim3 grid(static_cast<unsigned>(a.grid)), block(kWG);
  const bool i32 = p.itype == DRX_I32;
  const bool small = 2 * p.block_nnz <= 1024;
  DRX_REQUIRE(2 * p.block_nnz <= 2048,
              "fused residual needs a plan with block_nnz <= 1024");
#define DRX_FU(TV, TX, TY, TA)                                                 \
  do {                                                                         \
    if (i32 && small)                                                          \
      hipLaunchKernelGGL((k_fused_residual<1024, TV, TX, TY, TA, int32_t>),    \
                         grid, block, 0, s, a);                                \
    else if (i32)                                                              \
      hipLaunchKernelGGL((k_fused_residual<2048, TV, TX, TY, TA, int32_t>),    \
                         grid, block, 0, s, a);                                \
    else if (small)                                                            \
      hipLaunchKernelGGL((k_fused_residual<1024, TV, TX, TY, TA, int64_t>),    \
                         grid, block, 0, s, a);                                \
    else                                                                       \
      hipLaunchKernelGGL((k_fused_residual<2048, TV, TX, TY, TA, int64_t>),    \
                         grid, block, 0, s, a);                                \
  } while (0)
  if (vtype == DRX_F64 && xtype == DRX_F64 && ytype == DRX_F64)
    DRX_FU(double, double, double, double);
  else if (vtype == DRX_F32 && xtype == DRX_F32 && ytype == DRX_F32)
    DRX_FU(float, float, float, float);
  else if (vtype == DRX_BF16 && xtype == DRX_BF16 && ytype == DRX_F32)
    DRX_FU(bf16, bf16, float, float);
  else if (vtype == DRX_F16 && xtype == DRX_F16 && ytype == DRX_F32)
    DRX_FU(f16, f16, float, float);
  else
    throw Error(DRX_ERR_UNSUPPORTED, "unsupported (vtype, xtype, ytype)");
#undef DRX_FU
  DRX_HIP(hipGetLastError());
}

}  // namespace drx

// ---------------------------------------------------------------------------
// C entry points
// ---------------------------------------------------------------------------
using namespace drx;

extern "C" {

int drx_version(void) { return 1; }

const char *drx_last_error(void) { return g_last_error.c_str(); }

int drx_plan_create(drx_plan *out, int itype, const void *rowptr,
                    int64_t rowptr_base, int64_t nrows, int64_t ncols,
                    int64_t nnz, int has_sentinel, int64_t block_nnz,
                    drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(out != nullptr, "plan out-pointer is null");
    *out = nullptr;
    Plan *p = make_plan(itype, rowptr, rowptr_base, nrows, ncols, nnz,
                        has_sentinel, block_nnz,
                        static_cast<hipStream_t>(stream), true, false);
    *out = reinterpret_cast<drx_plan>(p);
  });
}

int drx_plan_destroy(drx_plan plan) {
  return guarded([&] {
    if (!plan) return;
    Plan *p = reinterpret_cast<Plan *>(plan);
    free_plan(p, nullptr);
  });
}

int drx_plan_info(drx_plan plan, int64_t *nblocks, int64_t *block_nnz) {
  return guarded([&] {
    DRX_REQUIRE(plan != nullptr, "plan is null");
    Plan *p = reinterpret_cast<Plan *>(plan);
    if (nblocks) *nblocks = p->nblocks;
    if (block_nnz) *block_nnz = p->block_nnz;
  });
}

int drx_spmv(drx_plan plan, int vtype, const void *colind, const void *vals,
             int xtype, const void *x, int64_t ldx, int ytype, void *y,
             int64_t ldy, int64_t k, double alpha, double beta,
             drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(plan != nullptr, "plan is null");
    run_spmv(*reinterpret_cast<Plan *>(plan), vtype, colind, vals, xtype, x,
             ldx, ytype, y, ldy, k, alpha, beta,
             static_cast<hipStream_t>(stream));
  });
}

int drx_spmv_csr(int itype, const void *rowptr, int64_t rowptr_base,
                 int64_t nrows, int64_t ncols, int64_t nnz, int has_sentinel,
                 const void *colind, int vtype, const void *vals, int xtype,
                 const void *x, int64_t ldx, int ytype, void *y, int64_t ldy,
                 int64_t k, double alpha, double beta, drx_stream_t stream) {
  return guarded([&] {
    hipStream_t s = static_cast<hipStream_t>(stream);
    Plan *p = make_plan(itype, rowptr, rowptr_base, nrows, ncols, nnz,
                        has_sentinel, 0, s, false, true);
    try {
      run_spmv(*p, vtype, colind, vals, xtype, x, ldx, ytype, y, ldy, k,
               alpha, beta, s);
    } catch (...) {
      free_plan(p, s);
      throw;
    }
    free_plan(p, s);
  });
}

int drx_csr_dvals(drx_plan plan, const void *colind, int xtype, const void *x,
                  int64_t ldx, int dytype, const void *dy, int64_t lddy,
                  int64_t k, int dvtype, void *dvals, double alpha,
                  double beta, drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(plan != nullptr, "plan is null");
    run_dvals(*reinterpret_cast<Plan *>(plan), colind, xtype, x, ldx, dytype,
              dy, lddy, k, dvtype, dvals, alpha, beta,
              static_cast<hipStream_t>(stream));
  });
}

int drx_spmv_fused_residual(drx_plan plan, int vtype, const void *colind,
                            const void *vals, int xtype, const void *x,
                            int ytype, const void *t, void *y, void *r,
                            void *dvals, drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(plan != nullptr, "plan is null");
    run_fused(*reinterpret_cast<Plan *>(plan), vtype, colind, vals, xtype, x,
              ytype, t, y, r, dvals, static_cast<hipStream_t>(stream));
  });
}

}  // extern "C"
