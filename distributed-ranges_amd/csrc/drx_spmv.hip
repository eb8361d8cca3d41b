// drx_spmv.hip -- CSR sparse gemv forward kernel for gfx950 (MI355X) and its
// C entry points.
//
// y[:, j] = alpha * A x[:, j] + beta * y[:, j] over one CSR share.
//
// One workgroup per row block of the plan (rows whose first entry lies in one
// block_nnz-sized window of the share):
//   1. read the block's 32-byte record (rows, entry window) -- nothing else
//      is needed before the streaming loads start;
//   2. stream vals/colind of the window with coalesced buffer loads (lane l
//      reads entry e0 + l + WG*u; 32-bit offsets from a wave-uniform
//      resource, non-temporal so L2 keeps x), and stage the window's row
//      offsets in LDS at the same time;
//   3. gather x (buffer loads, 32-bit offsets) and store the products to LDS;
//   4. reduce the rows from LDS with lane groups of G = 1..64 lanes picked
//      from the block's mean row length: DPP row sums inside 16 lanes,
//      shuffles across them; one write per row, no atomics;
//   5. a last row longer than the window is reduced by the whole workgroup
//      (strided coalesced loads, DPP wave sums, LDS across waves).
// Workgroup i takes block (i % 8) * ceil(nb / 8) + i / 8: each XCD (which
// receives every 8th workgroup) walks a contiguous range of rows, so its L2
// keeps the x lines that neighbouring rows share.
//
// Reference kernels replaced: csr_row_distribution.hpp:84-115,
// csr_eq_distribution.hpp:62-111, sp/algorithms/matrix/local_gemv.hpp:22-78.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "drx_common.hpp"
#include "drx_internal.hpp"

namespace drx {

// Kernel modes: plain CSR; vals read through a permutation (the transposed
// structure of the adjoint); rows written through a row map and accumulated
// (one column panel of a panel-layout plan).
enum FwdMode { kPlain = 0, kPerm = 1, kRowMap = 2 };

struct FwdArgs {
  const void *rowptr;
  int64_t base, nrows, nnz;
  int sent;
  const BlockRec *rec;
  int64_t nblocks, per_xcd, grid;
  const void *colind, *vals, *x;
  const void *perm;    // kPerm: vals[perm[e]]
  const void *rowmap;  // kRowMap: output row of share row r
  int64_t ldx;
  uint32_t xbytes;   // bytes of one x column (buffer gathers)
  void *y;
  int64_t ldy, k;
  double alpha, beta;
};

template <int WG, typename TA>
__device__ __forceinline__ TA block_sum(TA v, TA *red) {
  v = group_sum<64>(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  TA s = 0;
#pragma unroll
  for (int i = 0; i < WG / 64; i++) s += red[i];
  return s;
}

// Rows [0, nr) of the block (row r0 + i), products prod[0 .. n_s), row
// offsets in rp_s (relative to e0) or, for blocks with more rows than the
// LDS holds, read from rowptr.
template <int WG, int G, int MODE, typename TA, typename TY, typename TI>
__device__ __forceinline__ void reduce_rows(const TA *prod, const int *rp_s,
                                            const int64_t *map_s, const TA *ypre,
                                            bool rows_lds, const FwdArgs &a,
                                            int64_t r0, int64_t e0, int64_t nr,
                                            TY *yj, TA alpha, TA beta) {
  const TI *rowptr = static_cast<const TI *>(a.rowptr);
  const TI *rowmap = static_cast<const TI *>(a.rowmap);
  const int lane = threadIdx.x & (G - 1);
  constexpr int NG = WG / G;
  for (int64_t i = threadIdx.x / G; i < nr; i += NG) {
    int s, e;
    if (rows_lds) {
      s = rp_s[i];
      e = rp_s[i + 1];
    } else {
      s = static_cast<int>(row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r0 + i) - e0);
      e = static_cast<int>(row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r0 + i + 1) - e0);
    }
    TA acc = 0;
    for (int q = s + lane; q < e; q += G) acc += prod[q];
    acc = group_sum<G>(acc);
    if (lane == 0) {
      if constexpr (MODE == kRowMap) {
        if (ypre != nullptr && rows_lds) {
          yj[map_s[i]] = from_acc<TY>(ypre[i] + alpha * acc);
        } else {
          const int64_t row = static_cast<int64_t>(rowmap[r0 + i]);
          yj[row] = from_acc<TY>(to_acc<TA>(yj[row]) + alpha * acc);
        }
      } else {
        TA out = alpha * acc;
        if (beta != TA(0)) out += beta * to_acc<TA>(yj[r0 + i]);
        yj[r0 + i] = from_acc<TY>(out);
      }
    }
  }
}

template <int WG, int MODE, typename TA, typename TY, typename TI>
__device__ __forceinline__ void reduce_rows_any(int n_s, const TA *prod,
                                                const int *rp_s, const int64_t *map_s,
                                                const TA *ypre, bool rows_lds,
                                                const FwdArgs &a, int64_t r0,
                                                int64_t e0, int64_t nr, TY *yj,
                                                TA alpha, TA beta) {
  if (nr <= 0) return;
  const int64_t avg = n_s / nr;
#define DRX_RR(G) reduce_rows<WG, G, MODE, TA, TY, TI>(prod, rp_s, map_s, ypre, rows_lds, a, r0, e0, nr, yj, alpha, beta)
  if (avg <= 3)
    DRX_RR(1);
  else if (avg <= 6)
    DRX_RR(2);
  else if (avg <= 12)
    DRX_RR(4);
  else if (avg <= 24)
    DRX_RR(8);
  else if (avg <= 48)
    DRX_RR(16);
  else if (avg <= 96)
    DRX_RR(32);
  else
    DRX_RR(64);
#undef DRX_RR
}

template <int MODE, typename TV, typename TI>
__device__ __forceinline__ TV long_val(const TV *vals, const TI *perm,
                                       int64_t q) {
  if constexpr (MODE == kPerm)
    return vals[static_cast<int64_t>(ld_stream(perm + q))];
  else
    return ld_stream(vals + q);
}

template <int WG, int CAP, int MODE, bool BUFX, typename TV, typename TX,
          typename TY, typename TA, typename TI>
__global__ __launch_bounds__(WG) void k_spmv_fwd(FwdArgs a) {
  constexpr int T = CAP / WG;
  constexpr int ROWCAP = CAP / 2;
  constexpr int MAPCAP = MODE == kRowMap ? ROWCAP : 1;
  __shared__ TA prod[CAP];
  __shared__ int rp_s[ROWCAP + 1];
  __shared__ int64_t map_s[MAPCAP];  // kRowMap: output rows of the block
  __shared__ TA ypre[MAPCAP];        // kRowMap, k = 1: their current y
  __shared__ TA red[WG / 64];
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
    const BlockRec b = a.rec[blk];
    const int64_t r1 = a.rec[blk + 1].r0, e1 = a.rec[blk + 1].e0;
    if (b.r0 >= r1) continue;
    const int n_s = static_cast<int>(b.es - b.e0);
    const int64_t nr = (b.es == e1 ? r1 : r1 - 1) - b.r0;
    const bool rows_lds = nr < ROWCAP;

    // Stream the window (reads past n_s return 0 and touch no memory).
    TA v[T];
    uint32_t c[T];
    {
      const rsrc_t rc = make_rsrc(colind + b.e0, n_s * sizeof(TI));
      if constexpr (MODE == kPerm) {
        const rsrc_t rp = make_rsrc(perm + b.e0, n_s * sizeof(TI));
#pragma unroll
        for (int u = 0; u < T; u++) {
          const uint32_t idx = tid + u * WG;
          c[u] = col32(buf_load<TI, kNT>(rc, idx * sizeof(TI)));
          const int64_t pe = static_cast<int64_t>(buf_load<TI, kNT>(rp, idx * sizeof(TI)));
          v[u] = idx < static_cast<uint32_t>(n_s) ? to_acc<TA>(vals[pe]) : TA(0);
        }
      } else {
        const rsrc_t rv = make_rsrc(vals + b.e0, n_s * sizeof(TV));
#pragma unroll
        for (int u = 0; u < T; u++) {
          const uint32_t idx = tid + u * WG;
          v[u] = to_acc<TA>(buf_load<TV, kNT>(rv, idx * sizeof(TV)));
          c[u] = col32(buf_load<TI, kNT>(rc, idx * sizeof(TI)));
        }
      }
    }
    if (rows_lds) {
      for (int i = tid; i <= nr; i += WG) {
        rp_s[i] = static_cast<int>(
            row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, b.r0 + i) - b.e0);
        if constexpr (MODE == kRowMap) {
          if (i < nr) {
            const int64_t row = static_cast<int64_t>(static_cast<const TI *>(a.rowmap)[b.r0 + i]);
            map_s[i] = row;
            if (a.k == 1) ypre[i] = to_acc<TA>(y[row]);
          }
        }
      }
    }

    for (int64_t j = 0; j < a.k; j++) {
      const TX *xj = x + j * a.ldx;
      if constexpr (BUFX) {
        const rsrc_t rx = make_rsrc(xj, a.xbytes);
#pragma unroll
        for (int u = 0; u < T; u++) {
          const int idx = tid + u * WG;
          if (idx < n_s) prod[idx] = v[u] * to_acc<TA>(buf_load<TX>(rx, c[u] * sizeof(TX)));
        }
      } else {
#pragma unroll
        for (int u = 0; u < T; u++) {
          const int idx = tid + u * WG;
          if (idx < n_s) prod[idx] = v[u] * to_acc<TA>(xj[c[u]]);
        }
      }
      __syncthreads();
      reduce_rows_any<WG, MODE, TA, TY, TI>(n_s, prod, rp_s, map_s,
                                            MODE == kRowMap && a.k == 1 ? ypre : nullptr,
                                            rows_lds, a, b.r0, b.e0,
                                      nr, y + j * a.ldy, alpha, beta);
      __syncthreads();
    }

    if (b.es < e1) {
      // Long last row: the whole workgroup, strided coalesced loads.
      const int64_t last = r1 - 1;
      for (int64_t j = 0; j < a.k; j++) {
        const TX *xj = x + j * a.ldx;
        TA acc0 = 0, acc1 = 0;
        int64_t q = b.es + tid;
        for (; q + 3 * WG < e1; q += 4 * WG) {
          const TV w0 = long_val<MODE>(vals, perm, q);
          const TV w1 = long_val<MODE>(vals, perm, q + WG);
          const TV w2 = long_val<MODE>(vals, perm, q + 2 * WG);
          const TV w3 = long_val<MODE>(vals, perm, q + 3 * WG);
          const TI c0 = ld_stream(colind + q), c1 = ld_stream(colind + q + WG);
          const TI c2 = ld_stream(colind + q + 2 * WG), c3 = ld_stream(colind + q + 3 * WG);
          const TA x0 = to_acc<TA>(xj[static_cast<int64_t>(c0)]);
          const TA x1 = to_acc<TA>(xj[static_cast<int64_t>(c1)]);
          const TA x2 = to_acc<TA>(xj[static_cast<int64_t>(c2)]);
          const TA x3 = to_acc<TA>(xj[static_cast<int64_t>(c3)]);
          acc0 += to_acc<TA>(w0) * x0;
          acc1 += to_acc<TA>(w1) * x1;
          acc0 += to_acc<TA>(w2) * x2;
          acc1 += to_acc<TA>(w3) * x3;
        }
        for (; q < e1; q += WG)
          acc0 += to_acc<TA>(long_val<MODE>(vals, perm, q)) *
                  to_acc<TA>(xj[static_cast<int64_t>(ld_stream(colind + q))]);
        const TA s = block_sum<WG>(acc0 + acc1, red);
        if (tid == 0) {
          if constexpr (MODE == kRowMap) {
            const int64_t row = static_cast<int64_t>(static_cast<const TI *>(a.rowmap)[last]);
            y[row + j * a.ldy] = from_acc<TY>(to_acc<TA>(y[row + j * a.ldy]) + alpha * s);
          } else {
            TA out = alpha * s;
            if (beta != TA(0)) out += beta * to_acc<TA>(y[last + j * a.ldy]);
            y[last + j * a.ldy] = from_acc<TY>(out);
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Launch helpers and type dispatch
// ---------------------------------------------------------------------------
Grid grid_for(int64_t nblocks) {
  Grid g;
  g.per_xcd = (nblocks + kXcds - 1) / kXcds;
  const int64_t want = g.per_xcd * kXcds;
  const int64_t cap = int64_t(1) << 22;  // a multiple of 8; loop beyond it
  g.grid = std::max<int64_t>(kXcds, std::min(want, cap));
  return g;
}

template <int MODE, bool BUFX, typename TV, typename TX, typename TY,
          typename TA, typename TI>
static void launch_fwd(const Plan &p, const FwdArgs &a, hipStream_t s) {
  const dim3 grid(static_cast<unsigned>(a.grid));
  if (p.cap <= 1024)
    hipLaunchKernelGGL((k_spmv_fwd<256, 1024, MODE, BUFX, TV, TX, TY, TA, TI>),
                       grid, dim3(256), 0, s, a);
  else if (p.cap <= 2048)
    hipLaunchKernelGGL((k_spmv_fwd<256, 2048, MODE, BUFX, TV, TX, TY, TA, TI>),
                       grid, dim3(256), 0, s, a);
  else if (p.cap <= 4096)
    hipLaunchKernelGGL((k_spmv_fwd<512, 4096, MODE, BUFX, TV, TX, TY, TA, TI>),
                       grid, dim3(512), 0, s, a);
  else
    hipLaunchKernelGGL((k_spmv_fwd<1024, 8192, MODE, BUFX, TV, TX, TY, TA, TI>),
                       grid, dim3(1024), 0, s, a);
}

template <int MODE, typename TV, typename TX, typename TY, typename TA>
static void launch_fwd_t(const Plan &p, const FwdArgs &a, bool bufx,
                         hipStream_t s) {
  if constexpr (MODE == kRowMap) {
    // panels exist only when x fits a buffer resource
    if (p.itype == DRX_I32)
      launch_fwd<MODE, true, TV, TX, TY, TA, int32_t>(p, a, s);
    else
      launch_fwd<MODE, true, TV, TX, TY, TA, int64_t>(p, a, s);
    (void)bufx;
  } else if (p.itype == DRX_I32) {
    if (bufx)
      launch_fwd<MODE, true, TV, TX, TY, TA, int32_t>(p, a, s);
    else
      launch_fwd<MODE, false, TV, TX, TY, TA, int32_t>(p, a, s);
  } else {
    if (bufx)
      launch_fwd<MODE, true, TV, TX, TY, TA, int64_t>(p, a, s);
    else
      launch_fwd<MODE, false, TV, TX, TY, TA, int64_t>(p, a, s);
  }
}

template <int MODE>
static void run_spmv_impl(const Plan &p, int vtype, const void *colind,
                          const void *vals, const void *perm,
                          const void *rowmap, int xtype,
                          const void *x, int64_t ldx, int ytype, void *y,
                          int64_t ldy, int64_t k, double alpha, double beta,
                          hipStream_t s) {
  DRX_REQUIRE(k >= 1, "k must be >= 1");
  DRX_REQUIRE(p.nrows == 0 || y != nullptr, "y is null");
  DRX_REQUIRE(p.nnz == 0 || (colind && vals && x), "colind/vals/x is null");
  DRX_REQUIRE(MODE != kPerm || p.nnz == 0 || perm != nullptr, "perm is null");
  DRX_REQUIRE(k == 1 || (ldx >= p.ncols && ldy >= p.nrows),
              "ldx/ldy smaller than the operand");
  if (p.nrows == 0) return;
  FwdArgs a{};
  a.rowptr = p.rowptr;
  a.base = p.base;
  a.nrows = p.nrows;
  a.nnz = p.nnz;
  a.sent = p.has_sentinel;
  a.rec = p.rec;
  a.nblocks = p.nblocks;
  const Grid g = grid_for(p.nblocks);
  a.per_xcd = g.per_xcd;
  a.grid = g.grid;
  a.colind = colind;
  a.vals = vals;
  a.perm = perm;
  a.rowmap = rowmap;
  a.x = x;
  a.ldx = ldx;
  a.y = y;
  a.ldy = ldy;
  a.k = k;
  a.alpha = alpha;
  a.beta = beta;
  const uint64_t xb =
      static_cast<uint64_t>(std::max<int64_t>(p.ncols, 1)) * dtype_size(xtype);
  const bool bufx = xb < (uint64_t(1) << 32);
  a.xbytes = bufx ? static_cast<uint32_t>(xb) : 0;
  if (vtype == DRX_F64 && xtype == DRX_F64 && ytype == DRX_F64)
    launch_fwd_t<MODE, double, double, double, double>(p, a, bufx, s);
  else if (vtype == DRX_F32 && xtype == DRX_F32 && ytype == DRX_F32)
    launch_fwd_t<MODE, float, float, float, float>(p, a, bufx, s);
  else if (vtype == DRX_BF16 && xtype == DRX_BF16 && ytype == DRX_F32)
    launch_fwd_t<MODE, bf16, bf16, float, float>(p, a, bufx, s);
  else if (vtype == DRX_F16 && xtype == DRX_F16 && ytype == DRX_F32)
    launch_fwd_t<MODE, f16, f16, float, float>(p, a, bufx, s);
  else if (vtype == DRX_BF16 && xtype == DRX_F32 && ytype == DRX_F32)
    launch_fwd_t<MODE, bf16, float, float, float>(p, a, bufx, s);
  else if (vtype == DRX_F16 && xtype == DRX_F32 && ytype == DRX_F32)
    launch_fwd_t<MODE, f16, float, float, float>(p, a, bufx, s);
  else
    throw Error(DRX_ERR_UNSUPPORTED, "unsupported (vtype, xtype, ytype)");
  DRX_HIP(hipGetLastError());
}

void run_spmv(const Plan &p, int vtype, const void *colind, const void *vals,
              int xtype, const void *x, int64_t ldx, int ytype, void *y,
              int64_t ldy, int64_t k, double alpha, double beta,
              hipStream_t s) {
  if (p.panels) {
    run_spmv_panels(p, vtype, colind, vals, xtype, x, ldx, ytype, y, ldy, k,
                    alpha, beta, s);
    return;
  }
  run_spmv_impl<kPlain>(p, vtype, colind, vals, nullptr, nullptr, xtype, x,
                        ldx, ytype, y, ldy, k, alpha, beta, s);
}

// One column panel: rows through the panel's row map, accumulated into y.
void run_spmv_rowmap(const Plan &p, int vtype, const void *colind,
                     const void *vals, const void *rowmap, int xtype,
                     const void *x, int64_t ldx, int ytype, void *y,
                     int64_t ldy, int64_t k, double alpha, hipStream_t s) {
  run_spmv_impl<kRowMap>(p, vtype, colind, vals, nullptr, rowmap, xtype, x,
                         ldx, ytype, y, ldy, k, alpha, 1.0, s);
}

void run_spmv_perm(const Plan &p, int vtype, const void *colind,
                   const void *vals, const void *perm, int xtype,
                   const void *x, int64_t ldx, int ytype, void *y,
                   int64_t ldy, int64_t k, double alpha, double beta,
                   hipStream_t s) {
  run_spmv_impl<kPerm>(p, vtype, colind, vals, perm, nullptr, xtype, x, ldx,
                       ytype, y, ldy, k, alpha, beta, s);
}

}  // namespace drx

using namespace drx;

extern "C" {

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

}  // extern "C"
