// drx_dvals.hip -- gradient of y = A x with respect to the stored values
// (dvals), and the fused forward + residual + dvals pass, for gfx950.
//
// Both walk the plan's row blocks like the forward kernel. dvals streams the
// block's column indices, looks each entry's row up in LDS and writes dvals
// with coalesced stores (dy of the block's rows staged in LDS). The fused
// kernel stages both the products and the gathered x values of the window in
// LDS, so that after the row sums (and the residual r = y - t) the dvals of
// the row are written from LDS without reading colind or x a second time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "drx_common.hpp"
#include "drx_internal.hpp"

namespace drx {

constexpr int kWG = 256;

// ---------------------------------------------------------------------------
// dvals[e] = alpha * sum_j dy[row(e) + j lddy] x[col(e) + j ldx] + beta dvals[e]
// ---------------------------------------------------------------------------
struct DvalsArgs {
  const void *rowptr;
  int64_t base, nrows, nnz;
  int sent;
  const BlockRec *rec;
  int64_t nblocks, per_xcd, grid;
  const void *colind, *x;
  int64_t ldx;
  const void *dy;
  int64_t lddy, k;
  void *dvals;
  double alpha, beta;
};

// Row index (within the block) of every window entry, written by lane
// groups of G lanes per row; row offsets from LDS or, for blocks with more
// rows than the LDS holds, from rowptr.
template <int WG, int G, typename TI>
__device__ __forceinline__ void fill_rowid(const DvalsArgs &a, const int *rp_s,
                                           bool rows_lds, int64_t r0, int64_t e0,
                                           int64_t nr, int *rowid) {
  const TI *rowptr = static_cast<const TI *>(a.rowptr);
  const int lane = threadIdx.x & (G - 1);
  for (int64_t i = threadIdx.x / G; i < nr; i += WG / G) {
    int s, e;
    if (rows_lds) {
      s = rp_s[i];
      e = rp_s[i + 1];
    } else {
      s = static_cast<int>(row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r0 + i) - e0);
      e = static_cast<int>(row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r0 + i + 1) - e0);
    }
    for (int q = s + lane; q < e; q += G) rowid[q] = static_cast<int>(i);
  }
}

// One row block per workgroup, like the forward kernel: the window's column
// indices are streamed (coalesced, non-temporal) and x gathered into
// registers, each entry's row is looked up in LDS (rowid), dy of the
// block's rows is staged in LDS, and dvals is written with coalesced stores.
// A last row longer than the window is done by the whole workgroup.
// (Staging x windows in LDS measured slower here: band +-16 FP32 540 vs
// 431 us.)
template <int WG, int CAP, typename TX, typename TD, typename TA, typename TI>
__global__ __launch_bounds__(WG) void k_dvals(DvalsArgs a) {
  constexpr int T = CAP / WG;
  constexpr int ROWCAP = CAP / 2;
  __shared__ int rp_s[ROWCAP + 1];
  __shared__ int rowid[CAP];
  __shared__ TA dy_s[ROWCAP];
  const int tid = threadIdx.x;
  const TI *rowptr = static_cast<const TI *>(a.rowptr);
  const TI *colind = static_cast<const TI *>(a.colind);
  const TX *x = static_cast<const TX *>(a.x);
  const TX *dy = static_cast<const TX *>(a.dy);
  TD *dv = static_cast<TD *>(a.dvals);
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
    uint32_t c[T];
    {
      const rsrc_t rc = make_rsrc(colind + b.e0, n_s * sizeof(TI));
#pragma unroll
      for (int u = 0; u < T; u++)
        c[u] = col32(buf_load<TI, kNT>(rc, (tid + u * WG) * sizeof(TI)));
    }
    if (rows_lds) {
      for (int i = tid; i <= nr; i += WG)
        rp_s[i] = static_cast<int>(row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, b.r0 + i) - b.e0);
    }
    __syncthreads();
    if (nr > 0) {
      const int64_t avg = n_s / nr;
      if (avg <= 3)
        fill_rowid<WG, 1, TI>(a, rp_s, rows_lds, b.r0, b.e0, nr, rowid);
      else if (avg <= 12)
        fill_rowid<WG, 4, TI>(a, rp_s, rows_lds, b.r0, b.e0, nr, rowid);
      else if (avg <= 48)
        fill_rowid<WG, 16, TI>(a, rp_s, rows_lds, b.r0, b.e0, nr, rowid);
      else
        fill_rowid<WG, 64, TI>(a, rp_s, rows_lds, b.r0, b.e0, nr, rowid);
    }
    TA acc[T];
#pragma unroll
    for (int u = 0; u < T; u++) acc[u] = 0;
    for (int64_t j = 0; j < a.k; j++) {
      const TX *xj = x + j * a.ldx;
      const TX *dyj = dy + j * a.lddy + b.r0;
      if (rows_lds) {
        for (int i = tid; i < nr; i += WG) dy_s[i] = to_acc<TA>(dyj[i]);
      }
      __syncthreads();
      // (Issuing all x reads before the products, as the forward and fused
      // kernels do, needs 2T more registers here: one wave fewer, band
      // +-16 FP32 430 -> 461 us.)
#pragma unroll
      for (int u = 0; u < T; u++) {
        const int idx = tid + u * WG;
        if (idx < n_s) {
          const int ri = rowid[idx];
          const TA d = rows_lds ? dy_s[ri] : to_acc<TA>(dyj[ri]);
          acc[u] += d * to_acc<TA>(xj[c[u]]);
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < T; u++) {
      const int idx = tid + u * WG;
      if (idx < n_s) {
        TA out = alpha * acc[u];
        if (beta != TA(0)) out += beta * static_cast<TA>(dv[b.e0 + idx]);
        dv[b.e0 + idx] = static_cast<TD>(out);
      }
    }
    if (b.es < e1) {
      const int64_t last = r1 - 1;
      for (int64_t q = b.es + tid; q < e1; q += WG) {
        const int64_t col = static_cast<int64_t>(ld_stream(colind + q));
        TA s = 0;
        for (int64_t j = 0; j < a.k; j++)
          s += to_acc<TA>(dy[last + j * a.lddy]) * to_acc<TA>(x[col + j * a.ldx]);
        TA out = alpha * s;
        if (beta != TA(0)) out += beta * static_cast<TA>(dv[q]);
        dv[q] = static_cast<TD>(out);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Fused forward + residual + dvals (k = 1)
// ---------------------------------------------------------------------------
struct FusedArgs {
  const void *rowptr;
  int64_t base, nrows, nnz;
  int sent;
  const BlockRec *rec;
  int64_t nblocks, per_xcd, grid;
  const void *colind, *vals, *x, *t;
  void *y, *r, *dvals;
  int xstage;  // stage each block's x window in LDS
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
  const int lane = threadIdx.x & (G - 1);
  constexpr int NG = kWG / G;
  for (int64_t r = ra + threadIdx.x / G; r < rb; r += NG) {
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

// XS: the block's x window (its first kXsCap columns) is loaded once into
// LDS; entries read x from there, or from memory past it.
constexpr uint32_t kXsCap = 512;

template <int CAP, typename TV, typename TX, typename TY, typename TA,
          typename TI, bool XS = false>
__global__ __launch_bounds__(kWG) void k_fused_residual(FusedArgs a) {
  constexpr int T = CAP / kWG;
  __shared__ TA prod[CAP];
  __shared__ TA xg[CAP];
  __shared__ TA red[kWG / 64];
  __shared__ TX xs[XS ? kXsCap : 1];
  __shared__ uint32_t xmm[2 * (kWG / 64)];
  const int tid = threadIdx.x;
  const TI *colind = static_cast<const TI *>(a.colind);
  const TV *vals = static_cast<const TV *>(a.vals);
  const TX *x = static_cast<const TX *>(a.x);
  for (int64_t wi = blockIdx.x; wi < a.per_xcd * kXcds; wi += a.grid) {
    const int64_t blk = map_block(wi, a.per_xcd);
    if (blk >= a.nblocks) continue;
    const BlockRec b = a.rec[blk];
    const int64_t r1 = a.rec[blk + 1].r0, e1 = a.rec[blk + 1].e0;
    if (b.r0 >= r1) continue;
    const int64_t last = r1 - 1;
    const int64_t rs_end = b.es == e1 ? r1 : last;
    const int n_s = static_cast<int>(b.es - b.e0);
    // The window is streamed with range-checked buffer loads (slots past
    // n_s read 0), and every lane issues all of its x reads before it stores
    // a product: per-entry predicates made each wave wait for one load at a
    // time. Slots past n_s read x[0] and store products nothing reads.
    uint32_t c[T];
    TA w[T], xv[T];
    {
      const rsrc_t rv = make_rsrc(vals + b.e0, n_s * sizeof(TV));
      const rsrc_t rc = make_rsrc(colind + b.e0, n_s * sizeof(TI));
#pragma unroll
      for (int u = 0; u < T; u++) {
        const uint32_t idx = tid + u * kWG;
        w[u] = to_acc<TA>(buf_load<TV, kNT>(rv, idx * sizeof(TV)));
        c[u] = col32(buf_load<TI, kNT>(rc, idx * sizeof(TI)));
      }
    }
    if constexpr (XS) {
      uint32_t lo, hi, xn = 0;
      block_col_range<kWG, T>(c, n_s, xmm, lo, hi);
      if (lo <= hi) {
        xn = hi - lo + 1 < kXsCap ? hi - lo + 1 : kXsCap;
        for (uint32_t i = tid; i < xn; i += kWG) xs[i] = x[lo + i];
      }
      __syncthreads();
      bool far = false;
#pragma unroll
      for (int u = 0; u < T; u++) {
        const uint32_t off = c[u] - lo;
        xv[u] = to_acc<TA>(xs[off < xn ? off : 0]);
        far |= off >= xn && tid + u * kWG < n_s;
      }
      if (far) {
#pragma unroll
        for (int u = 0; u < T; u++)
          if (c[u] - lo >= xn && tid + u * kWG < n_s) xv[u] = to_acc<TA>(x[c[u]]);
      }
    } else if (n_s > 0) {
#pragma unroll
      for (int u = 0; u < T; u++) xv[u] = to_acc<TA>(x[c[u]]);
    } else {
#pragma unroll
      for (int u = 0; u < T; u++) xv[u] = 0;
    }
#pragma unroll
    for (int u = 0; u < T; u++) {
      const int idx = tid + u * kWG;
      xg[idx] = xv[u];
      prod[idx] = w[u] * xv[u];
    }
    __syncthreads();
    const int64_t nr = rs_end - b.r0;
    if (nr > 0) {
      const int64_t avg = n_s / nr;
      if (avg <= 3)
        fused_rows<1, TA, TY, TI>(prod, xg, a, b.r0, rs_end, b.e0);
      else if (avg <= 12)
        fused_rows<4, TA, TY, TI>(prod, xg, a, b.r0, rs_end, b.e0);
      else if (avg <= 48)
        fused_rows<16, TA, TY, TI>(prod, xg, a, b.r0, rs_end, b.e0);
      else
        fused_rows<64, TA, TY, TI>(prod, xg, a, b.r0, rs_end, b.e0);
    }
    __syncthreads();
    if (b.es < e1) {
      TA acc = 0;
      for (int64_t q = b.es + tid; q < e1; q += kWG)
        acc += to_acc<TA>(ld_stream(vals + q)) *
               to_acc<TA>(x[static_cast<int64_t>(ld_stream(colind + q))]);
      acc = group_sum<64>(acc);
      if ((tid & 63) == 0) red[tid >> 6] = acc;
      __syncthreads();
      TA s = 0;
#pragma unroll
      for (int i = 0; i < kWG / 64; i++) s += red[i];
      const TA rr = s - to_acc<TA>(static_cast<const TY *>(a.t)[last]);
      if (tid == 0) {
        static_cast<TY *>(a.y)[last] = from_acc<TY>(s);
        static_cast<TY *>(a.r)[last] = from_acc<TY>(rr);
      }
      TY *dv = static_cast<TY *>(a.dvals);
      for (int64_t q = b.es + tid; q < e1; q += kWG)
        dv[q] = from_acc<TY>(rr * to_acc<TA>(x[static_cast<int64_t>(colind[q])]));
      __syncthreads();
    }
  }
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
  a.rec = p.rec;
  a.nblocks = p.nblocks;
  const Grid g = grid_for(p.nblocks);
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
  const dim3 grid(static_cast<unsigned>(a.grid));
  const bool i32 = p.itype == DRX_I32;
#define DRX_DV_I(WG, CAP, TX, TD, TA)                                          \
  do {                                                                         \
    if (i32)                                                                   \
      hipLaunchKernelGGL((k_dvals<WG, CAP, TX, TD, TA, int32_t>), grid,        \
                         dim3(WG), 0, s, a);                                   \
    else                                                                       \
      hipLaunchKernelGGL((k_dvals<WG, CAP, TX, TD, TA, int64_t>), grid,        \
                         dim3(WG), 0, s, a);                                   \
  } while (0)
  // FP32 with 32-bit indices and rows of >= 16 entries on average: the
  // plan's 2048-entry blocks, 16 per lane (band +-16: 2575 vs 2331 GB/s; the
  // 7-entry rows of a 3-D Laplacian run slower that way, 2036 vs 2303)
  const bool wide = xtype == DRX_F32 && dvtype == DRX_F32 && i32 && p.rec_wide != nullptr &&
                    tuned_wide() != 0 && p.nnz >= 16 * p.nrows;
  if (wide) {
    a.rec = p.rec_wide;
    a.nblocks = p.nblocks_wide;
    const Grid gw = grid_for(p.nblocks_wide);
    a.per_xcd = gw.per_xcd;
    a.grid = gw.grid;
    hipLaunchKernelGGL((k_dvals<256, 4096, float, float, float, int32_t>),
                       dim3(static_cast<unsigned>(a.grid)), dim3(256), 0, s, a);
    DRX_HIP(hipGetLastError());
    return;
  }
#define DRX_DV(TX, TD, TA)                                                     \
  do {                                                                         \
    if (p.cap <= 1024)                                                         \
      DRX_DV_I(256, 1024, TX, TD, TA);                                         \
    else if (p.cap <= 2048)                                                    \
      DRX_DV_I(256, 2048, TX, TD, TA);                                         \
    else if (p.cap <= 4096)                                                    \
      DRX_DV_I(512, 4096, TX, TD, TA);                                         \
    else                                                                       \
      DRX_DV_I(1024, 8192, TX, TD, TA);                                        \
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
#undef DRX_DV_I
  DRX_HIP(hipGetLastError());
}

void run_fused(const Plan &p, int vtype, const void *colind, const void *vals,
               int xtype, const void *x, int ytype, const void *t, void *y,
               void *r, void *dvals, hipStream_t s) {
  DRX_REQUIRE(p.nrows == 0 || (y && r && t), "null output");
  DRX_REQUIRE(p.nnz == 0 || (colind && vals && x && dvals), "null operand");
  DRX_REQUIRE(p.cap == 1024 || p.cap == 2048,
              "the fused residual kernel needs a plan with block_nnz 512 or 1024");
  if (p.nrows == 0) return;
  FusedArgs a{};
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
  a.x = x;
  a.t = t;
  a.y = y;
  a.r = r;
  a.dvals = dvals;
  // x windows staged in LDS only when the knob forces it: with all of a
  // lane's x reads in flight, per-entry reads win on bands (+-16 FP32:
  // 498 vs 561 us per call).
  a.xstage = p.nnz > 0 && tuned_xstage() == 2 &&
             static_cast<uint64_t>(p.ncols) < (uint64_t(1) << 32);
  const dim3 grid(static_cast<unsigned>(a.grid)), block(kWG);
  const bool i32 = p.itype == DRX_I32;
  const bool small = p.cap == 1024;
#define DRX_FU_X(TV, TX, TY, TA, XS)                                           \
  do {                                                                         \
    if (i32 && small)                                                          \
      hipLaunchKernelGGL((k_fused_residual<1024, TV, TX, TY, TA, int32_t, XS>), \
                         grid, block, 0, s, a);                                \
    else if (i32)                                                              \
      hipLaunchKernelGGL((k_fused_residual<2048, TV, TX, TY, TA, int32_t, XS>), \
                         grid, block, 0, s, a);                                \
    else if (small)                                                            \
      hipLaunchKernelGGL((k_fused_residual<1024, TV, TX, TY, TA, int64_t, XS>), \
                         grid, block, 0, s, a);                                \
    else                                                                       \
      hipLaunchKernelGGL((k_fused_residual<2048, TV, TX, TY, TA, int64_t, XS>), \
                         grid, block, 0, s, a);                                \
  } while (0)
#define DRX_FU(TV, TX, TY, TA)                                                 \
  do {                                                                         \
    if (a.xstage && (std::is_same_v<TX, float> || std::is_same_v<TX, double>)) \
      DRX_FU_X(TV, TX, TY, TA, (std::is_same_v<TX, float> || std::is_same_v<TX, double>)); \
    else                                                                       \
      DRX_FU_X(TV, TX, TY, TA, false);                                         \
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
#undef DRX_FU_X
  DRX_HIP(hipGetLastError());
}

}  // namespace drx

using namespace drx;

extern "C" {

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
