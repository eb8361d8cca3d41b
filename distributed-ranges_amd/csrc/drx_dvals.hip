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
#include "drx_fwd.hpp"
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
  uint32_t xbytes;  // bytes of one x column (buffer gathers)
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

// alpha * acc + beta * old, with one explicit rounding order wherever dvals
// are written (so grouped and single-entry stores give the same bits).
template <typename TA> __device__ __forceinline__ TA dv_out(TA alpha, TA acc, TA beta, TA old) {
  const TA o = alpha * acc;
  return beta != TA(0) ? __builtin_fma(beta, old, o) : o;
}

// One row block per workgroup, like the forward kernel: the window's column
// indices are streamed (coalesced, non-temporal), x gathered into registers,
// each entry's row looked up in LDS (rowid), dy of the block's rows staged
// in LDS, and dvals written with coalesced stores. A last row longer than
// the window is done by the whole workgroup. (Staging x windows in LDS
// measured slower here: band +-16 FP32 540 vs 431 us.)
//
// BUFX: x gathered through a buffer resource, every lane's T gathers issued
// before the first product (past the window a lane gathers x[0], and entries
// outside the block are never stored); a gather inside the per-entry
// predicate made each wave wait for one gather at a time. VW = 4 (FP32
// dvals, 32-bit indices, colind and dvals aligned to a group): each lane
// streams and stores groups of four consecutive entries with 16-byte
// accesses, slot s holding entry e0 - sh + s as in the forward kernel.
template <int WG, int CAP, typename TX, typename TD, typename TA, typename TI, bool BUFX = false,
          int VW = 1>
__global__ __launch_bounds__(WG) void k_dvals(DvalsArgs a) {
  static_assert(VW == 1 || (VW == 4 && sizeof(TI) == 4 && sizeof(TD) == 4),
                "groups of four: 32-bit indices and 4-byte dvals");
  constexpr int T = CAP / WG;
  // (8 rows short of CAP / 2: a 4096-entry window's LDS then fits five
  // workgroups per CU, not four)
  constexpr int ROWCAP = CAP / 2 - 8;
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
    // entry of slot u of this lane, relative to e0 (in [0, n_s) inside the block)
    const int sh = VW > 1 ? static_cast<int>(b.e0 & (VW - 1)) : 0;
    auto entry = [&](int u) -> int {
      if constexpr (VW == 1)
        return tid + u * WG;
      else
        return (tid + (u / VW) * WG) * VW + (u % VW) - sh;
    };
    uint32_t c[T];
    if constexpr (VW > 1) {
      const uint32_t ngrp = static_cast<uint32_t>(n_s + sh + VW - 1) / VW;
      const rsrc_t rc = make_rsrc(colind + (b.e0 - sh), ngrp * VW * 4);
#pragma unroll
      for (int g = 0; g < T / VW; g++) {
        const Words<VW> cw = buf_load<Words<VW>, kNT>(rc, (tid + g * WG) * (VW * 4));
#pragma unroll
        for (int i = 0; i < VW; i++) c[g * VW + i] = cw.w[i];
      }
    } else {
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
      if constexpr (BUFX) {
        const rsrc_t rx = make_rsrc(xj, a.xbytes);
        TA xg[T];
#pragma unroll
        for (int u = 0; u < T; u++) xg[u] = to_acc<TA>(buf_load<TX>(rx, c[u] * sizeof(TX)));
#pragma unroll
        for (int u = 0; u < T; u++) {
          const int q = entry(u);
          if (q >= 0 && q < n_s) {
            const int ri = rowid[q];
            const TA d = rows_lds ? dy_s[ri] : to_acc<TA>(dyj[ri]);
            acc[u] += d * xg[u];
          }
        }
      } else {
#pragma unroll
        for (int u = 0; u < T; u++) {
          const int q = entry(u);
          if (q >= 0 && q < n_s) {
            const int ri = rowid[q];
            const TA d = rows_lds ? dy_s[ri] : to_acc<TA>(dyj[ri]);
            acc[u] += d * to_acc<TA>(xj[c[u]]);
          }
        }
      }
      __syncthreads();
    }
    if constexpr (VW > 1) {
      // a group inside the block: one 16-byte store; the first and last
      // groups of the block, entry by entry
#pragma unroll
      for (int g = 0; g < T / VW; g++) {
        const int q0 = (tid + g * WG) * VW - sh;
        if (q0 >= 0 && q0 + VW <= n_s) {
          TD *p = dv + b.e0 + q0;
          Words<VW> o;
          if (beta != TA(0)) {
            const Words<VW> old = *reinterpret_cast<const Words<VW> *>(p);
#pragma unroll
            for (int i = 0; i < VW; i++)
              o.w[i] = __builtin_bit_cast(uint32_t, static_cast<TD>(dv_out<TA>(
                  alpha, acc[g * VW + i], beta,
                  static_cast<TA>(__builtin_bit_cast(TD, old.w[i])))));
          } else {
#pragma unroll
            for (int i = 0; i < VW; i++)
              o.w[i] = __builtin_bit_cast(uint32_t, static_cast<TD>(alpha * acc[g * VW + i]));
          }
          *reinterpret_cast<Words<VW> *>(p) = o;
        } else {
#pragma unroll
          for (int i = 0; i < VW; i++) {
            const int q = q0 + i;
            if (q >= 0 && q < n_s) {
              const TA old = beta != TA(0) ? static_cast<TA>(dv[b.e0 + q]) : TA(0);
              dv[b.e0 + q] = static_cast<TD>(dv_out<TA>(alpha, acc[g * VW + i], beta, old));
            }
          }
        }
      }
    } else {
#pragma unroll
      for (int u = 0; u < T; u++) {
        const int idx = tid + u * WG;
        if (idx < n_s) {
          const TA old = beta != TA(0) ? static_cast<TA>(dv[b.e0 + idx]) : TA(0);
          dv[b.e0 + idx] = static_cast<TD>(dv_out<TA>(alpha, acc[u], beta, old));
        }
      }
    }
    if (b.es < e1) {
      const int64_t last = r1 - 1;
      for (int64_t q = b.es + tid; q < e1; q += WG) {
        const int64_t col = static_cast<int64_t>(ld_stream(colind + q));
        TA s = 0;
        for (int64_t j = 0; j < a.k; j++)
          s += to_acc<TA>(dy[last + j * a.lddy]) * to_acc<TA>(x[col + j * a.ldx]);
        const TA old = beta != TA(0) ? static_cast<TA>(dv[q]) : TA(0);
        dv[q] = static_cast<TD>(dv_out<TA>(alpha, s, beta, old));
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Row-ordered dvals (one right-hand side; shares of rows of at most 128
// entries): the block's column indices are copied straight into LDS, G lanes
// per row gather x in row order with the row's dy in a register (no row-id
// table, no dy staging) and put dy x in LDS at the entry's slot, and the
// block's dvals are then written in entry order with 16-byte stores.
// ---------------------------------------------------------------------------
template <int WG, int CAP, typename TI, typename TA> struct DvalsRowsSmem {
  static constexpr int ROWCAP = CAP / 4;
  alignas(16) TI c[CAP];
  alignas(16) TA o[CAP];
  int rp_s[ROWCAP + 1];
};

template <int WG, int G, int U, typename TX, typename TA, typename TI>
__device__ __forceinline__ void dvals_row_pass(const TI *c_s, TA *o_s, const int *rp_s,
                                               bool rows_lds, const DvalsArgs &a, int64_t r0,
                                               int64_t e0, int64_t nr) {
  const TI *rowptr = static_cast<const TI *>(a.rowptr);
  const TX *dy = static_cast<const TX *>(a.dy);
  const rsrc_t rx = make_rsrc(a.x, a.xbytes);
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
    const TA d = to_acc<TA>(dy[r0 + i]);
    for (int q = s + lane; q < e; q += G * U) {
      uint32_t cc[U];
      TA xx[U];
#pragma unroll
      for (int u = 0; u < U; u++) cc[u] = col32(c_s[q + u * G < e ? q + u * G : q]);
#pragma unroll
      for (int u = 0; u < U; u++) xx[u] = to_acc<TA>(buf_load<TX>(rx, cc[u] * sizeof(TX)));
#pragma unroll
      for (int u = 0; u < U; u++)
        if (q + u * G < e) o_s[q + u * G] = d * xx[u];
    }
  }
}

template <int WG, int CAP, typename TI, typename TA> constexpr int dvals_rows_waves() {
  constexpr int lds = static_cast<int>(sizeof(DvalsRowsSmem<WG, CAP, TI, TA>));
  constexpr int w = (160 * 1024) / lds * (WG / 256);
  return w > 8 ? 8 : (w < 1 ? 1 : w);
}

// VW4: FP32 dvals aligned to 16 bytes, stored in groups of four entries
// (the block's first and last groups entry by entry).
template <int WG, int CAP, typename TX, typename TD, typename TA, typename TI, bool VW4>
__global__ __launch_bounds__(WG)
__attribute__((amdgpu_waves_per_eu(dvals_rows_waves<WG, CAP, TI, TA>(), 8)))
void k_dvals_rows(DvalsArgs a) {
  using Sm = DvalsRowsSmem<WG, CAP, TI, TA>;
  constexpr int ROWCAP = Sm::ROWCAP;
  constexpr int NI = (ROWCAP + WG) / WG;
  static_assert(!VW4 || sizeof(TD) == 4, "groups of four 4-byte dvals");
  __shared__ Sm sm;
  int tid;
  asm volatile("" : "=v"(tid) : "0"(static_cast<int>(threadIdx.x)));
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
    const int shc = lds_window<WG, CAP>(colind, b.e0, n_s, sm.c, tid);
    if (rows_lds) {
      TI raw[NI];
      rows_offsets_load<WG, NI>(a, b.r0, nr, raw, tid);
      rows_offsets_store<WG, NI>(a, b.r0, b.e0, nr, raw, sm.rp_s, tid);
    }
    // the LDS copy counts as a vector-memory load
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    // o[slot] = dy[row] x[col] for slots [shc, shc + n_s)
    if (nr > 0) {
      const TI *cs = sm.c + shc;
      TA *os = sm.o + shc;
      const int64_t avg = n_s / nr;
      if (avg <= 12)
        dvals_row_pass<WG, 1, 8, TX, TA, TI>(cs, os, sm.rp_s, rows_lds, a, b.r0, b.e0, nr);
      else if (avg <= 40)
        dvals_row_pass<WG, 8, 5, TX, TA, TI>(cs, os, sm.rp_s, rows_lds, a, b.r0, b.e0, nr);
      else
        dvals_row_pass<WG, 16, 5, TX, TA, TI>(cs, os, sm.rp_s, rows_lds, a, b.r0, b.e0, nr);
    }
    __syncthreads();
    if constexpr (VW4) {
      // groups of four from the 16-byte-aligned entry at or below e0
      const int sh4 = static_cast<int>(b.e0 & 3);
      const int ng = (n_s + sh4 + 3) / 4;
      for (int g = tid; g < ng; g += WG) {
        const int q0 = g * 4 - sh4;  // block-relative entry of the group's first slot
        if (q0 >= 0 && q0 + 4 <= n_s) {
          TD *p = dv + b.e0 + q0;
          Words<4> o;
          if (beta != TA(0)) {
            const Words<4> old = *reinterpret_cast<const Words<4> *>(p);
#pragma unroll
            for (int i = 0; i < 4; i++)
              o.w[i] = __builtin_bit_cast(uint32_t, static_cast<TD>(dv_out<TA>(
                  alpha, sm.o[shc + q0 + i], beta,
                  static_cast<TA>(__builtin_bit_cast(TD, old.w[i])))));
          } else {
#pragma unroll
            for (int i = 0; i < 4; i++)
              o.w[i] = __builtin_bit_cast(uint32_t, static_cast<TD>(alpha * sm.o[shc + q0 + i]));
          }
          *reinterpret_cast<Words<4> *>(p) = o;
        } else {
#pragma unroll
          for (int i = 0; i < 4; i++) {
            const int q = q0 + i;
            if (q >= 0 && q < n_s) {
              const TA old = beta != TA(0) ? static_cast<TA>(dv[b.e0 + q]) : TA(0);
              dv[b.e0 + q] = static_cast<TD>(dv_out<TA>(alpha, sm.o[shc + q], beta, old));
            }
          }
        }
      }
    } else {
      for (int q = tid; q < n_s; q += WG) {
        const TA old = beta != TA(0) ? static_cast<TA>(dv[b.e0 + q]) : TA(0);
        dv[b.e0 + q] = static_cast<TD>(dv_out<TA>(alpha, sm.o[shc + q], beta, old));
      }
    }
    if (b.es < e1) {
      const int64_t last = r1 - 1;
      const TA d = to_acc<TA>(dy[last]);
      for (int64_t q = b.es + tid; q < e1; q += WG) {
        const TA old = beta != TA(0) ? static_cast<TA>(dv[q]) : TA(0);
        dv[q] = static_cast<TD>(
            dv_out<TA>(alpha, d * to_acc<TA>(x[static_cast<int64_t>(ld_stream(colind + q))]),
                       beta, old));
      }
    }
    __syncthreads();  // the next block overwrites the LDS window
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
  uint32_t xbytes;  // bytes of x (buffer gathers of the row-ordered kernel)
};

// Rows [ra, rb) of a block: row offsets (relative to e0) and targets from
// LDS (rows_lds) or, for blocks with more rows than the LDS holds, from
// memory.
template <int G, typename TA, typename TY, typename TI>
__device__ __forceinline__ void fused_rows(const TA *prod, const TA *xg, const int *rp_s,
                                           const TA *t_s, bool rows_lds,
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
    int64_t s, e;
    TA tv;
    if (rows_lds) {
      s = rp_s[r - ra];
      e = rp_s[r - ra + 1];
      tv = t_s[r - ra];
    } else {
      s = row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r) - e0;
      e = row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r + 1) - e0;
      tv = to_acc<TA>(t[r]);
    }
    TA acc = 0;
    for (int64_t q = s + lane; q < e; q += G) acc += prod[q];
    acc = group_sum<G>(acc);
    const TA rr = acc - tv;
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

// VW = 4 (not with XS; 32-bit indices, four values in at most 16 bytes,
// vals / colind aligned to a group): lanes stream groups of four
// consecutive entries with 16-byte loads, slot s of prod / xg holding entry
// e0 - sh + s as in the forward kernel.
template <int CAP, typename TV, typename TX, typename TY, typename TA,
          typename TI, bool XS = false, int VW = 1>
__global__ __launch_bounds__(kWG) void k_fused_residual(FusedArgs a) {
  static_assert(VW == 1 || (VW == 4 && !XS && sizeof(TI) == 4 && 4 * sizeof(TV) <= 16),
                "groups of four: no x staging, 32-bit indices, at most 16 bytes of values");
  constexpr int T = CAP / kWG;
  // row offsets and targets of blocks of fewer than CAP / 4 - 8 rows,
  // staged once per block: read per row pass from memory they cost each
  // pass a round trip (8 rows short, a 2048-entry window's LDS fits eight
  // workgroups per CU)
  constexpr int ROWCAP = CAP / 4 - 8;
  __shared__ TA prod[CAP];
  __shared__ TA xg[CAP];
  __shared__ int rp_s[ROWCAP + 1];
  __shared__ TA t_s[ROWCAP];
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
    const int sh = VW > 1 ? static_cast<int>(b.e0 & (VW - 1)) : 0;
    auto slot = [&](int u) -> int {
      if constexpr (VW == 1)
        return tid + u * kWG;
      else
        return (tid + (u / VW) * kWG) * VW + (u % VW);
    };
    if constexpr (VW > 1) {
      const int64_t ea = b.e0 - sh;
      const uint32_t ngrp = static_cast<uint32_t>(n_s + sh + VW - 1) / VW;
      constexpr uint32_t VB = VW * sizeof(TV);
      const rsrc_t rv = make_rsrc(vals + ea, ngrp * VB);
      const rsrc_t rc = make_rsrc(colind + ea, ngrp * VW * 4);
#pragma unroll
      for (int g = 0; g < T / VW; g++) {
        const uint32_t gi = tid + g * kWG;
        const Words<VW> cw = buf_load<Words<VW>, kNT>(rc, gi * (VW * 4));
        const Words<(VB + 3) / 4> vw = buf_load<Words<(VB + 3) / 4>, kNT>(rv, gi * VB);
        TV t[VW];
        __builtin_memcpy(t, vw.w, VB);
#pragma unroll
        for (int i = 0; i < VW; i++) {
          c[g * VW + i] = cw.w[i];
          w[g * VW + i] = to_acc<TA>(t[i]);
        }
      }
    } else {
      const rsrc_t rv = make_rsrc(vals + b.e0, n_s * sizeof(TV));
      const rsrc_t rc = make_rsrc(colind + b.e0, n_s * sizeof(TI));
#pragma unroll
      for (int u = 0; u < T; u++) {
        const uint32_t idx = tid + u * kWG;
        w[u] = to_acc<TA>(buf_load<TV, kNT>(rv, idx * sizeof(TV)));
        c[u] = col32(buf_load<TI, kNT>(rc, idx * sizeof(TI)));
      }
    }
    const int64_t nr = rs_end - b.r0;
    const bool rows_lds = nr < ROWCAP;
    if (rows_lds) {
      const TI *rowptr = static_cast<const TI *>(a.rowptr);
      const TY *t = static_cast<const TY *>(a.t);
      for (int i = tid; i <= nr; i += kWG) {
        rp_s[i] = static_cast<int>(row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, b.r0 + i) - b.e0);
        if (i < nr) t_s[i] = to_acc<TA>(t[b.r0 + i]);
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
      if constexpr (VW > 1) {
        // Buffer reads, range-checked: the last group's indices past the
        // block end can lie past the array's last entry (inside its
        // aligned 16 bytes) and hold any value.
        const rsrc_t rx = make_rsrc(x, a.xbytes);
#pragma unroll
        for (int u = 0; u < T; u++) xv[u] = to_acc<TA>(buf_load<TX>(rx, c[u] * sizeof(TX)));
      } else {
#pragma unroll
        for (int u = 0; u < T; u++) xv[u] = to_acc<TA>(x[c[u]]);
      }
    } else {
#pragma unroll
      for (int u = 0; u < T; u++) xv[u] = 0;
    }
#pragma unroll
    for (int u = 0; u < T; u++) {
      xg[slot(u)] = xv[u];
      prod[slot(u)] = w[u] * xv[u];
    }
    __syncthreads();
    if (nr > 0) {
      const int64_t avg = n_s / nr;
      const TA *pr = prod + sh, *xq = xg + sh;
      if (avg <= 3)
        fused_rows<1, TA, TY, TI>(pr, xq, rp_s, t_s, rows_lds, a, b.r0, rs_end, b.e0);
      else if (avg <= 12)
        fused_rows<4, TA, TY, TI>(pr, xq, rp_s, t_s, rows_lds, a, b.r0, rs_end, b.e0);
      else if (avg <= 48)
        fused_rows<16, TA, TY, TI>(pr, xq, rp_s, t_s, rows_lds, a, b.r0, rs_end, b.e0);
      else
        fused_rows<64, TA, TY, TI>(pr, xq, rp_s, t_s, rows_lds, a, b.r0, rs_end, b.e0);
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

// ---------------------------------------------------------------------------
// Row-ordered fused pass (shares of rows of at most 128 entries, the window
// copied straight into LDS as in k_spmv_rows): G lanes per row sum it in
// entry order (the forward kernel's order, so y equals its bits), the
// group's lane 0 writes y and r = y - t, and each lane puts r x of its own
// entries in LDS at their slots from the x values it gathered (rows longer
// than one round gather x again for the later rounds); the block's dvals
// are then stored in entry order.
// ---------------------------------------------------------------------------
template <int WG, int G, int U, typename TV, typename TX, typename TY, typename TA, typename TI>
__device__ __forceinline__ void fused_row_pass(const TV *v_s, const TI *c_s, TA *o_s,
                                               const int *rp_s, bool rows_lds,
                                               const FusedArgs &a, int64_t r0, int64_t e0,
                                               int64_t nr) {
  const TI *rowptr = static_cast<const TI *>(a.rowptr);
  const TY *t = static_cast<const TY *>(a.t);
  TY *y = static_cast<TY *>(a.y);
  TY *res = static_cast<TY *>(a.r);
  const rsrc_t rx = make_rsrc(a.x, a.xbytes);
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
    const TA tv = to_acc<TA>(t[r0 + i]);
    TA acc = 0;
    TA x0[U];
    const int q0 = s + lane;
    if (q0 < e) {
      uint32_t cc[U];
      TA vv[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int qq = q0 + u * G < e ? q0 + u * G : q0;
        cc[u] = col32(c_s[qq]);
        vv[u] = to_acc<TA>(v_s[qq]);
      }
#pragma unroll
      for (int u = 0; u < U; u++) x0[u] = to_acc<TA>(buf_load<TX>(rx, cc[u] * sizeof(TX)));
#pragma unroll
      for (int u = 0; u < U; u++)
        if (q0 + u * G < e) acc += vv[u] * x0[u];
    }
    for (int q = q0 + G * U; q < e; q += G * U) {
      uint32_t cc[U];
      TA vv[U], xx[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int qq = q + u * G < e ? q + u * G : q;
        cc[u] = col32(c_s[qq]);
        vv[u] = to_acc<TA>(v_s[qq]);
      }
#pragma unroll
      for (int u = 0; u < U; u++) xx[u] = to_acc<TA>(buf_load<TX>(rx, cc[u] * sizeof(TX)));
#pragma unroll
      for (int u = 0; u < U; u++)
        if (q + u * G < e) acc += vv[u] * xx[u];
    }
    acc = group_sum<G>(acc);
    const TA rr = acc - tv;
    if (lane == 0) {
      y[r0 + i] = from_acc<TY>(acc);
      res[r0 + i] = from_acc<TY>(rr);
    }
    if (q0 < e) {
#pragma unroll
      for (int u = 0; u < U; u++)
        if (q0 + u * G < e) o_s[q0 + u * G] = rr * x0[u];
    }
    for (int q = q0 + G * U; q < e; q += G * U) {
      uint32_t cc[U];
      TA xx[U];
#pragma unroll
      for (int u = 0; u < U; u++) cc[u] = col32(c_s[q + u * G < e ? q + u * G : q]);
#pragma unroll
      for (int u = 0; u < U; u++) xx[u] = to_acc<TA>(buf_load<TX>(rx, cc[u] * sizeof(TX)));
#pragma unroll
      for (int u = 0; u < U; u++)
        if (q + u * G < e) o_s[q + u * G] = rr * xx[u];
    }
  }
}

template <int WG, int CAP, typename TV, typename TI, typename TA> struct FusedRowsSmem {
  static constexpr int ROWCAP = CAP / 4;
  alignas(16) TV v[CAP];
  alignas(16) TI c[CAP];
  alignas(16) TA o[CAP];  // r x of every entry, stored in entry order
  int rp_s[ROWCAP + 1];
  TA red[WG / 64];
};

// As many workgroups as the LDS holds, at most 7 waves (6 for 16-bit
// values, 5 for 8-byte sums): the x values a lane keeps for its dvals spill
// registers at one wave more.
template <int WG, int CAP, typename TV, typename TI, typename TA> constexpr int fused_rows_waves() {
  constexpr int lds = static_cast<int>(sizeof(FusedRowsSmem<WG, CAP, TV, TI, TA>));
  constexpr int w = (160 * 1024) / lds * (WG / 256);
  constexpr int most = sizeof(TA) == 8 ? 5 : (sizeof(TV) == 2 ? 6 : 7);
  return w > most ? most : (w < 1 ? 1 : w);
}

template <int WG, int CAP, typename TV, typename TX, typename TY, typename TA, typename TI>
__global__ __launch_bounds__(WG)
__attribute__((amdgpu_waves_per_eu(fused_rows_waves<WG, CAP, TV, TI, TA>(), 8)))
void k_fused_rows(FusedArgs a) {
  using Sm = FusedRowsSmem<WG, CAP, TV, TI, TA>;
  constexpr int ROWCAP = Sm::ROWCAP;
  constexpr int NI = (ROWCAP + WG) / WG;
  __shared__ Sm sm;
  int tid;
  asm volatile("" : "=v"(tid) : "0"(static_cast<int>(threadIdx.x)));
  const TI *colind = static_cast<const TI *>(a.colind);
  const TV *vals = static_cast<const TV *>(a.vals);
  const TX *x = static_cast<const TX *>(a.x);
  for (int64_t wi = blockIdx.x; wi < a.per_xcd * kXcds; wi += a.grid) {
    const int64_t blk = map_block(wi, a.per_xcd);
    if (blk >= a.nblocks) continue;
    const BlockRec b = a.rec[blk];
    const int64_t r1 = a.rec[blk + 1].r0, e1 = a.rec[blk + 1].e0;
    if (b.r0 >= r1) continue;
    const int n_s = static_cast<int>(b.es - b.e0);
    const int64_t nr = (b.es == e1 ? r1 : r1 - 1) - b.r0;
    const bool rows_lds = nr < ROWCAP;
    const int shv = lds_window<WG, CAP>(vals, b.e0, n_s, sm.v, tid);
    const int shc = lds_window<WG, CAP>(colind, b.e0, n_s, sm.c, tid);
    if (rows_lds) {
      TI raw[NI];
      rows_offsets_load<WG, NI>(a, b.r0, nr, raw, tid);
      rows_offsets_store<WG, NI>(a, b.r0, b.e0, nr, raw, sm.rp_s, tid);
    }
    // the LDS copies count as vector-memory loads
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (nr > 0) {
      const TV *vs = sm.v + shv;
      const TI *cs = sm.c + shc;
      TA *os = sm.o;
      const int64_t avg = n_s / nr;
      if (avg <= 12)
        fused_row_pass<WG, 1, 8, TV, TX, TY, TA, TI>(vs, cs, os, sm.rp_s, rows_lds, a, b.r0, b.e0,
                                                     nr);
      else if (avg <= 40)
        fused_row_pass<WG, 8, 5, TV, TX, TY, TA, TI>(vs, cs, os, sm.rp_s, rows_lds, a, b.r0, b.e0,
                                                     nr);
      else
        fused_row_pass<WG, 16, 5, TV, TX, TY, TA, TI>(vs, cs, os, sm.rp_s, rows_lds, a, b.r0,
                                                      b.e0, nr);
    }
    __syncthreads();
    // dvals of the block's rows in entry order: groups of four (16-byte
    // stores) when they are 4-byte values aligned to 16 bytes
    {
      TY *dv = static_cast<TY *>(a.dvals);
      if (sizeof(TY) == 4 && reinterpret_cast<uintptr_t>(dv) % 16 == 0) {
        const int sh4 = static_cast<int>(b.e0 & 3);
        const int ng = (n_s + sh4 + 3) / 4;
        for (int g = tid; g < ng; g += WG) {
          const int q0 = g * 4 - sh4;
          if (q0 >= 0 && q0 + 4 <= n_s) {
            Words<4> o;
#pragma unroll
            for (int i = 0; i < 4; i++)
              o.w[i] = __builtin_bit_cast(uint32_t, static_cast<float>(from_acc<TY>(sm.o[q0 + i])));
            *reinterpret_cast<Words<4> *>(dv + b.e0 + q0) = o;
          } else {
#pragma unroll
            for (int i = 0; i < 4; i++)
              if (q0 + i >= 0 && q0 + i < n_s) dv[b.e0 + q0 + i] = from_acc<TY>(sm.o[q0 + i]);
          }
        }
      } else {
        for (int q = tid; q < n_s; q += WG) dv[b.e0 + q] = from_acc<TY>(sm.o[q]);
      }
    }
    if (b.es < e1) {
      // Long last row: the whole workgroup (the sum order of the forward
      // kernels' long-row path).
      const int64_t last = r1 - 1;
      TA acc0 = 0, acc1 = 0;
      int64_t q = b.es + tid;
      for (; q + 3 * WG < e1; q += 4 * WG) {
        const TV w0 = ld_stream(vals + q), w1 = ld_stream(vals + q + WG);
        const TV w2 = ld_stream(vals + q + 2 * WG), w3 = ld_stream(vals + q + 3 * WG);
        const TI c0 = ld_stream(colind + q), c1 = ld_stream(colind + q + WG);
        const TI c2 = ld_stream(colind + q + 2 * WG), c3 = ld_stream(colind + q + 3 * WG);
        const TA x0 = to_acc<TA>(x[static_cast<int64_t>(c0)]);
        const TA x1 = to_acc<TA>(x[static_cast<int64_t>(c1)]);
        const TA x2 = to_acc<TA>(x[static_cast<int64_t>(c2)]);
        const TA x3 = to_acc<TA>(x[static_cast<int64_t>(c3)]);
        acc0 += to_acc<TA>(w0) * x0;
        acc1 += to_acc<TA>(w1) * x1;
        acc0 += to_acc<TA>(w2) * x2;
        acc1 += to_acc<TA>(w3) * x3;
      }
      for (; q < e1; q += WG)
        acc0 += to_acc<TA>(ld_stream(vals + q)) *
                to_acc<TA>(x[static_cast<int64_t>(ld_stream(colind + q))]);
      const TA sum = block_sum<WG>(acc0 + acc1, sm.red);
      const TA rr = sum - to_acc<TA>(static_cast<const TY *>(a.t)[last]);
      if (tid == 0) {
        static_cast<TY *>(a.y)[last] = from_acc<TY>(sum);
        static_cast<TY *>(a.r)[last] = from_acc<TY>(rr);
      }
      TY *dv = static_cast<TY *>(a.dvals);
      for (int64_t q2 = b.es + tid; q2 < e1; q2 += WG)
        dv[q2] = from_acc<TY>(rr * to_acc<TA>(x[static_cast<int64_t>(colind[q2])]));
    }
    __syncthreads();  // the next block overwrites the LDS window
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
  // x gathered through a buffer resource when one column fits 32-bit offsets
  const uint64_t xb = static_cast<uint64_t>(std::max<int64_t>(p.ncols, 1)) * dtype_size(xtype);
  const bool bufx = xb < (uint64_t(1) << 32);
  a.xbytes = bufx ? static_cast<uint32_t>(xb) : 0;
  const dim3 grid(static_cast<unsigned>(a.grid));
  const bool i32 = p.itype == DRX_I32;
  // Row-ordered dvals for the shares the forward product sums in row order
  // (knob rowgather: 1 by the plan's rows, 3 always, 0 and 2 never), one
  // right-hand side, colind aligned to 16 bytes, the plan's 2048-entry
  // windows.
  {
    const int rg = tuned_rowgather();
    const bool rows = k == 1 && bufx && p.cap == 2048 &&
                      reinterpret_cast<uintptr_t>(colind) % 16 == 0 &&
                      (rg == 3 || (rg == 1 && p.lmax > 0 && p.lmax <= 128 &&
                                   p.nnz >= 4 * p.nrows));
    if (rows) {
      const bool v4 = dvtype == DRX_F32 && reinterpret_cast<uintptr_t>(dvals) % 16 == 0;
#define DRX_DR(TX, TD, TA)                                                                   \
  do {                                                                                       \
    if (i32 && v4)                                                                           \
      hipLaunchKernelGGL((k_dvals_rows<256, 2048, TX, TD, TA, int32_t, std::is_same_v<TD, float>>), \
                         grid, dim3(256), 0, s, a);                                          \
    else if (i32)                                                                            \
      hipLaunchKernelGGL((k_dvals_rows<256, 2048, TX, TD, TA, int32_t, false>), grid, dim3(256), \
                         0, s, a);                                                           \
    else if (v4)                                                                             \
      hipLaunchKernelGGL((k_dvals_rows<256, 2048, TX, TD, TA, int64_t, std::is_same_v<TD, float>>), \
                         grid, dim3(256), 0, s, a);                                          \
    else                                                                                     \
      hipLaunchKernelGGL((k_dvals_rows<256, 2048, TX, TD, TA, int64_t, false>), grid, dim3(256), \
                         0, s, a);                                                           \
  } while (0)
      if (xtype == DRX_F64 && dvtype == DRX_F64)
        DRX_DR(double, double, double);
      else if (xtype == DRX_F32 && dvtype == DRX_F32)
        DRX_DR(float, float, float);
      else if (xtype == DRX_BF16 && dvtype == DRX_F32)
        DRX_DR(bf16, float, float);
      else if (xtype == DRX_F16 && dvtype == DRX_F32)
        DRX_DR(f16, float, float);
      else
        throw Error(DRX_ERR_UNSUPPORTED, "unsupported (xtype, dvtype)");
#undef DRX_DR
      DRX_HIP(hipGetLastError());
      return;
    }
  }
#define DRX_DV_I(WG, CAP, TX, TD, TA)                                          \
  do {                                                                         \
    if (i32 && bufx)                                                           \
      hipLaunchKernelGGL((k_dvals<WG, CAP, TX, TD, TA, int32_t, true>), grid,  \
                         dim3(WG), 0, s, a);                                   \
    else if (i32)                                                              \
      hipLaunchKernelGGL((k_dvals<WG, CAP, TX, TD, TA, int32_t>), grid,        \
                         dim3(WG), 0, s, a);                                   \
    else if (bufx)                                                             \
      hipLaunchKernelGGL((k_dvals<WG, CAP, TX, TD, TA, int64_t, true>), grid,  \
                         dim3(WG), 0, s, a);                                   \
    else                                                                       \
      hipLaunchKernelGGL((k_dvals<WG, CAP, TX, TD, TA, int64_t>), grid,        \
                         dim3(WG), 0, s, a);                                   \
  } while (0)
  // FP32 with 32-bit indices and rows of >= 16 entries on average: windows
  // of 4096 entries, 16 per lane (band +-16: 2575 vs 2331 GB/s with gathers
  // one at a time; the 7-entry rows of a 3-D Laplacian ran slower that way,
  // 2036 vs 2303)
  const bool wide = xtype == DRX_F32 && dvtype == DRX_F32 && i32 && bufx &&
                    p.rec_wide != nullptr && tuned_wide() != 0 && p.nnz >= 16 * p.nrows;
  if (wide) {
    a.rec = p.rec_wide;
    a.nblocks = p.nblocks_wide;
    const Grid gw = grid_for(p.nblocks_wide);
    a.per_xcd = gw.per_xcd;
    a.grid = gw.grid;
    // groups of four entries per lane when colind and dvals are aligned to
    // a group (knob vec 0 or 4)
    const int vec = tuned_vec();
    const bool g4 = (vec == 0 || vec == 4) && reinterpret_cast<uintptr_t>(colind) % 16 == 0 &&
                    reinterpret_cast<uintptr_t>(dvals) % 16 == 0;
    if (g4)
      hipLaunchKernelGGL((k_dvals<256, 4096, float, float, float, int32_t, true, 4>),
                         dim3(static_cast<unsigned>(a.grid)), dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL((k_dvals<256, 4096, float, float, float, int32_t, true>),
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
  // Row-ordered pass with the window copied straight into LDS for the
  // shares the forward product sums in row order (knob rowgather: 1 by the
  // plan's rows, 3 always, 0 and 2 never). Measured against the
  // entry-ordered kernel (us per call, one box): band +-16 FP32 337-366 ->
  // 313-318, FP64 588 -> 577; 3-D Laplacian 256^3 FP32 379 -> 331, FP64 654
  // -> 603. (Storing each lane's dvals straight from the row pass instead of
  // staging them for entry-ordered stores measured slower than the
  // entry-ordered kernel: band FP32 374 vs 353 us.)
  {
    const int rg = tuned_rowgather();
    const uint64_t xb = static_cast<uint64_t>(std::max<int64_t>(p.ncols, 1)) * dtype_size(xtype);
    const uintptr_t vw = dtype_size(vtype) >= 4 ? 16 : 4;
    const bool rows = p.nnz > 0 && !a.xstage && p.cap == 2048 && xb < (uint64_t(1) << 32) &&
                      (rg == 3 || (rg == 1 && p.lmax > 0 && p.lmax <= 128 &&
                                   p.nnz >= 4 * p.nrows)) &&
                      reinterpret_cast<uintptr_t>(colind) % 16 == 0 &&
                      reinterpret_cast<uintptr_t>(vals) % vw == 0;
    if (rows) {
      a.xbytes = static_cast<uint32_t>(xb);
#define DRX_FR(TV, TX, TY, TA)                                                               \
  do {                                                                                       \
    if (i32)                                                                                 \
      hipLaunchKernelGGL((k_fused_rows<256, 2048, TV, TX, TY, TA, int32_t>), grid, block, 0, s, a); \
    else                                                                                     \
      hipLaunchKernelGGL((k_fused_rows<256, 2048, TV, TX, TY, TA, int64_t>), grid, block, 0, s, a); \
  } while (0)
      if (vtype == DRX_F64 && xtype == DRX_F64 && ytype == DRX_F64)
        DRX_FR(double, double, double, double);
      else if (vtype == DRX_F32 && xtype == DRX_F32 && ytype == DRX_F32)
        DRX_FR(float, float, float, float);
      else if (vtype == DRX_BF16 && xtype == DRX_BF16 && ytype == DRX_F32)
        DRX_FR(bf16, bf16, float, float);
      else if (vtype == DRX_F16 && xtype == DRX_F16 && ytype == DRX_F32)
        DRX_FR(f16, f16, float, float);
      else
        throw Error(DRX_ERR_UNSUPPORTED, "unsupported (vtype, xtype, ytype)");
#undef DRX_FR
      DRX_HIP(hipGetLastError());
      return;
    }
  }
  // groups of four entries per lane: 32-bit indices, vals / colind aligned
  // to a group, no x staging (knob vec 0 or 4), x within 4 GB (its reads
  // are range-checked buffer loads)
  const int vec = tuned_vec();
  const uint64_t xbytes = static_cast<uint64_t>(std::max<int64_t>(p.ncols, 1)) * dtype_size(xtype);
  if (xbytes < (uint64_t(1) << 32)) a.xbytes = static_cast<uint32_t>(xbytes);
  const bool g4 = i32 && !a.xstage && (vec == 0 || vec == 4) && dtype_size(vtype) <= 4 &&
                  xbytes < (uint64_t(1) << 32) &&
                  reinterpret_cast<uintptr_t>(colind) % 16 == 0 &&
                  reinterpret_cast<uintptr_t>(vals) % (4 * dtype_size(vtype)) == 0;
#define DRX_FU_X(TV, TX, TY, TA, XS)                                           \
  do {                                                                         \
    if constexpr (!(XS) && 4 * sizeof(TV) <= 16) {                             \
      if (g4) {                                                                \
        if (small)                                                             \
          hipLaunchKernelGGL((k_fused_residual<1024, TV, TX, TY, TA, int32_t, false, 4>), \
                             grid, block, 0, s, a);                            \
        else                                                                   \
          hipLaunchKernelGGL((k_fused_residual<2048, TV, TX, TY, TA, int32_t, false, 4>), \
                             grid, block, 0, s, a);                            \
        break;                                                                 \
      }                                                                        \
    }                                                                          \
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
