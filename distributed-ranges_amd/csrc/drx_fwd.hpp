// drx_fwd.hpp -- the forward (y = alpha*A*x + beta*y) row-block body shared
// by the per-share kernel (drx_spmv.hip) and the batched persistent kernel
// (drx_batch.hip).
//
// A workgroup processing one row block of a share:
//   1. reads the block's 32-byte record (rows, entry window) -- nothing else
//      is needed before the streaming loads start;
//   2. streams vals/colind of the window with coalesced buffer loads (lane l
//      reads entry e0 + l + WG*u; 32-bit offsets from a wave-uniform
//      resource, non-temporal so L2 keeps x), and stages the window's row
//      offsets in LDS at the same time;
//   3. gathers x (buffer loads, 32-bit offsets) and stores the products to
//      LDS;
//   4. reduces the rows from LDS with lane groups of G = 1..64 lanes picked
//      from the block's mean row length: DPP row sums inside 16 lanes,
//      shuffles across them; one write per row, no atomics;
//   5. a last row longer than the window is reduced by the whole workgroup
//      (strided coalesced loads, DPP wave sums, LDS across waves).
#pragma once

#include <hip/hip_runtime.h>

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
  int gshift;  // lane-group choice uses (mean row length >> gshift)
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
  const int64_t avg = (n_s / nr) >> a.gshift;
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

// LDS of one workgroup of the forward kernels.
template <int WG, int CAP, bool ROWMAP, typename TA> struct FwdSmem {
  static constexpr int ROWCAP = CAP / 2;
  static constexpr int MAPCAP = ROWMAP ? ROWCAP : 1;
  TA prod[CAP];
  int rp_s[ROWCAP + 1];
  int64_t map_s[MAPCAP];  // kRowMap: output rows of the block
  TA ypre[MAPCAP];        // kRowMap, k = 1: their current y
  TA red[WG / 64];
};

// One row block `blk` of the share described by `a` (all lanes of the
// workgroup call it together).
template <int WG, int CAP, int MODE, bool BUFX, typename TV, typename TX,
          typename TY, typename TA, typename TI>
__device__ __forceinline__ void process_block(const FwdArgs &a, int64_t blk,
                                              FwdSmem<WG, CAP, MODE == kRowMap, TA> &sm) {
  constexpr int T = CAP / WG;
  constexpr int ROWCAP = CAP / 2;
  TA *prod = sm.prod;
  int *rp_s = sm.rp_s;
  int64_t *map_s = sm.map_s;
  TA *ypre = sm.ypre;
  TA *red = sm.red;
  const int tid = threadIdx.x;
  const TI *rowptr = static_cast<const TI *>(a.rowptr);
  const TI *colind = static_cast<const TI *>(a.colind);
  const TV *vals = static_cast<const TV *>(a.vals);
  const TI *perm = static_cast<const TI *>(a.perm);
  const TX *x = static_cast<const TX *>(a.x);
  TY *y = static_cast<TY *>(a.y);
  const TA alpha = static_cast<TA>(a.alpha), beta = static_cast<TA>(a.beta);

  {
    const BlockRec b = a.rec[blk];
    const int64_t r1 = a.rec[blk + 1].r0, e1 = a.rec[blk + 1].e0;
    if (b.r0 >= r1) return;
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

template <int WG, int CAP, int MODE, bool BUFX, typename TV, typename TX,
          typename TY, typename TA, typename TI>
__global__ __launch_bounds__(WG) void k_spmv_fwd(FwdArgs a) {
  __shared__ FwdSmem<WG, CAP, MODE == kRowMap, TA> sm;
  for (int64_t wi = blockIdx.x; wi < a.per_xcd * kXcds; wi += a.grid) {
    const int64_t blk = map_block(wi, a.per_xcd);
    if (blk >= a.nblocks) continue;
    process_block<WG, CAP, MODE, BUFX, TV, TX, TY, TA, TI>(a, blk, sm);
  }
}

}  // namespace drx
