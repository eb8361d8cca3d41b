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

#include <type_traits>

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
  int xstage;  // kPlain, k = 1: stage each block's x window in LDS
  int rowgather;  // kPlain, k = 1: row-ordered gathers (kernels built with RG)
  int64_t lmax;   // longest row of the share (-1 unknown): row-ordered round size
  const void *xr;  // k >= 2: x in row-major chunks of KP columns
  int64_t ncols;   // columns of the share (rows of one xr chunk)
};

// Row-map entries of column panels carry in their top bit whether the
// segment is its row's first in panel order. That segment writes
// y = beta * y + alpha * sum (y not read when beta is 0); the row's later
// segments add alpha * sum to it.
template <typename TI> __device__ __forceinline__ int64_t map_row(TI m) {
  return static_cast<int64_t>(m & static_cast<TI>(~kFirstSeg<TI>));
}
// y of a segment's row before its sum is added.
template <typename TA, typename TY, typename TI>
__device__ __forceinline__ TA map_ypre(const TY *yj, TI m, TA beta) {
  const int64_t row = map_row(m);
  if ((m & kFirstSeg<TI>) == 0) return to_acc<TA>(yj[row]);
  return beta == TA(0) ? TA(0) : beta * to_acc<TA>(yj[row]);
}

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
                                            const TI *map_s, const TA *ypre,
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
          yj[map_row(map_s[i])] = from_acc<TY>(ypre[i] + alpha * acc);
        } else {
          const TI m = rowmap[r0 + i];
          yj[map_row(m)] = from_acc<TY>(map_ypre<TA>(yj, m, beta) + alpha * acc);
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
                                                const int *rp_s, const TI *map_s,
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

// Row-ordered gathers (plain CSR, one right-hand side): one lane per row,
// the block's rows in consecutive lanes. A lane reads its row's values and
// columns from LDS (slots [rp_s[i], rp_s[i + 1]) of v_s / c_s), U at a time,
// issues their U x gathers together and adds the products in entry order.
// Where the rows are short and neighbouring rows have neighbouring columns
// (stencils), one gather instruction of a wave reads x of 64 consecutive
// rows' u-th neighbours: a few cache lines, where entry-ordered gathers of
// the same entries span every neighbour's lines. Slots past a row's end
// repeat its first entry (in range, and its x is gathered anyway) and add
// nothing.
// PERM (the adjoint through a permutation): v_s holds each entry's position
// in vals, and the values are gathered too (through a buffer resource over
// the share's nnz values of type TV).
template <int WG, int U, typename TX, typename TY, typename TA, typename TI,
          typename VS = TA, typename CS = uint32_t, bool PERM = false, typename TV = TA>
__device__ __forceinline__ void row_gather_sums(const VS *v_s, const CS *c_s,
                                                const int *rp_s, bool rows_lds,
                                                const FwdArgs &a, int64_t r0,
                                                int64_t e0, int64_t nr, TY *y,
                                                TA alpha, TA beta) {
  const TI *rowptr = static_cast<const TI *>(a.rowptr);
  const rsrc_t rx = make_rsrc(a.x, a.xbytes);
  [[maybe_unused]] const rsrc_t rv =
      make_rsrc(a.vals, PERM ? static_cast<uint32_t>(a.nnz * sizeof(TV)) : 0u);
  for (int64_t i = threadIdx.x; i < nr; i += WG) {
    int s, e;
    if (rows_lds) {
      s = rp_s[i];
      e = rp_s[i + 1];
    } else {
      s = static_cast<int>(row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r0 + i) - e0);
      e = static_cast<int>(row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r0 + i + 1) - e0);
    }
    TA acc = 0;
    for (int q = s; q < e; q += U) {
      uint32_t cc[U];
      TA vv[U], xx[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int qq = q + u < e ? q + u : q;
        cc[u] = col32(c_s[qq]);
        if constexpr (PERM)
          vv[u] = to_acc<TA>(buf_load<TV>(rv, col32(v_s[qq]) * static_cast<uint32_t>(sizeof(TV))));
        else
          vv[u] = to_acc<TA>(v_s[qq]);
      }
#pragma unroll
      for (int u = 0; u < U; u++) xx[u] = to_acc<TA>(buf_load<TX>(rx, cc[u] * sizeof(TX)));
#pragma unroll
      for (int u = 0; u < U; u++)
        if (q + u < e) acc += vv[u] * xx[u];
    }
    TA out = alpha * acc;
    if (beta != TA(0)) out += beta * to_acc<TA>(y[r0 + i]);
    y[r0 + i] = from_acc<TY>(out);
  }
}

// G lanes per row (rows of more than a few entries): lane j of a row's group
// takes entries j, j + G, ..., U of them per round with their gathers in
// flight, and the group's sums are added with DPP / shuffles.
template <int WG, int G, int U, typename TX, typename TY, typename TA, typename TI,
          typename VS, typename CS, bool PERM = false, typename TV = TA>
__device__ __forceinline__ void row_group_sums(const VS *v_s, const CS *c_s, const int *rp_s,
                                               bool rows_lds, const FwdArgs &a, int64_t r0,
                                               int64_t e0, int64_t nr, TY *y, TA alpha,
                                               TA beta) {
  const TI *rowptr = static_cast<const TI *>(a.rowptr);
  const rsrc_t rx = make_rsrc(a.x, a.xbytes);
  [[maybe_unused]] const rsrc_t rv =
      make_rsrc(a.vals, PERM ? static_cast<uint32_t>(a.nnz * sizeof(TV)) : 0u);
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
    TA acc = 0;
    for (int q = s + lane; q < e; q += G * U) {
      uint32_t cc[U];
      TA vv[U], xx[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int qq = q + u * G < e ? q + u * G : q;
        cc[u] = col32(c_s[qq]);
        if constexpr (PERM)
          vv[u] = to_acc<TA>(buf_load<TV>(rv, col32(v_s[qq]) * static_cast<uint32_t>(sizeof(TV))));
        else
          vv[u] = to_acc<TA>(v_s[qq]);
      }
#pragma unroll
      for (int u = 0; u < U; u++) xx[u] = to_acc<TA>(buf_load<TX>(rx, cc[u] * sizeof(TX)));
#pragma unroll
      for (int u = 0; u < U; u++)
        if (q + u * G < e) acc += vv[u] * xx[u];
    }
    acc = group_sum<G>(acc);
    if (lane == 0) {
      TA out = alpha * acc;
      if (beta != TA(0)) out += beta * to_acc<TA>(y[r0 + i]);
      y[r0 + i] = from_acc<TY>(out);
    }
  }
}

// Row-ordered sums of a block by its mean row length: one lane per row up
// to 12 entries, else groups of 8 or 16 lanes. (One lane per row with eight
// entries per round measured faster than two rows of four per lane and one
// row of four: 3-D Laplacian 256^3 BF16 174 vs 184 / 188 us, FP32 212 vs
// 219 / 221 us.)
template <int WG, typename TX, typename TY, typename TA, typename TI, typename VS, typename CS,
          bool PERM = false, typename TV = TA>
__device__ __forceinline__ void row_sums_any(int n_s, const VS *v_s, const CS *c_s,
                                             const int *rp_s, bool rows_lds, const FwdArgs &a,
                                             int64_t r0, int64_t e0, int64_t nr, TY *y, TA alpha,
                                             TA beta) {
  const int64_t avg = nr > 0 ? n_s / nr : 0;
  // one lane per row: rounds of the share's longest row when that is 5 or
  // 7 (2-D and 3-D stencils), so that no gather slot of a round is idle;
  // the sums are the same either way (entry order)
#define DRX_RG(U)                                                                           \
  row_gather_sums<WG, U, TX, TY, TA, TI, VS, CS, PERM, TV>(v_s, c_s, rp_s, rows_lds, a, r0, e0, \
                                                           nr, y, alpha, beta)
#define DRX_RGG(G)                                                                          \
  row_group_sums<WG, G, 5, TX, TY, TA, TI, VS, CS, PERM, TV>(v_s, c_s, rp_s, rows_lds, a, r0,  \
                                                             e0, nr, y, alpha, beta)
  if (avg <= 12 && a.lmax > 0 && a.lmax <= 5)
    DRX_RG(5);
  else if (avg <= 12 && a.lmax > 5 && a.lmax <= 7)
    DRX_RG(7);
  else if (avg <= 12)
    DRX_RG(8);
  else if (avg <= 40)
    DRX_RGG(8);
  else
    DRX_RGG(16);
#undef DRX_RGG
#undef DRX_RG
}

template <int MODE, typename TV, typename TI>
__device__ __forceinline__ TV long_val(const TV *vals, const TI *perm,
                                       int64_t q) {
  if constexpr (MODE == kPerm)
    return vals[static_cast<int64_t>(ld_stream(perm + q))];
  else
    return ld_stream(vals + q);
}

// LDS of one workgroup of the forward kernels (row-map entries have the
// index type; RG: the window's column indices for row-ordered gathers).
template <int WG, int CAP, bool ROWMAP, typename TA, typename TI, bool RG = false>
struct FwdSmem {
  // rows of a block held in LDS (blocks with more read rowptr); row-ordered
  // gathers serve shares of at least four entries per row on average
  static constexpr int ROWCAP = RG ? CAP / 4 : CAP / 2;
  static constexpr int MAPCAP = ROWMAP ? ROWCAP : 1;
  TA prod[CAP];
  uint32_t cs[RG ? CAP : 0];  // empty unless RG (the other kernels keep their LDS)
  TA ypre[MAPCAP];        // kRowMap, k = 1: the current y of the block's rows
  TA red[WG / 64];
  TI map_s[MAPCAP];       // kRowMap: output rows of the block
  int rp_s[ROWCAP + 1];
  uint32_t xmm[2 * (WG / 64)];  // per-wave column min / max (x staging)
};

// Storage value from its raw bits (2- or 4-byte types).
template <typename TV, typename U> __device__ __forceinline__ TV bits_to(U u) {
  if constexpr (sizeof(TV) == 2)
    return __builtin_bit_cast(TV, static_cast<uint16_t>(u));
  else
    return __builtin_bit_cast(TV, static_cast<uint32_t>(u));
}

// One row block `blk` of the share described by `a` (all lanes of the
// workgroup call it together).
//
// VW = 2 or 4 (32-bit indices, plain or row-map mode, VW values in at most
// 16 bytes): each lane streams groups of VW consecutive entries, one 4 VW-
// byte colind load and one VW * sizeof(TV)-byte vals load per group, from the
// group-aligned entry ea below the block's first entry e0 (the caller
// guarantees vals / colind pointers aligned to a group). Slot s of the LDS
// product buffer holds entry ea + s; the block's entries are slots
// [e0 - ea, e0 - ea + n_s), inside the buffer because a window leaves
// kWinSlack entries free. The last group may extend up to VW - 1 entries
// past the block (and past the arrays' last entry, inside that same aligned
// group); such entries are read but never used.
template <int WG, int CAP, int MODE, bool BUFX, typename TV, typename TX,
          typename TY, typename TA, typename TI, int VW = 1, bool XS = false,
          bool RG = false>
__device__ __forceinline__ void process_block(const FwdArgs &a, int64_t blk,
                                              FwdSmem<WG, CAP, MODE == kRowMap, TA, TI, RG> &sm) {
  static_assert(VW == 1 || ((VW == 2 || VW == 4) && MODE != kPerm && sizeof(TI) == 4 &&
                            VW * sizeof(TV) <= 16 && VW <= kWinSlack),
                "groups of entries: plain or row-map mode, 32-bit indices, at most 16 bytes "
                "of values per group");
  static_assert(!XS || (MODE == kPlain && BUFX), "x staging: plain mode, buffer gathers");
  static_assert(!RG || (MODE == kPlain && BUFX && !XS), "row gathers: plain mode, buffer gathers");
  constexpr int T = CAP / WG;
  constexpr int ROWCAP = FwdSmem<WG, CAP, MODE == kRowMap, TA, TI, RG>::ROWCAP;
  TA *prod = sm.prod;
  int *rp_s = sm.rp_s;
  TI *map_s = sm.map_s;
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
    // LDS slot of entry u of a lane (the entry is ea + slot, ea = e0 - sh);
    // the block's entries are slots [sh, sh + n_s)
    const int sh = VW > 1 ? static_cast<int>(b.e0 & (VW - 1)) : 0;
    auto slot = [&](int u) -> int {
      if constexpr (VW == 1)
        return tid + u * WG;
      else
        return (tid + (u / VW) * WG) * VW + (u % VW);
    };

    constexpr bool kRowsFirst = MODE == kRowMap && sizeof(TV) == 4;
    if (kRowsFirst && rows_lds) {
      // Column panels of 4-byte values: row offsets, output rows and their
      // current y, before
      // the window is streamed so that their registers are free again. Every
      // load of a lane is issued before the first LDS write, with indices
      // clamped to valid rows: a per-row predicate around load and use made
      // each wave wait for one load at a time, three dependent round trips
      // per row slot, and panel blocks hold up to ROWCAP short rows.
      // nr < ROWCAP, so NI slots per lane cover rows 0..nr. Headline Zipf
      // FP32 panels: 10.31 -> 9.91 ms per step. (Plain blocks hold few rows
      // per lane and keep the simple loop below: batching spilled the
      // 4096-entry kernel. 16-bit panels keep it too: with it, their pair
      // unpacking waited on each streamed load, 9.2 -> 10.6 ms.)
      constexpr int NI = (ROWCAP + WG - 1) / WG;
      const TI *rowmap = static_cast<const TI *>(a.rowmap);
      TI raw[NI], row[NI];
      TA yv[NI];
      // Slots k with k * WG > nr hold no row of the block (a block-uniform
      // test): they load nothing.
#pragma unroll
      for (int k = 0; k < NI; k++) {
        if (k * WG > nr) break;
        const int64_t i = tid + k * WG;
        const int64_t r = b.r0 + (i <= nr ? i : nr);
        // row_begin without its branch: rowptr[nrows] exists only with a
        // sentinel; otherwise that row begins at nnz.
        raw[k] = rowptr[r < a.nrows || a.sent ? r : r - 1];
        row[k] = nr > 0 ? rowmap[b.r0 + (i < nr ? i : nr - 1)] : TI(0);
      }
      if (a.k == 1 && nr > 0) {
#pragma unroll
        for (int k = 0; k < NI; k++) {
          if (k * WG > nr) break;
          yv[k] = map_ypre<TA>(y, row[k], beta);
        }
      }
#pragma unroll
      for (int k = 0; k < NI; k++) {
        if (k * WG > nr) break;
        const int i = tid + k * WG;
        if (i <= nr) {
          const int64_t r = b.r0 + i;
          int64_t v0 = a.nnz;
          if (r < a.nrows || a.sent) {
            v0 = static_cast<int64_t>(raw[k]) - a.base;
            v0 = v0 < 0 ? 0 : (v0 > a.nnz ? a.nnz : v0);
          }
          rp_s[i] = static_cast<int>(v0 - b.e0);
        }
        if (i < nr) {
          map_s[i] = row[k];
          if (a.k == 1) ypre[i] = yv[k];
        }
      }
    }

    // Stream the window (reads past n_s return 0 and touch no memory).
    TA v[T];
    uint32_t c[T];
    if constexpr (VW > 1) {
      // the lanes' groups cover entries [ea, ea + CAP) and the block's
      // entries [e0, e0 + n_s) with sh + n_s <= CAP
      const int64_t ea = b.e0 - sh;
      const uint32_t ngrp = static_cast<uint32_t>(n_s + sh + VW - 1) / VW;
      constexpr uint32_t VB = VW * sizeof(TV);  // bytes of a group's values
      const rsrc_t rv = make_rsrc(vals + ea, ngrp * VB);
      const rsrc_t rc = make_rsrc(colind + ea, ngrp * VW * 4);
#pragma unroll
      for (int g = 0; g < T / VW; g++) {
        const uint32_t gi = tid + g * WG;
        const Words<VW> cw = buf_load<Words<VW>, kNT>(rc, gi * (VW * 4));
        const Words<(VB + 3) / 4> vw = buf_load<Words<(VB + 3) / 4>, kNT>(rv, gi * VB);
        TV t[VW];
        __builtin_memcpy(t, vw.w, VB);
#pragma unroll
        for (int i = 0; i < VW; i++) {
          c[g * VW + i] = cw.w[i];
          v[g * VW + i] = to_acc<TA>(t[i]);
        }
      }
    } else {
      const rsrc_t rc = make_rsrc(colind + b.e0, n_s * sizeof(TI));
      if constexpr (MODE == kPerm) {
        // Slots past the window read perm = 0, so every lane loads vals[0]
        // or its own entry unconditionally (a predicated load made each
        // wave wait for one value at a time); n_s > 0 guarantees vals[0].
        const rsrc_t rp = make_rsrc(perm + b.e0, n_s * sizeof(TI));
        TI pe[T];
#pragma unroll
        for (int u = 0; u < T; u++) {
          const uint32_t idx = tid + u * WG;
          c[u] = col32(buf_load<TI, kNT>(rc, idx * sizeof(TI)));
          pe[u] = buf_load<TI, kNT>(rp, idx * sizeof(TI));
        }
        if (n_s > 0) {
#pragma unroll
          for (int u = 0; u < T; u++) v[u] = to_acc<TA>(vals[static_cast<int64_t>(pe[u])]);
#pragma unroll
          for (int u = 0; u < T; u++)
            v[u] = tid + u * WG < static_cast<uint32_t>(n_s) ? v[u] : TA(0);
        } else {
#pragma unroll
          for (int u = 0; u < T; u++) v[u] = 0;
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
    if (!kRowsFirst && rows_lds) {
      for (int i = tid; i <= nr; i += WG) {
        rp_s[i] = static_cast<int>(
            row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, b.r0 + i) - b.e0);
        if constexpr (MODE == kRowMap) {
          if (i < nr) {
            const TI row = static_cast<const TI *>(a.rowmap)[b.r0 + i];
            map_s[i] = row;
            if (a.k == 1) ypre[i] = map_ypre<TA>(y, row, beta);
          }
        }
      }
    }

    // x staged in LDS (plain CSR, one right-hand side, columns local to the
    // block): the window's columns span [lo, hi]; x[lo, lo + xn) is loaded
    // once, coalesced, into the product buffer (xn <= its capacity), and each
    // entry reads its x from there -- or from memory past xn. The products
    // equal the gathered ones bit for bit.
    bool staged = false;
    if constexpr (RG) {
      // Row-ordered gathers through registers (k = 1; the host picks these
      // kernels for shares of rows of at most 128 entries whose vals or
      // colind are not aligned for k_spmv_rows' straight LDS copies): the
      // window's values and columns go to LDS instead of products.
#pragma unroll
      for (int u = 0; u < T; u++) {
        prod[slot(u)] = v[u];
        sm.cs[slot(u)] = c[u];
      }
      __syncthreads();
      row_sums_any<WG, TX, TY, TA, TI>(n_s, prod + sh, sm.cs + sh, rp_s, rows_lds, a, b.r0, b.e0,
                                       nr, y, alpha, beta);
      __syncthreads();
    }
    if constexpr (XS) {
      if (a.k == 1) {
        constexpr uint32_t XCAP = CAP * sizeof(TA) / sizeof(TX);
        auto in_block = [&](int u) {
          const int sl = slot(u);
          return sl >= sh && sl < sh + n_s;
        };
        uint32_t lo, hi;
        block_col_range<WG, T>(c, in_block, sm.xmm, lo, hi);
        if (lo <= hi) {
          const uint32_t xn = hi - lo + 1 < XCAP ? hi - lo + 1 : XCAP;
          TX *xs = reinterpret_cast<TX *>(prod);
          const rsrc_t rx = make_rsrc(x, a.xbytes);
          for (uint32_t i = tid; i < xn; i += WG) xs[i] = buf_load<TX>(rx, (lo + i) * sizeof(TX));
          __syncthreads();
          // All LDS reads first, then the rare columns past the window from
          // memory, then the products (slots outside the block hold
          // products the reduction never reads).
          TA xv[T];
          bool far = false;
#pragma unroll
          for (int u = 0; u < T; u++) {
            const uint32_t off = c[u] - lo;
            xv[u] = to_acc<TA>(xs[off < xn ? off : 0]);
            far |= off >= xn && in_block(u);
          }
          if (far) {
#pragma unroll
            for (int u = 0; u < T; u++)
              if (c[u] - lo >= xn && in_block(u))
                xv[u] = to_acc<TA>(buf_load<TX>(rx, c[u] * sizeof(TX)));
          }
          __syncthreads();
#pragma unroll
          for (int u = 0; u < T; u++) prod[slot(u)] = v[u] * xv[u];
          staged = true;
        }
      }
    }

    for (int64_t j = 0; j < (RG ? 0 : a.k); j++) {
      const TX *xj = x + j * a.ldx;
      if (staged) {
      } else if constexpr (BUFX) {
        const rsrc_t rx = make_rsrc(xj, a.xbytes);
        if constexpr (sizeof(TA) == 8 && T > 4) {
          // 8-byte sums over 8 or more entries per lane: the gathers in two
          // halves of T / 2, each half all in flight before its products
          // (every slot: past the window it gathers x[0] into a product the
          // reduction never reads). All T at once costs T more register
          // pairs; a predicated gather per slot made each wave wait for one
          // gather at a time. Measured at 6 waves, against one at a time at
          // 7: 3-D Laplacian 256^3 FP64 413 -> 392 us, band +-16 FP64 389 ->
          // 355 us, 2-D Laplacian 1000^2 23.5 -> 21.4 us per call.
          constexpr int H = T / 2;
#pragma unroll
          for (int h = 0; h < 2; h++) {
            TA xh[H];
#pragma unroll
            for (int i = 0; i < H; i++)
              xh[i] = to_acc<TA>(buf_load<TX>(rx, c[h * H + i] * sizeof(TX)));
#pragma unroll
            for (int i = 0; i < H; i++) prod[slot(h * H + i)] = v[h * H + i] * xh[i];
          }
        } else {
          // Every lane issues all T gathers before storing a product: a
          // per-entry predicate around load and store made the compiler
          // wait for each gather in turn (band +-16 FP32: 338 -> 273 us).
          // Slots outside the block (past the window: their column reads 0,
          // so they gather x[0]; before e0 in a group: an entry of the
          // previous block) go to product slots that the reduction never
          // reads.
          TA xg[T];
#pragma unroll
          for (int u = 0; u < T; u++) xg[u] = to_acc<TA>(buf_load<TX>(rx, c[u] * sizeof(TX)));
#pragma unroll
          for (int u = 0; u < T; u++) prod[slot(u)] = v[u] * xg[u];
        }
      } else {
#pragma unroll
        for (int u = 0; u < T; u++) {
          const int s = slot(u);
          if (s >= sh && s < sh + n_s) prod[s] = v[u] * to_acc<TA>(xj[c[u]]);
        }
      }
      __syncthreads();
      reduce_rows_any<WG, MODE, TA, TY, TI>(n_s, prod + sh, rp_s, map_s,
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
            const TI m = static_cast<const TI *>(a.rowmap)[last];
            TY *yj = y + j * a.ldy;
            yj[map_row(m)] = from_acc<TY>(map_ypre<TA>(yj, m, beta) + alpha * s);
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

// Waves per SIMD the forward kernel is compiled for: as many workgroups as
// the 160 KB of LDS per CU holds (one wave per SIMD per 256 lanes), at most
// 8. Measured on MI355X: 7-8 waves raise the structured forms by 7-17 %
// over the compiler's default (4-5 waves); trading LDS for an eighth FP64
// workgroup spilled registers and lost 30 %.
template <int WG, int CAP, int MODE, typename TA, typename TI, bool RG = false>
constexpr int fwd_waves() {
  constexpr int lds = static_cast<int>(sizeof(FwdSmem<WG, CAP, MODE == kRowMap, TA, TI, RG>));
  constexpr int wgs = (160 * 1024) / lds;
  constexpr int w = wgs * (WG / 256);
  return w > 8 ? 8 : (w < 1 ? 1 : w);
}

template <int WG, int CAP, int MODE, typename TV, typename TA, typename TI, bool RG = false>
constexpr int fwd_kernel_waves() {
  constexpr int w = fwd_waves<WG, CAP, MODE, TA, TI, RG>() -
                    ((sizeof(TV) == 2 && MODE == kPlain) || MODE == kPerm ? 1 : 0);
  return sizeof(TA) == 8 && CAP / WG >= 8 && w > 6 ? 6 : w;
}

template <int WG, int CAP, int MODE, bool BUFX, typename TV, typename TX,
          typename TY, typename TA, typename TI, int VW = 1, bool XS = false,
          bool RG = false>
__global__ __launch_bounds__(WG)
// (16-bit values in plain mode take one wave fewer: at the LDS-bound count
// their pair unpacking spilled registers; 3-D Laplacian BF16 390 -> 371 us.
// The panel kernel's LDS already holds it at 7 waves, where it does not.
// Permuted values take one wave fewer too: their T unconditional value
// loads spilled at 8; band +-16 FP32 adjoint 563 -> 471 us.)
// (FP64 sums with 8 or more entries per lane run at most 6 waves: their
// gathers in halves of four need the registers.)
__attribute__((amdgpu_waves_per_eu(fwd_kernel_waves<WG, CAP, MODE, TV, TA, TI, RG>(), 8)))
void k_spmv_fwd(FwdArgs a) {
  __shared__ FwdSmem<WG, CAP, MODE == kRowMap, TA, TI, RG> sm;
  for (int64_t wi = blockIdx.x; wi < a.per_xcd * kXcds; wi += a.grid) {
    const int64_t blk = map_block(wi, a.per_xcd);
    if (blk >= a.nblocks) continue;
    process_block<WG, CAP, MODE, BUFX, TV, TX, TY, TA, TI, VW, XS, RG>(a, blk, sm);
  }
}

// ---------------------------------------------------------------------------
// Row-ordered gathers with each block's window copied straight into LDS
// (plain CSR, one right-hand side): buffer loads with the LDS flag write a
// wave's 64 lanes x D bytes to a contiguous 64 D-byte span of LDS, so the
// window's values and column indices land in LDS exactly as they lie in
// memory, with no register round trip and no LDS write instructions. The
// rows are then summed as in row_gather_sums, reading the stored types.
// ---------------------------------------------------------------------------
template <int WG, int CAP, typename TV, typename TI, typename TA> struct RowsSmem {
  static constexpr int ROWCAP = CAP / 4;
  alignas(16) TV v[CAP];
  alignas(16) TI c[CAP];
  int rp_s[ROWCAP + 1];
  TA red[WG / 64];
};

// Bytes per lane of the LDS copy of an array of E: 16, or 4 for 2-byte
// elements (a 16-byte unit aligned to 8 bytes could cross a page past the
// array's end).
template <typename E> constexpr int lds_copy_width() { return sizeof(E) >= 4 ? 16 : 4; }

// Copies the block's elements [e0, e0 + n) of `src` into `dst` from the
// copy unit at or below e0 (returns the shift e0 - that unit's first
// element, below kWinSlack); whole units, so the last one may extend past
// the block and past the array's last element inside that aligned unit
// (the caller guarantees `src` aligned to a unit). Rounds past the
// block's units are skipped; lanes past them read nothing (range check).
template <int WG, int CAP, typename E>
__device__ __forceinline__ int lds_window(const E *src, int64_t e0, int n, E *dst, int tid) {
  constexpr int D = lds_copy_width<E>();
  constexpr int A = D / static_cast<int>(sizeof(E));
  constexpr int RB = WG * D;
  constexpr int R = CAP * static_cast<int>(sizeof(E)) / RB;
  static_assert(R * RB == CAP * static_cast<int>(sizeof(E)), "whole rounds per window");
  static_assert(A <= kWinSlack, "the shifted window fits CAP entries");
  const int sh = static_cast<int>(e0 & (A - 1));
  const uint32_t need =
      n > 0 ? static_cast<uint32_t>((sh + n) * sizeof(E) + D - 1) / D * D : 0u;
  const rsrc_t r = make_rsrc(src + (e0 - sh), need);
  char *base = reinterpret_cast<char *>(dst) + (tid >> 6) * 64 * D;
#pragma unroll
  for (int g = 0; g < R; g++) {
    if (static_cast<uint32_t>(g * RB) < need) {
      auto *l = (__attribute__((address_space(3))) void *)(base + g * RB);
      const uint32_t off = static_cast<uint32_t>((g * WG + tid) * D);
      if constexpr (D == 16)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(r, l, 16, off, 0, 0, kNT);
      else
        __builtin_amdgcn_raw_ptr_buffer_load_lds(r, l, 4, off, 0, 0, kNT);
    }
  }
  return sh;
}

// Row offsets 0..nr of a block into registers: every lane's loads issued
// together (indices clamped to valid rows; row_begin without its branch).
template <int WG, int NI, typename TI, typename Args>
__device__ __forceinline__ void rows_offsets_load(const Args &a, int64_t r0, int64_t nr,
                                                  TI (&raw)[NI], int tid) {
  const TI *rowptr = static_cast<const TI *>(a.rowptr);
#pragma unroll
  for (int k = 0; k < NI; k++) {
    if (k * WG > nr) break;
    const int64_t i = tid + k * WG;
    const int64_t r = r0 + (i <= nr ? i : nr);
    raw[k] = rowptr[r < a.nrows || a.sent ? r : r - 1];
  }
}

template <int WG, int NI, typename TI, typename Args>
__device__ __forceinline__ void rows_offsets_store(const Args &a, int64_t r0, int64_t e0,
                                                   int64_t nr, const TI (&raw)[NI], int *rp_s,
                                                   int tid) {
#pragma unroll
  for (int k = 0; k < NI; k++) {
    if (k * WG > nr) break;
    const int64_t i = tid + k * WG;
    if (i <= nr) {
      const int64_t r = r0 + i;
      int64_t v0 = a.nnz;
      if (r < a.nrows || a.sent) {
        v0 = static_cast<int64_t>(raw[k]) - a.base;
        v0 = v0 < 0 ? 0 : (v0 > a.nnz ? a.nnz : v0);
      }
      rp_s[i] = static_cast<int>(v0 - e0);
    }
  }
}

// A last row [es, e1) longer than the window (row r1 - 1): the whole
// workgroup, strided coalesced loads, in the sum order of process_block's
// long-row path (the same bits).
template <int WG, typename TV, typename TX, typename TY, typename TA, typename TI,
          int MODE = kPlain>
__device__ __forceinline__ void rows_long_last(const FwdArgs &a, int64_t es, int64_t e1,
                                               int64_t r1, TA *red, int tid) {
  const TI *colind = static_cast<const TI *>(a.colind);
  const TV *vals = static_cast<const TV *>(a.vals);
  const TI *perm = static_cast<const TI *>(a.perm);
  const TX *x = static_cast<const TX *>(a.x);
  TY *y = static_cast<TY *>(a.y);
  const TA alpha = static_cast<TA>(a.alpha), beta = static_cast<TA>(a.beta);
  const int64_t last = r1 - 1;
  TA acc0 = 0, acc1 = 0;
  int64_t q = es + tid;
  for (; q + 3 * WG < e1; q += 4 * WG) {
    const TV w0 = long_val<MODE>(vals, perm, q), w1 = long_val<MODE>(vals, perm, q + WG);
    const TV w2 = long_val<MODE>(vals, perm, q + 2 * WG);
    const TV w3 = long_val<MODE>(vals, perm, q + 3 * WG);
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
    acc0 += to_acc<TA>(long_val<MODE>(vals, perm, q)) *
            to_acc<TA>(x[static_cast<int64_t>(ld_stream(colind + q))]);
  const TA s = block_sum<WG>(acc0 + acc1, red);
  if (tid == 0) {
    TA out = alpha * s;
    if (beta != TA(0)) out += beta * to_acc<TA>(y[last]);
    y[last] = from_acc<TY>(out);
  }
}

// MODE kPerm (the adjoint through a permutation): the window's positions in
// vals (perm) are copied into LDS in place of the values, and each entry
// gathers its value: a wave's value gathers then read the u-th entries of
// 64 consecutive rows (of the transposed structure), which for stencils and
// bands sit next to each other in vals.
template <int MODE, typename TV, typename TI>
using RowsVal = std::conditional_t<MODE == kPerm, TI, TV>;

// A block's bounds: its record and the next one's first row and entry.
struct RowsRec {
  BlockRec b;
  int64_t r1, e1;
};
__device__ __forceinline__ RowsRec rows_rec(const FwdArgs &a, int64_t blk) {
  return RowsRec{a.rec[blk], a.rec[blk + 1].r0, a.rec[blk + 1].e0};
}

template <int WG, int CAP, typename TV, typename TX, typename TY, typename TA, typename TI,
          int MODE = kPlain>
__device__ __forceinline__ void process_block_rows(
    const FwdArgs &a, const RowsRec &rr, RowsSmem<WG, CAP, RowsVal<MODE, TV, TI>, TI, TA> &sm,
    int tid) {
  using VS = RowsVal<MODE, TV, TI>;
  constexpr int ROWCAP = RowsSmem<WG, CAP, VS, TI, TA>::ROWCAP;
  constexpr int NI = (ROWCAP + WG) / WG;
  const TI *colind = static_cast<const TI *>(a.colind);
  const VS *vsrc = static_cast<const VS *>(MODE == kPerm ? a.perm : a.vals);
  TY *y = static_cast<TY *>(a.y);
  const TA alpha = static_cast<TA>(a.alpha), beta = static_cast<TA>(a.beta);
  const BlockRec b = rr.b;
  const int64_t r1 = rr.r1, e1 = rr.e1;
  if (b.r0 >= r1) return;
  const int n_s = static_cast<int>(b.es - b.e0);
  const int64_t nr = (b.es == e1 ? r1 : r1 - 1) - b.r0;
  const bool rows_lds = nr < ROWCAP;
  const int shv = lds_window<WG, CAP>(vsrc, b.e0, n_s, sm.v, tid);
  const int shc = lds_window<WG, CAP>(colind, b.e0, n_s, sm.c, tid);
  if (rows_lds) {
    TI raw[NI];
    rows_offsets_load<WG, NI>(a, b.r0, nr, raw, tid);
    rows_offsets_store<WG, NI>(a, b.r0, b.e0, nr, raw, sm.rp_s, tid);
  }
  // the LDS copies count as vector-memory loads
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  row_sums_any<WG, TX, TY, TA, TI, VS, TI, MODE == kPerm, TV>(
      n_s, sm.v + shv, sm.c + shc, sm.rp_s, rows_lds, a, b.r0, b.e0, nr, y, alpha, beta);
  if (b.es < e1) rows_long_last<WG, TV, TX, TY, TA, TI, MODE>(a, b.es, e1, r1, sm.red, tid);
  __syncthreads();  // the next block overwrites the LDS window
}

// As many workgroups as the LDS holds, at most 8 waves (7 for 8-byte sums,
// which spill at 8).
template <int WG, int CAP, typename TV, typename TI, typename TA> constexpr int rows_waves() {
  constexpr int lds = static_cast<int>(sizeof(RowsSmem<WG, CAP, TV, TI, TA>));
  constexpr int w = (160 * 1024) / lds * (WG / 256);
  constexpr int most = sizeof(TA) == 8 ? 7 : 8;
  return w > most ? most : (w < 1 ? 1 : w);
}

template <int WG, int CAP, typename TV, typename TX, typename TY, typename TA, typename TI,
          int MODE = kPlain>
__global__ __launch_bounds__(WG)
__attribute__((amdgpu_waves_per_eu(rows_waves<WG, CAP, RowsVal<MODE, TV, TI>, TI, TA>(), 8)))
void k_spmv_rows(FwdArgs a) {
  __shared__ RowsSmem<WG, CAP, RowsVal<MODE, TV, TI>, TI, TA> sm;
  int tid;
  asm volatile("" : "=v"(tid) : "0"(static_cast<int>(threadIdx.x)));
  // Workgroups that walk several blocks (capped grids) load the next
  // block's record before the current block's window copies, so that its
  // round trip overlaps theirs.
  const int64_t total = a.per_xcd * kXcds;
  auto next_valid = [&](int64_t w) {
    for (; w < total; w += a.grid)
      if (map_block(w, a.per_xcd) < a.nblocks) break;
    return w;
  };
  int64_t wi = next_valid(blockIdx.x);
  if (wi >= total) return;
  RowsRec cur = rows_rec(a, map_block(wi, a.per_xcd));
  for (;;) {
    const int64_t wn = next_valid(wi + a.grid);
    RowsRec nxt = cur;
    if (wn < total) nxt = rows_rec(a, map_block(wn, a.per_xcd));
    process_block_rows<WG, CAP, TV, TX, TY, TA, TI, MODE>(a, cur, sm, tid);
    if (wn >= total) break;
    wi = wn;
    cur = nxt;
  }
}

}  // namespace drx

// ---------------------------------------------------------------------------
// Several right-hand sides (k >= 2).
//
// x is first copied into row-major chunks of KP columns (KP = 2, 4, 8 or 16
// by k; at most 8 for 8-byte x, whose 16 running sums would cost
// occupancy),
//   xr[(chunk * ncols + c) * KP + jj] = x[(chunk * KP + jj) * ldx + c]
// (zero padded past k), so that an entry gathers one contiguous KP-vector
// (16-128 bytes) instead of k values from k separate columns. The window's
// values and column indices are staged in LDS; groups of G lanes then walk
// their rows, each lane holding KP running sums, and the G lanes of a row
// write its KP outputs (column jj by lane jj mod G).
// ---------------------------------------------------------------------------
namespace drx {

template <int WG, int CAP, typename TA> struct MrhsSmem {
  TA v[CAP];
  uint32_t c[CAP];
  int rp_s[CAP / 2 + 1];
  TA red[16 * (WG / 64)];
};

// KP consecutive x values (one xr row) converted to the accumulator type.
template <int KP, typename TX, typename TA>
__device__ __forceinline__ void load_xrow(const TX *p, TA (&xv)[KP]) {
  constexpr int BYTES = KP * static_cast<int>(sizeof(TX));
  static_assert(BYTES % 4 == 0, "an xr row is a multiple of 4 bytes");
  if constexpr (BYTES >= 16) {
    constexpr int PER = 16 / static_cast<int>(sizeof(TX));
    const uint4 *q = reinterpret_cast<const uint4 *>(p);
#pragma unroll
    for (int i = 0; i < BYTES / 16; i++) {
      const uint4 w = q[i];
      TX t[PER];
      __builtin_memcpy(t, &w, 16);
#pragma unroll
      for (int s = 0; s < PER; s++) xv[i * PER + s] = to_acc<TA>(t[s]);
    }
  } else if constexpr (BYTES == 8) {
    const uint2 w = *reinterpret_cast<const uint2 *>(p);
    TX t[KP];
    __builtin_memcpy(t, &w, 8);
#pragma unroll
    for (int s = 0; s < KP; s++) xv[s] = to_acc<TA>(t[s]);
  } else {
    const uint32_t w = *reinterpret_cast<const uint32_t *>(p);
    TX t[KP];
    __builtin_memcpy(t, &w, 4);
#pragma unroll
    for (int s = 0; s < KP; s++) xv[s] = to_acc<TA>(t[s]);
  }
}

template <int WG, int G, int KP, typename TX, typename TY, typename TA,
          typename TI>
__device__ __forceinline__ void reduce_rows_mrhs(
    const TA *v_s, const uint32_t *c_s, const int *rp_s, bool rows_lds,
    const FwdArgs &a, int64_t r0, int64_t e0, int64_t nr, const TX *xrc,
    int64_t j0, int kk, TA alpha, TA beta, int tid) {
  const TI *rowptr = static_cast<const TI *>(a.rowptr);
  TY *y = static_cast<TY *>(a.y);
  const int lane = tid & (G - 1);
  for (int64_t i = tid / G; i < nr; i += WG / G) {
    int s, e;
    if (rows_lds) {
      s = rp_s[i];
      e = rp_s[i + 1];
    } else {
      s = static_cast<int>(row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r0 + i) - e0);
      e = static_cast<int>(row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r0 + i + 1) - e0);
    }
    TA acc[KP];
#pragma unroll
    for (int jj = 0; jj < KP; jj++) acc[jj] = 0;
    int q = s + lane;
    for (; q + G < e; q += 2 * G) {
      TA x0[KP], x1[KP];
      load_xrow<KP>(xrc + static_cast<size_t>(c_s[q]) * KP, x0);
      load_xrow<KP>(xrc + static_cast<size_t>(c_s[q + G]) * KP, x1);
      const TA v0 = v_s[q], v1 = v_s[q + G];
#pragma unroll
      for (int jj = 0; jj < KP; jj++) {
        acc[jj] += v0 * x0[jj];
        acc[jj] += v1 * x1[jj];
      }
    }
    if (q < e) {
      TA x0[KP];
      load_xrow<KP>(xrc + static_cast<size_t>(c_s[q]) * KP, x0);
      const TA v0 = v_s[q];
#pragma unroll
      for (int jj = 0; jj < KP; jj++) acc[jj] += v0 * x0[jj];
    }
#pragma unroll
    for (int jj = 0; jj < KP; jj++) acc[jj] = group_sum<G>(acc[jj]);
    const int64_t row = r0 + i;
#pragma unroll
    for (int jj = 0; jj < KP; jj++) {
      if (jj < kk && (jj & (G - 1)) == lane) {
        TY *yp = y + (j0 + jj) * a.ldy + row;
        TA out = alpha * acc[jj];
        if (beta != TA(0)) out += beta * to_acc<TA>(*yp);
        *yp = from_acc<TY>(out);
      }
    }
  }
}

// Wide chunks (an x row of KP values spans S = KP / KL 16-byte pieces, KL =
// 16 / sizeof(x)): S lanes share an entry, each loading and summing one
// 16-byte piece, so a lane holds KL sums instead of KP and keeps four
// entries' pieces in flight. Groups of G * S lanes take a row (G entry
// lanes), and the pieces' sums are added across the G entry lanes with
// lane shuffles (stride S). Holding the whole KP-vector per lane (16 sums
// and 16 values per entry in flight for FP32 k = 16) limited the batched
// kernel to two waves per SIMD with two entries in flight per lane.
template <int WG, int G, int KP, typename TX, typename TY, typename TA,
          typename TI>
__device__ __forceinline__ void reduce_rows_mrhs_split(
    const TA *v_s, const uint32_t *c_s, const int *rp_s, bool rows_lds,
    const FwdArgs &a, int64_t r0, int64_t e0, int64_t nr, const TX *xrc,
    int64_t j0, int kk, TA alpha, TA beta, int tid) {
  constexpr int KL = 16 / static_cast<int>(sizeof(TX));
  constexpr int S = KP / KL;
  constexpr int GS = G * S;
  static_assert(S >= 2 && GS <= 64 && KP % KL == 0, "pieces of 16 bytes within a wave");
  const TI *rowptr = static_cast<const TI *>(a.rowptr);
  TY *y = static_cast<TY *>(a.y);
  const int lg = tid & (GS - 1);
  const int sub = lg % S, el = lg / S;
  for (int64_t i = tid / GS; i < nr; i += WG / GS) {
    int s, e;
    if (rows_lds) {
      s = rp_s[i];
      e = rp_s[i + 1];
    } else {
      s = static_cast<int>(row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r0 + i) - e0);
      e = static_cast<int>(row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r0 + i + 1) - e0);
    }
    TA acc[KL];
#pragma unroll
    for (int j = 0; j < KL; j++) acc[j] = 0;
    const TX *xs = xrc + sub * KL;
    int q = s + el;
    for (; q + 3 * G < e; q += 4 * G) {
      TA x0[KL], x1[KL], x2[KL], x3[KL];
      load_xrow<KL>(xs + static_cast<size_t>(c_s[q]) * KP, x0);
      load_xrow<KL>(xs + static_cast<size_t>(c_s[q + G]) * KP, x1);
      load_xrow<KL>(xs + static_cast<size_t>(c_s[q + 2 * G]) * KP, x2);
      load_xrow<KL>(xs + static_cast<size_t>(c_s[q + 3 * G]) * KP, x3);
      const TA v0 = v_s[q], v1 = v_s[q + G], v2 = v_s[q + 2 * G], v3 = v_s[q + 3 * G];
#pragma unroll
      for (int j = 0; j < KL; j++) {
        acc[j] += v0 * x0[j];
        acc[j] += v1 * x1[j];
        acc[j] += v2 * x2[j];
        acc[j] += v3 * x3[j];
      }
    }
    for (; q < e; q += G) {
      TA x0[KL];
      load_xrow<KL>(xs + static_cast<size_t>(c_s[q]) * KP, x0);
      const TA v0 = v_s[q];
#pragma unroll
      for (int j = 0; j < KL; j++) acc[j] += v0 * x0[j];
    }
#pragma unroll
    for (int o = S; o < GS; o *= 2) {
#pragma unroll
      for (int j = 0; j < KL; j++) acc[j] += __shfl_xor(acc[j], o, 64);
    }
    if (el == 0) {
      const int64_t row = r0 + i;
#pragma unroll
      for (int j = 0; j < KL; j++) {
        const int col = sub * KL + j;
        if (col < kk) {
          TY *yp = y + (j0 + col) * a.ldy + row;
          TA out = alpha * acc[j];
          if (beta != TA(0)) out += beta * to_acc<TA>(*yp);
          *yp = from_acc<TY>(out);
        }
      }
    }
  }
}

// Sum of row entries [s, e) (LDS window, or memory for the long last row)
// times their x rows, by the whole workgroup, into red: lanes take 16-byte
// pieces of the x rows when KP spans two or more (S lanes per entry), else
// whole rows; red[c * NW + w] holds wave w's sum for column c of the chunk.
template <int WG, int KP, int MODE, typename TV, typename TX, typename TA, typename TI, bool LDSW>
__device__ __forceinline__ void wg_row_sums(const FwdArgs &a, const TA *v_s, const uint32_t *c_s,
                                            int64_t s, int64_t e, const TX *xrc, TA *red,
                                            int tid) {
  constexpr int NW = WG / 64;
  constexpr int KL0 = 16 / static_cast<int>(sizeof(TX));
  constexpr bool kSplit = (sizeof(TX) == 4 || sizeof(TX) == 8) && KP >= 2 * KL0;
  constexpr int KL = kSplit ? KL0 : KP;
  constexpr int S = KP / KL;
  const int sub = tid % S, el = tid / S;
  constexpr int NE = WG / S;  // entry lanes
  const TI *colind = static_cast<const TI *>(a.colind);
  const TV *vals = static_cast<const TV *>(a.vals);
  const TI *perm = static_cast<const TI *>(a.perm);
  auto entry = [&](int64_t q, TA &w, uint32_t &c) {
    if constexpr (LDSW) {
      w = v_s[q];
      c = c_s[q];
    } else {
      w = to_acc<TA>(long_val<MODE>(vals, perm, q));
      c = col32(ld_stream(colind + q));
    }
  };
  const TX *xs = xrc + sub * KL;
  TA acc[KL];
#pragma unroll
  for (int j = 0; j < KL; j++) acc[j] = 0;
  int64_t q = s + el;
  for (; q + NE < e; q += 2 * NE) {
    TA w0, w1;
    uint32_t c0, c1;
    entry(q, w0, c0);
    entry(q + NE, w1, c1);
    TA x0[KL], x1[KL];
    load_xrow<KL>(xs + static_cast<size_t>(c0) * KP, x0);
    load_xrow<KL>(xs + static_cast<size_t>(c1) * KP, x1);
#pragma unroll
    for (int j = 0; j < KL; j++) {
      acc[j] += w0 * x0[j];
      acc[j] += w1 * x1[j];
    }
  }
  if (q < e) {
    TA w0;
    uint32_t c0;
    entry(q, w0, c0);
    TA x0[KL];
    load_xrow<KL>(xs + static_cast<size_t>(c0) * KP, x0);
#pragma unroll
    for (int j = 0; j < KL; j++) acc[j] += w0 * x0[j];
  }
  // across the wave's entry lanes (stride S), then per wave into red
#pragma unroll
  for (int o = S; o < 64; o *= 2) {
#pragma unroll
    for (int j = 0; j < KL; j++) acc[j] += __shfl_xor(acc[j], o, 64);
  }
  __syncthreads();
  if ((tid & 63) < S) {
#pragma unroll
    for (int j = 0; j < KL; j++) red[(sub * KL + j) * NW + (tid >> 6)] = acc[j];
  }
  __syncthreads();
}

// Blocks of fewer long rows than waves (rows of a wide band: one row of
// ~2000 entries per block): every row by the whole workgroup, each lane
// walking its entries with two x rows in flight, the KP sums reduced
// across the waves through LDS. With a wave per row the other waves sat
// idle and each lane walked four times as many entries one after another.
template <int WG, int KP, typename TX, typename TY, typename TA, typename TI>
__device__ __forceinline__ void reduce_rows_mrhs_wg(
    const TA *v_s, const uint32_t *c_s, const int *rp_s, bool rows_lds, const FwdArgs &a,
    int64_t r0, int64_t e0, int64_t nr, const TX *xrc, int64_t j0, int kk, TA alpha, TA beta,
    TA *red, int tid) {
  constexpr int NW = WG / 64;
  static_assert(KP <= 16, "red holds 16 sums per wave");
  const TI *rowptr = static_cast<const TI *>(a.rowptr);
  TY *y = static_cast<TY *>(a.y);
  for (int64_t i = 0; i < nr; i++) {
    int s, e;
    if (rows_lds) {
      s = rp_s[i];
      e = rp_s[i + 1];
    } else {
      s = static_cast<int>(row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r0 + i) - e0);
      e = static_cast<int>(row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, r0 + i + 1) - e0);
    }
    wg_row_sums<WG, KP, kPlain, TA, TX, TA, TI, true>(a, v_s, c_s, s, e, xrc, red, tid);
    if (tid < kk) {
      TA sum = 0;
#pragma unroll
      for (int w = 0; w < NW; w++) sum += red[tid * NW + w];
      TY *yp = y + (j0 + tid) * a.ldy + r0 + i;
      TA out = alpha * sum;
      if (beta != TA(0)) out += beta * to_acc<TA>(*yp);
      *yp = from_acc<TY>(out);
    }
  }
  __syncthreads();  // red is reused by the next chunk of right-hand sides
}

template <int WG, int KP, typename TX, typename TY, typename TA, typename TI>
__device__ __forceinline__ void reduce_rows_mrhs_any(
    int n_s, const TA *v_s, const uint32_t *c_s, const int *rp_s,
    bool rows_lds, const FwdArgs &a, int64_t r0, int64_t e0, int64_t nr,
    const TX *xrc, int64_t j0, int kk, TA alpha, TA beta, TA *red, int tid) {
  if (nr <= 0) return;
  const int64_t avg = (n_s / nr) >> a.gshift;
  if (nr < WG / 64 && n_s / nr > 96) {
    reduce_rows_mrhs_wg<WG, KP, TX, TY, TA, TI>(v_s, c_s, rp_s, rows_lds, a, r0, e0, nr, xrc, j0,
                                               kk, alpha, beta, red, tid);
    return;
  }
  constexpr int KL = 16 / static_cast<int>(sizeof(TX));
  if constexpr ((sizeof(TX) == 4 || sizeof(TX) == 8) && KP >= 2 * KL) {
    // 16-byte pieces of the x rows, S lanes per entry (G entry lanes per
    // row, G * S <= 64)
    constexpr int GMAX = 64 / (KP / KL);
    constexpr int G2 = 2 < GMAX ? 2 : GMAX, G4 = 4 < GMAX ? 4 : GMAX;
    constexpr int G8 = 8 < GMAX ? 8 : GMAX, G16 = 16 < GMAX ? 16 : GMAX;
#define DRX_RS(G) reduce_rows_mrhs_split<WG, G, KP, TX, TY, TA, TI>(v_s, c_s, rp_s, rows_lds, a, r0, e0, nr, xrc, j0, kk, alpha, beta, tid)
    if (avg <= 3)
      DRX_RS(1);
    else if (avg <= 6)
      DRX_RS(G2);
    else if (avg <= 12)
      DRX_RS(G4);
    else if (avg <= 24)
      DRX_RS(G8);
    else
      DRX_RS(G16);
#undef DRX_RS
    return;
  }
#define DRX_RM(G) reduce_rows_mrhs<WG, G, KP, TX, TY, TA, TI>(v_s, c_s, rp_s, rows_lds, a, r0, e0, nr, xrc, j0, kk, alpha, beta, tid)
  if (avg <= 3)
    DRX_RM(1);
  else if (avg <= 6)
    DRX_RM(2);
  else if (avg <= 12)
    DRX_RM(4);
  else if (avg <= 24)
    DRX_RM(8);
  else if (avg <= 48)
    DRX_RM(16);
  else if (avg <= 96)
    DRX_RM(32);
  else
    DRX_RM(64);
#undef DRX_RM
}

// One row block for k >= 2 right-hand sides, x read from a.xr (all lanes of
// the workgroup call it together). MODE is kPlain or kPerm.
template <int WG, int CAP, int KP, int MODE, typename TV, typename TX,
          typename TY, typename TA, typename TI>
__device__ __forceinline__ void process_block_mrhs(const FwdArgs &a,
                                                   int64_t blk,
                                                   MrhsSmem<WG, CAP, TA> &sm) {
  static_assert(MODE == kPlain || MODE == kPerm, "row maps use k = 1 panels");
  constexpr int T = CAP / WG;
  constexpr int ROWCAP = CAP / 2;
  constexpr int NW = WG / 64;
  // The lane index as a value of this block: the batch kernel walks blocks
  // in a loop, and values derived from the lane index for every lane-group
  // size were hoisted out of it and held in registers (4 waves instead of 6).
  int tid;
  asm volatile("" : "=v"(tid) : "0"(static_cast<int>(threadIdx.x)));
  const TI *rowptr = static_cast<const TI *>(a.rowptr);
  const TI *colind = static_cast<const TI *>(a.colind);
  const TV *vals = static_cast<const TV *>(a.vals);
  const TI *perm = static_cast<const TI *>(a.perm);
  const TX *xr = static_cast<const TX *>(a.xr);
  TY *y = static_cast<TY *>(a.y);
  const TA alpha = static_cast<TA>(a.alpha), beta = static_cast<TA>(a.beta);

  const BlockRec b = a.rec[blk];
  const int64_t r1 = a.rec[blk + 1].r0, e1 = a.rec[blk + 1].e0;
  if (b.r0 >= r1) return;
  const int n_s = static_cast<int>(b.es - b.e0);
  const int64_t nr = (b.es == e1 ? r1 : r1 - 1) - b.r0;
  const bool rows_lds = nr < ROWCAP;

  // Stage the window in LDS (reads past n_s return 0 and touch no memory).
  {
    TA v[T];
    uint32_t c[T];
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
#pragma unroll
    for (int u = 0; u < T; u++) {
      const int idx = tid + u * WG;
      if (idx < n_s) {
        sm.v[idx] = v[u];
        sm.c[idx] = c[u];
      }
    }
  }
  if (rows_lds) {
    for (int i = tid; i <= nr; i += WG)
      sm.rp_s[i] = static_cast<int>(
          row_begin(rowptr, a.base, a.nrows, a.nnz, a.sent, b.r0 + i) - b.e0);
  }
  __syncthreads();

  const int64_t nchunk = (a.k + KP - 1) / KP;
  for (int64_t jc = 0; jc < nchunk; jc++) {
    const int kk = static_cast<int>(a.k - jc * KP < KP ? a.k - jc * KP : KP);
    reduce_rows_mrhs_any<WG, KP, TX, TY, TA, TI>(
        n_s, sm.v, sm.c, sm.rp_s, rows_lds, a, b.r0, b.e0, nr,
        xr + jc * a.ncols * KP, jc * KP, kk, alpha, beta, sm.red, tid);
  }

  if (b.es < e1) {
    // Long last row: the whole workgroup, strided coalesced loads.
    const int64_t last = r1 - 1;
    for (int64_t jc = 0; jc < nchunk; jc++) {
      const int kk = static_cast<int>(a.k - jc * KP < KP ? a.k - jc * KP : KP);
      const TX *xrc = xr + jc * a.ncols * KP;
      wg_row_sums<WG, KP, MODE, TV, TX, TA, TI, false>(a, nullptr, nullptr, b.es, e1, xrc,
                                                      sm.red, tid);
      if (tid < kk) {
        TA s = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) s += sm.red[tid * NW + w];
        TY *yp = y + (jc * KP + tid) * a.ldy + last;
        TA out = alpha * s;
        if (beta != TA(0)) out += beta * to_acc<TA>(*yp);
        *yp = from_acc<TY>(out);
      }
    }
  }
  __syncthreads();  // the next block overwrites the LDS window
}

// Waves per SIMD of the multi-RHS kernels: as many workgroups as the LDS
// holds, at most the most that compile without spilling: 6 for sums of up
// to 16 bytes per lane (73-78 VGPRs), 5 for wider ones taken in 16-byte
// pieces (95), 4 for 16-bit x in wide chunks (whole chunks per lane).
template <int WG, int CAP, int KP, typename TX, typename TA> constexpr int mrhs_waves() {
  constexpr int lds = static_cast<int>(sizeof(MrhsSmem<WG, CAP, TA>));
  constexpr int w = (160 * 1024) / lds * (WG / 256);
  constexpr int most = KP * sizeof(TA) <= 16 ? 6 : (sizeof(TX) == 2 ? 4 : 5);
  return w > most ? most : (w < 1 ? 1 : w);
}

template <int WG, int CAP, int KP, int MODE, typename TV, typename TX,
          typename TY, typename TA, typename TI>
__global__ __launch_bounds__(WG)
__attribute__((amdgpu_waves_per_eu(mrhs_waves<WG, CAP, KP, TX, TA>(), 8)))
void k_spmv_mrhs(FwdArgs a) {
  __shared__ MrhsSmem<WG, CAP, TA> sm;
  for (int64_t wi = blockIdx.x; wi < a.per_xcd * kXcds; wi += a.grid) {
    const int64_t blk = map_block(wi, a.per_xcd);
    if (blk >= a.nblocks) continue;
    process_block_mrhs<WG, CAP, KP, MODE, TV, TX, TY, TA, TI>(a, blk, sm);
  }
}

// Row-major chunks of x (layout above); one thread per (chunk, column).
template <int KP, typename TX>
__global__ void k_x_rowmajor(const TX *x, int64_t ldx, int64_t ncols,
                             int64_t k, TX *xr) {
  const int64_t nchunk = (k + KP - 1) / KP;
  const int64_t total = nchunk * ncols;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       t < total; t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t jc = t / ncols, c = t - jc * ncols;
    TX buf[KP];
#pragma unroll
    for (int jj = 0; jj < KP; jj++) {
      const int64_t j = jc * KP + jj;
      buf[jj] = j < k ? x[j * ldx + c] : TX{};
    }
    constexpr int BYTES = KP * static_cast<int>(sizeof(TX));
    if constexpr (BYTES >= 16) {
      uint4 *dst = reinterpret_cast<uint4 *>(xr + t * KP);
#pragma unroll
      for (int i = 0; i < BYTES / 16; i++) {
        uint4 w;
        __builtin_memcpy(&w, reinterpret_cast<const char *>(buf) + 16 * i, 16);
        dst[i] = w;
      }
    } else if constexpr (BYTES == 8) {
      uint2 w;
      __builtin_memcpy(&w, buf, 8);
      *reinterpret_cast<uint2 *>(xr + t * KP) = w;
    } else {
      uint32_t w;
      __builtin_memcpy(&w, buf, 4);
      *reinterpret_cast<uint32_t *>(xr + t * KP) = w;
    }
  }
}

// Host side (drx_mrhs.hip): copy x into row-major chunks and launch the
// multi-RHS kernel for a plain (mode kPlain) or permuted (kPerm) share.
void run_mrhs(const Plan &p, FwdArgs a, int mode, int vtype, int xtype,
              int ytype, hipStream_t s);

}  // namespace drx
