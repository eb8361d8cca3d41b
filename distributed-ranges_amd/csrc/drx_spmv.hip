// drx_spmv.hip -- CSR sparse gemv forward kernel for gfx950 (MI355X) and its
// C entry points.
//
// y[:, j] = alpha * A x[:, j] + beta * y[:, j] over one CSR share.
//
// One workgroup per row block of the plan; the block body is in
// drx_fwd.hpp.
// Workgroup i takes block (i % 8) * ceil(nb / 8) + i / 8: each XCD (which
// receives every 8th workgroup) walks a contiguous range of rows, so its L2
// keeps the x lines that neighbouring rows share.
//
// Reference kernels replaced: csr_row_distribution.hpp:84-115,
// csr_eq_distribution.hpp:62-111, sp/algorithms/matrix/local_gemv.hpp:22-78.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <type_traits>

#include "drx_common.hpp"
#include "drx_fwd.hpp"
#include "drx_internal.hpp"

namespace drx {

// ---------------------------------------------------------------------------
// Launch helpers and type dispatch
// ---------------------------------------------------------------------------
Grid grid_for(int64_t nblocks) {
  Grid g;
  g.per_xcd = (nblocks + kXcds - 1) / kXcds;
  const int64_t want = g.per_xcd * kXcds;
  // a multiple of 8 (workgroups keep their XCD as they loop past it)
  const int64_t tuned = tuned_grid_cap();
  const int64_t cap = tuned > 0 ? (tuned + kXcds - 1) / kXcds * kXcds : int64_t(1) << 22;
  g.grid = std::max<int64_t>(kXcds, std::min(want, cap));
  return g;
}

// Window 4096 on 256 lanes (16 entries per lane) for values of at most 4
// bytes with 32-bit indices in plain mode (the permuted-values variant
// spills registers): for FP32 measured faster than 512 lanes x 8 (band
// +-16: 3544 vs 3006 GB/s) and than 1024-entry blocks (3-D Laplacian: 3369
// vs 3094). 16-bit values take it when knob fwd_cap asks for 4096.
template <int MODE, typename TV, typename TA, typename TI>
constexpr bool wide_window() {
  return MODE == kPlain && sizeof(TV) <= 4 && sizeof(TA) == 4 && sizeof(TI) == 4;
}

template <int MODE, bool BUFX, typename TV, typename TX, typename TY,
          typename TA, typename TI, int VW, bool XS = false>
static void launch_fwd_vw(int64_t cap, const FwdArgs &a, hipStream_t s) {
  const dim3 grid(static_cast<unsigned>(a.grid));
  if (cap <= 1024)
    hipLaunchKernelGGL((k_spmv_fwd<256, 1024, MODE, BUFX, TV, TX, TY, TA, TI, VW, XS>),
                       grid, dim3(256), 0, s, a);
  else if (cap <= 2048)
    hipLaunchKernelGGL((k_spmv_fwd<256, 2048, MODE, BUFX, TV, TX, TY, TA, TI, VW, XS>),
                       grid, dim3(256), 0, s, a);
  else if (cap <= 4096) {
    if constexpr (wide_window<MODE, TV, TA, TI>())
      hipLaunchKernelGGL((k_spmv_fwd<256, 4096, MODE, BUFX, TV, TX, TY, TA, TI, VW, XS>),
                         grid, dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL((k_spmv_fwd<512, 4096, MODE, BUFX, TV, TX, TY, TA, TI, VW, XS>),
                         grid, dim3(512), 0, s, a);
  } else
    hipLaunchKernelGGL((k_spmv_fwd<1024, 8192, MODE, BUFX, TV, TX, TY, TA, TI, VW, XS>),
                       grid, dim3(1024), 0, s, a);
}

// The row-gather kernel that copies windows straight into LDS needs vals
// aligned to its copy unit and colind to 16 bytes.
template <typename TV> static bool rows_lds_copy(const FwdArgs &a) {
  return reinterpret_cast<uintptr_t>(a.vals) % lds_copy_width<TV>() == 0 &&
         reinterpret_cast<uintptr_t>(a.colind) % 16 == 0;
}

// Row-ordered gathers (plain CSR, k = 1) for the 2048- and 4096-entry
// windows.
template <bool BUFX, typename TV, typename TX, typename TY, typename TA, typename TI, int VW>
static bool launch_fwd_rg(int64_t cap, const FwdArgs &a, hipStream_t s) {
  const dim3 grid(static_cast<unsigned>(a.grid));
  if (cap == 2048) {
    hipLaunchKernelGGL((k_spmv_fwd<256, 2048, kPlain, BUFX, TV, TX, TY, TA, TI, VW, false, true>),
                       grid, dim3(256), 0, s, a);
    return true;
  }
  if (cap == 4096) {
    if constexpr (wide_window<kPlain, TV, TA, TI>())
      hipLaunchKernelGGL((k_spmv_fwd<256, 4096, kPlain, BUFX, TV, TX, TY, TA, TI, VW, false, true>),
                         grid, dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL((k_spmv_fwd<512, 4096, kPlain, BUFX, TV, TX, TY, TA, TI, VW, false, true>),
                         grid, dim3(512), 0, s, a);
    return true;
  }
  return false;
}

// Groups of VW consecutive entries per lane (32-bit indices, VW values in at
// most 16 bytes, vals / colind aligned to a group; not the permuted-values
// mode). Knob "vec": 0 (default) by type: groups of four for values of at
// most 4 bytes, single entries for FP64; 1 single entries; 2 pairs; 4
// groups of four, wherever eligible (else single entries). Measured with
// every gather in flight (us per call): 3-D Laplacian 256^3 FP32 271 -> 222,
// BF16 / FP16 234 / 260 -> 199, headline Zipf column panels 8.44 -> 8.12 ms
// per step; band +-16 FP32 211 -> 212 (its gathers, coalesced across lanes
// with one entry per lane, spread over four times the lines with groups);
// FP64 pairs (3-D Laplacian 403 -> 420) spill.
template <int MODE, bool BUFX, typename TV, typename TX, typename TY,
          typename TA, typename TI>
static void launch_fwd(const Plan &p, FwdArgs a, hipStream_t s) {
  int64_t cap = p.cap;
  int vw = 1;
  if constexpr (MODE != kPerm && sizeof(TI) == 4) {
    const int vec = tuned_vec();
    const int want = vec == 0 ? (sizeof(TV) <= 4 ? 4 : 1) : vec;
    auto aligned = [&](int w) {
      return reinterpret_cast<uintptr_t>(a.vals) % (w * sizeof(TV)) == 0 &&
             reinterpret_cast<uintptr_t>(a.colind) % (w * 4) == 0;
    };
    if (want == 4 && 4 * sizeof(TV) <= 16 && aligned(4))
      vw = 4;
    else if (want == 2 && 2 * sizeof(TV) <= 16 && aligned(2))
      vw = 2;
  }
  if constexpr (MODE == kPlain) {
    // Window per value type (knob fwd_cap overrides): 4096 for 4-byte
    // values and indices (16 entries per lane; knob wide = 0 keeps the
    // plan's own), else the plan's own (2048 with the default block size).
    // Measured on filled windows (3-D Laplacian 256^3 / band +-16, 4M rows):
    // FP64 2048: 400 / 388 us, 1024 (every gather in flight): 472 / 449 us,
    // 4096 (512 lanes): 436 / 405 us; BF16 2048: 234 us, 1024: 306 us,
    // 4096: 263 us; FP32 band 4096: 217 us, 2048: 244 us.
    // 16-bit values take 4096 with groups of four (band +-16 BF16: 189 vs
    // 195 us; 3-D Laplacian equal), 2048 with single entries (4096: 296 us).
    // Row-ordered gathers keep the plan's 2048-entry windows, except 16-bit
    // values copied straight into LDS, which take 4096 (3-D Laplacian 256^3,
    // us per call, windows of 1024 / 2048 / 4096 entries: FP64 361 / 351 /
    // 362, FP32 218 / 199 / 210, BF16 208 / 173 / 164; 2-D Laplacian 1000^2
    // FP64 20.5 / 19.8 / 22.8; through registers 4096 is slower for every
    // type).
    int64_t want = tuned_fwd_cap();
    if (want == 0 && a.rowgather == 1 && sizeof(TV) == 2 && rows_lds_copy<TV>(a))
      want = 4096;
    if (want == 0 && !a.rowgather && wide_window<MODE, TV, TA, TI>() && tuned_wide() != 0 &&
        (sizeof(TV) == 4 || vw == 4))
      want = 4096;
    const BlockRec *rec = nullptr;
    int64_t nb = 0;
    if (want == 4096 && p.rec_wide != nullptr) {
      rec = p.rec_wide;
      nb = p.nblocks_wide;
    }
    if (rec != nullptr) {
      a.rec = rec;
      a.nblocks = nb;
      const Grid g = grid_for(nb);
      a.per_xcd = g.per_xcd;
      a.grid = g.grid;
      cap = want;
    }
  }
  if constexpr (MODE == kPerm && BUFX) {
    // the adjoint through a permutation: row-ordered, perm copied into LDS
    // (perm and colind aligned to 16 bytes)
    if (a.rowgather && a.rowgather != 2 && a.k == 1 && cap == 2048 &&
        reinterpret_cast<uintptr_t>(a.perm) % 16 == 0 &&
        reinterpret_cast<uintptr_t>(a.colind) % 16 == 0) {
      hipLaunchKernelGGL((k_spmv_rows<256, 2048, TV, TX, TY, TA, TI, kPerm>), dim3(a.grid),
                         dim3(256), 0, s, a);
      return;
    }
  }
  if constexpr (MODE == kPlain && BUFX) {
    if (a.rowgather && a.k == 1) {
      // The window copied straight into LDS (vals aligned to its copy unit,
      // colind to 16 bytes) unless knob 2 asks for the register path.
      // Measured against the register path (us per call): 3-D Laplacian
      // 256^3 FP64 382 -> 354, FP32 208 -> 204, BF16 174 -> 172; 2-D
      // Laplacian 1000^2 FP64 21.2 -> 20.1.
      if (a.rowgather != 2 && (cap == 1024 || cap == 2048 || cap == 4096) &&
          rows_lds_copy<TV>(a)) {
        if (cap == 1024)
          hipLaunchKernelGGL((k_spmv_rows<256, 1024, TV, TX, TY, TA, TI>), dim3(a.grid),
                             dim3(256), 0, s, a);
        else if (cap == 2048)
          hipLaunchKernelGGL((k_spmv_rows<256, 2048, TV, TX, TY, TA, TI>), dim3(a.grid),
                             dim3(256), 0, s, a);
        else
          hipLaunchKernelGGL((k_spmv_rows<256, 4096, TV, TX, TY, TA, TI>), dim3(a.grid),
                             dim3(256), 0, s, a);
        return;
      }
      if constexpr (sizeof(TI) == 4 && 4 * sizeof(TV) <= 16) {
        if (vw == 4 && launch_fwd_rg<BUFX, TV, TX, TY, TA, TI, 4>(cap, a, s)) return;
      }
      if (vw == 1 && launch_fwd_rg<BUFX, TV, TX, TY, TA, TI, 1>(cap, a, s)) return;
    }
  }
  // x staged in LDS (knob xstage 2): plain CSR with x of the values' type
  // (instantiated apart, so the other kernels keep their registers).
  constexpr bool kCanStage = MODE == kPlain && BUFX && std::is_same_v<TV, TX>;
  if constexpr (kCanStage && sizeof(TI) == 4 && 4 * sizeof(TV) <= 16) {
    if (vw == 4) {
      if (a.xstage)
        launch_fwd_vw<MODE, BUFX, TV, TX, TY, TA, TI, 4, true>(cap, a, s);
      else
        launch_fwd_vw<MODE, BUFX, TV, TX, TY, TA, TI, 4>(cap, a, s);
      return;
    }
  }
  if constexpr (MODE != kPerm && sizeof(TI) == 4 && 4 * sizeof(TV) <= 16) {
    if (vw == 4) {
      launch_fwd_vw<MODE, BUFX, TV, TX, TY, TA, TI, 4>(cap, a, s);
      return;
    }
  }
  if constexpr (MODE != kPerm && sizeof(TI) == 4 && 2 * sizeof(TV) <= 16) {
    if (vw == 2) {
      launch_fwd_vw<MODE, BUFX, TV, TX, TY, TA, TI, 2>(cap, a, s);
      return;
    }
  }
  if constexpr (kCanStage) {
    if (a.xstage) {
      launch_fwd_vw<MODE, BUFX, TV, TX, TY, TA, TI, 1, true>(cap, a, s);
      return;
    }
  }
  launch_fwd_vw<MODE, BUFX, TV, TX, TY, TA, TI, 1>(cap, a, s);
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
  a.gshift = 0;
  a.lmax = p.lmax;
  const uint64_t xb =
      static_cast<uint64_t>(std::max<int64_t>(p.ncols, 1)) * dtype_size(xtype);
  const bool bufx = xb < (uint64_t(1) << 32);
  // Gathers are range-checked against xbytes, and lanes past a window
  // gather column 0: a share without entries (x may be null or empty) gets
  // an empty range.
  a.xbytes = bufx && p.nnz > 0 ? static_cast<uint32_t>(xb) : 0;
  a.ncols = p.ncols;
  if constexpr (MODE != kRowMap) {
    // Row-major chunks of x for k >= 2, except up to 4 right-hand sides over
    // short rows with local columns (stencils): their column gathers are
    // already coalesced and the copy of x is large next to the matrix
    // (2-D Laplacian FP64, k = 2 / 3 / 4: 3009 / 2891 / 2811 GB/s through
    // column gathers vs 2151 / 2245 / 2572 through chunks; a +-16 band, 33
    // entries per row, is faster through chunks at every k).
    const int mr = p.no_scratch ? 0 : tuned_mrhs();
    const bool stencil = k <= 4 && p.nnz <= 8 * p.nrows;
    if (k >= 2 && mr != 0 &&
        (mr == 2 || !stencil || (column_locality(p, colind, s) & 1) == 0)) {
      run_mrhs(p, a, MODE, vtype, xtype, ytype, s);
      return;
    }
  }
  // Staging each block's x window in LDS (one right-hand side) only when
  // the knob forces it: with all of a lane's gathers in flight, per-entry
  // gathers beat it on bands (+-16 FP32: 241 vs 286 us per call).
  if constexpr (MODE == kPerm) {
    // Row-ordered gathers of x and of the values through the permutation
    // (values gathered through a buffer resource: nnz values within 4 GB).
    if (k == 1 && bufx && p.nnz > 0 &&
        static_cast<uint64_t>(p.nnz) * dtype_size(vtype) < (uint64_t(1) << 32)) {
      const int rg = tuned_rowgather();
      a.rowgather = rg >= 2 ? rg
                    : (rg == 1 && p.lmax > 0 && p.lmax <= 128 && p.nnz >= 4 * p.nrows) ? 1
                                                                                     : 0;
    }
  }
  if constexpr (MODE == kPlain) {
    if (k == 1 && bufx && p.nnz > 0) a.xstage = tuned_xstage() == 2;
    // Row-ordered gathers for shares of rows of at most 128 entries (at
    // least 4 on average; knob rowgather), unless x staging is forced.
    // Measured with the window copied into LDS (us per call, against
    // entry-ordered gathers, one run): 3-D Laplacian 256^3 FP64 397 -> 355,
    // FP32 227 -> 210, BF16 201 -> 167; 2-D Laplacian 1000^2 FP64 21.6 ->
    // 19.9; band +-16 FP32 217 -> 210, FP64 350 -> 321, BF16 194 -> 179;
    // band +-40 FP32 283 -> 258; random columns, 16M rows of 4..12 entries
    // FP32 2369 -> 2379, 4M rows of 16..48 FP32 9130 -> 9168, BF16 8127 ->
    // 7947, 2M rows of 40..64 FP64 16143 -> 15813.
    if (k == 1 && bufx && p.nnz > 0) {
      const int rg = tuned_rowgather();
      a.rowgather = a.xstage ? 0
                    : rg >= 2 ? rg
                    : (rg == 1 && p.lmax > 0 && p.lmax <= 128 && p.nnz >= 4 * p.nrows) ? 1
                                                                                     : 0;
    }
  }
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
  // Column panels serve k = 1; with k >= 2 the row-major x chunks are KP
  // times wider than a panel's L2-sized slice, and the product runs over
  // the share's own arrays.
  if (p.panels && (k == 1 || tuned_mrhs() == 0)) {
    run_spmv_panels(p, vtype, colind, vals, xtype, x, ldx, ytype, y, ldy, k,
                    alpha, beta, s);
    return;
  }
  if (p.panels) check_panel_operands(p, vtype, colind, vals);
  run_spmv_impl<kPlain>(p, vtype, colind, vals, nullptr, nullptr, xtype, x,
                        ldx, ytype, y, ldy, k, alpha, beta, s);
}

void run_spmv_rowmap(const Plan &p, int vtype, const void *colind,
                     const void *vals, const void *rowmap, int xtype,
                     const void *x, int64_t ldx, int ytype, void *y,
                     int64_t ldy, int64_t k, double alpha, double beta,
                     hipStream_t s) {
  run_spmv_impl<kRowMap>(p, vtype, colind, vals, nullptr, rowmap, xtype, x,
                         ldx, ytype, y, ldy, k, alpha, beta, s);
}

void run_spmv_perm(const Plan &p, int vtype, const void *colind,
                   const void *vals, const void *perm, int xtype,
                   const void *x, int64_t ldx, int ytype, void *y,
                   int64_t ldy, int64_t k, double alpha, double beta,
                   hipStream_t s) {
  run_spmv_impl<kPerm>(p, vtype, colind, vals, perm, nullptr, xtype, x, ldx,
                       ytype, y, ldy, k, alpha, beta, s);
}

// ---------------------------------------------------------------------------
// Scratch for the entry point that allocates nothing (drx_spmv_csr_noalloc):
// the plan records of one call live in this file-scope device array, one
// copy per GPU, held until the process ends. Its device address comes from
// hipGetSymbolAddress once per device (with that device current); host code
// passes that address, never the array's name, to kernels.
// ---------------------------------------------------------------------------
constexpr int64_t kNoallocRecords = 262144;  // 8 MiB of 32-byte records
static __device__ __align__(256) BlockRec scratch[kNoallocRecords];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static BlockRec *g_scratch_ptr[kMaxDevices] = {};

static BlockRec *scratch_for_device(int dev) {
  DRX_REQUIRE(dev >= 0 && dev < kMaxDevices, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void *ptr = nullptr;
    DRX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    DRX_REQUIRE(ptr != nullptr, "no device address for the scratch records");
    g_scratch_ptr[dev] = static_cast<BlockRec *>(ptr);
  }
  return g_scratch_ptr[dev];
}

}  // namespace drx

using namespace drx;

extern "C" {

int64_t drx_spmv_csr_noalloc_max_nnz(void) {
  // largest nnz with records_needed(nnz, 1024) <= kNoallocRecords
  int64_t lo = 0, hi = kNoallocRecords * int64_t(1024);
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo + 1) / 2;
    if (records_needed(mid, 1024) <= kNoallocRecords)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

int drx_spmv_csr_noalloc(int itype, const void *rowptr, int64_t rowptr_base,
                         int64_t nrows, int64_t ncols, int64_t nnz,
                         int has_sentinel, const void *colind, int vtype,
                         const void *vals, int xtype, const void *x,
                         int64_t ldx, int ytype, void *y, int64_t ldy,
                         int64_t k, double alpha, double beta,
                         drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(nnz >= 0 && records_needed(nnz, 1024) <= kNoallocRecords,
                "drx_spmv_csr_noalloc: the share exceeds the scratch records "
                "(see drx_spmv_csr_noalloc_max_nnz)");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int dev = 0;
    DRX_HIP(hipGetDevice(&dev));
    BlockRec *buf = scratch_for_device(dev);
    Plan p;
    init_plan_in(p, itype, rowptr, rowptr_base, nrows, ncols, nnz, has_sentinel,
                 buf, kNoallocRecords, s);
    run_spmv(p, vtype, colind, vals, xtype, x, ldx, ytype, y, ldy, k, alpha,
             beta, s);
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

}  // extern "C"
