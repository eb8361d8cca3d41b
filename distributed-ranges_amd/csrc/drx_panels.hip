// drx_panels.hip -- column-panel layout of a plan, for matrices whose x does
// not fit the L2 of an XCD (4 MB) and whose columns are irregular.
//
// Random 4-byte gathers on MI355X run at ~220-250 G/s while the gathered
// range fits L2 and at ~55 G/s beyond it (MALL or HBM alike), so a matrix
// with uniformly random columns over a 64 MB x is bound by its gathers. The
// panel layout splits the columns into panels of W columns (W * sizeof(x)
// about 2 MB) and re-orders a copy of the entries panel-major: a stable radix
// sort by panel keeps rows, and columns within rows, in order. Inside a panel
// the rows that have entries are compressed into segments (row id, first
// entry), which form one CSR share per panel with a row map. A product runs
// one forward launch per panel: a row's first segment in panel order writes
// y[row] = beta * y[row] + alpha * sum (its row-map entry carries a flag
// bit), later segments add alpha * sum, and rows without entries are scaled
// by a small launch over their list. Every launch gathers from one L2-sized
// slice of x. Summation order is fixed (panel by panel), so results are
// deterministic.
//
// The copy is tied to the colind/vals given to drx_plan_set_panels; new
// values are re-packed through the kept permutation by
// drx_plan_update_values.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <vector>

#include "drx_common.hpp"
#include "drx_internal.hpp"

namespace drx {

struct PanelSet {
  int64_t width = 0, npanels = 0, nseg = 0;
  int vtype = 0;
  const void *src_colind = nullptr, *src_vals = nullptr;
  void *pcol = nullptr, *pval = nullptr, *perm = nullptr;
  void *seg_row = nullptr, *seg_ptr = nullptr;
  // rows of the share without entries (y = beta * y for them; every other
  // row is written by its first segment)
  void *empty_rows = nullptr;
  int64_t nempty = 0;
  std::vector<int64_t> pstart, pseg;  // npanels + 1 each (host)
  std::vector<Plan *> plans;
};

void free_panels(PanelSet *ps) {
  if (!ps) return;
  for (auto *q : ps->plans) free_plan(q, nullptr);
  for (void *ptr : {ps->pcol, ps->pval, ps->perm, ps->seg_row, ps->seg_ptr, ps->empty_rows})
    if (ptr) (void)hipFree(ptr);
  delete ps;
}

static dim3 grid_of(int64_t n, int tpb = 256) {
  int64_t b = (n + tpb - 1) / tpb;
  if (b > (int64_t(1) << 20)) b = int64_t(1) << 20;
  return dim3(static_cast<unsigned>(b < 1 ? 1 : b));
}

#define DRX_GRID_LOOP(i, n)                                                    \
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; \
       i < (n); i += static_cast<int64_t>(gridDim.x) * blockDim.x)

// Local row of every entry and its panel key.
template <typename TI>
__global__ void k_row_key(const TI *rowptr, int64_t base, int64_t nrows,
                          int64_t nnz, int sent, const TI *colind,
                          int64_t width, TI *rowid, uint32_t *key, TI *iota) {
  DRX_GRID_LOOP(e, nnz) {
    int64_t lo = 0, hi = nrows - 1;  // last row with row_begin <= e
    while (lo < hi) {
      const int64_t mid = lo + (hi - lo + 1) / 2;
      if (row_begin(rowptr, base, nrows, nnz, sent, mid) <= e)
        lo = mid;
      else
        hi = mid - 1;
    }
    rowid[e] = static_cast<TI>(lo);
    key[e] = static_cast<uint32_t>(static_cast<int64_t>(colind[e]) / width);
    iota[e] = static_cast<TI>(e);
  }
}

template <typename TI, typename TV>
__global__ void k_gather_entries(const TI *perm, const TI *colind,
                                 const TV *vals, const TI *rowid, int64_t nnz,
                                 TI *pcol, TV *pval, TI *prow) {
  DRX_GRID_LOOP(i, nnz) {
    const int64_t e = static_cast<int64_t>(perm[i]);
    pcol[i] = colind[e];
    pval[i] = vals[e];
    prow[i] = rowid[e];
  }
}

template <typename TI>
__global__ void k_seg_flags(const uint32_t *key, const TI *prow, int64_t nnz,
                            int32_t *flag) {
  DRX_GRID_LOOP(i, nnz) {
    flag[i] = (i == 0 || key[i] != key[i - 1] || prow[i] != prow[i - 1]) ? 1 : 0;
  }
}

template <typename TI>
__global__ void k_seg_fill(const int32_t *flag, const int32_t *segid,
                           const TI *prow, int64_t nnz, TI *seg_row,
                           TI *seg_ptr) {
  DRX_GRID_LOOP(i, nnz) {
    if (flag[i]) {
      const int64_t s = segid[i] - 1;
      seg_row[s] = prow[i];
      seg_ptr[s] = static_cast<TI>(i);
    }
  }
}

__global__ void k_panel_bounds(const uint32_t *key, const int32_t *segid,
                               int64_t nnz, int64_t nseg, int64_t npanels,
                               int64_t *pstart, int64_t *pseg) {
  DRX_GRID_LOOP(p, npanels + 1) {
    int64_t lo = 0, hi = nnz;  // first i with key[i] >= p
    while (lo < hi) {
      const int64_t mid = lo + (hi - lo) / 2;
      if (static_cast<int64_t>(key[mid]) >= p)
        hi = mid;
      else
        lo = mid + 1;
    }
    pstart[p] = lo;
    pseg[p] = lo < nnz ? segid[lo] - 1 : nseg;
  }
}

// Lowest panel holding each row (segment starts only).
template <typename TI>
__global__ void k_row_first_panel(const int32_t *flag, const uint32_t *key, const TI *prow,
                                  int64_t nnz, uint32_t *rowfirst) {
  DRX_GRID_LOOP(i, nnz) {
    if (flag[i]) atomicMin(&rowfirst[static_cast<int64_t>(prow[i])], key[i]);
  }
}

// Marks each row's first segment in panel order (top bit of its row id).
template <typename TI>
__global__ void k_mark_first(const int32_t *flag, const int32_t *segid, const uint32_t *key,
                             const TI *prow, int64_t nnz, const uint32_t *rowfirst,
                             TI *seg_row) {
  DRX_GRID_LOOP(i, nnz) {
    if (flag[i] && key[i] == rowfirst[static_cast<int64_t>(prow[i])]) {
      const int64_t sg = segid[i] - 1;
      seg_row[sg] = static_cast<TI>(seg_row[sg] | kFirstSeg<TI>);
    }
  }
}

// Rows without segments: counted, and listed when `list` is given.
template <typename TI>
__global__ void k_empty_rows(const uint32_t *rowfirst, int64_t nrows,
                             unsigned long long *count, TI *list) {
  DRX_GRID_LOOP(r, nrows) {
    if (rowfirst[r] == 0xffffffffu) {
      const unsigned long long at = atomicAdd(count, 1ull);
      if (list != nullptr) list[at] = static_cast<TI>(r);
    }
  }
}

template <typename TI, typename TV>
__global__ void k_repack(const TI *perm, const TV *vals, int64_t nnz, TV *pval) {
  DRX_GRID_LOOP(i, nnz) { pval[i] = vals[static_cast<int64_t>(perm[i])]; }
}

template <typename T> static T *dmalloc(int64_t n) {
  T *p = nullptr;
  DRX_HIP(hipMalloc(reinterpret_cast<void **>(&p), static_cast<size_t>(std::max<int64_t>(n, 1)) * sizeof(T)));
  return p;
}

template <typename TI, typename TV>
static void build_panels(Plan &p, PanelSet &ps, const TI *colind,
                         const TV *vals, hipStream_t s) {
  const int64_t nnz = p.nnz;
  TI *rowid = nullptr, *iota = nullptr, *prow = nullptr;
  uint32_t *key = nullptr, *key_sorted = nullptr;
  int32_t *flag = nullptr, *segid = nullptr;
  void *temp = nullptr;
  int64_t *d_pstart = nullptr, *d_pseg = nullptr;
  uint32_t *rowfirst = nullptr;
  unsigned long long *cnt = nullptr;
  auto cleanup = [&] {
    for (void *q : {(void *)rowid, (void *)iota, (void *)prow, (void *)key,
                    (void *)key_sorted, (void *)flag, (void *)segid, temp,
                    (void *)d_pstart, (void *)d_pseg, (void *)rowfirst, (void *)cnt})
      if (q) (void)hipFree(q);
  };
  try {
    rowid = dmalloc<TI>(nnz);
    iota = dmalloc<TI>(nnz);
    key = dmalloc<uint32_t>(nnz);
    key_sorted = dmalloc<uint32_t>(nnz);
    ps.perm = dmalloc<TI>(nnz);
    hipLaunchKernelGGL(k_row_key<TI>, grid_of(nnz), dim3(256), 0, s,
                       static_cast<const TI *>(p.rowptr), p.base, p.nrows, nnz,
                       p.has_sentinel, colind, ps.width, rowid, key, iota);
    DRX_HIP(hipGetLastError());
    int end_bit = 1;
    while ((int64_t(1) << end_bit) < std::max<int64_t>(ps.npanels, 2)) end_bit++;
    size_t tb = 0;
    DRX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key, key_sorted, iota,
                                               static_cast<TI *>(ps.perm),
                                               static_cast<int>(nnz), 0, end_bit, s));
    DRX_HIP(hipMalloc(&temp, std::max<size_t>(tb, 1)));
    DRX_HIP(hipcub::DeviceRadixSort::SortPairs(temp, tb, key, key_sorted, iota,
                                               static_cast<TI *>(ps.perm),
                                               static_cast<int>(nnz), 0, end_bit, s));
    DRX_HIP(hipStreamSynchronize(s));
    (void)hipFree(temp);
    temp = nullptr;
    (void)hipFree(key);
    key = nullptr;
    (void)hipFree(iota);
    iota = nullptr;
    ps.pcol = dmalloc<TI>(nnz);
    ps.pval = dmalloc<TV>(nnz);
    prow = dmalloc<TI>(nnz);
    hipLaunchKernelGGL((k_gather_entries<TI, TV>), grid_of(nnz), dim3(256), 0, s,
                       static_cast<const TI *>(ps.perm), colind, vals, rowid, nnz,
                       static_cast<TI *>(ps.pcol), static_cast<TV *>(ps.pval), prow);
    DRX_HIP(hipGetLastError());
    DRX_HIP(hipStreamSynchronize(s));
    (void)hipFree(rowid);
    rowid = nullptr;
    flag = dmalloc<int32_t>(nnz);
    segid = dmalloc<int32_t>(nnz);
    hipLaunchKernelGGL(k_seg_flags<TI>, grid_of(nnz), dim3(256), 0, s, key_sorted,
                       prow, nnz, flag);
    DRX_HIP(hipGetLastError());
    tb = 0;
    DRX_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb, flag, segid,
                                             static_cast<int>(nnz), s));
    DRX_HIP(hipMalloc(&temp, std::max<size_t>(tb, 1)));
    DRX_HIP(hipcub::DeviceScan::InclusiveSum(temp, tb, flag, segid,
                                             static_cast<int>(nnz), s));
    int32_t nseg32 = 0;
    DRX_HIP(hipMemcpyAsync(&nseg32, segid + nnz - 1, sizeof(int32_t),
                           hipMemcpyDeviceToHost, s));
    DRX_HIP(hipStreamSynchronize(s));
    (void)hipFree(temp);
    temp = nullptr;
    ps.nseg = nseg32;
    ps.seg_row = dmalloc<TI>(ps.nseg);
    ps.seg_ptr = dmalloc<TI>(ps.nseg + 1);
    hipLaunchKernelGGL(k_seg_fill<TI>, grid_of(nnz), dim3(256), 0, s, flag, segid,
                       prow, nnz, static_cast<TI *>(ps.seg_row),
                       static_cast<TI *>(ps.seg_ptr));
    const TI end = static_cast<TI>(nnz);
    DRX_HIP(hipMemcpyAsync(static_cast<TI *>(ps.seg_ptr) + ps.nseg, &end, sizeof(TI),
                           hipMemcpyHostToDevice, s));
    {
      // first segment of every row in panel order, and the rows without any
      rowfirst = dmalloc<uint32_t>(p.nrows);
      cnt = dmalloc<unsigned long long>(1);
      DRX_HIP(hipMemsetAsync(rowfirst, 0xff, static_cast<size_t>(std::max<int64_t>(p.nrows, 1)) * 4, s));
      DRX_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), s));
      hipLaunchKernelGGL(k_row_first_panel<TI>, grid_of(nnz), dim3(256), 0, s, flag, key_sorted,
                         prow, nnz, rowfirst);
      hipLaunchKernelGGL(k_mark_first<TI>, grid_of(nnz), dim3(256), 0, s, flag, segid,
                         key_sorted, prow, nnz, rowfirst, static_cast<TI *>(ps.seg_row));
      hipLaunchKernelGGL(k_empty_rows<TI>, grid_of(p.nrows), dim3(256), 0, s, rowfirst, p.nrows,
                         cnt, static_cast<TI *>(nullptr));
      DRX_HIP(hipGetLastError());
      unsigned long long ne = 0;
      DRX_HIP(hipMemcpyAsync(&ne, cnt, sizeof(ne), hipMemcpyDeviceToHost, s));
      DRX_HIP(hipStreamSynchronize(s));
      ps.nempty = static_cast<int64_t>(ne);
      if (ps.nempty > 0) {
        ps.empty_rows = dmalloc<TI>(ps.nempty);
        DRX_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), s));
        hipLaunchKernelGGL(k_empty_rows<TI>, grid_of(p.nrows), dim3(256), 0, s, rowfirst,
                           p.nrows, cnt, static_cast<TI *>(ps.empty_rows));
        DRX_HIP(hipGetLastError());
      }
    }
    d_pstart = dmalloc<int64_t>(ps.npanels + 1);
    d_pseg = dmalloc<int64_t>(ps.npanels + 1);
    hipLaunchKernelGGL(k_panel_bounds, grid_of(ps.npanels + 1), dim3(256), 0, s,
                       key_sorted, segid, nnz, ps.nseg, ps.npanels, d_pstart, d_pseg);
    DRX_HIP(hipGetLastError());
    ps.pstart.resize(ps.npanels + 1);
    ps.pseg.resize(ps.npanels + 1);
    DRX_HIP(hipMemcpyAsync(ps.pstart.data(), d_pstart, (ps.npanels + 1) * sizeof(int64_t),
                           hipMemcpyDeviceToHost, s));
    DRX_HIP(hipMemcpyAsync(ps.pseg.data(), d_pseg, (ps.npanels + 1) * sizeof(int64_t),
                           hipMemcpyDeviceToHost, s));
    DRX_HIP(hipStreamSynchronize(s));
    cleanup();
    rowid = iota = prow = nullptr;
    key = key_sorted = nullptr;
    flag = segid = nullptr;
    d_pstart = d_pseg = nullptr;
    rowfirst = nullptr;
    cnt = nullptr;
    // One plan per panel over its segments (absolute entry offsets, sentinel
    // = next panel's first entry). The plan's base is the panel's first
    // entry rounded down to a multiple of 4, so that the panel's arrays start
    // aligned to a group of entries (2 or 4 entries per lane per load); a
    // panel starting elsewhere then has up to 3 leading entries that belong
    // to none of its rows.
    const TI *segp = static_cast<const TI *>(ps.seg_ptr);
    // A panel without entries gets no plan (the product skips it): its
    // rounded-down base would give it stray entries and no rows.
    for (int64_t q = 0; q < ps.npanels; q++) {
      const int64_t rows = ps.pseg[q + 1] - ps.pseg[q];
      if (ps.pstart[q + 1] == ps.pstart[q]) {
        ps.plans.push_back(nullptr);
        continue;
      }
      const int64_t base = ps.pstart[q] & ~int64_t(3);
      const int64_t pn = ps.pstart[q + 1] - base;
      ps.plans.push_back(make_plan(p.itype, segp + ps.pseg[q], base, rows, p.ncols, pn, 1,
                                   p.block_nnz, s, false, false, true));
    }
    DRX_HIP(hipStreamSynchronize(s));
  } catch (...) {
    cleanup();
    throw;
  }
}

// y[r] = beta * y[r] (0 when beta is 0) for the listed rows.
template <typename TY, typename TI>
__global__ void k_scale_rows(const TI *rows, int64_t n, TY *y, int64_t ldy, int64_t k,
                             double beta) {
  DRX_GRID_LOOP(i, n * k) {
    const int64_t r = static_cast<int64_t>(rows[i % n]), j = i / n;
    TY &v = y[r + j * ldy];
    if (beta == 0.0)
      v = TY(0);
    else
      v = static_cast<TY>(static_cast<double>(v) * beta);
  }
}

template <typename TY>
__global__ void k_scale_y(TY *y, int64_t nrows, int64_t ldy, int64_t k,
                          double beta) {
  DRX_GRID_LOOP(i, nrows * k) {
    const int64_t r = i % nrows, j = i / nrows;
    TY &v = y[r + j * ldy];
    if (beta == 0.0)
      v = TY(0);
    else
      v = static_cast<TY>(static_cast<double>(v) * beta);
  }
}

void check_panel_operands(const Plan &p, int vtype, const void *colind,
                          const void *vals) {
  const PanelSet &ps = *p.panels;
  DRX_REQUIRE(vtype == ps.vtype, "vtype differs from the panel layout's");
  DRX_REQUIRE(vals == ps.src_vals && colind == ps.src_colind,
              "the panel layout holds a copy of the colind/vals it was built "
              "from: pass those pointers (drx_plan_update_values re-packs new values)");
}

void run_spmv_panels(const Plan &p, int vtype, const void *colind,
                     const void *vals, int xtype, const void *x, int64_t ldx,
                     int ytype, void *y, int64_t ldy, int64_t k, double alpha,
                     double beta, hipStream_t s) {
  const PanelSet &ps = *p.panels;
  check_panel_operands(p, vtype, colind, vals);
  DRX_REQUIRE(k >= 1 && y != nullptr, "bad output");
  DRX_REQUIRE(k == 1 || (ldx >= p.ncols && ldy >= p.nrows), "ldx/ldy too small");
  DRX_REQUIRE(static_cast<uint64_t>(p.ncols) * dtype_size(xtype) < (uint64_t(1) << 32),
              "the panel layout needs ncols * sizeof(x) < 4 GiB");
  if (p.nrows == 0) return;
  DRX_REQUIRE(ytype == DRX_F64 || ytype == DRX_F32, "unsupported ytype");
  // Rows with entries are written by their first segment (beta * y +
  // alpha * sum); only rows without entries are scaled apart.
  if (beta != 1.0 && p.nnz == 0) {
    if (ytype == DRX_F64)
      hipLaunchKernelGGL(k_scale_y<double>, grid_of(p.nrows * k), dim3(256), 0, s,
                         static_cast<double *>(y), p.nrows, ldy, k, beta);
    else
      hipLaunchKernelGGL(k_scale_y<float>, grid_of(p.nrows * k), dim3(256), 0, s,
                         static_cast<float *>(y), p.nrows, ldy, k, beta);
    DRX_HIP(hipGetLastError());
  } else if (beta != 1.0 && ps.nempty > 0) {
#define DRX_SR(TY, TI)                                                         \
  hipLaunchKernelGGL((k_scale_rows<TY, TI>), grid_of(ps.nempty * k), dim3(256), 0, s, \
                     static_cast<const TI *>(ps.empty_rows), ps.nempty,        \
                     static_cast<TY *>(y), ldy, k, beta)
    if (ytype == DRX_F64) {
      if (p.itype == DRX_I32) DRX_SR(double, int32_t); else DRX_SR(double, int64_t);
    } else {
      if (p.itype == DRX_I32) DRX_SR(float, int32_t); else DRX_SR(float, int64_t);
    }
#undef DRX_SR
    DRX_HIP(hipGetLastError());
  }
  const size_t isz = p.itype == DRX_I32 ? 4 : 8, vsz = dtype_size(vtype);
  for (int64_t q = 0; q < ps.npanels; q++) {
    if (ps.pstart[q + 1] == ps.pstart[q]) continue;
    const int64_t base = ps.pstart[q] & ~int64_t(3);  // the panel plan's base
    run_spmv_rowmap(*ps.plans[q], vtype,
                    static_cast<const char *>(ps.pcol) + base * isz,
                    static_cast<const char *>(ps.pval) + base * vsz,
                    static_cast<const char *>(ps.seg_row) + ps.pseg[q] * isz, xtype,
                    x, ldx, ytype, y, ldy, k, alpha, beta, s);
  }
}

// Builds the panel layout of p (drx_plan_set_panels).
static void set_panels(Plan &p, int vtype, const void *colind, const void *vals,
                     int64_t panel_cols, hipStream_t s) {
  DRX_REQUIRE(p.panels == nullptr, "the plan already has a panel layout");
  DRX_REQUIRE(p.nnz == 0 || (colind && vals), "colind/vals is null");
  DRX_REQUIRE(p.nnz <= INT32_MAX, "the panel layout is limited to 2^31-1 entries");
  DRX_REQUIRE(vtype >= DRX_F64 && vtype <= DRX_F16, "bad vtype");
  PanelSet *ps = new PanelSet();
  ps->vtype = vtype;
  ps->src_colind = colind;
  ps->src_vals = vals;
  ps->width = panel_cols > 0 ? panel_cols
                             : std::max<int64_t>(4096, (int64_t(2) << 20) /
                                                           static_cast<int64_t>(dtype_size(vtype)));
  ps->npanels = std::max<int64_t>(1, (p.ncols + ps->width - 1) / ps->width);
  DRX_REQUIRE(ps->npanels <= (int64_t(1) << 20), "too many panels");
  try {
    if (p.nnz > 0) {
#define DRX_BP(TI, TV)                                                         \
  build_panels<TI, TV>(p, *ps, static_cast<const TI *>(colind),                \
                       static_cast<const TV *>(vals), s)
      const bool i32 = p.itype == DRX_I32;
      switch (vtype) {
        case DRX_F64: if (i32) DRX_BP(int32_t, double); else DRX_BP(int64_t, double); break;
        case DRX_F32: if (i32) DRX_BP(int32_t, float); else DRX_BP(int64_t, float); break;
        default: if (i32) DRX_BP(int32_t, uint16_t); else DRX_BP(int64_t, uint16_t); break;
      }
#undef DRX_BP
    } else {
      ps->npanels = 0;
    }
  } catch (...) {
    free_panels(ps);
    throw;
  }
  p.panels = ps;
}

// Column profile of a share for the automatic panel choice: up to 2048 row
// blocks spread over the share, one wave each; a wave compares up to 256
// entries of its block's window with the window's middle column and counts
// those within `near` columns of it. out = {near, sampled}.
template <typename TI>
__global__ void k_col_profile(const BlockRec *rec, int64_t nblocks, const TI *colind,
                              int64_t near, unsigned long long *out) {
  const int64_t S = nblocks < 2048 ? nblocks : 2048;
  const int lane = threadIdx.x & 63;
  unsigned long long n_near = 0, n_seen = 0;
  for (int64_t i = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / 64; i < S;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x / 64) {
    const BlockRec b = rec[i * nblocks / S];
    const int64_t n = b.es - b.e0;
    if (n < 2) continue;
    const int64_t mid = static_cast<int64_t>(colind[b.e0 + n / 2]);
    const int64_t step = n > 256 ? n / 256 : 1;
    for (int64_t j = lane; j < 256 && j * step < n; j += 64) {
      const int64_t c = static_cast<int64_t>(colind[b.e0 + j * step]);
      const int64_t d = c > mid ? c - mid : mid - c;
      n_near += d <= near ? 1 : 0;
      n_seen += 1;
    }
  }
  atomicAdd(&out[0], n_near);
  atomicAdd(&out[1], n_seen);
}

// Whether the panel layout pays for plan p (see drx_plan_auto_panels).
static bool panels_pay(const Plan &p, int vtype, const void *colind, hipStream_t s) {
  const int64_t sx = static_cast<int64_t>(dtype_size(vtype));
  if (tuned_auto_panels() == 0) return false;
  if (p.panels != nullptr || p.nnz == 0 || p.nnz > INT32_MAX) return false;
  if (p.ncols * sx <= (int64_t(4) << 20)) return false;          // x fits one XCD's L2
  if (p.nnz < 4 * p.nrows) return false;                         // y updates would dominate
  if (static_cast<uint64_t>(p.ncols) * sx >= (uint64_t(1) << 32)) return false;
  // device memory for the copy (entries, permutation, segments, build
  // scratch): about 4 index arrays and 1 value array of nnz, with a margin
  const int64_t si = p.itype == DRX_I32 ? 4 : 8;
  size_t free_b = 0, total_b = 0;
  DRX_HIP(hipMemGetInfo(&free_b, &total_b));
  if (static_cast<double>(free_b) < 1.25 * static_cast<double>(p.nnz) * (4 * si + sx + 8))
    return false;
  unsigned long long *d = static_cast<unsigned long long *>(
      scratch_alloc(2 * sizeof(unsigned long long), s));
  unsigned long long h[2] = {0, 0};
  try {
    DRX_HIP(hipMemsetAsync(d, 0, 2 * sizeof(unsigned long long), s));
    const int64_t near = (int64_t(1) << 20) / sx;  // 1 MiB of x either side
    if (p.itype == DRX_I32)
      hipLaunchKernelGGL(k_col_profile<int32_t>, dim3(32), dim3(256), 0, s, p.rec, p.nblocks,
                         static_cast<const int32_t *>(colind), near, d);
    else
      hipLaunchKernelGGL(k_col_profile<int64_t>, dim3(32), dim3(256), 0, s, p.rec, p.nblocks,
                         static_cast<const int64_t *>(colind), near, d);
    DRX_HIP(hipGetLastError());
    DRX_HIP(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s));
    DRX_HIP(hipStreamSynchronize(s));
  } catch (...) {
    scratch_free(d, s);
    throw;
  }
  scratch_free(d, s);
  // scattered: fewer than half of the sampled entries near their block's
  // middle column
  return h[1] > 0 && 2 * h[0] < h[1];
}

}  // namespace drx

using namespace drx;

extern "C" {

int drx_plan_set_panels(drx_plan plan, int vtype, const void *colind,
                        const void *vals, int64_t panel_cols,
                        drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(plan != nullptr, "plan is null");
    set_panels(*reinterpret_cast<Plan *>(plan), vtype, colind, vals, panel_cols,
               static_cast<hipStream_t>(stream));
  });
}

int drx_plan_auto_panels(drx_plan plan, int vtype, const void *colind,
                         const void *vals, drx_stream_t stream, int *built) {
  return guarded([&] {
    if (built) *built = 0;
    DRX_REQUIRE(plan != nullptr, "plan is null");
    Plan &p = *reinterpret_cast<Plan *>(plan);
    DRX_REQUIRE(vtype >= DRX_F64 && vtype <= DRX_F16, "bad vtype");
    DRX_REQUIRE(p.nnz == 0 || (colind && vals), "colind/vals is null");
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    DRX_HIP(hipStreamIsCapturing(s, &cs));
    DRX_REQUIRE(cs == hipStreamCaptureStatusNone, "the stream is being captured");
    if (!panels_pay(p, vtype, colind, s)) return;
    set_panels(p, vtype, colind, vals, 0, s);
    if (built) *built = 1;
  });
}

int drx_plan_update_values(drx_plan plan, const void *vals, drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(plan != nullptr, "plan is null");
    Plan &p = *reinterpret_cast<Plan *>(plan);
    DRX_REQUIRE(p.panels != nullptr, "the plan has no panel layout");
    DRX_REQUIRE(p.nnz == 0 || vals != nullptr, "vals is null");
    PanelSet &ps = *p.panels;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (p.nnz > 0) {
      const bool i32 = p.itype == DRX_I32;
#define DRX_RP(TI, TV)                                                         \
  hipLaunchKernelGGL((k_repack<TI, TV>), grid_of(p.nnz), dim3(256), 0, s,      \
                     static_cast<const TI *>(ps.perm),                         \
                     static_cast<const TV *>(vals), p.nnz,                     \
                     static_cast<TV *>(ps.pval))
      switch (ps.vtype) {
        case DRX_F64: if (i32) DRX_RP(int32_t, double); else DRX_RP(int64_t, double); break;
        case DRX_F32: if (i32) DRX_RP(int32_t, float); else DRX_RP(int64_t, float); break;
        default: if (i32) DRX_RP(int32_t, uint16_t); else DRX_RP(int64_t, uint16_t); break;
      }
#undef DRX_RP
      DRX_HIP(hipGetLastError());
    }
    ps.src_vals = vals;
  });
}

int drx_plan_panels(drx_plan plan, int64_t *npanels, int64_t *panel_cols,
                    int64_t *segments) {
  return guarded([&] {
    DRX_REQUIRE(plan != nullptr, "plan is null");
    const Plan &p = *reinterpret_cast<Plan *>(plan);
    const PanelSet *ps = p.panels;
    if (npanels) *npanels = ps ? ps->npanels : 0;
    if (panel_cols) *panel_cols = ps ? ps->width : 0;
    if (segments) *segments = ps ? ps->nseg : 0;
  });
}

}  // extern "C"
