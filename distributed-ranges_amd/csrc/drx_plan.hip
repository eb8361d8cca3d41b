// drx_plan.hip -- the plan of one CSR share (row blocks and their entry
// windows), library status strings, and the plan entry points.
//
// Block c holds the rows whose first entry lies in [c*P, (c+1)*P) of the
// share (P = the pitch, block_pitch), found by one binary search on rowptr
// per block boundary. Its record also stores where the LDS window ends (es):
// all of the block's entries when they fit the window less its slack
// (kWinSlack), otherwise everything before the block's last row, which is
// then reduced by the whole workgroup. Computing this once lets the kernels
// start streaming vals and colind as soon as they read their 32-byte record.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "drx_common.hpp"
#include "drx_internal.hpp"

namespace drx {

static thread_local std::string g_last_error;
void set_last_error(const char *msg) { g_last_error = msg ? msg : ""; }
const char *last_error() { return g_last_error.c_str(); }

template <typename TI>
__device__ int64_t first_row_at(const TI *rowptr, int64_t base, int64_t nrows,
                                int64_t nnz, int sent, int64_t target) {
  int64_t lo = 0, hi = nrows;  // first r with row_begin(r) >= target
  while (lo < hi) {
    int64_t mid = lo + (hi - lo) / 2;
    if (row_begin(rowptr, base, nrows, nnz, sent, mid) >= target)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

template <typename TI>
__global__ void k_plan_records(const TI *rowptr, int64_t base, int64_t nrows,
                               int64_t nnz, int sent, int64_t nb_nnz,
                               int64_t cap, int64_t nblocks, BlockRec *rec) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c > nblocks) return;
  auto r_of = [&](int64_t cc) -> int64_t {
    if (cc == 0) return 0;
    if (cc >= nblocks) return nrows;
    return first_row_at(rowptr, base, nrows, nnz, sent, cc * nb_nnz);
  };
  const int64_t r0 = r_of(c);
  const int64_t e0 = row_begin(rowptr, base, nrows, nnz, sent, r0);
  BlockRec out;
  out.r0 = r0;
  out.e0 = e0;
  out.es = e0;
  out.pad = 0;
  if (c < nblocks) {
    const int64_t r1 = r_of(c + 1);
    const int64_t e1 = row_begin(rowptr, base, nrows, nnz, sent, r1);
    if (r1 > r0 && e1 - e0 > cap - kWinSlack)
      out.es = row_begin(rowptr, base, nrows, nnz, sent, r1 - 1);
    else
      out.es = e1;
  }
  rec[c] = out;
}

// Longest row of the share (clamped to it): one partial maximum per
// workgroup, then one atomic.
template <typename TI>
__global__ void k_max_rowlen(const TI *rowptr, int64_t base, int64_t nrows,
                             int64_t nnz, int sent, unsigned long long *out) {
  __shared__ unsigned long long red[4];
  unsigned long long m = 0;
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < nrows;
       r += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t len = row_begin(rowptr, base, nrows, nnz, sent, r + 1) -
                        row_begin(rowptr, base, nrows, nnz, sent, r);
    m = len > static_cast<int64_t>(m) ? static_cast<unsigned long long>(len) : m;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long v = __shfl_down(m, o, 64);
    m = v > m ? v : m;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < static_cast<int>(blockDim.x >> 6); i++) m = red[i] > m ? red[i] : m;
    atomicMax(out, m);
  }
}

// Entries between the starts of consecutive blocks of a window of `cap`
// entries. Block c holds the rows whose first entry lies in
// [c * pitch, (c + 1) * pitch), so it spans at most pitch + lmax - 1
// entries: with the longest row known and at most cap / 2, the pitch fills
// the window up to that row and the window's slack (cap - kWinSlack -
// lmax); otherwise (unknown, long rows, or fewer than 2.5 entries per row,
// whose blocks would hold more rows than the LDS row table) it is cap / 2
// and a last row that does not fit is reduced by the whole workgroup. Knob
// "fill" 0 keeps cap / 2.
int64_t block_pitch(const Plan &p, int64_t cap) {
  if (tuned_fill() == 0 || p.lmax < 0 || p.lmax > cap / 2 || 2 * p.nnz < 5 * p.nrows)
    return cap / 2;
  return cap - kWinSlack - std::max<int64_t>(p.lmax, 1);
}

int64_t default_block_nnz(int itype) {
  int64_t v = tuned_block_nnz();
  if (v > 0) return v;
  (void)itype;
  return 1024;
}

// Records of blocks starting every `nb` entries with windows of `cap`
// entries into a new array, or into `into` (caller-owned, at least
// nblocks + 1 records) when it is given.
static BlockRec *make_records(const Plan &p, int64_t nb, int64_t cap, int64_t *nblocks,
                              hipStream_t s, bool async_alloc, BlockRec *into = nullptr) {
  *nblocks = p.nnz / nb + 1;
  const size_t bytes = static_cast<size_t>(*nblocks + 1) * sizeof(BlockRec);
  BlockRec *rec = into;
  if (rec != nullptr) {
  } else if (async_alloc)
    DRX_HIP(hipMallocAsync(reinterpret_cast<void **>(&rec), bytes, s));
  else
    DRX_HIP(hipMalloc(reinterpret_cast<void **>(&rec), bytes));
  const int tpb = 256;
  const int64_t n = *nblocks + 1;
  const dim3 grid(static_cast<unsigned>((n + tpb - 1) / tpb));
  if (p.itype == DRX_I32)
    hipLaunchKernelGGL(k_plan_records<int32_t>, grid, dim3(tpb), 0, s,
                       static_cast<const int32_t *>(p.rowptr), p.base, p.nrows,
                       p.nnz, p.has_sentinel, nb, cap, *nblocks, rec);
  else
    hipLaunchKernelGGL(k_plan_records<int64_t>, grid, dim3(tpb), 0, s,
                       static_cast<const int64_t *>(p.rowptr), p.base, p.nrows,
                       p.nnz, p.has_sentinel, nb, cap, *nblocks, rec);
  DRX_HIP(hipGetLastError());
  return rec;
}

// The plan's records, and the records of twice as large blocks (window
// 4096) that the forward kernel uses for 4-byte values when the plan's
// block is 1024 entries.
// The plan's records (windows of 2 block_nnz entries), and for the default
// 1024-entry blocks the records of windows of 4096 that the forward kernel
// uses for 4-byte values.
static void build_records(Plan &p, hipStream_t s, bool async_alloc) {
  p.cap = 2 * p.block_nnz;
  p.owns_async = async_alloc;
  p.rec = make_records(p, block_pitch(p, p.cap), p.cap, &p.nblocks, s, async_alloc);
  if (p.block_nnz == 1024)
    p.rec_wide = make_records(p, block_pitch(p, 4096), 4096, &p.nblocks_wide, s, async_alloc);
}

// Longest row of p's share (waits for the stream).
static int64_t max_row_len(const Plan &p, hipStream_t s) {
  if (p.nrows == 0) return 0;
  unsigned long long *d = static_cast<unsigned long long *>(
      scratch_alloc(sizeof(unsigned long long), s));
  unsigned long long h = 0;
  try {
    DRX_HIP(hipMemsetAsync(d, 0, sizeof(unsigned long long), s));
    const int64_t nb = std::min<int64_t>((p.nrows + 255) / 256, 4096);
    if (p.itype == DRX_I32)
      hipLaunchKernelGGL(k_max_rowlen<int32_t>, dim3(static_cast<unsigned>(nb)), dim3(256), 0, s,
                         static_cast<const int32_t *>(p.rowptr), p.base, p.nrows, p.nnz,
                         p.has_sentinel, d);
    else
      hipLaunchKernelGGL(k_max_rowlen<int64_t>, dim3(static_cast<unsigned>(nb)), dim3(256), 0, s,
                         static_cast<const int64_t *>(p.rowptr), p.base, p.nrows, p.nnz,
                         p.has_sentinel, d);
    DRX_HIP(hipGetLastError());
    DRX_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, s));
    DRX_HIP(hipStreamSynchronize(s));
  } catch (...) {
    scratch_free(d, s);
    throw;
  }
  scratch_free(d, s);
  return static_cast<int64_t>(h);
}

int64_t records_needed(int64_t nnz, int64_t block_nnz) {
  int64_t n = nnz / block_nnz + 2;
  if (block_nnz == 1024) n += nnz / 2048 + 2;
  return n;
}

static void validate_plan_args(int itype, const void *rowptr, int64_t nrows,
                               int64_t ncols, int64_t nnz,
                               int64_t block_nnz) {
  DRX_REQUIRE(itype == DRX_I32 || itype == DRX_I64, "itype must be I32 or I64");
  DRX_REQUIRE(nrows >= 0 && ncols >= 0 && nnz >= 0, "negative size");
  DRX_REQUIRE(nrows == 0 || rowptr != nullptr, "rowptr is null");
  DRX_REQUIRE(nrows > 0 || nnz == 0, "a share without rows has no entries");
  DRX_REQUIRE(block_nnz <= 0 || (block_nnz >= 64 && block_nnz <= 4096 &&
                                 (block_nnz & (block_nnz - 1)) == 0),
              "block_nnz must be a power of two in [64, 4096]");
  if (itype == DRX_I32) {
    DRX_REQUIRE(nnz <= INT32_MAX && ncols <= static_cast<int64_t>(INT32_MAX) + 1,
                "sizes exceed the 32-bit index type");
  }
}

Plan *make_plan(int itype, const void *rowptr, int64_t base, int64_t nrows,
                int64_t ncols, int64_t nnz, int sent, int64_t block_nnz,
                hipStream_t s, bool validate, bool async_alloc, bool fill) {
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
  try {
    DRX_HIP(hipGetDevice(&p->device));
    if (validate && nrows > 0) {
      // The first and last row offsets must bracket the share.
      const size_t isz = itype == DRX_I32 ? 4 : 8;
      const int64_t lastidx = sent ? nrows : nrows - 1;
      unsigned char buf[16] = {0};
      DRX_HIP(hipMemcpyAsync(buf, rowptr, isz, hipMemcpyDeviceToHost, s));
      DRX_HIP(hipMemcpyAsync(buf + 8,
                             static_cast<const char *>(rowptr) + lastidx * isz,
                             isz, hipMemcpyDeviceToHost, s));
      DRX_HIP(hipStreamSynchronize(s));
      int64_t first = 0, lastv = 0;
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
    if (fill) p->lmax = max_row_len(*p, s);
    build_records(*p, s, async_alloc);
  } catch (...) {
    free_plan(p, s);
    throw;
  }
  return p;
}

void init_plan_in(Plan &p, int itype, const void *rowptr, int64_t base, int64_t nrows,
                  int64_t ncols, int64_t nnz, int sent, BlockRec *buf,
                  int64_t buf_records, hipStream_t s) {
  validate_plan_args(itype, rowptr, nrows, ncols, nnz, 0);
  p.itype = itype;
  p.rowptr = rowptr;
  p.base = base;
  p.nrows = nrows;
  p.ncols = ncols;
  p.nnz = nnz;
  p.has_sentinel = sent ? 1 : 0;
  p.block_nnz = 1024;
  p.cap = 2 * p.block_nnz;
  p.no_scratch = true;
  DRX_REQUIRE(records_needed(nnz, p.block_nnz) <= buf_records,
              "the share needs more plan records than the buffer holds");
  DRX_HIP(hipGetDevice(&p.device));
  p.rec = make_records(p, p.block_nnz, p.cap, &p.nblocks, s, false, buf);
  p.rec_wide = make_records(p, 2048, 4096, &p.nblocks_wide, s, false, buf + p.nblocks + 1);
}

void free_plan(Plan *p, hipStream_t s) {
  if (!p) return;
  if (p->panels) free_panels(p->panels);
  for (BlockRec *r : {p->rec, p->rec_wide}) {
    if (!r) continue;
    if (p->owns_async)
      (void)hipFreeAsync(r, s);
    else
      (void)hipFree(r);
  }
  delete p;
}

}  // namespace drx

using namespace drx;

extern "C" {

int drx_version(void) { return 2; }

const char *drx_last_error(void) { return drx::last_error(); }

int drx_plan_create(drx_plan *out, int itype, const void *rowptr,
                    int64_t rowptr_base, int64_t nrows, int64_t ncols,
                    int64_t nnz, int has_sentinel, int64_t block_nnz,
                    drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(out != nullptr, "plan out-pointer is null");
    *out = nullptr;
    Plan *p = make_plan(itype, rowptr, rowptr_base, nrows, ncols, nnz,
                        has_sentinel, block_nnz,
                        static_cast<hipStream_t>(stream), true, false, true);
    *out = reinterpret_cast<drx_plan>(p);
  });
}

int drx_plan_destroy(drx_plan plan) {
  return guarded([&] {
    if (!plan) return;
    free_plan(reinterpret_cast<Plan *>(plan), nullptr);
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

}  // extern "C"
