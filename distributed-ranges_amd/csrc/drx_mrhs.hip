// drx_mrhs.hip -- products with several right-hand sides (k >= 2), and the
// library's stream-ordered scratch pool.
//
// x (column-major, k columns of ldx) is copied into row-major chunks of KP
// columns (KP = 2, 4, 8 or 16 by k, at most 8 for FP64), then the multi-RHS block body of
// drx_fwd.hpp gathers one contiguous KP-vector per entry. With column-major
// x each entry costs k gathers from k separate columns; the copy costs one
// streaming pass over x. A plan's column panels serve k = 1 only: for k >= 2
// the product runs over the share's own CSR arrays.
//
// Reference kernels replaced: the k-column loops of
// csr_row_distribution.hpp:84-115 and csr_eq_distribution.hpp:62-111 (slim
// matrix right-hand sides, broadcasted_slim_matrix.hpp:10-83).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "drx_common.hpp"
#include "drx_fwd.hpp"
#include "drx_internal.hpp"

namespace drx {

// ---------------------------------------------------------------------------
// Scratch: one memory pool per device, created on first use, that keeps its
// memory across calls (release threshold = max), so that steady-state calls
// reuse it without reaching the driver's allocator. Allocation and release
// are ordered on the caller's stream.
// ---------------------------------------------------------------------------
constexpr int kMaxDevices = 64;
static std::mutex g_pool_mu;
static hipMemPool_t g_pools[kMaxDevices];

static hipMemPool_t pool_for(int dev) {
  std::lock_guard<std::mutex> lk(g_pool_mu);
  DRX_REQUIRE(dev >= 0 && dev < kMaxDevices, "device ordinal out of range");
  if (!g_pools[dev]) {
    hipMemPoolProps props{};
    props.allocType = hipMemAllocationTypePinned;
    props.handleTypes = hipMemHandleTypeNone;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = dev;
    hipMemPool_t pool = nullptr;
    DRX_HIP(hipMemPoolCreate(&pool, &props));
    uint64_t keep = UINT64_MAX;
    DRX_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
    g_pools[dev] = pool;
  }
  return g_pools[dev];
}

void *scratch_alloc(size_t bytes, hipStream_t s) {
  int dev = 0;
  DRX_HIP(hipGetDevice(&dev));
  void *p = nullptr;
  DRX_HIP(hipMallocFromPoolAsync(&p, std::max<size_t>(bytes, 256), pool_for(dev), s));
  return p;
}

void scratch_free(void *p, hipStream_t s) {
  if (p) (void)hipFreeAsync(p, s);
}

// Samples up to 4096 row blocks: the span between the first and the last
// column index of each window, against its entry count (tight: at most
// 4 n + 64) and against ncols / 16 (loose). out = {tight, loose, sampled}.
template <typename TI>
__global__ void k_col_local(const BlockRec *rec, int64_t nblocks, int64_t ncols,
                            const TI *colind, int *out) {
  __shared__ int cnt[3];
  if (threadIdx.x < 3) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t S = nblocks < 4096 ? nblocks : 4096;
  int tight = 0, loose = 0, seen = 0;
  for (int64_t i = threadIdx.x; i < S; i += blockDim.x) {
    const BlockRec b = rec[i * nblocks / S];
    const int64_t n = b.es - b.e0;
    if (n >= 2) {
      const int64_t lo = static_cast<int64_t>(colind[b.e0]);
      const int64_t hi = static_cast<int64_t>(colind[b.es - 1]);
      const int64_t span = hi > lo ? hi - lo : lo - hi;
      seen++;
      if (span <= 4 * n + 64) tight++;
      if (span <= 4 * n + 64 || span <= ncols / 16) loose++;
    }
  }
  atomicAdd(&cnt[0], tight);
  atomicAdd(&cnt[1], loose);
  atomicAdd(&cnt[2], seen);
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = cnt[0];
    out[1] = cnt[1];
    out[2] = cnt[2];
  }
}

int column_locality(const Plan &p, const void *colind, hipStream_t s) {
  if (p.nnz == 0 || p.nblocks == 0 || colind == nullptr) return 0;
  {
    const int cached = p.loc_local.load(std::memory_order_acquire);
    if (cached >= 0 && p.loc_colind.load(std::memory_order_acquire) == colind) return cached;
  }
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  DRX_HIP(hipStreamIsCapturing(s, &st));
  if (st != hipStreamCaptureStatusNone) return 0;
  int *d = static_cast<int *>(scratch_alloc(3 * sizeof(int), s));
  int h[3] = {0, 0, 0};
  try {
    if (p.itype == DRX_I32)
      hipLaunchKernelGGL(k_col_local<int32_t>, dim3(1), dim3(256), 0, s, p.rec, p.nblocks,
                         p.ncols, static_cast<const int32_t *>(colind), d);
    else
      hipLaunchKernelGGL(k_col_local<int64_t>, dim3(1), dim3(256), 0, s, p.rec, p.nblocks,
                         p.ncols, static_cast<const int64_t *>(colind), d);
    DRX_HIP(hipGetLastError());
    DRX_HIP(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s));
    DRX_HIP(hipStreamSynchronize(s));
  } catch (...) {
    scratch_free(d, s);
    throw;
  }
  scratch_free(d, s);
  int loc = 0;
  if (h[2] > 0 && 10 * h[0] >= 9 * h[2]) loc |= 1;
  if (h[2] > 0 && 10 * h[1] >= 9 * h[2]) loc |= 2;
  // Invalidate, then publish pointer and value.
  p.loc_local.store(-1, std::memory_order_release);
  p.loc_colind.store(colind, std::memory_order_release);
  p.loc_local.store(loc, std::memory_order_release);
  return loc;
}

int mrhs_kp(int64_t k, int xtype, bool loose_local) {
  if (k <= 2) return 2;
  if (k <= 4) return 4;
  if (k <= 8 || dtype_size(xtype) == 8) return 8;
  const int cap = tuned_mrhs_kp_max();
  return (cap == 8 || (cap == 0 && loose_local)) ? 8 : 16;
}

size_t mrhs_xr_bytes(int64_t ncols, int64_t k, int kp, int xtype) {
  const int64_t nchunk = (k + kp - 1) / kp;
  return static_cast<size_t>(nchunk) * static_cast<size_t>(std::max<int64_t>(ncols, 1)) *
         kp * dtype_size(xtype);
}

// x -> xr on stream s (only the element size matters: the zero padding is
// all-zero bits, +0 in every floating type).
void launch_x_rowmajor(int xtype, const void *x, int64_t ldx, int64_t ncols,
                       int64_t k, int kp, void *xr, hipStream_t s) {
  const int64_t total = ((k + kp - 1) / kp) * ncols;
  if (total == 0) return;
  const dim3 g(static_cast<unsigned>(std::min<int64_t>((total + 255) / 256, 1 << 16)));
#define DRX_XR(KP, T)                                                          \
  hipLaunchKernelGGL((k_x_rowmajor<KP, T>), g, dim3(256), 0, s,                \
                     static_cast<const T *>(x), ldx, ncols, k, static_cast<T *>(xr))
#define DRX_XR_T(T)                                                            \
  do {                                                                         \
    if (kp == 2)                                                               \
      DRX_XR(2, T);                                                            \
    else if (kp == 4)                                                          \
      DRX_XR(4, T);                                                            \
    else if (kp == 8)                                                          \
      DRX_XR(8, T);                                                            \
    else                                                                       \
      DRX_XR(16, T);                                                           \
  } while (0)
  const size_t xs = dtype_size(xtype);
  if (xs == 8)
    DRX_XR_T(uint64_t);
  else if (xs == 4)
    DRX_XR_T(uint32_t);
  else
    DRX_XR_T(uint16_t);
#undef DRX_XR_T
#undef DRX_XR
  DRX_HIP(hipGetLastError());
}

template <int KP, int MODE, typename TV, typename TX, typename TY, typename TA,
          typename TI>
static void launch_mrhs(const Plan &p, const FwdArgs &a, hipStream_t s) {
  const dim3 grid(static_cast<unsigned>(a.grid));
  if (p.cap <= 2048)
    hipLaunchKernelGGL((k_spmv_mrhs<256, 2048, KP, MODE, TV, TX, TY, TA, TI>),
                       grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((k_spmv_mrhs<512, 8192, KP, MODE, TV, TX, TY, TA, TI>),
                       grid, dim3(512), 0, s, a);
}

template <int KP, int MODE, typename TV, typename TX, typename TY, typename TA>
static void launch_mrhs_i(const Plan &p, const FwdArgs &a, hipStream_t s) {
  if (p.itype == DRX_I32)
    launch_mrhs<KP, MODE, TV, TX, TY, TA, int32_t>(p, a, s);
  else
    launch_mrhs<KP, MODE, TV, TX, TY, TA, int64_t>(p, a, s);
}

template <typename TV, typename TX, typename TY, typename TA>
static void launch_mrhs_t(const Plan &p, const FwdArgs &a, int mode, int kp,
                          hipStream_t s) {
#define DRX_KP(MODE)                                                           \
  do {                                                                         \
    if (kp == 2)                                                               \
      launch_mrhs_i<2, MODE, TV, TX, TY, TA>(p, a, s);                         \
    else if (kp == 4)                                                          \
      launch_mrhs_i<4, MODE, TV, TX, TY, TA>(p, a, s);                         \
    else if (kp == 8)                                                          \
      launch_mrhs_i<8, MODE, TV, TX, TY, TA>(p, a, s);                         \
    else                                                                       \
      launch_mrhs_i<16, MODE, TV, TX, TY, TA>(p, a, s);                        \
  } while (0)
  if (mode == kPerm)
    DRX_KP(kPerm);
  else
    DRX_KP(kPlain);
#undef DRX_KP
}

void run_mrhs(const Plan &p, FwdArgs a, int mode, int vtype, int xtype,
              int ytype, hipStream_t s) {
  DRX_REQUIRE(mode == kPlain || mode == kPerm, "bad mode for k >= 2");
  // one chunk width for the copy and the kernel (the knob may change);
  // 8-wide chunks for k > 8 when the columns are loosely local (measured on
  // the 3-D Laplacian 128^3 FP32, k = 16: 2293 vs 1982 GB/s, while a matrix
  // of random columns runs half as fast with them: 323 vs 618 GB/s)
  const bool loose = a.k > 8 && dtype_size(xtype) == 4 && tuned_mrhs_kp_max() == 0 &&
                     (column_locality(p, a.colind, s) & 2) != 0;
  const int kp = mrhs_kp(a.k, xtype, loose);
  void *xr = scratch_alloc(mrhs_xr_bytes(p.ncols, a.k, kp, xtype), s);
  try {
    launch_x_rowmajor(xtype, a.x, a.ldx, p.ncols, a.k, kp, xr, s);
    a.xr = xr;
    a.ncols = p.ncols;
    a.gshift = tuned_mrhs_gshift();
    if (vtype == DRX_F64 && xtype == DRX_F64 && ytype == DRX_F64)
      launch_mrhs_t<double, double, double, double>(p, a, mode, kp, s);
    else if (vtype == DRX_F32 && xtype == DRX_F32 && ytype == DRX_F32)
      launch_mrhs_t<float, float, float, float>(p, a, mode, kp, s);
    else if (vtype == DRX_BF16 && xtype == DRX_BF16 && ytype == DRX_F32)
      launch_mrhs_t<bf16, bf16, float, float>(p, a, mode, kp, s);
    else if (vtype == DRX_F16 && xtype == DRX_F16 && ytype == DRX_F32)
      launch_mrhs_t<f16, f16, float, float>(p, a, mode, kp, s);
    else if (vtype == DRX_BF16 && xtype == DRX_F32 && ytype == DRX_F32)
      launch_mrhs_t<bf16, float, float, float>(p, a, mode, kp, s);
    else if (vtype == DRX_F16 && xtype == DRX_F32 && ytype == DRX_F32)
      launch_mrhs_t<f16, float, float, float>(p, a, mode, kp, s);
    else
      throw Error(DRX_ERR_UNSUPPORTED, "unsupported (vtype, xtype, ytype)");
    DRX_HIP(hipGetLastError());
  } catch (...) {
    scratch_free(xr, s);
    throw;
  }
  scratch_free(xr, s);
}

}  // namespace drx
