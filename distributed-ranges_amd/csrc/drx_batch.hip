// drx_batch.hip -- many independent CSR products (forward and adjoint items)
// in one persistent launch, capturable in a hipGraph, with an autotuner.
//
// Every item is re-planned inside the batch with the batch's block size, so
// that all items are cut into row blocks of about block_nnz entries. The
// work units (item, block) are numbered item by item; the persistent grid
// (wg_per_cu workgroups per CU) splits the numbering into 8 contiguous
// ranges, one per XCD (workgroup i is dealt to XCD i % 8), and the
// workgroups of an XCD stride through its range, so neighbouring blocks of
// one matrix stay on one XCD's L2. Forward items run the plain block body,
// adjoint items the permuted-values body over their transposed structure.
// Nothing is allocated or cleared during a run, so the launch can be
// captured in a graph and replayed.
//
// With k >= 2 right-hand sides each item's x is first copied into row-major
// chunks (one copy kernel over a job table, before the persistent kernel and
// captured with it); the chunk buffers are allocated when the batch is
// created.
//
// Tunables: block_nnz (entries per block, i.e. rows per workgroup), gshift
// (the lane-group thresholds: a block of mean row length L uses G lanes per
// row chosen from L >> gshift), wg_per_cu (persistent grid size). The
// autotuner times every combination on the batch itself.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>
#include <vector>

#include "drx_common.hpp"
#include "drx_fwd.hpp"
#include "drx_internal.hpp"

namespace drx {

struct BatchItemHost {
  const Plan *src = nullptr;  // the user's plan (forward) or the transpose's
  int mode = kPlain;
  const void *colind = nullptr, *vals = nullptr, *perm = nullptr;
  const void *vals_override = nullptr;  // adjoint: the transpose's own values
  const void *x = nullptr;
  int64_t ldx = 0;
  void *y = nullptr;
  int64_t ldy = 0;
  void *xr = nullptr;  // k >= 2: row-major chunks of x
};

// One x -> xr copy of a batch with k >= 2: threads [start, start + nchunk *
// ncols) of the copy kernel.
struct XrJob {
  const void *x;
  int64_t ldx, ncols;
  void *xr;
  int64_t start;
};

struct Batch {
  std::vector<BatchItemHost> items;
  int itype = 0, vtype = 0, xtype = 0, ytype = 0;
  int64_t k = 1;
  double alpha = 1, beta = 0;
  int64_t block_nnz = 1024;
  int wg_per_cu = 4;
  int gshift = 0;
  double tuned_ms = 0;
  std::vector<Plan *> plans;
  FwdArgs *d_args = nullptr;
  int *d_mode = nullptr;
  int64_t *d_off = nullptr;
  int64_t total = 0;
  int num_cus = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int vw = 1;                   // entries per lane and load of forward items
  bool mrhs = false;            // k >= 2 through row-major chunks of x
  int kp = 0;                   // their width, fixed when the batch is made
  void *xr_all = nullptr;       // all items' chunks
  XrJob *d_jobs = nullptr;
  int64_t njobs = 0, xr_threads = 0;
};

// A record every lane of the workgroup reads at the same address, copied
// word by word through readfirstlane so that it lives in scalar registers.
__device__ __forceinline__ FwdArgs load_uniform(const FwdArgs *p) {
  static_assert(sizeof(FwdArgs) % 4 == 0, "word copy");
  FwdArgs out;
  const uint32_t *src = reinterpret_cast<const uint32_t *>(p);
  uint32_t *dst = reinterpret_cast<uint32_t *>(&out);
#pragma unroll
  for (int i = 0; i < static_cast<int>(sizeof(FwdArgs) / 4); i++)
    dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
  return out;
}

// Waves per SIMD the batch kernel is compiled for: the most that needs no
// register spills (5 for 4-byte accumulation in 256-lane workgroups, else
// the compiler's 4).
template <int WG, typename TA> constexpr int batch_waves() {
  return (sizeof(TA) == 8 || WG >= 512) ? 4 : 5;
}

// VW: consecutive entries per lane and load of the forward items (4 when
// every forward item's vals / colind are aligned to a group of four).
template <int WG, int CAP, typename TV, typename TX, typename TY, typename TA,
          typename TI, int VW>
__global__ __launch_bounds__(WG)
__attribute__((amdgpu_waves_per_eu(batch_waves<WG, TA>(), 8)))
void k_batch(const FwdArgs *items,
                                              const int *modes,
                                              const int64_t *off,
                                              int64_t nitems, int64_t total) {
  __shared__ FwdSmem<WG, CAP, false, TA, TI> sm;
  const int64_t per_xcd = (total + kXcds - 1) / kXcds;
  const int64_t xcd = blockIdx.x % kXcds;
  const int64_t lw = blockIdx.x / kXcds, nw = gridDim.x / kXcds;
  const int64_t lo = xcd * per_xcd;
  const int64_t hi = lo + per_xcd < total ? lo + per_xcd : total;
  int64_t w = lo + lw;
  if (w >= hi) return;
  int64_t it = 0, itb = nitems;  // last item with off[it] <= w
  while (it + 1 < itb) {
    const int64_t mid = (it + itb) / 2;
    if (off[mid] <= w)
      it = mid;
    else
      itb = mid;
  }
  // The item's argument record is copied once per item change: read from
  // global memory inside the block body it would be re-loaded after every
  // store to y.
  FwdArgs cur = load_uniform(items + it);
  int mode = modes[it];
  for (; w < hi; w += nw) {
    if (off[it + 1] <= w) {
      while (off[it + 1] <= w) it++;
      cur = load_uniform(items + it);
      mode = modes[it];
    }
    const int64_t blk = w - off[it];
    if (mode == kPerm)
      process_block<WG, CAP, kPerm, true, TV, TX, TY, TA, TI>(cur, blk, sm);
    else
      process_block<WG, CAP, kPlain, true, TV, TX, TY, TA, TI, VW>(cur, blk, sm);
  }
}

template <int WG, int CAP, typename TV, typename TX, typename TY, typename TA,
          typename TI>
static const void *batch_kernel(int vw) {
  if constexpr (sizeof(TI) == 4 && 4 * sizeof(TV) <= 16) {
    if (vw == 4) return reinterpret_cast<const void *>(k_batch<WG, CAP, TV, TX, TY, TA, TI, 4>);
  }
  (void)vw;
  return reinterpret_cast<const void *>(k_batch<WG, CAP, TV, TX, TY, TA, TI, 1>);
}

// k >= 2: the same walk over (item, block) with the multi-RHS block body.
template <int WG, int CAP, int KP, typename TV, typename TX, typename TY,
          typename TA, typename TI>
__global__ __launch_bounds__(WG)
__attribute__((amdgpu_waves_per_eu(mrhs_waves<WG, CAP, KP, TX, TA>(), 8)))
void k_batch_mrhs(const FwdArgs *items,
                                                   const int *modes,
                                                   const int64_t *off,
                                                   int64_t nitems, int64_t total) {
  __shared__ MrhsSmem<WG, CAP, TA> sm;
  const int64_t per_xcd = (total + kXcds - 1) / kXcds;
  const int64_t xcd = blockIdx.x % kXcds;
  const int64_t lw = blockIdx.x / kXcds, nw = gridDim.x / kXcds;
  const int64_t lo = xcd * per_xcd;
  const int64_t hi = lo + per_xcd < total ? lo + per_xcd : total;
  int64_t w = lo + lw;
  if (w >= hi) return;
  int64_t it = 0, itb = nitems;
  while (it + 1 < itb) {
    const int64_t mid = (it + itb) / 2;
    if (off[mid] <= w)
      it = mid;
    else
      itb = mid;
  }
  FwdArgs cur = load_uniform(items + it);
  int mode = modes[it];
  for (; w < hi; w += nw) {
    if (off[it + 1] <= w) {
      while (off[it + 1] <= w) it++;
      cur = load_uniform(items + it);
      mode = modes[it];
    }
    const int64_t blk = w - off[it];
    if (mode == kPerm)
      process_block_mrhs<WG, CAP, KP, kPerm, TV, TX, TY, TA, TI>(cur, blk, sm);
    else
      process_block_mrhs<WG, CAP, KP, kPlain, TV, TX, TY, TA, TI>(cur, blk, sm);
  }
}

// All items' x -> xr copies (layout of k_x_rowmajor), job by binary search.
template <int KP, typename T>
__global__ void k_batch_xr(const XrJob *jobs, int64_t njobs, int64_t total,
                           int64_t k) {
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       t < total; t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    int64_t lo = 0, hi = njobs;  // last job with start <= t
    while (lo + 1 < hi) {
      const int64_t mid = (lo + hi) / 2;
      if (jobs[mid].start <= t)
        lo = mid;
      else
        hi = mid;
    }
    const XrJob jb = jobs[lo];
    const int64_t u = t - jb.start;
    const int64_t jc = u / jb.ncols, c = u - jc * jb.ncols;
    const T *x = static_cast<const T *>(jb.x);
    T buf[KP];
#pragma unroll
    for (int jj = 0; jj < KP; jj++) {
      const int64_t j = jc * KP + jj;
      buf[jj] = j < k ? x[j * jb.ldx + c] : T(0);
    }
    constexpr int BYTES = KP * static_cast<int>(sizeof(T));
    T *dst = static_cast<T *>(jb.xr) + u * KP;
    if constexpr (BYTES >= 16) {
#pragma unroll
      for (int i = 0; i < BYTES / 16; i++) {
        uint4 v;
        __builtin_memcpy(&v, reinterpret_cast<const char *>(buf) + 16 * i, 16);
        reinterpret_cast<uint4 *>(dst)[i] = v;
      }
    } else if constexpr (BYTES == 8) {
      uint2 v;
      __builtin_memcpy(&v, buf, 8);
      *reinterpret_cast<uint2 *>(dst) = v;
    } else {
      uint32_t v;
      __builtin_memcpy(&v, buf, 4);
      *reinterpret_cast<uint32_t *>(dst) = v;
    }
  }
}

// Resolve the kernel for the batch's types and window; returns its address
// and workgroup size.
template <int KP, typename TV, typename TX, typename TY, typename TA,
          typename TI>
static void pick_mrhs_kp(const Batch &b, const void **fn, int *wg) {
  if (2 * b.block_nnz <= 2048) {
    *fn = reinterpret_cast<const void *>(k_batch_mrhs<256, 2048, KP, TV, TX, TY, TA, TI>);
    *wg = 256;
  } else {
    *fn = reinterpret_cast<const void *>(k_batch_mrhs<512, 8192, KP, TV, TX, TY, TA, TI>);
    *wg = 512;
  }
}

template <typename TV, typename TX, typename TY, typename TA, typename TI>
static void pick_kernel_t(const Batch &b, const void **fn, int *wg) {
  const int64_t cap = 2 * b.block_nnz;
  if (b.mrhs) {
    const int kp = b.kp;
    if (kp == 2)
      pick_mrhs_kp<2, TV, TX, TY, TA, TI>(b, fn, wg);
    else if (kp == 4)
      pick_mrhs_kp<4, TV, TX, TY, TA, TI>(b, fn, wg);
    else if (kp == 8)
      pick_mrhs_kp<8, TV, TX, TY, TA, TI>(b, fn, wg);
    else
      pick_mrhs_kp<16, TV, TX, TY, TA, TI>(b, fn, wg);
  } else if (cap <= 1024) {
    *fn = batch_kernel<256, 1024, TV, TX, TY, TA, TI>(b.vw);
    *wg = 256;
  } else if (cap <= 2048) {
    *fn = batch_kernel<256, 2048, TV, TX, TY, TA, TI>(b.vw);
    *wg = 256;
  } else {
    *fn = batch_kernel<512, 4096, TV, TX, TY, TA, TI>(b.vw);
    *wg = 512;
  }
}

template <typename TV, typename TX, typename TY, typename TA>
static void pick_kernel_i(const Batch &b, const void **fn, int *wg) {
  if (b.itype == DRX_I32)
    pick_kernel_t<TV, TX, TY, TA, int32_t>(b, fn, wg);
  else
    pick_kernel_t<TV, TX, TY, TA, int64_t>(b, fn, wg);
}

static void pick_kernel(const Batch &b, const void **fn, int *wg) {
  if (b.vtype == DRX_F64 && b.xtype == DRX_F64 && b.ytype == DRX_F64)
    pick_kernel_i<double, double, double, double>(b, fn, wg);
  else if (b.vtype == DRX_F32 && b.xtype == DRX_F32 && b.ytype == DRX_F32)
    pick_kernel_i<float, float, float, float>(b, fn, wg);
  else if (b.vtype == DRX_BF16 && b.xtype == DRX_BF16 && b.ytype == DRX_F32)
    pick_kernel_i<bf16, bf16, float, float>(b, fn, wg);
  else if (b.vtype == DRX_F16 && b.xtype == DRX_F16 && b.ytype == DRX_F32)
    pick_kernel_i<f16, f16, float, float>(b, fn, wg);
  else
    throw Error(DRX_ERR_UNSUPPORTED, "unsupported (vtype, xtype, ytype) for a batch");
}

static void free_chunks(Batch &b) {
  if (b.xr_all) (void)hipFree(b.xr_all);
  if (b.d_jobs) (void)hipFree(b.d_jobs);
  b.xr_all = nullptr;
  b.d_jobs = nullptr;
}

// k >= 2: one buffer for all items' row-major chunks and the copy jobs.
static void make_chunks(Batch &b, hipStream_t s) {
  const int64_t n = static_cast<int64_t>(b.items.size());
  const int kp = b.kp;
  const int64_t nchunk = (b.k + kp - 1) / kp;
  std::vector<size_t> at(n);
  size_t bytes = 0;
  for (int64_t i = 0; i < n; i++) {
    at[i] = bytes;
    bytes += (mrhs_xr_bytes(b.items[i].src->ncols, b.k, kp, b.xtype) + 255) & ~size_t(255);
  }
  DRX_HIP(hipMalloc(&b.xr_all, std::max<size_t>(bytes, 256)));
  std::vector<XrJob> jobs;
  int64_t start = 0;
  for (int64_t i = 0; i < n; i++) {
    BatchItemHost &h = b.items[i];
    h.xr = static_cast<char *>(b.xr_all) + at[i];
    const int64_t nc = h.src->ncols;
    if (nc == 0 || h.src->nnz == 0) continue;
    jobs.push_back(XrJob{h.x, h.ldx, nc, h.xr, start});
    start += nchunk * nc;
  }
  b.njobs = static_cast<int64_t>(jobs.size());
  b.xr_threads = start;
  DRX_HIP(hipMalloc(&b.d_jobs, std::max<int64_t>(b.njobs, 1) * sizeof(XrJob)));
  if (b.njobs > 0) {
    DRX_HIP(hipMemcpyAsync(b.d_jobs, jobs.data(), b.njobs * sizeof(XrJob),
                           hipMemcpyHostToDevice, s));
    DRX_HIP(hipStreamSynchronize(s));
  }
}

static void launch_chunks(const Batch &b, hipStream_t s) {
  if (b.xr_threads == 0) return;
  const int kp = b.kp;
  const dim3 g(static_cast<unsigned>(std::min<int64_t>((b.xr_threads + 255) / 256, 1 << 16)));
  const size_t xs = dtype_size(b.xtype);
#define DRX_BXR(KP, T)                                                         \
  hipLaunchKernelGGL((k_batch_xr<KP, T>), g, dim3(256), 0, s, b.d_jobs,        \
                     b.njobs, b.xr_threads, b.k)
#define DRX_BXR_T(T)                                                           \
  do {                                                                         \
    if (kp == 2)                                                               \
      DRX_BXR(2, T);                                                           \
    else if (kp == 4)                                                          \
      DRX_BXR(4, T);                                                           \
    else if (kp == 8)                                                          \
      DRX_BXR(8, T);                                                           \
    else                                                                       \
      DRX_BXR(16, T);                                                          \
  } while (0)
  if (xs == 8)
    DRX_BXR_T(uint64_t);
  else if (xs == 4)
    DRX_BXR_T(uint32_t);
  else
    DRX_BXR_T(uint16_t);
#undef DRX_BXR_T
#undef DRX_BXR
  DRX_HIP(hipGetLastError());
}

static void free_plans(Batch &b) {
  for (auto *p : b.plans) free_plan(p, nullptr);
  b.plans.clear();
  for (void *q : {(void *)b.d_args, (void *)b.d_mode, (void *)b.d_off})
    if (q) (void)hipFree(q);
  b.d_args = nullptr;
  b.d_mode = nullptr;
  b.d_off = nullptr;
}

static void drop_graph(Batch &b) {
  if (b.exec) (void)hipGraphExecDestroy(b.exec);
  if (b.graph) (void)hipGraphDestroy(b.graph);
  b.exec = nullptr;
  b.graph = nullptr;
}

// (Re)build the per-item plans and the device item table for block_nnz.
static void replan(Batch &b, int64_t block_nnz, hipStream_t s) {
  drop_graph(b);
  free_plans(b);
  b.block_nnz = block_nnz;
  const int64_t n = static_cast<int64_t>(b.items.size());
  std::vector<FwdArgs> args(n);
  std::vector<int> modes(n);
  std::vector<int64_t> off(n + 1, 0);
  const size_t xs = dtype_size(b.xtype);
  for (int64_t i = 0; i < n; i++) {
    const BatchItemHost &it = b.items[i];
    const Plan &src = *it.src;
    Plan *p = make_plan(src.itype, src.rowptr, src.base, src.nrows, src.ncols,
                        src.nnz, src.has_sentinel, block_nnz, s, false, false, true);
    b.plans.push_back(p);
    FwdArgs a{};
    a.rowptr = p->rowptr;
    a.base = p->base;
    a.nrows = p->nrows;
    a.nnz = p->nnz;
    a.sent = p->has_sentinel;
    a.rec = p->rec;
    a.nblocks = p->nblocks;
    a.colind = it.colind;
    a.vals = it.vals;
    a.perm = it.perm;
    a.x = it.x;
    a.ldx = it.ldx;
    a.xbytes = p->nnz > 0 ? static_cast<uint32_t>(std::max<int64_t>(p->ncols, 1) * xs) : 0;
    a.y = it.y;
    a.ldy = it.ldy;
    a.k = b.k;
    a.alpha = b.alpha;
    a.beta = b.beta;
    a.gshift = b.gshift;
    a.xr = it.xr;
    a.ncols = p->ncols;
    args[i] = a;
    modes[i] = it.mode;
    off[i + 1] = off[i] + (p->nrows > 0 ? p->nblocks : 0);
  }
  b.total = off[n];
  // Groups of four consecutive entries per lane and load for the forward
  // items (knob vec 0 or 4) when their type allows it and every forward
  // item's vals / colind are aligned to a group.
  {
    const size_t sv = dtype_size(b.vtype);
    const int vec = tuned_vec();
    bool four = (vec == 0 || vec == 4) && b.itype == DRX_I32 && 4 * sv <= 16;
    for (int64_t i = 0; four && i < n; i++)
      if (modes[i] == kPlain)
        four = reinterpret_cast<uintptr_t>(args[i].vals) % (4 * sv) == 0 &&
               reinterpret_cast<uintptr_t>(args[i].colind) % 16 == 0;
    b.vw = four ? 4 : 1;
  }
  DRX_HIP(hipMalloc(&b.d_args, std::max<int64_t>(n, 1) * sizeof(FwdArgs)));
  DRX_HIP(hipMalloc(&b.d_mode, std::max<int64_t>(n, 1) * sizeof(int)));
  DRX_HIP(hipMalloc(&b.d_off, (n + 1) * sizeof(int64_t)));
  DRX_HIP(hipMemcpyAsync(b.d_args, args.data(), n * sizeof(FwdArgs), hipMemcpyHostToDevice, s));
  DRX_HIP(hipMemcpyAsync(b.d_mode, modes.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
  DRX_HIP(hipMemcpyAsync(b.d_off, off.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
  DRX_HIP(hipStreamSynchronize(s));
}

// Write gshift into the device table (host copy rebuilt from the plans).
static void set_gshift(Batch &b, int gshift, hipStream_t s) {
  if (b.gshift == gshift) return;
  drop_graph(b);
  b.gshift = gshift;
  const int64_t n = static_cast<int64_t>(b.items.size());
  std::vector<FwdArgs> args(n);
  DRX_HIP(hipMemcpyAsync(args.data(), b.d_args, n * sizeof(FwdArgs), hipMemcpyDeviceToHost, s));
  DRX_HIP(hipStreamSynchronize(s));
  for (auto &a : args) a.gshift = gshift;
  DRX_HIP(hipMemcpyAsync(b.d_args, args.data(), n * sizeof(FwdArgs), hipMemcpyHostToDevice, s));
  DRX_HIP(hipStreamSynchronize(s));
}

static void launch(Batch &b, hipStream_t s) {
  if (b.total == 0) return;
  const void *fn = nullptr;
  int wg = 256;
  pick_kernel(b, &fn, &wg);
  int per_cu = 0;
  DRX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, wg, 0));
  per_cu = std::max(1, std::min(per_cu, b.wg_per_cu));
  const int64_t grid = static_cast<int64_t>(b.num_cus) * per_cu;  // multiple of 8
  int64_t nitems = static_cast<int64_t>(b.items.size());
  if (b.mrhs) launch_chunks(b, s);
  void *params[] = {&b.d_args, &b.d_mode, &b.d_off, &nitems, &b.total};
  DRX_HIP(hipLaunchKernel(fn, dim3(static_cast<unsigned>(grid)), dim3(wg), params, 0, s));
}

static double time_ms(Batch &b, hipStream_t s, int iters) {
  hipEvent_t e0, e1;
  DRX_HIP(hipEventCreate(&e0));
  DRX_HIP(hipEventCreate(&e1));
  launch(b, s);
  DRX_HIP(hipEventRecord(e0, s));
  for (int i = 0; i < iters; i++) launch(b, s);
  DRX_HIP(hipEventRecord(e1, s));
  DRX_HIP(hipEventSynchronize(e1));
  float ms = 0;
  DRX_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / iters;
}

static std::string tuning_json(const Batch &b) {
  char buf[256];
  std::snprintf(buf, sizeof(buf),
                "{\"block_nnz\": %lld, \"gshift\": %d, \"wg_per_cu\": %d, \"ms\": %.6f}",
                static_cast<long long>(b.block_nnz), b.gshift, b.wg_per_cu, b.tuned_ms);
  return buf;
}

}  // namespace drx

using namespace drx;

extern "C" {

int drx_batch_create(drx_batch *out, int64_t nitems, const drx_batch_item *items,
                     int vtype, int xtype, int ytype, int64_t k, double alpha,
                     double beta, drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(out != nullptr, "batch out-pointer is null");
    *out = nullptr;
    DRX_REQUIRE(nitems >= 0 && (nitems == 0 || items != nullptr), "bad items");
    DRX_REQUIRE(k >= 1, "k must be >= 1");
    hipStream_t s = static_cast<hipStream_t>(stream);
    Batch *b = new Batch();
    try {
      b->vtype = vtype;
      b->xtype = xtype;
      b->ytype = ytype;
      b->k = k;
      b->mrhs = k >= 2 && tuned_mrhs() != 0;
      b->kp = b->mrhs ? mrhs_kp(k, xtype, tuned_mrhs_kp_max() == 8) : 0;
      b->alpha = alpha;
      b->beta = beta;
      int dev = 0;
      DRX_HIP(hipGetDevice(&dev));
      DRX_HIP(hipDeviceGetAttribute(&b->num_cus, hipDeviceAttributeMultiprocessorCount, dev));
      b->num_cus = std::max(kXcds, (b->num_cus / kXcds) * kXcds);
      const size_t xs = dtype_size(xtype);
      for (int64_t i = 0; i < nitems; i++) {
        const drx_batch_item &u = items[i];
        DRX_REQUIRE((u.plan != nullptr) != (u.transpose != nullptr),
                    "each batch item names a plan (forward) or a transpose (adjoint)");
        BatchItemHost h;
        if (u.plan) {
          const Plan *p = reinterpret_cast<const Plan *>(u.plan);
          DRX_REQUIRE(p->panels == nullptr, "batch items use plain plans");
          h.src = p;
          h.mode = kPlain;
          h.colind = u.colind;
        } else {
          const Transpose *t = reinterpret_cast<const Transpose *>(u.transpose);
          h.src = t->plan;
          h.colind = t->rowind;
          if (u.vals == nullptr) {  // the transpose's stored CSC-order values
            DRX_REQUIRE(t->tvals != nullptr && t->tvtype == vtype,
                        "adjoint item without vals needs drx_transpose_set_values");
            h.mode = kPlain;
            h.vals_override = t->tvals;
          } else {
            h.mode = kPerm;
            h.perm = t->perm;
          }
        }
        const Plan &p = *h.src;
        if (i == 0) b->itype = p.itype;
        DRX_REQUIRE(p.itype == b->itype, "batch items must share an index type");
        DRX_REQUIRE(p.nnz == 0 || (h.colind && (u.vals || h.vals_override) && u.x),
                    "null operand in a batch item");
        DRX_REQUIRE(p.nrows == 0 || u.y, "null output in a batch item");
        DRX_REQUIRE(k == 1 || (u.ldx >= p.ncols && u.ldy >= p.nrows), "ldx/ldy too small");
        DRX_REQUIRE(static_cast<uint64_t>(std::max<int64_t>(p.ncols, 1)) * xs < (uint64_t(1) << 32),
                    "batch items need ncols * sizeof(x) < 4 GiB");
        h.vals = h.vals_override ? h.vals_override : u.vals;
        h.x = u.x;
        h.ldx = u.ldx > 0 ? u.ldx : p.ncols;
        h.y = u.y;
        h.ldy = u.ldy > 0 ? u.ldy : p.nrows;
        b->items.push_back(h);
      }
      const void *fn;
      int wg;
      pick_kernel(*b, &fn, &wg);  // validates the type combination
      if (b->mrhs) make_chunks(*b, s);
      // The batch kernels take windows of at most 4096 entries (blocks of
      // at most 2048, as drx_batch_set_tuning enforces); a larger tuned
      // default is clamped instead of streaming only part of a window.
      const int64_t bn = std::min<int64_t>(tuned_block_nnz() > 0 ? tuned_block_nnz() : 1024, 2048);
      replan(*b, bn, s);
    } catch (...) {
      drop_graph(*b);
      free_plans(*b);
      free_chunks(*b);
      delete b;
      throw;
    }
    *out = reinterpret_cast<drx_batch>(b);
  });
}

int drx_batch_destroy(drx_batch batch) {
  return guarded([&] {
    if (!batch) return;
    Batch *b = reinterpret_cast<Batch *>(batch);
    drop_graph(*b);
    free_plans(*b);
    free_chunks(*b);
    delete b;
  });
}

int drx_batch_run(drx_batch batch, drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(batch != nullptr, "batch is null");
    launch(*reinterpret_cast<Batch *>(batch), static_cast<hipStream_t>(stream));
  });
}

int drx_batch_graph(drx_batch batch) {
  return guarded([&] {
    DRX_REQUIRE(batch != nullptr, "batch is null");
    Batch &b = *reinterpret_cast<Batch *>(batch);
    drop_graph(b);
    hipStream_t cs = nullptr;
    DRX_HIP(hipStreamCreate(&cs));
    try {
      DRX_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      try {
        launch(b, cs);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(cs, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      DRX_HIP(hipStreamEndCapture(cs, &b.graph));
      DRX_HIP(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
    } catch (...) {
      (void)hipStreamDestroy(cs);
      drop_graph(b);
      throw;
    }
    (void)hipStreamDestroy(cs);
  });
}

int drx_batch_graph_launch(drx_batch batch, drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(batch != nullptr, "batch is null");
    Batch &b = *reinterpret_cast<Batch *>(batch);
    DRX_REQUIRE(b.exec != nullptr, "no graph: call drx_batch_graph first");
    DRX_HIP(hipGraphLaunch(b.exec, static_cast<hipStream_t>(stream)));
  });
}

int drx_batch_set_tuning(drx_batch batch, int64_t block_nnz, int gshift,
                         int wg_per_cu, drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(batch != nullptr, "batch is null");
    DRX_REQUIRE(block_nnz >= 64 && block_nnz <= 2048 && (block_nnz & (block_nnz - 1)) == 0,
                "block_nnz must be a power of two in [64, 2048]");
    DRX_REQUIRE(gshift >= 0 && gshift <= 3 && wg_per_cu >= 1 && wg_per_cu <= 32,
                "bad gshift / wg_per_cu");
    Batch &b = *reinterpret_cast<Batch *>(batch);
    hipStream_t s = static_cast<hipStream_t>(stream);
    b.wg_per_cu = wg_per_cu;
    drop_graph(b);
    b.gshift = gshift;
    if (block_nnz != b.block_nnz || b.plans.empty())
      replan(b, block_nnz, s);
    else {
      const int g = b.gshift;
      b.gshift = -1;
      set_gshift(b, g, s);
    }
  });
}

int drx_batch_autotune(drx_batch batch, int iters, char *json, int64_t json_cap,
                       drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(batch != nullptr, "batch is null");
    Batch &b = *reinterpret_cast<Batch *>(batch);
    hipStream_t s = static_cast<hipStream_t>(stream);
    iters = std::max(1, iters);
    // Timing runs write the items' outputs; with beta forced to 0 they hold
    // alpha * A x (never a repeatedly accumulated beta * y) afterwards. The
    // caller's beta is restored for the final plan.
    const double beta0 = b.beta;
    struct Restore {
      Batch &b;
      double beta;
      ~Restore() { b.beta = beta; }
    } restore{b, beta0};
    b.beta = 0;
    double best = 1e30;
    int64_t best_bn = b.block_nnz;
    int best_g = 0, best_w = b.wg_per_cu;
    for (int64_t bn : {256, 512, 1024, 2048}) {
      replan(b, bn, s);
      for (int g : {0, 1, 2}) {
        if (g == 2 && !b.mrhs) continue;
        set_gshift(b, g, s);
        for (int w : {2, 4, 8, 16}) {
          b.wg_per_cu = w;
          const double ms = time_ms(b, s, iters);
          if (ms < best) {
            best = ms;
            best_bn = bn;
            best_g = g;
            best_w = w;
          }
        }
      }
    }
    b.beta = beta0;
    replan(b, best_bn, s);
    set_gshift(b, best_g, s);
    b.wg_per_cu = best_w;
    b.tuned_ms = best;
    if (json && json_cap > 0) {
      const std::string j = tuning_json(b);
      std::snprintf(json, static_cast<size_t>(json_cap), "%s", j.c_str());
    }
  });
}

int drx_batch_info(drx_batch batch, int64_t *work_units, int64_t *block_nnz,
                   int *gshift, int *wg_per_cu) {
  return guarded([&] {
    DRX_REQUIRE(batch != nullptr, "batch is null");
    const Batch &b = *reinterpret_cast<Batch *>(batch);
    if (work_units) *work_units = b.total;
    if (block_nnz) *block_nnz = b.block_nnz;
    if (gshift) *gshift = b.gshift;
    if (wg_per_cu) *wg_per_cu = b.wg_per_cu;
  });
}

}  // extern "C"
