// drx_gen.hip -- device-side synthetic inputs (counter-based, reproducible on
// any box), dtype conversion, and the tuning knobs of the library.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "drx_common.hpp"
#include "drx_internal.hpp"

namespace drx {

// ---------------------------------------------------------------------------
// Tuning knobs (process-wide; 0 = default). The environment variable
// DRX_BLOCK_NNZ seeds block_nnz when nothing was set through the API.
// ---------------------------------------------------------------------------
static std::atomic<int64_t> g_block_nnz{0};
static std::atomic<int> g_mrhs{1};
static std::atomic<int> g_mrhs_gshift{1};
static std::atomic<int> g_vec{0};
static std::atomic<int> g_wide{1};
int tuned_wide() { return g_wide.load(); }
static std::atomic<int64_t> g_grid_cap{0};

int tuned_mrhs() { return g_mrhs.load(); }
int tuned_mrhs_gshift() { return g_mrhs_gshift.load(); }
static std::atomic<int> g_mrhs_kp_max{0};
int tuned_mrhs_kp_max() { return g_mrhs_kp_max.load(); }
int tuned_vec() { return g_vec.load(); }
int64_t tuned_grid_cap() { return g_grid_cap.load(); }
static std::atomic<int> g_gemm_tile{0};
int tuned_gemm_tile() { return g_gemm_tile.load(); }
static std::atomic<int> g_xstage{1};
int tuned_xstage() { return g_xstage.load(); }
static std::atomic<int> g_auto_panels{1};
int tuned_auto_panels() { return g_auto_panels.load(); }
static std::atomic<int> g_fill{1};
int tuned_fill() { return g_fill.load(); }
static std::atomic<int> g_fwd_cap{0};
int tuned_fwd_cap() { return g_fwd_cap.load(); }
static std::atomic<int> g_rowgather{1};
int tuned_rowgather() { return g_rowgather.load(); }

int64_t tuned_block_nnz() {
  int64_t v = g_block_nnz.load();
  if (v > 0) return v;
  if (const char *e = std::getenv("DRX_BLOCK_NNZ")) {
    long long w = std::atoll(e);
    if (w >= 64 && w <= 4096 && (w & (w - 1)) == 0) return w;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Counter-based randomness: splitmix64 finalizer.
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
// Stream `seed`, counter i.
__host__ __device__ __forceinline__ uint64_t rnd(uint64_t seed, uint64_t i) {
  return mix64(mix64(seed) ^ (i * 0xD1B54A32D192ED03ull));
}
__host__ __device__ __forceinline__ double u01(uint64_t h) {
  return static_cast<double>(h >> 11) * (1.0 / 9007199254740992.0);
}

template <typename T> __device__ __forceinline__ T store_as(double v) {
  if constexpr (std::is_same_v<T, bf16>)
    return f32_to_bf16(static_cast<float>(v));
  else if constexpr (std::is_same_v<T, f16>)
    return static_cast<f16>(static_cast<float>(v));
  else
    return static_cast<T>(v);
}

template <typename T>
__global__ void k_uniform(T *out, int64_t n, double lo, double hi,
                          uint64_t seed) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       i < n; i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[i] = store_as<T>(lo + (hi - lo) * u01(rnd(seed, i)));
}

template <typename T>
__global__ void k_normal(T *out, int64_t n, uint64_t seed) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       i < n; i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const double u1 = (static_cast<double>(rnd(seed, 2 * i) >> 11) + 1.0) *
                      (1.0 / 9007199254740992.0);
    const double u2 = u01(rnd(seed, 2 * i + 1));
    out[i] = store_as<T>(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
  }
}

// out[i] = x[idx[i]]: streamed (non-temporal) indices, buffer gathers of x
// with 32-bit offsets, 4 independent gathers per lane in flight -- the
// memory pattern of a sparse product's gathers without its reduction.
template <typename T>
__global__ __launch_bounds__(256) void k_gather(const int32_t *idx, int64_t n,
                                                const void *x, uint32_t xbytes, T *out) {
  const rsrc_t rx = make_rsrc(x, xbytes);
  const int64_t step = static_cast<int64_t>(gridDim.x) * 1024;
  for (int64_t base = static_cast<int64_t>(blockIdx.x) * 1024; base < n; base += step) {
    int64_t e[4];
    uint32_t c[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      e[u] = base + threadIdx.x + u * 256;
      c[u] = e[u] < n ? static_cast<uint32_t>(__builtin_nontemporal_load(idx + e[u])) : 0u;
    }
    T v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) v[u] = buf_load<T>(rx, c[u] * static_cast<uint32_t>(sizeof(T)));
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (e[u] < n) out[e[u]] = v[u];
  }
}

// Read-rate probe: every workgroup streams 16-byte chunks (non-temporal
// loads, 8 in flight per lane) over its share of [0, n16) and writes one
// 4-byte digest per workgroup, so that the loads cannot be dropped. The
// benchmark times it over a buffer far larger than the caches to measure the
// HBM read ceiling of the box.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_stream_read(const u32x4 *src, int64_t n16,
                                                     uint32_t *out) {
  uint32_t acc = 0;
  const int64_t step = static_cast<int64_t>(gridDim.x) * 256 * 8;
  for (int64_t base = static_cast<int64_t>(blockIdx.x) * 256 * 8 + threadIdx.x; base < n16;
       base += step) {
    u32x4 v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int64_t i = base + u * 256;
      v[u] = i < n16 ? __builtin_nontemporal_load(src + i) : u32x4{0, 0, 0, 0};
    }
#pragma unroll
    for (int u = 0; u < 8; u++) acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
  }
  for (int o = 32; o > 0; o >>= 1) acc ^= __shfl_xor(acc, o, 64);
  __shared__ uint32_t red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = red[0] ^ red[1] ^ red[2] ^ red[3];
}

static dim3 grid1d(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65536 * 8) b = 65536 * 8;
  if (b < 1) b = 1;
  return dim3(static_cast<unsigned>(b));
}

// ---------------------------------------------------------------------------
// Row pointers from row lengths: len[1..n] in a 64-bit buffer, inclusive
// scan, then narrowed to the index type.
// ---------------------------------------------------------------------------
template <typename TI>
__global__ void k_narrow(const int64_t *src, TI *dst, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       i < n; i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[i] = static_cast<TI>(src[i]);
}

// rp64[0..nrows] holds 0 followed by the lengths; scans it in place and
// writes the index-typed rowptr. Returns the total (synchronizes).
static int64_t finish_rowptr(int64_t *rp64, int64_t nrows, int itype,
                             void *rowptr, hipStream_t s) {
  size_t temp_bytes = 0;
  const int64_t n = nrows + 1;
  DRX_REQUIRE(n <= INT32_MAX, "too many rows for the scan");
  DRX_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, temp_bytes, rp64, rp64,
                                           static_cast<int>(n), s));
  void *temp = nullptr;
  DRX_HIP(hipMalloc(&temp, std::max<size_t>(temp_bytes, 1)));
  DRX_HIP(hipcub::DeviceScan::InclusiveSum(temp, temp_bytes, rp64, rp64,
                                           static_cast<int>(n), s));
  int64_t total = 0;
  DRX_HIP(hipMemcpyAsync(&total, rp64 + nrows, sizeof(int64_t),
                         hipMemcpyDeviceToHost, s));
  DRX_HIP(hipStreamSynchronize(s));
  (void)hipFree(temp);
  if (itype == DRX_I32) {
    DRX_REQUIRE(total <= INT32_MAX, "nnz exceeds the 32-bit index type");
    hipLaunchKernelGGL(k_narrow<int32_t>, grid1d(n), dim3(256), 0, s, rp64,
                       static_cast<int32_t *>(rowptr), n);
  } else {
    DRX_HIP(hipMemcpyAsync(rowptr, rp64, n * sizeof(int64_t),
                           hipMemcpyDeviceToDevice, s));
  }
  DRX_HIP(hipGetLastError());
  DRX_HIP(hipStreamSynchronize(s));
  return total;
}

// ---------------------------------------------------------------------------
// Zipf row lengths: P(L = l) = l^-s / zeta(s) for l < lmax, the remaining
// mass at lmax (a Zipf draw clipped to [1, lmax]).
// ---------------------------------------------------------------------------
__global__ void k_zipf_len(int64_t *rp64, int64_t nrows, const double *cdf,
                           int64_t lmax, uint64_t seed) {
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       r < nrows; r += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const double u = u01(rnd(seed, r));
    int64_t lo = 0, hi = lmax - 1;  // first index with cdf[idx] > u
    while (lo < hi) {
      int64_t mid = (lo + hi) / 2;
      if (cdf[mid] > u)
        hi = mid;
      else
        lo = mid + 1;
    }
    rp64[r + 1] = lo + 1;
    if (r == 0) rp64[0] = 0;
  }
}

__global__ void k_uniform_len(int64_t *rp64, int64_t nrows, int64_t lmin,
                              int64_t lmax, uint64_t seed) {
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       r < nrows; r += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const uint64_t span = static_cast<uint64_t>(lmax - lmin + 1);
    rp64[r + 1] = lmin + static_cast<int64_t>(rnd(seed, r) % span);
    if (r == 0) rp64[0] = 0;
  }
}

static double zeta(double s) {
  const int N = 100000;
  double acc = 0;
  for (int i = N - 1; i >= 1; i--) acc += std::pow(static_cast<double>(i), -s);
  const double n = N;
  acc += std::pow(n, 1 - s) / (s - 1) + 0.5 * std::pow(n, -s) +
         s * std::pow(n, -s - 1) / 12.0;
  return acc;
}

// ---------------------------------------------------------------------------
// Stratified sorted distinct columns: entry i of a row of length L takes
// column floor(i*n/L) + (hash % (floor((i+1)*n/L) - floor(i*n/L))).
// One wavefront per row.
// ---------------------------------------------------------------------------
template <typename TI>
__global__ void k_stratified(const TI *rowptr, int64_t nrows, int64_t ncols,
                             TI *colind, uint64_t seed) {
  const int64_t wave = (static_cast<int64_t>(blockIdx.x) * blockDim.x +
                        threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = static_cast<int64_t>(gridDim.x) * blockDim.x / 64;
  for (int64_t r = wave; r < nrows; r += nwaves) {
    const int64_t b = rowptr[r], e = rowptr[r + 1], L = e - b;
    for (int64_t i = lane; i < L; i += 64) {
      const int64_t lo = (i * ncols) / L, hi = ((i + 1) * ncols) / L;
      const uint64_t h = rnd(seed, static_cast<uint64_t>(b + i));
      const int64_t w = hi > lo ? hi - lo : 1;  // rows longer than ncols
      colind[b + i] = static_cast<TI>(lo + static_cast<int64_t>(
                                               h % static_cast<uint64_t>(w)));
    }
  }
}

// ---------------------------------------------------------------------------
// Laplacians (lexicographic rows, sorted columns) and band matrices.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void lap_coords(int dim, int64_t g, int64_t r,
                                           int64_t c[3]) {
  c[0] = c[1] = c[2] = 0;
  if (dim == 2) {
    c[0] = r / g;
    c[1] = r % g;
  } else {
    c[0] = r / (g * g);
    c[1] = (r / g) % g;
    c[2] = r % g;
  }
}

__global__ void k_lap_len(int dim, int64_t g, int64_t nrows, int64_t *rp64) {
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       r < nrows; r += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    int64_t c[3];
    lap_coords(dim, g, r, c);
    int64_t len = 1;
    for (int d = 0; d < dim; d++) len += (c[d] > 0) + (c[d] < g - 1);
    rp64[r + 1] = len;
    if (r == 0) rp64[0] = 0;
  }
}

template <typename TI, typename TV>
__global__ void k_lap_fill(int dim, int64_t g, int64_t nrows,
                           const TI *rowptr, TI *colind, TV *vals) {
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       r < nrows; r += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    int64_t c[3];
    lap_coords(dim, g, r, c);
    int64_t q = rowptr[r];
    // strides of coordinates 0..dim-1 in row numbering
    const int64_t st[3] = {dim == 2 ? g : g * g, dim == 2 ? 1 : g, 1};
    // lower neighbours, from the farthest (largest stride) inward
    for (int d = 0; d < dim; d++)
      if (c[d] > 0) {
        colind[q] = static_cast<TI>(r - st[d]);
        vals[q++] = store_as<TV>(-1.0);
      }
    colind[q] = static_cast<TI>(r);
    vals[q++] = store_as<TV>(2.0 * dim);
    for (int d = dim - 1; d >= 0; d--)
      if (c[d] < g - 1) {
        colind[q] = static_cast<TI>(r + st[d]);
        vals[q++] = store_as<TV>(-1.0);
      }
  }
}

__global__ void k_band_len(int64_t n, int64_t lo, int64_t hi, int64_t *rp64) {
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       r < n; r += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t a = r - lo < 0 ? 0 : r - lo;
    const int64_t b = r + hi > n - 1 ? n - 1 : r + hi;
    rp64[r + 1] = b - a + 1;
    if (r == 0) rp64[0] = 0;
  }
}

template <typename TI>
__global__ void k_band_fill(int64_t n, int64_t lo, const TI *rowptr,
                            TI *colind) {
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       r < n; r += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t a = r - lo < 0 ? 0 : r - lo;
    const int64_t b = rowptr[r], e = rowptr[r + 1];
    for (int64_t q = b; q < e; q++) colind[q] = static_cast<TI>(a + (q - b));
  }
}

template <typename TS, typename TD>
__global__ void k_convert(const TS *src, TD *dst, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       i < n; i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    double v;
    if constexpr (std::is_same_v<TS, bf16> || std::is_same_v<TS, f16>)
      v = static_cast<double>(to_f32(src[i]));
    else
      v = static_cast<double>(src[i]);
    if constexpr (std::is_same_v<TD, bf16>)
      dst[i] = f32_to_bf16(static_cast<float>(v));
    else if constexpr (std::is_same_v<TD, f16>)
      dst[i] = static_cast<f16>(static_cast<float>(v));
    else
      dst[i] = static_cast<TD>(v);
  }
}

}  // namespace drx

using namespace drx;

extern "C" {

int drx_tune_set(const char *key, int64_t value) {
  return guarded([&] {
    DRX_REQUIRE(key != nullptr, "key is null");
    if (std::strcmp(key, "block_nnz") == 0) {
      DRX_REQUIRE(value == 0 || (value >= 64 && value <= 4096 &&
                                 (value & (value - 1)) == 0),
                  "block_nnz must be 0 or a power of two in [64, 4096]");
      g_block_nnz.store(value);
    } else if (std::strcmp(key, "mrhs") == 0) {
      DRX_REQUIRE(value >= 0 && value <= 2, "mrhs must be 0, 1 or 2");
      g_mrhs.store(static_cast<int>(value));
    } else if (std::strcmp(key, "mrhs_gshift") == 0) {
      DRX_REQUIRE(value >= 0 && value <= 3, "mrhs_gshift must be in [0, 3]");
      g_mrhs_gshift.store(static_cast<int>(value));
    } else if (std::strcmp(key, "mrhs_kp_max") == 0) {
      DRX_REQUIRE(value == 0 || value == 8 || value == 16, "mrhs_kp_max must be 0, 8 or 16");
      g_mrhs_kp_max.store(static_cast<int>(value));
    } else if (std::strcmp(key, "wide") == 0) {
      DRX_REQUIRE(value == 0 || value == 1, "wide must be 0 or 1");
      g_wide.store(static_cast<int>(value));
    } else if (std::strcmp(key, "vec") == 0) {
      DRX_REQUIRE(value == 0 || value == 1 || value == 2 || value == 4,
                  "vec must be 0, 1, 2 or 4");
      g_vec.store(static_cast<int>(value));
    } else if (std::strcmp(key, "grid_cap") == 0) {
      DRX_REQUIRE(value >= 0 && value <= (int64_t(1) << 30), "grid_cap out of range");
      g_grid_cap.store(value);
    } else if (std::strcmp(key, "xstage") == 0) {
      DRX_REQUIRE(value >= 0 && value <= 2, "xstage must be 0, 1 or 2");
      g_xstage.store(static_cast<int>(value));
    } else if (std::strcmp(key, "fill") == 0) {
      DRX_REQUIRE(value == 0 || value == 1, "fill must be 0 or 1");
      g_fill.store(static_cast<int>(value));
    } else if (std::strcmp(key, "fwd_cap") == 0) {
      DRX_REQUIRE(value == 0 || value == 2048 || value == 4096,
                  "fwd_cap must be 0, 2048 or 4096");
      g_fwd_cap.store(static_cast<int>(value));
    } else if (std::strcmp(key, "rowgather") == 0) {
      DRX_REQUIRE(value >= 0 && value <= 3, "rowgather must be in [0, 3]");
      g_rowgather.store(static_cast<int>(value));
    } else if (std::strcmp(key, "auto_panels") == 0) {
      DRX_REQUIRE(value == 0 || value == 1, "auto_panels must be 0 or 1");
      g_auto_panels.store(static_cast<int>(value));
    } else if (std::strcmp(key, "gemm_tile") == 0) {
      DRX_REQUIRE(value >= 0 && value <= 7, "gemm_tile must be in [0, 7]");
      g_gemm_tile.store(static_cast<int>(value));
    } else {
      throw Error(DRX_ERR_INVALID, std::string("unknown tuning key: ") + key);
    }
  });
}

int drx_tune_get(const char *key, int64_t *value) {
  return guarded([&] {
    DRX_REQUIRE(key != nullptr && value != nullptr, "null argument");
    if (std::strcmp(key, "block_nnz") == 0)
      *value = default_block_nnz(DRX_I32);
    else if (std::strcmp(key, "mrhs") == 0)
      *value = g_mrhs.load();
    else if (std::strcmp(key, "mrhs_gshift") == 0)
      *value = g_mrhs_gshift.load();
    else if (std::strcmp(key, "vec") == 0)
      *value = g_vec.load();
    else if (std::strcmp(key, "wide") == 0)
      *value = g_wide.load();
    else if (std::strcmp(key, "mrhs_kp_max") == 0)
      *value = g_mrhs_kp_max.load();
    else if (std::strcmp(key, "grid_cap") == 0)
      *value = g_grid_cap.load();
    else if (std::strcmp(key, "gemm_tile") == 0)
      *value = g_gemm_tile.load();
    else if (std::strcmp(key, "xstage") == 0)
      *value = g_xstage.load();
    else if (std::strcmp(key, "auto_panels") == 0)
      *value = g_auto_panels.load();
    else if (std::strcmp(key, "fill") == 0)
      *value = g_fill.load();
    else if (std::strcmp(key, "fwd_cap") == 0)
      *value = g_fwd_cap.load();
    else if (std::strcmp(key, "rowgather") == 0)
      *value = g_rowgather.load();
    else
      throw Error(DRX_ERR_INVALID, std::string("unknown tuning key: ") + key);
  });
}

int drx_gen_uniform(int dtype, void *out, int64_t n, double lo, double hi,
                    uint64_t seed, drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(n >= 0 && (n == 0 || out), "bad output");
    if (n == 0) return;
    hipStream_t s = static_cast<hipStream_t>(stream);
    dim3 g = grid1d(n);
    switch (dtype) {
      case DRX_F64: hipLaunchKernelGGL(k_uniform<double>, g, dim3(256), 0, s, static_cast<double *>(out), n, lo, hi, seed); break;
      case DRX_F32: hipLaunchKernelGGL(k_uniform<float>, g, dim3(256), 0, s, static_cast<float *>(out), n, lo, hi, seed); break;
      case DRX_BF16: hipLaunchKernelGGL(k_uniform<bf16>, g, dim3(256), 0, s, static_cast<bf16 *>(out), n, lo, hi, seed); break;
      case DRX_F16: hipLaunchKernelGGL(k_uniform<f16>, g, dim3(256), 0, s, static_cast<f16 *>(out), n, lo, hi, seed); break;
      default: throw Error(DRX_ERR_INVALID, "bad dtype");
    }
    DRX_HIP(hipGetLastError());
  });
}

int drx_gen_normal(int dtype, void *out, int64_t n, uint64_t seed,
                   drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(n >= 0 && (n == 0 || out), "bad output");
    if (n == 0) return;
    hipStream_t s = static_cast<hipStream_t>(stream);
    dim3 g = grid1d(n);
    switch (dtype) {
      case DRX_F64: hipLaunchKernelGGL(k_normal<double>, g, dim3(256), 0, s, static_cast<double *>(out), n, seed); break;
      case DRX_F32: hipLaunchKernelGGL(k_normal<float>, g, dim3(256), 0, s, static_cast<float *>(out), n, seed); break;
      case DRX_BF16: hipLaunchKernelGGL(k_normal<bf16>, g, dim3(256), 0, s, static_cast<bf16 *>(out), n, seed); break;
      case DRX_F16: hipLaunchKernelGGL(k_normal<f16>, g, dim3(256), 0, s, static_cast<f16 *>(out), n, seed); break;
      default: throw Error(DRX_ERR_INVALID, "bad dtype");
    }
    DRX_HIP(hipGetLastError());
  });
}

int drx_gen_zipf_rowptr(int itype, void *rowptr, int64_t nrows, double s_exp,
                        int64_t lmax, uint64_t seed, int64_t *nnz_out,
                        drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(itype == DRX_I32 || itype == DRX_I64, "bad itype");
    DRX_REQUIRE(nrows >= 1 && rowptr && nnz_out, "bad argument");
    DRX_REQUIRE(s_exp > 1.0 && lmax >= 1 && lmax <= (1 << 24),
                "need s > 1 and 1 <= lmax <= 2^24");
    hipStream_t s = static_cast<hipStream_t>(stream);
    // cdf[l-1] = P(L <= l); cdf[lmax-1] = 1 (the clipped tail).
    std::vector<double> cdf(static_cast<size_t>(lmax));
    const double z = zeta(s_exp);
    double acc = 0;
    for (int64_t l = 1; l < lmax; l++) {
      acc += std::pow(static_cast<double>(l), -s_exp) / z;
      cdf[l - 1] = acc;
    }
    cdf[lmax - 1] = 2.0;  // every u in [0, 1) lands at or before lmax
    double *d_cdf = nullptr;
    int64_t *rp64 = nullptr;
    DRX_HIP(hipMalloc(&d_cdf, lmax * sizeof(double)));
    DRX_HIP(hipMalloc(&rp64, (nrows + 1) * sizeof(int64_t)));
    try {
      DRX_HIP(hipMemcpyAsync(d_cdf, cdf.data(), lmax * sizeof(double),
                             hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(k_zipf_len, grid1d(nrows), dim3(256), 0, s, rp64,
                         nrows, d_cdf, lmax, seed);
      DRX_HIP(hipGetLastError());
      *nnz_out = finish_rowptr(rp64, nrows, itype, rowptr, s);
    } catch (...) {
      (void)hipFree(d_cdf);
      (void)hipFree(rp64);
      throw;
    }
    (void)hipFree(d_cdf);
    (void)hipFree(rp64);
  });
}

int drx_gen_uniform_rowptr(int itype, void *rowptr, int64_t nrows,
                           int64_t lmin, int64_t lmax, uint64_t seed,
                           int64_t *nnz_out, drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(itype == DRX_I32 || itype == DRX_I64, "bad itype");
    DRX_REQUIRE(nrows >= 1 && rowptr && nnz_out, "bad argument");
    DRX_REQUIRE(lmin >= 0 && lmax >= lmin, "need 0 <= lmin <= lmax");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int64_t *rp64 = nullptr;
    DRX_HIP(hipMalloc(&rp64, (nrows + 1) * sizeof(int64_t)));
    try {
      hipLaunchKernelGGL(k_uniform_len, grid1d(nrows), dim3(256), 0, s, rp64,
                         nrows, lmin, lmax, seed);
      DRX_HIP(hipGetLastError());
      *nnz_out = finish_rowptr(rp64, nrows, itype, rowptr, s);
    } catch (...) {
      (void)hipFree(rp64);
      throw;
    }
    (void)hipFree(rp64);
  });
}

int drx_gen_stratified_colind(int itype, const void *rowptr, int64_t nrows,
                              int64_t ncols, void *colind, uint64_t seed,
                              drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(itype == DRX_I32 || itype == DRX_I64, "bad itype");
    DRX_REQUIRE(nrows >= 0 && ncols >= 1 && ncols < (int64_t(1) << 40),
                "bad size");
    DRX_REQUIRE(nrows == 0 || (rowptr && colind), "null argument");
    if (nrows == 0) return;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int64_t waves = std::min<int64_t>(nrows, int64_t(1) << 20);
    dim3 g(static_cast<unsigned>((waves * 64 + 255) / 256));
    if (itype == DRX_I32)
      hipLaunchKernelGGL(k_stratified<int32_t>, g, dim3(256), 0, s,
                         static_cast<const int32_t *>(rowptr), nrows, ncols,
                         static_cast<int32_t *>(colind), seed);
    else
      hipLaunchKernelGGL(k_stratified<int64_t>, g, dim3(256), 0, s,
                         static_cast<const int64_t *>(rowptr), nrows, ncols,
                         static_cast<int64_t *>(colind), seed);
    DRX_HIP(hipGetLastError());
  });
}

int64_t drx_gen_laplacian_nnz(int dim, int64_t g) {
  if (g < 1 || (dim != 2 && dim != 3)) return -1;
  if (dim == 2) return 5 * g * g - 4 * g;
  return 7 * g * g * g - 6 * g * g;
}

int drx_gen_laplacian(int dim, int64_t g, int itype, void *rowptr,
                      void *colind, int vtype, void *vals,
                      drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
    DRX_REQUIRE(g >= 1 && rowptr && colind && vals, "bad argument");
    DRX_REQUIRE(itype == DRX_I32 || itype == DRX_I64, "bad itype");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t nrows = dim == 2 ? g * g : g * g * g;
    int64_t *rp64 = nullptr;
    DRX_HIP(hipMalloc(&rp64, (nrows + 1) * sizeof(int64_t)));
    try {
      hipLaunchKernelGGL(k_lap_len, grid1d(nrows), dim3(256), 0, s, dim, g,
                         nrows, rp64);
      DRX_HIP(hipGetLastError());
      const int64_t nnz = finish_rowptr(rp64, nrows, itype, rowptr, s);
      DRX_REQUIRE(nnz == drx_gen_laplacian_nnz(dim, g), "laplacian nnz");
#define DRX_LAP(TI, TV)                                                        \
  hipLaunchKernelGGL((k_lap_fill<TI, TV>), grid1d(nrows), dim3(256), 0, s,     \
                     dim, g, nrows, static_cast<const TI *>(rowptr),           \
                     static_cast<TI *>(colind), static_cast<TV *>(vals))
      const bool i32 = itype == DRX_I32;
      switch (vtype) {
        case DRX_F64: if (i32) DRX_LAP(int32_t, double); else DRX_LAP(int64_t, double); break;
        case DRX_F32: if (i32) DRX_LAP(int32_t, float); else DRX_LAP(int64_t, float); break;
        case DRX_BF16: if (i32) DRX_LAP(int32_t, bf16); else DRX_LAP(int64_t, bf16); break;
        case DRX_F16: if (i32) DRX_LAP(int32_t, f16); else DRX_LAP(int64_t, f16); break;
        default: throw Error(DRX_ERR_INVALID, "bad vtype");
      }
#undef DRX_LAP
      DRX_HIP(hipGetLastError());
      DRX_HIP(hipStreamSynchronize(s));
    } catch (...) {
      (void)hipFree(rp64);
      throw;
    }
    (void)hipFree(rp64);
  });
}

int64_t drx_gen_band_nnz(int64_t n, int64_t lo, int64_t hi) {
  if (n < 1 || lo < 0 || hi < 0) return -1;
  int64_t total = 0;
  // sum over rows of min(n-1, r+hi) - max(0, r-lo) + 1
  const int64_t l = std::min(lo, n - 1), h = std::min(hi, n - 1);
  total = n * (1 + l + h) - l * (l + 1) / 2 - h * (h + 1) / 2;
  return total;
}

int drx_gen_band(int64_t n, int64_t lo, int64_t hi, int itype, void *rowptr,
                 void *colind, int vtype, void *vals, uint64_t seed,
                 drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(n >= 1 && lo >= 0 && hi >= 0 && rowptr && colind,
                "bad argument");
    DRX_REQUIRE(itype == DRX_I32 || itype == DRX_I64, "bad itype");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int64_t *rp64 = nullptr;
    DRX_HIP(hipMalloc(&rp64, (n + 1) * sizeof(int64_t)));
    int64_t nnz = 0;
    try {
      hipLaunchKernelGGL(k_band_len, grid1d(n), dim3(256), 0, s, n, lo, hi,
                         rp64);
      DRX_HIP(hipGetLastError());
      nnz = finish_rowptr(rp64, n, itype, rowptr, s);
      DRX_REQUIRE(nnz == drx_gen_band_nnz(n, lo, hi), "band nnz");
      if (itype == DRX_I32)
        hipLaunchKernelGGL(k_band_fill<int32_t>, grid1d(n), dim3(256), 0, s, n,
                           lo, static_cast<const int32_t *>(rowptr),
                           static_cast<int32_t *>(colind));
      else
        hipLaunchKernelGGL(k_band_fill<int64_t>, grid1d(n), dim3(256), 0, s, n,
                           lo, static_cast<const int64_t *>(rowptr),
                           static_cast<int64_t *>(colind));
      DRX_HIP(hipGetLastError());
      DRX_HIP(hipStreamSynchronize(s));
    } catch (...) {
      (void)hipFree(rp64);
      throw;
    }
    (void)hipFree(rp64);
    if (vals) {
      int st = drx_gen_normal(vtype, vals, nnz, seed, stream);
      if (st != DRX_OK) throw Error(st, drx_last_error());
    }
  });
}

int drx_gather(const int32_t *idx, int64_t n, int xtype, const void *x,
               int64_t nx, void *out, drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(n >= 0 && nx >= 0 && (n == 0 || (idx && x && out)), "bad argument");
    DRX_REQUIRE(xtype >= DRX_F64 && xtype <= DRX_F16, "bad xtype");
    DRX_REQUIRE(static_cast<uint64_t>(std::max<int64_t>(nx, 1)) * dtype_size(xtype) <
                    (uint64_t(1) << 32),
                "x must be below 4 GiB");
    if (n == 0) return;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t xb = static_cast<uint32_t>(std::max<int64_t>(nx, 1) * dtype_size(xtype));
    const int64_t per = 1024;  // indices per workgroup (4 per lane)
    const dim3 g(static_cast<unsigned>(std::min<int64_t>((n + per - 1) / per, int64_t(1) << 22)));
    switch (dtype_size(xtype)) {
      case 8: hipLaunchKernelGGL(k_gather<uint64_t>, g, dim3(256), 0, s, idx, n, x, xb, static_cast<uint64_t *>(out)); break;
      case 4: hipLaunchKernelGGL(k_gather<uint32_t>, g, dim3(256), 0, s, idx, n, x, xb, static_cast<uint32_t *>(out)); break;
      default: hipLaunchKernelGGL(k_gather<uint16_t>, g, dim3(256), 0, s, idx, n, x, xb, static_cast<uint16_t *>(out)); break;
    }
    DRX_HIP(hipGetLastError());
  });
}

int drx_stream_read(const void *src, int64_t bytes, void *digest,
                    int64_t digest_len, drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(bytes >= 0 && bytes % 16 == 0, "bytes must be a multiple of 16");
    DRX_REQUIRE(bytes == 0 || (src && digest), "bad argument");
    DRX_REQUIRE(reinterpret_cast<uintptr_t>(src) % 16 == 0, "src must be 16-byte aligned");
    DRX_REQUIRE(digest_len >= 1, "digest_len must be at least 1");
    if (bytes == 0) return;
    const int64_t n16 = bytes / 16;
    const int64_t per = 256 * 8;
    const int64_t g = std::min<int64_t>(std::min<int64_t>((n16 + per - 1) / per, digest_len),
                                        int64_t(1) << 20);
    hipLaunchKernelGGL(k_stream_read, dim3(static_cast<unsigned>(g)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), static_cast<const u32x4 *>(src), n16,
                       static_cast<uint32_t *>(digest));
    DRX_HIP(hipGetLastError());
  });
}

int drx_copy_device(void *dst, const void *src, int64_t bytes,
                    drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(bytes >= 0 && (bytes == 0 || (dst && src)), "bad argument");
    if (bytes == 0) return;
    hipStream_t s = static_cast<hipStream_t>(stream);
    DRX_HIP(hipMemcpyAsync(dst, src, static_cast<size_t>(bytes),
                           hipMemcpyDefault, s));
    DRX_HIP(hipStreamSynchronize(s));
  });
}

int drx_convert(int src_dtype, const void *src, int dst_dtype, void *dst,
                int64_t n, drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(n >= 0 && (n == 0 || (src && dst)), "bad argument");
    if (n == 0) return;
    hipStream_t s = static_cast<hipStream_t>(stream);
    dim3 g = grid1d(n);
#define DRX_CV2(TS, TD)                                                        \
  hipLaunchKernelGGL((k_convert<TS, TD>), g, dim3(256), 0, s,                  \
                     static_cast<const TS *>(src), static_cast<TD *>(dst), n)
#define DRX_CV1(TS)                                                            \
  switch (dst_dtype) {                                                         \
    case DRX_F64: DRX_CV2(TS, double); break;                                  \
    case DRX_F32: DRX_CV2(TS, float); break;                                   \
    case DRX_BF16: DRX_CV2(TS, bf16); break;                                   \
    case DRX_F16: DRX_CV2(TS, f16); break;                                     \
    default: throw Error(DRX_ERR_INVALID, "bad dst dtype");                    \
  }
    switch (src_dtype) {
      case DRX_F64: DRX_CV1(double); break;
      case DRX_F32: DRX_CV1(float); break;
      case DRX_BF16: DRX_CV1(bf16); break;
      case DRX_F16: DRX_CV1(f16); break;
      default: throw Error(DRX_ERR_INVALID, "bad src dtype");
    }
#undef DRX_CV1
#undef DRX_CV2
    DRX_HIP(hipGetLastError());
  });
}

}  // extern "C"
