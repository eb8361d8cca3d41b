// drx_gemm.hip -- dense row-major product C = alpha * A * B + beta * C on the
// CDNA4 matrix cores: the local tile product behind dr::sp::gemm
// (include/dr/sp/algorithms/matrix/local_gemm.hpp:15-41 custom_gemm,
// c[i][j] += a[i][k] * b[k][j]; :45-60 mkl_gemm, a row-major gemm with
// alpha = beta = 1 and the views' leading dimensions).
//
// Each workgroup owns a BM x BN tile of C; its waves form a WM x WN grid and
// each wave owns (BM/WM) x (BN/WN) as 16 x 16 MFMA tiles
// (v_mfma_f32_16x16x4_f32 / v_mfma_f64_16x16x4_f64). K advances in BK-deep
// stages through two LDS buffers: while the waves multiply out of one buffer,
// the next stage's A and B slices are in flight to registers and are written
// to the other buffer, so each stage costs one barrier.
//
// MFMA 16x16x4 operand layout (64 lanes, g = lane / 16, r = lane % 16):
//   A operand  lane -> A[r][g],   B operand lane -> B[g][r],
//   result     lane, element v -> D[4 g + v][r] (f32), D[g + 4 v][r] (f64).
// The k index the hardware calls g is free to map to any real k as long as A
// and B agree, so lane group g takes the KK consecutive k of its slot: its A
// operands for KK successive instructions come from one 16-byte LDS read.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "drx_common.hpp"
#include "drx_internal.hpp"
#include "drx_spmv.h"

namespace drx {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<float> {
  using acc_t = f32x4;
  static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // Row of the 16 x 16 result that element v of lane group g holds.
  static __device__ __forceinline__ int row(int g, int v) { return 4 * g + v; }
};
template <> struct Mfma<double> {
  using acc_t = f64x4;
  static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int v) { return g + 4 * v; }
};

template <typename T, int BM, int BN, int BK, int WM, int WN>
struct GemmCfg {
  static constexpr int kThreads = 64 * WM * WN;
  static constexpr int V = 16 / static_cast<int>(sizeof(T));  // elements per 16-byte access
  static constexpr int KK = V;                                 // k per lane per A read
  static constexpr int SK = 4 * KK;                            // k per LDS sub-step
  static constexpr int LDA_S = BK + V;  // A tile row stride (elements): one 16-byte pad
  static constexpr int LDB_S = BN + V;  // B tile row stride
  static constexpr int TM = BM / WM, TN = BN / WN;
  static constexpr int FM = TM / 16, FN = TN / 16;
  static constexpr int A_TPR = BK / V, B_TPR = BN / V;  // 16-byte slots per tile row
  static constexpr int A_VECS = BM * BK / V / kThreads;
  static constexpr int B_VECS = BK * BN / V / kThreads;
  static constexpr int STAGE = BM * LDA_S + BK * LDB_S;  // elements per buffer
  static_assert(BK % SK == 0, "BK must hold whole sub-steps");
  static_assert(TM % 16 == 0 && TN % 16 == 0, "wave tile must be whole MFMA tiles");
  static_assert(A_VECS * V * kThreads == BM * BK, "A stage must split evenly");
  static_assert(B_VECS * V * kThreads == BK * BN, "B stage must split evenly");
  static constexpr int ACC_REGS = FM * FN * 4 * static_cast<int>(sizeof(T)) / 4;
  static_assert(2 * STAGE * sizeof(T) <= 160 * 1024, "two stages must fit 160 KB of LDS");
};

// Tile rows are taken GROUP_M at a time and swept column by column, so the
// workgroups resident on one XCD share A row panels and B column panels in
// that XCD's L2.
constexpr int kGroupM = 8;

// Waves per SIMD the register budget is held to, from the accumulator
// registers per lane: 64 -> 4 waves with eight-wave workgroups, 3 with
// four-wave ones (their lanes carry more of each stage's loads; 4 would
// spill); 128 -> 2; more -> 1.
template <int ACC_REGS, int THREADS> constexpr int gemm_waves() {
  return ACC_REGS <= 64 ? (THREADS >= 512 ? 4 : 3) : ACC_REGS <= 128 ? 2 : 1;
}

template <typename T, int BM, int BN, int BK, int WM, int WN>
__global__ void __launch_bounds__(64 * WM * WN)
    __attribute__((amdgpu_waves_per_eu(
        gemm_waves<GemmCfg<T, BM, BN, BK, WM, WN>::ACC_REGS, 64 * WM * WN>())))
    k_gemm(int64_t m, int64_t n, int64_t K, T alpha, const T *__restrict__ A,
           int64_t lda, const T *__restrict__ B, int64_t ldb, T beta,
           T *__restrict__ C, int64_t ldc, int64_t tiles_m, int64_t tiles_n,
           int64_t per_xcd, int vec_ok) {
  using Cfg = GemmCfg<T, BM, BN, BK, WM, WN>;
  using vec_t = T __attribute__((ext_vector_type(Cfg::V)));
  using acc_t = typename Mfma<T>::acc_t;
  __shared__ __attribute__((aligned(16))) T smem[2 * Cfg::STAGE];

  const int64_t logical = map_block(blockIdx.x, per_xcd);
  if (logical >= tiles_m * tiles_n) return;  // whole workgroup leaves together
  const int64_t per_group = int64_t(kGroupM) * tiles_n;
  const int64_t first = logical / per_group * kGroupM;
  const int64_t gsz = tiles_m - first < kGroupM ? tiles_m - first : kGroupM;
  const int64_t in_g = logical % per_group;
  const int64_t m0 = (first + in_g % gsz) * BM;
  const int64_t n0 = (in_g / gsz) * BN;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, r = lane & 15;
  const int wm = wave / WN, wn = wave % WN;
  const bool inner_mn = vec_ok && m0 + BM <= m && n0 + BN <= n;

  vec_t ra[Cfg::A_VECS], rb[Cfg::B_VECS];
  auto load = [&](int64_t k0) {
    if (inner_mn && k0 + BK <= K) {
#pragma unroll
      for (int p = 0; p < Cfg::A_VECS; ++p) {
        const int q = tid + p * Cfg::kThreads;
        const int row = q / Cfg::A_TPR, kv = q % Cfg::A_TPR;
        ra[p] = *reinterpret_cast<const vec_t *>(A + (m0 + row) * lda + k0 + kv * Cfg::V);
      }
#pragma unroll
      for (int p = 0; p < Cfg::B_VECS; ++p) {
        const int q = tid + p * Cfg::kThreads;
        const int kr = q / Cfg::B_TPR, jv = q % Cfg::B_TPR;
        rb[p] = *reinterpret_cast<const vec_t *>(B + (k0 + kr) * ldb + n0 + jv * Cfg::V);
      }
    } else {  // edge tile, partial last stage or unaligned rows: per 16-byte slot
#pragma unroll
      for (int p = 0; p < Cfg::A_VECS; ++p) {
        const int q = tid + p * Cfg::kThreads;
        const int row = q / Cfg::A_TPR, kv = q % Cfg::A_TPR;
        const int64_t gi = m0 + row, gk = k0 + kv * Cfg::V;
        if (vec_ok && gi < m && gk + Cfg::V <= K) {
          ra[p] = *reinterpret_cast<const vec_t *>(A + gi * lda + gk);
        } else {
#pragma unroll
          for (int e = 0; e < Cfg::V; ++e)
            ra[p][e] = (gi < m && gk + e < K) ? A[gi * lda + gk + e] : T(0);
        }
      }
#pragma unroll
      for (int p = 0; p < Cfg::B_VECS; ++p) {
        const int q = tid + p * Cfg::kThreads;
        const int kr = q / Cfg::B_TPR, jv = q % Cfg::B_TPR;
        const int64_t gk = k0 + kr, gj = n0 + jv * Cfg::V;
        if (vec_ok && gk < K && gj + Cfg::V <= n) {
          rb[p] = *reinterpret_cast<const vec_t *>(B + gk * ldb + gj);
        } else {
#pragma unroll
          for (int e = 0; e < Cfg::V; ++e)
            rb[p][e] = (gk < K && gj + e < n) ? B[gk * ldb + gj + e] : T(0);
        }
      }
    }
  };
  auto store = [&](T *st) {
    T *As = st, *Bs = st + BM * Cfg::LDA_S;
#pragma unroll
    for (int p = 0; p < Cfg::A_VECS; ++p) {
      const int q = tid + p * Cfg::kThreads;
      *reinterpret_cast<vec_t *>(As + (q / Cfg::A_TPR) * Cfg::LDA_S +
                                 (q % Cfg::A_TPR) * Cfg::V) = ra[p];
    }
#pragma unroll
    for (int p = 0; p < Cfg::B_VECS; ++p) {
      const int q = tid + p * Cfg::kThreads;
      *reinterpret_cast<vec_t *>(Bs + (q / Cfg::B_TPR) * Cfg::LDB_S +
                                 (q % Cfg::B_TPR) * Cfg::V) = rb[p];
    }
  };

  acc_t acc[Cfg::FM][Cfg::FN];
#pragma unroll
  for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
    for (int j = 0; j < Cfg::FN; ++j) acc[i][j] = acc_t{};

  const int64_t nk = (K + BK - 1) / BK;
  if (nk > 0) {
    load(0);
    store(smem);
  }
  __syncthreads();
  for (int64_t kt = 0; kt < nk; ++kt) {
    const T *As = smem + (kt & 1) * Cfg::STAGE;
    const T *Bs = As + BM * Cfg::LDA_S;
    if (kt + 1 < nk) load((kt + 1) * BK);
#pragma unroll
    for (int s = 0; s < BK / Cfg::SK; ++s) {
      vec_t a[Cfg::FM];
      T b[Cfg::FN][Cfg::KK];
#pragma unroll
      for (int i = 0; i < Cfg::FM; ++i)
        a[i] = *reinterpret_cast<const vec_t *>(
            As + (wm * Cfg::TM + i * 16 + r) * Cfg::LDA_S + s * Cfg::SK + g * Cfg::KK);
#pragma unroll
      for (int j = 0; j < Cfg::FN; ++j)
#pragma unroll
        for (int kk = 0; kk < Cfg::KK; ++kk)
          b[j][kk] = Bs[(s * Cfg::SK + g * Cfg::KK + kk) * Cfg::LDB_S + wn * Cfg::TN + j * 16 + r];
#pragma unroll
      for (int kk = 0; kk < Cfg::KK; ++kk)
#pragma unroll
        for (int i = 0; i < Cfg::FM; ++i)
#pragma unroll
          for (int j = 0; j < Cfg::FN; ++j)
            acc[i][j] = Mfma<T>::op(a[i][kk], b[j][kk], acc[i][j]);
    }
    if (kt + 1 < nk) store(smem + ((kt + 1) & 1) * Cfg::STAGE);
    __syncthreads();
  }

  // Epilogue: element v of the lane holds D[Mfma<T>::row(g, v)][r] of each
  // MFMA tile. beta == 0 does not read C (so C may hold anything, NaN
  // included), as in BLAS.
#pragma unroll
  for (int i = 0; i < Cfg::FM; ++i) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t gi = m0 + wm * Cfg::TM + i * 16 + Mfma<T>::row(g, v);
      if (gi >= m) continue;
      T *crow = C + gi * ldc;
#pragma unroll
      for (int j = 0; j < Cfg::FN; ++j) {
        const int64_t gj = n0 + wn * Cfg::TN + j * 16 + r;
        if (gj >= n) continue;
        const T prod = alpha * acc[i][j][v];
        crow[gj] = beta == T(0) ? prod : prod + beta * crow[gj];
      }
    }
  }
}

template <typename T, int BM, int BN, int BK, int WM, int WN>
void launch_gemm(int64_t m, int64_t n, int64_t k, T alpha, const T *a, int64_t lda,
                 const T *b, int64_t ldb, T beta, T *c, int64_t ldc, hipStream_t s) {
  using Cfg = GemmCfg<T, BM, BN, BK, WM, WN>;
  const int64_t tiles_m = (m + BM - 1) / BM, tiles_n = (n + BN - 1) / BN;
  const int64_t per_xcd = (tiles_m * tiles_n + kXcds - 1) / kXcds;
  DRX_REQUIRE(per_xcd * kXcds <= int64_t(0xffffffff), "product too large for one launch");
  const bool aligned = reinterpret_cast<uintptr_t>(a) % 16 == 0 &&
                       reinterpret_cast<uintptr_t>(b) % 16 == 0 &&
                       lda % Cfg::V == 0 && ldb % Cfg::V == 0;
  hipLaunchKernelGGL((k_gemm<T, BM, BN, BK, WM, WN>),
                     dim3(static_cast<unsigned>(per_xcd * kXcds)), dim3(Cfg::kThreads), 0, s,
                     m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, tiles_m, tiles_n,
                     per_xcd, aligned ? 1 : 0);
  DRX_HIP(hipGetLastError());
}

// Workgroup tiles (gemm_tile knob):
//   1: 128 x 128, four waves of 64 x 64, BK = 16 (FP32) / 8 (FP64)
//   2: 256 x 128, eight waves of 64 x 64, same depth
//   3: 128 x 128, twice the depth          4: 256 x 128, twice the depth
//   5: 256 x 256, eight waves of 64 x 128 (FP32 only; FP64 takes 4)
// 0 picks by type and size (measured on 4096^3 .. 16384^3): FP64 takes 3;
// FP32 takes 2 once C holds 2048 tiles of 256 x 128 (about 8192^2), 5 while
// it holds at least 256 of 256 x 256 (about 4096^2), else 1.
inline int default_gemm_tile(bool fp64, int64_t m, int64_t n) {
  if (fp64) return 3;
  if (((m + 255) / 256) * ((n + 127) / 128) >= 2048) return 2;
  return ((m + 255) / 256) * ((n + 255) / 256) >= 256 ? 5 : 1;
}

template <typename T>
void run_gemm(int64_t m, int64_t n, int64_t k, T alpha, const T *a, int64_t lda,
              const T *b, int64_t ldb, T beta, T *c, int64_t ldc, hipStream_t s) {
  constexpr int BK = sizeof(T) == 4 ? 16 : 8;
  int t = tuned_gemm_tile();
  if (t == 0) t = default_gemm_tile(sizeof(T) == 8, m, n);
  switch (t) {
    case 1: launch_gemm<T, 128, 128, BK, 2, 2>(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s); break;
    case 2: launch_gemm<T, 256, 128, BK, 4, 2>(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s); break;
    case 3: launch_gemm<T, 128, 128, 2 * BK, 2, 2>(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s); break;
    case 5:
      if constexpr (sizeof(T) == 4) {
        launch_gemm<T, 256, 256, BK, 4, 2>(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s);
        break;
      }
      [[fallthrough]];
    default: launch_gemm<T, 256, 128, 2 * BK, 4, 2>(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s); break;
  }
}

}  // namespace
}  // namespace drx

using namespace drx;

extern "C" int drx_gemm(int vtype, int64_t m, int64_t n, int64_t k, double alpha,
                        const void *a, int64_t lda, const void *b, int64_t ldb,
                        double beta, void *c, int64_t ldc, drx_stream_t stream) {
  return guarded([&] {
    DRX_REQUIRE(m >= 0 && n >= 0 && k >= 0, "negative dimension");
    DRX_REQUIRE(vtype == DRX_F32 || vtype == DRX_F64, "vtype must be DRX_F32 or DRX_F64");
    DRX_REQUIRE(lda >= std::max<int64_t>(k, 1) && ldb >= std::max<int64_t>(n, 1) &&
                    ldc >= std::max<int64_t>(n, 1),
                "leading dimension below the row length");
    if (m == 0 || n == 0) return;
    DRX_REQUIRE(c != nullptr && (k == 0 || (a != nullptr && b != nullptr)), "null operand");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (vtype == DRX_F32)
      run_gemm<float>(m, n, k, static_cast<float>(alpha), static_cast<const float *>(a), lda,
                      static_cast<const float *>(b), ldb, static_cast<float>(beta),
                      static_cast<float *>(c), ldc, s);
    else
      run_gemm<double>(m, n, k, alpha, static_cast<const double *>(a), lda,
                       static_cast<const double *>(b), ldb, beta, static_cast<double *>(c),
                       ldc, s);
  });
}
