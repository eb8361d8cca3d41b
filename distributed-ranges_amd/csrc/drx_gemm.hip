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

// Direct-to-LDS variant for FP32, BK = 16, BN = 128 (A stage rows of 64
// bytes, B stage rows of 512 bytes). global_load_lds_dwordx4 writes lane i of
// a wave's instruction to 16 bytes at (wave's 1 KB block) + 16 i, so the LDS
// image is unpadded and each lane picks the global slot that belongs at its
// position. Slot c of a 16-row A block's row r sits at c ^ fa(r); slot c of B
// stage row k at c ^ hb(k). Both swizzles keep the MFMA operand reads free of
// bank conflicts (16 lanes per 256 bytes or 8 per 128 bytes for the 16-byte A
// reads; 64 lanes per 256 bytes for the 4-byte B reads). Full stages load
// this way, with no register round trip; rows past m read row m - 1 instead
// (their products are never stored). Tiles on the n edge and a partial last
// stage load through registers into the same positions, after the stage's
// products, so that those registers are never live beside the operands.
__device__ __forceinline__ int dl_fa(int r) { return ((r >> 1) ^ (r >> 3)) & 3; }
// One 16-byte global -> LDS copy per lane into the wave's 1 KB block at lds.
__device__ __forceinline__ void lds_dma16(const void *src, char *lds) {
  __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void *)lds, 16, 0, 0);
}
__device__ __forceinline__ int dl_hb(int k) { return 4 * ((k >> 2) & 3); }

// Register budget: 3 waves per SIMD (the kernel needs about 136 VGPRs; at
// 128 it spills).
template <int BM, int WM, int WN>
__global__ void __launch_bounds__(64 * WM * WN) __attribute__((amdgpu_waves_per_eu(3)))
    k_gemm_dl(int64_t m, int64_t n, int64_t K, float alpha, const float *__restrict__ A,
              int64_t lda, const float *__restrict__ B, int64_t ldb, float beta,
              float *__restrict__ C, int64_t ldc, int64_t tiles_m, int64_t tiles_n,
              int64_t per_xcd, int vec_ok) {
  constexpr int BN = 128, BK = 16, NW = WM * WN;
  constexpr int A_BYTES = BM * BK * 4, B_BYTES = BK * BN * 4, STAGE = A_BYTES + B_BYTES;
  constexpr int A_INSTR = A_BYTES / 1024 / NW, B_INSTR = B_BYTES / 1024 / NW;
  constexpr int TM = BM / WM, TN = BN / WN, FM = TM / 16, FN = TN / 16;
  static_assert(A_INSTR * NW * 1024 == A_BYTES && B_INSTR * NW * 1024 == B_BYTES,
                "stages must split into whole 1 KB wave blocks");
  static_assert(TM % 16 == 0 && TN % 16 == 0, "wave tile must be whole MFMA tiles");
  __shared__ __attribute__((aligned(1024))) char smem[2 * STAGE];

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
  const bool n_inner = vec_ok && n0 + BN <= n;

  // The (row, slot) each lane moves: A block p * NW + wave, row block*16 +
  // lane/4, slot (lane%4) ^ fa(lane/4); B block p * NW + wave, row
  // 2 block + lane/32, slot (lane%32) ^ hb(row).
  auto a_row = [&](int p) { return (p * NW + wave) * 16 + (lane >> 2); };
  auto a_col = [&](int) { return 4 * ((lane & 3) ^ dl_fa(lane >> 2)); };
  auto b_row = [&](int p) { return (p * NW + wave) * 2 + (lane >> 5); };
  auto b_col = [&](int p) { return 4 * ((lane & 31) ^ dl_hb(b_row(p))); };
  const float *a_src[A_INSTR];
  const float *b_src[B_INSTR];
#pragma unroll
  for (int p = 0; p < A_INSTR; ++p) {
    const int64_t gi = m0 + a_row(p) < m ? m0 + a_row(p) : m - 1;
    a_src[p] = A + gi * lda + a_col(p);
  }
#pragma unroll
  for (int p = 0; p < B_INSTR; ++p) b_src[p] = B + b_row(p) * ldb + n0 + b_col(p);

  auto issue_direct = [&](int64_t k0, char *st) {
#pragma unroll
    for (int p = 0; p < A_INSTR; ++p)
      lds_dma16(a_src[p] + k0, st + (p * NW + wave) * 1024);
#pragma unroll
    for (int p = 0; p < B_INSTR; ++p)
      lds_dma16(b_src[p] + k0 * ldb, st + A_BYTES + (p * NW + wave) * 1024);
  };

  f32x4 ra[A_INSTR], rb[B_INSTR];
  auto load_regs = [&](int64_t k0) {
#pragma unroll
    for (int p = 0; p < A_INSTR; ++p) {
      const int64_t gi = m0 + a_row(p), gk = k0 + a_col(p);
      if (vec_ok && gi < m && gk + 4 <= K) {
        ra[p] = *reinterpret_cast<const f32x4 *>(A + gi * lda + gk);
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) ra[p][e] = (gi < m && gk + e < K) ? A[gi * lda + gk + e] : 0.f;
      }
    }
#pragma unroll
    for (int p = 0; p < B_INSTR; ++p) {
      const int64_t gk = k0 + b_row(p), gj = n0 + b_col(p);
      if (vec_ok && gk < K && gj + 4 <= n) {
        rb[p] = *reinterpret_cast<const f32x4 *>(B + gk * ldb + gj);
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) rb[p][e] = (gk < K && gj + e < n) ? B[gk * ldb + gj + e] : 0.f;
      }
    }
  };
  auto store_regs = [&](char *st) {
#pragma unroll
    for (int p = 0; p < A_INSTR; ++p)
      *reinterpret_cast<f32x4 *>(st + (p * NW + wave) * 1024 + lane * 16) = ra[p];
#pragma unroll
    for (int p = 0; p < B_INSTR; ++p)
      *reinterpret_cast<f32x4 *>(st + A_BYTES + (p * NW + wave) * 1024 + lane * 16) = rb[p];
  };

  // Operand offsets (bytes within a stage) of this lane. Row block i of the
  // wave's A tile is a_off + 1024 i; B row 4 g + kk of column tile j is
  // b_off[j] + 512 kk (hb is the same for the four rows of lane group g).
  const int a_off = wm * TM / 16 * 1024 + r * 64 + (g ^ dl_fa(r)) * 16;
  int b_off[FN];
#pragma unroll
  for (int j = 0; j < FN; ++j) {
    const int col = wn * TN + j * 16 + r;
    b_off[j] = A_BYTES + 4 * g * 512 + (((col >> 2) ^ dl_hb(4 * g)) * 16) + (col & 3) * 4;
  }

  f32x4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = f32x4{};

  const int64_t nk = (K + BK - 1) / BK;
  if (nk > 0) {
    if (n_inner && BK <= K) {
      issue_direct(0, smem);
    } else {
      load_regs(0);
      store_regs(smem);
    }
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  for (int64_t kt = 0; kt < nk; ++kt) {
    const char *cur = smem + (kt & 1) * STAGE;
    char *nxt = smem + ((kt + 1) & 1) * STAGE;
    const bool more = kt + 1 < nk;
    const bool direct = more && n_inner && (kt + 2) * BK <= K;
    if (direct) issue_direct((kt + 1) * BK, nxt);
    f32x4 a[FM];
#pragma unroll
    for (int i = 0; i < FM; ++i) a[i] = *reinterpret_cast<const f32x4 *>(cur + a_off + i * 1024);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      float b[FN];
#pragma unroll
      for (int j = 0; j < FN; ++j) b[j] = *reinterpret_cast<const float *>(cur + b_off[j] + kk * 512);
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = Mfma<float>::op(a[i][kk], b[j], acc[i][j]);
    }
    if (more && !direct) {
      load_regs((kt + 1) * BK);
      store_regs(nxt);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < FM; ++i) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t gi = m0 + wm * TM + i * 16 + Mfma<float>::row(g, v);
      if (gi >= m) continue;
      float *crow = C + gi * ldc;
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int64_t gj = n0 + wn * TN + j * 16 + r;
        if (gj >= n) continue;
        const float prod = alpha * acc[i][j][v];
        crow[gj] = beta == 0.f ? prod : prod + beta * crow[gj];
      }
    }
  }
}

// FP64 direct-to-LDS variant, BK = 16, BN = 128: A stage rows of 128 bytes
// (8 slots; a wave's 1 KB block holds 8 rows), B stage rows of 1 KB (one per
// block). Slot c of A row r (r mod 16 within its MFMA tile) sits at
// c ^ fa64(r); slot c of B row k at c ^ hb64(k). The 16-byte A reads (two
// k per lane) and the 8-byte B reads stay free of bank conflicts for 8 or
// 16 lanes per 128 or 256 bytes.
__device__ __forceinline__ int dl64_fa(int r) { return (r ^ (r >> 3)) & 7; }
__device__ __forceinline__ int dl64_hb(int k) { return 8 * ((k >> 1) & 1); }

template <int BM, int WM, int WN>
__global__ void __launch_bounds__(64 * WM * WN) __attribute__((amdgpu_waves_per_eu(2)))
    k_gemm_dl64(int64_t m, int64_t n, int64_t K, double alpha, const double *__restrict__ A,
                int64_t lda, const double *__restrict__ B, int64_t ldb, double beta,
                double *__restrict__ C, int64_t ldc, int64_t tiles_m, int64_t tiles_n,
                int64_t per_xcd, int vec_ok) {
  typedef double f64x2 __attribute__((ext_vector_type(2)));
  constexpr int BN = 128, BK = 16, NW = WM * WN;
  constexpr int A_BYTES = BM * BK * 8, B_BYTES = BK * BN * 8, STAGE = A_BYTES + B_BYTES;
  constexpr int A_INSTR = A_BYTES / 1024 / NW, B_INSTR = B_BYTES / 1024 / NW;
  constexpr int TM = BM / WM, TN = BN / WN, FM = TM / 16, FN = TN / 16;
  static_assert(A_INSTR * NW * 1024 == A_BYTES && B_INSTR * NW * 1024 == B_BYTES,
                "stages must split into whole 1 KB wave blocks");
  static_assert(TM % 16 == 0 && TN % 16 == 0, "wave tile must be whole MFMA tiles");
  __shared__ __attribute__((aligned(1024))) char smem[2 * STAGE];

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
  const bool n_inner = vec_ok && n0 + BN <= n;

  // A block p * NW + wave: rows block*8 + lane/8, slot (lane%8) ^ fa64(row);
  // B block p * NW + wave: row = block, slot lane ^ hb64(row).
  auto a_row = [&](int p) { return (p * NW + wave) * 8 + (lane >> 3); };
  auto a_col = [&](int p) { return 2 * ((lane & 7) ^ dl64_fa(a_row(p) & 15)); };
  auto b_row = [&](int p) { return p * NW + wave; };
  auto b_col = [&](int p) { return 2 * (lane ^ dl64_hb(b_row(p))); };
  const double *a_src[A_INSTR];
  const double *b_src[B_INSTR];
#pragma unroll
  for (int p = 0; p < A_INSTR; ++p) {
    const int64_t gi = m0 + a_row(p) < m ? m0 + a_row(p) : m - 1;
    a_src[p] = A + gi * lda + a_col(p);
  }
#pragma unroll
  for (int p = 0; p < B_INSTR; ++p) b_src[p] = B + b_row(p) * ldb + n0 + b_col(p);

  auto issue_direct = [&](int64_t k0, char *st) {
#pragma unroll
    for (int p = 0; p < A_INSTR; ++p) lds_dma16(a_src[p] + k0, st + (p * NW + wave) * 1024);
#pragma unroll
    for (int p = 0; p < B_INSTR; ++p)
      lds_dma16(b_src[p] + k0 * ldb, st + A_BYTES + (p * NW + wave) * 1024);
  };
  auto load_store_regs = [&](int64_t k0, char *st) {
#pragma unroll
    for (int p = 0; p < A_INSTR; ++p) {
      const int64_t gi = m0 + a_row(p), gk = k0 + a_col(p);
      f64x2 v;
      if (vec_ok && gi < m && gk + 2 <= K) {
        v = *reinterpret_cast<const f64x2 *>(A + gi * lda + gk);
      } else {
#pragma unroll
        for (int e = 0; e < 2; ++e) v[e] = (gi < m && gk + e < K) ? A[gi * lda + gk + e] : 0.0;
      }
      *reinterpret_cast<f64x2 *>(st + (p * NW + wave) * 1024 + lane * 16) = v;
    }
#pragma unroll
    for (int p = 0; p < B_INSTR; ++p) {
      const int64_t gk = k0 + b_row(p), gj = n0 + b_col(p);
      f64x2 v;
      if (vec_ok && gk < K && gj + 2 <= n) {
        v = *reinterpret_cast<const f64x2 *>(B + gk * ldb + gj);
      } else {
#pragma unroll
        for (int e = 0; e < 2; ++e) v[e] = (gk < K && gj + e < n) ? B[gk * ldb + gj + e] : 0.0;
      }
      *reinterpret_cast<f64x2 *>(st + A_BYTES + (p * NW + wave) * 1024 + lane * 16) = v;
    }
  };

  // Operand offsets (bytes within a stage). A row block i of the wave at
  // a_off[s] + 2048 i for sub-step s (slot 4 s + g, i.e. bit 2 flipped for
  // s = 1); B row 8 s + 2 g + kk of column tile j at b_off[j] + 1024 (8 s + kk)
  // (hb64 is the same for all of lane group g's rows).
  const int a_base = (wm * TM + r) * 128;
  const int a_off0 = a_base + ((g ^ dl64_fa(r)) * 16);
  const int a_off1 = a_base + (((g ^ dl64_fa(r)) ^ 4) * 16);
  int b_off[FN];
#pragma unroll
  for (int j = 0; j < FN; ++j) {
    const int col = wn * TN + j * 16 + r;
    b_off[j] = A_BYTES + 2 * g * 1024 + (((col >> 1) ^ dl64_hb(2 * g)) * 16) + (col & 1) * 8;
  }

  f64x4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = f64x4{};

  const int64_t nk = (K + BK - 1) / BK;
  if (nk > 0) {
    if (n_inner && BK <= K)
      issue_direct(0, smem);
    else
      load_store_regs(0, smem);
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  for (int64_t kt = 0; kt < nk; ++kt) {
    const char *cur = smem + (kt & 1) * STAGE;
    char *nxt = smem + ((kt + 1) & 1) * STAGE;
    const bool more = kt + 1 < nk;
    const bool direct = more && n_inner && (kt + 2) * BK <= K;
    if (direct) issue_direct((kt + 1) * BK, nxt);
#pragma unroll
    for (int s2 = 0; s2 < 2; ++s2) {
      f64x2 a[FM];
#pragma unroll
      for (int i = 0; i < FM; ++i)
        a[i] = *reinterpret_cast<const f64x2 *>(cur + (s2 ? a_off1 : a_off0) + i * 2048);
#pragma unroll
      for (int kk = 0; kk < 2; ++kk) {
        double b[FN];
#pragma unroll
        for (int j = 0; j < FN; ++j)
          b[j] = *reinterpret_cast<const double *>(cur + b_off[j] + (8 * s2 + kk) * 1024);
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) acc[i][j] = Mfma<double>::op(a[i][kk], b[j], acc[i][j]);
      }
    }
    if (more && !direct) load_store_regs((kt + 1) * BK, nxt);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < FM; ++i) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t gi = m0 + wm * TM + i * 16 + Mfma<double>::row(g, v);
      if (gi >= m) continue;
      double *crow = C + gi * ldc;
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int64_t gj = n0 + wn * TN + j * 16 + r;
        if (gj >= n) continue;
        const double prod = alpha * acc[i][j][v];
        crow[gj] = beta == 0.0 ? prod : prod + beta * crow[gj];
      }
    }
  }
}

inline void launch_gemm_dl64(int64_t m, int64_t n, int64_t k, double alpha, const double *a,
                             int64_t lda, const double *b, int64_t ldb, double beta, double *c,
                             int64_t ldc, hipStream_t s) {
  const int64_t tiles_m = (m + 127) / 128, tiles_n = (n + 127) / 128;
  const int64_t per_xcd = (tiles_m * tiles_n + kXcds - 1) / kXcds;
  DRX_REQUIRE(per_xcd * kXcds <= int64_t(0xffffffff), "product too large for one launch");
  const bool aligned = reinterpret_cast<uintptr_t>(a) % 16 == 0 &&
                       reinterpret_cast<uintptr_t>(b) % 16 == 0 && lda % 2 == 0 && ldb % 2 == 0;
  hipLaunchKernelGGL((k_gemm_dl64<128, 2, 2>), dim3(static_cast<unsigned>(per_xcd * kXcds)),
                     dim3(256), 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, tiles_m,
                     tiles_n, per_xcd, aligned ? 1 : 0);
  DRX_HIP(hipGetLastError());
}

template <int BM, int WM, int WN>
void launch_gemm_dl(int64_t m, int64_t n, int64_t k, float alpha, const float *a, int64_t lda,
                    const float *b, int64_t ldb, float beta, float *c, int64_t ldc,
                    hipStream_t s) {
  const int64_t tiles_m = (m + BM - 1) / BM, tiles_n = (n + 127) / 128;
  const int64_t per_xcd = (tiles_m * tiles_n + kXcds - 1) / kXcds;
  DRX_REQUIRE(per_xcd * kXcds <= int64_t(0xffffffff), "product too large for one launch");
  const bool aligned = reinterpret_cast<uintptr_t>(a) % 16 == 0 &&
                       reinterpret_cast<uintptr_t>(b) % 16 == 0 && lda % 4 == 0 && ldb % 4 == 0;
  hipLaunchKernelGGL((k_gemm_dl<BM, WM, WN>), dim3(static_cast<unsigned>(per_xcd * kXcds)),
                     dim3(64 * WM * WN), 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                     tiles_m, tiles_n, per_xcd, aligned ? 1 : 0);
  DRX_HIP(hipGetLastError());
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
//   6, 7: 128 x 128 and 256 x 128 loading stages straight into LDS (FP32;
//      FP64 has the 128 x 128 one, BK 16, and takes 4 for 7)
// 0 picks by type and size (measured on 4096^3 .. 16384^3): FP64 takes 6
// (its direct-to-LDS kernel: 64-68 TFLOP/s against 57-65 for 3); FP32 takes
// 6 once C holds 4096 tiles of 128 x 128 (about 8192^2), 5 while it holds at
// least 256 of 256 x 256 (about 4096^2), else 1.
inline int default_gemm_tile(bool fp64, int64_t m, int64_t n) {
  if (fp64) return 6;
  if (((m + 127) / 128) * ((n + 127) / 128) >= 4096) return 6;
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
    case 6:
    case 7:
      if constexpr (sizeof(T) == 4) {
        if (t == 6)
          launch_gemm_dl<128, 2, 2>(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s);
        else
          launch_gemm_dl<256, 4, 2>(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s);
      } else if (t == 6) {
        launch_gemm_dl64(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s);
      } else {
        launch_gemm<T, 256, 128, 2 * BK, 4, 2>(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s);
      }
      break;
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
