// drx_common.hpp -- shared device helpers for the gfx950 sparse gemv kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>

#include "drx_spmv.h"

namespace drx {

// ---------------------------------------------------------------------------
// Errors: every C entry point returns a status and never throws; the message
// of the last failure on the calling thread is kept for drx_last_error().
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

void set_last_error(const char *msg);

#define DRX_REQUIRE(cond, msg)                                                 \
  do {                                                                         \
    if (!(cond)) throw ::drx::Error(DRX_ERR_INVALID, msg);                     \
  } while (0)

#define DRX_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      throw ::drx::Error(e_ == hipErrorOutOfMemory ? DRX_ERR_NOMEM             \
                                                   : DRX_ERR_HIP,              \
                         std::string(#call) + ": " + hipGetErrorString(e_));   \
  } while (0)

template <typename F> int guarded(F &&f) {
  try {
    f();
    return DRX_OK;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.status;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return DRX_ERR_NOMEM;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return DRX_ERR_INVALID;
  } catch (...) {
    set_last_error("unknown error");
    return DRX_ERR_INVALID;
  }
}

// ---------------------------------------------------------------------------
// Storage types
// ---------------------------------------------------------------------------
struct bf16 {
  uint16_t bits;
};
using f16 = _Float16;

__device__ __forceinline__ float to_f32(bf16 v) {
  return __uint_as_float(static_cast<uint32_t>(v.bits) << 16);
}
__device__ __forceinline__ float to_f32(f16 v) { return static_cast<float>(v); }
__device__ __forceinline__ float to_f32(float v) { return v; }

template <typename TA, typename T>
__device__ __forceinline__ TA to_acc(T v) {
  if constexpr (std::is_same_v<T, bf16> || std::is_same_v<T, f16>)
    return static_cast<TA>(to_f32(v));
  else
    return static_cast<TA>(v);
}

// Round-to-nearest-even float -> bf16 (NaN kept quiet).
__device__ __forceinline__ bf16 f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu))
    return bf16{static_cast<uint16_t>((u >> 16) | 0x40)};
  u += 0x7fffu + ((u >> 16) & 1u);
  return bf16{static_cast<uint16_t>(u >> 16)};
}

template <typename TY, typename TA>
__device__ __forceinline__ TY from_acc(TA v) {
  if constexpr (std::is_same_v<TY, bf16>)
    return f32_to_bf16(static_cast<float>(v));
  else if constexpr (std::is_same_v<TY, f16>)
    return static_cast<f16>(static_cast<float>(v));
  else
    return static_cast<TY>(v);
}

// Streaming (read-once) loads of the matrix arrays: non-temporal, so that the
// L2 keeps the gathered x instead of vals/colind.
template <typename T> __device__ __forceinline__ T ld_stream(const T *p) {
  if constexpr (sizeof(T) == 4 || sizeof(T) == 8)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}

// ---------------------------------------------------------------------------
// Lane-group sums with DPP (within 16-lane rows) and cross-row shuffles.
// G lanes, aligned to G; every lane of the group ends with the group sum.
// ---------------------------------------------------------------------------
template <int CTRL> __device__ __forceinline__ float dpp_f32(float v) {
  return __int_as_float(
      __builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL> __device__ __forceinline__ double dpp_f64(double v) {
  int64_t b = __double_as_longlong(v);
  int lo = static_cast<int>(b & 0xffffffff);
  int hi = static_cast<int>(b >> 32);
  lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xf, 0xf, false);
  hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xf, 0xf, false);
  return __longlong_as_double((static_cast<int64_t>(hi) << 32) |
                              static_cast<uint32_t>(lo));
}
template <int CTRL, typename T> __device__ __forceinline__ T dpp(T v) {
  if constexpr (std::is_same_v<T, double>)
    return dpp_f64<CTRL>(v);
  else
    return dpp_f32<CTRL>(v);
}

// DPP controls (GFX9 encoding).
constexpr int kQuadXor1 = 0xB1;      // quad_perm [1,0,3,2]
constexpr int kQuadXor2 = 0x4E;      // quad_perm [2,3,0,1]
constexpr int kRowHalfMirror = 0x141;  // lane i <- 7-i within 8
constexpr int kRowMirror = 0x140;      // lane i <- 15-i within 16

template <int G, typename T> __device__ __forceinline__ T group_sum(T v) {
  if constexpr (G >= 2) v += dpp<kQuadXor1>(v);
  if constexpr (G >= 4) v += dpp<kQuadXor2>(v);
  if constexpr (G >= 8) v += dpp<kRowHalfMirror>(v);
  if constexpr (G >= 16) v += dpp<kRowMirror>(v);
  if constexpr (G >= 32) v += __shfl_xor(v, 16, 64);
  if constexpr (G >= 64) v += __shfl_xor(v, 32, 64);
  return v;
}

// Wave-wide minimum and maximum of a 32-bit value: DPP inside 16-lane rows,
// then the four rows' results through readlane (scalar, wave-uniform).
template <int CTRL> __device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
  return static_cast<uint32_t>(
      __builtin_amdgcn_mov_dpp(static_cast<int>(v), CTRL, 0xf, 0xf, false));
}
__device__ __forceinline__ void wave_min_max(uint32_t &lo, uint32_t &hi) {
  uint32_t t;
  t = dpp_u32<kQuadXor1>(lo); lo = t < lo ? t : lo;
  t = dpp_u32<kQuadXor1>(hi); hi = t > hi ? t : hi;
  t = dpp_u32<kQuadXor2>(lo); lo = t < lo ? t : lo;
  t = dpp_u32<kQuadXor2>(hi); hi = t > hi ? t : hi;
  t = dpp_u32<kRowHalfMirror>(lo); lo = t < lo ? t : lo;
  t = dpp_u32<kRowHalfMirror>(hi); hi = t > hi ? t : hi;
  t = dpp_u32<kRowMirror>(lo); lo = t < lo ? t : lo;
  t = dpp_u32<kRowMirror>(hi); hi = t > hi ? t : hi;
  uint32_t l = 0xffffffffu, h = 0;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const uint32_t lr = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(lo), 16 * r));
    const uint32_t hr = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(hi), 16 * r));
    l = lr < l ? lr : l;
    h = hr > h ? hr : h;
  }
  lo = l;
  hi = h;
}

// Smallest and largest column index of a workgroup's window over the slots
// u (c[u] for u < T) for which in_block(u) holds. xmm holds 2 * WG / 64
// words of LDS. Ends after a barrier with the same lo, hi in every lane (lo >
// hi when the window is empty).
template <int WG, int T, typename F>
__device__ __forceinline__ void block_col_range(const uint32_t (&c)[T], F in_block, uint32_t *xmm,
                                                uint32_t &lo, uint32_t &hi) {
  const int tid = threadIdx.x;
  lo = 0xffffffffu;
  hi = 0;
#pragma unroll
  for (int u = 0; u < T; u++) {
    if (in_block(u)) {
      lo = c[u] < lo ? c[u] : lo;
      hi = c[u] > hi ? c[u] : hi;
    }
  }
  wave_min_max(lo, hi);
  if ((tid & 63) == 0) {
    xmm[tid >> 6] = lo;
    xmm[WG / 64 + (tid >> 6)] = hi;
  }
  __syncthreads();
#pragma unroll
  for (int w = 0; w < WG / 64; w++) {
    lo = xmm[w] < lo ? xmm[w] : lo;
    hi = xmm[WG / 64 + w] > hi ? xmm[WG / 64 + w] : hi;
  }
}

// The same over lane entries tid + u * WG < n.
template <int WG, int T>
__device__ __forceinline__ void block_col_range(const uint32_t (&c)[T], int n, uint32_t *xmm,
                                                uint32_t &lo, uint32_t &hi) {
  block_col_range<WG, T>(c, [&](int u) { return static_cast<int>(threadIdx.x) + u * WG < n; },
                         xmm, lo, hi);
}

// ---------------------------------------------------------------------------
// Row spans of a CSR share. rowptr holds absolute offsets; base is the
// share's first nnz; the row's span is clamped into [0, nnz). Row `nrows` is
// the end: rowptr[nrows] with a sentinel, else nnz.
// ---------------------------------------------------------------------------
template <typename TI>
__device__ __forceinline__ int64_t row_begin(const TI *rowptr, int64_t base,
                                             int64_t nrows, int64_t nnz,
                                             int has_sentinel, int64_t r) {
  if (r < nrows || has_sentinel) {
    int64_t v = static_cast<int64_t>(rowptr[r]) - base;
    return v < 0 ? 0 : (v > nnz ? nnz : v);
  }
  return nnz;
}

// Top bit of a column panel's row-map entry: the segment is its row's first
// in panel order (row ids are non-negative, so the bit is free).
template <typename TI> constexpr TI kFirstSeg = static_cast<TI>(TI(1) << (sizeof(TI) * 8 - 1));

// Entries a block's window leaves free: a window of W entries holds at most
// W - kWinSlack of the block's entries, so that lanes streaming groups of up
// to kWinSlack consecutive entries from the group-aligned entry below the
// block's first cover the whole block within W slots.
constexpr int kWinSlack = 4;

// Number of XCDs that a launch's workgroups are dealt over round-robin.
constexpr int kXcds = 8;

// Workgroup i (dealt to XCD i % 8) takes logical block
// (i % 8) * per_xcd + i / 8, so each XCD walks one contiguous range of rows.
__device__ __forceinline__ int64_t map_block(int64_t i, int64_t per_xcd) {
  return (i % kXcds) * per_xcd + i / kXcds;
}

// ---------------------------------------------------------------------------
// Buffer addressing: a wave-uniform 128-bit resource (base, size) and 32-bit
// per-lane byte offsets. Offsets at or past `bytes` read 0 (hardware range
// check) instead of touching memory.
// ---------------------------------------------------------------------------
using rsrc_t = __amdgpu_buffer_rsrc_t;
constexpr int kRsrcWord3 = 0x00020000;  // raw buffer, gfx9 family
constexpr int kNT = 2;                   // non-temporal (streamed once)

__device__ __forceinline__ rsrc_t make_rsrc(const void *base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0,
                                           static_cast<int>(bytes), kRsrcWord3);
}

template <typename T, int AUX = 0>
__device__ __forceinline__ T buf_load(rsrc_t r, uint32_t off) {
  if constexpr (sizeof(T) == 16)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, AUX));
  else if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, AUX));
  else if constexpr (sizeof(T) == 4)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, AUX));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, AUX));
}

// N consecutive 4-byte words (one buffer load of 4 N bytes).
template <int N> struct Words {
  uint32_t w[N];
};

// Column index as a 32-bit element offset (the kernels that use buffer
// gathers require ncols * sizeof(x) < 2^32).
template <typename TI> __device__ __forceinline__ uint32_t col32(TI c) {
  return static_cast<uint32_t>(c);
}

}  // namespace drx
