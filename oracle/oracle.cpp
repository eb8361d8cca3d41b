// oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of Distributed Ranges' sparse gemv path and an extended-
// precision oracle. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this library, and only as the checker (or, for the
// cpu_baseline leg, as the timed CPU restatement of the reference's CPU path).
// The product (libdrx_spmv.so) never calls into it.
//
// Citations are to files of oneapi-src/distributed-ranges:
//   include/dr/mp/containers/matrix_formats/csr_row_distribution.hpp  (csr_row)
//   include/dr/mp/containers/matrix_formats/csr_eq_distribution.hpp   (csr_eq)
//   include/dr/mp/containers/distributed_vector.hpp                   (dv)
//   include/dr/mp/algorithms/matrix/gemv.hpp                          (gemv)
//   include/dr/detail/generate_csr.hpp                                (gen)
//   include/dr/detail/matrix_io.hpp                                   (io)
//
// Pinning (what checks each function; the tests live in tests/test_oracle.py):
//   orc_gen_random_csr_*      bit-exact against fixtures written by the
//                             reference's own generator compiled from source
//                             (oracle/_ref/ref_gen) and against the FNV-1a64
//                             hashes recorded for the reference generator.
//   orc_gen_band_csr_*        bit-exact against the same fixtures; closed form
//                             nnz formula; all-ones row counts.
//   orc_row_partition/_eq_partition
//                             brute-force partition invariants (every nnz in
//                             exactly one share, shares contiguous) and the
//                             reference's own formulas re-derived by hand for
//                             small P; ragged and empty cases.
//   orc_mp_gemv_{row,eq}      known answers recorded for the reference tests
//                             (row sums c[0..2], sum of the width-20 product),
//                             dense brute force, scipy.sparse, closed forms
//                             (Laplacian with x = 1), linearity, P-invariance.
//   orc_spmv_ref / orc_spmv_t_ref / orc_dvals_ref
//                             dense brute force (numpy), scipy.sparse, the
//                             dot-product identity <Ax,w> = <x,A^T w>, symmetric
//                             matrices (A^T = A), x = e_j returns column j.
//   orc_csr_transpose         scipy.sparse .T.tocsr() structure (bit-exact).
//   orc_gemveq_cpu_bench      same output as orc_mp_gemv_eq (bit-exact) for
//                             every P; it is the same arithmetic, threaded.
//   orc_gemm_ref              closed form of the sp gemm example's inputs
//                             (a = (i + k) / m, b = (k + j) / k), identity
//                             operands, numpy's float64 matmul, alpha/beta
//                             linearity.
#include <algorithm>
#include <atomic>
#include <barrier>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <random>
#include <thread>
#include <type_traits>
#include <unordered_set>
#include <utility>
#include <vector>

#include <pthread.h>
#include <sched.h>

namespace {

// ---------------------------------------------------------------------------
// Partitions
// ---------------------------------------------------------------------------

// distributed_vector segment size with the default distribution (granularity
// 1, no halo): dv:310-311  segment_size = gran * max(ceil(size/gran/P), 0, 0).
inline std::size_t dv_segment_size(std::size_t size, std::size_t P) {
  return (size + P - 1) / P;
}

// csr_row::init, csr_row:230-262. Rank i owns rows [i*seg, min((i+1)*seg, m)),
// nnz [val_offsets[i], val_offsets[i] + val_sizes[i]).
template <typename I>
void row_partition(std::size_t m, std::size_t nnz, const I *rowptr,
                   std::size_t P, std::size_t *seg_out,
                   std::size_t *val_offsets, std::size_t *val_sizes) {
  std::size_t seg = dv_segment_size(m, P);
  for (std::size_t i = 0; i < P; i++) {
    std::size_t first_index = i * seg;  // rows_data_->get_segment_offset(i)
    if (first_index > m) {              // csr_row:237-241
      val_offsets[i] = nnz;
      val_sizes[i] = 0;
      continue;
    }
    std::size_t lower_limit = static_cast<std::size_t>(rowptr[first_index]);
    std::size_t higher_limit = nnz;
    if ((i + 1) * seg < m) {  // csr_row:244-247
      higher_limit = static_cast<std::size_t>(rowptr[(i + 1) * seg]);
    }
    val_offsets[i] = lower_limit;
    val_sizes[i] = higher_limit - lower_limit;
  }
  *seg_out = seg;
}

// std::upper_bound(rowptr, rowptr + m, value) - rowptr with the reference's
// mixed size_t / I comparison (value < element).
template <typename I>
std::size_t upper_bound_idx(const I *rowptr, std::size_t m, std::size_t value) {
  auto it = std::upper_bound(rowptr, rowptr + m, value,
                             [](std::size_t v, const I &e) {
                               return v < static_cast<std::size_t>(e);
                             });
  return static_cast<std::size_t>(it - rowptr);
}

// csr_eq::init, csr_eq:248-284. Rank i owns nnz [i*seg, min((i+1)*seg, nnz))
// and rows [row_offsets[i], row_offsets[i] + row_sizes[i]).
template <typename I>
void eq_partition(std::size_t m, std::size_t nnz, const I *rowptr,
                  std::size_t P, std::size_t *seg_out, std::size_t *row_offsets,
                  std::size_t *row_sizes, std::size_t *max_row_size) {
  std::size_t seg = dv_segment_size(nnz, P);  // vals_data_->segment_size()
  std::size_t total = 0;
  for (std::size_t i = 0; i < P; i++) {
    std::size_t first_index = i * seg;
    std::size_t last_index = (i + 1) * seg - 1;  // wraps when seg == 0
    std::size_t lower_limit = upper_bound_idx(rowptr, m, first_index) - 1;
    std::size_t higher_limit = upper_bound_idx(rowptr, m, last_index);
    row_offsets[i] = lower_limit;
    row_sizes[i] = higher_limit - lower_limit;
    total += row_sizes[i];  // csr_eq:271 ("max_row_size_" is the sum)
  }
  *seg_out = seg;
  *max_row_size = total;
}

// ---------------------------------------------------------------------------
// Per-rank CPU loops
// ---------------------------------------------------------------------------

// csr_row::local_gemv CPU branch, csr_row:116-133. `local_rows` is the rank's
// segment of the global rowptr (absolute offsets, no end sentinel); `size` is
// the number of rows of the segment. res is [k][seg] column-major. The
// reference reads local_rows[row_i + 1] one past the segment when it steps
// onto the last row; that value is never used, so it is not read here.
template <typename T, typename I>
void row_local_gemv(T *res, std::size_t seg, const I *local_rows,
                    std::size_t size, std::size_t position,
                    std::size_t val_count, const T *vals_data,
                    const I *cols_data, const T *x, std::size_t n,
                    std::size_t k) {
  std::size_t row_i = 0;
  std::size_t current_row_position =
      size > 1 ? static_cast<std::size_t>(local_rows[1]) : 0;
  for (std::size_t i = 0; i < val_count; i++) {
    while (row_i + 1 < size && position + i >= current_row_position) {
      row_i++;
      if (row_i + 1 < size)
        current_row_position = static_cast<std::size_t>(local_rows[row_i + 1]);
    }
    for (std::size_t j = 0; j < k; j++) {
      res[row_i + j * seg] += vals_data[i] * x[cols_data[i] + j * n];
    }
  }
}

// csr_eq::local_gemv CPU branch, csr_eq:112-130. `rows_data` is
// rowptr[row_offsets[rank] .. + row_sizes[rank]) (absolute, no sentinel);
// res is [k][row_sizes[rank]].
template <typename T, typename I>
void eq_local_gemv(T *res, std::size_t res_col_len, const I *rows_data,
                   std::size_t size, std::size_t position,
                   std::size_t elem_count, const T *local_vals,
                   const I *local_cols, const T *x, std::size_t n,
                   std::size_t k) {
  long long row_i = -1;
  std::size_t current_row_position = static_cast<std::size_t>(rows_data[0]);
  for (std::size_t i = 0; i < elem_count; i++) {
    while (static_cast<std::size_t>(row_i + 1) < size &&
           position + i >= current_row_position) {
      row_i++;
      if (static_cast<std::size_t>(row_i + 1) < size)
        current_row_position =
            static_cast<std::size_t>(rows_data[row_i + 1]);
    }
    for (std::size_t j = 0; j < k; j++) {
      res[row_i + j * res_col_len] +=
          local_vals[i] * x[local_cols[i] + j * n];
    }
  }
}

// ---------------------------------------------------------------------------
// Gathers (root side)
// ---------------------------------------------------------------------------

// csr_row::gather_gemv_vector, csr_row:164-199. scratch = concatenation of
// every rank's [k][seg] partial (MPI_Gather of equal counts).
template <typename T>
void row_gather(T *res, std::size_t m, std::size_t seg, std::size_t P,
                std::size_t k, const T *scratch) {
  for (std::size_t j = 0; j < P; j++) {
    if (j * seg >= m) break;
    std::size_t comm_segment_size = std::min(seg, m - j * seg);
    for (std::size_t i = 0; i < k; i++) {
      const T *piece_start = scratch + j * k * seg + i * seg;
      std::copy(piece_start, piece_start + comm_segment_size,
                res + m * i + j * seg);
    }
  }
}

// csr_eq::gather_gemv_vector, csr_eq:168-206. gathered = concatenation of
// each rank's [k][row_sizes[i]] partial (MPI_Gatherv in rank order).
template <typename T>
void eq_gather(T *res, std::size_t m, std::size_t P, std::size_t k,
               const std::size_t *row_offsets, const std::size_t *row_sizes,
               const T *gathered) {
  std::fill(res, res + m * k, T(0));
  for (std::size_t kk = 0; kk < k; kk++) {
    std::size_t current_offset = 0;
    for (std::size_t i = 0; i < P; i++) {
      std::size_t first_row = row_offsets[i];
      std::size_t last_row = row_offsets[i] + row_sizes[i];
      std::size_t row_size = row_sizes[i];
      if (first_row < last_row) {
        res[first_row + kk * m] += gathered[k * current_offset + kk * row_size];
      }
      if (first_row < last_row - 1) {
        const T *piece_start =
            gathered + k * current_offset + kk * row_size + 1;
        std::copy(piece_start, piece_start + (last_row - first_row - 1),
                  res + first_row + kk * m + 1);
      }
      current_offset += row_sizes[i];
    }
  }
}

// ---------------------------------------------------------------------------
// Whole mp::gemv emulation with P ranks executed in sequence
// ---------------------------------------------------------------------------

// gemv(root, res, a, b) -> csr_row::local_gemv_and_collect (csr_row:136-153):
// each rank allocates a zeroed [k][seg] partial, runs local_gemv, the root
// gathers and reorders. All-ranks gemv (gemv.hpp:41-71) then broadcasts the
// same buffer, so its values are identical.
template <typename T, typename I>
void mp_gemv_row(std::size_t m, std::size_t n, std::size_t nnz,
                 const I *rowptr, const I *colind, const T *vals, const T *x,
                 std::size_t k, std::size_t P, T *res) {
  std::vector<std::size_t> off(P), sz(P);
  std::size_t seg = 0;
  row_partition(m, nnz, rowptr, P, &seg, off.data(), sz.data());
  std::vector<T> scratch(seg * P * k, T(0));
  for (std::size_t rank = 0; rank < P; rank++) {
    T *partial = scratch.data() + rank * seg * k;  // zero-filled (csr_row:145)
    if (m <= seg * rank) continue;                 // csr_row:65-66
    std::size_t size = std::min(seg, m - seg * rank);
    row_local_gemv(partial, seg, rowptr + seg * rank, size, off[rank],
                   sz[rank], vals + off[rank], colind + off[rank], x, n, k);
  }
  row_gather(res, m, seg, P, k, scratch.data());
}

// gemv(root, res, a, b) -> csr_eq::local_gemv_and_collect (csr_eq:133-153).
template <typename T, typename I>
void mp_gemv_eq(std::size_t m, std::size_t n, std::size_t nnz, const I *rowptr,
                const I *colind, const T *vals, const T *x, std::size_t k,
                std::size_t P, T *res) {
  std::vector<std::size_t> roff(P), rsz(P);
  std::size_t seg = 0, total = 0;
  eq_partition(m, nnz, rowptr, P, &seg, roff.data(), rsz.data(), &total);
  std::vector<T> gathered(total * k, T(0));
  std::size_t cur = 0;
  for (std::size_t rank = 0; rank < P; rank++) {
    T *partial = gathered.data() + cur * k;  // zero-filled (csr_eq:145-146)
    cur += rsz[rank];
    if (nnz <= seg * rank) continue;  // csr_eq:56-58
    std::size_t position = seg * rank;
    std::size_t elem_count = std::min(seg, nnz - seg * rank);
    eq_local_gemv(partial, rsz[rank], rowptr + roff[rank], rsz[rank],
                  position, elem_count, vals + position, colind + position, x,
                  n, k);
  }
  eq_gather(res, m, P, k, roff.data(), rsz.data(), gathered.data());
}

// ---------------------------------------------------------------------------
// Extended-precision oracle (long double accumulation, x86 80-bit)
// ---------------------------------------------------------------------------

// Row r spans [clamp(rowptr[r]-base), clamp(end(r)-base)) where end(r) is
// rowptr[r+1] for r + 1 < nrows, rowptr[nrows] if a sentinel exists, else
// base + nnz. This one rule covers full matrices (base 0, sentinel), csr_row
// shares (absolute offsets, no sentinel) and csr_eq shares (boundary rows cut).
template <typename I>
inline void row_span(const I *rowptr, std::int64_t base, std::int64_t nrows,
                     std::int64_t nnz, int has_sentinel, std::int64_t r,
                     std::int64_t *b, std::int64_t *e) {
  auto clampv = [&](std::int64_t v) {
    return std::min<std::int64_t>(std::max<std::int64_t>(v, 0), nnz);
  };
  *b = clampv(static_cast<std::int64_t>(rowptr[r]) - base);
  if (r + 1 < nrows || has_sentinel)
    *e = clampv(static_cast<std::int64_t>(rowptr[r + 1]) - base);
  else
    *e = nnz;
}

template <typename T, typename I>
void spmv_ref(const I *rowptr, std::int64_t base, std::int64_t nrows,
              std::int64_t nnz, int has_sentinel, const I *colind,
              const T *vals, const T *x, std::int64_t ldx, std::int64_t k,
              double *y, std::int64_t ldy, double *absy) {
  for (std::int64_t r = 0; r < nrows; r++) {
    std::int64_t b, e;
    row_span(rowptr, base, nrows, nnz, has_sentinel, r, &b, &e);
    for (std::int64_t j = 0; j < k; j++) {
      long double s = 0, sa = 0;
      for (std::int64_t q = b; q < e; q++) {
        long double p = static_cast<long double>(vals[q]) *
                        static_cast<long double>(x[colind[q] + j * ldx]);
        s += p;
        sa += p < 0 ? -p : p;
      }
      y[r + j * ldy] = static_cast<double>(s);
      if (absy) absy[r + j * ldy] = static_cast<double>(sa);
    }
  }
}

// dx = A^T dy over the same row spans.
template <typename T, typename I>
void spmv_t_ref(const I *rowptr, std::int64_t base, std::int64_t nrows,
                std::int64_t ncols, std::int64_t nnz, int has_sentinel,
                const I *colind, const T *vals, const T *dy, std::int64_t lddy,
                std::int64_t k, double *dx, std::int64_t lddx, double *absdx) {
  std::vector<long double> acc(static_cast<std::size_t>(ncols * k), 0),
      acca(static_cast<std::size_t>(ncols * k), 0);
  for (std::int64_t r = 0; r < nrows; r++) {
    std::int64_t b, e;
    row_span(rowptr, base, nrows, nnz, has_sentinel, r, &b, &e);
    for (std::int64_t j = 0; j < k; j++) {
      long double d = static_cast<long double>(dy[r + j * lddy]);
      for (std::int64_t q = b; q < e; q++) {
        long double p = static_cast<long double>(vals[q]) * d;
        acc[colind[q] + j * ncols] += p;
        acca[colind[q] + j * ncols] += p < 0 ? -p : p;
      }
    }
  }
  for (std::int64_t j = 0; j < k; j++)
    for (std::int64_t c = 0; c < ncols; c++) {
      dx[c + j * lddx] = static_cast<double>(acc[c + j * ncols]);
      if (absdx) absdx[c + j * lddx] = static_cast<double>(acca[c + j * ncols]);
    }
}

// dvals[e] = sum_j dy[row(e) + j*lddy] * x[col(e) + j*ldx].
template <typename T, typename I>
void dvals_ref(const I *rowptr, std::int64_t base, std::int64_t nrows,
               std::int64_t nnz, int has_sentinel, const I *colind, const T *x,
               std::int64_t ldx, const T *dy, std::int64_t lddy,
               std::int64_t k, double *dvals) {
  for (std::int64_t r = 0; r < nrows; r++) {
    std::int64_t b, e;
    row_span(rowptr, base, nrows, nnz, has_sentinel, r, &b, &e);
    for (std::int64_t q = b; q < e; q++) {
      long double s = 0;
      for (std::int64_t j = 0; j < k; j++)
        s += static_cast<long double>(dy[r + j * lddy]) *
             static_cast<long double>(x[colind[q] + j * ldx]);
      dvals[q] = static_cast<double>(s);
    }
  }
}

// Stable CSR -> CSC (counting sort by column, rows ascending within a column).
// perm[p] is the CSR position of CSC entry p.
template <typename I>
void csr_transpose(const I *rowptr, std::int64_t base, std::int64_t nrows,
                   std::int64_t ncols, std::int64_t nnz, int has_sentinel,
                   const I *colind, std::int64_t *colptr, std::int64_t *rowind,
                   std::int64_t *perm) {
  std::fill(colptr, colptr + ncols + 1, 0);
  for (std::int64_t q = 0; q < nnz; q++) colptr[colind[q] + 1]++;
  for (std::int64_t c = 0; c < ncols; c++) colptr[c + 1] += colptr[c];
  std::vector<std::int64_t> next(colptr, colptr + ncols);
  for (std::int64_t r = 0; r < nrows; r++) {
    std::int64_t b, e;
    row_span(rowptr, base, nrows, nnz, has_sentinel, r, &b, &e);
    for (std::int64_t q = b; q < e; q++) {
      std::int64_t p = next[colind[q]]++;
      rowind[p] = r;
      perm[p] = q;
    }
  }
}

// ---------------------------------------------------------------------------
// Threaded GemvEq_DR CPU path (the reference's CPU baseline, restated)
// ---------------------------------------------------------------------------

// One "rank" per std::thread, pinned to one allowed core, owning a private
// copy of its csr_eq share (the reference's ranks own their MPI windows).
// One timed call is exactly what gemv(0, res, m, x) does without SYCL
// (benchmarks/gbench/mp/gemv.cpp:60-63, csr_eq:133-153):
//   every rank: allocate + zero its partial, run the sequential local loop;
//   gather to root (Gatherv == root copies the partials in rank order);
//   root: fill(res, 0), boundary-row add and copies (csr_eq:185-205);
//   fence (barrier).
// Dense tile product, the step of dr::sp::gemm (sp gemm). custom_gemm
// (include/dr/sp/algorithms/matrix/local_gemm.hpp:15-41) adds a[i][k] * b[k][j]
// into c[i][j] for every (i, k, j) of row-major views; mkl_gemm (:45-60) is
// the same product, alpha = beta = 1. Here out = alpha * sum + beta * c in
// long double (c not read when beta == 0) and absout = |alpha| sum |a||b| +
// |beta||c|, the scale of the rounding-error bound.
template <typename T>
void gemm_ref(std::int64_t m, std::int64_t n, std::int64_t k, double alpha,
              const T *a, std::int64_t lda, const T *b, std::int64_t ldb,
              double beta, const T *c, std::int64_t ldc, double *out,
              double *absout) {
  std::vector<long double> s(n), sa(n);
  for (std::int64_t i = 0; i < m; i++) {
    std::fill(s.begin(), s.end(), 0.0L);
    std::fill(sa.begin(), sa.end(), 0.0L);
    for (std::int64_t q = 0; q < k; q++) {
      const long double av = a[i * lda + q];
      const long double aa = av < 0 ? -av : av;
      const T *brow = b + q * ldb;
      for (std::int64_t j = 0; j < n; j++) {
        const long double bv = brow[j];
        s[j] += av * bv;
        sa[j] += aa * (bv < 0 ? -bv : bv);
      }
    }
    for (std::int64_t j = 0; j < n; j++) {
      long double v = static_cast<long double>(alpha) * s[j];
      long double w = std::fabs(static_cast<long double>(alpha)) * sa[j];
      if (beta != 0) {
        const long double cv = c[i * ldc + j];
        v += static_cast<long double>(beta) * cv;
        w += std::fabs(static_cast<long double>(beta)) * (cv < 0 ? -cv : cv);
      }
      out[i * n + j] = static_cast<double>(v);
      if (absout) absout[i * n + j] = static_cast<double>(w);
    }
  }
}

template <typename T, typename I>
double gemveq_cpu_bench(std::size_t m, std::size_t n, std::size_t nnz,
                        const I *rowptr, const I *colind, const T *vals,
                        const T *x, std::size_t k, std::size_t P, int warmup,
                        int reps, T *res, double *times) {
  std::vector<std::size_t> roff(P), rsz(P);
  std::size_t seg = 0, total = 0;
  eq_partition(m, nnz, rowptr, P, &seg, roff.data(), rsz.data(), &total);

  cpu_set_t allowed;
  CPU_ZERO(&allowed);
  sched_getaffinity(0, sizeof(allowed), &allowed);
  std::vector<int> cores;
  for (int c = 0; c < CPU_SETSIZE; c++)
    if (CPU_ISSET(c, &allowed)) cores.push_back(c);

  std::vector<T> gathered(total * k);
  std::vector<T *> partials(P, nullptr);
  std::barrier sync(static_cast<std::ptrdiff_t>(P));
  std::vector<double> t_rank0;

  auto rank_main = [&](std::size_t rank) {
    if (!cores.empty()) {
      cpu_set_t one;
      CPU_ZERO(&one);
      CPU_SET(cores[rank % cores.size()], &one);
      pthread_setaffinity_np(pthread_self(), sizeof(one), &one);
    }
    // Share storage, first-touched by the owning rank.
    std::size_t position = seg * rank;
    std::size_t elem_count =
        nnz > position ? std::min(seg, nnz - position) : 0;
    std::vector<T> lv(vals + std::min(position, nnz),
                      vals + std::min(position, nnz) + elem_count);
    std::vector<I> lc(colind + std::min(position, nnz),
                      colind + std::min(position, nnz) + elem_count);
    std::vector<I> lr(rowptr + std::min(roff[rank], m),
                      rowptr + std::min(roff[rank] + rsz[rank], m));
    std::vector<T> lx(x, x + n * k);  // broadcasted_slim_matrix copy
    for (int it = 0; it < warmup + reps; it++) {
      sync.arrive_and_wait();
      auto t0 = std::chrono::steady_clock::now();
      std::vector<T> partial(rsz[rank] * k);  // alloc + zero (csr_eq:138-146)
      if (nnz > seg * rank && !lr.empty()) {
        eq_local_gemv(partial.data(), rsz[rank], lr.data(), rsz[rank],
                      position, elem_count, lv.data(), lc.data(), lx.data(),
                      n, k);
      }
      partials[rank] = partial.data();
      sync.arrive_and_wait();  // MPI_Gatherv to the root
      if (rank == 0) {
        std::size_t cur = 0;
        for (std::size_t r = 0; r < P; r++) {
          std::memcpy(gathered.data() + cur * k, partials[r],
                      rsz[r] * k * sizeof(T));
          cur += rsz[r];
        }
        eq_gather(res, m, P, k, roff.data(), rsz.data(), gathered.data());
      }
      sync.arrive_and_wait();  // fence()
      if (rank == 0) {
        double dt = std::chrono::duration<double>(
                        std::chrono::steady_clock::now() - t0)
                        .count();
        if (it >= warmup) t_rank0.push_back(dt);
      }
    }
  };
  // Every rank, rank 0 included, runs on a thread of its own, so pinning a
  // rank never changes the caller's own CPU affinity.
  std::vector<std::thread> th;
  for (std::size_t r = 0; r < P; r++) th.emplace_back(rank_main, r);
  for (auto &t : th) t.join();
  for (std::size_t i = 0; i < t_rank0.size(); i++) times[i] = t_rank0[i];
  std::vector<double> s = t_rank0;
  std::sort(s.begin(), s.end());
  return s.empty() ? 0.0 : s[s.size() / 2];
}

// ---------------------------------------------------------------------------
// Generators (gen:40-103 and gen:108-153), same std:: engine and
// distributions as the reference, so the same standard library gives the same
// arrays. T = float/double, I = std::size_t (the mp tests) or int64_t / long.
// ---------------------------------------------------------------------------

template <typename T> struct udist {
  using type = std::uniform_int_distribution<T>;
};
template <> struct udist<float> {
  using type = std::uniform_real_distribution<float>;
};
template <> struct udist<double> {
  using type = std::uniform_real_distribution<double>;
};

struct pair_hash_u64 {
  std::size_t operator()(const std::pair<std::uint64_t, std::uint64_t> &v)
      const {
    return v.first * 31 + v.second;
  }
};

// Returns nnz; fills arrays sized by the caller (query with values == null).
template <typename T, typename I>
std::size_t gen_random_csr(std::size_t m, std::size_t n, double density,
                           unsigned int seed, T *values, I *rowptr,
                           I *colind) {
  std::size_t nnz = density * m * n;  // gen:48 (double -> size_t truncation)
  if (values == nullptr) return nnz;
  std::unordered_set<std::pair<std::uint64_t, std::uint64_t>, pair_hash_u64>
      tuples{};
  std::vector<std::pair<std::pair<I, I>, T>> entries;
  entries.reserve(nnz);
  std::mt19937 gen(seed);
  std::uniform_int_distribution<I> row(0, static_cast<I>(m - 1));
  std::uniform_int_distribution<I> column(0, static_cast<I>(n - 1));
  typename udist<T>::type value_gen(0, 1);
  while (tuples.size() < nnz) {
    auto i = row(gen);
    auto j = column(gen);
    std::pair<std::uint64_t, std::uint64_t> key(static_cast<std::uint64_t>(i),
                                                static_cast<std::uint64_t>(j));
    if (tuples.find(key) == tuples.end()) {
      T value = value_gen(gen);
      tuples.insert(key);
      entries.push_back({{i, j}, value});
    }
  }
  rowptr[0] = 0;
  std::size_t r = 0, c = 0;
  std::sort(entries.begin(), entries.end());
  for (auto &&[index, value] : entries) {
    auto &&[i, j] = index;
    values[c] = value;
    colind[c] = j;
    while (r < static_cast<std::size_t>(i)) {
      rowptr[r + 1] = static_cast<I>(c);
      r++;
    }
    c++;
  }
  for (; r < m; r++) rowptr[r + 1] = static_cast<I>(nnz);
  return nnz;
}

template <typename T, typename I>
std::size_t gen_band_csr(std::size_t size, std::size_t up_band,
                         std::size_t down_band, T *values, I *rowptr,
                         I *colind) {
  std::size_t nnz = (1 + up_band + down_band) * size -
                    (up_band * (up_band + 1) / 2) -
                    (down_band * (down_band + 1) / 2);  // gen:111-113
  if (values == nullptr) return nnz;
  rowptr[0] = 0;
  std::size_t r = 0, c = 0;
  for (long long i = 0; i < static_cast<long long>(size); i++) {
    for (long long j = std::max(i - static_cast<long long>(down_band), 0LL);
         j < i; j++) {
      values[c] = 1;
      colind[c] = static_cast<I>(j);
      c++;
    }
    values[c] = 1;
    colind[c] = static_cast<I>(i);
    c++;
    for (long long j = i + 1; j <= i + static_cast<long long>(up_band); j++) {
      if (j >= static_cast<long long>(size)) continue;
      values[c] = 1;
      colind[c] = static_cast<I>(j);
      c++;
    }
    rowptr[r + 1] = static_cast<I>(c);
    r++;
  }
  for (; r < size; r++) rowptr[r + 1] = static_cast<I>(nnz);
  return nnz;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI (ctypes). Suffixes: _f32/_f64 = value type; index type int64_t unless
// the name says _i32 or _u64.
// ---------------------------------------------------------------------------
extern "C" {

int orc_version() { return 1; }

void orc_row_partition(std::size_t m, std::size_t nnz, const std::int64_t *rp,
                       std::size_t P, std::size_t *seg, std::size_t *off,
                       std::size_t *sz) {
  row_partition(m, nnz, rp, P, seg, off, sz);
}
void orc_eq_partition(std::size_t m, std::size_t nnz, const std::int64_t *rp,
                      std::size_t P, std::size_t *seg, std::size_t *roff,
                      std::size_t *rsz, std::size_t *total) {
  eq_partition(m, nnz, rp, P, seg, roff, rsz, total);
}

#define ORC_MP(T, TN)                                                          \
  void orc_mp_gemv_row_##TN(std::size_t m, std::size_t n, std::size_t nnz,     \
                            const std::int64_t *rp, const std::int64_t *ci,    \
                            const T *v, const T *x, std::size_t k,             \
                            std::size_t P, T *res) {                           \
    mp_gemv_row(m, n, nnz, rp, ci, v, x, k, P, res);                           \
  }                                                                            \
  void orc_mp_gemv_eq_##TN(std::size_t m, std::size_t n, std::size_t nnz,      \
                           const std::int64_t *rp, const std::int64_t *ci,     \
                           const T *v, const T *x, std::size_t k,              \
                           std::size_t P, T *res) {                            \
    mp_gemv_eq(m, n, nnz, rp, ci, v, x, k, P, res);                            \
  }                                                                            \
  void orc_mp_gemv_eq_i32_##TN(std::size_t m, std::size_t n, std::size_t nnz,  \
                               const std::int32_t *rp, const std::int32_t *ci, \
                               const T *v, const T *x, std::size_t k,          \
                               std::size_t P, T *res) {                        \
    mp_gemv_eq(m, n, nnz, rp, ci, v, x, k, P, res);                            \
  }                                                                            \
  void orc_spmv_ref_##TN(const std::int64_t *rp, std::int64_t base,            \
                         std::int64_t nrows, std::int64_t nnz, int sent,       \
                         const std::int64_t *ci, const T *v, const T *x,       \
                         std::int64_t ldx, std::int64_t k, double *y,          \
                         std::int64_t ldy, double *absy) {                     \
    spmv_ref(rp, base, nrows, nnz, sent, ci, v, x, ldx, k, y, ldy, absy);      \
  }                                                                            \
  void orc_spmv_t_ref_##TN(const std::int64_t *rp, std::int64_t base,          \
                           std::int64_t nrows, std::int64_t ncols,             \
                           std::int64_t nnz, int sent, const std::int64_t *ci, \
                           const T *v, const T *dy, std::int64_t lddy,         \
                           std::int64_t k, double *dx, std::int64_t lddx,      \
                           double *absdx) {                                    \
    spmv_t_ref(rp, base, nrows, ncols, nnz, sent, ci, v, dy, lddy, k, dx,      \
               lddx, absdx);                                                   \
  }                                                                            \
  void orc_dvals_ref_##TN(const std::int64_t *rp, std::int64_t base,           \
                          std::int64_t nrows, std::int64_t nnz, int sent,      \
                          const std::int64_t *ci, const T *x,                  \
                          std::int64_t ldx, const T *dy, std::int64_t lddy,    \
                          std::int64_t k, double *dv) {                        \
    dvals_ref(rp, base, nrows, nnz, sent, ci, x, ldx, dy, lddy, k, dv);        \
  }                                                                            \
  double orc_gemveq_cpu_bench_i32_##TN(                                        \
      std::size_t m, std::size_t n, std::size_t nnz, const std::int32_t *rp,   \
      const std::int32_t *ci, const T *v, const T *x, std::size_t k,           \
      std::size_t P, int warmup, int reps, T *res, double *times) {            \
    return gemveq_cpu_bench(m, n, nnz, rp, ci, v, x, k, P, warmup, reps, res,  \
                            times);                                            \
  }                                                                            \
  double orc_gemveq_cpu_bench_##TN(                                            \
      std::size_t m, std::size_t n, std::size_t nnz, const std::int64_t *rp,   \
      const std::int64_t *ci, const T *v, const T *x, std::size_t k,           \
      std::size_t P, int warmup, int reps, T *res, double *times) {            \
    return gemveq_cpu_bench(m, n, nnz, rp, ci, v, x, k, P, warmup, reps, res,  \
                            times);                                            \
  }

ORC_MP(float, f32)
ORC_MP(double, f64)
#undef ORC_MP

void orc_gemm_ref_f32(std::int64_t m, std::int64_t n, std::int64_t k, double alpha,
                      const float *a, std::int64_t lda, const float *b,
                      std::int64_t ldb, double beta, const float *c,
                      std::int64_t ldc, double *out, double *absout) {
  gemm_ref(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, out, absout);
}
void orc_gemm_ref_f64(std::int64_t m, std::int64_t n, std::int64_t k, double alpha,
                      const double *a, std::int64_t lda, const double *b,
                      std::int64_t ldb, double beta, const double *c,
                      std::int64_t ldc, double *out, double *absout) {
  gemm_ref(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, out, absout);
}

void orc_csr_transpose(const std::int64_t *rp, std::int64_t base,
                       std::int64_t nrows, std::int64_t ncols,
                       std::int64_t nnz, int sent, const std::int64_t *ci,
                       std::int64_t *colptr, std::int64_t *rowind,
                       std::int64_t *perm) {
  csr_transpose(rp, base, nrows, ncols, nnz, sent, ci, colptr, rowind, perm);
}

// generate_random_csr<float, std::size_t> (the mp tests' call) and
// generate_random_csr<float, int64_t> (sp::sparse_matrix<float> default I).
std::size_t orc_gen_random_csr_f32_u64(std::size_t m, std::size_t n,
                                       double density, unsigned seed,
                                       float *v, std::uint64_t *rp,
                                       std::uint64_t *ci) {
  return gen_random_csr<float, std::uint64_t>(m, n, density, seed, v, rp, ci);
}
std::size_t orc_gen_random_csr_f32_i64(std::size_t m, std::size_t n,
                                       double density, unsigned seed,
                                       float *v, std::int64_t *rp,
                                       std::int64_t *ci) {
  return gen_random_csr<float, std::int64_t>(m, n, density, seed, v, rp, ci);
}
std::size_t orc_gen_band_csr_f64_i64(std::size_t size, std::size_t up,
                                     std::size_t down, double *v,
                                     std::int64_t *rp, std::int64_t *ci) {
  return gen_band_csr<double, std::int64_t>(size, up, down, v, rp, ci);
}

}  // extern "C"
