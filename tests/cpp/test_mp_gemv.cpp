// test_mp_gemv.cpp -- the reference's dr::mp gemv tests
// (test/gtest/mp/gemv.cpp:6-184) replayed through this host interface, for
// P = 1, 2, 3, 4 ranks sharing the GPU, plus the reference benchmark's
// GemvEq_DR input (band matrix, width 8, x[i + j*n] = i*j + 1) checked for
// exact equality against a host loop (examples/mp/sparse_matrix.cpp:84-102).
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <iterator>
#include <ranges>
#include <utility>
#include <vector>

#include "dr/mp.hpp"
#include "harness.hpp"

namespace mp = dr::mp;

template <typename M> void testMatrixGemv(std::size_t m, std::size_t k, M &a) {
  std::vector<float> base_b(k, 1.f);
  std::vector<float> c(m, 0.f);
  mp::broadcasted_vector<float> allocated_b;
  allocated_b.broadcast_data(k, 0, base_b, mp::default_comm());
  mp::gemv(c, a, allocated_b);
  std::vector<float> c_ref(m, 0.f);
  for (auto &&[index, v] : a) c_ref[index[0]] += v;
  EXPECT_TRUE(harness::fp_equal(c_ref, c));
  // rooted form: only the root's result is defined
  std::vector<float> c0(m, 0.f);
  mp::gemv(0, c0, a, allocated_b);
  if (mp::default_comm().rank() == 0) EXPECT_TRUE(harness::fp_equal(c_ref, c0));
  allocated_b.destroy_data();
}

template <typename M>
void testMatrixGemm(std::size_t m, std::size_t n, M &a, std::size_t width) {
  std::vector<float> base_b(n * width);
  std::vector<float> c(m * width, 0.f);
  for (std::size_t i = 0; i < n * width; i++) base_b[i] = static_cast<float>(i);
  mp::broadcasted_slim_matrix<float> allocated_b;
  allocated_b.broadcast_data(n, width, 0, base_b, mp::default_comm());
  mp::gemv(c, a, allocated_b);
  std::vector<float> c_ref(m * width, 0.f);
  for (auto &&[index, v] : a)
    for (std::size_t j = 0; j < width; j++) c_ref[index[0] + j * m] += v * base_b[index[1] + j * n];
  EXPECT_TRUE(harness::fp_equal(c_ref, c));
  allocated_b.destroy_data();
}

template <template <typename, typename, typename> class Distr>
void gemv_case(std::size_t m, std::size_t k, bool gemm) {
  for (std::size_t P = 1; P <= 4; P++) {
    auto csr = dr::generate_random_csr<float, unsigned long>(dr::index<unsigned long>(m, k), 0.1f);
    mp::run_ranks(P, [&] {
      mp::distributed_sparse_matrix<float, unsigned long, mp::MpiBackend,
                                    Distr<float, unsigned long, mp::MpiBackend>>
          a(csr, 0);
      if (gemm)
        testMatrixGemm(m, k, a, 20);
      else
        testMatrixGemv(m, k, a);
    });
    dr::destroy_csr_matrix_view(csr);
  }
}

TEST(SparseMatrix, GemvRow) { gemv_case<mp::csr_row_distribution>(100, 100, false); }
TEST(SparseMatrix, GemvEq) { gemv_case<mp::csr_eq_distribution>(100, 100, false); }
TEST(SparseMatrix, GemvRowNotSquare) { gemv_case<mp::csr_row_distribution>(1000, 10, false); }
TEST(SparseMatrix, GemvEqNotSquare) { gemv_case<mp::csr_eq_distribution>(1000, 10, false); }
TEST(SparseMatrix, GemvRowNotSquareDifferentAxis) { gemv_case<mp::csr_row_distribution>(10, 1000, false); }
TEST(SparseMatrix, GemvEqNotSquareDifferentAxis) { gemv_case<mp::csr_eq_distribution>(10, 1000, false); }
TEST(SparseMatrix, GemmRow) { gemv_case<mp::csr_row_distribution>(100, 100, true); }
TEST(SparseMatrix, GemmEq) { gemv_case<mp::csr_eq_distribution>(100, 100, true); }
TEST(SparseMatrix, GemmRowNotSquare) { gemv_case<mp::csr_row_distribution>(1000, 10, true); }
TEST(SparseMatrix, GemmEqNotSquare) { gemv_case<mp::csr_eq_distribution>(1000, 10, true); }
TEST(SparseMatrix, GemmRowNotSquareDifferentAxis) { gemv_case<mp::csr_row_distribution>(10, 1000, true); }
TEST(SparseMatrix, GemmEqNotSquareDifferentAxis) { gemv_case<mp::csr_eq_distribution>(10, 1000, true); }

// Distributed iteration reproduces the input CSR (test/gtest/mp/sparse_matrix.cpp).
TEST(SparseMatrix, IterationMatchesInput) {
  auto csr = dr::generate_random_csr<float, unsigned long>(dr::index<unsigned long>(100, 100), 0.1f);
  for (std::size_t P = 1; P <= 4; P++) {
    mp::run_ranks(P, [&] {
      mp::distributed_sparse_matrix<float, unsigned long, mp::MpiBackend,
                                    mp::csr_row_distribution<float, unsigned long, mp::MpiBackend>>
          a(csr, 0);
      mp::distributed_sparse_matrix<float, unsigned long> b(csr, 0);  // csr_eq default
      std::vector<dr::matrix_entry<float, unsigned long>> in(csr.begin(), csr.end());
      std::vector<dr::matrix_entry<float, unsigned long>> ga(a.begin(), a.end());
      std::vector<dr::matrix_entry<float, unsigned long>> gb(b.begin(), b.end());
      EXPECT_EQ(ga.size(), in.size());
      EXPECT_EQ(gb.size(), in.size());
      bool same = ga.size() == in.size() && gb.size() == in.size();
      for (std::size_t i = 0; same && i < in.size(); i++)
        same = ga[i].idx == in[i].idx && ga[i].val == in[i].val && gb[i].idx == in[i].idx &&
               gb[i].val == in[i].val;
      EXPECT_TRUE(same);
    });
  }
  dr::destroy_csr_matrix_view(csr);
}

// The distributions satisfy the plug-in concepts, and the matrix is a
// forward range with segments (test/gtest/mp/sparse_matrix.cpp:34-70).
template <template <typename, typename, typename> class Distr> void static_assert_case() {
  using Dist = Distr<float, unsigned long, mp::MpiBackend>;
  static_assert(mp::matrix_distibution<Dist>);
  static_assert(mp::vector_multiplicable<Dist>);
  using Matrix = mp::distributed_sparse_matrix<float, unsigned long, mp::MpiBackend, Dist>;
  static_assert(std::forward_iterator<decltype(std::declval<Matrix &>().begin())>);
  static_assert(std::forward_iterator<decltype(std::declval<Matrix &>().end())>);
  static_assert(std::forward_iterator<decltype(std::ranges::begin(std::declval<Matrix &>()))>);
  static_assert(std::forward_iterator<decltype(std::ranges::end(std::declval<Matrix &>()))>);
  static_assert(std::ranges::forward_range<Matrix>);
  static_assert(dr::distributed_range<Matrix>);
  // the default distribution of the reference's test (csr_eq) as well
  static_assert(dr::distributed_range<mp::distributed_sparse_matrix<float, unsigned long>>);
  // and a constructed matrix iterates as one
  auto csr = dr::generate_random_csr<float, unsigned long>(dr::index<unsigned long>(100, 100), 0.1f);
  mp::run_ranks(2, [&] {
    Matrix a(csr, 0);
    std::size_t n = 0;
    for (auto it = std::ranges::begin(a); it != std::ranges::end(a); ++it) n++;
    EXPECT_EQ(n, csr.size());
    EXPECT_EQ(dr::ranges::segments(a).size(), std::size_t(2));
  });
  dr::destroy_csr_matrix_view(csr);
}

TEST(SparseMatrix, staticAssertEq) { static_assert_case<mp::csr_eq_distribution>(); }
TEST(SparseMatrix, staticAssertRow) { static_assert_case<mp::csr_row_distribution>(); }

// GemvEq_DR's input at the mp-bench ctest size (n = 866, band +-17, f64/long,
// width 8, x[i + j*n] = i*j + 1): all-ones band, so every product is an
// integer and the result is exact.
TEST(SparseMatrix, GemvEqBenchInputExact) {
  const long n = 866;
  auto csr = dr::generate_band_csr<double, long>(n, 17, 17);
  EXPECT_EQ(csr.size(), std::size_t(30004));
  const std::size_t width = 8;
  std::vector<double> base_a(n * width);
  for (std::size_t j = 0; j < width; j++)
    for (long i = 0; i < n; i++) base_a[i + j * n] = double(i) * double(j) + 1;
  std::vector<double> ref(n * width, 0.0);
  for (auto &&[index, v] : csr)
    for (std::size_t j = 0; j < width; j++) ref[index[0] + j * n] += v * base_a[index[1] + j * n];
  for (std::size_t P = 1; P <= 4; P++) {
    mp::run_ranks(P, [&] {
      mp::distributed_sparse_matrix<double, long, mp::MpiBackend,
                                    mp::csr_eq_distribution<double, long, mp::MpiBackend>>
          m(csr, 0);
      mp::broadcasted_slim_matrix<double> allocated_a;
      allocated_a.broadcast_data(n, width, 0, base_a, mp::default_comm());
      std::vector<double> res(n * width);
      mp::gemv(0, res, m, allocated_a);
      if (mp::default_comm().rank() == 0) EXPECT_TRUE(res == ref);
      mp::distributed_sparse_matrix<double, long, mp::MpiBackend,
                                    mp::csr_row_distribution<double, long, mp::MpiBackend>>
          r(csr, 0);
      std::vector<double> res2(n * width);
      mp::gemv(res2, r, allocated_a);
      EXPECT_TRUE(res2 == ref);
      allocated_a.destroy_data();
    });
  }
  dr::destroy_csr_matrix_view(csr);
}

// int32 indices (as the benchmark workloads use) and empty rows, mixed signs.
TEST(SparseMatrix, Int32IndicesEmptyRows) {
  const int m = 500, n = 300;
  std::vector<std::pair<std::pair<int, int>, double>> cells;
  unsigned s = 7;
  for (int i = 0; i < m; i++) {
    if (i % 3 == 1) continue;  // empty rows
    for (int j = (i * 7) % 11; j < n; j += 13 + (i % 5)) {
      s = s * 1103515245u + 12345u;
      cells.push_back({{i, j}, ((s >> 8) % 2001) / 1000.0 - 1.0});
    }
  }
  auto csr = dr::__detail::convert_to_csr(cells, {std::size_t(m), std::size_t(n)}, cells.size(),
                                          std::allocator<double>{});
  std::vector<double> x(n);
  for (int j = 0; j < n; j++) x[j] = std::sin(0.37 * j);
  std::vector<double> ref(m, 0.0);
  for (auto &&[index, v] : csr) ref[index[0]] += v * x[index[1]];
  for (std::size_t P = 1; P <= 4; P++) {
    mp::run_ranks(P, [&] {
      mp::distributed_sparse_matrix<double, int> a(csr, 0);
      mp::broadcasted_vector<double> b;
      b.broadcast_data(n, 0, x, mp::default_comm());
      std::vector<double> c(m);
      mp::gemv(c, a, b);
      bool ok = true;
      for (int i = 0; i < m; i++) ok = ok && std::abs(c[i] - ref[i]) <= 1e-12 * (1 + std::abs(ref[i]));
      EXPECT_TRUE(ok);
      b.destroy_data();
    });
  }
  std::allocator<double>{}.deallocate(csr.values_data(), csr.size());
  std::allocator<int>{}.deallocate(csr.rowptr_data(), m + 1);
  std::allocator<int>{}.deallocate(csr.colind_data(), csr.size());
}

// A Zipf(1.5) matrix (rows clipped to [1, lmax], stratified sorted columns,
// values N(0, 1)) generated on the GPU by the library's generators and
// copied into host arrays (release with dr::destroy_csr_matrix_view).
static dr::views::csr_matrix_view<float, int> zipf_host(int m, int n, int lmax,
                                                        std::uint64_t seed) {
  dr::drx::device_buffer<int> rp(m + 1);
  std::int64_t nnz = 0;
  dr::drx::check(drx_gen_zipf_rowptr(DRX_I32, rp.data(), m, 1.5, lmax, seed, &nnz, nullptr),
                 "drx_gen_zipf_rowptr");
  dr::drx::device_buffer<int> ci(nnz);
  dr::drx::device_buffer<float> v(nnz);
  dr::drx::check(drx_gen_stratified_colind(DRX_I32, rp.data(), m, n, ci.data(), seed + 1, nullptr),
                 "drx_gen_stratified_colind");
  dr::drx::check(drx_gen_normal(DRX_F32, v.data(), nnz, seed + 2, nullptr), "drx_gen_normal");
  int *hrp = new int[m + 1];
  int *hci = new int[nnz];
  float *hv = new float[nnz];
  rp.download(hrp, m + 1);
  ci.download(hci, nnz);
  v.download(hv, nnz);
  return dr::views::csr_matrix_view<float, int>(hv, hrp, hci, dr::index<int>(m, n), nnz, 0);
}

// Componentwise error max |y - y*| / (|A||x|) against a long double loop.
static double zipf_err(const dr::views::csr_matrix_view<float, int> &a,
                       const std::vector<float> &x, const std::vector<float> &y) {
  const int m = static_cast<int>(a.shape()[0]);
  double worst = 0;
  for (int r = 0; r < m; r++) {
    long double s = 0, t = 0;
    for (int q = a.rowptr_data()[r]; q < a.rowptr_data()[r + 1]; q++) {
      const long double p = static_cast<long double>(a.values_data()[q]) * x[a.colind_data()[q]];
      s += p;
      t += p < 0 ? -p : p;
    }
    const double d = std::abs(static_cast<double>(y[r] - s));
    if (t > 0) worst = std::max(worst, d / static_cast<double>(t));
    else if (d != 0) worst = 1;
  }
  return worst;
}

// The headline speed through the reference's seams: a share whose columns
// scatter over an x larger than one XCD's L2 (2M columns of FP32, 8 MB)
// takes the column-panel layout when its distribution builds the share
// (csr_row_distribution::init / csr_eq_distribution::init), and
// dr::mp::gemv then runs the panel kernels; results within the FP32
// tolerance (componentwise 1e-5) of a long double loop, for P = 1 and 2.
TEST(SparseMatrix, ZipfShareTakesColumnPanels) {
  const int m = 2000000, n = 2000000;
  auto csr = zipf_host(m, n, 256, 77);
  std::vector<float> x(n);
  for (int j = 0; j < n; j++) x[j] = static_cast<float>(std::sin(0.001 * j + 0.3));
  for (std::size_t P = 1; P <= 2; P++) {
    mp::run_ranks(P, [&] {
      mp::distributed_sparse_matrix<float, int, mp::MpiBackend,
                                    mp::csr_row_distribution<float, int, mp::MpiBackend>>
          r(csr, 0);
      mp::distributed_sparse_matrix<float, int> e(csr, 0);  // csr_eq
      EXPECT_TRUE(r.matrix_distribution().panels() >= 4);
      EXPECT_TRUE(e.matrix_distribution().panels() >= 4);
      mp::broadcasted_vector<float> b;
      b.broadcast_data(n, 0, x, mp::default_comm());
      std::vector<float> yr(m), ye(m);
      mp::gemv(yr, r, b);
      mp::gemv(0, ye, e, b);
      EXPECT_TRUE(zipf_err(csr, x, yr) < 1e-5);
      if (mp::default_comm().rank() == 0) EXPECT_TRUE(zipf_err(csr, x, ye) < 1e-5);
      b.destroy_data();
    });
  }
  // the knob turns the layout off; results stay within tolerance
  dr::drx::check(drx_tune_set("auto_panels", 0), "drx_tune_set");
  mp::run_ranks(1, [&] {
    mp::distributed_sparse_matrix<float, int> e(csr, 0);
    EXPECT_TRUE(e.matrix_distribution().panels() == 0);
    mp::broadcasted_vector<float> b;
    b.broadcast_data(n, 0, x, mp::default_comm());
    std::vector<float> y(m);
    mp::gemv(y, e, b);
    EXPECT_TRUE(zipf_err(csr, x, y) < 1e-5);
    b.destroy_data();
  });
  dr::drx::check(drx_tune_set("auto_panels", 1), "drx_tune_set");
  dr::destroy_csr_matrix_view(csr);
}

// Local columns (a band) over the same x size keep plain CSR.
TEST(SparseMatrix, BandShareKeepsPlainCsr) {
  const long n = 2000000;
  auto csr = dr::generate_band_csr<float, int>(n, 8, 8);
  mp::run_ranks(1, [&] {
    mp::distributed_sparse_matrix<float, int> e(csr, 0);
    EXPECT_TRUE(e.matrix_distribution().panels() == 0);
  });
  dr::destroy_csr_matrix_view(csr);
}

int main(int argc, char **argv) { return harness::run_all(argc > 1 ? argv[1] : nullptr); }
