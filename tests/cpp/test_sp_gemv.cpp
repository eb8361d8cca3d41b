// test_sp_gemv.cpp -- the reference's dr::sp gemv tests
// (test/gtest/sp/gemv.cpp:7-184) replayed through this host interface with
// 1, 2, 3 and 4 logical devices on the one GPU (the reference's
// --devicesCount duplication, test/gtest/sp/sp-tests.cpp:31-42).
#include <cmath>
#include <cstdint>
#include <vector>

#include "dr/sp.hpp"
#include "harness.hpp"

namespace sp = dr::sp;

template <typename T, typename I>
void check_rowsums(long m, long k, sp::sparse_matrix<T, I> &a) {
  sp::distributed_vector<T> b(k, 1.f);
  sp::distributed_vector<T> c(m, 0.f);
  sp::gemv(c, a, b);
  std::vector<T> c_local(m);
  sp::copy(c.begin(), c.end(), c_local.begin());
  std::vector<T> c_ref(m, 0.f);
  for (auto &&[index, v] : a) c_ref[index[0]] += v;
  EXPECT_TRUE(harness::fp_equal(c_ref, c_local));
}

void random_case(long m, long k, double density) {
  for (std::size_t P = 1; P <= 4; P++) {
    sp::init(P);
    sp::sparse_matrix<float> a({m, k}, density,
                               sp::block_cyclic({sp::tile::div, sp::tile::div}, {sp::nprocs(), 1}));
    check_rowsums(m, k, a);
    sp::finalize();
  }
}

TEST(SparseMatrix, Gemv) { random_case(100, 100, 0.1f); }
TEST(SparseMatrix, NotSquareMatrix) { random_case(10, 1000, 0.1f); }
TEST(SparseMatrix, NotSquareMatrixOtherAxis) { random_case(1000, 10, 0.1f); }
TEST(SparseMatrix, VerySparseMatrix) { random_case(100, 100, 0.001f); }

TEST(SparseMatrix, EmptyGemv) {
  for (std::size_t P = 1; P <= 4; P++) {
    sp::init(P);
    std::vector<std::pair<std::pair<int, int>, float>> base;
    auto csr = dr::__detail::convert_to_csr(base, {100, 100}, base.size(), std::allocator<float>{});
    auto a = sp::create_distributed(csr, sp::row_cyclic());
    sp::distributed_vector<float> b(100, 1.f);
    sp::distributed_vector<float> c(100, 0.f);
    sp::gemv(c, a, b);
    std::vector<float> c_local(100);
    sp::copy(c.begin(), c.end(), c_local.begin());
    EXPECT_TRUE(harness::fp_equal(std::vector<float>(100, 0.f), c_local));
    std::allocator<float>{}.deallocate(csr.values_data(), 0);
    std::allocator<int>{}.deallocate(csr.rowptr_data(), 101);
    std::allocator<int>{}.deallocate(csr.colind_data(), 0);
    sp::finalize();
  }
}

TEST(SparseMatrix, ZeroVector) {
  for (std::size_t P = 1; P <= 4; P++) {
    sp::init(P);
    std::vector<std::pair<std::pair<int, int>, float>> base;
    for (int i = 0; i < 100; i++)
      for (int j = 0; j < 100; j++) base.push_back({{i, j}, static_cast<float>(i + j)});
    auto csr = dr::__detail::convert_to_csr(base, {100, 100}, base.size(), std::allocator<float>{});
    auto a = sp::create_distributed(csr, sp::row_cyclic());
    sp::distributed_vector<float> b(100, 0.f);
    sp::distributed_vector<float> c(100, 0.f);
    sp::gemv(c, a, b);
    std::vector<float> c_local(100);
    sp::copy(c.begin(), c.end(), c_local.begin());
    EXPECT_TRUE(harness::fp_equal(std::vector<float>(100, 0.f), c_local));
    // and with b = 1: row sums of a_ij = i + j are 100 i + 4950 (exact)
    sp::distributed_vector<float> b1(100, 1.f);
    sp::distributed_vector<float> c1(100, 0.f);
    sp::gemv(c1, a, b1);
    sp::copy(c1.begin(), c1.end(), c_local.begin());
    bool ok = true;
    for (int i = 0; i < 100; i++) ok = ok && c_local[i] == 100.f * i + 4950.f;
    EXPECT_TRUE(ok);
    std::allocator<float>{}.deallocate(csr.values_data(), csr.size());
    std::allocator<int>{}.deallocate(csr.rowptr_data(), 101);
    std::allocator<int>{}.deallocate(csr.colind_data(), csr.size());
    sp::finalize();
  }
}

// gemv accumulates into c (alpha = beta = 1, as the reference's MKL call).
TEST(SparseMatrix, GemvAccumulates) {
  sp::init(2);
  sp::sparse_matrix<float> a({100, 100}, 0.1f);
  sp::distributed_vector<float> b(100, 1.f);
  sp::distributed_vector<float> c(100, 0.f);
  sp::gemv(c, a, b);
  sp::gemv(c, a, b);
  std::vector<float> c_local(100);
  sp::copy(c.begin(), c.end(), c_local.begin());
  std::vector<float> c_ref(100, 0.f);
  for (auto &&[index, v] : a) c_ref[index[0]] += v;
  for (auto &v : c_ref) v *= 2;
  EXPECT_TRUE(harness::fp_equal(c_ref, c_local));
  sp::finalize();
}

// The 2-D tile grid forms (gemv.hpp:126-207: gemv_square, gemv_square_copy)
// and flat_gemv (gemv.hpp:19-76), as examples/sp/gemv_benchmark.cpp calls
// them: row sums of a random matrix, accumulated twice.
template <typename F> void square_case(long m, long k, F &&fn, bool square_grid) {
  for (std::size_t P = 1; P <= 4; P++) {
    sp::init(P);
    const std::size_t g1 = square_grid ? sp::nprocs() : 1;
    sp::sparse_matrix<float> a({m, k}, 0.05f,
                               sp::block_cyclic({sp::tile::div, sp::tile::div}, {sp::nprocs(), g1}));
    sp::distributed_vector<float> b(k, 1.f);
    sp::distributed_vector<float> c(m, 0.f);
    fn(c, a, b);
    fn(c, a, b);
    std::vector<float> c_local(m);
    sp::copy(c.begin(), c.end(), c_local.begin());
    std::vector<float> c_ref(m, 0.f);
    for (auto &&[index, v] : a) c_ref[index[0]] += 2 * v;
    EXPECT_TRUE(harness::fp_equal(c_ref, c_local));
    sp::finalize();
  }
}

TEST(SparseMatrix, GemvSquare) {
  square_case(400, 400, [](auto &c, auto &a, auto &b) { sp::gemv_square(c, a, b); }, true);
}
TEST(SparseMatrix, GemvSquareCopy) {
  square_case(400, 400, [](auto &c, auto &a, auto &b) { sp::gemv_square_copy(c, a, b); }, true);
}
TEST(SparseMatrix, FlatGemv) {
  square_case(300, 500, [](auto &c, auto &a, auto &b) { sp::flat_gemv(c, a, b); }, false);
}

// Tiles keep their plans across products, and a tile whose columns scatter
// over an x larger than one XCD's L2 (2M columns of FP32) takes the column
// panel layout: two products with the same b add up to twice one product
// (c += A b accumulates), within the FP32 tolerance of a long double loop.
TEST(SparseMatrix, ZipfTilesTakeColumnPanels) {
  const int m = 1000000, n = 2000000;
  dr::drx::device_buffer<int> rp(m + 1);
  std::int64_t nnz = 0;
  dr::drx::check(drx_gen_zipf_rowptr(DRX_I32, rp.data(), m, 1.5, 256, 5, &nnz, nullptr), "zipf");
  dr::drx::device_buffer<int> ci(nnz);
  dr::drx::device_buffer<float> v(nnz);
  dr::drx::check(drx_gen_stratified_colind(DRX_I32, rp.data(), m, n, ci.data(), 6, nullptr), "cols");
  dr::drx::check(drx_gen_normal(DRX_F32, v.data(), nnz, 7, nullptr), "vals");
  std::vector<int> hrp(m + 1), hci(nnz);
  std::vector<float> hv(nnz);
  rp.download(hrp.data(), m + 1);
  ci.download(hci.data(), nnz);
  v.download(hv.data(), nnz);
  dr::views::csr_matrix_view<float, int> csr(hv.data(), hrp.data(), hci.data(),
                                             dr::index<int>(m, n), nnz, 0);
  std::vector<float> x(n);
  for (int j = 0; j < n; j++) x[j] = static_cast<float>(std::cos(0.002 * j));
  for (std::size_t P = 1; P <= 2; P++) {
    sp::init(P);
    auto a = sp::create_distributed(csr, sp::row_cyclic());
    sp::distributed_vector<float> b(n, 0.f);
    sp::copy(x.begin(), x.end(), b.begin());
    sp::distributed_vector<float> c(m, 0.f);
    sp::gemv(c, a, b);
    for (std::size_t i = 0; i < a.grid_shape()[0]; i++) {
      std::int64_t np = 0;
      dr::drx::check(drx_plan_panels(a.tile_plan(dr::index<int>(i, 0), nullptr), &np, nullptr,
                                     nullptr), "drx_plan_panels");
      EXPECT_TRUE(np >= 4);
    }
    sp::gemv(c, a, b);
    std::vector<float> y(m);
    sp::copy(c.begin(), c.end(), y.begin());
    double worst = 0;
    for (int r = 0; r < m; r++) {
      long double s = 0, t = 0;
      for (int q = hrp[r]; q < hrp[r + 1]; q++) {
        const long double p = static_cast<long double>(hv[q]) * x[hci[q]];
        s += p;
        t += p < 0 ? -p : p;
      }
      const double d = std::abs(static_cast<double>(y[r] - 2 * s));
      if (t > 0) worst = std::max(worst, d / static_cast<double>(2 * t));
    }
    EXPECT_TRUE(worst < 2e-5);
    sp::finalize();
  }
}

int main(int argc, char **argv) { return harness::run_all(argc > 1 ? argv[1] : nullptr); }
