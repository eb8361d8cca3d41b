// test_sp_gemm.cpp -- dr::sp::gemm over dense tile grids, checked the way
// the reference's example checks it (examples/sp/gemm_example.cpp:12-95:
// a[i][k] = (i + k) / m, b[k][j] = (k + j) / k, against a serial triple
// loop with a relative tolerance of 128 epsilon), with 1, 2, 3 and 4 logical
// devices on the one GPU and the partitions partition_matmul chooses.
#include <cmath>
#include <limits>
#include <vector>

#include "dr/sp.hpp"
#include "harness.hpp"

namespace sp = dr::sp;

template <typename T>
std::vector<T> serial(std::size_t m, std::size_t n, std::size_t k, const std::vector<T> &a,
                      const std::vector<T> &b) {
  std::vector<T> c(m * n, T(0));
  for (std::size_t i = 0; i < m; i++)
    for (std::size_t q = 0; q < k; q++)
      for (std::size_t j = 0; j < n; j++) c[i * n + j] += a[i * k + q] * b[q * n + j];
  return c;
}

template <typename T> bool close(const std::vector<T> &x, const std::vector<T> &y) {
  const T eps = 128 * std::numeric_limits<T>::epsilon();
  for (std::size_t i = 0; i < x.size(); i++) {
    if (x[i] == y[i]) continue;
    const T d = std::abs(x[i] - y[i]);
    const T norm = std::min(std::abs(x[i]) + std::abs(y[i]), std::numeric_limits<T>::max());
    if (d >= eps * std::max(norm, std::numeric_limits<T>::min())) return false;
  }
  return true;
}

template <typename T, typename Gemm>
void example_case(std::size_t m, std::size_t n, std::size_t k, Gemm &&gemm) {
  std::vector<T> a_h(m * k), b_h(k * n);
  for (std::size_t i = 0; i < m; i++)
    for (std::size_t q = 0; q < k; q++) a_h[i * k + q] = static_cast<T>(i + q) / m;
  for (std::size_t q = 0; q < k; q++)
    for (std::size_t j = 0; j < n; j++) b_h[q * n + j] = static_cast<T>(q + j) / k;
  const auto c_ref = serial(m, n, k, a_h, b_h);
  for (std::size_t P = 1; P <= 4; P++) {
    sp::init(P);
    auto parts = sp::partition_matmul(m, n, k);
    sp::distributed_dense_matrix<T> a({m, k}, parts[0]);
    sp::distributed_dense_matrix<T> b({k, n}, parts[1]);
    sp::distributed_dense_matrix<T> c({m, n}, parts[2]);
    a.copy_from_host(a_h.data());
    b.copy_from_host(b_h.data());
    gemm(a, b, c);
    std::vector<T> c_h(m * n);
    c.copy_to_host(c_h.data());
    EXPECT_TRUE(close(c_ref, c_h));
    sp::finalize();
  }
}

TEST(DenseMatrix, GemmExample) {
  example_case<float>(100, 100, 100, [](auto &a, auto &b, auto &c) { sp::gemm(a, b, c); });
  example_case<float>(130, 70, 210, [](auto &a, auto &b, auto &c) { sp::gemm(a, b, c); });
  example_case<double>(97, 301, 45, [](auto &a, auto &b, auto &c) { sp::gemm(a, b, c); });
}

TEST(DenseMatrix, GemmInplaceAndBufferedAsync) {
  example_case<float>(64, 200, 96, [](auto &a, auto &b, auto &c) { sp::gemm_inplace(a, b, c); });
  example_case<double>(200, 64, 96,
                       [](auto &a, auto &b, auto &c) { sp::gemm_buffered_async(a, b, c); });
}

TEST(DenseMatrix, GemmAccumulatesIntoC) {
  // c += a b, as oneMKL's gemm with beta = 1 (and custom_gemm) do
  sp::init(2);
  auto parts = sp::partition_matmul(40, 40, 40);
  sp::distributed_dense_matrix<double> a({40, 40}, parts[0]);
  sp::distributed_dense_matrix<double> b({40, 40}, parts[1]);
  sp::distributed_dense_matrix<double> c({40, 40}, parts[2]);
  for (std::size_t i = 0; i < 40; i++) {
    a[{i, i}] = 2.0;
    b[{i, (i + 1) % 40}] = 1.5;
    c[{i, 0}] = 1.0;
  }
  sp::gemm(a, b, c);
  std::vector<double> h(1600);
  c.copy_to_host(h.data());
  bool ok = true;
  for (std::size_t i = 0; i < 40; i++)
    for (std::size_t j = 0; j < 40; j++) {
      const double want = (j == (i + 1) % 40 ? 3.0 : 0.0) + (j == 0 ? 1.0 : 0.0);
      ok = ok && h[i * 40 + j] == want;
    }
  EXPECT_TRUE(ok);
  EXPECT_TRUE(static_cast<double>(c[{5, 6}]) == 3.0);
  sp::finalize();
}

TEST(DenseMatrix, TilesAndPartitions) {
  sp::init(4);
  EXPECT_TRUE(sp::detail::factor(4)[0] == 2 && sp::detail::factor(4)[1] == 2);
  EXPECT_TRUE(sp::detail::factor(3)[0] == 3 && sp::detail::factor(3)[1] == 1);
  sp::distributed_dense_matrix<float> d({10, 7});
  EXPECT_TRUE(d.grid_shape()[0] == 2 && d.grid_shape()[1] == 2);
  EXPECT_TRUE(d.tile_shape()[0] == 5 && d.tile_shape()[1] == 4);
  auto t = d.tile({1, 1});
  EXPECT_TRUE(t.shape()[0] == 5 && t.shape()[1] == 3 && t.ld() % 32 == 0);
  std::size_t ranks = 0;
  for (auto &v : d.tiles()) ranks |= std::size_t(1) << v.rank();
  EXPECT_TRUE(ranks == 0xf);
  sp::finalize();
}

int main(int argc, char **argv) { return harness::run_all(argc > 1 ? argv[1] : nullptr); }
