// test_mmio.cpp -- CPU only: dr::read_csr through the library's reader, for
// the value/index types the reference's callers use (read_csr<double, long>
// in examples/mp/sparse_benchmark.cpp:58; T = float, I = int in
// examples/sp/gemv_benchmark.cpp). Prints the arrays of the file named on the
// command line; the pytest wrapper compares them with the restated reader.
#include <cstdio>
#include <memory>
#include <string>

#include "dr/detail/matrix_io.hpp"

template <typename T, typename I> static int dump(const char *path) {
  auto v = dr::read_csr<T, I>(path);
  std::printf("shape %lld %lld nnz %zu\n", (long long)v.shape()[0], (long long)v.shape()[1],
              v.size());
  std::printf("rowptr");
  for (std::size_t r = 0; r <= static_cast<std::size_t>(v.shape()[0]); r++)
    std::printf(" %lld", (long long)v.rowptr_data()[r]);
  std::printf("\ncolind");
  for (std::size_t e = 0; e < v.size(); e++) std::printf(" %lld", (long long)v.colind_data()[e]);
  std::printf("\nvals");
  for (std::size_t e = 0; e < v.size(); e++) std::printf(" %.17g", (double)v.values_data()[e]);
  std::printf("\n");
  dr::__detail::destroy_csr_matrix_view(v, std::allocator<T>{});
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 3) return 2;
  const std::string kind = argv[2];
  try {
    if (kind == "double_long") return dump<double, long>(argv[1]);
    if (kind == "float_int") return dump<float, int>(argv[1]);
    if (kind == "double_size_t") return dump<double, std::size_t>(argv[1]);
  } catch (const std::exception &e) {
    std::printf("error %s\n", e.what());
    return 1;
  }
  return 2;
}
