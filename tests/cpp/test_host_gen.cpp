// test_host_gen.cpp -- CPU only: the host interface's generators
// (dr::generate_random_csr / dr::generate_band_csr) print FNV-1a64 hashes of
// their arrays for the matrices the reference's tests use; the pytest
// wrapper compares them with tests/golden/manifest.json, written from the
// reference's own generator.
#include <cstdint>
#include <cstdio>
#include <string>

#include "dr/detail/generate_csr.hpp"

static std::uint64_t fnv(const void *p, std::size_t n) {
  std::uint64_t h = 0xcbf29ce484222325ull;
  auto *b = static_cast<const unsigned char *>(p);
  for (std::size_t i = 0; i < n; i++) {
    h ^= b[i];
    h *= 0x100000001b3ull;
  }
  return h;
}

template <typename V> static void emit(const std::string &name, V csr, std::size_t m) {
  using T = typename V::scalar_type;
  using I = typename V::index_type;
  std::printf("%s %zu %016llx %016llx %016llx\n", name.c_str(), csr.size(),
              (unsigned long long)fnv(csr.rowptr_data(), (m + 1) * sizeof(I)),
              (unsigned long long)fnv(csr.colind_data(), csr.size() * sizeof(I)),
              (unsigned long long)fnv(csr.values_data(), csr.size() * sizeof(T)));
  dr::destroy_csr_matrix_view(csr);
}

int main() {
  const std::size_t shapes[3][2] = {{100, 100}, {1000, 10}, {10, 1000}};
  for (auto &s : shapes)
    emit("mp_random_" + std::to_string(s[0]) + "x" + std::to_string(s[1]),
         dr::generate_random_csr<float, std::size_t>(dr::index<std::size_t>(s[0], s[1]), 0.1f),
         s[0]);
  for (long P = 1; P <= 4; P++) {
    long tm = (100 + P - 1) / P;
    for (long i = 0; i * tm < 100; i++) {
      long rows = std::min(tm, 100 - i * tm);
      emit("sp_gemv_P" + std::to_string(P) + "_t" + std::to_string(i),
           dr::generate_random_csr<float, std::int64_t>(dr::index<std::int64_t>(rows, 100), 0.1f,
                                                        static_cast<unsigned>(i)),
           rows);
    }
  }
  emit("band_866_17_17", dr::generate_band_csr<double, long>(866, 17, 17), 866);
  emit("band_50_3_5", dr::generate_band_csr<double, long>(50, 3, 5), 50);
  return 0;
}
