// ref_gen.cpp -- TEST INFRASTRUCTURE ONLY (fixture writer).
//
// A small driver, written for this repository, that includes the reference's
// own generator headers from their place under the reference tree and writes
// the matrices its gemv tests use, as raw little-endian arrays, so that the
// oracle's restated generators can be checked bit for bit. Built by
// oracle/Makefile into oracle/_ref/ref_gen (git-ignored); the fixtures it
// writes are committed under tests/golden/ by tests/golden/make_fixtures.py.
//
// Reference headers used: include/dr/detail/generate_csr.hpp
// (generate_random_csr :40-103, generate_band_csr :108-153) and
// include/dr/detail/matrix_io.hpp (convert_to_csr :26-72).
#include <cassert>
#include <cstdio>
#include <cstdint>
#include <string>
#include <vector>

#include <dr/detail/ranges_shim.hpp>
#include <dr/detail/iterator_adaptor.hpp>
#include <dr/detail/generate_csr.hpp>
#include <dr/detail/matrix_io.hpp>

static std::string outdir;
static FILE *manifest;

template <typename T, typename I>
static void dump(const std::string &name, std::size_t m, std::size_t n,
                 std::size_t nnz, const T *v, const I *rp, const I *ci,
                 const char *vt, const char *it) {
  auto wr = [&](const std::string &suffix, const void *p, std::size_t bytes) {
    FILE *f = std::fopen((outdir + "/" + name + suffix).c_str(), "wb");
    if (bytes) std::fwrite(p, 1, bytes, f);
    std::fclose(f);
  };
  wr(".vals", v, nnz * sizeof(T));
  wr(".rowptr", rp, (m + 1) * sizeof(I));
  wr(".colind", ci, nnz * sizeof(I));
  std::fprintf(manifest, "%s %zu %zu %zu %s %s\n", name.c_str(), m, n, nnz, vt,
               it);
}

template <typename V> static void release(V &csr) {
  delete[] csr.values_data();
  delete[] csr.rowptr_data();
  delete[] csr.colind_data();
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: ref_gen <outdir>\n");
    return 2;
  }
  outdir = argv[1];
  manifest = std::fopen((outdir + "/manifest.txt").c_str(), "w");

  // mp tests: dr::generate_random_csr({m, k}, 0.1f), T=float, I=size_t, seed 0
  // (test/gtest/mp/gemv.cpp:57,68,79,90,101,112).
  const std::size_t shapes[3][2] = {{100, 100}, {1000, 10}, {10, 1000}};
  for (auto &s : shapes) {
    auto csr = dr::generate_random_csr(dr::index<std::size_t>(s[0], s[1]),
                                       0.1f);
    dump("mp_random_" + std::to_string(s[0]) + "x" + std::to_string(s[1]),
         s[0], s[1], csr.size(), csr.values_data(), csr.rowptr_data(),
         csr.colind_data(), "f32", "u64");
    release(csr);
  }

  // sp tests: sparse_matrix<float> (I = int64_t) with a {P,1} row grid; tile i
  // is generate_random_csr<float,int64_t>({tm_i, n}, density, seed = i)
  // (include/dr/sp/containers/sparse_matrix.hpp:348-403,
  // test/gtest/sp/gemv.cpp:7-35,96-184).
  struct spcase { const char *name; long m, n; double density; };
  const spcase sp[4] = {{"sp_gemv", 100, 100, 0.1f},
                        {"sp_notsquare", 10, 1000, 0.1f},
                        {"sp_notsquare_other", 1000, 10, 0.1f},
                        {"sp_verysparse", 100, 100, 0.001f}};
  for (auto &c : sp) {
    for (long P = 1; P <= 4; P++) {
      long tm = (c.m + P - 1) / P;
      for (long i = 0; i * tm < c.m; i++) {
        long rows = std::min(tm, c.m - i * tm);
        auto csr = dr::generate_random_csr<float, std::int64_t>(
            dr::index<std::int64_t>(rows, c.n), c.density,
            static_cast<unsigned>(i));
        dump(std::string(c.name) + "_P" + std::to_string(P) + "_t" +
                 std::to_string(i),
             rows, c.n, csr.size(), csr.values_data(), csr.rowptr_data(),
             csr.colind_data(), "f32", "i64");
        release(csr);
      }
    }
  }

  // mp-bench ctest matrix: generate_band_csr<double,long>(866, 17, 17)
  // (benchmarks/gbench/mp/gemv.cpp:22-39 at --vector-size 30000), and a small
  // asymmetric band.
  {
    auto csr = dr::generate_band_csr<double, long>(866, 17, 17);
    dump("band_866_17_17", 866, 866, csr.size(), csr.values_data(),
         csr.rowptr_data(), csr.colind_data(), "f64", "i64");
    release(csr);
    auto csr2 = dr::generate_band_csr<double, long>(50, 3, 5);
    dump("band_50_3_5", 50, 50, csr2.size(), csr2.values_data(),
         csr2.rowptr_data(), csr2.colind_data(), "f64", "i64");
    release(csr2);
  }

  // sp ZeroVector matrix: dense 100x100, a_ij = i + j, through convert_to_csr
  // (test/gtest/sp/gemv.cpp:64-94).
  {
    std::vector<std::pair<std::pair<int, int>, float>> base;
    for (int i = 0; i < 100; i++)
      for (int j = 0; j < 100; j++)
        base.push_back({{i, j}, static_cast<float>(i + j)});
    auto csr = dr::__detail::convert_to_csr(base, {100, 100}, base.size(),
                                            std::allocator<float>{});
    dump("sp_zerovector_dense", 100, 100, csr.size(), csr.values_data(),
         csr.rowptr_data(), csr.colind_data(), "f32", "i32");
    std::allocator<float>{}.deallocate(csr.values_data(), csr.size());
    std::allocator<int>{}.deallocate(csr.rowptr_data(), 101);
    std::allocator<int>{}.deallocate(csr.colind_data(), csr.size());
  }
  std::fclose(manifest);
  return 0;
}
