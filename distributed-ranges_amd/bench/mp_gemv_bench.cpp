// mp_gemv_bench.cpp -- times the reference benchmark's call shape,
// dr::mp::gemv(0, res, m, b) (benchmarks/gbench/mp/gemv.cpp:41-64, GemvEq_DR:
// a csr_eq distributed_sparse_matrix built from a host CSR, a
// broadcasted_vector b, res a host std::vector on the root), through this
// repository's host interface on one GPU (one rank).
//
// The matrix is the 16M x 16M Zipf(1.5) matrix of the headline workload
// (rows clipped to [1, 4096], stratified sorted columns, values N(0, 1),
// x U[-1, 1]), generated on the GPU by the library's generators with the
// headline's seeds and copied to the host, as a reference caller would
// hold it. Building the matrix (upload, plan, column-panel layout) is not
// timed, like the reference's benchmark fixture. One timed call includes
// what the reference's call does: the share's product, the partial's copy
// to the host, the gather into res on the root and the fence.
//
// Usage: mp_gemv_bench [rows] [reps] [lmax]. Prints one JSON line.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "dr/mp.hpp"

namespace mp = dr::mp;

int main(int argc, char **argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 16000000;
  const int reps = argc > 2 ? std::atoi(argv[2]) : 20;
  const int lmax = argc > 3 ? std::atoi(argv[3]) : 4096;
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };

  // the headline matrix and x (seeds 0 / 1, as bench.py generates them)
  std::int64_t nnz = 0;
  std::vector<int> hrp(static_cast<std::size_t>(n) + 1);
  std::vector<int> hci;
  std::vector<float> hv, x(n);
  {
    dr::drx::device_buffer<int> rp(static_cast<std::size_t>(n) + 1);
    dr::drx::check(drx_gen_zipf_rowptr(DRX_I32, rp.data(), n, 1.5, lmax, 0, &nnz, nullptr),
                   "drx_gen_zipf_rowptr");
    dr::drx::device_buffer<int> ci(static_cast<std::size_t>(nnz));
    dr::drx::device_buffer<float> v(static_cast<std::size_t>(nnz));
    dr::drx::device_buffer<float> xd(static_cast<std::size_t>(n));
    dr::drx::check(drx_gen_stratified_colind(DRX_I32, rp.data(), n, n, ci.data(), 1, nullptr),
                   "drx_gen_stratified_colind");
    dr::drx::check(drx_gen_normal(DRX_F32, v.data(), nnz, 2, nullptr), "drx_gen_normal");
    dr::drx::check(drx_gen_uniform(DRX_F32, xd.data(), n, -1.0, 1.0, 1, nullptr),
                   "drx_gen_uniform");
    hci.resize(static_cast<std::size_t>(nnz));
    hv.resize(static_cast<std::size_t>(nnz));
    rp.download(hrp.data(), hrp.size());
    ci.download(hci.data(), hci.size());
    v.download(hv.data(), hv.size());
    xd.download(x.data(), x.size());
  }
  dr::views::csr_matrix_view<float, int> csr(hv.data(), hrp.data(), hci.data(),
                                             dr::index<int>(n, n), nnz, 0);

  mp::init();
  const auto t0 = now();
  mp::distributed_sparse_matrix<float, int> m(csr, 0);  // csr_eq, as GemvEq_DR
  mp::broadcasted_vector<float> b;
  b.broadcast_data(n, 0, x, mp::default_comm());
  const auto t1 = now();
  std::vector<float> res(n);
  for (int i = 0; i < 3; i++) mp::gemv(0, res, m, b);  // warm-up
  std::vector<double> t(reps);
  for (int i = 0; i < reps; i++) {
    const auto a = now();
    mp::gemv(0, res, m, b);
    t[i] = ms(a, now());
  }
  std::sort(t.begin(), t.end());
  const double med = t[reps / 2];

  // sampled rows (incl. the longest) against a long double loop
  std::mt19937_64 rng(9);
  std::vector<int> rows(2000);
  for (auto &r : rows) r = static_cast<int>(rng() % static_cast<std::uint64_t>(n));
  int longest = 0;
  for (int r = 0; r < n; r++)
    if (hrp[r + 1] - hrp[r] > hrp[longest + 1] - hrp[longest]) longest = r;
  rows.push_back(longest);
  double worst = 0;
  for (int r : rows) {
    long double s = 0, a = 0;
    for (int q = hrp[r]; q < hrp[r + 1]; q++) {
      const long double p = static_cast<long double>(hv[q]) * x[hci[q]];
      s += p;
      a += p < 0 ? -p : p;
    }
    if (a > 0) worst = std::max(worst, std::abs(static_cast<double>(res[r] - s)) / static_cast<double>(a));
  }
  const double bytes = static_cast<double>(nnz) * 8 + 4.0 * (n + 1) + 4.0 * n + 4.0 * n;
  std::printf("{\"rows\": %d, \"nnz\": %lld, \"reps\": %d, \"ms_per_call\": %.4f, "
              "\"ms_min\": %.4f, \"GBps\": %.2f, \"panels\": %lld, \"setup_s\": %.2f, "
              "\"max_componentwise_err_sampled\": %.3e}\n",
              n, static_cast<long long>(nnz), reps, med, t[0], bytes / med / 1e6,
              static_cast<long long>(m.matrix_distribution().panels()), ms(t0, t1) / 1e3, worst);
  b.destroy_data();
  return worst < 1e-5 ? 0 : 1;
}
