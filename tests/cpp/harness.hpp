// harness.hpp -- a minimal test harness for the C++ host-interface tests
// (no gtest in this image). TEST(suite, name) registers a case; main() runs
// every case, prints PASS/FAIL lines and exits non-zero on any failure.
#pragma once

#include <cmath>
#include <cstdio>
#include <functional>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

namespace harness {
struct test_case {
  std::string name;
  std::function<void()> fn;
};
inline std::vector<test_case> &registry() {
  static std::vector<test_case> r;
  return r;
}
inline int &failures() {
  static int f = 0;
  return f;
}
inline std::string &current() {
  static std::string c;
  return c;
}
struct registrar {
  registrar(const char *n, std::function<void()> f) { registry().push_back({n, std::move(f)}); }
};
inline void fail(const char *file, int line, const std::string &msg) {
  std::printf("  failure %s:%d: %s\n", file, line, msg.c_str());
  failures()++;
}

// The reference's fp_equal (test/gtest/include/common-tests.hpp:82-111): each
// element equal or |a - b| < max(min_normal, eps * (|a| + |b|)).
template <typename T>
bool fp_equal(const std::vector<T> &a, const std::vector<T> &b,
              T eps = 128 * std::numeric_limits<T>::epsilon()) {
  if (a.size() != b.size()) return false;
  for (std::size_t i = 0; i < a.size(); i++) {
    if (a[i] == b[i]) continue;
    T norm = std::min<T>(std::abs(a[i]) + std::abs(b[i]), std::numeric_limits<float>::max());
    if (!(std::abs(a[i] - b[i]) < std::max(std::numeric_limits<T>::min(), eps * norm))) return false;
  }
  return true;
}

inline int run_all(const char *filter = nullptr) {
  int n = 0;
  for (auto &t : registry()) {
    if (filter && t.name.find(filter) == std::string::npos) continue;
    current() = t.name;
    const int before = failures();
    std::printf("[ RUN  ] %s\n", t.name.c_str());
    std::fflush(stdout);
    try {
      t.fn();
    } catch (const std::exception &e) {
      fail(__FILE__, __LINE__, std::string("exception: ") + e.what());
    }
    std::printf("[ %s ] %s\n", failures() == before ? "PASS" : "FAIL", t.name.c_str());
    std::fflush(stdout);
    n++;
  }
  std::printf("%d tests, %d failures\n", n, failures());
  return failures() ? 1 : 0;
}
}  // namespace harness

#define TEST(suite, name)                                                      \
  static void suite##_##name();                                                \
  static harness::registrar reg_##suite##_##name(#suite "." #name,             \
                                                 suite##_##name);              \
  static void suite##_##name()

#define EXPECT_TRUE(c)                                                         \
  do {                                                                         \
    if (!(c)) harness::fail(__FILE__, __LINE__, #c);                           \
  } while (0)
#define EXPECT_EQ(a, b)                                                        \
  do {                                                                         \
    if (!((a) == (b))) {                                                       \
      std::ostringstream os_;                                                  \
      os_ << #a " == " #b " (" << (a) << " vs " << (b) << ")";                 \
      harness::fail(__FILE__, __LINE__, os_.str());                            \
    }                                                                          \
  } while (0)
