// drx_mmio.cpp -- parallel Matrix Market reader producing host CSR arrays.
//
// Semantics follow dr::__detail::read_csr_matrix_base and
// convert_csr_base_to_csr (include/dr/detail/matrix_io.hpp:119-241):
//   * header "%%MatrixMarket matrix coordinate <field> <symmetry>"; field
//     "pattern" gives value 1, any other field is parsed as a number (the
//     first number after the indices); symmetry "general" or "symmetric"
//     (an off-diagonal entry (i, j) is stored at (i, j) and (j, i));
//   * "%" lines before the size line are comments; the size line is
//     "m n entries";
//   * indices are 1-based unless one_indexed == 0; an index outside the
//     shape is an error; more entry lines than announced is an error, fewer
//     is accepted;
//   * rows are sorted by column. Entries with the same (row, column) are
//     kept (not summed), in file order, a mirrored entry right after its
//     entry (the reference leaves their order to std::sort).
// Differences: header words match without regard to case (the format's own
// rule; the reference accepts lower case only); on inputs the reference does
// not define, blank lines are skipped, a line without its value is an error,
// and an index below the base is reported as out of bounds.
//
// The file is memory-mapped and its body split at line boundaries over
// `nthreads` threads; each thread parses its lines with std::from_chars
// (values parsed directly in the output type, so an F32 value is the
// correctly rounded decimal, as istream >> float gives). Rows are counted
// with atomic increments, entries scattered to their rows, and each row
// sorted by (column, position in file).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <charconv>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "drx_spmv.h"

namespace drx {
void set_last_error(const char *msg);
}

namespace {

struct MmError {
  int status;
  std::string msg;
};

[[noreturn]] void fail(int status, const std::string &m) { throw MmError{status, m}; }

struct Mapped {
  const char *p = nullptr;
  size_t n = 0;
  int fd = -1;
  ~Mapped() {
    if (p && n) munmap(const_cast<char *>(p), n);
    if (fd >= 0) close(fd);
  }
};

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r'; }

// Next whitespace-separated token of [p, e) (e = end of line).
inline bool token(const char *&p, const char *e, const char *&t0, const char *&t1) {
  while (p < e && is_space(*p)) p++;
  if (p >= e) return false;
  t0 = p;
  while (p < e && !is_space(*p)) p++;
  t1 = p;
  return true;
}

template <typename V> bool parse_num(const char *t0, const char *t1, V &out) {
  if (t0 < t1 && *t0 == '+') t0++;
  auto r = std::from_chars(t0, t1, out);
  return r.ec == std::errc();
}

// A value token: the longest numeric prefix, as istream extraction reads
// it. Results below the normal range (from_chars reports them out of range)
// take strtod/strtof's gradual underflow, as istream's do.
template <typename V> bool parse_value(const char *t0, const char *t1, V &out) {
  if (t0 < t1 && *t0 == '+') t0++;
  const auto r = std::from_chars(t0, t1, out);
  if (r.ec == std::errc()) return true;
  if (r.ec != std::errc::result_out_of_range || t1 - t0 > 200) return false;
  char buf[256];
  std::memcpy(buf, t0, t1 - t0);
  buf[t1 - t0] = 0;
  char *stop = nullptr;
  if constexpr (std::is_same_v<V, float>)
    out = std::strtof(buf, &stop);
  else
    out = std::strtod(buf, &stop);
  return stop != buf;
}

struct Entry {
  int64_t col;
  int64_t seq;  // 2 * line + (mirrored ? 1 : 0)
  double val;
};

struct Parsed {
  std::vector<int64_t> row, col;
  std::vector<double> val;
  int64_t lines = 0;
  bool bad_line = false, oob = false;
};

template <typename V>
void parse_chunk(const char *p, const char *e, bool pattern, bool symmetric,
                 bool one_indexed, int64_t m, int64_t n, Parsed &out) {
  while (p < e) {
    const char *le = static_cast<const char *>(std::memchr(p, '\n', e - p));
    if (!le) le = e;
    const char *q = p;
    const char *a0, *a1;
    if (token(q, le, a0, a1)) {
      const char *b0, *b1;
      int64_t i = 0, j = 0;
      V v = V(1);
      bool ok = parse_num(a0, a1, i) && token(q, le, b0, b1) && parse_num(b0, b1, j);
      if (ok && !pattern) {
        const char *c0, *c1;
        ok = token(q, le, c0, c1);
        if (ok) ok = parse_value(c0, c1, v);
      }
      if (!ok) {
        out.bad_line = true;
        return;
      }
      if (one_indexed) {
        i--;
        j--;
      }
      if (i < 0 || j < 0 || i >= m || j >= n) out.oob = true;
      if (symmetric && (j >= m || i >= n)) out.oob = true;  // mirrored entry
      out.row.push_back(i);
      out.col.push_back(j);
      out.val.push_back(static_cast<double>(v));
      out.lines++;
    }
    p = le + 1;
  }
}

template <typename TI, typename TV>
void emit(const std::vector<int64_t> &rowptr, const std::vector<Entry> &ent,
          void **rp_out, void **ci_out, void **v_out) {
  const size_t m1 = rowptr.size(), nnz = ent.size();
  TI *rp = static_cast<TI *>(std::malloc(std::max<size_t>(m1, 1) * sizeof(TI)));
  TI *ci = static_cast<TI *>(std::malloc(std::max<size_t>(nnz, 1) * sizeof(TI)));
  TV *vv = static_cast<TV *>(std::malloc(std::max<size_t>(nnz, 1) * sizeof(TV)));
  if (!rp || !ci || !vv) {
    std::free(rp);
    std::free(ci);
    std::free(vv);
    fail(DRX_ERR_NOMEM, "host allocation failed");
  }
  for (size_t r = 0; r < m1; r++) rp[r] = static_cast<TI>(rowptr[r]);
  for (size_t e = 0; e < nnz; e++) {
    ci[e] = static_cast<TI>(ent[e].col);
    vv[e] = static_cast<TV>(ent[e].val);
  }
  *rp_out = rp;
  *ci_out = ci;
  *v_out = vv;
}

void read_csr(const char *path, int one_indexed, int itype, int vtype,
              int nthreads, int64_t *nrows, int64_t *ncols, int64_t *nnz_out,
              void **rowptr_out, void **colind_out, void **vals_out) {
  if (!path || !nrows || !ncols || !nnz_out || !rowptr_out || !colind_out || !vals_out)
    fail(DRX_ERR_INVALID, "null argument");
  if (itype != DRX_I32 && itype != DRX_I64) fail(DRX_ERR_INVALID, "itype must be I32 or I64");
  if (vtype != DRX_F64 && vtype != DRX_F32) fail(DRX_ERR_INVALID, "vtype must be F64 or F32");
  const std::string fp(path);
  Mapped f;
  f.fd = open(path, O_RDONLY);
  if (f.fd < 0) fail(DRX_ERR_INVALID, "mmread: cannot open " + fp);
  struct stat st;
  if (fstat(f.fd, &st) != 0) fail(DRX_ERR_INVALID, "mmread: cannot open " + fp);
  f.n = static_cast<size_t>(st.st_size);
  if (f.n > 0) {
    void *m = mmap(nullptr, f.n, PROT_READ, MAP_PRIVATE, f.fd, 0);
    if (m == MAP_FAILED) fail(DRX_ERR_INVALID, "mmread: cannot map " + fp);
    f.p = static_cast<const char *>(m);
    madvise(m, f.n, MADV_SEQUENTIAL);
  }
  const char *p = f.p, *end = f.p + f.n;
  auto next_line = [&](const char *&lb, const char *&le) -> bool {
    if (p >= end) return false;
    lb = p;
    le = static_cast<const char *>(std::memchr(p, '\n', end - p));
    if (!le) le = end;
    p = le < end ? le + 1 : end;
    return true;
  };
  const std::string not_mm = fp + " could not be parsed as a Matrix Market file.";
  const char *lb, *le;
  if (!next_line(lb, le)) fail(DRX_ERR_INVALID, not_mm);
  std::vector<std::string> words;
  {
    const char *q = lb, *t0, *t1;
    while (token(q, le, t0, t1)) words.emplace_back(t0, t1);
  }
  auto lower = [](std::string s) {
    for (auto &c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
  };
  if (words.size() < 3 || words[0] != "%%MatrixMarket" || lower(words[1]) != "matrix" ||
      lower(words[2]) != "coordinate")
    fail(DRX_ERR_INVALID, not_mm);
  const bool pattern = words.size() > 3 && lower(words[3]) == "pattern";
  const std::string sym = words.size() > 4 ? lower(words[4]) : "";
  bool symmetric;
  if (sym == "general")
    symmetric = false;
  else if (sym == "symmetric")
    symmetric = true;
  else
    fail(DRX_ERR_INVALID, fp + " has an unsupported matrix type");

  // comments, then the size line
  int64_t m = -1, n = -1, announced = -1;
  for (;;) {
    if (!next_line(lb, le)) fail(DRX_ERR_INVALID, not_mm);
    const char *q = lb, *t0, *t1;
    if (!token(q, le, t0, t1)) continue;  // blank
    if (*t0 == '%') continue;
    bool ok = parse_num(t0, t1, m);
    const char *u0, *u1, *w0, *w1;
    ok = ok && token(q, le, u0, u1) && parse_num(u0, u1, n) && token(q, le, w0, w1) &&
         parse_num(w0, w1, announced);
    if (!ok || m < 0 || n < 0 || announced < 0) fail(DRX_ERR_INVALID, not_mm);
    break;
  }
  if (itype == DRX_I32 && (m > INT32_MAX || n > INT32_MAX))
    fail(DRX_ERR_INVALID, "shape exceeds the 32-bit index type");

  // split the body at line boundaries
  const char *body = p;
  unsigned T = nthreads > 0 ? static_cast<unsigned>(nthreads) : std::thread::hardware_concurrency();
  T = std::max(1u, std::min(T, 256u));
  const size_t blen = static_cast<size_t>(end - body);
  if (blen < (size_t(1) << 20)) T = 1;
  std::vector<const char *> cut(T + 1);
  cut[0] = body;
  cut[T] = end;
  for (unsigned t = 1; t < T; t++) {
    const char *c = body + blen * t / T;
    if (c < cut[t - 1]) c = cut[t - 1];
    const char *nl = static_cast<const char *>(std::memchr(c, '\n', end - c));
    cut[t] = nl ? nl + 1 : end;
  }
  std::vector<Parsed> parts(T);
  {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; t++)
      th.emplace_back([&, t] {
        if (vtype == DRX_F32)
          parse_chunk<float>(cut[t], cut[t + 1], pattern, symmetric, one_indexed != 0, m, n,
                             parts[t]);
        else
          parse_chunk<double>(cut[t], cut[t + 1], pattern, symmetric, one_indexed != 0, m, n,
                              parts[t]);
      });
    for (auto &x : th) x.join();
  }
  int64_t lines = 0;
  std::vector<int64_t> line0(T);
  for (unsigned t = 0; t < T; t++) {
    if (parts[t].bad_line) fail(DRX_ERR_INVALID, fp + ": malformed entry line");
    if (parts[t].oob) fail(DRX_ERR_INVALID, "read_MatrixMarket: file has nonzero out of bounds.");
    line0[t] = lines;
    lines += parts[t].lines;
  }
  if (lines > announced)
    fail(DRX_ERR_INVALID,
         "read_MatrixMarket: error reading Matrix Market file, file has more "
         "nonzeros than reported.");

  // rows: counts (atomic), offsets, scatter, per-row sort
  std::vector<int64_t> count(static_cast<size_t>(m) + 1, 0);
  auto each_entry = [&](unsigned t, auto &&fn) {
    const Parsed &P = parts[t];
    for (int64_t e = 0; e < P.lines; e++) {
      const int64_t i = P.row[e], j = P.col[e], seq = 2 * (line0[t] + e);
      fn(i, j, seq, P.val[e]);
      if (symmetric && i != j) fn(j, i, seq + 1, P.val[e]);
    }
  };
  {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; t++)
      th.emplace_back([&, t] {
        each_entry(t, [&](int64_t i, int64_t, int64_t, double) {
          std::atomic_ref<int64_t>(count[i]).fetch_add(1, std::memory_order_relaxed);
        });
      });
    for (auto &x : th) x.join();
  }
  std::vector<int64_t> rowptr(static_cast<size_t>(m) + 1, 0);
  for (int64_t r = 0; r < m; r++) rowptr[r + 1] = rowptr[r] + count[r];
  const int64_t nnz = rowptr[m];
  if (itype == DRX_I32 && nnz > INT32_MAX)
    fail(DRX_ERR_INVALID, "entries exceed the 32-bit index type");
  std::vector<Entry> ent(static_cast<size_t>(nnz));
  std::copy(rowptr.begin(), rowptr.end() - 1, count.begin());  // cursors
  {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; t++)
      th.emplace_back([&, t] {
        each_entry(t, [&](int64_t i, int64_t j, int64_t seq, double v) {
          const int64_t at =
              std::atomic_ref<int64_t>(count[i]).fetch_add(1, std::memory_order_relaxed);
          ent[at] = Entry{j, seq, v};
        });
      });
    for (auto &x : th) x.join();
  }
  parts.clear();
  parts.shrink_to_fit();
  {
    std::atomic<int64_t> next{0};
    const int64_t step = 4096;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; t++)
      th.emplace_back([&] {
        for (int64_t r0 = next.fetch_add(step); r0 < m; r0 = next.fetch_add(step))
          for (int64_t r = r0; r < std::min(m, r0 + step); r++)
            std::sort(ent.begin() + rowptr[r], ent.begin() + rowptr[r + 1],
                      [](const Entry &a, const Entry &b) {
                        return a.col != b.col ? a.col < b.col : a.seq < b.seq;
                      });
      });
    for (auto &x : th) x.join();
  }
  void *rp = nullptr, *ci = nullptr, *vv = nullptr;
  if (itype == DRX_I32) {
    if (vtype == DRX_F64)
      emit<int32_t, double>(rowptr, ent, &rp, &ci, &vv);
    else
      emit<int32_t, float>(rowptr, ent, &rp, &ci, &vv);
  } else {
    if (vtype == DRX_F64)
      emit<int64_t, double>(rowptr, ent, &rp, &ci, &vv);
    else
      emit<int64_t, float>(rowptr, ent, &rp, &ci, &vv);
  }
  *nrows = m;
  *ncols = n;
  *nnz_out = nnz;
  *rowptr_out = rp;
  *colind_out = ci;
  *vals_out = vv;
}

}  // namespace

extern "C" {

int drx_mm_read_csr(const char *path, int one_indexed, int itype, int vtype,
                    int nthreads, int64_t *nrows, int64_t *ncols, int64_t *nnz,
                    void **rowptr, void **colind, void **vals) {
  try {
    read_csr(path, one_indexed, itype, vtype, nthreads, nrows, ncols, nnz, rowptr,
             colind, vals);
    return DRX_OK;
  } catch (const MmError &e) {
    drx::set_last_error(e.msg.c_str());
    return e.status;
  } catch (const std::bad_alloc &) {
    drx::set_last_error("host allocation failed");
    return DRX_ERR_NOMEM;
  } catch (const std::exception &e) {
    drx::set_last_error(e.what());
    return DRX_ERR_INVALID;
  }
}

void drx_host_free(void *p) { std::free(p); }

}  // extern "C"
