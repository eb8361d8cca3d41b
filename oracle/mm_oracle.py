"""Matrix Market reader oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module; the product reader is drx_mm_read_csr in
libdrx_spmv.so.

Restates dr::__detail::read_csr_matrix_base and convert_csr_base_to_csr
(include/dr/detail/matrix_io.hpp:119-241) line by line, in pure Python loops
(small files only):
  * header checks, matrix_io.hpp:137-177 (first word "%%MatrixMarket", then
    "matrix", "coordinate"; field "pattern" -> value 1; symmetry "general" or
    "symmetric", anything else is unsupported);
  * comment skip and the size line, :179-191;
  * entry loop, :197-221 (1-based indices decremented; bounds check; push at
    (i, j) and, for a symmetric file with i != j, at (j, i); more entry lines
    than announced is an error);
  * per-row sort by column, csr_matrix_base.hpp:69-76 (std::sort, whose
    order among equal columns is unspecified: restated with a stable sort,
    i.e. push order);
  * CSR assembly, matrix_io.hpp:74-111.
Case: the reference compares header words case-sensitively; this oracle
does too (the library accepts any case, the format's own rule).

Pins (tests/test_mmio.py): scipy.io.mmread (an independent reader) on
duplicate-free general / symmetric / pattern / integer files, and
hand-computed answers for duplicate and mirrored entries and the error
cases.
"""


class MMError(ValueError):
    pass


def read_csr(path, one_indexed=True, value=float):
    with open(path) as f:
        lines = f.read().split("\n")
    pos = 0
    words = lines[pos].split()
    pos += 1
    if len(words) < 3 or words[0] != "%%MatrixMarket" or words[1] != "matrix" \
            or words[2] != "coordinate":
        raise MMError("could not be parsed as a Matrix Market file")
    pattern = len(words) > 3 and words[3] == "pattern"
    sym = words[4] if len(words) > 4 else ""
    if sym == "general":
        symmetric = False
    elif sym == "symmetric":
        symmetric = True
    else:
        raise MMError("unsupported matrix type")
    while lines[pos].startswith("%"):
        pos += 1
    m, n, nnz = (int(t) for t in lines[pos].split()[:3])
    pos += 1
    rows = [[] for _ in range(m)]
    c = 0
    for line in lines[pos:]:
        t = line.split()
        if not t:
            continue  # blank (the reference reads only the file's real lines)
        i, j = int(t[0]), int(t[1])
        v = value(1) if pattern else value(t[2])
        if one_indexed:
            i -= 1
            j -= 1
        if i < 0 or j < 0 or i >= m or j >= n:
            raise MMError("file has nonzero out of bounds")
        rows[i].append((v, j))
        if symmetric and i != j:
            rows[j].append((v, i))
        c += 1
        if c > nnz:
            raise MMError("file has more nonzeros than reported")
    for r in rows:
        r.sort(key=lambda e: e[1])
    rowptr, colind, vals = [0], [], []
    for r in rows:
        for v, j in r:
            vals.append(v)
            colind.append(j)
        rowptr.append(len(vals))
    return rowptr, colind, vals, (m, n)
