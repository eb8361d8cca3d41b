"""The Matrix Market reader (drx_mm_read_csr, host code) against the
restated reference reader (oracle/mm_oracle.py) and scipy.io.mmread.

CPU only: the reader does no device work.
"""
import os
import sys

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

import drx

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import mm_oracle  # noqa: E402


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def same(got, ref):
    rp, ci, v, shape = got
    rrp, rci, rv, rshape = ref
    assert shape == rshape
    assert np.array_equal(rp, np.asarray(rrp, rp.dtype))
    assert np.array_equal(ci, np.asarray(rci, ci.dtype))
    assert np.array_equal(v, np.asarray(rv, v.dtype))


@pytest.mark.parametrize("kind", ["general", "symmetric"])
@pytest.mark.parametrize("field", ["real", "integer", "pattern"])
def test_against_scipy_and_oracle(tmp_path, kind, field):
    rng = np.random.default_rng(len(kind) * 10 + len(field))
    A = sp.random(300, 300 if kind == "symmetric" else 250, density=0.03, random_state=rng,
                  format="coo")
    if kind == "symmetric":
        A = sp.triu(A + A.T).tocoo()
    if field == "integer":
        A.data = rng.integers(-50, 50, A.nnz).astype(float)
    elif field == "real":
        A.data = rng.standard_normal(A.nnz)
    path = str(tmp_path / "a.mtx")
    scipy.io.mmwrite(path, A if field != "pattern" else (A != 0).astype(float).tocoo(),
                     field=field, symmetry=kind)
    ref = scipy.io.mmread(path).tocsr()
    ref.sort_indices()
    orc = mm_oracle.read_csr(path)
    # the oracle equals scipy on duplicate-free files (scipy sums duplicates)
    assert np.array_equal(orc[0], ref.indptr) and np.array_equal(orc[1], ref.indices)
    assert np.array_equal(np.asarray(orc[2]), ref.data)
    for nt in (1, 3, 0):
        got = drx.read_csr_host(path, np.float64, np.int64, nthreads=nt)
        same(got, orc)
    got32 = drx.read_csr_host(path, np.float32, np.int32)
    same(got32, mm_oracle.read_csr(path, value=np.float32))


def test_duplicates_mirrors_and_comments(tmp_path):
    text = """%%MatrixMarket matrix coordinate real symmetric
% a comment
%another
4 4 6
2 1 1.5
1 1 2.0
2 1 -0.25
4 2 3
3 3 +4e-1
4 4 7
"""
    path = write(tmp_path, "d.mtx", text)
    got = drx.read_csr_host(path)
    # row 0: (0,0)=2, mirrored (0,1)=1.5 then (0,1)=-0.25 (file order)
    assert got[0].tolist() == [0, 3, 6, 7, 9]
    assert got[1].tolist() == [0, 1, 1, 0, 0, 3, 2, 1, 3]
    assert got[2].tolist() == [2.0, 1.5, -0.25, 1.5, -0.25, 3.0, 0.4, 3.0, 7.0]
    same(got, mm_oracle.read_csr(path))


def test_zero_indexed_and_fewer_lines(tmp_path):
    path = write(tmp_path, "z.mtx", "%%MatrixMarket matrix coordinate real general\n"
                                    "3 5 4\n0 4 1\n2 0 2\n")
    got = drx.read_csr_host(path, one_indexed=False)
    assert got[3] == (3, 5)
    assert got[0].tolist() == [0, 1, 1, 2] and got[1].tolist() == [4, 0]
    same(got, mm_oracle.read_csr(path, one_indexed=False))


def test_float32_values_are_correctly_rounded(tmp_path):
    vals = ["0.1", "3.4028234e38", "1.00000005960464477539", "-7.006492e-46", "16777217"]
    body = "".join("1 %d %s\n" % (i + 1, v) for i, v in enumerate(vals))
    path = write(tmp_path, "f.mtx", "%%MatrixMarket matrix coordinate real general\n"
                                    "1 5 5\n" + body)
    got = drx.read_csr_host(path, np.float32, np.int32)
    assert got[2].tolist() == [float(np.float32(v)) for v in vals]


@pytest.mark.parametrize("text,msg", [
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", "could not be parsed"),
    ("%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 1\n", "could not be parsed"),
    ("%%MatrixMarket matrix coordinate real hermitian\n1 1 1\n1 1 1\n", "unsupported"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n", "out of bounds"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n0 1 1\n", "out of bounds"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 1\n2 2 1\n", "more nonzeros"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1\n", "malformed"),
])
def test_errors(tmp_path, text, msg):
    path = write(tmp_path, "e.mtx", text)
    with pytest.raises(drx.DrxError, match=msg):
        drx.read_csr_host(path)
    if "malformed" not in msg:
        with pytest.raises(mm_oracle.MMError):
            mm_oracle.read_csr(path)
    with pytest.raises(drx.DrxError, match="cannot open"):
        drx.read_csr_host(str(tmp_path / "missing.mtx"))


def test_large_file_threads_agree(tmp_path):
    """A file of several MB is split over threads at line boundaries; every
    thread count gives the same arrays, equal to scipy's."""
    rng = np.random.default_rng(7)
    A = sp.random(20000, 15000, density=0.002, random_state=rng, format="coo")
    A.data = rng.standard_normal(A.nnz)
    path = str(tmp_path / "big.mtx")
    scipy.io.mmwrite(path, A, field="real", symmetry="general")
    assert os.path.getsize(path) > (1 << 20)
    ref = A.tocsr()
    ref.sort_indices()
    base = drx.read_csr_host(path, nthreads=1)
    assert np.array_equal(base[0], ref.indptr) and np.array_equal(base[1], ref.indices)
    assert np.array_equal(base[2], ref.data)
    for nt in (2, 7, 16):
        same(drx.read_csr_host(path, nthreads=nt), base)
