"""Pins of the CPU oracle (oracle/oracle.cpp) -- CPU only.

Each restated function is checked against something that does not share its
formula: fixtures written by the reference's own generator (compiled from the
reference sources into oracle/_ref/ref_gen), the known answers recorded for
the reference tests, scipy.sparse, dense numpy products, closed forms and
invariants. The last tests plant plausible mistakes in oracle outputs and
show that these checks catch them.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest
import scipy.sparse as sp

import _oracle
from _oracle import componentwise_err, fp_equal, load_fixture, manifest

MP_SHAPES = [(100, 100), (1000, 10), (10, 1000)]

# Known answers recorded for the reference's generator as the mp tests call it
# (generate_random_csr({m, n}, 0.1f), float, size_t, seed 0): sum of vals, the
# first three row sums (x = 1, float accumulation), and the sum of the
# width-20 product with b[i] = i (test/gtest/mp/gemv.cpp:27-52).
KNOWN = {
    (100, 100): (508.4906764, [2.29662418, 4.79853439, 3.79471779], 10153711.53),
    (1000, 10): (518.5280002, [0.326641172, 0.196699843, 0.0], 1031494.961),
    (10, 1000): (511.236304, [48.8772545, 46.3883629, 44.3156395], 102268213.3),
}


def rows_of(rowptr):
    rp = np.asarray(rowptr, dtype=np.int64)
    return np.repeat(np.arange(rp.size - 1), np.diff(rp))


def to_scipy(rowptr, colind, vals, shape):
    return sp.csr_matrix((np.asarray(vals, np.float64),
                          np.asarray(colind, np.int64),
                          np.asarray(rowptr, np.int64)), shape=shape)


def random_csr(m, n, density, seed, empty_rows=False, dtype=np.float64):
    rng = np.random.default_rng(seed)
    A = sp.random(m, n, density=density, random_state=rng, format="csr",
                  dtype=np.float64)
    A.data = rng.uniform(-1, 1, A.nnz)
    if empty_rows and m > 3:
        keep = np.ones(m, bool)
        keep[rng.choice(m, size=m // 3, replace=False)] = False
        A = sp.diags(keep.astype(float)) @ A
        A = A.tocsr()
        A.eliminate_zeros()
    A.sort_indices()
    return (A.indptr.astype(np.int64), A.indices.astype(np.int64),
            A.data.astype(dtype), (m, n))


# ---------------------------------------------------------------------------
# Generators vs. the reference's own generator
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape", MP_SHAPES)
def test_random_csr_matches_reference_fixture(orc, shape):
    m, n = shape
    fx = load_fixture("mp_random_%dx%d" % shape)
    rp, ci, v = orc.gen_random_csr(m, n, np.float32(0.1).item(), 0, "u64")
    assert np.array_equal(rp, fx["rowptr"])
    assert np.array_equal(ci, fx["colind"])
    assert np.array_equal(v.view(np.uint32), fx["vals"].view(np.uint32))


def test_sp_tiles_match_reference_fixture(orc):
    cases = {"sp_gemv": (100, 100, 0.1), "sp_notsquare": (10, 1000, 0.1),
             "sp_notsquare_other": (1000, 10, 0.1),
             "sp_verysparse": (100, 100, 0.001)}
    man = manifest()
    checked = 0
    for name, (m, n, dens) in cases.items():
        for P in range(1, 5):
            tm = -(-m // P)
            for i in range((m + tm - 1) // tm):
                rows = min(tm, m - i * tm)
                fx = load_fixture("%s_P%d_t%d" % (name, P, i))
                rp, ci, v = orc.gen_random_csr(rows, n, np.float32(dens).item(), i, "i64")
                assert np.array_equal(rp, fx["rowptr"]) and np.array_equal(ci, fx["colind"])
                assert np.array_equal(v.view(np.uint32), fx["vals"].view(np.uint32))
                assert man["%s_P%d_t%d" % (name, P, i)]["nnz"] == ci.size
                checked += 1
    assert checked == sum(len(range(0, m, -(-m // P))) for (m, n, d) in cases.values()
                          for P in range(1, 5))


@pytest.mark.parametrize("args", [(866, 17, 17), (50, 3, 5)])
def test_band_matches_reference_fixture(orc, args):
    fx = load_fixture("band_%d_%d_%d" % args)
    rp, ci, v = orc.gen_band_csr(args[0], args[1], args[2])
    assert np.array_equal(rp, fx["rowptr"]) and np.array_equal(ci, fx["colind"])
    assert np.array_equal(v, fx["vals"])
    n, up, down = args
    # closed form: row i holds columns [max(0, i-down), min(n-1, i+up)]
    lens = np.minimum(n - 1, np.arange(n) + up) - np.maximum(0, np.arange(n) - down) + 1
    assert np.array_equal(np.diff(rp), lens)
    assert ci.size == 30004 if args == (866, 17, 17) else True


@pytest.mark.parametrize("shape", MP_SHAPES)
def test_fixture_known_answers(shape):
    fx = load_fixture("mp_random_%dx%d" % shape)
    m, n = shape
    vsum, rows3, gsum = KNOWN[shape]
    assert abs(fx["vals"].astype(np.float64).sum() - vsum) <= 1e-9 * vsum
    # float row sums in entry order, as the reference test's c_ref loop does
    c = np.zeros(m, np.float32)
    rp = fx["rowptr"].astype(np.int64)
    for r in range(3):
        for q in range(rp[r], rp[r + 1]):
            c[r] = np.float32(c[r] + fx["vals"][q])
    assert np.allclose(c[:3], rows3, rtol=1e-8, atol=0)
    A = to_scipy(fx["rowptr"], fx["colind"], fx["vals"], shape)
    b = np.arange(n * 20, dtype=np.float64).reshape(20, n)
    assert abs((A @ b.T).sum() - gsum) <= 1e-6 * gsum


def test_fixtures_match_reference_build():
    ref_gen = os.path.join(_oracle.REPO, "oracle", "_ref", "ref_gen")
    if not os.path.exists(ref_gen):
        pytest.skip("oracle/_ref/ref_gen not built here (needs the reference tree)")
    with tempfile.TemporaryDirectory() as td:
        subprocess.check_call([ref_gen, td])
        man = manifest()
        for line in open(os.path.join(td, "manifest.txt")):
            name = line.split()[0]
            fx = load_fixture(name)
            for key in ("vals", "rowptr", "colind"):
                raw = np.fromfile(os.path.join(td, name + "." + key), dtype=fx[key].dtype)
                assert np.array_equal(raw.view(np.uint8), fx[key].view(np.uint8)), (name, key)
            assert man[name]["nnz"] == fx["colind"].size


# ---------------------------------------------------------------------------
# Partitions (csr_row / csr_eq init)
# ---------------------------------------------------------------------------
PART_CASES = [
    np.array([0, 3, 3, 7, 8, 8, 12], np.int64),          # empty rows inside
    np.array([0, 0, 0, 5, 5], np.int64),                 # leading/trailing empty
    np.array([0, 1, 2, 3], np.int64),                    # P > m for large P
    np.array([0, 10], np.int64),                         # one long row
    np.array([0, 0, 0, 0], np.int64),                    # no entries
]


@pytest.mark.parametrize("case", range(len(PART_CASES)))
@pytest.mark.parametrize("P", [1, 2, 3, 4, 7])
def test_row_partition_owns_its_rows_entries(orc, case, P):
    rp = PART_CASES[case]
    m, nnz = rp.size - 1, int(rp[-1])
    seg, off, sz = orc.row_partition(rp, nnz, P)
    assert seg == -(-m // P)
    rows = rows_of(rp)
    for i in range(P):
        mine = np.nonzero((rows >= i * seg) & (rows < (i + 1) * seg))[0]
        if i * seg > m:
            assert off[i] == nnz and sz[i] == 0
        elif mine.size:
            assert off[i] == mine[0] and sz[i] == mine.size
        else:
            assert sz[i] == 0
    assert sz.sum() == nnz


@pytest.mark.parametrize("case", range(len(PART_CASES)))
@pytest.mark.parametrize("P", [1, 2, 3, 4, 7])
def test_eq_partition_rows_cover_entries(orc, case, P):
    rp = PART_CASES[case]
    m, nnz = rp.size - 1, int(rp[-1])
    seg, roff, rsz, tot = orc.eq_partition(rp, nnz, P)
    assert seg == -(-nnz // P)
    assert tot == rsz.sum()
    rows = rows_of(rp)
    for i in range(P):
        lo, hi = i * seg, min((i + 1) * seg, nnz)
        assert rsz[i] >= 1
        if lo < hi:
            assert roff[i] == rows[lo]
            # the share's rows cover its entries; rows past the one holding
            # its last entry can only be empty trailing rows (the reference's
            # last index is (i+1)*seg - 1 even past nnz)
            last = roff[i] + rsz[i] - 1
            assert last >= rows[hi - 1]
            assert all(rp[r] == rp[r + 1] for r in range(rows[hi - 1] + 1, last + 1))


# ---------------------------------------------------------------------------
# The restated mp gemv
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape", MP_SHAPES)
@pytest.mark.parametrize("kind", ["row", "eq"])
@pytest.mark.parametrize("P", [1, 2, 3, 4])
def test_mp_gemv_replays_reference_tests(orc, shape, kind, P):
    """test/gtest/mp/gemv.cpp: x = 1 (row sums) and width-20 b[i] = i."""
    fx = load_fixture("mp_random_%dx%d" % shape)
    m, n = shape
    rp, ci, v = fx["rowptr"].astype(np.int64), fx["colind"].astype(np.int64), fx["vals"]
    c = orc.mp_gemv(kind, rp, ci, v, np.ones(n, np.float32), shape, P)
    rows = rows_of(rp)
    c_ref = np.zeros(m, np.float32)
    np.add.at(c_ref, rows, v)
    assert fp_equal(c, c_ref, 128 * np.finfo(np.float32).eps)
    assert np.allclose(c[:3], KNOWN[shape][1], rtol=2e-7, atol=0)
    b = np.arange(n * 20, dtype=np.float32).reshape(20, n)
    g = orc.mp_gemv(kind, rp, ci, v, b, shape, P)
    dense = to_scipy(rp, ci, v, shape).toarray()
    g64 = (dense @ b.astype(np.float64).T).T
    _, absg = orc.spmv(rp, ci, v, b)
    assert componentwise_err(g, g64, absg) < 1e-5


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_mp_gemv_row_is_rank_invariant_bitwise(orc, seed):
    """Each csr_row rank reduces whole rows in entry order, so every P gives
    bit-identical results, equal to csr_eq at P = 1."""
    rp, ci, v, shape = random_csr(57, 41, 0.2, seed, empty_rows=True)
    x = np.random.default_rng(seed).uniform(-1, 1, (3, shape[1]))
    base = orc.mp_gemv("row", rp, ci, v, x, shape, 1)
    for P in (2, 3, 4, 5, 8, 60):
        assert np.array_equal(orc.mp_gemv("row", rp, ci, v, x, shape, P), base)
    assert np.array_equal(orc.mp_gemv("eq", rp, ci, v, x, shape, 1), base)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
@pytest.mark.parametrize("P", [1, 2, 3, 4, 6])
def test_mp_gemv_matches_scipy_mixed_signs(orc, seed, P):
    rp, ci, v, shape = random_csr(73, 59, 0.15, seed, empty_rows=seed % 2 == 1)
    x = np.random.default_rng(100 + seed).uniform(-1, 1, (2, shape[1]))
    A = to_scipy(rp, ci, v, shape)
    y_sp = (A @ x.T).T
    _, absy = orc.spmv(rp, ci, v, x)
    for kind in ("row", "eq"):
        y = orc.mp_gemv(kind, rp, ci, v, x, shape, P)
        assert componentwise_err(y, y_sp, absy) < 1e-14


def test_mp_gemv_empty_matrix(orc):
    rp = np.zeros(101, np.int64)
    ci = np.zeros(0, np.int64)
    v = np.zeros(0, np.float32)
    for kind in ("row", "eq"):
        for P in (1, 3):
            y = orc.mp_gemv(kind, rp, ci, v, np.ones(100, np.float32), (100, 100), P)
            assert not y.any()


def test_laplacian_closed_form(orc):
    g = 6
    A = laplacian2d(g)
    for kind in ("row", "eq"):
        for P in (1, 2, 5):
            y = orc.mp_gemv(kind, A.indptr, A.indices, A.data, np.ones(g * g), (g * g, g * g), P)
            Y = y.reshape(g, g)
            assert (Y[1:-1, 1:-1] == 0).all()
            assert Y[0, 0] == Y[0, -1] == Y[-1, 0] == Y[-1, -1] == 2
            assert (Y[0, 1:-1] == 1).all() and (Y[1:-1, 0] == 1).all()


def laplacian2d(g):
    T = sp.diags([-1, 2, -1], [-1, 0, 1], shape=(g, g))
    I = sp.identity(g)
    A = (sp.kron(I, T) + sp.kron(T, I)).tocsr()
    A.sort_indices()
    return A


# ---------------------------------------------------------------------------
# Extended-precision oracle
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("seed", range(4))
def test_spmv_ref_dense_brute_force(orc, seed):
    rng = np.random.default_rng(seed)
    m, n = rng.integers(1, 64, 2)
    D = rng.uniform(-1, 1, (m, n)) * (rng.random((m, n)) < 0.3)
    A = sp.csr_matrix(D)
    x = rng.uniform(-1, 1, (3, n))
    y, absy = orc.spmv(A.indptr, A.indices, A.data, x)
    assert np.allclose(y, (D @ x.T).T, rtol=0, atol=1e-14)
    assert np.allclose(absy, (np.abs(D) @ np.abs(x).T).T, rtol=1e-15, atol=1e-15)


def test_spmv_ref_unit_vectors_return_columns(orc):
    rp, ci, v, shape = random_csr(30, 20, 0.3, 7)
    D = to_scipy(rp, ci, v, shape).toarray()
    for j in range(shape[1]):
        e = np.zeros(shape[1])
        e[j] = 1
        y, _ = orc.spmv(rp, ci, v, e)
        assert np.array_equal(y, D[:, j])


def test_spmv_ref_share_semantics(orc):
    """Rows of a nnz share (cut at both ends, no sentinel) give the partial
    sums the csr_eq gather adds back together."""
    rp, ci, v, shape = random_csr(40, 40, 0.3, 3)
    x = np.random.default_rng(3).uniform(-1, 1, shape[1])
    y_full, _ = orc.spmv(rp, ci, v, x)
    nnz = ci.size
    for P in (2, 3, 4):
        seg, roff, rsz, _ = orc.eq_partition(rp, nnz, P)
        acc = np.zeros(shape[0])
        for i in range(P):
            lo, hi = i * seg, min((i + 1) * seg, nnz)
            if lo >= hi:
                continue
            ys, _ = orc.spmv(rp[roff[i]:roff[i] + rsz[i]], ci[lo:hi], v[lo:hi], x,
                             nrows=int(rsz[i]), base=lo, nnz=hi - lo, sentinel=False)
            acc[roff[i]:roff[i] + rsz[i]] += ys
        assert np.allclose(acc, y_full, rtol=0, atol=1e-15)


@pytest.mark.parametrize("seed", range(3))
def test_adjoint_dot_product_identity(orc, seed):
    rp, ci, v, shape = random_csr(50, 35, 0.2, seed, empty_rows=True)
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, shape[1])
    w = rng.uniform(-1, 1, shape[0])
    y, _ = orc.spmv(rp, ci, v, x)
    z, _ = orc.spmv_t(rp, ci, v, w, shape[1])
    assert abs(np.dot(y, w) - np.dot(x, z)) < 1e-13
    A = to_scipy(rp, ci, v, shape)
    assert np.allclose(z, A.T @ w, rtol=0, atol=1e-14)


def test_adjoint_of_symmetric_is_forward(orc):
    A = laplacian2d(7)
    x = np.random.default_rng(0).uniform(-1, 1, 49)
    y, _ = orc.spmv(A.indptr, A.indices, A.data, x)
    z, _ = orc.spmv_t(A.indptr, A.indices, A.data, x, 49)
    assert np.array_equal(y, z)


def test_dvals_closed_form(orc):
    rp, ci, v, shape = random_csr(30, 25, 0.3, 5)
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, (2, shape[1]))
    dy = rng.uniform(-1, 1, (2, shape[0]))
    d = orc.dvals(rp, ci, x, dy)
    rows = rows_of(rp)
    expect = (dy[:, rows] * x[:, ci]).sum(axis=0)
    assert np.allclose(d, expect, rtol=0, atol=1e-15)


def test_transpose_matches_scipy(orc):
    rp, ci, v, shape = random_csr(45, 33, 0.2, 9, empty_rows=True)
    cp, ri, pm = orc.transpose(rp, ci, shape[1])
    At = to_scipy(rp, ci, v, shape).T.tocsr()
    At.sort_indices()
    assert np.array_equal(cp, At.indptr) and np.array_equal(ri, At.indices)
    assert np.array_equal(v[pm], At.data)


# ---------------------------------------------------------------------------
# Threaded GemvEq_DR CPU path == sequential restatement
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("P", [1, 2, 3, 4])
def test_gemveq_cpu_bench_equals_restatement(orc, P):
    fx = load_fixture("band_866_17_17")
    rp, ci, v = fx["rowptr"], fx["colind"], fx["vals"]
    n = 866
    # GemvEq_DR's x: x[i + j*n] = i*j + 1, width 8 (benchmarks/gbench/mp/gemv.cpp:47-55)
    x = (np.arange(n)[None, :] * np.arange(8)[:, None] + 1).astype(np.float64)
    y, med, times = orc.gemveq_cpu_bench(rp, ci, v, x, (n, n), P, warmup=1, reps=2)
    assert np.array_equal(y, orc.mp_gemv("eq", rp, ci, v, x, (n, n), P))
    assert med > 0 and times.size == 2
    # all-ones band with x = i*j + 1: row sums in closed form
    rows = rows_of(rp)
    expect = np.zeros((8, n))
    for j in range(8):
        np.add.at(expect[j], rows, x[j, ci])
    assert np.array_equal(y, expect)


# ---------------------------------------------------------------------------
# Planted mistakes: the checks above must fail on plausible oracle bugs.
# ---------------------------------------------------------------------------
def test_checks_catch_planted_mistakes(orc):
    rp, ci, v, shape = random_csr(40, 31, 0.25, 11)
    x = np.random.default_rng(11).uniform(-1, 1, shape[1])
    A = to_scipy(rp, ci, v, shape)
    y_sp = A @ x
    y, absy = orc.spmv(rp, ci, v, x)
    assert componentwise_err(y, y_sp, absy) < 1e-14
    # dropped term: the last entry of every row lost
    rows = rows_of(rp)
    dropped = y.copy()
    last = rp[1:] - 1
    nz = np.diff(rp) > 0
    dropped[nz] -= v[last[nz]] * x[ci[last[nz]]]
    assert componentwise_err(dropped, y_sp, absy) > 1e-6
    # wrong sign on one row
    flipped = y.copy()
    r = int(np.argmax(np.abs(y)))
    flipped[r] = -flipped[r]
    assert componentwise_err(flipped, y_sp, absy) > 1e-6
    # off-by-one row index (row advance one entry late)
    shifted = np.zeros_like(y)
    late = np.minimum(rows + np.isin(np.arange(ci.size), rp[1:-1]), shape[0] - 1)
    np.add.at(shifted, late, v * x[ci])
    assert componentwise_err(shifted, y_sp, absy) > 1e-6
    # transposed operand: the dot-product identity notices A in place of A^T
    w = np.random.default_rng(12).uniform(-1, 1, shape[0])
    wrong_t = (to_scipy(rp, ci, v, shape)[:, :shape[0]] if shape[1] >= shape[0] else None)
    z, _ = orc.spmv_t(rp, ci, v, w, shape[1])
    assert abs(np.dot(y, w) - np.dot(x, z)) < 1e-13
    zz = np.resize(A @ np.resize(w, shape[1]), shape[1])
    assert abs(np.dot(y, w) - np.dot(x, zz)) > 1e-6
    del wrong_t


# ---------------------------------------------------------------------------
# Dense tile product (sp gemm step): closed form, identities, numpy, linearity.
# ---------------------------------------------------------------------------
def _gemm_example_inputs(m, n, k):
    # examples/sp/gemm_example.cpp:19-29 fills a[i][q] = (i + q) / m and
    # b[q][j] = (q + j) / k.
    i = np.arange(m)[:, None]
    q = np.arange(k)
    a = (i + q[None, :]) / m
    b = (q[:, None] + np.arange(n)[None, :]) / k
    return a, b


def _gemm_example_closed_form(m, n, k):
    # sum_q (i + q)(q + j) = k i j + (i + j) k(k-1)/2 + (k-1)k(2k-1)/6
    i = np.arange(m, dtype=np.float64)[:, None]
    j = np.arange(n, dtype=np.float64)[None, :]
    s1 = k * (k - 1) / 2.0
    s2 = (k - 1) * k * (2 * k - 1) / 6.0
    return (k * i * j + (i + j) * s1 + s2) / (m * k)


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (7, 5, 3), (40, 33, 70)])
def test_gemm_closed_form(orc, m, n, k):
    a, b = _gemm_example_inputs(m, n, k)
    out, absout = orc.gemm(a, b)
    assert componentwise_err(out, _gemm_example_closed_form(m, n, k), absout) < 1e-14


def test_gemm_identity_and_numpy(orc):
    rng = np.random.default_rng(5)
    for dt in (np.float32, np.float64):
        a = rng.standard_normal((23, 17)).astype(dt)
        b = rng.standard_normal((17, 29)).astype(dt)
        assert np.array_equal(orc.gemm(a, np.eye(17, dtype=dt))[0], a.astype(np.float64))
        assert np.array_equal(orc.gemm(np.eye(23, dtype=dt), a)[0], a.astype(np.float64))
        out, absout = orc.gemm(a, b)
        ref = a.astype(np.float64) @ b.astype(np.float64)
        assert componentwise_err(out, ref, absout) < 1e-13


def test_gemm_alpha_beta(orc):
    rng = np.random.default_rng(6)
    a = rng.standard_normal((9, 12))
    b = rng.standard_normal((12, 10))
    c = rng.standard_normal((9, 10))
    base, absbase = orc.gemm(a, b)
    out, _ = orc.gemm(a, b, c, alpha=-0.5, beta=2.0)
    assert componentwise_err(out, -0.5 * base + 2.0 * c, np.abs(base) + np.abs(c) + absbase) < 1e-14
    # beta == 0 does not read c (NaN there stays out of the result)
    nan_c = np.full((9, 10), np.nan)
    out0, _ = orc.gemm(a, b, nan_c, alpha=1.0, beta=0.0)
    assert np.isfinite(out0).all()


def test_gemm_checks_catch_planted_mistakes(orc):
    m, n, k = 12, 9, 15
    a, b = _gemm_example_inputs(m, n, k)
    out, absout = orc.gemm(a, b)
    exact = _gemm_example_closed_form(m, n, k)
    # dropped term: the last k index lost
    dropped = out - np.outer(a[:, -1], b[-1, :])
    assert componentwise_err(dropped, exact, absout) > 1e-6
    # off-by-one row of a
    shifted = np.vstack([out[1:], out[-1:]])
    assert componentwise_err(shifted, exact, absout) > 1e-6
    # transposed operand: random (non-symmetric) inputs, b read as b^T
    rng = np.random.default_rng(7)
    ar = rng.standard_normal((10, 10))
    br = rng.standard_normal((10, 10))
    good, absr = orc.gemm(ar, br)
    wrong = ar @ br.T
    assert componentwise_err(good, ar @ br, absr) < 1e-13
    assert componentwise_err(wrong, ar @ br, absr) > 1e-6
