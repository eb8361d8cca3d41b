"""GPU parity of the forward kernel (drx_spmv / drx_spmv_csr) against the CPU
oracle, through the C ABI.

Tolerances (componentwise, scaled by |A||x|, against the extended-precision
oracle): 1e-12 for FP64, 1e-5 for FP32; BF16/FP16 storage with FP32
accumulation: 1e-5 against the oracle on the dequantized inputs (accumulation
error only), total error vs the FP64 inputs reported and bounded by the
storage rounding (2^-8 for BF16, 2^-11 for FP16, per unit of |A||x|).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import drx
from _oracle import componentwise_err, fp_equal, load_fixture

pytestmark = pytest.mark.gpu

TOL = {torch.float64: 1e-12, torch.float32: 1e-5}
DEV = "cuda"


def upload(rp, ci, v, shape, itype=torch.int32, vdtype=None, block_nnz=0):
    return drx.from_host(np.asarray(rp, np.int64), np.asarray(ci, np.int64), v,
                         shape, itype=itype, vdtype=vdtype, block_nnz=block_nnz)


def np_rand_csr(m, n, density, seed, empty_rows=False):
    rng = np.random.default_rng(seed)
    A = sp.random(m, n, density=density, random_state=rng, format="csr")
    A.data = rng.uniform(-1, 1, A.nnz)
    if empty_rows and m > 3:
        keep = (rng.random(m) > 0.3).astype(float)
        A = (sp.diags(keep) @ A).tocsr()
        A.eliminate_zeros()
    A.sort_indices()
    return A


def run(A_dev, x_np, dtype, k_major=False):
    x = torch.as_tensor(np.ascontiguousarray(x_np)).to(dtype).to(DEV)
    y = A_dev.matvec(x)
    torch.cuda.synchronize()
    return y.double().cpu().numpy()


@pytest.mark.parametrize("name", ["mp_random_100x100", "mp_random_1000x10",
                                  "mp_random_10x1000"])
@pytest.mark.parametrize("itype", [torch.int32, torch.int64])
def test_reference_mp_fixtures(orc, name, itype):
    """test/gtest/mp/gemv.cpp: x = 1 (fp_equal vs the reference's c_ref) and
    the width-20 slim matrix b[i] = i."""
    fx = load_fixture(name)
    m, n = [int(t) for t in fx["shape"]]
    A = upload(fx["rowptr"], fx["colind"], fx["vals"], (m, n), itype=itype)
    y = run(A, np.ones(n, np.float32), torch.float32)
    c_ref = orc.mp_gemv("eq", fx["rowptr"], fx["colind"], fx["vals"],
                        np.ones(n, np.float32), (m, n), 1)
    assert fp_equal(y.astype(np.float32), c_ref, 128 * np.finfo(np.float32).eps)
    b = np.arange(n * 20, dtype=np.float32).reshape(20, n)
    g = run(A, b, torch.float32)
    ystar, absy = orc.spmv(fx["rowptr"], fx["colind"], fx["vals"], b)
    assert componentwise_err(g, ystar, absy) < TOL[torch.float32]
    gm = orc.mp_gemv("row", fx["rowptr"], fx["colind"], fx["vals"], b, (m, n), 2)
    assert fp_equal(g.astype(np.float32), gm, 128 * np.finfo(np.float32).eps)


@pytest.mark.parametrize("name,P", [("sp_gemv", 1), ("sp_gemv", 3),
                                    ("sp_notsquare", 2), ("sp_notsquare_other", 4),
                                    ("sp_verysparse", 1), ("sp_verysparse", 4)])
def test_reference_sp_tiles_accumulate(orc, name, P):
    """test/gtest/sp/gemv.cpp: each row tile i is generated with seed i; the
    per-tile call accumulates (c += A_i b, alpha = beta = 1) into c = 0."""
    m = {"sp_gemv": 100, "sp_notsquare": 10, "sp_notsquare_other": 1000,
         "sp_verysparse": 100}[name]
    n = {"sp_gemv": 100, "sp_notsquare": 1000, "sp_notsquare_other": 10,
         "sp_verysparse": 100}[name]
    tm = -(-m // P)
    c = torch.zeros(m, dtype=torch.float32, device=DEV)
    b = torch.ones(n, dtype=torch.float32, device=DEV)
    c_ref = np.zeros(m, np.float32)
    for i in range((m + tm - 1) // tm):
        fx = load_fixture("%s_P%d_t%d" % (name, P, i))
        rows = fx["rowptr"].size - 1
        rp = torch.as_tensor(fx["rowptr"]).to(DEV)
        ci = torch.as_tensor(fx["colind"]).to(DEV)
        va = torch.as_tensor(fx["vals"]).to(DEV)
        seg = c[i * tm:i * tm + rows]
        drx.check(drx.lib().drx_spmv_csr(
            drx.I64, drx.core._p(rp), 0, rows, n, ci.numel(), 1, drx.core._p(ci),
            drx.F32, drx.core._p(va), drx.F32, drx.core._p(b), n, drx.F32,
            drx.core._p(seg), rows, 1, 1.0, 1.0, drx.core._s()))
        c_ref[i * tm:i * tm + rows] = orc.mp_gemv(
            "eq", fx["rowptr"], fx["colind"], fx["vals"], np.ones(n, np.float32),
            (rows, n), 1)
    torch.cuda.synchronize()
    assert fp_equal(c.cpu().numpy(), c_ref, 128 * np.finfo(np.float32).eps)


def test_empty_and_zero_vector():
    """sp EmptyGemv (no entries) and ZeroVector (dense a_ij = i + j, x = 0)."""
    A = upload(np.zeros(101, np.int64), np.zeros(0, np.int64),
               np.zeros(0, np.float32), (100, 100))
    y = torch.full((100,), 7.0, device=DEV)
    A.matvec(torch.ones(100, device=DEV), y=y)
    torch.cuda.synchronize()
    assert not y.any()
    fx = load_fixture("sp_zerovector_dense")
    A = upload(fx["rowptr"], fx["colind"], fx["vals"], (100, 100))
    y = run(A, np.zeros(100, np.float32), torch.float32)
    assert not y.any()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("itype", [torch.int32, torch.int64])
@pytest.mark.parametrize("block_nnz", [64, 256, 1024, 4096])
@pytest.mark.parametrize("k", [1, 3, 16])
def test_random_mixed_sign(orc, dtype, itype, block_nnz, k):
    A = np_rand_csr(3000, 2500, 0.004, block_nnz + k, empty_rows=True)
    Ad = upload(A.indptr, A.indices, A.data.astype(np.float64 if dtype == torch.float64
                                                  else np.float32),
                A.shape, itype=itype, block_nnz=block_nnz)
    x = np.random.default_rng(k).uniform(-1, 1, (k, A.shape[1]))
    xs = x.astype(np.float64 if dtype == torch.float64 else np.float32)
    y = run(Ad, xs if k > 1 else xs[0], dtype)
    ystar, absy = orc.spmv(A.indptr, A.indices, Ad.vals.cpu().numpy(), xs)
    ystar = ystar if k > 1 else ystar[0]
    absy = absy if k > 1 else absy[0]
    assert componentwise_err(y, ystar, absy) < TOL[dtype]


@pytest.mark.parametrize("lens", [[0, 0, 5000, 0, 3], [70000], [1, 9000, 1, 9000, 2],
                                  [3000] * 7, [0] * 50])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_long_and_empty_rows(orc, lens, dtype):
    """Rows longer than the LDS window go to the whole-workgroup path."""
    rng = np.random.default_rng(len(lens))
    n = 80000
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ci = np.concatenate([np.sort(rng.choice(n, L, replace=False)) for L in lens]
                        or [np.zeros(0)]).astype(np.int64)
    v = rng.uniform(-1, 1, ci.size).astype(np.float64 if dtype == torch.float64 else np.float32)
    for bn in (64, 1024):
        Ad = upload(rp, ci, v, (len(lens), n), block_nnz=bn)
        x = rng.uniform(-1, 1, n).astype(v.dtype)
        y = run(Ad, x, dtype)
        ystar, absy = orc.spmv(rp, ci, v, x)
        assert componentwise_err(y, ystar, absy) < TOL[dtype]


def _ragged_with_long_rows(m, n, seed, long_rows=(7, 700), long_len=9000):
    """Rows of 0..60 entries (some empty) plus rows longer than the largest
    LDS window (8192 entries), sorted distinct columns."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 61, m)
    lens[rng.random(m) < 0.2] = 0
    for r in long_rows:
        lens[r] = long_len
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ci = np.concatenate([np.sort(rng.choice(n, L, replace=False)) for L in lens]).astype(np.int64)
    v = rng.uniform(-1, 1, ci.size)
    return rp, ci, v


@pytest.mark.parametrize("mrhs", [2, 0])
@pytest.mark.parametrize("k", [2, 3, 4, 5, 8, 9, 16, 20, 33])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_multi_rhs_paths(orc, mrhs, k, dtype):
    """k right-hand sides through row-major chunks of x (mrhs=2: chunk widths
    2, 4, 8, 16 and several chunks for k > 16) and through direct column
    gathers (mrhs=0), with empty rows and rows longer than every window;
    alpha/beta and padded leading dimensions."""
    rp, ci, v = _ragged_with_long_rows(1500, 10000, k)
    v = v.astype(np.float64 if dtype == torch.float64 else np.float32)
    m, n = 1500, 10000
    rng = np.random.default_rng(100 + k)
    ldx, ldy = n + 3, m + 5
    x = np.zeros((k, ldx), v.dtype)
    x[:, :n] = rng.uniform(-1, 1, (k, n))
    y0 = rng.uniform(-1, 1, (k, ldy)).astype(v.dtype)
    ys, absy = orc.spmv(rp, ci, v, x[:, :n])
    with drx.tuned(mrhs=mrhs):
        for itype, bn in ((torch.int32, 64), (torch.int64, 1024), (torch.int32, 4096)):
            Ad = upload(rp, ci, v, (m, n), itype=itype, block_nnz=bn)
            xd = torch.as_tensor(x).to(DEV)
            y = torch.as_tensor(y0).to(DEV)
            code = drx.F64 if dtype == torch.float64 else drx.F32
            drx.check(drx.lib().drx_spmv(Ad.plan.handle, code, drx.core._p(Ad.colind),
                                         drx.core._p(Ad.vals), code, drx.core._p(xd), ldx,
                                         code, drx.core._p(y), ldy, k, 1.5, -0.25,
                                         drx.core._s()))
            torch.cuda.synchronize()
            got = y.double().cpu().numpy()
            expect = 1.5 * ys - 0.25 * y0[:, :m].astype(np.float64)
            den = 1.5 * absy + 0.25 * np.abs(y0[:, :m])
            assert (np.abs(got[:, :m] - expect) <= TOL[dtype] * den + 1e-300).all(), (itype, bn)
            assert np.array_equal(got[:, m:], y0[:, m:].astype(np.float64))  # rows past m
            Ad.close()


@pytest.mark.parametrize("vdt", [torch.bfloat16, torch.float16])
@pytest.mark.parametrize("xdt_same", [True, False])
def test_low_precision_multi_rhs(orc, vdt, xdt_same):
    """16-bit storage with FP32 accumulation for k = 2, 6 and 17 right-hand
    sides (row-major chunks of 16-bit x: 4-, 16- and 32-byte gathers)."""
    A = drx.gen.laplacian(3, 24, vdtype=torch.float32)
    vq = A.vals.to(vdt)
    rp, ci = A.rowptr.cpu().numpy(), A.colind.cpu().numpy()
    vd = vq.double().cpu().numpy()
    for k in (2, 6, 17):
        x32 = drx.gen.uniform(k * A.shape[1], torch.float32, -1, 1, seed=k).reshape(k, -1)
        xq = x32.to(vdt) if xdt_same else x32
        y = A.matvec(xq, vals=vq)
        torch.cuda.synchronize()
        ystar, absy = orc.spmv(rp, ci, vd, xq.double().cpu().numpy())
        assert componentwise_err(y.cpu().numpy(), ystar, absy) < 1e-5


def test_multi_rhs_on_panel_plan(orc):
    """A plan with column panels serves k >= 2 from the share's own arrays
    (same operand contract: the pointers the layout was built from)."""
    A = np_rand_csr(3000, 5000, 0.004, 77)
    Ad = upload(A.indptr, A.indices, A.data.astype(np.float32), A.shape)
    Ad.set_panels(512)
    x = np.random.default_rng(5).uniform(-1, 1, (7, A.shape[1])).astype(np.float32)
    y = Ad.matvec(torch.as_tensor(x).to(DEV))
    torch.cuda.synchronize()
    ys, absy = orc.spmv(A.indptr, A.indices, Ad.vals.cpu().numpy(), x)
    assert componentwise_err(y.cpu().numpy(), ys, absy) < 1e-5
    with pytest.raises(drx.DrxError):
        Ad.matvec(torch.as_tensor(x).to(DEV), vals=Ad.vals.clone())


def test_alpha_beta_and_ldx(orc):
    A = np_rand_csr(500, 400, 0.02, 1)
    Ad = upload(A.indptr, A.indices, A.data, A.shape)
    k, ldx, ldy = 3, 450, 520
    x = torch.zeros(k, ldx, dtype=torch.float64, device=DEV)
    x[:, :400] = torch.rand(k, 400, dtype=torch.float64, device=DEV) - 0.5
    y0 = torch.rand(k, ldy, dtype=torch.float64, device=DEV)
    y = y0.clone()
    drx.check(drx.lib().drx_spmv(Ad.plan.handle, drx.F64, drx.core._p(Ad.colind),
                                 drx.core._p(Ad.vals), drx.F64, drx.core._p(x), ldx,
                                 drx.F64, drx.core._p(y), ldy, k, 2.0, -0.5,
                                 drx.core._s()))
    torch.cuda.synchronize()
    ys, absy = orc.spmv(A.indptr, A.indices, A.data, x.cpu().numpy(), ldx=ldx)
    expect = 2.0 * ys - 0.5 * y0[:, :500].cpu().numpy()
    got = y[:, :500].cpu().numpy()
    assert np.abs(got - expect).max() <= 1e-12 * (2 * absy + 0.5).max()
    assert torch.equal(y[:, 500:], y0[:, 500:])  # untouched past nrows


@pytest.mark.parametrize("kind", ["row", "eq"])
@pytest.mark.parametrize("P", [1, 2, 3, 4, 7])
def test_share_emulation(orc, kind, P):
    """Each rank's share (absolute rowptr offsets, no sentinel) runs through
    the kernel on the one GPU; the partials are assembled with the
    reference's gather rules (csr_row_distribution.hpp:158-204,
    csr_eq_distribution.hpp:158-216)."""
    A = np_rand_csr(2000, 1500, 0.01, P, empty_rows=True)
    rp = A.indptr.astype(np.int64)
    ci = A.indices.astype(np.int64)
    v = A.data.astype(np.float64)
    m, n = A.shape
    nnz = ci.size
    k = 2
    x = np.random.default_rng(P).uniform(-1, 1, (k, n))
    xd = torch.as_tensor(x).to(DEV)
    rp_d = torch.as_tensor(rp.astype(np.int32)).to(DEV)
    ci_d = torch.as_tensor(ci.astype(np.int32)).to(DEV)
    v_d = torch.as_tensor(v).to(DEV)
    res = np.zeros((k, m))
    if kind == "row":
        seg, off, sz = orc.row_partition(rp, nnz, P)
        for r in range(P):
            if m <= seg * r:
                continue
            rows = min(seg, m - seg * r)
            part = torch.zeros(k, seg, dtype=torch.float64, device=DEV)
            sh = drx.CsrMatrix(rp_d[seg * r:seg * r + rows],
                               ci_d[off[r]:off[r] + sz[r]], v_d[off[r]:off[r] + sz[r]],
                               (rows, n), base=int(off[r]), has_sentinel=False)
            sh.matvec(xd, y=part)
            p = part.cpu().numpy()
            res[:, seg * r:seg * r + rows] = p[:, :rows]
    else:
        seg, roff, rsz, _ = orc.eq_partition(rp, nnz, P)
        res[:] = 0
        for r in range(P):
            if nnz <= seg * r:
                continue
            lo, hi = seg * r, min(seg * (r + 1), nnz)
            part = torch.zeros(k, rsz[r], dtype=torch.float64, device=DEV)
            sh = drx.CsrMatrix(rp_d[roff[r]:roff[r] + rsz[r]], ci_d[lo:hi], v_d[lo:hi],
                               (int(rsz[r]), n), base=lo, has_sentinel=False)
            sh.matvec(xd, y=part)
            p = part.cpu().numpy()
            res[:, roff[r]] += p[:, 0]
            res[:, roff[r] + 1:roff[r] + rsz[r]] = p[:, 1:]
    torch.cuda.synchronize()
    expect = orc.mp_gemv(kind, rp, ci, v, x, (m, n), P)
    _, absy = orc.spmv(rp, ci, v, x)
    assert componentwise_err(res, expect, absy) < 1e-12


def test_laplacian_2d_config1(orc):
    """2-D 5-point Laplacian on 1000 x 1000 (1e6 rows, 4,996,000 entries),
    FP64, x ~ U[-1, 1]: <= 1e-12 vs the oracle; x = 1 closed form."""
    A = drx.gen.laplacian(2, 1000, vdtype=torch.float64)
    assert A.nnz == 4996000
    x = drx.gen.uniform(A.shape[1], torch.float64, -1, 1, seed=20261202)
    y = A.matvec(x)
    torch.cuda.synchronize()
    rp, ci, v = A.rowptr.cpu().numpy(), A.colind.cpu().numpy(), A.vals.cpu().numpy()
    ystar, absy = orc.spmv(rp, ci, v, x.cpu().numpy())
    assert componentwise_err(y.cpu().numpy(), ystar, absy) < 1e-12
    Y = A.matvec(torch.ones(A.shape[1], dtype=torch.float64, device=DEV)).cpu().numpy()
    Y = Y.reshape(1000, 1000)
    assert (Y[1:-1, 1:-1] == 0).all() and Y[0, 0] == 2 and (Y[0, 1:-1] == 1).all()


def test_zipf_small_full_parity(orc):
    A = drx.gen.zipf_csr(200000, 200000, s=1.5, lmax=4096, seed=5)
    x = drx.gen.uniform(A.shape[1], torch.float32, -1, 1, seed=6)
    y = A.matvec(x)
    torch.cuda.synchronize()
    rp, ci, v = A.rowptr.cpu().numpy(), A.colind.cpu().numpy(), A.vals.cpu().numpy()
    ystar, absy = orc.spmv(rp, ci, v, x.cpu().numpy())
    assert componentwise_err(y.cpu().numpy(), ystar, absy) < 1e-5


@pytest.mark.parametrize("vdt", [torch.bfloat16, torch.float16])
@pytest.mark.parametrize("xdt_same", [True, False])
def test_low_precision_storage(orc, vdt, xdt_same):
    A = drx.gen.laplacian(3, 40, vdtype=torch.float32)
    x32 = drx.gen.uniform(A.shape[1], torch.float32, -1, 1, seed=3)
    vq = A.vals.to(vdt)
    xq = x32.to(vdt) if xdt_same else x32
    y = A.matvec(xq, vals=vq)
    torch.cuda.synchronize()
    rp, ci = A.rowptr.cpu().numpy(), A.colind.cpu().numpy()
    vd = vq.double().cpu().numpy()
    xd = xq.double().cpu().numpy()
    ystar, absy = orc.spmv(rp, ci, vd, xd)
    assert componentwise_err(y.cpu().numpy(), ystar, absy) < 1e-5
    y64, abs64 = orc.spmv(rp, ci, A.vals.double().cpu().numpy(), x32.double().cpu().numpy())
    unit = 2.0 ** -8 if vdt == torch.bfloat16 else 2.0 ** -11
    assert componentwise_err(y.cpu().numpy(), y64, abs64) < (2 if xdt_same else 1) * unit + 1e-5


@pytest.mark.parametrize("vdtype", [torch.float64, torch.float32])
def test_share_beyond_2_31_entries(vdtype):
    """A share of 2^24 rows and 2.16e9 > 2^31 entries (band +-64, int64
    indices; values 1): every output row equals its closed form exactly.
    x_i = i (FP64) or i mod 1024 (FP32, every partial sum an integer below
    2^24), so rounding cannot hide an error. The CSC build and the panel
    layout (radix sort, 32-bit segment ids) refuse such a share with a clear
    error instead of running."""
    n, hw = 1 << 24, 64
    A = drx.gen.band(n, hw, hw, vdtype=vdtype, itype=torch.int64, seed=5)
    assert A.nnz > 2 ** 31
    A.vals.fill_(1.0)
    i = torch.arange(n, device=DEV, dtype=torch.int64)
    xi = i if vdtype == torch.float64 else i % 1024
    y = A.matvec(xi.to(vdtype))
    lo = torch.clamp(i - hw, min=0)
    hi = torch.clamp(i + hw, max=n - 1)
    if vdtype == torch.float64:
        exact = (lo + hi) * (hi - lo + 1) // 2
    else:
        cs = torch.cumsum(xi, 0)
        exact = cs[hi] - torch.where(lo > 0, cs[(lo - 1).clamp_min(0)], torch.zeros_like(lo))
    torch.cuda.synchronize()
    assert torch.equal(y.to(torch.int64), exact)
    with pytest.raises(drx.DrxError, match="2\\^31"):
        A.transpose()
    with pytest.raises(drx.DrxError, match="2\\^31"):
        A.set_panels()
    A.close()
    del A, y, exact, i, xi
    torch.cuda.empty_cache()


def test_zipf_config2_full_size_properties(orc):
    """16M x 16M, Zipf(1.5) rows clipped to [1, 4096], FP32: 3050 sampled
    rows (incl. the 50 longest) against the oracle, and exact
    homogeneity y(2x) = 2 y(x) (scaling by 2 is exact in FP32)."""
    n = 16_000_000
    A = drx.gen.zipf_csr(n, n, s=1.5, lmax=4096, seed=0)
    lens = torch.diff(A.rowptr.long())
    assert int(lens.min()) >= 1 and int(lens.max()) <= 4096
    mean = A.nnz / n
    assert 90 < mean < 105, mean
    x = drx.gen.uniform(n, torch.float32, -1, 1, seed=1)
    y = A.matvec(x)
    y2 = A.matvec(2 * x)
    torch.cuda.synchronize()
    assert torch.equal(y2, 2 * y)
    rng = np.random.default_rng(0)
    rows = np.concatenate([rng.choice(n, 3000, replace=False),
                           torch.topk(lens, 50).indices.cpu().numpy()])
    rp = A.rowptr.long().cpu().numpy()
    # the sampled rows as a small CSR on the host, checked by the oracle
    srp = [0]
    sci, sv = [], []
    for r in rows:
        b, e = int(rp[r]), int(rp[r + 1])
        sci.append(A.colind[b:e].cpu().numpy())
        sv.append(A.vals[b:e].cpu().numpy())
        srp.append(srp[-1] + e - b)
    ystar, absy = orc.spmv(np.array(srp), np.concatenate(sci), np.concatenate(sv),
                           x.cpu().numpy())
    got = y.cpu().numpy()[rows]
    assert componentwise_err(got, ystar, absy) < 1e-5
    A.close()


@pytest.mark.parametrize("panel_cols", [64, 333, 1000])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("k", [1, 3])
def test_panel_layout(orc, panel_cols, dtype, k):
    """Column-panel layout: same products as plain CSR within tolerance, for
    several panel widths (empty panels and rows included), k right-hand
    sides, alpha/beta, and new values re-packed through the permutation."""
    A = np_rand_csr(4000, 3000, 0.003, panel_cols + k, empty_rows=True)
    v = A.data.astype(np.float64 if dtype == torch.float64 else np.float32)
    Ad = upload(A.indptr, A.indices, v, A.shape, block_nnz=256)
    npan, w, nseg = Ad.set_panels(panel_cols)
    assert w == panel_cols and npan == -(-3000 // panel_cols) and nseg > 0
    x = np.random.default_rng(k).uniform(-1, 1, (k, A.shape[1])).astype(v.dtype)
    xd = torch.as_tensor(x if k > 1 else x[0]).to(DEV)
    y0 = torch.rand((k, 4000) if k > 1 else (4000,), dtype=dtype, device=DEV)
    y = y0.clone()
    Ad.matvec(xd, y=y, alpha=2.0, beta=0.5)
    torch.cuda.synchronize()
    ys, absy = orc.spmv(A.indptr, A.indices, v, x if k > 1 else x[0])
    expect = 2.0 * ys + 0.5 * y0.double().cpu().numpy()
    tol = TOL[dtype]
    den = 2.0 * absy + 0.5 * np.abs(y0.double().cpu().numpy())
    assert (np.abs(y.double().cpu().numpy() - expect) <= tol * den + 1e-300).all()
    # beta = 0 overwrites, new values through drx_plan_update_values
    v2 = torch.as_tensor(v * -3).to(DEV)
    Ad.update_values(v2)
    y2 = Ad.matvec(xd)
    torch.cuda.synchronize()
    ys2, absy2 = orc.spmv(A.indptr, A.indices, v * -3, x if k > 1 else x[0])
    assert componentwise_err(y2.double().cpu().numpy(), ys2, absy2) < tol
    # the layout refuses other vals pointers (it holds a copy)
    with pytest.raises(drx.DrxError):
        Ad.matvec(xd, vals=torch.as_tensor(v).to(DEV))


def test_panel_layout_zipf(orc):
    A = drx.gen.zipf_csr(300000, 300000, s=1.5, lmax=4096, seed=9)
    A.set_panels(32768)
    x = drx.gen.uniform(A.shape[1], torch.float32, -1, 1, seed=10)
    y = A.matvec(x)
    torch.cuda.synchronize()
    rp, ci, v = A.rowptr.cpu().numpy(), A.colind.cpu().numpy(), A.vals.cpu().numpy()
    ystar, absy = orc.spmv(rp, ci, v, x.cpu().numpy())
    assert componentwise_err(y.cpu().numpy(), ystar, absy) < 1e-5


def _batch_mats(nmat, seed, rows=(500, 4000), lens=(3, 64)):
    rng = np.random.default_rng(seed)
    mats = []
    for i in range(nmat):
        m = int(rng.integers(*rows))
        mats.append(drx.gen.uniform_rows_csr(m, m, lens[0], lens[1], seed=seed * 1000 + i,
                                             block_nnz=1024))
    return mats


@pytest.mark.parametrize("k", [1, 4, 17])
def test_batch_matches_single_calls_and_oracle(orc, k):
    """Forward + adjoint items of uneven matrices in one persistent launch:
    bit-identical to the single calls at the same block size (1024 entries:
    the single FP32 calls' 2048-entry blocks turned off) and lane-group
    thresholds, and within tolerance of the oracle."""
    mats = _batch_mats(9, 3 + k)
    def vec(n, s):
        return drx.gen.uniform(k * n, torch.float32, -1, 1, seed=s).reshape(k, n) if k > 1 \
            else drx.gen.uniform(n, torch.float32, -1, 1, seed=s)
    xs = [vec(A.shape[1], 100 + i) for i, A in enumerate(mats)]
    dys = [vec(A.shape[0], 200 + i) for i, A in enumerate(mats)]
    ys = [torch.zeros_like(x) for x in xs]
    dxs = [torch.zeros_like(d) for d in dys]
    b = drx.Batch(forward=[(A, x, y) for A, x, y in zip(mats, xs, ys)],
                  adjoint=[(A, d, dx) for A, d, dx in zip(mats, dys, dxs)], k=k)
    # same lane-group thresholds as the single calls (gshift 1 for k >= 2)
    b.set_tuning(1024, drx.tune_get("mrhs_gshift") if k > 1 else 0, 4)
    b.run()
    torch.cuda.synchronize()
    for A, x, y, d, dx in zip(mats, xs, ys, dys, dxs):
        # single calls on the batch's 1024-entry blocks and entry-ordered gathers
        with drx.tuned(wide=0, rowgather=0):
            assert torch.equal(y, A.matvec(x))
            assert torch.equal(dx, A.rmatvec(d))
        rp, ci, v = A.rowptr.cpu().numpy(), A.colind.cpu().numpy(), A.vals.cpu().numpy()
        ys_, absy = orc.spmv(rp, ci, v, x.cpu().numpy())
        assert componentwise_err(y.cpu().numpy(), ys_, absy) < 1e-5
        dx_, absx = orc.spmv_t(rp, ci, v, d.cpu().numpy(), A.shape[1])
        assert componentwise_err(dx.cpu().numpy(), dx_, absx) < 1e-5
    b.close()


def test_batch_graph_replays_and_autotune(orc, tmp_path):
    """The captured graph is replayed several times with eager work and
    synchronizations between replays (new x each time); every replay equals
    the eager launch. The autotuned batch still matches the single calls."""
    mats = _batch_mats(6, 11)
    xs = [drx.gen.uniform(A.shape[1], torch.float32, -1, 1, seed=i) for i, A in enumerate(mats)]
    ys = [torch.zeros(A.shape[0], device=DEV) for A in mats]
    b = drx.Batch(forward=[(A, x, y) for A, x, y in zip(mats, xs, ys)])
    t = b.autotune(iters=2, cache=str(tmp_path / "tune.json"), key="t6")
    assert t["block_nnz"] in (256, 512, 1024, 2048) and t["wg_per_cu"] >= 1
    t2 = b.autotune(cache=str(tmp_path / "tune.json"), key="t6")
    assert t2.get("cached") and t2["block_nnz"] == t["block_nnz"]
    b.graph()
    for rep in range(4):
        for i, x in enumerate(xs):
            x.copy_(drx.gen.uniform(x.numel(), torch.float32, -1, 1, seed=50 + rep * 10 + i))
        torch.cuda.synchronize()
        b.graph_launch()
        torch.cuda.synchronize()
        got = [y.clone() for y in ys]
        for y in ys:
            y.zero_()
        b.run()
        torch.cuda.synchronize()
        for g, y, A, x in zip(got, ys, mats, xs):
            assert torch.equal(g, y)
            rp, ci, v = A.rowptr.cpu().numpy(), A.colind.cpu().numpy(), A.vals.cpu().numpy()
            ys_, absy = orc.spmv(rp, ci, v, x.cpu().numpy())
            assert componentwise_err(g.cpu().numpy(), ys_, absy) < 1e-5
    b.close()


def test_mmread_then_gemv(orc, tmp_path):
    """A Matrix Market file read by the library's reader and multiplied on
    the GPU equals the oracle on the scipy-read matrix."""
    import scipy.io
    rng = np.random.default_rng(31)
    A = sp.random(2000, 1500, density=0.01, random_state=rng, format="coo")
    A.data = rng.standard_normal(A.nnz)
    path = str(tmp_path / "m.mtx")
    scipy.io.mmwrite(path, A, field="real", symmetry="general")
    Ad = drx.mmread(path, dtype=torch.float64)
    ref = scipy.io.mmread(path).tocsr()
    ref.sort_indices()
    x = rng.uniform(-1, 1, 1500)
    y = run(Ad, x, torch.float64)
    ys, absy = orc.spmv(ref.indptr, ref.indices, ref.data, x)
    assert componentwise_err(y, ys, absy) < 1e-12


def _odd_window_rows(nb):
    """Row lengths that give, at block_nnz nb (window 2 nb), a block whose
    window is exactly full and starts at an odd entry ([nb + 1, 3 nb + 1):
    rows of nb - 2 and nb + 2 entries starting in block 1, the next row
    starting in block 3), next to long, empty and odd-length rows."""
    return [nb + 1, nb - 2, nb + 2, 0, 3, 1, 300, 7, 0, 129, 5, 64, 63, 1, 2, 9000, 11]


@pytest.mark.parametrize("vdt", [torch.float32, torch.bfloat16, torch.float16, torch.float64])
@pytest.mark.parametrize("block_nnz", [64, 512, 1024])
def test_paired_entry_loads_bitexact(orc, vdt, block_nnz):
    """vec = 2 and 4 (groups of two or four consecutive entries per lane and
    load, from the group-aligned entry below each block) give the same bits
    as one entry per lane: windows starting at every offset within a group,
    full windows, vals starting one to three elements into a larger buffer
    (not group-aligned: single entries), column panels. FP64 values take
    pairs (four would exceed 16 bytes per load: single entries)."""
    rng = np.random.default_rng(block_nnz)
    lens = _odd_window_rows(block_nnz) + list(rng.integers(0, 40, 3000))
    n = 20000
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ci = np.concatenate([np.sort(rng.choice(n, L, replace=False)) for L in lens]).astype(np.int64)
    v = rng.uniform(-1, 1, ci.size)
    Ad = upload(rp, ci, v.astype(np.float64 if vdt == torch.float64 else np.float32),
                (len(lens), n), block_nnz=block_nnz)
    vq = Ad.vals.to(vdt)
    xdt = torch.float64 if vdt == torch.float64 else torch.float32
    x = drx.gen.uniform(n, xdt, -1, 1, seed=3)
    xq = x.to(vdt) if vdt not in (torch.float32, torch.float64) else x
    # the same blocks for every load layout: the window is fixed (16-bit
    # values otherwise take 4096 with groups of four and 2048 without)
    caps = (2048, 4096) if vdt != torch.float64 else (2048,)
    ys_all = []
    for cap in caps:
        with drx.tuned(vec=1, fwd_cap=cap):
            y1 = Ad.matvec(xq, vals=vq)
        for vec in (2, 4):
            with drx.tuned(vec=vec, fwd_cap=cap):
                ys_all.append((y1, Ad.matvec(xq, vals=vq)))
                for off in (1, 2, 3):  # vals not aligned to a group: single entries
                    buf = torch.zeros(vq.numel() + off, dtype=vdt, device=DEV)
                    buf[off:] = vq
                    ys_all.append((y1, Ad.matvec(xq, vals=buf[off:])))
    torch.cuda.synchronize()
    for ya, yv in ys_all:
        assert torch.equal(ya, yv)
    ys, absy = orc.spmv(rp, ci, vq.double().cpu().numpy(), xq.double().cpu().numpy())
    tol = 1e-12 if vdt == torch.float64 else 1e-5
    for ya, _ in ys_all[::4]:
        assert componentwise_err(ya.cpu().numpy(), ys, absy) < tol
    if vdt == torch.float32:
        Ad.set_panels(4096)
        pv = []
        for vec in (1, 2, 4):
            with drx.tuned(vec=vec):
                pv.append(Ad.matvec(x))
        torch.cuda.synchronize()
        assert torch.equal(pv[0], pv[1]) and torch.equal(pv[0], pv[2])
        assert componentwise_err(pv[2].cpu().numpy(), ys, absy) < 1e-5
    Ad.close()


def _staging_cases(rng):
    """(name, rowptr, colind, vals, shape): a band (every block's columns fit
    its LDS window), a 2-D 5-point stencil (windows a little wider than the
    FP64 LDS capacity: some entries read x from memory), random columns
    (spans far wider), empty rows, rows longer than a window, one column."""
    out = []
    n = 30000
    A = sp.diags([rng.uniform(-1, 1, n - abs(o)) for o in range(-16, 17)], list(range(-16, 17)),
                 format="csr")
    out.append(("band", A))
    g = 300
    T = sp.diags([4.0 * np.ones(g), -np.ones(g - 1), -np.ones(g - 1)], [0, -1, 1])
    I = sp.identity(g)
    L = (sp.kron(I, T) + sp.kron(sp.diags([-np.ones(g - 1), -np.ones(g - 1)], [-1, 1]), I)).tocsr()
    out.append(("lap2d", L))
    out.append(("random", np_rand_csr(5000, 40000, 0.002, 7, empty_rows=True)))
    lens = [0, 5000, 3, 0, 9000, 1, 2, 700] + list(rng.integers(0, 30, 2000))
    m = len(lens)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ci = np.concatenate([np.sort(rng.choice(12000, L, replace=False)) for L in lens]).astype(np.int64)
    out.append(("long_rows", sp.csr_matrix((rng.uniform(-1, 1, ci.size), ci, rp), shape=(m, 12000))))
    C = sp.csr_matrix((rng.uniform(-1, 1, 4000), np.zeros(4000, np.int64), np.arange(4001)),
                      shape=(4000, 1))
    out.append(("one_column", C))
    return out


@pytest.mark.parametrize("vdt", [torch.float64, torch.float32])
@pytest.mark.parametrize("itype", [torch.int32, torch.int64])
def test_x_staging_bitexact(orc, vdt, itype):
    """xstage = 2 (every block stages its x window in LDS) gives the same
    bits as per-entry gathers (xstage = 0) and the default (1); all within
    the oracle tolerance."""
    rng = np.random.default_rng(17)
    for name, A in _staging_cases(rng):
        A = A.tocsr()
        A.sort_indices()
        rp, ci = A.indptr.astype(np.int64), A.indices.astype(np.int64)
        v = A.data.astype(np.float32 if vdt == torch.float32 else np.float64)
        Ad = upload(rp, ci, v, A.shape, itype=itype)
        x = drx.gen.uniform(A.shape[1], vdt, -1, 1, seed=11)
        ys = {}
        for xs in (0, 2, 1):
            with drx.tuned(xstage=xs, rowgather=0):
                ys[xs] = Ad.matvec(x)
        torch.cuda.synchronize()
        assert torch.equal(ys[0], ys[2]), name
        assert torch.equal(ys[0], ys[1]), name
        yo, absy = orc.spmv(rp, ci, v, x.cpu().numpy())
        assert componentwise_err(ys[2].cpu().numpy(), yo, absy) < TOL[vdt], name
        Ad.close()


@pytest.mark.parametrize("cap", [8, 64, 1000])
def test_grid_cap_loops_over_blocks(orc, cap):
    """A capped grid (workgroups loop over the row blocks, keeping their XCD)
    gives the same bits as one workgroup per block."""
    A = np_rand_csr(20000, 15000, 0.002, cap)
    Ad = upload(A.indptr, A.indices, A.data.astype(np.float32), A.shape, block_nnz=256)
    x = drx.gen.uniform(15000, torch.float32, -1, 1, seed=4)
    y0 = Ad.matvec(x)
    with drx.tuned(grid_cap=cap):
        y1 = Ad.matvec(x)
    torch.cuda.synchronize()
    assert torch.equal(y0, y1)


def test_gather_probe():
    """drx_gather (the benchmark's gather-rate probe) equals torch indexing."""
    x = torch.randn(70001, device=DEV)
    idx = torch.randint(0, 70001, (123457,), dtype=torch.int32, device=DEV)
    out = drx.gen.gather(idx, x)
    assert torch.equal(out, x[idx.long()])
    xd = x.double()
    assert torch.equal(drx.gen.gather(idx, xd), xd[idx.long()])
    xh = x.to(torch.bfloat16)
    assert torch.equal(drx.gen.gather(idx, xh), xh[idx.long()])


@pytest.mark.parametrize("k", [2, 4, 5])
def test_multi_rhs_path_choice_by_column_locality(k):
    """The default (mrhs=1) takes column gathers for k <= 4 over short rows
    with local columns (a 2-D Laplacian) and row-major chunks otherwise (a
    +-16 band, a random matrix, k = 5): its bits equal the forced path's."""
    lap = drx.gen.laplacian(2, 150, vdtype=torch.float64)
    band = drx.gen.band(20000, 16, 16, vdtype=torch.float64, seed=5)
    rnd = drx.gen.uniform_rows_csr(20000, 20000, 3, 12, seed=6, vdtype=torch.float64)
    for A, cols in ((lap, k <= 4), (band, False), (rnd, False)):
        n = A.shape[1]
        x = drx.gen.uniform(k * n, torch.float64, -1, 1, seed=k).reshape(k, -1)
        y_auto = A.matvec(x)
        with drx.tuned(mrhs=0):
            y_cols = A.matvec(x)
        with drx.tuned(mrhs=2):
            y_chunks = A.matvec(x)
        torch.cuda.synchronize()
        assert torch.equal(y_auto, y_cols if cols else y_chunks), (A.shape, k)


def test_multi_rhs_chunk_width_by_locality(orc):
    """k > 8 FP32: 8-wide chunks over loosely local columns (a 3-D stencil),
    16-wide over random columns; the default's bits equal the forced width's,
    and both are within tolerance of the oracle."""
    lap = drx.gen.laplacian(3, 24, vdtype=torch.float32)
    rnd = drx.gen.uniform_rows_csr(14000, 14000, 3, 30, seed=8)
    for A, kp in ((lap, 8), (rnd, 16)):
        k = 12
        x = drx.gen.uniform(k * A.shape[1], torch.float32, -1, 1, seed=9).reshape(k, -1)
        y_auto = A.matvec(x)
        with drx.tuned(mrhs_kp_max=kp):
            y_forced = A.matvec(x)
        torch.cuda.synchronize()
        assert torch.equal(y_auto, y_forced), kp
        rp, ci, v = A.rowptr.cpu().numpy(), A.colind.cpu().numpy(), A.vals.cpu().numpy()
        ys, absy = orc.spmv(rp, ci, v, x.cpu().numpy())
        assert componentwise_err(y_auto.cpu().numpy(), ys, absy) < 1e-5


@pytest.mark.parametrize("wide", [1, 0])
@pytest.mark.parametrize("lens_kind", ["ragged", "long"])
def test_fp32_wide_blocks(orc, wide, lens_kind):
    """FP32 with 32-bit indices streams the plan's 2048-entry blocks (window
    4096, 16 entries per lane) or its own 1024-entry blocks: both within
    tolerance of the oracle, with long rows (whole-workgroup path), empty
    rows, alpha and beta."""
    if lens_kind == "ragged":
        rp, ci, v = _ragged_with_long_rows(6000, 30000, 11, long_rows=(), long_len=0)
        m, n = 6000, 30000
    else:
        rp, ci, v = _ragged_with_long_rows(3000, 30000, 12, long_rows=(5, 2000), long_len=9000)
        m, n = 3000, 30000
    v = v.astype(np.float32)
    Ad = upload(rp, ci, v, (m, n))
    x = np.random.default_rng(3).uniform(-1, 1, n).astype(np.float32)
    y0 = np.random.default_rng(4).uniform(-1, 1, m).astype(np.float32)
    y = torch.as_tensor(y0).to(DEV)
    with drx.tuned(wide=wide):
        Ad.matvec(torch.as_tensor(x).to(DEV), y=y, alpha=1.5, beta=-0.5)
    torch.cuda.synchronize()
    ys, absy = orc.spmv(rp, ci, v, x)
    expect = 1.5 * ys - 0.5 * y0.astype(np.float64)
    den = 1.5 * absy + 0.5 * np.abs(y0)
    assert (np.abs(y.double().cpu().numpy() - expect) <= 1e-5 * den + 1e-300).all()


def _short_rows(m, n, seed, lmax, local, empty=0.15, extra=()):
    """Rows of 1..lmax entries (a share `empty` of them empty), sorted
    distinct columns: within 64 columns around the row's diagonal (local) or
    anywhere; `extra` appends rows of the given lengths."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(1, lmax + 1, m)
    lens[rng.random(m) < empty] = 0
    lens = list(lens) + list(extra)
    cols = []
    for r, L in enumerate(lens):
        if local and L <= 64:
            lo = min(max(0, r * n // len(lens) - 32), n - 64)
            cols.append(np.sort(rng.choice(64, L, replace=False)) + lo)
        else:
            cols.append(np.sort(rng.choice(n, L, replace=False)))
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    return rp, np.concatenate(cols).astype(np.int64), rng.uniform(-1, 1, int(rp[-1])), len(lens)


def _row_gather_cases():
    """(name, rowptr, colind, vals, shape, auto): `auto` says whether the
    default picks row-ordered gathers (longest row <= 128, >= 4 per row)."""
    out = []
    L = drx.gen.laplacian(3, 22, vdtype=torch.float64)
    out.append(("lap3d", L.rowptr.cpu().numpy(), L.colind.cpu().numpy(), L.vals.cpu().numpy(),
                L.shape, True))
    rp, ci, v, m = _short_rows(20000, 30000, 1, 16, True, empty=0.1)
    out.append(("short_local", rp, ci, v, (m, 30000), True))
    rp, ci, v, m = _short_rows(20000, 50000, 2, 12, False, empty=0.05)
    out.append(("short_random", rp, ci, v, (m, 50000), True))
    # rows of 17..64 entries: lane groups of 8 and 16 per row
    B = drx.gen.band(30000, 16, 16, vdtype=torch.float64, seed=5)
    out.append(("band16", B.rowptr.cpu().numpy(), B.colind.cpu().numpy(), B.vals.cpu().numpy(),
                B.shape, True))
    rp, ci, v, m = _short_rows(8000, 40000, 6, 64, False, empty=0.02)
    out.append(("random_to_64", rp, ci, v, (m, 40000), True))
    # rows longer than eight entries (several rounds per lane), longer than
    # the window (the whole-workgroup path) and empty: forced only
    rp, ci, v, m = _short_rows(6000, 30000, 3, 40, True, extra=(9000, 3, 0, 2500, 17))
    out.append(("long_rows", rp, ci, v, (m, 30000), False))
    # rows of 0..2 entries: more rows per block than the LDS holds (row
    # offsets read from rowptr)
    rp, ci, v, m = _short_rows(30000, 30000, 4, 2, True, empty=0.3)
    out.append(("tiny_rows", rp, ci, v, (m, 30000), False))
    return out


@pytest.mark.parametrize("vdt", [torch.float64, torch.float32, torch.bfloat16, torch.float16])
@pytest.mark.parametrize("itype", [torch.int32, torch.int64])
def test_row_ordered_gathers(orc, vdt, itype):
    """Row-ordered gathers (knob rowgather: values and columns of each block
    staged in LDS, one lane per row) against the oracle, with alpha and beta:
    stencils, short local and random rows, rows longer than eight entries
    and than the window, empty rows, blocks with more rows than the LDS
    holds; lane groups of 8 and 16 per row for blocks of longer rows. The
    window copied straight into LDS (3) and through registers
    (2: groups of four entries per lane, or single entries when vals start
    one element into a larger buffer) give the same bits. The default (1)
    picks them exactly for the eligible shares: the straight copy for
    aligned vals, the register path otherwise, entry-ordered gathers (0) for
    the rest."""
    low = vdt in (torch.bfloat16, torch.float16)
    sdt = np.float64 if vdt == torch.float64 else np.float32
    tol = 1e-12 if vdt == torch.float64 else 1e-5
    for name, rp, ci, v, shape, auto in _row_gather_cases():
        m, n = shape
        Ad = upload(rp, ci, v.astype(sdt), shape, itype=itype)
        vq = Ad.vals.to(vdt)
        xdt = torch.float64 if vdt == torch.float64 else torch.float32
        xq = drx.gen.uniform(n, xdt, -1, 1, seed=21).to(vdt)
        ydt = torch.float32 if low else vdt
        y0 = drx.gen.uniform(m, ydt, -1, 1, seed=22)
        outs = {}
        for rg, vals_off in ((3, 0), (2, 0), (2, 1), (1, 0), (1, 1), (0, 0), (0, 1)):
            vv = vq
            if vals_off:
                buf = torch.zeros(vq.numel() + vals_off, dtype=vdt, device=DEV)
                buf[vals_off:] = vq
                vv = buf[vals_off:]
            y = y0.clone()
            with drx.tuned(rowgather=rg):
                Ad.matvec(xq, y=y, vals=vv, alpha=1.5, beta=-0.5)
            outs[(rg, vals_off)] = y
        torch.cuda.synchronize()
        assert torch.equal(outs[(3, 0)], outs[(2, 0)]), name
        assert torch.equal(outs[(1, 0)], outs[(3, 0)] if auto else outs[(0, 0)]), name
        assert torch.equal(outs[(1, 1)], outs[(2, 1)] if auto else outs[(0, 1)]), name
        vd = vq.double().cpu().numpy()
        xd = xq.double().cpu().numpy()
        ys, absy = orc.spmv(rp, ci, vd, xd)
        y0d = y0.double().cpu().numpy()
        expect = 1.5 * ys - 0.5 * y0d
        den = 1.5 * absy + 0.5 * np.abs(y0d)
        for key, y in outs.items():
            got = y.double().cpu().numpy()
            assert (np.abs(got - expect) <= tol * den + 1e-300).all(), (name, key)
        Ad.close()


@pytest.mark.parametrize("vdt", [torch.float64, torch.float32, torch.bfloat16])
def test_row_ordered_gathers_1024_windows(orc, vdt):
    """Row-ordered gathers on a plan of 512-entry blocks (1024-entry windows,
    `k_spmv_rows<256, 1024>`), default and forced, against the oracle."""
    sdt = np.float64 if vdt == torch.float64 else np.float32
    tol = 1e-12 if vdt == torch.float64 else 1e-5
    for name, rp, ci, v, shape, _ in _row_gather_cases()[:4]:
        m, n = shape
        Ad = upload(rp, ci, v.astype(sdt), shape, block_nnz=512)
        vq = Ad.vals.to(vdt)
        xdt = torch.float64 if vdt == torch.float64 else torch.float32
        xq = drx.gen.uniform(n, xdt, -1, 1, seed=31).to(vdt)
        ys = {}
        for rg in (1, 3):
            with drx.tuned(rowgather=rg):
                ys[rg] = Ad.matvec(xq, vals=vq)
        torch.cuda.synchronize()
        ystar, absy = orc.spmv(rp, ci, vq.double().cpu().numpy(), xq.double().cpu().numpy())
        for rg, y in ys.items():
            assert componentwise_err(y.double().cpu().numpy(), ystar, absy) < tol, (name, rg)
        Ad.close()
