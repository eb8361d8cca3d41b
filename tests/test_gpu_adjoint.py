"""GPU parity of the adjoint path: transposed structure (bit-exact vs the
oracle's stable counting sort), dx = A^T dy, dvals, the fused residual form
and the torch.autograd.Function (central finite differences in FP64).

Tolerances: componentwise vs the extended-precision oracle, 1e-12 (FP64) and
1e-5 (FP32). dvals are single products (k = 1) or k-term sums: 1e-15 / 1e-6
relative.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import drx
from _oracle import componentwise_err

pytestmark = pytest.mark.gpu
DEV = "cuda"


def rand_csr(m, n, density, seed):
    rng = np.random.default_rng(seed)
    A = sp.random(m, n, density=density, random_state=rng, format="csr")
    A.data = rng.uniform(-1, 1, A.nnz)
    keep = (rng.random(m) > 0.2).astype(float)
    A = (sp.diags(keep) @ A).tocsr()
    A.eliminate_zeros()
    A.sort_indices()
    return A


@pytest.mark.parametrize("itype", [torch.int32, torch.int64])
def test_transpose_structure_bitexact(orc, itype):
    A = rand_csr(3000, 2000, 0.003, 1)
    Ad = drx.from_host(A.indptr, A.indices, A.data, A.shape, itype=itype)
    cp, ri, pm = Ad.transpose().arrays()
    ocp, ori, opm = orc.transpose(A.indptr, A.indices, A.shape[1])
    assert np.array_equal(cp.cpu().numpy(), ocp)
    assert np.array_equal(ri.cpu().numpy(), ori)
    assert np.array_equal(pm.cpu().numpy(), opm)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("k", [1, 4, 11])
def test_adjoint_matches_oracle(orc, dtype, k):
    A = rand_csr(5000, 4000, 0.002, k)
    v = A.data.astype(np.float64 if dtype == torch.float64 else np.float32)
    Ad = drx.from_host(A.indptr, A.indices, v, A.shape)
    dy = np.random.default_rng(k).uniform(-1, 1, (k, A.shape[0])).astype(v.dtype)
    dyd = torch.as_tensor(dy if k > 1 else dy[0]).to(DEV)
    dx = Ad.rmatvec(dyd)
    torch.cuda.synchronize()
    ref, absr = orc.spmv_t(A.indptr, A.indices, v, dy if k > 1 else dy[0], A.shape[1])
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    assert componentwise_err(dx.cpu().numpy(), ref, absr) < tol
    # new vals need no rebuild: the permutation reads them in place
    v2 = (v * 3).astype(v.dtype)
    dx2 = Ad.rmatvec(dyd, vals=torch.as_tensor(v2).to(DEV))
    ref2, absr2 = orc.spmv_t(A.indptr, A.indices, v2, dy if k > 1 else dy[0], A.shape[1])
    assert componentwise_err(dx2.cpu().numpy(), ref2, absr2) < tol


def test_dot_product_identity_gpu():
    """<A x, w> == <x, A^T w> with both sides computed by the GPU kernels."""
    A = drx.gen.band(200000, 16, 16, vdtype=torch.float64, seed=3)
    x = drx.gen.uniform(A.shape[1], torch.float64, -1, 1, seed=4)
    w = drx.gen.uniform(A.shape[0], torch.float64, -1, 1, seed=5)
    lhs = torch.dot(A.matvec(x), w).item()
    rhs = torch.dot(x, A.rmatvec(w)).item()
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_symmetric_adjoint_equals_forward():
    A = drx.gen.laplacian(3, 30, vdtype=torch.float64)
    x = drx.gen.uniform(A.shape[1], torch.float64, -1, 1, seed=8)
    y = A.matvec(x)
    z = A.rmatvec(x)
    assert torch.allclose(y, z, rtol=0, atol=1e-13)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("k", [1, 3])
def test_dvals_matches_oracle(orc, dtype, k):
    A = rand_csr(4000, 3000, 0.003, 10 + k)
    v = A.data.astype(np.float64 if dtype == torch.float64 else np.float32)
    Ad = drx.from_host(A.indptr, A.indices, v, A.shape)
    rng = np.random.default_rng(k)
    x = rng.uniform(-1, 1, (k, A.shape[1])).astype(v.dtype)
    dy = rng.uniform(-1, 1, (k, A.shape[0])).astype(v.dtype)
    sq = (lambda a: a) if k > 1 else (lambda a: a[0])
    d = Ad.dvals(torch.as_tensor(sq(x)).to(DEV), torch.as_tensor(sq(dy)).to(DEV))
    ref = orc.dvals(A.indptr, A.indices, sq(x), sq(dy))
    tol = 1e-15 if dtype == torch.float64 else 1e-6
    scale = np.abs(dy).max() * np.abs(x).max() * k
    assert np.abs(d.cpu().numpy() - ref).max() <= tol * scale


@pytest.mark.parametrize("block_nnz", [64, 1024, 4096])
@pytest.mark.parametrize("k", [1, 2])
def test_dvals_long_short_and_empty_rows(orc, block_nnz, k):
    """Rows longer than the window (whole-workgroup path), runs of empty and
    one-entry rows (more rows than the window's row table: row ids from
    rowptr), alpha and beta."""
    rng = np.random.default_rng(block_nnz + k)
    lens = [0] * 300 + [1] * 500 + [9000, 0, 3, 20000] + list(rng.integers(0, 50, 2000)) + [0] * 100
    n = 30000
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ci = np.concatenate([np.sort(rng.choice(n, L, replace=False)) for L in lens]).astype(np.int64)
    v = rng.uniform(-1, 1, ci.size)
    Ad = drx.from_host(rp, ci, v, (len(lens), n), block_nnz=block_nnz)
    x = rng.uniform(-1, 1, (k, n))
    dy = rng.uniform(-1, 1, (k, len(lens)))
    sq = (lambda a: a) if k > 1 else (lambda a: a[0])
    d0 = rng.uniform(-1, 1, ci.size)
    out = torch.as_tensor(d0).to(DEV)
    Ad.dvals(torch.as_tensor(sq(x)).to(DEV), torch.as_tensor(sq(dy)).to(DEV), out=out,
             alpha=2.0, beta=-0.5)
    ref = 2.0 * orc.dvals(rp, ci, sq(x), sq(dy)) - 0.5 * d0
    assert np.abs(out.cpu().numpy() - ref).max() <= 1e-14 * (2 * k + 1)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_fused_residual_matches_separate(orc, dtype):
    A = drx.gen.band(100000, 16, 16, vdtype=dtype, seed=11)
    x = drx.gen.uniform(A.shape[1], dtype, -1, 1, seed=12)
    t = drx.gen.uniform(A.shape[0], dtype, -1, 1, seed=13)
    y, r, dv = A.fused_residual(x, t)
    torch.cuda.synchronize()
    rp, ci, v = A.rowptr.cpu().numpy(), A.colind.cpu().numpy(), A.vals.cpu().numpy()
    ystar, absy = orc.spmv(rp, ci, v, x.cpu().numpy())
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    assert componentwise_err(y.cpu().numpy(), ystar, absy) < tol
    assert torch.equal(r, y - t)
    ref = orc.dvals(rp, ci, x.cpu().numpy(), r.cpu().numpy())
    assert np.abs(dv.cpu().numpy() - ref).max() <= (1e-15 if dtype == torch.float64 else 1e-6) * np.abs(ref).max()


@pytest.mark.parametrize("vdt", [torch.bfloat16, torch.float16])
def test_fused_residual_16bit_storage(orc, vdt):
    """16-bit vals and x, FP32 sums, y / r / dvals in FP32: against the
    oracle on the dequantized inputs, groups of four entries per lane and
    single entries bit for bit."""
    A = drx.gen.band(50000, 16, 16, vdtype=torch.float32, seed=41)
    A.vals = A.vals.to(vdt)
    x = drx.gen.uniform(A.shape[1], torch.float32, -1, 1, seed=42).to(vdt)
    t = drx.gen.uniform(A.shape[0], torch.float32, -1, 1, seed=43)
    y, r, dv = A.fused_residual(x, t)
    with drx.tuned(vec=1):
        y1, r1, dv1 = A.fused_residual(x, t)
    torch.cuda.synchronize()
    assert torch.equal(y, y1) and torch.equal(r, r1) and torch.equal(dv, dv1)
    rp, ci = A.rowptr.cpu().numpy(), A.colind.cpu().numpy()
    v64, x64 = A.vals.double().cpu().numpy(), x.double().cpu().numpy()
    ystar, absy = orc.spmv(rp, ci, v64, x64)
    assert componentwise_err(y.cpu().numpy(), ystar, absy) < 1e-5
    assert torch.equal(r, y - t)
    ref = orc.dvals(rp, ci, x64, r.double().cpu().numpy())
    assert np.abs(dv.double().cpu().numpy() - ref).max() <= 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("kind", ["band", "wide_band", "random"])
def test_fused_residual_x_staging_bitexact(orc, dtype, kind):
    """xstage = 2 (each block's x window staged in LDS; windows wider than
    the staging buffer read the rest from memory), the default, groups of
    four entries per lane and unaligned vals give the same bits as per-entry
    gathers (xstage = 0)."""
    if kind == "band":
        A = drx.gen.band(60000, 16, 16, vdtype=dtype, seed=31)
    elif kind == "wide_band":
        A = drx.gen.band(60000, 400, 400, vdtype=dtype, seed=32)
    else:
        A = drx.gen.zipf_csr(40000, 50000, s=1.5, lmax=300, seed=33, vdtype=dtype)
    x = drx.gen.uniform(A.shape[1], dtype, -1, 1, seed=34)
    t = drx.gen.uniform(A.shape[0], dtype, -1, 1, seed=35)
    outs = {}
    # (entry-ordered kernels throughout: rowgather 0)
    for xs in (0, 2, 1):
        with drx.tuned(xstage=xs, vec=1, rowgather=0):
            outs[xs] = A.fused_residual(x, t)
    # groups of four entries per lane (FP32, default vec) and single entries
    # when vals start one element into a larger buffer
    with drx.tuned(rowgather=0):
        outs["g4"] = A.fused_residual(x, t)
    buf = torch.zeros(A.nnz + 1, dtype=dtype, device=DEV)
    buf[1:] = A.vals
    v0 = A.vals
    A.vals = buf[1:]
    with drx.tuned(rowgather=0):
        outs["off"] = A.fused_residual(x, t)
    A.vals = v0
    torch.cuda.synchronize()
    for xs in (2, 1, "g4", "off"):
        for a, b in zip(outs[0], outs[xs]):
            assert torch.equal(a, b), (kind, xs)
    rp, ci, v = A.rowptr.cpu().numpy(), A.colind.cpu().numpy(), A.vals.cpu().numpy()
    ystar, absy = orc.spmv(rp, ci, v, x.cpu().numpy())
    assert componentwise_err(outs[2][0].cpu().numpy(), ystar, absy) < (1e-12 if dtype == torch.float64 else 1e-5)


def test_autograd_finite_differences_fp64():
    """Banded 4096-row instance (band +-16, 33 diagonals), loss
    0.5 ||A x - t||^2: the autograd gradients (drx adjoint + dvals kernels)
    against central finite differences in FP64 and against the closed form
    dx = A^T (A x - t), dvals = r[row] x[col]."""
    A = drx.gen.band(4096, 16, 16, vdtype=torch.float64, seed=21)
    assert A.nnz == 134896
    x = drx.gen.uniform(4096, torch.float64, -1, 1, seed=22).requires_grad_(True)
    t = drx.gen.uniform(4096, torch.float64, -1, 1, seed=23)
    vals = A.vals.clone().requires_grad_(True)

    def loss_fn(vv, xx):
        r = drx.spmv(A, xx, vals=vv) - t
        return 0.5 * (r * r).sum()

    L = loss_fn(vals, x)
    L.backward()
    gx, gv = x.grad.detach().clone(), vals.grad.detach().clone()
    # closed form through the oracle-checked kernels
    with torch.no_grad():
        r = A.matvec(x.detach()) - t
        assert torch.allclose(gx, A.rmatvec(r), rtol=0, atol=1e-12)
        assert torch.allclose(gv, A.dvals(x.detach(), r), rtol=0, atol=1e-12)
    # central finite differences on sampled coordinates. The loss is quadratic
    # in every single coordinate, so the central difference has no truncation
    # error; a large step keeps the cancellation error (eps*|L|/h) small.
    rng = np.random.default_rng(0)
    h = 1e-3
    with torch.no_grad():
        for i in rng.choice(4096, 12, replace=False):
            xp, xm = x.detach().clone(), x.detach().clone()
            xp[i] += h
            xm[i] -= h
            fd = (loss_fn(vals.detach(), xp) - loss_fn(vals.detach(), xm)).item() / (2 * h)
            assert abs(fd - gx[i].item()) <= 1e-8 * max(1.0, abs(fd))
        for e in rng.choice(A.nnz, 12, replace=False):
            vp, vm = vals.detach().clone(), vals.detach().clone()
            vp[e] += h
            vm[e] -= h
            fd = (loss_fn(vp, x.detach()) - loss_fn(vm, x.detach())).item() / (2 * h)
            assert abs(fd - gv[e].item()) <= 1e-8 * max(1.0, abs(fd))


def test_autograd_gradcheck_small():
    A = drx.from_host(*(lambda M: (M.indptr, M.indices, M.data))(rand_csr(40, 30, 0.2, 3)),
                      (40, 30))
    x = torch.rand(30, dtype=torch.float64, device=DEV, requires_grad=True)
    vals = A.vals.clone().requires_grad_(True)
    assert torch.autograd.gradcheck(lambda vv, xx: drx.spmv(A, xx, vals=vv), (vals, x),
                                    eps=1e-6, atol=1e-8)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_adjoint_with_stored_transposed_values(orc, dtype):
    """Values stored in CSC order give the same adjoint, bit for bit, as
    reading them through the permutation: by default both sum the short rows
    of the transposed structure in row order; with entry-ordered gathers
    (rowgather 0) on the same 1024-entry blocks (the stored-values FP32
    path otherwise streams 2048-entry blocks, checked against the oracle).
    Changing the values in place (new tensor version) falls back to the
    permutation."""
    A = rand_csr(5000, 4000, 0.002, 21)
    v = A.data.astype(np.float64 if dtype == torch.float64 else np.float32)
    Ad = drx.from_host(A.indptr, A.indices, v, A.shape)
    dy = torch.as_tensor(np.random.default_rng(2).uniform(-1, 1, A.shape[0]).astype(v.dtype)).to(DEV)
    via_perm = Ad.rmatvec(dy)
    with drx.tuned(rowgather=0):
        via_perm0 = Ad.rmatvec(dy)
    Ad.set_transposed_values()
    assert Ad.stored_transposed_values(Ad.vals)
    with drx.tuned(wide=0, rowgather=0):
        stored0 = Ad.rmatvec(dy)
    stored_wide = Ad.rmatvec(dy)
    torch.cuda.synchronize()
    assert torch.equal(stored0, via_perm0)
    assert torch.equal(stored_wide, via_perm)
    ref_p, absr_p = orc.spmv_t(A.indptr, A.indices, v, dy.cpu().numpy(), A.shape[1])
    tol_p = 1e-12 if dtype == torch.float64 else 1e-5
    assert componentwise_err(via_perm.cpu().numpy(), ref_p, absr_p) < tol_p
    ref, absr = orc.spmv_t(A.indptr, A.indices, v, dy.cpu().numpy(), A.shape[1])
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    assert componentwise_err(stored_wide.cpu().numpy(), ref, absr) < tol
    # several right-hand sides: both value sources, bit for bit
    dyk = torch.as_tensor(np.random.default_rng(3).uniform(-1, 1, (9, A.shape[0])).astype(v.dtype)).to(DEV)
    stored_k = Ad.rmatvec(dyk)
    via_perm_k = Ad.rmatvec(dyk, vals=Ad.vals.clone())
    torch.cuda.synchronize()
    assert torch.equal(stored_k, via_perm_k)
    Ad.vals.mul_(2)  # in-place update: the stored copy is stale now
    assert not Ad.stored_transposed_values(Ad.vals)
    dx2 = Ad.rmatvec(dy)
    ref, absr = orc.spmv_t(A.indptr, A.indices, v * 2, dy.cpu().numpy(), A.shape[1])
    assert componentwise_err(dx2.cpu().numpy(), ref, absr) < (1e-12 if dtype == torch.float64 else 1e-5)


@pytest.mark.parametrize("beta", [0.0, -0.75])
def test_dvals_grouped_entries_bitexact(orc, beta):
    """FP32 dvals over 4096-entry windows with groups of four entries per
    lane (16-byte colind loads and dvals stores; vec 0 or 4) give the same
    bits as one entry per lane (vec 1), with windows starting at every
    offset inside a group, long and empty rows, runs of short rows (blocks
    of more rows than lanes, and of more than the LDS holds), and output
    buffers starting one to three elements into a larger allocation (single
    entries); beta reads the old values."""
    rng = np.random.default_rng(51)
    lens = ([0] * 7 + list(rng.integers(16, 70, 4000)) + [9000, 0, 3, 20000, 17, 1] + [5] * 600
            + [1] * 5000 + list(rng.integers(16, 70, 300)))
    n = 40000
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ci = np.concatenate([np.sort(rng.choice(n, L, replace=False)) for L in lens]).astype(np.int64)
    v = rng.uniform(-1, 1, ci.size).astype(np.float32)
    Ad = drx.from_host(rp, ci, v, (len(lens), n))
    x = torch.as_tensor(rng.uniform(-1, 1, n).astype(np.float32)).to(DEV)
    dy = torch.as_tensor(rng.uniform(-1, 1, len(lens)).astype(np.float32)).to(DEV)
    d0 = torch.as_tensor(rng.uniform(-1, 1, ci.size).astype(np.float32)).to(DEV)
    outs = []
    for vec in (1, 4, 0):
        for off in (0, 1, 3):
            buf = torch.zeros(ci.size + off, dtype=torch.float32, device=DEV)
            o = buf[off:]
            o.copy_(d0)
            with drx.tuned(vec=vec):
                Ad.dvals(x, dy, out=o, alpha=1.5, beta=beta)
            outs.append(o)
    torch.cuda.synchronize()
    for o in outs[1:]:
        assert torch.equal(outs[0], o)
    ref = 1.5 * orc.dvals(rp, ci, x.cpu().numpy(), dy.cpu().numpy()) + beta * d0.cpu().numpy()
    assert np.abs(outs[0].cpu().numpy() - ref).max() <= 1e-5
    Ad.close()


@pytest.mark.parametrize("wide", [1, 0])
def test_dvals_fp32_wide_blocks(orc, wide):
    """FP32 dvals over rows of >= 16 entries stream 4096-entry windows (or
    the plan's own with wide = 0), with long and empty rows."""
    rng = np.random.default_rng(40 + wide)
    lens = [0] * 50 + list(rng.integers(10, 60, 3000)) + [9000, 0, 20000]
    n = 30000
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ci = np.concatenate([np.sort(rng.choice(n, L, replace=False)) for L in lens]).astype(np.int64)
    v = rng.uniform(-1, 1, ci.size).astype(np.float32)
    Ad = drx.from_host(rp, ci, v, (len(lens), n))
    x = rng.uniform(-1, 1, n).astype(np.float32)
    dy = rng.uniform(-1, 1, len(lens)).astype(np.float32)
    with drx.tuned(wide=wide):
        d = Ad.dvals(torch.as_tensor(x).to(DEV), torch.as_tensor(dy).to(DEV))
    ref = orc.dvals(rp, ci, x, dy)
    assert np.abs(d.cpu().numpy() - ref).max() <= 1e-6


def _fused_rows_cases():
    """(name, rowptr, colind, vals, shape, eligible): eligible shares are
    the ones the forward product sums in row order by default (longest row
    <= 128, >= 4 per row)."""
    rng = np.random.default_rng(61)
    out = []
    B = drx.gen.band(40000, 16, 16, vdtype=torch.float64, seed=62)
    out.append(("band16", B.rowptr.cpu().numpy(), B.colind.cpu().numpy(), B.vals.cpu().numpy(),
                B.shape, True))
    L = drx.gen.laplacian(3, 20, vdtype=torch.float64)
    out.append(("lap3d", L.rowptr.cpu().numpy(), L.colind.cpu().numpy(), L.vals.cpu().numpy(),
                L.shape, True))
    n = 30000
    for name, lens, auto in (
            ("rows_to_120", list(rng.integers(0, 121, 3000)), True),
            # rows longer than one round, than the window, empty; tiny rows
            ("long_rows", list(rng.integers(0, 200, 2000)) + [9000, 0, 3, 2500, 17], False),
            ("tiny_rows", list(rng.integers(0, 3, 20000)), False)):
        rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        ci = np.concatenate([np.sort(rng.choice(n, L_, replace=False)) for L_ in lens]).astype(np.int64)
        out.append((name, rp, ci, rng.uniform(-1, 1, ci.size), (len(lens), n), auto))
    return out


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32, torch.bfloat16])
@pytest.mark.parametrize("itype", [torch.int32, torch.int64])
def test_fused_residual_row_ordered(orc, dtype, itype):
    """The row-ordered fused kernel (window copied into LDS, lane groups per
    row, r x staged in LDS and stored in entry order): y equals the forward
    product's bits under the same knob (the same row-ordered sums), r = y -
    t exactly, dvals within the oracle's bound. The default picks it exactly
    for the eligible shares, the entry-ordered kernel (rowgather 0) for the
    others; that kernel runs on every share too (nnz not a multiple of four:
    groups of four read indices past the arrays' end)."""
    low = dtype == torch.bfloat16
    sdt = np.float64 if dtype == torch.float64 else np.float32
    for name, rp, ci, v, shape, eligible in _fused_rows_cases():
        m, n = shape
        A = drx.from_host(rp, ci, v.astype(sdt), shape, itype=itype)
        if low:
            A.vals = A.vals.to(dtype)
        xdt = torch.float64 if dtype == torch.float64 else torch.float32
        x = drx.gen.uniform(n, xdt, -1, 1, seed=63).to(dtype)
        t = drx.gen.uniform(m, xdt, -1, 1, seed=64)
        res = {}
        for rg in (3, 1, 0):
            with drx.tuned(rowgather=rg):
                res[rg] = A.fused_residual(x, t)
                if rg == 3:
                    yf = A.matvec(x)
        torch.cuda.synchronize()
        y, r, dv = res[3]
        assert torch.equal(y, yf), name
        assert torch.equal(r, y - t), name
        for a_, b_ in zip(res[1], res[3] if eligible else res[0]):
            assert torch.equal(a_, b_), name
        y0, r0, dv0 = res[0]
        assert torch.equal(r0, y0 - t), name
        vd = A.vals.double().cpu().numpy()
        xd = x.double().cpu().numpy()
        ystar, absy = orc.spmv(rp, ci, vd, xd)
        tol = 1e-12 if dtype == torch.float64 else 1e-5
        for yy, rr, dd in (res[3], res[0]):
            assert componentwise_err(yy.double().cpu().numpy(), ystar, absy) < tol, name
            ref = orc.dvals(rp, ci, xd, rr.double().cpu().numpy())
            bound = (1e-15 if dtype == torch.float64 else 1e-6) * max(np.abs(ref).max(), 1e-300)
            assert np.abs(dd.double().cpu().numpy() - ref).max() <= bound, name
        A.close()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32, torch.bfloat16])
@pytest.mark.parametrize("itype", [torch.int32, torch.int64])
def test_dvals_row_ordered(orc, dtype, itype):
    """Row-ordered dvals (column indices copied into LDS, lane groups per
    row, dy x staged in LDS and stored in entry order, in groups of four
    when dvals is aligned) give the same bits as the entry-ordered kernel
    (the same one product per entry), with alpha and beta, and match the
    oracle; stencils, bands, rows of up to 120 entries, and forced on rows
    longer than the window and on blocks of more rows than the LDS holds."""
    sdt = np.float64 if dtype == torch.float64 else np.float32
    ddt = torch.float64 if dtype == torch.float64 else torch.float32
    for name, rp, ci, v, shape, _ in _fused_rows_cases():
        m, n = shape
        A = drx.from_host(rp, ci, v.astype(sdt), shape, itype=itype)
        x = drx.gen.uniform(n, ddt, -1, 1, seed=71).to(dtype)
        dy = drx.gen.uniform(m, ddt, -1, 1, seed=72).to(dtype)
        d0 = drx.gen.uniform(A.nnz, ddt, -1, 1, seed=73)
        outs = {}
        for rg in (3, 1, 0):
            for off in (0, 1):
                buf = torch.zeros(A.nnz + off, dtype=ddt, device=DEV)
                o = buf[off:]
                o.copy_(d0)
                with drx.tuned(rowgather=rg):
                    A.dvals(x, dy, out=o, alpha=1.5, beta=-0.5)
                outs[(rg, off)] = o
        torch.cuda.synchronize()
        for key, o in outs.items():
            assert torch.equal(o, outs[(0, 0)]), (name, key)
        ref = 1.5 * orc.dvals(rp, ci, x.double().cpu().numpy(), dy.double().cpu().numpy()) \
            - 0.5 * d0.double().cpu().numpy()
        tol = 1e-15 if dtype == torch.float64 else 1e-6
        assert np.abs(outs[(3, 0)].double().cpu().numpy() - ref).max() <= tol * 4, name
        A.close()
