"""GPU parity of the benchmark configurations' exact forms, through the C ABI:

- the 16M x 16M Zipf matrix (Zipf(1.5) row lengths clipped to [1, 4096], FP32)
  through the column-panel plan the benchmark times (default panel width);
- a Zipf matrix of 4M columns with BF16 / FP16 storage through the panel
  kernel at its default width (16-bit values: 1M-column panels);
- batches shaped like the batched configuration (rows U[1k, 200k], 3..64
  entries per row, forward + adjoint items, k = 4 and 16) captured as one
  hipGraph and replayed with new right-hand sides;
- the batch's clamp of a tuned default block size above its kernels' window,
  and panel layouts with empty panels at odd entry offsets.

Tolerances are componentwise, |y - y*| / (|A||x|), against the oracle
(extended precision): 1e-5 for FP32 accumulation (the reference's tests
compare with fp_equal at 128 eps, test/gtest/include/common-tests.hpp:82-95;
componentwise 1e-5 is stricter for long rows). 16-bit storage: 1e-5 against
the oracle on the dequantized values and x (accumulation error), and the
storage rounding bound against the oracle on the FP32 inputs: per row,
(2 u + u^2) |A||x| plus eta ((1 + u)(|v| + |x|) + eta) summed over the row's entries
(u = 2^-8 for BF16, 2^-11 for FP16; eta = half the subnormal spacing, which
bounds the rounding of values below the smallest normal), plus the
accumulation tolerance.
"""
import numpy as np
import pytest
import torch

import drx
from _oracle import componentwise_err

pytestmark = pytest.mark.gpu

DEV = "cuda"
U16 = {torch.bfloat16: 2.0 ** -8, torch.float16: 2.0 ** -11}
# Half the spacing of the subnormals: the absolute rounding error of values
# below the smallest normal (FP16: 6.1e-5; BF16 shares FP32's exponent range).
ETA16 = {torch.bfloat16: 2.0 ** -134, torch.float16: 2.0 ** -25}


def _host(A, vals=None):
    vals = A.vals if vals is None else vals
    return (A.rowptr.long().cpu().numpy(), A.colind.long().cpu().numpy(),
            vals.double().cpu().numpy() if vals.dtype in (torch.bfloat16, torch.float16)
            else vals.cpu().numpy())


def _sampled_rows(A, rows, vals=None):
    """The given rows of A as a small host CSR (rowptr, colind, vals)."""
    rp = A.rowptr.long().cpu().numpy()
    vals = A.vals if vals is None else vals
    srp, sci, sv = [0], [], []
    for r in rows:
        b, e = int(rp[r]), int(rp[r + 1])
        sci.append(A.colind[b:e].long().cpu().numpy())
        sv.append(vals[b:e].cpu().numpy())
        srp.append(srp[-1] + e - b)
    return np.array(srp), np.concatenate(sci), np.concatenate(sv)


@pytest.mark.parametrize("vdt", [torch.bfloat16, torch.float16])
def test_zipf_16bit_storage_column_panels(orc, vdt):
    """Config 5's Zipf leg at reduced size: 1M rows x 4M columns, Zipf(1.5)
    rows clipped to [1, 4096], values (and x) stored in 16 bits with FP32
    accumulation, through the 16-bit column-panel kernel at the default panel
    width (2 MB of x: 1M columns, 4 panels)."""
    m, n = 1_000_000, 4_000_000
    A = drx.gen.zipf_csr(m, n, s=1.5, lmax=4096, seed=21)
    vq = A.vals.to(vdt)
    Aq = drx.CsrMatrix(A.rowptr, A.colind, vq, A.shape)
    npan, width, nseg = Aq.set_panels(0)
    assert width == 1 << 20 and npan == 4 and nseg > m
    x32 = drx.gen.uniform(n, torch.float32, -1, 1, seed=22)
    xq = x32.to(vdt)
    y16 = Aq.matvec(xq)   # 16-bit x
    y32x = Aq.matvec(x32)  # 16-bit values, FP32 x
    torch.cuda.synchronize()
    rp, ci, vq_h = _host(A, vq)
    # accumulation error: oracle on the dequantized inputs
    ys, absy = orc.spmv(rp, ci, vq_h, xq.double().cpu().numpy())
    assert componentwise_err(y16.cpu().numpy(), ys, absy) < 1e-5
    ys2, absy2 = orc.spmv(rp, ci, vq_h, x32.double().cpu().numpy())
    assert componentwise_err(y32x.cpu().numpy(), ys2, absy2) < 1e-5
    # total error against the FP32 inputs: bounded by the storage rounding
    v32 = A.vals.cpu().numpy()
    yo, abso = orc.spmv(rp, ci, v32, x32.cpu().numpy())
    u, eta = U16[vdt], ETA16[vdt]
    xa = np.abs(x32.double().cpu().numpy())[ci]
    va = np.abs(v32.astype(np.float64))
    # storage rounding of both operands, and of the values only (FP32 x)
    sub_both = np.add.reduceat(eta * ((1 + u) * (xa + va) + eta), rp[:-1])
    sub_vals = np.add.reduceat(eta * xa, rp[:-1])
    d16 = np.abs(y16.cpu().numpy().astype(np.float64) - yo)
    d32 = np.abs(y32x.cpu().numpy().astype(np.float64) - yo)
    assert bool((d16 <= (2 * u + u * u + 1e-5) * abso + sub_both).all())
    assert bool((d32 <= (u + 1e-5) * abso + sub_vals).all())
    # the panel plan and the plain plan agree within the accumulation bound
    plain = drx.CsrMatrix(A.rowptr, A.colind, vq, A.shape)
    yp = plain.matvec(xq)
    torch.cuda.synchronize()
    assert componentwise_err(yp.cpu().numpy(), ys, absy) < 1e-5
    plain.close()
    Aq.close()
    A.close()


def test_zipf_config2_headline_panel_plan(orc):
    """The exact plan the benchmark times: 16M x 16M Zipf FP32 with the
    default column panels (31 panels of 524,288 columns). Exact homogeneity
    y(2x) = 2 y(x), 3050 sampled rows (incl. the 50 longest) against the
    oracle, and every row within the FP32 tolerance of the plain plan."""
    n = 16_000_000
    A = drx.gen.zipf_csr(n, n, s=1.5, lmax=4096, seed=0)
    x = drx.gen.uniform(n, torch.float32, -1, 1, seed=1)
    y_plain = A.matvec(x)
    npan, width, nseg = A.set_panels(0)
    assert (npan, width) == (31, 524288) and nseg > n
    y = A.matvec(x)
    y2 = A.matvec(2 * x)
    torch.cuda.synchronize()
    assert torch.equal(y2, 2 * y)
    lens = torch.diff(A.rowptr.long())
    rng = np.random.default_rng(5)
    rows = np.concatenate([rng.choice(n, 3000, replace=False),
                           torch.topk(lens, 50).indices.cpu().numpy()])
    srp, sci, sv = _sampled_rows(A, rows)
    ystar, absy = orc.spmv(srp, sci, sv, x.cpu().numpy())
    assert componentwise_err(y.cpu().numpy()[rows], ystar, absy) < 1e-5
    # every row: |y_panels - y_plain| <= 2e-5 |A||x| (|A||x| from the kernel
    # on |vals| and |x|, exact to rounding)
    absA = drx.CsrMatrix(A.rowptr, A.colind, A.vals.abs(), A.shape)
    ax = absA.matvec(x.abs())
    torch.cuda.synchronize()
    assert bool(((y - y_plain).abs() <= 2e-5 * ax + 1e-30).all())
    absA.close()
    A.close()
    del A, x, y, y2, y_plain, lens, ax
    torch.cuda.empty_cache()


def _config4_mats(nmat, seed):
    rng = np.random.default_rng(seed)
    rows = rng.integers(1000, 200001, nmat)
    return [drx.gen.uniform_rows_csr(int(m), int(m), 3, 64, seed=seed * 1000 + i)
            for i, m in enumerate(rows)]


@pytest.mark.parametrize("k", [4, 16])
def test_batch_config4_graph_replays(orc, k):
    """Forward + adjoint items of config-4-shaped matrices (rows U[1k, 200k],
    3..64 entries per row, FP32) with k right-hand sides, autotuned and
    captured as one hipGraph (x copied into row-major chunks, then the
    persistent kernel); replayed 3 times with new x and dy, eager work and
    synchronizations between replays. Every replay equals an eager run bit for
    bit and the oracle within 1e-5."""
    mats = _config4_mats(8, 40 + k)
    xs = [torch.empty(k, A.shape[1], device=DEV) for A in mats]
    dys = [torch.empty(k, A.shape[0], device=DEV) for A in mats]
    ys = [torch.zeros_like(x) for x in xs]
    dxs = [torch.zeros_like(d) for d in dys]
    for i, (x, d) in enumerate(zip(xs, dys)):
        x.copy_(drx.gen.uniform(x.numel(), torch.float32, -1, 1, seed=i).reshape(x.shape))
        d.copy_(drx.gen.uniform(d.numel(), torch.float32, -1, 1, seed=900 + i).reshape(d.shape))
    b = drx.Batch(forward=list(zip(mats, xs, ys)), adjoint=list(zip(mats, dys, dxs)), k=k)
    b.autotune(iters=1)
    b.graph()
    host = [_host(A) for A in mats]
    for rep in range(3):
        for i, (x, d) in enumerate(zip(xs, dys)):
            x.copy_(drx.gen.uniform(x.numel(), torch.float32, -1, 1,
                                    seed=100 * (rep + 1) + i).reshape(x.shape))
            d.copy_(drx.gen.uniform(d.numel(), torch.float32, -1, 1,
                                    seed=5000 + 100 * rep + i).reshape(d.shape))
        torch.cuda.synchronize()
        b.graph_launch()
        torch.cuda.synchronize()
        got_y = [y.clone() for y in ys]
        got_dx = [dx.clone() for dx in dxs]
        for t in ys + dxs:
            t.fill_(7.0)
        b.run()
        torch.cuda.synchronize()
        for i, (A, (rp, ci, v)) in enumerate(zip(mats, host)):
            assert torch.equal(got_y[i], ys[i]) and torch.equal(got_dx[i], dxs[i])
            yst, absy = orc.spmv(rp, ci, v, xs[i].cpu().numpy())
            assert componentwise_err(got_y[i].cpu().numpy(), yst, absy) < 1e-5, (rep, i)
            dxst, absx = orc.spmv_t(rp, ci, v, dys[i].cpu().numpy(), A.shape[1])
            assert componentwise_err(got_dx[i].cpu().numpy(), dxst, absx) < 1e-5, (rep, i)
    b.close()


def test_batch_clamps_tuned_block_size(orc):
    """A tuned default block size of 4096 entries is clamped to the batch
    kernels' 2048 (windows of 4096): a k = 1 forward + adjoint batch over
    rows of up to 3000 entries still matches the oracle."""
    mats = [drx.gen.uniform_rows_csr(m, m + 100, 1, 3000, seed=70 + m) for m in (700, 1300, 2500)]
    xs = [drx.gen.uniform(A.shape[1], torch.float32, -1, 1, seed=i) for i, A in enumerate(mats)]
    dys = [drx.gen.uniform(A.shape[0], torch.float32, -1, 1, seed=50 + i) for i, A in enumerate(mats)]
    ys = [torch.zeros(A.shape[0], device=DEV) for A in mats]
    dxs = [torch.zeros(A.shape[1], device=DEV) for A in mats]
    with drx.tuned(block_nnz=4096):
        b = drx.Batch(forward=list(zip(mats, xs, ys)), adjoint=list(zip(mats, dys, dxs)))
    assert b.info()["block_nnz"] == 2048
    b.run()
    torch.cuda.synchronize()
    for A, x, y, d, dx in zip(mats, xs, ys, dys, dxs):
        rp, ci, v = _host(A)
        yst, absy = orc.spmv(rp, ci, v, x.cpu().numpy())
        assert componentwise_err(y.cpu().numpy(), yst, absy) < 1e-5
        dxst, absx = orc.spmv_t(rp, ci, v, d.cpu().numpy(), A.shape[1])
        assert componentwise_err(dx.cpu().numpy(), dxst, absx) < 1e-5
    b.close()


@pytest.mark.parametrize("vdt", [torch.float32, torch.bfloat16])
def test_panel_layout_empty_panels_at_odd_offsets(orc, vdt):
    """Columns only in panels 0 and 3 of 10 (64 columns each), an odd number
    of entries in panel 0: the empty middle panels and the empty trailing
    panels start at odd entry offsets. The layout builds, and its products
    match the oracle."""
    rng = np.random.default_rng(3)
    m, w = 900, 64
    p0 = [sorted(rng.choice(w, rng.integers(0, 6), replace=False).tolist()) for _ in range(m)]
    p3 = [sorted((rng.choice(w, rng.integers(0, 4), replace=False) + 3 * w).tolist())
          for _ in range(m)]
    if sum(map(len, p0)) % 2 == 0:  # odd entry count in panel 0
        i = next(i for i in range(m) if p0[i])
        p0[i] = p0[i][1:]
    if (sum(map(len, p0)) + sum(map(len, p3))) % 2 == 0:  # odd total
        i = next(i for i in range(m) if p3[i])
        p3[i] = p3[i][1:]
    rows = [a + b for a, b in zip(p0, p3)]
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64)
    ci = np.array([c for r in rows for c in r], np.int64)
    assert int((ci < w).sum()) % 2 == 1 and ci.size % 2 == 1
    v = rng.uniform(-1, 1, ci.size).astype(np.float32)
    Ad = drx.from_host(rp, ci, v, (m, 10 * w))
    Aq = drx.CsrMatrix(Ad.rowptr, Ad.colind, Ad.vals.to(vdt), Ad.shape)
    npan, width, nseg = Aq.set_panels(w)
    assert (npan, width) == (10, w)
    x = drx.gen.uniform(10 * w, torch.float32, -1, 1, seed=9)
    xq = x.to(vdt)
    y = Aq.matvec(xq)
    torch.cuda.synchronize()
    ys, absy = orc.spmv(rp, ci, Aq.vals.double().cpu().numpy(), xq.double().cpu().numpy())
    assert componentwise_err(y.cpu().numpy(), ys, absy) < 1e-5
    Aq.close()
    Ad.close()
