"""GPU parity of the dense tile product (drx_gemm, the step of dr::sp::gemm)
against the CPU oracle, through the C ABI.

Tolerance: componentwise, |c - c*| <= (k + 2) u (|alpha| |A||B| + |beta| |C|)
with u = 2^-24 (FP32) or 2^-53 (FP64) -- the deterministic rounding bound of
a k-term dot product in any summation order plus the alpha/beta update.
Small-integer inputs, whose every partial sum is exact in the working type,
must match the oracle bit for bit; they pin the MFMA operand and result
layouts independently of rounding.
"""
import numpy as np
import pytest
import torch

import drx
from _oracle import componentwise_err

pytestmark = pytest.mark.gpu

DEV = "cuda"
U = {torch.float32: 2.0 ** -24, torch.float64: 2.0 ** -53}
NP = {torch.float32: np.float32, torch.float64: np.float64}

SHAPES = [(1, 1, 1), (16, 16, 4), (128, 128, 16), (256, 128, 16), (130, 70, 37),
          (300, 257, 129), (513, 129, 200), (64, 700, 9)]


def _dev(x, dtype):
    return torch.as_tensor(np.ascontiguousarray(x)).to(dtype).to(DEV)


def _bound(k, dtype):
    return (k + 2) * U[dtype]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("tile", [1, 2, 3, 4, 5, 6, 7])
@pytest.mark.parametrize("m,n,k", SHAPES)
def test_gemm_integer_inputs_exact(orc, dtype, tile, m, n, k):
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    a = rng.integers(-3, 4, (m, k)).astype(NP[dtype])
    b = rng.integers(-3, 4, (k, n)).astype(NP[dtype])
    c = rng.integers(-5, 6, (m, n)).astype(NP[dtype])
    ref, _ = orc.gemm(a, b, c, alpha=1.0, beta=1.0)
    cd = _dev(c, dtype)
    with drx.tuned(gemm_tile=tile):
        drx.gemm(_dev(a, dtype), _dev(b, dtype), cd)
    torch.cuda.synchronize()
    assert np.array_equal(cd.double().cpu().numpy(), ref)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("alpha,beta", [(1.0, 1.0), (1.0, 0.0), (-0.5, 2.0)])
@pytest.mark.parametrize("m,n,k", [(130, 70, 37), (300, 257, 129), (257, 384, 1000)])
def test_gemm_random_within_bound(orc, dtype, alpha, beta, m, n, k):
    rng = np.random.default_rng(k)
    a = rng.standard_normal((m, k)).astype(NP[dtype])
    b = rng.standard_normal((k, n)).astype(NP[dtype])
    c = rng.standard_normal((m, n)).astype(NP[dtype])
    ref, absref = orc.gemm(a, b, c, alpha=alpha, beta=beta)
    cd = _dev(c, dtype)
    drx.gemm(_dev(a, dtype), _dev(b, dtype), cd, alpha=alpha, beta=beta)
    torch.cuda.synchronize()
    assert componentwise_err(cd.double().cpu().numpy(), ref, absref) <= _bound(k, dtype)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("offset", [0, 1])
def test_gemm_leading_dimensions_and_unaligned_views(orc, dtype, offset):
    """Views into wider row-major buffers (lda > k, ldb > n, ldc > n), as
    dense_matrix_view tiles are; offset 1 makes the rows misaligned for
    16-byte loads."""
    m, n, k = 200, 150, 90
    rng = np.random.default_rng(3 + offset)
    A = rng.standard_normal((m, k + 13)).astype(NP[dtype])
    B = rng.standard_normal((k, n + 7)).astype(NP[dtype])
    C = rng.standard_normal((m, n + 5)).astype(NP[dtype])
    a, b, c = A[:, offset:offset + k], B[:, offset:offset + n], C[:, offset:offset + n]
    ref, absref = orc.gemm(a, b, c)
    Ad, Bd, Cd = _dev(A, dtype), _dev(B, dtype), _dev(C, dtype)
    cview = Cd[:, offset:offset + n]
    drx.gemm(Ad[:, offset:offset + k], Bd[:, offset:offset + n], cview)
    torch.cuda.synchronize()
    out = Cd.double().cpu().numpy()
    assert componentwise_err(out[:, offset:offset + n], ref, absref) <= _bound(k, dtype)
    # columns outside the view are untouched
    keep = np.ones(n + 5, bool)
    keep[offset:offset + n] = False
    assert np.array_equal(out[:, keep], C[:, keep].astype(np.float64))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_gemm_beta_zero_does_not_read_c(orc, dtype):
    rng = np.random.default_rng(9)
    a = rng.standard_normal((70, 40)).astype(NP[dtype])
    b = rng.standard_normal((40, 90)).astype(NP[dtype])
    ref, absref = orc.gemm(a, b)
    cd = torch.full((70, 90), float("nan"), dtype=dtype, device=DEV)
    drx.gemm(_dev(a, dtype), _dev(b, dtype), cd, beta=0.0)
    torch.cuda.synchronize()
    assert componentwise_err(cd.double().cpu().numpy(), ref, absref) <= _bound(40, dtype)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_gemm_k_zero_scales_c(dtype):
    c = torch.arange(12, dtype=dtype, device=DEV).reshape(3, 4)
    want = 3.0 * c
    a = torch.empty(3, 0, dtype=dtype, device=DEV)
    b = torch.empty(0, 4, dtype=dtype, device=DEV)
    drx.gemm(a, b, c, alpha=1.0, beta=3.0)
    torch.cuda.synchronize()
    assert torch.equal(c, want)


def test_gemm_python_shim_rejects_mixed_dtypes():
    a = torch.zeros(4, 5, device=DEV)
    b = torch.zeros(5, 6, device=DEV, dtype=torch.float64)
    with pytest.raises(ValueError):
        drx.gemm(a, b)
    with pytest.raises(ValueError):
        drx.gemm(a, a)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_gemm_reference_example_inputs(dtype):
    """examples/sp/gemm_example.cpp inputs a[i][q] = (i + q) / m,
    b[q][j] = (q + j) / k against their closed form."""
    m, n, k = 384, 320, 500
    i = torch.arange(m, dtype=torch.float64)[:, None]
    q = torch.arange(k, dtype=torch.float64)
    j = torch.arange(n, dtype=torch.float64)[None, :]
    a = ((i + q[None, :]) / m).to(dtype).to(DEV)
    b = ((q[:, None] + j) / k).to(dtype).to(DEV)
    c = drx.gemm(a, b)
    torch.cuda.synchronize()
    s1, s2 = k * (k - 1) / 2.0, (k - 1) * k * (2 * k - 1) / 6.0
    exact = (k * i * j + (i + j) * s1 + s2) / (m * k)
    # inputs are rounded to dtype first: bound the product of rounded inputs
    # by (k + 2) u plus the input rounding, 2u per product term
    err = (c.double().cpu() - exact).abs() / exact.clamp_min(1e-300)
    assert float(err.max()) <= (k + 4) * U[dtype]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_gemm_large_against_fp64_torch(dtype):
    """4096^3 (FP32 on 256 x 256 tiles, FP64 on 128 x 128 tiles loaded
    straight into LDS: the defaults there) against torch's float64 product
    of the same inputs, within the componentwise bound (|A||B| from float64
    too)."""
    n = 4096
    g = torch.Generator(device=DEV).manual_seed(1)
    a = torch.randn(n, n, device=DEV, dtype=dtype, generator=g)
    b = torch.randn(n, n, device=DEV, dtype=dtype, generator=g)
    c = drx.gemm(a, b)
    ref = a.double() @ b.double()
    absref = a.double().abs() @ b.double().abs()
    torch.cuda.synchronize()
    err = float(((c.double() - ref).abs() / absref).max())
    assert err <= _bound(n, dtype)
