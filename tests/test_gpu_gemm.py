"""GEMM engine parity (GPU): every transpose combination, batching, split-K and
ragged edges against float64 numpy."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda")


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (17, 33, 5), (64, 64, 64), (130, 70, 257),
                                   (300, 200, 100), (40, 40, 23104 // 8)])
def test_gemm_f64(ta, tb, m, n, k):
    from mpqc_amd.ccsd import gemm
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((n, k) if tb else (k, n))
    C = rng.standard_normal((m, n))
    ref = 0.7 * ((A.T if ta else A) @ (B.T if tb else B)) - 0.3 * C
    dA, dB, dC = _dev(A), _dev(B), _dev(C)
    gemm(ta, tb, m, n, k, 0.7, dA.data_ptr(), A.shape[1], dB.data_ptr(), B.shape[1], -0.3,
         dC.data_ptr(), n)
    torch.cuda.synchronize()
    got = dC.cpu().numpy()
    scale = np.abs(A).max() * np.abs(B).max() * k
    assert np.max(np.abs(got - ref)) <= 1e-14 * scale + 1e-13


def test_gemm_batched_strided():
    from mpqc_amd.ccsd import gemm
    rng = np.random.default_rng(7)
    batch, m, n, k = 37, 21, 19, 45
    A = rng.standard_normal((batch, k, m))      # transposed A
    B = rng.standard_normal((batch, k, n))
    C = np.zeros((batch, m, n))
    ref = np.einsum("bkm,bkn->bmn", A, B)
    dA, dB, dC = _dev(A), _dev(B), _dev(C)
    gemm(1, 0, m, n, k, 1.0, dA.data_ptr(), m, dB.data_ptr(), n, 0.0, dC.data_ptr(), n,
         stride_a=k * m, stride_b=k * n, stride_c=m * n, batch=batch)
    torch.cuda.synchronize()
    assert np.max(np.abs(dC.cpu().numpy() - ref)) < 1e-12


def test_gemm_gemv_splitk():
    from mpqc_amd.ccsd import gemm
    rng = np.random.default_rng(8)
    m, k = 333, 20000
    A = rng.standard_normal((m, k))
    x = rng.standard_normal((k, 1))
    y = rng.standard_normal((m, 1))
    ref = A @ x + y
    dA, dx, dy = _dev(A), _dev(x), _dev(y)
    gemm(0, 0, m, 1, k, 1.0, dA.data_ptr(), k, dx.data_ptr(), 1, 1.0, dy.data_ptr(), 1)
    torch.cuda.synchronize()
    assert np.max(np.abs(dy.cpu().numpy() - ref)) < 1e-10


@pytest.mark.parametrize("beta", [0.0, 1.0, -0.3])
@pytest.mark.parametrize("m,n,k,pad,odd_ldb", [(152, 4800, 40, 0, 0), (1, 2048, 1, 0, 0), (17, 2049, 3, 3, 0),
                                               (16, 3000, 4, 0, 1), (129, 2100, 63, 5, 0), (256, 2048, 64, 0, 0),
                                               (255, 2501, 33, 1, 1), (128, 6000, 10, 0, 0), (70, 2111, 57, 2, 1)])
def test_gemm_thin_k(m, n, k, pad, odd_ldb, beta):
    """Short contractions (K <= 64) of a small A (M <= 256) with a long B: the
    streaming kernel that keeps A in registers and stages B through LDS (ragged
    M, N and K, leading dimensions above the minimum, 16-byte B loads with an
    odd N and 8-byte ones with an odd ldb; beta = 0 must not read C)."""
    from mpqc_amd.ccsd import gemm
    rng = np.random.default_rng(m * 7 + n + k)
    lda, ldc = k + pad, n + pad
    ldb = n + 2 * pad
    if ldb % 2 != odd_ldb:
        ldb += 1
    A = rng.standard_normal((m, lda))
    B = rng.standard_normal((k, ldb))
    C = rng.standard_normal((m, ldc))
    if beta == 0.0:
        C[:, :n] = np.nan
    ref = C.copy()
    ref[:, :n] = 0.7 * (A[:, :k] @ B[:, :n]) + (beta * C[:, :n] if beta != 0.0 else 0.0)
    dA, dB, dC = _dev(A), _dev(B), _dev(C)
    gemm(0, 0, m, n, k, 0.7, dA.data_ptr(), lda, dB.data_ptr(), ldb, beta, dC.data_ptr(), ldc)
    torch.cuda.synchronize()
    got = dC.cpu().numpy()
    assert np.all(np.isfinite(got[:, :n]))
    assert np.array_equal(got[:, n:], C[:, n:])       # padding columns untouched
    assert np.max(np.abs(got - ref)) <= 1e-14 * k * 16 + 1e-13


@pytest.mark.parametrize("beta", [0.0, 1.0])
@pytest.mark.parametrize("m,n,k,pad", [(5000, 40, 152, 0), (4096, 1, 3, 0), (4100, 17, 33, 1),
                                       (10000, 64, 120, 0), (6001, 40, 150, 3), (4500, 48, 16, 2)])
def test_gemm_thin_n(m, n, k, pad, beta):
    """A tall A (M >= 4096, K <= 256) against a small B (N <= 64): the streaming kernel
    with B in LDS (ragged M, N and K, an odd lda takes 8-byte loads, beta = 0 must not
    read C)."""
    from mpqc_amd.ccsd import gemm
    rng = np.random.default_rng(m + 3 * n + k)
    lda, ldb, ldc = k + pad, n + pad, n + 2 * pad
    A = rng.standard_normal((m, lda))
    B = rng.standard_normal((k, ldb))
    C = rng.standard_normal((m, ldc))
    if beta == 0.0:
        C[:, :n] = np.nan
    ref = C.copy()
    ref[:, :n] = -0.9 * (A[:, :k] @ B[:, :n]) + (beta * C[:, :n] if beta != 0.0 else 0.0)
    dA, dB, dC = _dev(A), _dev(B), _dev(C)
    gemm(0, 0, m, n, k, -0.9, dA.data_ptr(), lda, dB.data_ptr(), ldb, beta, dC.data_ptr(), ldc)
    torch.cuda.synchronize()
    got = dC.cpu().numpy()
    assert np.all(np.isfinite(got[:, :n]))
    assert np.array_equal(got[:, n:], C[:, n:])
    assert np.max(np.abs(got - ref)) <= 1e-14 * k * 16 + 1e-13


@pytest.mark.parametrize("ta", [0, 1])
@pytest.mark.parametrize("m,n,k,batch", [(40, 40, 40, 600), (17, 33, 9, 300), (64, 64, 64, 256), (5, 3, 2, 1000)])
def test_gemm_small_batched_shared_a(ta, m, n, k, batch):
    """Many small products against one A (stride 0): one wave per product, A's
    fragments held for the whole launch (ragged M, N, K)."""
    from mpqc_amd.ccsd import gemm
    rng = np.random.default_rng(m * n + k + batch)
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((batch, k, n))
    C = rng.standard_normal((batch, m, n))
    ref = -1.0 * np.einsum("mk,bkn->bmn", A.T if ta else A, B) + 0.5 * C
    dA, dB, dC = _dev(A), _dev(B), _dev(C)
    gemm(ta, 0, m, n, k, -1.0, dA.data_ptr(), A.shape[1], dB.data_ptr(), n, 0.5, dC.data_ptr(), n,
         stride_a=0, stride_b=k * n, stride_c=m * n, batch=batch)
    torch.cuda.synchronize()
    assert np.max(np.abs(dC.cpu().numpy() - ref)) < 1e-12


@pytest.mark.parametrize("m,k,tb,incx,incy,beta", [(6080, 6080, 0, 1, 1, 1.0), (300, 1001, 1, 1, 1, 0.0),
                                                   (257, 4000, 0, 3, 2, -0.5), (64, 256, 0, 1, 1, 0.0)])
def test_gemm_gemv(m, k, tb, incx, incy, beta):
    """N = 1 with long rows: one wave per row (16-byte pieces when x is contiguous,
    strided x and y otherwise, odd K)."""
    from mpqc_amd.ccsd import gemm
    rng = np.random.default_rng(m + k)
    A = rng.standard_normal((m, k))
    x = rng.standard_normal(k * incx)
    y = rng.standard_normal(m * incy)
    if beta == 0.0:
        y[::incy] = np.nan
    ref = y.copy()
    ref[::incy] = 0.8 * (A @ x[::incx]) + (beta * y[::incy] if beta != 0.0 else 0.0)
    dA, dx, dy = _dev(A), _dev(x), _dev(y)
    # op(B) = B: x[k] at B[k * ldb]; op(B) = B^T: x[k] at B[k]
    gemm(0, tb, m, 1, k, 0.8, dA.data_ptr(), k, dx.data_ptr(), k if tb else incx, beta, dy.data_ptr(), incy)
    torch.cuda.synchronize()
    got = dy.cpu().numpy()
    assert np.all(np.isfinite(got[::incy]))
    assert np.max(np.abs(got - ref)) < 1e-10


def test_gemm_thin_k_batched():
    from mpqc_amd.ccsd import gemm
    rng = np.random.default_rng(11)
    batch, m, n, k = 3, 40, 2500, 12
    A = rng.standard_normal((batch, m, k))
    B = rng.standard_normal((batch, k, n))
    C = rng.standard_normal((batch, m, n))
    ref = -1.0 * np.einsum("bmk,bkn->bmn", A, B) + C
    dA, dB, dC = _dev(A), _dev(B), _dev(C)
    gemm(0, 0, m, n, k, -1.0, dA.data_ptr(), k, dB.data_ptr(), n, 1.0, dC.data_ptr(), n,
         stride_a=m * k, stride_b=k * n, stride_c=m * n, batch=batch)
    torch.cuda.synchronize()
    assert np.max(np.abs(dC.cpu().numpy() - ref)) < 1e-12


def test_gemm_rejects_bad_sizes():
    from mpqc_amd.ccsd import gemm
    from mpqc_amd._lib import CCSDError
    with pytest.raises(CCSDError):
        gemm(0, 0, 4, 4, 0, 1.0, 1, 4, 1, 4, 0.0, 1, 4)


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("m,n,k", [(17, 33, 5), (130, 70, 257), (300, 200, 3000), (40, 40, 20000)])
def test_gemm_f32_mfma_error_bound(ta, tb, m, n, k):
    """FP32-MFMA contraction: operands rounded to FP32, FP32 partial sums over
    64-wide K chunks, FP64 across chunks.  Bound: |err| <= 2^-21 sum|a||b|."""
    from mpqc_amd.ccsd import gemm
    rng = np.random.default_rng(m + n + k)
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((n, k) if tb else (k, n))
    Ao, Bo = (A.T if ta else A), (B.T if tb else B)
    ref = Ao @ Bo
    bound = 2.0 ** -21 * (np.abs(Ao) @ np.abs(Bo))
    dA, dB = _dev(A), _dev(B)
    dC = torch.zeros((m, n), dtype=torch.float64, device="cuda")
    gemm(ta, tb, m, n, k, 1.0, dA.data_ptr(), A.shape[1], dB.data_ptr(), B.shape[1], 0.0,
         dC.data_ptr(), n, precision="f32")
    torch.cuda.synchronize()
    err = np.abs(dC.cpu().numpy() - ref)
    assert np.all(err <= bound + 1e-300), float((err / bound).max())
    # and it really ran in reduced precision (FP64 would be ~1e-16 relative)
    assert err.max() > 1e-12 * np.abs(ref).max()


SPLIT_BOUNDS = {"bf16x3": 2.0 ** -14, "f16x3": 2.0 ** -19, "bf16x6": 2.0 ** -19}


@pytest.mark.parametrize("prec", ["bf16x3", "f16x3", "bf16x6"])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("m,n,k", [(17, 33, 5), (130, 70, 257), (300, 200, 3000), (40, 40, 20000),
                                   (512, 768, 3000), (768, 512, 1000)])
def test_gemm_split_error_bound(prec, ta, tb, m, n, k):
    """Split-precision contractions: |C - A B| <= bound(prec) * |A| |B| elementwise
    (plus a tiny absolute term for FP16 subnormal parts), and the result is
    measurably not FP64.  The two last shapes are whole 256 x 256 tiles of the
    two-plane deep-pipeline kernel: two FP32 K chunks added in FP64 by the split-K
    pass, and one chunk written in FP64 by the kernel itself."""
    from mpqc_amd.ccsd import gemm
    rng = np.random.default_rng(m * 7 + n + k)
    A = rng.standard_normal((k, m) if ta else (m, k)) * np.exp(rng.uniform(-3, 3, (k, m) if ta else (m, k)))
    B = rng.standard_normal((n, k) if tb else (k, n)) * 1e-3
    Ao, Bo = (A.T if ta else A), (B.T if tb else B)
    ref = Ao @ Bo
    absprod = np.abs(Ao) @ np.abs(Bo)
    bound = SPLIT_BOUNDS[prec] * absprod + 2.0 ** -30 * np.abs(A).max() * np.abs(B).max() * k
    dA, dB = _dev(A), _dev(B)
    dC = torch.full((m, n), 7.0, dtype=torch.float64, device="cuda")
    gemm(ta, tb, m, n, k, 1.0, dA.data_ptr(), A.shape[1], dB.data_ptr(), B.shape[1], 0.0,
         dC.data_ptr(), n, precision=prec)
    torch.cuda.synchronize()
    err = np.abs(dC.cpu().numpy() - ref)
    assert np.all(err <= bound), float((err / bound).max())
    assert err.max() > 1e-15 * np.abs(ref).max()
