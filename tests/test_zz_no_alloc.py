"""The allocation-free GEMM entry (mpqc_ccsd_gemm_noalloc) against the
allocating one (mpqc_ccsd_gemm) through the C-ABI: two calls in a row, the
second on a smaller problem, in FP64 and in a split-precision form.

At these sizes both entries plan the same split-K factor (the allocating entry's
workspace and the fixed scratch are both 64 MiB), so the kernels run the same
sums in the same order and the outputs must be bitwise equal; both are also
checked against numpy's FP64 product (1e-13 relative to |A||B| for FP64, the
split form's bound of 2^-14 relative to |A||B|).  The second call, whose partial sums are
fewer than the first call's, would read the first call's leftovers if a partial
were accumulated instead of written.  Device memory in use does not change
across the allocation-free calls.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(entry, prec, ta, tb, dA, dB, dC, m, n, k, alpha, beta):
    import torch
    from mpqc_amd import _lib as L
    from mpqc_amd.ccsd import PRECISIONS
    lib = L.load()
    fn = getattr(lib, entry)
    st = fn(ta, tb, m, n, k, alpha, ctypes.c_void_p(dA.data_ptr()), dA.shape[1], 0,
            ctypes.c_void_p(dB.data_ptr()), dB.shape[1], 0, beta, ctypes.c_void_p(dC.data_ptr()), n, 0, 1,
            PRECISIONS[prec], None)
    L.check(st)
    torch.cuda.synchronize()


def _case(seed, ta, tb, m, n, k):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((n, k) if tb else (k, n))
    C0 = rng.standard_normal((m, n))
    return A, B, C0


@pytest.mark.parametrize("prec", ["f64", "bf16x3"])
def test_noalloc_gemm_matches_allocating_entry_twice(prec):
    import torch
    cases = [(11, 0, 1, 96, 80, 6000), (12, 1, 0, 40, 24, 1500)]   # second call: smaller, fewer partials
    for i, (seed, ta, tb, m, n, k) in enumerate(cases):
        A, B, C0 = _case(seed, ta, tb, m, n, k)
        alpha, beta = 0.75, -0.5
        dA, dB = torch.from_numpy(A).to("cuda"), torch.from_numpy(B).to("cuda")
        dC_alloc, dC = torch.from_numpy(C0).to("cuda"), torch.from_numpy(C0).to("cuda")
        _call("mpqc_ccsd_gemm", prec, ta, tb, dA, dB, dC_alloc, m, n, k, alpha, beta)
        free_before = torch.cuda.mem_get_info()[0]
        _call("mpqc_ccsd_gemm_noalloc", prec, ta, tb, dA, dB, dC, m, n, k, alpha, beta)
        if i == 1:   # (the first call may load the library's code object holding the scratch)
            assert torch.cuda.mem_get_info()[0] == free_before
        out, ref_alloc = dC.cpu().numpy(), dC_alloc.cpu().numpy()
        Ao, Bo = (A.T if ta else A), (B.T if tb else B)
        exact = alpha * (Ao @ Bo) + beta * C0
        if prec == "f64":
            tol = 1e-13 * (np.abs(Ao) @ np.abs(Bo)) + 1e-15
        else:
            tol = 2.0 ** -14 * (np.abs(Ao) @ np.abs(Bo)) + 1e-15
        err = np.abs(out - exact)
        print(f"{prec} call {i + 1} ({m}x{n}x{k}): max |noalloc - alloc| "
              f"{np.abs(out - ref_alloc).max():.1e}, max err / bound {(err / tol).max():.2e}")
        assert np.array_equal(out, ref_alloc)
        assert np.all(err <= tol)


def test_noalloc_gemm_rejects_bad_arguments():
    from mpqc_amd import _lib as L
    lib = L.load()
    st = lib.mpqc_ccsd_gemm_noalloc(0, 0, 4, 4, 0, 1.0, None, 4, 0, None, 4, 0, 0.0, None, 4, 0, 1, 0, None)
    assert st == L.ERR_ARG
    assert b"invalid GEMM sizes" in lib.mpqc_ccsd_last_error(None)
