"""CPU pinning of the oracle's complex GEMM (cpu_gemm restatement) and of the
power epilogue's float formula."""
import numpy as np
import pytest


@pytest.mark.parametrize("M,N,K,B", [(1, 1, 1, 1), (5, 4, 7, 2), (33, 17, 64, 1)])
def test_gemm_vs_numpy(oracle, M, N, K, B):
    rng = np.random.default_rng(M * 100 + N)
    A = rng.standard_normal((B, M, K)) + 1j * rng.standard_normal((B, M, K))
    Bm = rng.standard_normal((B, N, K)) + 1j * rng.standard_normal((B, N, K))
    C = oracle.gemm_c64(A, Bm)
    ref = A @ np.transpose(Bm, (0, 2, 1))  # B^T, no conjugate
    assert np.allclose(C, ref, rtol=1e-12, atol=1e-12)
    # planted mistake: a conjugate on B is a different product
    if K > 1:
        assert not np.allclose(C, A @ np.conj(np.transpose(Bm, (0, 2, 1))))


def test_gemm_closed_forms(oracle):
    rng = np.random.default_rng(3)
    A = (rng.integers(-7, 8, (2, 6, 5)) + 1j * rng.integers(-7, 8, (2, 6, 5))).astype(np.complex128)
    eye = np.broadcast_to(np.eye(5, dtype=np.complex128), (2, 5, 5)).copy()
    assert np.array_equal(oracle.gemm_c64(A, eye), A)            # B = I
    assert np.array_equal(oracle.gemm_c64(A, 1j * eye), 1j * A)  # B = iI
    # integer data: exact (the reference test's data, test_gemm_api.cu:93-110)
    Bm = (rng.integers(-7, 8, (2, 4, 5)) + 1j * rng.integers(-7, 8, (2, 4, 5))).astype(np.complex128)
    C = oracle.gemm_c64(A, Bm)
    exact = np.einsum("bik,bjk->bij", A.real.astype(np.int64), Bm.real.astype(np.int64)) - \
        np.einsum("bik,bjk->bij", A.imag.astype(np.int64), Bm.imag.astype(np.int64))
    assert np.array_equal(C.real.astype(np.int64), exact)


def test_power_formula(oracle):
    rng = np.random.default_rng(4)
    re = rng.integers(-2896, 2897, 1000).astype(np.float32)
    im = rng.integers(-2896, 2897, 1000).astype(np.float32)
    # re^2 + im^2 < 2^24: every step exact, so it equals the FP64 value
    p = oracle.power_f32(re, im)
    assert np.array_equal(p.astype(np.float64), re.astype(np.float64) ** 2 + im.astype(np.float64) ** 2)
    # large values: fma(re, re, fl(im * im)) then * alpha, each rounded once in float32
    re = (rng.standard_normal(1000) * 3e5).astype(np.float32)
    im = (rng.standard_normal(1000) * 3e5).astype(np.float32)
    old = rng.integers(-1000, 1000, 1000).astype(np.float32)  # beta * old + p exact in FP64
    got = oracle.power_f32(re, im, alpha=0.5, beta=2.0, old=old)
    imim = (im.astype(np.float32) * im.astype(np.float32)).astype(np.float64)
    fma1 = (re.astype(np.float64) ** 2 + imim).astype(np.float32)   # exact in FP64, one rounding
    p = (np.float32(0.5) * fma1).astype(np.float32)
    want = (np.float64(2.0) * old.astype(np.float64) + p.astype(np.float64)).astype(np.float32)
    assert np.array_equal(got, want)
