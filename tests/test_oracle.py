"""CPU tests that pin the oracle (oracle/herk_oracle.c) before it is trusted.

Each check is independent of the oracle's own code: known answers from the
C++ standard and the reference's LUT, torch's float8_e4m3fn and numpy's float16
decoders, numpy complex128 products, closed forms and invariants, and the
committed golden fixtures.
"""
import hashlib
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_mt19937_64_known_answer(oracle):
    # C++11 [rand.predef]: the 10000th invocation of a default-constructed
    # mt19937_64 (seed 5489) produces 9981545732273789042.
    assert oracle.mt64_nth(5489, 10000) == 9981545732273789042


def test_qc_fp8_table_matches_reference_lut(oracle):
    # Nibble table of the reference LUT (cutlass_interface/cutlass_gemm_api.cu:1443-1451)
    ref_nib = [0x00, 0x38, 0x40, 0x44, 0x48, 0x4A, 0x4C, 0x4E,
               0x80, 0xB8, 0xC0, 0xC4, 0xC8, 0xCA, 0xCC, 0xCE]
    for b in range(256):
        re_b, im_b = oracle.qc_to_fp8_pair(b)
        assert re_b == ref_nib[b >> 4] and im_b == ref_nib[b & 15], hex(b)


def test_fp8_bytes_decode_to_the_integer(oracle):
    import torch
    bytes_ = np.array([oracle.qc_to_fp8_pair(b)[k] for b in range(256) for k in (0, 1)],
                      dtype=np.uint8)
    vals = torch.from_numpy(bytes_).view(torch.float8_e4m3fn).float().numpy()
    expect = []
    for b in range(256):
        r, i = oracle.decode_qc(b)
        expect += [r, i]
    assert np.array_equal(vals, np.array(expect, dtype=np.float32))
    # the oracle's own E4M3 decoder agrees with torch's on all 256 codes
    allb = np.arange(256, dtype=np.uint8)
    tv = torch.from_numpy(allb).view(torch.float8_e4m3fn).float().numpy()
    ov = np.array([oracle.fp8e4m3_to_float(int(b)) for b in allb], dtype=np.float32)
    nan = np.isnan(tv)
    assert np.array_equal(np.isnan(ov), nan)
    assert np.array_equal(ov[~nan], tv[~nan])


def test_qc_codec(oracle):
    # sign-magnitude, high nibble = Re (test_herk_int4.cu:39-45)
    assert oracle.decode_qc(0x9A) == (-1, -2)
    assert oracle.decode_qc(0x7F) == (7, -7)
    assert oracle.decode_qc(0x80) == (0, 0)     # -0 decodes to 0
    assert oracle.decode_qc(0x08) == (0, 0)
    assert oracle.encode_qc(-3, 5) == 0xB5
    assert oracle.encode_qc(0, 0) == 0x00       # +0 encoding for zero
    for r in range(-7, 8):
        for i in range(-7, 8):
            assert oracle.decode_qc(oracle.encode_qc(r, i)) == (r, i)


def test_generator_distribution_and_determinism(oracle):
    a = oracle.gen_qc_uniform(3, 200000)
    assert np.array_equal(a, oracle.gen_qc_uniform(3, 200000))
    assert not np.array_equal(a, oracle.gen_qc_uniform(4, 200000))
    nib = np.array([0, 1, 2, 3, 4, 5, 6, 7, 0, -1, -2, -3, -4, -5, -6, -7])
    re, im = nib[a >> 4], nib[a & 15]
    for v in (re, im):
        counts = np.bincount(v + 7, minlength=15)
        assert counts.size == 15 and counts.min() > 0
        expected = v.size / 15
        chi2 = ((counts - expected) ** 2 / expected).sum()
        assert chi2 < 60  # 14 dof; p ~ 1e-7
    # uniform generator never emits the -0 nibble; the raw-byte generator does
    assert not np.any((a & 15) == 8) and not np.any((a >> 4) == 8)
    rb = oracle.gen_bytes(42, 4096)
    assert np.any((rb & 15) == 8)


def _decode(A):
    nib = np.array([0, 1, 2, 3, 4, 5, 6, 7, 0, -1, -2, -3, -4, -5, -6, -7], dtype=np.int64)
    return nib[A >> 4] + 1j * nib[A & 15]


def _numpy_herk(A, N, K, B):
    X = _decode(A.astype(np.int64)).reshape(B, N, K)
    out = np.zeros((B, N * (N + 1) // 2, 2), dtype=np.int64)
    il, jl = np.tril_indices(N)
    for b in range(B):
        G = X[b] @ X[b].conj().T
        out[b, :, 0] = np.rint(G[il, jl].real)
        im = np.rint(G[il, jl].imag).astype(np.int64)
        im[il == jl] = 0
        out[b, :, 1] = im
    return out


def test_hand_computed_case(oracle):
    # a0 = 1+2i, a1 = 3-i, K = 1: C00 = 5, C10 = a1 conj(a0) = (3-i)(1-2i) = 1-7i, C11 = 10
    A = np.array([oracle.encode_qc(1, 2), oracle.encode_qc(3, -1)], dtype=np.uint8)
    C = oracle.herk_int64(A, 2, 1, 1)
    assert C[0].tolist() == [[5, 0], [1, -7], [10, 0]]


@pytest.mark.parametrize("N,K,B,raw", [(1, 1, 1, False), (7, 3, 2, False), (33, 64, 3, False),
                                        (50, 129, 1, True), (64, 17, 2, True)])
def test_oracle_vs_numpy(oracle, N, K, B, raw):
    A = (oracle.gen_bytes if raw else oracle.gen_qc_uniform)(N * 31 + K, N * K * B)
    assert np.array_equal(oracle.herk_int64(A, N, K, B), _numpy_herk(A, N, K, B))


def test_closed_forms(oracle):
    N, K = 24, 300
    A = np.full(N * K, oracle.encode_qc(7, 7), dtype=np.uint8)
    C = oracle.herk_int64(A, N, K, 1)
    assert np.all(C[..., 0] == 98 * K) and np.all(C[..., 1] == 0)
    # K = 1: outer product a a^H
    a = oracle.gen_qc_uniform(5, 40)
    x = _decode(a.astype(np.int64))
    C = oracle.herk_int64(a, 40, 1, 1)
    il, jl = np.tril_indices(40)
    outer = np.outer(x, x.conj())[il, jl]
    assert np.array_equal(C[0, :, 0], outer.real.astype(np.int64))
    im = outer.imag.astype(np.int64)
    im[il == jl] = 0
    assert np.array_equal(C[0, :, 1], im)
    # zero input
    assert not np.any(oracle.herk_int64(np.zeros(30 * 8, np.uint8), 30, 8, 1))


def test_invariants(oracle):
    N, K, B = 40, 200, 2
    A = oracle.gen_bytes(9, N * K * B)
    C = oracle.herk_int64(A, N, K, B)
    X = _decode(A.astype(np.int64)).reshape(B, N, K)
    d = np.arange(N) * (np.arange(N) + 1) // 2 + np.arange(N)
    assert np.array_equal(C[:, d, 0], (np.abs(X) ** 2).sum(-1).round().astype(np.int64))
    assert np.all(C[:, d, 1] == 0)
    assert np.all(C[:, d, 0] >= 0)


def test_row_subset_matches_full(oracle):
    N, K, B = 70, 40, 2
    A = oracle.gen_qc_uniform(6, N * K * B)
    full = oracle.herk_int64(A, N, K, B)
    part = oracle.herk_int64_rows(A, N, K, B, 20, 45)
    s, e = 20 * 21 // 2, 45 * 46 // 2
    assert np.array_equal(part[:, s:e], full[:, s:e])
    assert not np.any(part[:, :s]) and not np.any(part[:, e:])


def test_f32_epilogue_alpha_beta(oracle):
    # C = alpha A A^H + beta C with Im(diag) = 0 before alpha (cpu_herk_ab,
    # test_gemm_api.cu:726-751); values chosen so float64 evaluation is exact.
    N, K, B = 20, 50, 1
    A = oracle.gen_qc_uniform(12, N * K * B)
    acc = oracle.herk_int64(A, N, K, B).astype(np.float64)
    rng = np.random.default_rng(0)
    C0 = rng.integers(-500, 500, size=(B, N * (N + 1) // 2, 2)).astype(np.float32)
    for alpha, beta in [(1.0, 0.0), (0.5, 1.0), (-2.0, 0.25), (0.0, -3.0)]:
        got = oracle.herk_f32(A, N, K, B, alpha=alpha, beta=beta, C=C0)
        exp = acc * alpha + (beta * C0.astype(np.float64) if beta != 0 else 0.0)
        assert np.array_equal(got, exp.astype(np.float32)), (alpha, beta)


def test_f16_rounding_vs_numpy(oracle):
    rng = np.random.default_rng(1)
    vals = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * np.float32(1000.0),
        rng.standard_normal(5000).astype(np.float32) * np.float32(1e-5),   # subnormal range
        np.array([65504, 65519.99, 65520, 1e6, -1e6, 2 ** -24, 2 ** -25, 3 * 2 ** -26,
                  0.0, -0.0, 1.0 + 2 ** -11, 1.0 + 3 * 2 ** -11, 2049, 2051, 4097],
                 dtype=np.float32),
    ])
    got = np.array([oracle.f32_to_f16_rne(float(v)) for v in vals], dtype=np.uint16)
    exp = vals.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, exp)


def test_f16_output_path(oracle):
    N, K, B = 30, 400, 1
    A = oracle.gen_qc_uniform(13, N * K * B)
    h = oracle.herk_f16(A, N, K, B)
    exp = oracle.herk_int64(A, N, K, B).astype(np.float32).astype(np.float16).view(np.uint16)
    assert np.array_equal(h, exp)


def test_golden_fixtures(oracle):
    data = np.load(os.path.join(HERE, "golden", "herk_fixtures.npz"), allow_pickle=False)
    names = sorted({k.split("__")[0] for k in data.files})
    assert "cfg1" in names and "raw_bytes_negzero" in names
    for name in names:
        N, K, B, seed = (int(v) for v in data[f"{name}__meta"])
        gen = str(data[f"{name}__gen"])
        if f"{name}__A" in data.files:
            A = data[f"{name}__A"]
            if gen == "uniform":
                assert np.array_equal(A, oracle.gen_qc_uniform(seed, N * K * B))
            elif gen == "bytes":
                assert np.array_equal(A, oracle.gen_bytes(seed, N * K * B))
            C = oracle.herk_int64(A, N, K, B)
            assert np.array_equal(C, data[f"{name}__C"]), name
            assert np.array_equal(C, _numpy_herk(A, N, K, B)), name
        else:
            A = oracle.gen_qc_uniform(seed, N * K * B)

            def sha(a):
                return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()

            assert sha(A) == str(data[f"{name}__sha_A"])
            C = oracle.herk_int64(A, N, K, B)
            assert sha(C) == str(data[f"{name}__sha_C"])
            assert sha(oracle.herk_f32(A, N, K, B)) == str(data[f"{name}__sha_F32"])


def test_planted_mistakes_are_caught(oracle):
    """A plausible oracle mistake (conjugate on the wrong operand, swapped nibbles)
    must fail the independent numpy check."""
    N, K, B = 12, 9, 1
    A = oracle.gen_qc_uniform(14, N * K * B)
    C = oracle.herk_int64(A, N, K, B)
    swapped = ((A & 15) << 4) | (A >> 4)
    assert not np.array_equal(oracle.herk_int64(swapped.astype(np.uint8), N, K, B),
                              _numpy_herk(A, N, K, B))
    wrong_conj = C.copy()
    wrong_conj[..., 1] *= -1
    assert not np.array_equal(wrong_conj, _numpy_herk(A, N, K, B))
