"""CPU pinning of the MX (block-scaled FP6/FP4) part of the oracle.

The element formats are fixed by the OCP Microscaling Formats v1.0 spec: the
decode tables below are typed from the spec's value lists, not from the oracle.
Rounding is checked against a brute-force nearest search with ties to the even
code; the scale rule against the reference's ceil(log2(amax / max))
(cutlass_interface/gemm_complex_sm100/mxfp_kernels.hpp:147-166) on exact cases.
"""
import math

import numpy as np
import pytest

FP4_E2M1 = [0.0, 0.5, 1.0, 1.5, 2.0, 3.0, 4.0, 6.0]
FP6_E3M2 = [0.0, 0.0625, 0.125, 0.1875, 0.25, 0.3125, 0.375, 0.4375, 0.5, 0.625, 0.75, 0.875,
            1.0, 1.25, 1.5, 1.75, 2.0, 2.5, 3.0, 3.5, 4.0, 5.0, 6.0, 7.0, 8.0, 10.0, 12.0, 14.0,
            16.0, 20.0, 24.0, 28.0]
FMTS = [(0, FP4_E2M1, 3, 6.0), (1, FP6_E3M2, 5, 28.0)]


@pytest.mark.parametrize("fmt,table,sbit,maxv", FMTS)
def test_decode_table_matches_spec(oracle, fmt, table, sbit, maxv):
    n = len(table)
    for c in range(n):
        assert oracle.mx_decode(c, fmt) == table[c]
        assert oracle.mx_decode(c | (1 << sbit), fmt) == -table[c]


@pytest.mark.parametrize("fmt,table,sbit,maxv", FMTS)
def test_encode_round_trip_and_saturation(oracle, fmt, table, sbit, maxv):
    for c, v in enumerate(table):
        assert oracle.mx_encode(v, fmt) == c
        if c:
            assert oracle.mx_encode(-v, fmt) == c | (1 << sbit)
    top = len(table) - 1
    for v in (maxv, maxv * 1.01, maxv * 10, 1e30):
        assert oracle.mx_encode(v, fmt) == top
        assert oracle.mx_encode(-v, fmt) == top | (1 << sbit)


@pytest.mark.parametrize("fmt,table,sbit,maxv", FMTS)
def test_encode_nearest_ties_to_even(oracle, fmt, table, sbit, maxv):
    t = np.array(table)
    # every midpoint goes to the neighbour with the even code
    for c in range(len(table) - 1):
        mid = (table[c] + table[c + 1]) / 2
        want = c if c % 2 == 0 else c + 1
        assert oracle.mx_encode(mid, fmt) == want, (mid, c)
    # random values: nearest representable (brute force)
    rng = np.random.default_rng(fmt)
    for x in rng.uniform(0, maxv, 2000).astype(np.float32):
        d = np.abs(t - float(x))
        best = np.flatnonzero(d == d.min())
        code = oracle.mx_encode(float(x), fmt)
        assert code in best, (x, code, best)


@pytest.mark.parametrize("fmt,maxv", [(0, 6.0), (1, 28.0)])
def test_scale_rule(oracle, fmt, maxv):
    assert oracle.mx_scale_exponent(0.0, fmt) == 0
    for k in range(-20, 21):
        exact = maxv * 2.0 ** k
        assert oracle.mx_scale_exponent(exact, fmt) == 127 + k          # power of 2: no round up
        assert oracle.mx_scale_exponent(exact * 1.001, fmt) == 128 + k  # just above: next power
        assert oracle.mx_scale_exponent(exact * 0.75, fmt) == 127 + k   # in between: ceil
    big = float(np.float32(1e38))
    assert oracle.mx_scale_exponent(big, fmt) == 127 + math.ceil(math.log2(big / maxv))
    assert oracle.mx_scale_exponent(1e-44, fmt) == 0  # below 2^-127 * max: clamped to 0
    rng = np.random.default_rng(7)
    for a in rng.lognormal(0, 4, 500).astype(np.float32):
        e = oracle.mx_scale_exponent(float(a), fmt) - 127
        assert float(a) / 2.0 ** e <= maxv and float(a) / 2.0 ** (e - 1) > maxv * (1 - 1e-6)


def test_quantize_block_properties(oracle):
    rng = np.random.default_rng(1)
    rows, K = 6, 100
    X = (rng.standard_normal((rows, K, 2)) * 3).astype(np.float32)
    X[0, 5, 1] = 90.0  # outlier in one block
    for fmt, maxv, table in [(0, 6.0, FP4_E2M1), (1, 28.0, FP6_E3M2)]:
        codes, scales = oracle.mx_quantize(X, fmt)
        assert codes.shape == (rows, 4, 2, 32) and scales.shape == (rows, 4, 2)
        deq = oracle.mx_dequantize(codes, scales, fmt, K)
        for r in range(rows):
            for b in range(4):
                for part in range(2):
                    blk = X[r, b * 32:min(K, b * 32 + 32), part].astype(np.float64)
                    s = 2.0 ** (int(scales[r, b, part]) - 127)
                    assert np.abs(blk).max() <= maxv * s * (1 + 1e-7)
                    got = (deq.real if part == 0 else deq.imag)[r, b * 32:b * 32 + blk.size]
                    # nearest representable: error within half the local spacing
                    step = np.diff(np.array(table)).max() * s
                    assert np.all(np.abs(got - blk) <= step / 2 + 1e-12)
        # tail samples past K quantize to zero
        assert not np.any(codes[:, 3, :, K - 96:])


def test_int4_is_lossless_in_fp6(oracle):
    A = oracle.gen_qc_uniform(3, 4 * 64)
    nib = np.array([0, 1, 2, 3, 4, 5, 6, 7, 0, -1, -2, -3, -4, -5, -6, -7], dtype=np.float32)
    X = np.stack([nib[A >> 4], nib[A & 15]], axis=-1).reshape(4, 64, 2)
    codes, scales = oracle.mx_quantize(X, 1)
    deq = oracle.mx_dequantize(codes, scales, 1, 64)
    assert np.array_equal(deq, X[..., 0] + 1j * X[..., 1])


def test_herk_c64_vs_numpy(oracle):
    rng = np.random.default_rng(2)
    N, K, B = 9, 13, 2
    X = rng.standard_normal((B, N, K)) + 1j * rng.standard_normal((B, N, K))
    C = oracle.herk_c64(X, N, K, B)
    il, jl = np.tril_indices(N)
    for b in range(B):
        G = X[b] @ X[b].conj().T
        assert np.allclose(C[b, :, 0], G[il, jl].real, rtol=1e-13, atol=1e-12)
        im = G[il, jl].imag.copy()
        im[il == jl] = 0
        assert np.allclose(C[b, :, 1], im, rtol=1e-13, atol=1e-12)


# ---- the reference's own element converters (oracle.mx_encode_ref) ----
# Known answers worked by hand from the reference's arithmetic
# (cutlass_interface/gemm_complex_sm100/fp6_kernels.hpp:27-58, fp4_kernels.hpp:7-30):
# clamp, flush below 0.25 to a signed zero, frexp exponent clamped to the normal
# range, E3M2 mantissa rint(4 * (frac - 1)) clamped to 3 (no carry), E2M1 mantissa
# bit frac - 1 >= 0.5 (truncation).  Each value is (x, expected code).
REF_E3M2_CASES = [
    (0.0, 0x00), (-0.0, 0x00),            # -0.0 < 0 is false: no sign bit
    (0.2, 0x00), (-0.2, 0x20),            # flushed, sign of x < 0 kept
    (0.0625, 0x00), (0.24, 0x00),         # OCP subnormals: flushed here
    (0.25, 0x04), (0.3, 0x05),            # 0.3 = 1.2 * 2^-2 -> rint(0.8) = 1 -> 0.3125
    (0.475, 0x07),                        # 1.9 * 2^-2 -> rint(3.6) = 4 -> 3: 0.4375, not 0.5
    (1.125, 0x0C), (1.375, 0x0E),         # ties: rint(0.5) = 0, rint(1.5) = 2
    (1.9, 0x0F), (3.9, 0x13),             # no carry: 1.75, 3.5
    (27.9, 0x1F), (28.0, 0x1F), (30.0, 0x1F), (1e30, 0x1F), (-1e30, 0x3F),
    (float("nan"), 0x3F),                 # fmaxf(NaN, -28) = -28
]
REF_E2M1_CASES = [
    (0.0, 0x0), (0.2, 0x0), (-0.2, 0x8),
    (0.25, 0x2), (0.3, 0x2), (0.4, 0x3),  # [0.25, 1): exponent clamped up to 1.0 / 1.5
    (0.5, 0x2), (0.75, 0x3),              # 0.5 -> 1.0 (OCP keeps 0.5)
    (1.0, 0x2), (1.4, 0x2), (1.5, 0x3), (1.99, 0x3),   # truncation within the binade
    (2.5, 0x4), (3.5, 0x5), (5.9, 0x6),   # 2.5 -> 2, 3.5 -> 3 (OCP: 4), 5.9 -> 4
    (6.0, 0x7), (7.0, 0x7), (-7.0, 0xF), (float("nan"), 0xF),
]


@pytest.mark.parametrize("fmt,cases", [(1, REF_E3M2_CASES), (0, REF_E2M1_CASES)])
def test_ref_encode_known_answers(oracle, fmt, cases):
    for x, want in cases:
        assert oracle.mx_encode_ref(x, fmt) == want, (x, hex(oracle.mx_encode_ref(x, fmt)), hex(want))


def _ref_value_by_search(x: float, fmt: int) -> float:
    """The reference converter's value found by searching the format's grid, not by
    its bit arithmetic: E3M2 nearest of the four values of x's binade (ties to the
    even mantissa, 2^(e+1) not a candidate); E2M1 the largest grid value <= x from
    1.0 up, and 1.0 / 1.5 by the first significand bit below 1."""
    if math.isnan(x):
        x = -1e30
    maxv = 28.0 if fmt == 1 else 6.0
    c = min(abs(x), maxv)
    if c < 0.25:
        return 0.0
    m, e = math.frexp(c)  # c = m * 2^e, m in [0.5, 1)
    e -= 1
    if fmt == 1:
        cands = [2.0 ** e * (1 + j / 4) for j in range(4)]
        d = [abs(c - v) for v in cands]
        best = min(d)
        js = [j for j in range(4) if d[j] == best]
        return cands[js[0] if len(js) == 1 else [j for j in js if j % 2 == 0][0]]
    if c < 1.0:
        return 1.5 if 2 * m >= 1.5 else 1.0
    return max(v for v in (1.0, 1.5, 2.0, 3.0, 4.0, 6.0) if v <= c)


@pytest.mark.parametrize("fmt,table,sbit", [(1, FP6_E3M2, 5), (0, FP4_E2M1, 3)])
def test_ref_encode_every_fp16_value(oracle, fmt, table, sbit):
    """All finite FP16 values times the power-of-two factors the block scale applies."""
    h = np.arange(0, 0x7C00, dtype=np.uint16).view(np.float16).astype(np.float32)
    rng = np.random.default_rng(11)
    for f in (1.0, 2.0 ** -3, 2.0 ** 5):
        xs = np.concatenate([h, -h]) * np.float32(f)
        for x in rng.choice(xs, 6000, replace=False):
            x = float(x)
            code = oracle.mx_encode_ref(x, fmt)
            want = _ref_value_by_search(x, fmt)
            assert table[code & ((1 << sbit) - 1)] == want, (x, hex(code), want)
            assert (code >> sbit) & 1 == (1 if x < 0 else 0), (x, hex(code))


def test_ref_quantize_int4_blocks(oracle):
    """INT4 samples through the reference's quantizer: lossless in FP6 E3M2 (the
    reference's claim for ComputePrecision::FP6, cutlass_gemm_api.h:45); in FP4 a
    block whose largest magnitude is 7 (scale 2) maps 1..7 to 2, 2, 3, 4, 4, 6, 6,
    and one whose largest is 5 (scale 1) maps 1..5 to 1, 2, 3, 4, 4."""
    A = oracle.gen_qc_uniform(3, 4 * 64)
    nib = np.array([0, 1, 2, 3, 4, 5, 6, 7, 0, -1, -2, -3, -4, -5, -6, -7], dtype=np.float32)
    X = np.stack([nib[A >> 4], nib[A & 15]], axis=-1).reshape(4, 64, 2)
    codes, scales = oracle.mx_quantize(X, 1, oracle.MX_ROUND_REFERENCE)
    deq = oracle.mx_dequantize(codes, scales, 1, 64)
    assert np.array_equal(deq, X[..., 0] + 1j * X[..., 1])
    blk = np.zeros((2, 32, 2), dtype=np.float32)
    blk[0, :7, 0] = np.arange(1, 8)
    blk[0, 7:14, 1] = -np.arange(1, 8)
    blk[1, :5, 0] = np.arange(1, 6)
    codes, scales = oracle.mx_quantize(blk, 0, oracle.MX_ROUND_REFERENCE)
    deq = oracle.mx_dequantize(codes, scales, 0, 32)
    assert list(scales[0, 0]) == [128, 128] and list(scales[1, 0]) == [127, 0]
    assert list(deq[0, :7].real) == [2, 2, 3, 4, 4, 6, 6]
    assert list(deq[0, 7:14].imag) == [-2, -2, -3, -4, -4, -6, -6]
    assert list(deq[1, :5].real) == [1, 2, 3, 4, 4]


def test_reference_study_input_generator(reference_study_input):
    """First draws of std::mt19937(42) through uniform_real_distribution<float>(-0.5, 0.5)
    as a g++ (libstdc++) build prints them, rounded to FP16 as the reference's study does."""
    x = reference_study_input(1, 1, [16])[16][:5].astype(np.float32)
    want = np.array([-0.12545988, 0.296543002, 0.45071429, -0.316565216, 0.231993914], dtype=np.float32)
    assert np.array_equal(x, want.astype(np.float16).astype(np.float32))
