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
