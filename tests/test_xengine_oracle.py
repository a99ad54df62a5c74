"""CPU pinning of the oracle's GGP XEngine (M^H M, upper-packed / full) and FTD
raw-order HERK against independent computations: numpy complex matmul on a numpy
decode, closed forms, the pinned batch-first HERK oracle through a numpy
transpose, and planted mistakes that each check must reject."""
import numpy as np
import pytest


def _np_twos(b):
    """Two's-complement nibbles by sign extension through int8 (low nibble = Re)."""
    b = np.asarray(b, dtype=np.uint8)
    re = ((b << 4).astype(np.int8) >> 4).astype(np.int64)
    im = (b.astype(np.int8) >> 4).astype(np.int64)
    return re, im


def _np_xengine(data, N, K, B):
    re, im = _np_twos(data.reshape(B, K, N))
    M = re + 1j * im                            # [B][K][N]
    return np.conj(np.transpose(M, (0, 2, 1))) @ M  # M^H M, [B][N][N]


def _pack_upper(C):
    B, N, _ = C.shape
    iu, ju = np.triu_indices(N)                 # row-major over the upper triangle
    return C[:, iu, ju]


def _as_c(x):
    return x[..., 0] + 1j * x[..., 1]


@pytest.mark.parametrize("N,K,B", [(1, 1, 1), (5, 3, 2), (13, 40, 3), (64, 17, 1)])
def test_xengine_vs_numpy_matmul(oracle, N, K, B):
    data = oracle.gen_bytes(N * 7 + K, N * K * B)  # all 256 byte values, -8 included
    ref = _np_xengine(data, N, K, B)
    tri = oracle.xengine_ggp_i64(data, N, K, B, herm=False)
    assert np.array_equal(_as_c(tri), _pack_upper(ref))
    full = oracle.xengine_ggp_i64(data, N, K, B, herm=True)
    assert np.array_equal(_as_c(full), ref)
    # planted mistakes: the transposed product and the conjugate-swapped product differ
    wrong = np.transpose(ref, (0, 2, 1))
    if N > 1 and K > 1:
        assert not np.array_equal(_as_c(full), wrong)


def test_xengine_closed_forms(oracle):
    N, K = 6, 9
    # all -8 - 8i: every conj(x_q) x_p = 128
    full = oracle.xengine_ggp_i64(np.full(N * K, 0x88, np.uint8), N, K, 1, herm=True)
    assert np.all(full[..., 0] == 128 * K) and np.all(full[..., 1] == 0)
    # 0x7F: Re = -1 (low nibble F), Im = 7 -> |x|^2 = 50
    full = oracle.xengine_ggp_i64(np.full(N * K, 0x7F, np.uint8), N, K, 1, herm=True)
    assert np.all(full[..., 0] == 50 * K) and np.all(full[..., 1] == 0)
    # K = 1: outer product conj(x_q) x_p, Hermitian, real diagonal
    x = np.array([0x01, 0x10, 0xF0, 0x0F, 0x88, 0x77], np.uint8)
    re, im = _np_twos(x)
    v = re + 1j * im
    full = _as_c(oracle.xengine_ggp_i64(x, 6, 1, 1, herm=True))[0]
    assert np.array_equal(full, np.outer(np.conj(v), v))
    assert np.array_equal(full, full.conj().T)
    assert np.all(np.diag(full).imag == 0)


def test_xengine_tri_equals_transposed_herk_oracle(oracle):
    """On [-7, 7] data the GGP upper triangle at (q, p >= q) equals the HERK lower
    triangle at (p, q) of the same voltages (two oracle functions written apart)."""
    N, K, B = 37, 50, 2
    qc = oracle.gen_qc_uniform(11, N * K * B)            # [B][N][K], QC bytes
    ggp = oracle.convert_qc_to_ftd(qc).reshape(B, N, K).transpose(0, 2, 1).copy()
    tri = oracle.xengine_ggp_i64(ggp, N, K, B, herm=False)
    low = oracle.herk_int64(qc, N, K, B)
    for q in range(N):
        for p in range(q, N):
            assert np.array_equal(tri[:, q * N + p - q * (q + 1) // 2],
                                  low[:, p * (p + 1) // 2 + q])


def _ftd_to_batch_first(raw, n_ant, n_chan, n_time, n_ti, n_pol):
    K = n_time // n_ti
    x = raw[:K * n_ti * n_ant * n_chan * n_pol].reshape(K, n_ti, n_ant, n_chan, n_pol)
    return np.ascontiguousarray(x.transpose(3, 4, 1, 2, 0)).reshape(-1)  # [chan][pol][ti][ant][K]


@pytest.mark.parametrize("n_ant,n_chan,n_time,n_ti,n_pol",
                         [(5, 2, 8, 2, 2), (16, 3, 12, 3, 1), (9, 1, 7, 2, 2), (4, 4, 16, 1, 2)])
def test_ftd_vs_batch_first_oracle(oracle, n_ant, n_chan, n_time, n_ti, n_pol):
    raw = oracle.gen_qc_uniform(n_ant + n_time, n_time * n_ant * n_chan * n_pol)
    got = oracle.herk_ftd_i64(raw, n_ant, n_chan, n_time, n_ti, n_pol, twos=False)
    K = n_time // n_ti
    ref = oracle.herk_int64(_ftd_to_batch_first(raw, n_ant, n_chan, n_time, n_ti, n_pol), n_ant, K,
                            n_chan * n_pol * n_ti)
    assert np.array_equal(got, ref)
    # the FTD two's-complement form of the same voltages gives the same result
    got2 = oracle.herk_ftd_i64(oracle.convert_qc_to_ftd(raw), n_ant, n_chan, n_time, n_ti, n_pol,
                               twos=True)
    assert np.array_equal(got2, ref)
    # planted mistake: swapping the pol and ti roles of the batch index changes the result
    if n_pol > 1 and n_ti > 1:
        bad = ref.reshape(n_chan, n_pol, n_ti, -1, 2).transpose(0, 2, 1, 3, 4).reshape(ref.shape)
        assert not np.array_equal(got, bad)


def test_ftd_twos_with_minus_eight_vs_numpy(oracle):
    n_ant, n_chan, n_time, n_ti, n_pol = 7, 2, 10, 2, 2
    raw = oracle.gen_bytes(5, n_time * n_ant * n_chan * n_pol)
    got = _as_c(oracle.herk_ftd_i64(raw, n_ant, n_chan, n_time, n_ti, n_pol, twos=True))
    re, im = _np_twos(_ftd_to_batch_first(raw, n_ant, n_chan, n_time, n_ti, n_pol))
    X = (re + 1j * im).reshape(-1, n_ant, n_time // n_ti)
    il, jl = np.tril_indices(n_ant)
    ref = (X @ np.conj(np.transpose(X, (0, 2, 1))))[:, il, jl]
    assert np.array_equal(got, ref)
