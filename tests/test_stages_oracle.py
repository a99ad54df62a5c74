"""CPU pinning of the oracle's stage converters (GGP convention, casts, packing)."""
import numpy as np


def test_promote_data_closed_forms(oracle):
    # two's complement nibbles, low nibble = Re (include/kernels/promote_data.cuh:80-81)
    d = np.array([0x00, 0x8F, 0x7F, 0xF8, 0x21, 0x12], dtype=np.uint8)
    got = oracle.promote_data(d)
    assert got.tolist() == [[0, 0], [-1, -8], [-1, 7], [-8, -1], [1, 2], [2, 1]]


def test_qc_to_ftd_then_promote_is_decode(oracle):
    """convert (QC sign-magnitude -> FTD two's complement) composed with the GGP
    promote gives back the QC decode: the two conventions meet exactly."""
    allb = np.arange(256, dtype=np.uint8)
    ftd = oracle.convert_qc_to_ftd(allb)
    vals = oracle.promote_data(ftd)
    for b in range(256):
        assert tuple(vals[b]) == oracle.decode_qc(b)
    # spot values from the reference's converter (visibility_pipeline_test.cpp:105-117)
    assert ftd[oracle.encode_qc(-3, 5)] == ((5 & 15) << 4) | (-3 & 15)


def test_qc_nibble_fp16_table(oracle):
    # qc_to_fp16_nib_lut (cutlass_interface/cutlass_gemm_api.cu:1548-1551)
    ref = [0x0000, 0x3C00, 0x4000, 0x4200, 0x4400, 0x4500, 0x4600, 0x4700,
           0x8000, 0xBC00, 0xC000, 0xC200, 0xC400, 0xC500, 0xC600, 0xC700]
    assert [oracle.qc_nibble_fp16(n) for n in range(16)] == ref


def test_fp16_to_e4m3_vs_torch(oracle):
    import torch
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(4000) * 50, rng.standard_normal(2000) * 1e-2,
                        np.array([0.0, -0.0, 448, 449, 1e4, -1e4, 2 ** -9, 2 ** -10, 3 * 2 ** -11,
                                  240, 248, 256, 464])]).astype(np.float16)
    clamped = np.clip(x.astype(np.float32), -448, 448)
    ref = torch.from_numpy(clamped).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    got = np.array([oracle.f32_to_e4m3_clamped(float(v)) for v in x.astype(np.float32)],
                   dtype=np.uint8)
    assert np.array_equal(got, ref)


def test_fp16_to_e4m3_every_fp16_value(oracle):
    """All 65536 FP16 bit patterns: finite values and infinities clamp to +-448 and round
    as torch's float8_e4m3fn; a NaN becomes -448 (0xFE), what the reference's
    fminf(fmaxf(x, -448), 448) makes of it (cast_kernels_common.hpp:36-37)."""
    import torch
    h = np.arange(65536, dtype=np.uint32).astype(np.uint16).view(np.float16).astype(np.float32)
    nan = np.isnan(h)
    clamped = np.clip(np.where(nan, 0, h), -448, 448).astype(np.float32)
    ref = torch.from_numpy(clamped).to(torch.float8_e4m3fn).view(torch.uint8).numpy().copy()
    ref[nan] = 0xFE
    got = np.array([oracle.f32_to_e4m3_clamped(float(v)) for v in h], dtype=np.uint8)
    assert np.array_equal(got, ref)


def test_tri_from_herm_upper_row_major(oracle):
    N, B = 7, 2
    full = np.zeros((B, N, N, 2))
    for b in range(B):
        for i in range(N):
            for j in range(N):
                full[b, i, j] = (1000 * b + 10 * i + j, -(1000 * b + 10 * i + j))
    tri = oracle.tri_from_herm(full, N, B)
    for b in range(B):
        k = 0
        for i in range(N):
            for j in range(i, N):  # row-major over the upper triangle
                assert tuple(tri[b, k]) == (1000 * b + 10 * i + j, -(1000 * b + 10 * i + j))
                assert k == i * N + j - i * (i + 1) // 2
                k += 1
