"""GPU parity tests of the plain FP8 HERK (libftdx.so through its C ABI).

The bar is bit-exact: the FP32 (and FP16) output bytes must equal the CPU
oracle's (int64 HERK, then the reference epilogue in FP32) for inputs in the
exact domain 98*K < 2^24.  Full-size (4096 inputs x 96 channels x K=2048)
parity is shown through exact row slices plus size-independent weighted
checksums (sum_ij w_i conj(w_j) C_ij = sum_k |sum_i w_i a_ik|^2).
Cases mirror the reference's HERK tests (cutlass_interface/test_herk_int4.cu,
test_gemm_api.cu herk_* cases, tests/visibility_pipeline_test.cpp:1190-1262).
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _herk(dev, A_np, N, K, B, output="fp32", alpha=1.0, beta=0.0, C0=None, offset=0):
    import torch
    from ftd_amd import ftdx
    if offset:
        buf = torch.zeros(A_np.size + 64, dtype=torch.uint8, device=dev)
        A = buf[offset:offset + A_np.size]
        A.copy_(torch.from_numpy(A_np))
    else:
        A = torch.from_numpy(np.ascontiguousarray(A_np)).to(dev)
    C = None
    if C0 is not None:
        C = torch.from_numpy(np.ascontiguousarray(C0)).to(dev)
    out = ftdx.herk_int4(A, N, K, B, C=C, output=output, alpha=alpha, beta=beta)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _assert_bits_equal(got, ref):
    g = np.ascontiguousarray(got)
    r = np.ascontiguousarray(ref)
    assert g.shape == r.shape, (g.shape, r.shape)
    gv = g.view(np.uint32 if g.dtype == np.float32 else np.uint16)
    rv = r.view(np.uint32 if r.dtype == np.float32 else np.uint16)
    bad = np.nonzero(gv.reshape(-1) != rv.reshape(-1))[0]
    assert bad.size == 0, (f"{bad.size} mismatches; first at flat {bad[:5]}: "
                           f"got {g.reshape(-1)[bad[:5]]} ref {r.reshape(-1)[bad[:5]]}")


def test_cfg1_bit_exact(oracle, dev):
    """128 inputs (64 dual-pol antennas), 16 channels, K=1024, seed 1."""
    N, K, B = 128, 1024, 16
    A = oracle.gen_qc_uniform(1, N * K * B)
    got = _herk(dev, A, N, K, B)
    _assert_bits_equal(got, oracle.herk_f32(A, N, K, B))


@pytest.mark.parametrize("N,K,B", [
    (1, 1, 1), (2, 1, 3), (5, 7, 2), (31, 16, 2), (33, 64, 1), (64, 128, 2),
    (127, 100, 2), (128, 96, 1), (129, 65, 2), (130, 128, 1), (255, 48, 1),
    (257, 33, 1), (300, 1000, 1), (384, 256, 2), (64, 15, 3), (96, 17, 2),
])
def test_ragged_shapes(oracle, dev, N, K, B):
    """N not a multiple of the tile, K not a multiple of 16 or of the stage."""
    A = oracle.gen_qc_uniform(1000 + N * 7 + K, N * K * B)
    _assert_bits_equal(_herk(dev, A, N, K, B), oracle.herk_f32(A, N, K, B))


def test_k_zero(oracle, dev):
    N, K, B = 40, 0, 2
    A = np.zeros(0, dtype=np.uint8)
    got = _herk(dev, A, N, K, B)
    assert not np.any(got.view(np.uint32))


def test_negative_zero_and_raw_bytes(oracle, dev):
    """Raw random bytes include the -0 nibble (0x8), as the reference's
    CutlassHerkCorrectnessTest input (tests/visibility_pipeline_test.cpp:1203-1206)."""
    N, K, B = 96, 160, 2
    A = oracle.gen_bytes(77, N * K * B)
    assert np.any((A & 0xF) == 8) and np.any((A >> 4) == 8)
    _assert_bits_equal(_herk(dev, A, N, K, B), oracle.herk_f32(A, N, K, B))


def test_unaligned_input_pointer(oracle, dev):
    """A not 16-byte aligned: the rows are copied to aligned scratch first; same bytes."""
    N, K, B = 70, 64, 2
    A = oracle.gen_qc_uniform(5, N * K * B)
    _assert_bits_equal(_herk(dev, A, N, K, B, offset=3), oracle.herk_f32(A, N, K, B))


@pytest.mark.parametrize("alpha,beta", [(0.5, 1.0), (-2.0, 0.25), (1.0, -1.0), (0.0, 3.0)])
def test_alpha_beta(oracle, dev, alpha, beta):
    """C = alpha A A^H + beta C (test_gemm_api.cu herk_beta, cpu_herk_ab :726-751)."""
    N, K, B = 100, 128, 2
    A = oracle.gen_qc_uniform(11, N * K * B)
    rng = np.random.default_rng(3)
    C0 = rng.integers(-1000, 1000, size=(B, N * (N + 1) // 2, 2)).astype(np.float32)
    got = _herk(dev, A, N, K, B, alpha=alpha, beta=beta, C0=C0)
    ref = oracle.herk_f32(A, N, K, B, alpha=alpha, beta=beta, C=C0)
    _assert_bits_equal(got, ref)


def test_fp16_output(oracle, dev):
    N, K, B = 100, 300, 3  # 98*300 < 65504: finite everywhere
    A = oracle.gen_qc_uniform(21, N * K * B)
    got = _herk(dev, A, N, K, B, output="fp16")
    _assert_bits_equal(got, oracle.herk_f16(A, N, K, B).view(np.float16))


def test_fp16_output_overflow_is_inf(oracle, dev):
    """Full-scale input overflows FP16 beyond K = 668 (lib/cutlass_interface.cu:240-242)."""
    N, K, B = 64, 1024, 1
    A = np.full(N * K * B, oracle.encode_qc(7, 7), dtype=np.uint8)
    got = _herk(dev, A, N, K, B, output="fp16")
    _assert_bits_equal(got, oracle.herk_f16(A, N, K, B).view(np.float16))
    assert np.all(np.isinf(got[..., 0]))


def test_full_scale_closed_form(oracle, dev):
    """All samples 7+7i: every Re = 98K, Im = 0 (the exactness worst case)."""
    N, K, B = 256, 4096, 2
    A = np.full(N * K * B, oracle.encode_qc(7, 7), dtype=np.uint8)
    got = _herk(dev, A, N, K, B)
    assert np.all(got[..., 0] == 98.0 * K)
    assert np.all(got[..., 1] == 0.0)


def test_one_hot_rows(oracle, dev):
    """Row i has a single non-zero sample at k = i mod K: C is diagonal-sparse."""
    N, K, B = 200, 64, 1
    A = np.zeros((B, N, K), dtype=np.uint8)
    for i in range(N):
        A[0, i, i % K] = oracle.encode_qc((i % 15) - 7, ((3 * i) % 15) - 7)
    A = A.reshape(-1)
    _assert_bits_equal(_herk(dev, A, N, K, B), oracle.herk_f32(A, N, K, B))


def test_non_default_stream(oracle, dev):
    import torch
    from ftd_amd import ftdx
    N, K, B = 150, 256, 2
    A_np = oracle.gen_qc_uniform(9, N * K * B)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        A = torch.from_numpy(A_np).to(dev, non_blocking=False)
        C = ftdx.herk_int4(A, N, K, B, stream=s)
    s.synchronize()
    _assert_bits_equal(C.cpu().numpy(), oracle.herk_f32(A_np, N, K, B))


def test_class_api_and_errors(oracle, dev):
    import torch
    from ftd_amd import ftdx
    g = ftdx.CutlassComplexGemm()
    N, K, B = 64, 64, 2
    A_np = oracle.gen_qc_uniform(4, N * K * B)
    A = torch.from_numpy(A_np).to(dev)
    C = torch.empty((B, N * (N + 1) // 2, 2), dtype=torch.float32, device=dev)
    rc = g.herk(A, C, N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                ftdx.OutputPrecision.FP32, tune=False)
    torch.cuda.synchronize()
    assert rc == 0
    _assert_bits_equal(C.cpu().numpy(), oracle.herk_f32(A_np, N, K, B))
    # invalid sizes -> non-zero status, nothing launched
    assert g.herk(A, C, -1, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                  ftdx.OutputPrecision.FP32) == ftdx.FTDX_ERR_INVALID_ARG
    # FP32 output with FP4 compute throws (cutlass_gemm_api.cu:1721-1723)
    with pytest.raises(ftdx.UnsupportedError):
        g.herk(A, C, N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP4,
               ftdx.OutputPrecision.FP32)
    # GGP glue
    ftdx.initCutlassApi(0, None, None, False, 0)
    C.fill_(-1.0)
    assert ftdx.herkBatchedCutlassQC(A, C, N, K, B) == 0
    torch.cuda.synchronize()
    _assert_bits_equal(C.cpu().numpy(), oracle.herk_f32(A_np, N, K, B))


def test_cpp_drop_in_api(dev):
    """C++ CutlassComplexGemm / ggp::herkBatchedCutlassQC through libftdx_api.so."""
    exe = os.path.join(ROOT, "tests", "cpp", "test_herk_api")
    assert os.path.exists(exe), "tests/cpp/test_herk_api not built"
    r = subprocess.run([exe, "130", "300", "3", "7"], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout


def _weighted_checksum_check(torch, A_dev, C_dev, N, K, seed):
    """sum_{i,j} w_i conj(w_j) C_ij == sum_k |sum_i w_i a_ik|^2, exact in float64."""
    lut = torch.tensor([0, 1, 2, 3, 4, 5, 6, 7, 0, -1, -2, -3, -4, -5, -6, -7],
                       dtype=torch.float64, device=A_dev.device)
    a = A_dev.reshape(N, K).long()
    ar = lut[a >> 4]
    ai = lut[a & 15]
    g = torch.Generator(device="cpu").manual_seed(seed)
    wr = torch.randint(-3, 4, (N,), generator=g).double().to(A_dev.device)
    wi = torch.randint(-3, 4, (N,), generator=g).double().to(A_dev.device)
    # s_k = sum_i w_i a_ik
    sr = wr @ ar - wi @ ai
    si = wr @ ai + wi @ ar
    rhs = float((sr * sr + si * si).sum())
    c = C_dev.reshape(-1, 2).double()
    ii = torch.arange(N, device=A_dev.device, dtype=torch.int64)
    rows = torch.repeat_interleave(ii, ii + 1)
    cols = torch.arange(rows.numel(), device=A_dev.device) - rows * (rows + 1) // 2
    diag = rows == cols
    # w_i conj(w_j) = (wr_i wr_j + wi_i wi_j) + i (wi_i wr_j - wr_i wi_j)
    pr = wr[rows] * wr[cols] + wi[rows] * wi[cols]
    pi = wi[rows] * wr[cols] - wr[rows] * wi[cols]
    # Re(p * C) = pr cr - pi ci ; lower part counted twice (C_ji = conj C_ij)
    t = pr * c[:, 0] - pi * c[:, 1]
    lhs = float(2.0 * t[~diag].sum() + t[diag].sum())
    return lhs, rhs


@pytest.mark.parametrize("tile_config", [None, 4, 5], ids=["default", "t128x128_k64_p2", "t256x128_k64_p2"])
def test_full_size_cfg2(oracle, dev, tile_config):
    """Full-size plain HERK (2048 dual-pol antennas, 96 channels, K=2048, seed 2) on the
    bench's own input (ftd_amd.selfcheck.qc_uniform), on the library default tile and on
    the two tiles the autotuner picks at this shape: exact row slices vs the oracle and
    exact weighted checksums for every channel."""
    import torch
    from ftd_amd import ftdx, selfcheck
    N, K, B = 4096, 2048, 96
    A = selfcheck.qc_uniform(N * K * B, 2, dev)
    if tile_config is None:
        C = ftdx.herk_int4(A, N, K, B)
    else:
        ctx = ftdx.CutlassComplexGemm()
        ctx.set_launch_config(tile_config, 1)
        C = torch.empty((B, N * (N + 1) // 2, 2), dtype=torch.float32, device=dev)
        rc = ctx.herk(A, C, N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                      ftdx.OutputPrecision.FP32, tune=False)
        assert rc == 0
        assert ctx.query_config(N, K, B)[0] == tile_config
    torch.cuda.synchronize()
    rows = [0, 1, 31, 32, 127, 128, 129, 255, 256, 1000, 2047, 2048, 4094, 4095]
    A2 = A.view(B, N * K)
    for b in (0, 47, 95):
        Ab = A2[b].cpu().numpy()
        ref = np.zeros((1, N * (N + 1) // 2, 2), dtype=np.int64)
        for r in rows:
            oracle.herk_int64_rows(Ab, N, K, 1, r, r + 1, out=ref)
        for r in rows:
            s, e = r * (r + 1) // 2, (r + 1) * (r + 2) // 2
            got = C[b, s:e].cpu().numpy()
            assert np.array_equal(got, ref[0, s:e].astype(np.float32)), (b, r)
    for b in range(B):
        lhs, rhs = _weighted_checksum_check(torch, A2[b], C[b], N, K, seed=100 + b)
        assert lhs == rhs, (b, lhs, rhs)


def test_precast_paths_in_chunks(dev):
    """The pre-cast paths (FP16 input, ForceBaseline INT4, E3M2 operand tiles) convert and
    multiply chunks of at most 320 MB of operand rows.  2048 inputs x 90 channels x K=1024
    (4 MB of rows per channel) runs them in several chunks, the last one partial; every
    path's output bytes must equal the fused INT4 HERK's, which is checked exactly
    (weighted checksums of every channel, exact rows)."""
    import torch
    from ftd_amd import ftdx, selfcheck
    N, K, B = 2048, 1024, 90
    A = selfcheck.qc_uniform(N * K * B, 11, dev)
    I4, F16 = ftdx.InputPrecision.INT4, ftdx.InputPrecision.FP16
    F8, O32 = ftdx.ComputePrecision.FP8, ftdx.OutputPrecision.FP32
    P = N * (N + 1) // 2
    ctx = ftdx.CutlassComplexGemm()
    ctx.set_launch_config(5, 1)

    def run(X, inp):
        C = torch.empty((B, P, 2), dtype=torch.float32, device=dev)
        assert ctx.herk(X, C, N, K, B, inp, F8, O32, tune=False) == 0
        torch.cuda.synchronize()
        return C

    ref = run(A, I4)
    ok = selfcheck.check_herk(torch, A, ref, N, K, B, rows=[0, 1, 1023, 2047], row_channels=[0, 63, 64, 89])
    assert ok["ok"], ok
    ctx.set_launch_config(5 | 16, 1)  # E3M2 operand rows
    assert torch.equal(run(A, I4).view(torch.int32), ref.view(torch.int32))
    ctx.set_launch_config(5, 1)
    ctx.set_herk_mode(ftdx.HerkMode.ForceBaseline)  # FP8 operand rows
    assert torch.equal(run(A, I4).view(torch.int32), ref.view(torch.int32))
    ctx.set_herk_mode(ftdx.HerkMode.Auto)
    re, im = selfcheck._decode(torch, A)
    X16 = torch.stack([re, im], dim=-1).to(torch.float16).contiguous()
    del re, im
    assert torch.equal(run(X16, F16).view(torch.int32), ref.view(torch.int32))
    ctx.close()


def test_padded_rows_in_chunks(dev):
    """QC rows that 16-byte loads cannot take (K % 16 != 0, or A not 16-byte aligned) are
    copied zero-padded (to 1024 samples here: the tile's stage, 1/64 more than K rounded up
    to 16), at most 320 MB of rows per chunk, and the fused tile runs on the copy.  2048
    inputs x 170 channels x K=1000 (2 MB of padded rows per channel) runs in several
    chunks.  Zero samples add nothing, so the output bytes must equal those of the same
    rows padded on the host and run at K=1008 (16-byte loads, no copy, a partial last
    stage); the weighted checksums and exact rows check that output.  The same holds at
    an unaligned address, and for three replays of a captured call after init_herk."""
    import torch
    from ftd_amd import ftdx, selfcheck
    N, K, B = 2048, 1000, 170
    Kq = 1008
    A = selfcheck.qc_uniform(N * K * B, 17, dev)
    Ap = torch.zeros((B, N, Kq), dtype=torch.uint8, device=dev)
    Ap[:, :, :K] = A.view(B, N, K)
    I4, F8, O32 = ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8, ftdx.OutputPrecision.FP32
    P = N * (N + 1) // 2
    ctx = ftdx.CutlassComplexGemm()
    ctx.set_launch_config(5, 1)

    def run(c, X, k):
        C = torch.empty((B, P, 2), dtype=torch.float32, device=dev)
        assert c.herk(X, C, N, k, B, I4, F8, O32, tune=False) == 0
        torch.cuda.synchronize()
        return C

    ref = run(ctx, Ap.view(-1), Kq)
    got = run(ctx, A, K)
    assert torch.equal(got.view(torch.int32), ref.view(torch.int32))
    ok = selfcheck.check_herk(torch, A, got, N, K, B, rows=[0, 1, 1023, 2047], row_channels=[0, 154, 155, 169])
    assert ok["ok"], ok
    del got
    buf = torch.zeros(Ap.numel() + 64, dtype=torch.uint8, device=dev)
    Au = buf[1:1 + Ap.numel()]
    Au.copy_(Ap.view(-1))
    assert torch.equal(run(ctx, Au, Kq).view(torch.int32), ref.view(torch.int32))
    del buf, Au
    ctx.close()
    g = ftdx.CutlassComplexGemm()
    g.set_herk_mode(ftdx.HerkMode.ForceDirect)
    g.set_launch_config(5, 1)
    g.set_direct_herk_graph(True)
    g.init_herk(N, K, B)
    C = torch.empty_like(ref)
    for rep in range(3):
        C.fill_(float("nan"))
        torch.cuda.synchronize()
        assert g.herk(A, C, N, K, B, I4, F8, O32, tune=False) == 0
        torch.cuda.synchronize()
        assert torch.equal(C.view(torch.int32), ref.view(torch.int32)), rep
    cached, captures = g.graph_stats()
    assert captures == 1, (cached, captures)
    g.end_herk()
    g.close()


@pytest.mark.parametrize("cfg", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("K", [80, 112, 160, 1040])
def test_partial_last_stage(oracle, dev, cfg, K):
    """K a multiple of 16 but not of the stage (16, 32 or 48 samples left over; 16 to 112 on
    the 128-sample configuration): the branch-free loop's partial-last-stage variant, on
    edge tiles (N = 300), FP32 with alpha / beta and FP16, bit-exact against the oracle."""
    import torch
    from ftd_amd import ftdx
    N, B = 300, 2
    A_np = oracle.gen_bytes(3000 + 7 * cfg + K, N * K * B)
    A = torch.from_numpy(A_np).to(dev)
    ctx = ftdx.CutlassComplexGemm()
    ctx.set_launch_config(cfg, cfg & 1)
    P = N * (N + 1) // 2
    C0 = np.random.default_rng(K + cfg).integers(-30, 30, (B, P, 2)).astype(np.float32)
    C = torch.from_numpy(C0).to(dev)
    assert ctx.herk(A, C, N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                    ftdx.OutputPrecision.FP32, alpha=0.5, beta=2.0, tune=False) == 0
    torch.cuda.synchronize()
    _assert_bits_equal(C.cpu().numpy(), oracle.herk_f32(A_np, N, K, B, alpha=0.5, beta=2.0, C=C0))
    C16 = torch.empty((B, P, 2), dtype=torch.float16, device=dev)
    assert ctx.herk(A, C16, N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                    ftdx.OutputPrecision.FP16, tune=False) == 0
    torch.cuda.synchronize()
    _assert_bits_equal(C16.cpu().numpy(), oracle.herk_f16(A_np, N, K, B))
    ctx.close()
