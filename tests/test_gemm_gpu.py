"""GPU parity of the complex GEMM with prepared B (prepare_b / gemm_prepared* /
gemm, cutlass_gemm_api.h:215-470) against the oracle.  The reference tests draw
integers in [-7, 7] (test_gemm_api.cu:93-110): FP8 holds them exactly and the FP32
sums are exact, so those comparisons are bit-exact.  General FP16 values are
compared with the FP64 product of their FP8 quantization (FP32 accumulation)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view({4: np.uint32, 2: np.uint16}[a.dtype.itemsize])


def _int_planes(rng, shape):
    re = rng.integers(-7, 8, shape).astype(np.float16)
    im = rng.integers(-7, 8, shape).astype(np.float16)
    return re, im


def _dev(dev, *arrs):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in arrs]


# the last three hold interior tiles (direct global-to-LDS loads) of the 128- and 256-row tiles
SHAPES = [(64, 128, 96, 2), (100, 130, 50, 3), (1, 1, 1, 1), (257, 33, 300, 1), (128, 256, 2048, 2),
          (256, 384, 256, 2), (300, 200, 192, 1)]


@pytest.mark.parametrize("M,N,K,B", SHAPES)
def test_prepared_direct_integer_bit_exact(oracle, dev, M, N, K, B):
    import torch
    from ftd_amd import ftdx
    rng = np.random.default_rng(M + N + K)
    a_re, a_im = _int_planes(rng, (B, M, K))
    b_re, b_im = _int_planes(rng, (B, N, K))
    dA_re, dA_im, dB_re, dB_im = _dev(dev, a_re, a_im, b_re, b_im)
    C_re = torch.full((B, M, N), float("nan"), dtype=torch.float32, device=dev)
    C_im = torch.full_like(C_re, float("nan"))
    g = ftdx.CutlassComplexGemm()
    print(f"gemm direct M={M} N={N} K={K} B={B}", flush=True)
    g.prepare_b(dB_re, dB_im, N, K, B)
    assert g.gemm_prepared_direct(dA_re, dA_im, C_re, C_im, M, N, K, B) == 0
    torch.cuda.synchronize()
    ref = oracle.gemm_c64(a_re.astype(np.float64) + 1j * a_im, b_re.astype(np.float64) + 1j * b_im)
    assert np.array_equal(_bits(C_re.cpu().numpy()), _bits(ref.real.astype(np.float32)))
    assert np.array_equal(_bits(C_im.cpu().numpy()), _bits(ref.imag.astype(np.float32)))


@pytest.mark.parametrize("output", ["fp32", "fp16"])
@pytest.mark.parametrize("M,N,K,B", [(96, 80, 64, 2), (256, 256, 128, 1)])
def test_gemm_prepared_outputs_and_alpha_beta(oracle, dev, output, M, N, K, B):
    import torch
    from ftd_amd import ftdx
    rng = np.random.default_rng(9)
    a_re, a_im = _int_planes(rng, (B, M, K))
    b_re, b_im = _int_planes(rng, (B, N, K))
    dA_re, dA_im, dB_re, dB_im = _dev(dev, a_re, a_im, b_re, b_im)
    dt = torch.float32 if output == "fp32" else torch.float16
    old_re = rng.integers(-50, 50, (B, M, N)).astype(np.float32)
    old_im = rng.integers(-50, 50, (B, M, N)).astype(np.float32)
    C_re = torch.from_numpy(old_re).to(dev).to(dt)
    C_im = torch.from_numpy(old_im).to(dev).to(dt)
    g = ftdx.CutlassComplexGemm()
    g.prepare_b(dB_re, dB_im, N, K, B)
    op = ftdx.OutputPrecision.FP32 if output == "fp32" else ftdx.OutputPrecision.FP16
    assert g.gemm_prepared(dA_re, dA_im, C_re, C_im, M, N, K, B, output=op, alpha=0.5,
                           beta=2.0) == 0
    torch.cuda.synchronize()
    ref = oracle.gemm_c64(a_re.astype(np.float64) + 1j * a_im, b_re.astype(np.float64) + 1j * b_im)
    for got, r, o in ((C_re, ref.real, old_re), (C_im, ref.imag, old_im)):
        want = (0.5 * r + 2.0 * o)  # exact in FP32 (small integers and halves)
        assert np.array_equal(got.float().cpu().numpy(), want.astype(np.float32).astype(
            np.float16 if output == "fp16" else np.float32).astype(np.float32))


@pytest.mark.parametrize("M,N,K,B", [(64, 128, 96, 2), (100, 130, 50, 3), (256, 256, 128, 2)])
def test_prepared_int4_direct_and_power(oracle, dev, M, N, K, B):
    import torch
    from ftd_amd import ftdx
    qc = oracle.gen_bytes(M * 7 + K, M * K * B)  # raw QC bytes, -0 nibbles included
    Aq = oracle.decode_qc_array(qc).reshape(B, M, K)
    rng = np.random.default_rng(N)
    b_re, b_im = _int_planes(rng, (B, N, K))
    dA, dB_re, dB_im = _dev(dev, qc, b_re, b_im)
    g = ftdx.CutlassComplexGemm()
    g.prepare_b(dB_re, dB_im, N, K, B)
    C_re = torch.empty((B, M, N), dtype=torch.float32, device=dev)
    C_im = torch.empty_like(C_re)
    P = torch.empty_like(C_re)
    print(f"gemm int4 M={M} N={N} K={K} B={B}", flush=True)
    assert g.gemm_prepared_direct_int4(dA, C_re, C_im, M, N, K, B) == 0
    assert g.gemm_prepared_power_int4(dA, P, M, N, K, B) == 0
    torch.cuda.synchronize()
    ref = oracle.gemm_c64(Aq, b_re.astype(np.float64) + 1j * b_im)
    re32, im32 = ref.real.astype(np.float32), ref.imag.astype(np.float32)
    assert np.array_equal(_bits(C_re.cpu().numpy()), _bits(re32))
    assert np.array_equal(_bits(C_im.cpu().numpy()), _bits(im32))
    pw = oracle.power_f32(re32, im32).reshape(B, M, N)
    assert np.array_equal(_bits(P.cpu().numpy()), _bits(pw))


def test_power_fp16_and_fp8_inputs_with_beta(oracle, dev):
    import torch
    from ftd_amd import ftdx
    M, N, K, B = 70, 96, 128, 2
    rng = np.random.default_rng(21)
    a_re, a_im = _int_planes(rng, (B, M, K))
    b_re, b_im = _int_planes(rng, (B, N, K))
    dA_re, dA_im, dB_re, dB_im = _dev(dev, a_re, a_im, b_re, b_im)
    # FP8 interleaved A: the E4M3 bytes of the same integers, [re, im] pairs
    qc = np.array([oracle.encode_qc(int(r), int(i)) for r, i in
                   zip(a_re.reshape(-1).astype(int), a_im.reshape(-1).astype(int))], dtype=np.uint8)
    fp8 = np.array([oracle.qc_to_fp8_pair(int(b)) for b in qc], dtype=np.uint8).reshape(-1)
    dA8, = _dev(dev, fp8)
    old = rng.integers(-100, 100, (B, M, N)).astype(np.float32)
    P16 = torch.from_numpy(old).to(dev)
    P8 = torch.from_numpy(old).to(dev)
    g = ftdx.CutlassComplexGemm()
    g.prepare_b(dB_re, dB_im, N, K, B)
    assert g.gemm_prepared_power(dA_re, dA_im, P16, M, N, K, B, alpha=0.25, beta=3.0) == 0
    assert g.gemm_prepared_power_fp8(dA8, P8, M, N, K, B, alpha=0.25, beta=3.0) == 0
    torch.cuda.synchronize()
    ref = oracle.gemm_c64(a_re.astype(np.float64) + 1j * a_im, b_re.astype(np.float64) + 1j * b_im)
    pw = oracle.power_f32(ref.real.astype(np.float32), ref.imag.astype(np.float32), alpha=0.25,
                          beta=3.0, old=old).reshape(B, M, N)
    assert np.array_equal(_bits(P16.cpu().numpy()), _bits(pw))
    assert np.array_equal(_bits(P8.cpu().numpy()), _bits(pw))


def test_general_fp16_values_vs_fp64_of_fp8(oracle, dev):
    """Non-integer FP16 operands: FP8 E4M3 quantization (clamp +-448, nearest-even),
    then the product within FP32 accumulation of the FP64 result."""
    import torch
    from ftd_amd import ftdx
    M, N, K, B = 64, 64, 512, 1
    rng = np.random.default_rng(5)
    a_re, a_im, b_re, b_im = [(rng.standard_normal((B, d, K)) * 3).astype(np.float16)
                              for d in (M, M, N, N)]
    a_re[0, 0, 0] = 600.0  # clamps to 448
    q = lambda x: np.vectorize(lambda v: oracle.fp8e4m3_to_float(
        oracle.f32_to_e4m3_clamped(float(v))))(x.astype(np.float32))
    ref = oracle.gemm_c64(q(a_re) + 1j * q(a_im), q(b_re) + 1j * q(b_im))
    dA_re, dA_im, dB_re, dB_im = _dev(dev, a_re, a_im, b_re, b_im)
    C_re = torch.empty((B, M, N), dtype=torch.float32, device=dev)
    C_im = torch.empty_like(C_re)
    g = ftdx.CutlassComplexGemm()
    assert g.gemm(dA_re, dA_im, dB_re, dB_im, C_re, C_im, M, N, K, B,
                  output=ftdx.OutputPrecision.FP32) == 0
    torch.cuda.synchronize()
    got = C_re.cpu().numpy().astype(np.float64) + 1j * C_im.cpu().numpy()
    scale = np.abs(ref).max()
    # The FP8 MFMA sums each 64-product step at reduced internal precision: exact on
    # integer data (every partial sum representable, tests above), about 1e-4 of the
    # largest element here (measured 9e-5); 3e-4 bounds it.
    assert np.abs(got - ref).max() <= 3e-4 * scale


def test_errors_and_prepared_b_isolation(oracle, dev):
    import torch
    from ftd_amd import ftdx
    M, N, K, B = 32, 48, 40, 1
    rng = np.random.default_rng(6)
    a_re, a_im = _int_planes(rng, (B, M, K))
    b_re, b_im = _int_planes(rng, (B, N, K))
    c_re, c_im = _int_planes(rng, (B, N, K))
    dA_re, dA_im, dB_re, dB_im, dC_re, dC_im = _dev(dev, a_re, a_im, b_re, b_im, c_re, c_im)
    out_re = torch.empty((B, M, N), dtype=torch.float32, device=dev)
    out_im = torch.empty_like(out_re)
    g = ftdx.CutlassComplexGemm()
    # not prepared / mismatched shape: invalid argument, no launch
    assert g.gemm_prepared_direct(dA_re, dA_im, out_re, out_im, M, N, K, B) == ftdx.FTDX_ERR_INVALID_ARG
    g.prepare_b(dB_re, dB_im, N, K, B)
    assert g.gemm_prepared_direct(dA_re, dA_im, out_re, out_im, M, N + 1, K, B) == \
        ftdx.FTDX_ERR_INVALID_ARG
    # FP6 / FP4 prepare_b is the block-scaled path; the direct and power variants need
    # FP8-prepared B (cutlass_gemm_api.h:319, 346, 374, 396)
    g.prepare_b(dB_re, dB_im, N, K, B, compute=ftdx.ComputePrecision.FP6)
    with pytest.raises(ftdx.UnsupportedError):
        g.gemm_prepared_direct(dA_re, dA_im, out_re, out_im, M, N, K, B)
    g.prepare_b(dB_re, dB_im, N, K, B)
    # a one-shot gemm with another B leaves the prepared B untouched
    tmp_re = torch.empty((B, M, N), dtype=torch.float32, device=dev)
    tmp_im = torch.empty_like(tmp_re)
    assert g.gemm(dA_re, dA_im, dC_re, dC_im, tmp_re, tmp_im, M, N, K, B,
                  output=ftdx.OutputPrecision.FP32) == 0
    assert g.gemm_prepared_direct(dA_re, dA_im, out_re, out_im, M, N, K, B) == 0
    torch.cuda.synchronize()
    ref = oracle.gemm_c64(a_re.astype(np.float64) + 1j * a_im, b_re.astype(np.float64) + 1j * b_im)
    assert np.array_equal(out_re.cpu().numpy(), ref.real.astype(np.float32))
    ref2 = oracle.gemm_c64(a_re.astype(np.float64) + 1j * a_im, c_re.astype(np.float64) + 1j * c_im)
    assert np.array_equal(tmp_im.cpu().numpy(), ref2.imag.astype(np.float32))
    # M = 0 is a no-op; K = 0 leaves beta * C
    assert g.gemm_prepared_direct(dA_re, dA_im, out_re, out_im, 0, N, K, B) == 0
    g.prepare_b(dB_re, dB_im, N, 0, B)
    out_re.fill_(4.0)
    out_im.fill_(-2.0)
    assert g.gemm_prepared_direct(dA_re, dA_im, out_re, out_im, M, N, 0, B, alpha=1.0, beta=0.5) == 0
    torch.cuda.synchronize()
    assert torch.all(out_re == 2.0) and torch.all(out_im == -1.0)


def _gauss_planes(rng, shape, scale):
    return ((rng.standard_normal(shape) * scale).astype(np.float16),
            (rng.standard_normal(shape) * scale).astype(np.float16))


def _mx_dequant(oracle, re, im, fmt, rounding=0):
    """FP64 values the MFMA sees after MX quantization of [batch, rows, K] planes (the
    contexts' default element rounding: the reference's converters)."""
    Bt, R, K = re.shape
    X = np.stack([re, im], axis=-1).astype(np.float32).reshape(Bt * R, K, 2)
    codes, scales = oracle.mx_quantize(X, fmt, rounding)
    return oracle.mx_dequantize(codes, scales, fmt, K).reshape(Bt, R, K)


@pytest.mark.parametrize("compute", ["FP6", "FP4"])
@pytest.mark.parametrize("output", ["fp32", "fp16"])
@pytest.mark.parametrize("M,N,K,B", [(96, 130, 200, 2), (256, 256, 256, 1)])
@pytest.mark.parametrize("rounding", [0, 1])
def test_mx_gemm_prepared_matches_fp64_of_quantized(oracle, dev, compute, output, M, N, K, B, rounding):
    """FP6 / FP4 prepare_b + gemm_prepared (block-scaled, cutlass_gemm_api.h:40-47, 215-299):
    the product equals the FP64 GEMM of the MX-dequantized A and B, to the MFMA's
    accumulation (same bounds as the MX HERK), with alpha and beta applied; both element
    roundings (0: the reference's converters, 1: nearest even)."""
    import torch
    from ftd_amd import ftdx
    fmt = 0 if compute == "FP4" else 1
    rng = np.random.default_rng(M + N + K + fmt)
    a_re, a_im = _gauss_planes(rng, (B, M, K), 1.0)
    b_re, b_im = _gauss_planes(rng, (B, N, K), 0.3)
    dA_re, dA_im, dB_re, dB_im = _dev(dev, a_re, a_im, b_re, b_im)
    dt = torch.float32 if output == "fp32" else torch.float16
    old_re = rng.integers(-4, 4, (B, M, N)).astype(np.float32)
    old_im = rng.integers(-4, 4, (B, M, N)).astype(np.float32)
    C_re = torch.from_numpy(old_re).to(dev).to(dt)
    C_im = torch.from_numpy(old_im).to(dev).to(dt)
    g = ftdx.CutlassComplexGemm()
    g.set_mx_rounding(ftdx.MxRounding(rounding))
    g.prepare_b(dB_re, dB_im, N, K, B, compute=getattr(ftdx.ComputePrecision, compute))
    op = ftdx.OutputPrecision.FP32 if output == "fp32" else ftdx.OutputPrecision.FP16
    print(f"mx gemm {compute} {output} M={M} N={N} K={K} B={B} rounding={rounding}", flush=True)
    assert g.gemm_prepared(dA_re, dA_im, C_re, C_im, M, N, K, B, output=op, alpha=0.5, beta=2.0) == 0
    torch.cuda.synchronize()
    Aq = _mx_dequant(oracle, a_re, a_im, fmt, rounding)
    Bq = _mx_dequant(oracle, b_re, b_im, fmt, rounding)
    ref = 0.5 * np.einsum("bmk,bnk->bmn", Aq, Bq) + 2.0 * (old_re + 1j * old_im)
    got = C_re.double().cpu().numpy() + 1j * C_im.double().cpu().numpy()
    scale = np.abs(ref).max()
    tol = {("fp32", "FP4"): 2e-6, ("fp32", "FP6"): 5e-4}.get((output, compute), 2e-3)
    assert np.abs(got - ref).max() <= tol * scale, np.abs(got - ref).max() / scale


def test_mx_gemm_int4_a_and_one_shot(oracle, dev):
    """INT4 A is exact in FP6 E3M2; B is block-scaled.  The one-shot gemm() with FP4
    compute matches its prepared form and leaves a prepared FP6 B untouched."""
    import torch
    from ftd_amd import ftdx
    M, N, K, B = 128, 160, 256, 2
    qc = oracle.gen_qc_uniform(31, M * K * B)
    Aq = oracle.decode_qc_array(qc).reshape(B, M, K)
    rng = np.random.default_rng(5)
    b_re, b_im = _gauss_planes(rng, (B, N, K), 0.3)
    c_re, c_im = _gauss_planes(rng, (B, N, K), 0.3)
    dA, dB_re, dB_im, dC_re, dC_im = _dev(dev, qc, b_re, b_im, c_re, c_im)
    g = ftdx.CutlassComplexGemm()
    g.prepare_b(dB_re, dB_im, N, K, B, compute=ftdx.ComputePrecision.FP6)
    out = [torch.empty((B, M, N), dtype=torch.float32, device=dev) for _ in range(4)]
    assert g.gemm_prepared(dA, None, out[0], out[1], M, N, K, B, input=ftdx.InputPrecision.INT4) == 0
    # one-shot FP4 gemm on another B (FP16 A: the decoded integers)
    a_re = Aq.real.astype(np.float16)
    a_im = Aq.imag.astype(np.float16)
    dA_re, dA_im = _dev(dev, a_re, a_im)
    assert g.gemm(dA_re, dA_im, dC_re, dC_im, out[2], out[3], M, N, K, B,
                  compute=ftdx.ComputePrecision.FP4, output=ftdx.OutputPrecision.FP32) == 0
    torch.cuda.synchronize()
    ref6 = np.einsum("bmk,bnk->bmn", Aq, _mx_dequant(oracle, b_re, b_im, 1))
    got6 = out[0].double().cpu().numpy() + 1j * out[1].double().cpu().numpy()
    assert np.abs(got6 - ref6).max() <= 5e-4 * np.abs(ref6).max()
    ref4 = np.einsum("bmk,bnk->bmn", _mx_dequant(oracle, a_re, a_im, 0), _mx_dequant(oracle, c_re, c_im, 0))
    got4 = out[2].double().cpu().numpy() + 1j * out[3].double().cpu().numpy()
    assert np.abs(got4 - ref4).max() <= 2e-6 * np.abs(ref4).max()
    # the prepared FP6 B is still in place
    again = [torch.empty_like(out[0]) for _ in range(2)]
    assert g.gemm_prepared(dA, None, again[0], again[1], M, N, K, B, input=ftdx.InputPrecision.INT4) == 0
    torch.cuda.synchronize()
    assert torch.equal(again[0], out[0]) and torch.equal(again[1], out[1])


@pytest.mark.parametrize("fuse", [0, 2, 4])
def test_gemm_prepared_fused_matches_per_batch(oracle, dev, fuse):
    """gemm_prepared_fused (cutlass_gemm_api.h:440-470): with the prepared weight sets one B
    repeated, any fuse factor gives gemm_prepared's bytes; with distinct sets, fused batch e
    uses set e (gemm_complex_fp8.hpp:418-472)."""
    import torch
    from ftd_amd import ftdx
    M, N, K, B = 48, 96, 160, 8
    rng = np.random.default_rng(30 + fuse)
    a_re, a_im = _int_planes(rng, (B, M, K))
    b_re, b_im = _int_planes(rng, (1, N, K))
    dA_re, dA_im = _dev(dev, a_re, a_im)
    g = ftdx.CutlassComplexGemm()
    rep_re, rep_im = _dev(dev, np.repeat(b_re, B, 0), np.repeat(b_im, B, 0))
    g.prepare_b(rep_re, rep_im, N, K, B)
    C1 = [torch.empty((B, M, N), dtype=torch.float32, device=dev) for _ in range(4)]
    assert g.gemm_prepared(dA_re, dA_im, C1[0], C1[1], M, N, K, B) == 0
    assert g.gemm_prepared_fused(dA_re, dA_im, C1[2], C1[3], M, N, K, B, fuse_factor=fuse) == 0
    torch.cuda.synchronize()
    assert torch.equal(C1[0], C1[2]) and torch.equal(C1[1], C1[3])
    ref = oracle.gemm_c64(a_re.astype(np.float64) + 1j * a_im, np.repeat(b_re, B, 0) + 1j * np.repeat(b_im, B, 0))
    assert np.array_equal(C1[2].cpu().numpy(), ref.real.astype(np.float32))
    # distinct weight sets: fused batch e = b // ff multiplies by set e
    if fuse:
        sets_re, sets_im = _int_planes(rng, (B, N, K))
        s_re, s_im = _dev(dev, sets_re, sets_im)
        g.prepare_b(s_re, s_im, N, K, B)
        Cr, Ci = torch.empty((B, M, N), device=dev), torch.empty((B, M, N), device=dev)
        assert g.gemm_prepared_fused(dA_re, dA_im, Cr, Ci, M, N, K, B, fuse_factor=fuse) == 0
        torch.cuda.synchronize()
        Bsel = (sets_re.astype(np.float64) + 1j * sets_im)[np.arange(B) // fuse]
        ref = oracle.gemm_c64(a_re.astype(np.float64) + 1j * a_im, Bsel)
        assert np.array_equal(Cr.cpu().numpy(), ref.real.astype(np.float32))
        assert np.array_equal(Ci.cpu().numpy(), ref.imag.astype(np.float32))


def test_gemm_prepared_fused_errors_and_fp6_diagnostic(dev):
    import torch
    from ftd_amd import ftdx
    M, N, K, B = 32, 64, 64, 6
    g = ftdx.CutlassComplexGemm()
    z = torch.zeros((B, N, K), dtype=torch.float16, device=dev)
    g.prepare_b(z, z, N, K, B)
    A = torch.zeros((B, M, K), dtype=torch.float16, device=dev)
    C = torch.empty((B, M, N), dtype=torch.float32, device=dev)
    with pytest.raises(ftdx.FtdxError, match="divide"):
        g.gemm_prepared_fused(A, A, C, C, M, N, K, B, fuse_factor=4)
    with pytest.raises(ftdx.UnsupportedError, match="FP32"):
        g.gemm_prepared_fused(A, A, C, C, M, N, K, B, fuse_factor=2, output=ftdx.OutputPrecision.FP16)
    assert g.debug_fp6_real_gemm(64, 96, 160) == 0


def test_blas_glue_herk_and_gemm_fp32(oracle, dev):
    """quda::stridedBatchGEMMCutlass's HERK case and gemmBatchedCutlass
    (lib/cutlass_interface.cu:316-459) through ftdx_blas_*: FP32 interleaved integers, exact
    through FP16 and FP8 (|C| <= 2048 keeps the FP16 triangle exact)."""
    import ctypes
    import torch
    from ftd_amd import ftdx
    L = ftdx.lib()
    g = ftdx.CutlassComplexGemm()
    rng = np.random.default_rng(44)
    N, K, B = 48, 16, 3
    X = rng.integers(-7, 8, (B, N, K, 2)).astype(np.float32)
    dX = torch.from_numpy(X).to(dev)
    C = torch.zeros((B, N * (N + 1) // 2, 2), dtype=torch.float32, device=dev)
    assert L.ftdx_blas_herk_fp32(g._h, dX.data_ptr(), C.data_ptr(), N, K, B, 0) == 0
    Xc = X[..., 0].astype(np.float64) + 1j * X[..., 1]
    full = np.einsum("bik,bjk->bij", Xc, Xc.conj())
    il, jl = np.tril_indices(N)
    ref = np.stack([full[:, il, jl].real, full[:, il, jl].imag], axis=-1)
    ref[:, il == jl, 1] = 0
    assert np.array_equal(C.cpu().numpy(), ref.astype(np.float32))
    M, N2, K2 = 40, 56, 96
    A = rng.integers(-7, 8, (B, M, K2, 2)).astype(np.float32)
    Bm = rng.integers(-7, 8, (B, N2, K2, 2)).astype(np.float32)
    dA, dB = torch.from_numpy(A).to(dev), torch.from_numpy(Bm).to(dev)
    Ac = A[..., 0].astype(np.float64) + 1j * A[..., 1]
    Bc = Bm[..., 0].astype(np.float64) + 1j * Bm[..., 1]
    for conj_b in (0, 1):
        Cg = torch.zeros((B, M, N2, 2), dtype=torch.float32, device=dev)
        assert L.ftdx_blas_gemm_fp32(g._h, dA.data_ptr(), dB.data_ptr(), Cg.data_ptr(), M, N2, K2, B, conj_b, 0) == 0
        want = oracle.gemm_c64(Ac, Bc if conj_b else Bc.conj())  # Im negated unless conjugate_b
        assert np.array_equal(Cg.cpu().numpy(), np.stack([want.real, want.imag], -1).astype(np.float32)), conj_b
    del ctypes
