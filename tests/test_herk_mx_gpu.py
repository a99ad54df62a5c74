"""GPU tests of the MX block-scaled HERK (FP6 E3M2 / FP4 E2M1, 32-sample E8M0
scales along K, FP32 accumulate).

- The device quantizer's codes and scales equal the oracle's, byte for byte, under
  both element roundings: the reference's converters (the default,
  gemm_complex_sm100/fp6_kernels.hpp:27-58, fp4_kernels.hpp:7-30) and OCP nearest
  even.
- The MX HERK equals the FP64 HERK of the *dequantized* input to the MFMA's
  accumulation precision (FP32 rounding for FP4; ~1e-4 of the largest entry for
  FP6 E3M2, whose wide product range exceeds the MFMA's internal sum width).
  This is a parity check of the scaled-MFMA path.
- Against the FP64 HERK of the unquantized input the error is reported, not
  pinned: relative Frobenius and max-element error per channel, as the
  reference's precision study does (cutlass_interface/test_precision_accuracy.cu,
  thresholds 25% / 55% of peak for FP6 / FP4 at :149-160).
- INT4 input is exact in FP6 E3M2, so that case is bit-exact against the
  integer oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _voltages(seed, rows, K, outlier_frac=0.01, outlier_mag=30.0):
    """Complex Gaussian (E|x|^2 = 1) with a fraction of impulsive outliers of
    magnitude `outlier_mag` and uniform phase; returned as FP16 pairs."""
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal((rows, K)) + 1j * rng.standard_normal((rows, K))) / np.sqrt(2)
    mask = rng.random((rows, K)) < outlier_frac
    x[mask] = outlier_mag * np.exp(1j * rng.uniform(-np.pi, np.pi, mask.sum()))
    return np.stack([x.real, x.imag], axis=-1).astype(np.float16)


def _pack(codes, fmt, nblkp):
    """Oracle codes [rows, nblk, 2, 32] (one per byte) -> packed device layout
    [rows, nblkp, 2, PB]: FP4 nibble i at bits 4i, FP6 field i at bits 6i (LSB first)."""
    rows, nblk = codes.shape[0], codes.shape[1]
    bits = 4 if fmt == 0 else 6
    pb = 32 * bits // 8
    out = np.zeros((rows, nblkp, 2, pb), dtype=np.uint8)
    c = codes.astype(np.uint64)
    for r in range(rows):
        for b in range(nblk):
            for p in range(2):
                v = 0
                for i in range(32):
                    v |= int(c[r, b, p, i]) << (bits * i)
                out[r, b, p] = np.frombuffer(v.to_bytes(pb, "little"), dtype=np.uint8)
    return out


@pytest.mark.parametrize("compute", ["FP4", "FP6"])
@pytest.mark.parametrize("inp", ["FP16", "INT4", "FP16-edges"])
@pytest.mark.parametrize("rounding", [0, 1])
def test_quantizer_bytes_match_oracle(oracle, dev, compute, inp, rounding):
    """FP16-edges: Gaussian rows with one block per row holding the format maximum
    (scale 1, so planted values reach the converter unscaled) and values at the
    converters' edges: the flush threshold, rounding ties, the no-carry region, -0; and,
    for the reference rounding, a NaN (the reference's fmaxf clamp makes it -max)."""
    import torch
    from ftd_amd import ftdx
    fmt = 0 if compute == "FP4" else 1
    rows, K = 37, 200
    if inp == "FP16-edges":
        X16 = _voltages(13, rows, K, outlier_frac=0.0)
        rng = np.random.default_rng(14)
        maxv = 6.0 if fmt == 0 else 28.0
        edges = np.array([0.0, -0.0, 0.0625, 0.24, 0.25, 0.26, 0.5, 0.75, 1.125, 1.375, 1.9, 1.99,
                          2.5, 3.5, 3.9, 5.9, 6.0, 14.5, 27.9, 28.0], dtype=np.float16)
        edges = edges[np.abs(edges) <= maxv]
        for r in range(rows):
            b = r % (K // 32)
            X16[r, b * 32, 0] = maxv
            X16[r, b * 32 + 1:b * 32 + 1 + edges.size, 0] = edges * rng.choice([-1, 1], edges.size).astype(np.float16)
            X16[r, b * 32 + 1:b * 32 + 1 + edges.size, 1] = -edges
        if rounding == 0:
            X16[3, 100, 1] = np.float16(np.nan)
        X = torch.from_numpy(X16).to(dev)
        Xf = X16.astype(np.float32)
        inp = "FP16"
    elif inp == "FP16":
        X16 = _voltages(11, rows, K)
        X = torch.from_numpy(X16).to(dev)
        Xf = X16.astype(np.float32)
    else:
        A = oracle.gen_bytes(12, rows * K)  # raw bytes: -0 nibbles included
        X = torch.from_numpy(A).to(dev)
        nib = np.array([0, 1, 2, 3, 4, 5, 6, 7, 0, -1, -2, -3, -4, -5, -6, -7], dtype=np.float32)
        Xf = np.stack([nib[A >> 4], nib[A & 15]], axis=-1).reshape(rows, K, 2)
    codes, scales = ftdx.mx_quantize(X, getattr(ftdx.InputPrecision, inp),
                                     getattr(ftdx.ComputePrecision, compute), rows, K,
                                     rounding=ftdx.MxRounding(rounding))
    torch.cuda.synchronize()
    ref_codes, ref_scales = oracle.mx_quantize(Xf, fmt, rounding)
    nblkp = codes.shape[1]
    assert np.array_equal(scales.cpu().numpy()[:, :ref_scales.shape[1]], ref_scales)
    assert not np.any(scales.cpu().numpy()[:, ref_scales.shape[1]:])
    assert np.array_equal(codes.cpu().numpy(), _pack(ref_codes, fmt, nblkp))


@pytest.mark.parametrize("compute", ["FP4", "FP6"])
@pytest.mark.parametrize("output", ["fp32", "fp16"])
@pytest.mark.parametrize("N,K,B", [(200, 300, 2), (129, 64, 1), (33, 1000, 1), (256, 300, 2), (512, 1000, 1)])
@pytest.mark.parametrize("rounding", [0, 1])
def test_mx_herk_matches_fp64_of_quantized(oracle, dev, compute, output, N, K, B, rounding):
    import torch
    from ftd_amd import ftdx
    fmt = 0 if compute == "FP4" else 1
    X16 = _voltages(21 + N, B * N, K)
    ctx = ftdx.CutlassComplexGemm()
    ctx.set_mx_rounding(ftdx.MxRounding(rounding))
    C = ftdx.herk_mx(ctx, torch.from_numpy(X16).to(dev), N, K, B, ftdx.InputPrecision.FP16,
                     getattr(ftdx.ComputePrecision, compute), output=output)
    torch.cuda.synchronize()
    got = C.cpu().numpy().astype(np.float64)
    codes, scales = oracle.mx_quantize(X16.astype(np.float32), fmt, rounding)
    Xq = oracle.mx_dequantize(codes, scales, fmt, K).reshape(B, N, K)
    ref = oracle.herk_c64(Xq, N, K, B)
    scale = np.abs(ref).max()
    # FP4 products accumulate to FP32 rounding.  FP6 E3M2 spans 2^16 in product
    # magnitude and the MFMA's internal sum keeps fewer bits: ~1e-4 of the largest
    # entry, measured on MI355X (a single 32x32x64 step with random E3M2 codes
    # already differs by 2e-5 of its largest output).  FP16 output adds 2^-11.
    tol = {("fp32", "FP4"): 2e-6, ("fp32", "FP6"): 5e-4}.get((output, compute), 2e-3)
    assert np.abs(got - ref).max() <= tol * scale, np.abs(got - ref).max() / scale


@pytest.mark.parametrize("N", [200, 256])
@pytest.mark.parametrize("compute", ["FP4", "FP6"])
@pytest.mark.parametrize("cfg", [0, 1, 2, 3, 4, 5])
def test_mx_herk_every_tile_config(oracle, dev, compute, cfg, N):
    """The MX operand rows on each forced tile configuration: the FP64 HERK of the
    quantized voltages, to the same FP32 tolerance as the default configuration.  N = 256
    with FP4 takes the dense E2M1 rows on the 64-sample-stage configurations."""
    import torch
    from ftd_amd import ftdx
    K, B = 300, 2
    fmt = 0 if compute == "FP4" else 1
    X16 = _voltages(77, B * N, K)
    ctx = ftdx.CutlassComplexGemm()
    ctx.set_launch_config(cfg, 1)
    C = ftdx.herk_mx(ctx, torch.from_numpy(X16).to(dev), N, K, B, ftdx.InputPrecision.FP16,
                     getattr(ftdx.ComputePrecision, compute), output="fp32")
    torch.cuda.synchronize()
    codes, scales = oracle.mx_quantize(X16.astype(np.float32), fmt, oracle.MX_ROUND_REFERENCE)
    ref = oracle.herk_c64(oracle.mx_dequantize(codes, scales, fmt, K).reshape(B, N, K), N, K, B)
    tol = 2e-6 if compute == "FP4" else 5e-4
    got = C.cpu().numpy().astype(np.float64)
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()


@pytest.mark.parametrize("N", [130, 256])
def test_mx_fp6_int4_bit_exact(oracle, dev, N):
    """INT4 voltages are exact in FP6 E3M2 (scale 2^-2): the MX HERK is bit-exact, with
    edge tiles (N = 130) and with every tile interior (N = 256)."""
    import torch
    from ftd_amd import ftdx
    K, B = 256, 2
    A = oracle.gen_qc_uniform(31, N * K * B)
    ctx = ftdx.CutlassComplexGemm()
    C = ftdx.herk_mx(ctx, torch.from_numpy(A).to(dev), N, K, B, ftdx.InputPrecision.INT4,
                     ftdx.ComputePrecision.FP6, output="fp32")
    torch.cuda.synchronize()
    ref = oracle.herk_f32(A, N, K, B)
    assert np.array_equal(C.cpu().numpy().view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("compute,bound_fro,bound_max", [("FP6", 0.08, 0.25), ("FP4", 0.30, 0.55)])
def test_mx_error_vs_fp64_reported(oracle, dev, compute, bound_fro, bound_max):
    """Gaussian voltages with 1% 30-sigma outliers: error vs the FP64 HERK of
    the unquantized FP16 values, per channel (reported; loose sanity bounds)."""
    import torch
    from ftd_amd import ftdx
    N, K, B = 96, 512, 3
    X16 = _voltages(4, B * N, K)
    ctx = ftdx.CutlassComplexGemm()
    C = ftdx.herk_mx(ctx, torch.from_numpy(X16).to(dev), N, K, B, ftdx.InputPrecision.FP16,
                     getattr(ftdx.ComputePrecision, compute), output="fp32")
    torch.cuda.synchronize()
    got = C.cpu().numpy().astype(np.float64)
    X = (X16[..., 0].astype(np.float64) + 1j * X16[..., 1].astype(np.float64)).reshape(B, N, K)
    ref = oracle.herk_c64(X, N, K, B)
    for b in range(B):
        d = got[b] - ref[b]
        fro = np.sqrt((d ** 2).sum() / (ref[b] ** 2).sum())
        mx = np.abs(d).max() / np.abs(ref[b]).max()
        print(f"MX {compute} channel {b}: rel Frobenius {fro:.4e}, max-element {mx:.4e}")
        assert fro < bound_fro and mx < bound_max


def test_herk_dispatches_mx_and_keeps_reference_rules(oracle, dev):
    import torch
    from ftd_amd import ftdx
    N, K, B = 64, 96, 2
    X16 = _voltages(5, B * N, K)
    X = torch.from_numpy(X16).to(dev)
    g = ftdx.CutlassComplexGemm()
    C16 = torch.empty((B, N * (N + 1) // 2, 2), dtype=torch.float16, device=dev)
    assert g.herk(X, C16, N, K, B, ftdx.InputPrecision.FP16, ftdx.ComputePrecision.FP6,
                  ftdx.OutputPrecision.FP16, tune=False) == 0
    ref = ftdx.herk_mx(g, X, N, K, B, ftdx.InputPrecision.FP16, ftdx.ComputePrecision.FP6,
                       output="fp16")
    torch.cuda.synchronize()
    assert torch.equal(C16.view(torch.int16), ref.view(torch.int16))
    C32 = torch.empty((B, N * (N + 1) // 2, 2), dtype=torch.float32, device=dev)
    with pytest.raises(ftdx.UnsupportedError):
        g.herk(X, C32, N, K, B, ftdx.InputPrecision.FP16, ftdx.ComputePrecision.FP4,
               ftdx.OutputPrecision.FP32)


@pytest.mark.parametrize("N,K,B", [(300, 256, 2), (129, 70, 3)])
@pytest.mark.parametrize("tune", [False, True])
def test_herk_int4_fp6_is_lossless(oracle, dev, N, K, B, tune):
    """INT4 input with FP6 compute (lossless in the reference's E3M2 convention,
    cutlass_gemm_api.h:45) runs the E3M2 operand tiles: bit-exact FP16 output with
    alpha and beta, through the autotuner too."""
    import torch
    from ftd_amd import ftdx
    A = oracle.gen_qc_uniform(N + K + B, N * K * B)
    rng = np.random.default_rng(N)
    C0 = rng.integers(-50, 50, size=(B, N * (N + 1) // 2, 2)).astype(np.float16)
    g = ftdx.CutlassComplexGemm()
    g.set_strategy_tune_verbosity(0)
    C = torch.from_numpy(C0.copy()).to(dev)
    assert g.herk(torch.from_numpy(A).to(dev), C, N, K, B, ftdx.InputPrecision.INT4,
                  ftdx.ComputePrecision.FP6, ftdx.OutputPrecision.FP16, alpha=0.5, beta=2.0,
                  tune=tune) == 0
    torch.cuda.synchronize()
    ref = oracle.herk_f16(A, N, K, B, alpha=0.5, beta=2.0, C=C0)
    assert np.array_equal(C.cpu().numpy().view(np.uint16), ref)


@pytest.mark.parametrize("compute", ["FP4"])
@pytest.mark.parametrize("inp", ["FP16", "INT4"])
@pytest.mark.parametrize("K", [300, 512])
def test_mx_dense_rows_equal_inline_rows(oracle, dev, compute, inp, K):
    """Dense E2M1 rows (32 code bytes per block, scales in a stage-major array) on the
    default 256 x 128 tile give the same bytes as the 64-byte inline-scale rows on the
    64 x 64 tile: every accumulator sees the same block-by-block sequence of scaled MFMAs.
    Both input branches of the dense quantizer (FP16 and INT4 QC bytes)."""
    import torch
    from ftd_amd import ftdx
    N, B = 512, 2
    if inp == "FP16":
        X = torch.from_numpy(_voltages(5 + K, B * N, K)).to(dev)
    else:
        X = torch.from_numpy(oracle.gen_bytes(6 + K, B * N * K)).to(dev)
    comp = getattr(ftdx.ComputePrecision, compute)
    ip = getattr(ftdx.InputPrecision, inp)
    dense_ctx = ftdx.CutlassComplexGemm()          # default tile: all tiles interior at N = 512
    inline_ctx = ftdx.CutlassComplexGemm()
    inline_ctx.set_launch_config(2, 1)              # 64 x 64 tile: no dense rows
    Cd = ftdx.herk_mx(dense_ctx, X, N, K, B, ip, comp, output="fp32")
    Ci = ftdx.herk_mx(inline_ctx, X, N, K, B, ip, comp, output="fp32")
    torch.cuda.synchronize()
    assert torch.equal(Cd.view(torch.int32), Ci.view(torch.int32))


def test_mx_fp6_rows_past_4gib(oracle, dev):
    """FP6 MX HERK whose operand rows span more than 2^32 bytes per batch (256 inputs,
    K = 8,499,200: 4.35 GB of 64-byte-per-block rows, row 255 wholly past 4 GiB), so the
    tile loads take the 64-bit addressing path.  The voltages are zero except at 48 sample
    positions spread over K (the last ones within the final 100k samples), where every
    input gets an INT4 value; the result must equal the oracle HERK of those 48 columns bit
    for bit (INT4 is exact in E3M2 and the sums stay far below 2^24)."""
    import torch
    from ftd_amd import ftdx
    N, K, B = 256, 8_499_200, 1
    rng = np.random.default_rng(42)
    pos = np.concatenate([rng.choice(1_000_000, 16, replace=False),
                          4_000_000 + rng.choice(1_000_000, 16, replace=False),
                          K - 100_000 + rng.choice(100_000, 16, replace=False)])
    pos = np.sort(pos)
    sub = oracle.gen_qc_uniform(9, N * pos.size).reshape(N, pos.size)
    A = torch.zeros((N, K), dtype=torch.uint8, device=dev)
    A[:, torch.from_numpy(pos).to(dev)] = torch.from_numpy(sub).to(dev)
    ctx = ftdx.CutlassComplexGemm()
    print("herk_mx FP6 INT4 N=256 K=8499200 (rows past 4 GiB)", flush=True)
    C = ftdx.herk_mx(ctx, A.view(-1), N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP6,
                     output="fp32")
    torch.cuda.synchronize()
    ref = oracle.herk_f32(np.ascontiguousarray(sub).reshape(-1), N, pos.size, B)
    assert np.array_equal(C.cpu().numpy().view(np.uint32), ref.view(np.uint32))
    del A, C
    ctx.end_herk()
    torch.cuda.empty_cache()


@pytest.mark.parametrize("rounding", [0, 1])
@pytest.mark.parametrize("N", [130, 256])
def test_mx_fp4_int4_bit_exact(oracle, dev, rounding, N):
    """INT4 voltages in FP4 E2M1: every block scale is 2^k with k in -2..1 and every
    dequantized value an integer, so the products sum exactly in FP32 and the MX HERK is
    bit-exact against the FP64 HERK of the oracle's dequantized values; under the
    reference rounding a block whose largest magnitude is 7 maps 1 -> 2 and 7 -> 6."""
    import torch
    from ftd_amd import ftdx
    K, B = 320, 2
    A = oracle.gen_qc_uniform(41 + N, N * K * B)
    ctx = ftdx.CutlassComplexGemm()
    ctx.set_mx_rounding(ftdx.MxRounding(rounding))
    C = ftdx.herk_mx(ctx, torch.from_numpy(A).to(dev), N, K, B, ftdx.InputPrecision.INT4,
                     ftdx.ComputePrecision.FP4, output="fp32")
    torch.cuda.synchronize()
    nib = np.array([0, 1, 2, 3, 4, 5, 6, 7, 0, -1, -2, -3, -4, -5, -6, -7], dtype=np.float32)
    Xf = np.stack([nib[A >> 4], nib[A & 15]], axis=-1).reshape(B * N, K, 2)
    codes, scales = oracle.mx_quantize(Xf, 0, rounding)
    Xq = oracle.mx_dequantize(codes, scales, 0, K)
    assert np.array_equal(Xq, np.round(Xq))
    ref = oracle.herk_c64(Xq.reshape(B, N, K), N, K, B).astype(np.float32)
    assert np.array_equal(C.cpu().numpy().view(np.uint32), ref.view(np.uint32))


def test_reference_precision_study(oracle, dev, reference_study_input):
    """The reference's precision sweep (cutlass_interface/test_precision_accuracy.cu:172-357)
    on its own inputs: N = 512, batch 4, FP16 input, FP16 output, each precision's
    peak-relative error max|err| / max|ref| against the FP64 HERK of the FP16 values must
    stay under the reference's thresholds (FP8 15%, FP6 25%, FP4 55%; :149-160).  MX runs
    the reference's element converters (the default).  The K sweep is the reference's,
    every K's data drawn in turn so each K sees the reference's bytes; K = 16 .. 4096 are
    run at 256, 1024 and 4096."""
    import torch
    from ftd_amd import ftdx
    N, batch = 512, 4
    sweep = [16, 32, 64, 128, 256, 512, 1024, 2048, 4096]
    data = reference_study_input(N, batch, sweep)
    g = ftdx.CutlassComplexGemm()
    g.set_strategy_tune_verbosity(0)
    thresholds = {"FP8": 0.15, "FP6": 0.25, "FP4": 0.55}
    for K in (256, 1024, 4096):
        X16 = data[K].reshape(batch, N, K, 2)
        X = (X16[..., 0].astype(np.float64) + 1j * X16[..., 1].astype(np.float64))
        ref = oracle.herk_c64(X, N, K, batch)
        max_ref = np.abs(ref).max()
        dX = torch.from_numpy(X16.reshape(-1)).to(dev)
        for prec, thr in thresholds.items():
            C = torch.zeros((batch, N * (N + 1) // 2, 2), dtype=torch.float16, device=dev)
            assert g.herk(dX, C, N, K, batch, ftdx.InputPrecision.FP16, getattr(ftdx.ComputePrecision, prec),
                          ftdx.OutputPrecision.FP16, tune=False) == 0
            torch.cuda.synchronize()
            err = np.abs(C.cpu().numpy().astype(np.float64) - ref).max()
            print(f"K={K} {prec}: max_abs {err:.4e} peak-relative {err / max_ref:.4e}", flush=True)
            assert err / max_ref < thr


@pytest.mark.parametrize("rounding", [0, 1])
def test_mx_fp6_one_step_sum_width(oracle, dev, rounding):
    """Where the FP6 tolerance comes from: one 32-sample block (K = 32, one scaled MFMA step
    per output) of values that E3M2 holds exactly at scale 1, so quantization is exact.
      narrow: every value in {16, 20, 24, 28} (sign random): products within a factor of 3
              of each other, their sums exact in FP32 -- the output equals FP64 bit for bit;
      wide:   values from 0.25 to 28 in one step: products span 2^13.6, the sums are still
              multiples of 1/16 below 2^16 (exact in FP32), yet the MFMA's internal sum may
              round them -- the deviation stays within the 5e-4 x max bound the larger
              tests use, and is printed."""
    import torch
    from ftd_amd import ftdx
    N, K, B = 32, 32, 1
    rng = np.random.default_rng(90 + rounding)
    sign = rng.choice([-1.0, 1.0], size=(N, K, 2))
    narrow = rng.choice([16.0, 20.0, 24.0, 28.0], size=(N, K, 2)) * sign
    wide = rng.choice([0.25, 0.3125, 0.4375, 1.5, 3.5, 7.0, 14.0, 28.0], size=(N, K, 2)) * sign
    for X in (narrow, wide):
        X[:, 0, :] = 28.0  # every part's block maximum is 28: scale 2^0
    ctx = ftdx.CutlassComplexGemm()
    ctx.set_mx_rounding(ftdx.MxRounding(rounding))
    errs = {}
    for name, X in (("narrow", narrow), ("wide", wide)):
        X16 = X.astype(np.float16)
        assert np.array_equal(X16.astype(np.float64), X)
        C = ftdx.herk_mx(ctx, torch.from_numpy(X16.reshape(-1)).to(dev), N, K, B, ftdx.InputPrecision.FP16,
                         ftdx.ComputePrecision.FP6, output="fp32")
        torch.cuda.synchronize()
        codes, scales = oracle.mx_quantize(X16.astype(np.float32).reshape(N, K, 2), 1, rounding)
        assert np.all(scales[:, 0, :] == 127)
        Xq = oracle.mx_dequantize(codes, scales, 1, K)
        assert np.array_equal(Xq, X[..., 0] + 1j * X[..., 1])  # quantization exact
        ref = oracle.herk_c64(Xq.reshape(B, N, K), N, K, B)
        got = C.cpu().numpy().astype(np.float64)
        errs[name] = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"one-step FP6 MX, rounding {rounding}: narrow max err {errs['narrow']:.3e}, "
          f"wide max err {errs['wide']:.3e} (of the largest output)")
    assert errs["narrow"] == 0.0
    assert errs["wide"] <= 5e-4
