"""Seeded random shapes through the HERK entry points, each checked against the oracle.

Shapes are drawn from a fixed seed so a failure names a reproducible case: N from 1 to
700 (edge tiles of every configuration), K from 1 to 1100 (tails of every stage length),
1 to 3 channels, QC bytes with -0 nibbles, alpha / beta with exactly representable
values, and every tile configuration forced in turn.
- INT4 -> FP8 -> FP32 / FP16: bit-exact against the integer oracle (with its epilogue).
- FP16 input holding the same integers: bit-exact, through the FP16 -> E4M3 pass.
- MX FP6 / FP4 on FP16 voltages, both element roundings: within the MFMA's accumulation of
  the FP64 HERK of the oracle's dequantized values.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# seeds per test family; FTDX_RANDOM_CASES widens the sweep (160: 972 cases)
CASES = int(os.environ.get("FTDX_RANDOM_CASES", "32"))


def _draw(seed):
    rng = np.random.default_rng(seed)
    N = int(rng.integers(1, 701))
    K = int(rng.integers(1, 1101))
    B = int(rng.integers(1, 4))
    cfg = int(rng.integers(0, 6))
    alpha = float(rng.choice([1.0, 0.5, -2.0]))
    beta = float(rng.choice([0.0, 1.0, 0.25]))
    return rng, N, K, B, cfg, alpha, beta


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view({4: np.uint32, 2: np.uint16}[a.dtype.itemsize])


@pytest.mark.parametrize("case", range(CASES))
@pytest.mark.parametrize("output", ["fp32", "fp16"])
def test_int4_random_shape(oracle, dev, case, output):
    import torch
    from ftd_amd import ftdx
    rng, N, K, B, cfg, alpha, beta = _draw(1000 + case)
    if output == "fp16":
        K = min(K, 250)  # |alpha| * 98 K + |beta| * 40 stays below the FP16 maximum
    A = oracle.gen_bytes(2000 + case, N * K * B)
    P = N * (N + 1) // 2
    C0 = rng.integers(-40, 40, size=(B, P, 2)).astype(np.float32)
    g = ftdx.CutlassComplexGemm()
    g.set_launch_config(cfg, int(case % 2))
    odt = torch.float32 if output == "fp32" else torch.float16
    C = torch.from_numpy(C0).to(dev).to(odt)
    print(f"random INT4 case {case}: N={N} K={K} B={B} cfg={cfg} alpha={alpha} beta={beta} {output}",
          flush=True)
    out = ftdx.OutputPrecision.FP32 if output == "fp32" else ftdx.OutputPrecision.FP16
    assert g.herk(torch.from_numpy(A).to(dev), C, N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                  out, alpha=alpha, beta=beta, tune=False) == 0
    torch.cuda.synchronize()
    if output == "fp32":
        ref = oracle.herk_f32(A, N, K, B, alpha=alpha, beta=beta, C=C0)
    else:
        ref = oracle.herk_f16(A, N, K, B, alpha=alpha, beta=beta, C=C0.astype(np.float16))
    assert np.array_equal(_bits(C.cpu().numpy()), _bits(ref))


@pytest.mark.parametrize("case", range(CASES))
def test_fp16_input_random_shape(oracle, dev, case):
    """The QC integers as interleaved FP16 pairs: the FP16 -> E4M3 pass is exact on them."""
    import torch
    from ftd_amd import ftdx
    rng, N, K, B, cfg, alpha, beta = _draw(3000 + case)
    A = oracle.gen_qc_uniform(4000 + case, N * K * B)
    nib = np.array([0, 1, 2, 3, 4, 5, 6, 7, 0, -1, -2, -3, -4, -5, -6, -7], dtype=np.float16)
    X16 = np.stack([nib[A >> 4], nib[A & 15]], axis=-1)
    g = ftdx.CutlassComplexGemm()
    g.set_launch_config(cfg, 1)
    C = torch.empty((B, N * (N + 1) // 2, 2), dtype=torch.float32, device=dev)
    print(f"random FP16-input case {case}: N={N} K={K} B={B} cfg={cfg}", flush=True)
    assert g.herk(torch.from_numpy(X16.reshape(-1)).to(dev), C, N, K, B, ftdx.InputPrecision.FP16,
                  ftdx.ComputePrecision.FP8, ftdx.OutputPrecision.FP32, tune=False) == 0
    torch.cuda.synchronize()
    assert np.array_equal(_bits(C.cpu().numpy()), _bits(oracle.herk_f32(A, N, K, B)))


@pytest.mark.parametrize("case", range(CASES // 3))
@pytest.mark.parametrize("compute", ["FP6", "FP4"])
@pytest.mark.parametrize("rounding", [0, 1])
def test_mx_random_shape(oracle, dev, case, compute, rounding):
    import torch
    from ftd_amd import ftdx
    rng, N, K, B, cfg, _, _ = _draw(5000 + case)
    N, K = min(N, 400), min(K, 700)
    fmt = 0 if compute == "FP4" else 1
    x = (rng.standard_normal((B * N, K)) + 1j * rng.standard_normal((B * N, K))) * (0.5 ** 0.5)
    X16 = np.stack([x.real, x.imag], axis=-1).astype(np.float16)
    ctx = ftdx.CutlassComplexGemm()
    ctx.set_launch_config(cfg, 1)
    ctx.set_mx_rounding(ftdx.MxRounding(rounding))
    print(f"random MX case {case}: {compute} rounding={rounding} N={N} K={K} B={B} cfg={cfg}", flush=True)
    C = ftdx.herk_mx(ctx, torch.from_numpy(X16).to(dev), N, K, B, ftdx.InputPrecision.FP16,
                     getattr(ftdx.ComputePrecision, compute), output="fp32")
    torch.cuda.synchronize()
    codes, scales = oracle.mx_quantize(X16.astype(np.float32), fmt, rounding)
    ref = oracle.herk_c64(oracle.mx_dequantize(codes, scales, fmt, K).reshape(B, N, K), N, K, B)
    got = C.cpu().numpy().astype(np.float64)
    tol = 2e-6 if compute == "FP4" else 5e-4
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()


@pytest.mark.parametrize("case", range(CASES // 2))
def test_gemm_random_shape(oracle, dev, case):
    """Complex GEMM with prepared conj(B) (prepare_b, gemm_prepared_direct[_int4],
    gemm_prepared_power_int4) on integer data: planar FP32 outputs bit-exact against the
    oracle's complex product, the power against its single-rounding float steps."""
    import torch
    from ftd_amd import ftdx
    rng, N, K, B, cfg, _, _ = _draw(7000 + case)
    M = int(rng.integers(1, 400))
    re = rng.integers(-7, 8, (B, M, K)).astype(np.float16)
    im = rng.integers(-7, 8, (B, M, K)).astype(np.float16)
    b_re = rng.integers(-7, 8, (B, N, K)).astype(np.float16)
    b_im = rng.integers(-7, 8, (B, N, K)).astype(np.float16)
    qc = oracle.gen_bytes(8000 + case, M * K * B)
    Aq = oracle.decode_qc_array(qc).reshape(B, M, K)
    dA_re, dA_im, dB_re, dB_im, dqc = (torch.from_numpy(np.ascontiguousarray(x)).to(dev)
                                       for x in (re, im, b_re, b_im, qc))
    g = ftdx.CutlassComplexGemm()
    g.set_launch_config(cfg, int(case % 2))
    print(f"random GEMM case {case}: M={M} N={N} K={K} B={B} cfg={cfg}", flush=True)
    g.prepare_b(dB_re, dB_im, N, K, B)
    C = [torch.empty((B, M, N), dtype=torch.float32, device=dev) for _ in range(5)]
    assert g.gemm_prepared_direct(dA_re, dA_im, C[0], C[1], M, N, K, B) == 0
    assert g.gemm_prepared_direct_int4(dqc, C[2], C[3], M, N, K, B) == 0
    assert g.gemm_prepared_power_int4(dqc, C[4], M, N, K, B) == 0
    torch.cuda.synchronize()
    Bc = b_re.astype(np.float64) + 1j * b_im
    ref = oracle.gemm_c64(re.astype(np.float64) + 1j * im, Bc)
    assert np.array_equal(_bits(C[0].cpu().numpy()), _bits(ref.real.astype(np.float32)))
    assert np.array_equal(_bits(C[1].cpu().numpy()), _bits(ref.imag.astype(np.float32)))
    ref4 = oracle.gemm_c64(Aq, Bc)
    re32, im32 = ref4.real.astype(np.float32), ref4.imag.astype(np.float32)
    assert np.array_equal(_bits(C[2].cpu().numpy()), _bits(re32))
    assert np.array_equal(_bits(C[3].cpu().numpy()), _bits(im32))
    assert np.array_equal(_bits(C[4].cpu().numpy()), _bits(oracle.power_f32(re32, im32).reshape(B, M, N)))


@pytest.mark.parametrize("case", range(CASES // 4))
def test_grouped_random_list(oracle, dev, case):
    """A random list of 1 to 14 problems (N 1-600, K 1-900, raw QC bytes, alpha / beta) as
    one grouped plan on a random tile configuration, run twice; every problem bit-exact."""
    import torch
    from ftd_amd import ftdx
    rng = np.random.default_rng(9000 + case)
    count = int(rng.integers(1, 15))
    cfg = int(rng.integers(-1, 6))
    probs, refs = [], []
    for i in range(count):
        N, K = int(rng.integers(1, 601)), int(rng.integers(1, 901))
        alpha = float(rng.choice([1.0, 0.5, -2.0]))
        beta = float(rng.choice([0.0, 1.0]))
        A_np = oracle.gen_bytes(9100 + 20 * case + i, N * K)
        C0 = rng.integers(-40, 40, size=(N * (N + 1) // 2, 2)).astype(np.float32)
        probs.append((torch.from_numpy(A_np).to(dev), torch.from_numpy(C0).to(dev), N, K, alpha, beta))
        refs.append((C0, oracle.herk_f32(A_np, N, K, 1, alpha=alpha, beta=beta, C=C0[None])[0]))
    print(f"random grouped case {case}: {count} problems, cfg={cfg}: "
          + " ".join(f"{p[2]}x{p[3]}" for p in probs), flush=True)
    ctx = ftdx.CutlassComplexGemm()
    plan = ftdx.GroupedPlan(ctx, probs, tile_config=cfg, use_graph=bool(case % 2))
    for rep in range(2):
        for (A, C, N, K, _, _), (C0, _) in zip(probs, refs):
            C.copy_(torch.from_numpy(C0))  # beta reads the old values again
        plan.run()
        torch.cuda.synchronize()
        for (A, C, N, K, _, _), (_, ref) in zip(probs, refs):
            assert np.array_equal(_bits(C.cpu().numpy()), _bits(ref)), (rep, N, K)
    plan.close()


@pytest.mark.parametrize("case", range(CASES // 4))
def test_pol_reduce_and_grid_random(oracle, dev, case):
    """The polarization-summed HERK (polarizations as K segments) and the fused gridding
    epilogue on random shapes: bit-exact against the oracle's int64 HERK, polarization sum
    and pillbox scatter.  N * K * n_pol * 98 stays below 2^24, so every grid cell (the
    centre collects every auto-correlation) is an exact FP32 sum."""
    import torch
    from ftd_amd import ftdx
    rng = np.random.default_rng(11000 + case)
    N, K = int(rng.integers(1, 201)), int(rng.integers(1, 201))
    n_ch, n_pol, n_ti = int(rng.integers(1, 4)), int(rng.integers(1, 4)), int(rng.integers(1, 3))
    cfg = int(rng.integers(0, 6))
    qc = oracle.gen_bytes(12000 + case, N * K * n_ch * n_pol * n_ti)
    A = torch.from_numpy(qc).to(dev)
    g = ftdx.CutlassComplexGemm()
    g.set_launch_config(cfg, int(case % 2))
    print(f"random pol/grid case {case}: N={N} K={K} ch={n_ch} pol={n_pol} ti={n_ti} cfg={cfg}", flush=True)
    fused = ftdx.herk_pol_reduce(g, A, N, K, n_ch, n_pol, n_ti)
    torch.cuda.synchronize()
    per = oracle.herk_int64(qc, N, K, n_ch * n_pol * n_ti).reshape(n_ch, n_pol, n_ti, -1, 2).sum(axis=1)
    assert np.array_equal(_bits(fused.cpu().numpy()), _bits(per.reshape(n_ch * n_ti, -1, 2).astype(np.float32)))
    # gridding on the first time-inner slice layout: n_ch channels of n_pol polarizations
    Ng, cell, alpha = 64, 2e-3, 0.5
    qg = qc[:N * K * n_ch * n_pol]
    pos = rng.standard_normal((N, 2)) * 0.015
    ii, jj = np.tril_indices(N)
    uv = (pos[ii] - pos[jj]).astype(np.float32)
    freqs = np.array([0.9e9, 1.1e9, 1.3e9])
    grid = torch.zeros((n_ch, Ng, Ng, 2), dtype=torch.float32, device=dev)
    ftdx.herk_pol_grid(g, torch.from_numpy(qg).to(dev), N, K, n_ch, n_pol, torch.from_numpy(uv).to(dev),
                       torch.from_numpy(freqs).to(dev), grid, Ng, cell, alpha=alpha)
    torch.cuda.synchronize()
    pg = oracle.herk_int64(qg, N, K, n_ch * n_pol).reshape(n_ch, n_pol, -1, 2).sum(axis=1)
    vis = (pg.astype(np.float32) * np.float32(alpha)).astype(np.float32)
    vis[:, [i * (i + 1) // 2 + i for i in range(N)], 1] = 0.0
    ref = oracle.pillbox_grid(vis, uv, freqs, N * (N + 1) // 2, Ng, n_ch, cell)
    assert np.array_equal(_bits(grid.cpu().numpy()), _bits(ref))


@pytest.mark.parametrize("case", range(CASES // 2))
def test_grouped_random_aligned_list(oracle, dev, case):
    """A random list whose problems all take 16-byte loads (K a multiple of 16, each A its own
    allocation), so the plan runs the 16-byte-only grouped kernel at the tile's own prefetch
    depth; N from 1 to 600 with some multiples of 256 (interior tiles), alpha / beta, replayed
    twice; every problem bit-exact."""
    import torch
    from ftd_amd import ftdx
    rng = np.random.default_rng(15000 + case)
    count = int(rng.integers(1, 13))
    cfg = int(rng.integers(-1, 6))
    probs, refs = [], []
    for i in range(count):
        N = int(rng.choice([256, 512])) if rng.random() < 0.3 else int(rng.integers(1, 601))
        K = 16 * int(rng.integers(1, 65))
        alpha = float(rng.choice([1.0, 0.5, -2.0]))
        beta = float(rng.choice([0.0, 1.0]))
        A_np = oracle.gen_bytes(15100 + 20 * case + i, N * K)
        C0 = rng.integers(-40, 40, size=(N * (N + 1) // 2, 2)).astype(np.float32)
        probs.append((torch.from_numpy(A_np).to(dev), torch.from_numpy(C0).to(dev), N, K, alpha, beta))
        refs.append((C0, oracle.herk_f32(A_np, N, K, 1, alpha=alpha, beta=beta, C=C0[None])[0]))
    print(f"random aligned grouped case {case}: {count} problems, cfg={cfg}: "
          + " ".join(f"{p[2]}x{p[3]}" for p in probs), flush=True)
    ctx = ftdx.CutlassComplexGemm()
    plan = ftdx.GroupedPlan(ctx, probs, tile_config=cfg, use_graph=bool(case % 2))
    for rep in range(2):
        for (A, C, N, K, _, _), (C0, _) in zip(probs, refs):
            C.copy_(torch.from_numpy(C0))
        plan.run()
        torch.cuda.synchronize()
        for (A, C, N, K, _, _), (_, ref) in zip(probs, refs):
            assert np.array_equal(_bits(C.cpu().numpy()), _bits(ref)), (rep, N, K)
    plan.close()


@pytest.mark.parametrize("case", range(CASES // 4))
def test_pol_reduce_interior_random(oracle, dev, case):
    """The polarization-summed HERK with N a multiple of 128 and K a multiple of 64, so the
    tiles of the 64-sample configurations are interior: one or two segments run the
    branch-free loop, three the bounds-checked one.  Bit-exact against the oracle."""
    import torch
    from ftd_amd import ftdx
    rng = np.random.default_rng(16000 + case)
    N = 128 * int(rng.integers(1, 5))
    K = 64 * int(rng.integers(1, 7))
    n_ch, n_pol, n_ti = int(rng.integers(1, 3)), int(rng.integers(1, 4)), int(rng.integers(1, 3))
    cfg = int(rng.integers(0, 6))
    qc = oracle.gen_qc_uniform(16100 + case, N * K * n_ch * n_pol * n_ti)
    A = torch.from_numpy(qc).to(dev)
    g = ftdx.CutlassComplexGemm()
    g.set_launch_config(cfg, int(case % 2))
    print(f"random interior pol case {case}: N={N} K={K} ch={n_ch} pol={n_pol} ti={n_ti} cfg={cfg}", flush=True)
    fused = ftdx.herk_pol_reduce(g, A, N, K, n_ch, n_pol, n_ti)
    torch.cuda.synchronize()
    per = oracle.herk_int64(qc, N, K, n_ch * n_pol * n_ti).reshape(n_ch, n_pol, n_ti, -1, 2).sum(axis=1)
    assert np.array_equal(_bits(fused.cpu().numpy()), _bits(per.reshape(n_ch * n_ti, -1, 2).astype(np.float32)))
    g.close()
