"""GPU parity of the visibility consumers: polarization sum (stand-alone and fused
into the HERK as K segments) and pillbox UV gridding, bit-exact against the oracle
on integer visibilities (exact FP32 sums, so atomic order does not matter)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view({4: np.uint32, 2: np.uint16}[a.dtype.itemsize])


@pytest.mark.parametrize("dtype", [np.float32, np.float16])
@pytest.mark.parametrize("n_pol", [2, 3])
def test_pol_reduce_triangle(oracle, dev, dtype, n_pol):
    import torch
    from ftd_amd import ftdx
    N, n_ch, n_ti = 37, 3, 2
    rng = np.random.default_rng(n_pol)
    P = N * (N + 1) // 2
    t = (rng.standard_normal((n_ch * n_pol * n_ti, P, 2)) * 100).astype(dtype)
    if dtype == np.float32:
        t = np.round(t)  # integers: exact sums for any pol count
    got = ftdx.pol_reduce_triangle(torch.from_numpy(t).to(dev), N, n_ch, n_pol, n_ti)
    torch.cuda.synchronize()
    ref = oracle.pol_reduce(t, n_ch, n_pol, n_ti)
    if dtype == np.float16 and n_pol > 2:
        # FP16 sums of three values round twice on the GPU (float accumulate) and once here
        assert np.allclose(got.cpu().numpy().astype(np.float64), ref.astype(np.float64), rtol=2e-3)
    else:
        assert np.array_equal(_bits(got.cpu().numpy()), _bits(ref))


CASES = [(64, 128, 2, 2, 1), (96, 80, 3, 2, 2), (130, 50, 1, 3, 2), (33, 17, 2, 2, 3)]


@pytest.mark.parametrize("N,K,n_ch,n_pol,n_ti", CASES)
def test_fused_herk_pol_reduce_bit_exact(oracle, dev, N, K, n_ch, n_pol, n_ti):
    """K % 16 == 0 (16-byte loads) and ragged K (rows copied zero-padded) across segments."""
    import torch
    from ftd_amd import ftdx
    B = n_ch * n_pol * n_ti
    qc = oracle.gen_bytes(N + K + n_pol, N * K * B)
    A = torch.from_numpy(qc).to(dev)
    g = ftdx.CutlassComplexGemm()
    print(f"herk_pol_reduce N={N} K={K} ch={n_ch} pol={n_pol} ti={n_ti}", flush=True)
    fused = ftdx.herk_pol_reduce(g, A, N, K, n_ch, n_pol, n_ti)
    full = ftdx.herk_int4(A, N, K, B)
    unfused = ftdx.pol_reduce_triangle(full, N, n_ch, n_pol, n_ti)
    torch.cuda.synchronize()
    per = oracle.herk_int64(qc, N, K, B).reshape(n_ch, n_pol, n_ti, -1, 2).sum(axis=1)
    ref = per.reshape(n_ch * n_ti, -1, 2).astype(np.float32)
    assert np.array_equal(_bits(fused.cpu().numpy()), _bits(ref))
    assert np.array_equal(_bits(unfused.cpu().numpy()), _bits(ref))


@pytest.mark.parametrize("cfg", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("n_pol", [2, 3])
def test_fused_pol_reduce_interior_tiles(oracle, dev, cfg, n_pol):
    """Every tile interior (N a multiple of the tile rows, K a whole number of stages, three
    stages per segment on the 64-sample configurations): two segments run the branch-free
    interior loop, three the bounds-checked one; both bit-exact against the oracle."""
    import torch
    from ftd_amd import ftdx
    N, n_ch, n_ti = 512, 2, 2
    K = 384 if cfg == 1 else 192
    B = n_ch * n_pol * n_ti
    qc = oracle.gen_qc_uniform(40 + cfg, N * K * B)
    A = torch.from_numpy(qc).to(dev)
    g = ftdx.CutlassComplexGemm()
    g.set_launch_config(cfg, 1)
    print(f"herk_pol_reduce interior N={N} K={K} pol={n_pol} cfg={cfg}", flush=True)
    fused = ftdx.herk_pol_reduce(g, A, N, K, n_ch, n_pol, n_ti)
    torch.cuda.synchronize()
    per = oracle.herk_int64(qc, N, K, B).reshape(n_ch, n_pol, n_ti, -1, 2).sum(axis=1)
    ref = per.reshape(n_ch * n_ti, -1, 2).astype(np.float32)
    assert np.array_equal(_bits(fused.cpu().numpy()), _bits(ref))
    g.close()


@pytest.mark.parametrize("cfg", [0, 1, 2, 3, 4, 5])
def test_fused_pol_reduce_every_tile_config_fp16_alpha_beta(oracle, dev, cfg):
    import torch
    from ftd_amd import ftdx
    N, K, n_ch, n_pol, n_ti = 140, 96, 2, 2, 2
    B = n_ch * n_pol * n_ti
    qc = oracle.gen_qc_uniform(cfg + 9, N * K * B)
    A = torch.from_numpy(qc).to(dev)
    g = ftdx.CutlassComplexGemm()
    g.set_launch_config(cfg, cfg & 1)
    old = np.random.default_rng(cfg).integers(-20, 20, (n_ch * n_ti, N * (N + 1) // 2, 2))
    C = torch.from_numpy(old.astype(np.float16)).to(dev)
    ftdx.herk_pol_reduce(g, A, N, K, n_ch, n_pol, n_ti, output="fp16", C=C, alpha=0.5, beta=2.0)
    torch.cuda.synchronize()
    s = oracle.herk_int64(qc, N, K, B).reshape(n_ch, n_pol, n_ti, -1, 2).sum(axis=1)
    s = s.reshape(n_ch * n_ti, -1, 2).astype(np.float64)
    diag = np.array([i * (i + 1) // 2 + i for i in range(N)])
    want = 0.5 * s + 2.0 * old  # FP32 exact (ints and halves), then one rounding to FP16
    want[:, diag, 1] = 2.0 * old[:, diag, 1]  # Im(diag) = 0 before alpha / beta
    assert np.array_equal(_bits(C.cpu().numpy()), _bits(want.astype(np.float32).astype(np.float16)))


@pytest.mark.parametrize("vis_dtype", ["fp32", "fp16"])
def test_pillbox_grid_from_herk_output(oracle, dev, vis_dtype):
    import torch
    from ftd_amd import ftdx
    N, K, Nf, Ng, cell = 48, 32, 3, 64, 2e-3
    qc = oracle.gen_qc_uniform(77, N * K * Nf)
    vis = ftdx.herk_int4(torch.from_numpy(qc).to(dev), N, K, Nf, output=vis_dtype)
    n_bl = N * (N + 1) // 2
    rng = np.random.default_rng(5)
    uv = (rng.standard_normal((n_bl, 2)) * 0.02).astype(np.float32)
    freqs = np.array([0.9e9, 1.0e9, 1.1e9, 1.3e9])
    grid = torch.zeros((Nf, Ng, Ng, 2), dtype=torch.float32, device=dev)
    ftdx.pillbox_grid(vis, torch.from_numpy(uv).to(dev), torch.from_numpy(freqs).to(dev), grid, n_bl,
                      Ng, Nf, cell, freq_offset=1)
    torch.cuda.synchronize()
    vref = vis.float().cpu().numpy()  # FP16 values widen exactly
    ref = oracle.pillbox_grid(vref, uv, freqs, n_bl, Ng, Nf, cell, freq_offset=1)
    assert np.abs(ref).sum() > 0
    assert np.array_equal(grid.cpu().numpy(), ref)


def test_pillbox_grid_large_baseline_index(oracle, dev):
    """n_bl up to 4.5M (3000 inputs): the exact baseline decode still finds the autos."""
    import torch
    from ftd_amd import ftdx
    N, Ng, cell = 3000, 128, 1e-3
    n_bl = N * (N + 1) // 2
    rng = np.random.default_rng(6)
    uv = (rng.standard_normal((n_bl, 2)) * 0.01).astype(np.float32)
    vis = rng.integers(-3, 4, (1, n_bl, 2)).astype(np.float32)
    freqs = np.array([1.2e9])
    grid = torch.zeros((1, Ng, Ng, 2), dtype=torch.float32, device=dev)
    ftdx.pillbox_grid(torch.from_numpy(vis).to(dev), torch.from_numpy(uv).to(dev),
                      torch.from_numpy(freqs).to(dev), grid, n_bl, Ng, 1, cell)
    torch.cuda.synchronize()
    ref = oracle.pillbox_grid(vis, uv, freqs, n_bl, Ng, 1, cell)
    assert np.array_equal(grid.cpu().numpy(), ref)


def _grid_ref(oracle, qc, N, K, n_ch, n_pol, uv, freqs, Ng, cell, freq_offset, alpha):
    """Oracle: exact int64 HERK summed over polarizations, alpha applied in FP32, then the
    oracle's pillbox scatter of each channel's triangle."""
    per = oracle.herk_int64(qc, N, K, n_ch * n_pol).reshape(n_ch, n_pol, -1, 2).sum(axis=1)
    vis = (per.astype(np.float32) * np.float32(alpha)).astype(np.float32)
    vis[:, [i * (i + 1) // 2 + i for i in range(N)], 1] = 0.0
    return oracle.pillbox_grid(vis, uv, freqs, N * (N + 1) // 2, Ng, n_ch, cell, freq_offset=freq_offset)


GRID_CASES = [(96, 80, 3, 2, 0), (130, 64, 2, 2, 1), (64, 37, 2, 3, 2), (257, 128, 1, 1, 0),
              (256, 192, 2, 2, 4)]


@pytest.mark.parametrize("N,K,n_ch,n_pol,cfg", GRID_CASES)
def test_fused_herk_pol_grid_bit_exact(oracle, dev, N, K, n_ch, n_pol, cfg):
    """HERK + pol sum + gridding in one pass equals the oracle's HERK, pol sum and scatter
    and the unfused GPU sequence (ftdx_herk_pol_reduce then ftdx_pillbox_grid), bit for bit:
    integer visibilities keep every FP32 cell sum exact, so the atomic order is irrelevant.
    UV from antenna positions (autos at the grid centre) plus baselines off the grid."""
    import torch
    from ftd_amd import ftdx
    Ng, cell, alpha, fo = 64, 2e-3, 0.5, 1
    qc = oracle.gen_qc_uniform(N * 7 + K, N * K * n_ch * n_pol)
    A = torch.from_numpy(qc).to(dev)
    rng = np.random.default_rng(N + K)
    pos = rng.standard_normal((N, 2)) * 0.015
    ii, jj = np.tril_indices(N)
    uv = (pos[ii] - pos[jj]).astype(np.float32)
    uv[::97] *= 40.0  # some baselines land outside the grid and are dropped
    freqs = np.array([0.8e9, 0.95e9, 1.1e9, 1.25e9, 1.4e9])
    uv_d, f_d = torch.from_numpy(uv).to(dev), torch.from_numpy(freqs).to(dev)
    g = ftdx.CutlassComplexGemm()
    g.set_launch_config(cfg, cfg & 1)
    print(f"herk_pol_grid N={N} K={K} ch={n_ch} pol={n_pol} cfg={cfg}", flush=True)
    fused = torch.zeros((n_ch, Ng, Ng, 2), dtype=torch.float32, device=dev)
    ftdx.herk_pol_grid(g, A, N, K, n_ch, n_pol, uv_d, f_d, fused, Ng, cell, freq_offset=fo, alpha=alpha)
    tri = ftdx.herk_pol_reduce(g, A, N, K, n_ch, n_pol, 1, alpha=alpha)
    unfused = torch.zeros_like(fused)
    ftdx.pillbox_grid(tri, uv_d, f_d, unfused, N * (N + 1) // 2, Ng, n_ch, cell, freq_offset=fo)
    torch.cuda.synchronize()
    ref = _grid_ref(oracle, qc, N, K, n_ch, n_pol, uv, freqs, Ng, cell, fo, alpha)
    assert np.abs(ref).sum() > 0
    assert np.array_equal(_bits(fused.cpu().numpy()), _bits(ref))
    assert np.array_equal(_bits(unfused.cpu().numpy()), _bits(ref))


def test_fused_herk_pol_grid_accumulates_and_validates(oracle, dev):
    """The grid is accumulated into (two calls give twice the values) and bad arguments
    are rejected before any launch."""
    import torch
    from ftd_amd import ftdx
    N, K, n_ch, n_pol, Ng, cell = 64, 64, 1, 2, 32, 4e-3
    qc = oracle.gen_qc_uniform(3, N * K * n_ch * n_pol)
    A = torch.from_numpy(qc).to(dev)
    rng = np.random.default_rng(1)
    uv = torch.from_numpy((rng.standard_normal((N * (N + 1) // 2, 2)) * 0.01).astype(np.float32)).to(dev)
    f = torch.tensor([1.0e9], dtype=torch.float64, device=dev)
    g = ftdx.CutlassComplexGemm()
    grid = torch.zeros((n_ch, Ng, Ng, 2), dtype=torch.float32, device=dev)
    ftdx.herk_pol_grid(g, A, N, K, n_ch, n_pol, uv, f, grid, Ng, cell)
    once = grid.clone()
    ftdx.herk_pol_grid(g, A, N, K, n_ch, n_pol, uv, f, grid, Ng, cell)
    torch.cuda.synchronize()
    assert torch.equal(grid, 2 * once)
    with pytest.raises(ftdx.FtdxError):
        ftdx.herk_pol_grid(g, A, N, K, n_ch, 0, uv, f, grid, Ng, cell)
    with pytest.raises(ftdx.FtdxError):
        ftdx.herk_pol_grid(g, A, N, K, n_ch, n_pol, None, f, grid, Ng, cell)


def test_ragged_k_rows_padded_in_chunks(dev):
    """Rows of K % 16 != 0 are copied zero-padded to K rounded up to 16, at most 320 MB per
    chunk of whole polarization groups (channels), before the fused polarization-sum and
    gridding HERKs.  Several chunks here: each output must match the same rows padded
    beforehand and run at the padded K.  The polarization sum is compared bit for bit.  The
    grid is compared to 1e-5 of its largest cell: its atomic order differs between runs, and
    the central cells sum past 2^24 (7e7, where one FP32 step is 8), so their sums round
    differently by some tens of steps.  A chunk gridded at the wrong frequencies or into the
    wrong grids would differ by about 1e-2 of the largest cell."""
    import torch
    from ftd_amd import ftdx, selfcheck
    g = ftdx.CutlassComplexGemm()
    g.set_launch_config(5, 1)
    # polarization sum: 2048 inputs x 40 channels x 2 pol x 2 time-inner, K=1000 (8 MB per channel)
    N, K, Kq, n_ch, n_pol, n_ti = 2048, 1000, 1008, 40, 2, 2
    B = n_ch * n_pol * n_ti
    A = selfcheck.qc_uniform(N * K * B, 23, dev)
    Ap = torch.zeros((B, N, Kq), dtype=torch.uint8, device=dev)
    Ap[:, :, :K] = A.view(B, N, K)
    print(f"pol_reduce ragged N={N} K={K} ch={n_ch}", flush=True)
    got = ftdx.herk_pol_reduce(g, A, N, K, n_ch, n_pol, n_ti)
    ref = ftdx.herk_pol_reduce(g, Ap.view(-1), N, Kq, n_ch, n_pol, n_ti)
    torch.cuda.synchronize()
    assert torch.equal(got.view(torch.int32), ref.view(torch.int32))
    del A, Ap, got, ref
    torch.cuda.empty_cache()
    # gridding: 1024 inputs x 160 channels x 2 pol, K=1000 (2 MB per channel), values in {-1, 0, 1}
    N, K, Kq, n_ch, n_pol, Ng, cell, fo = 1024, 1000, 1008, 160, 2, 64, 2e-3, 3
    B = n_ch * n_pol
    gen = torch.Generator(device=dev)
    gen.manual_seed(29)
    re = torch.randint(-1, 2, (B * N * K,), generator=gen, device=dev, dtype=torch.int16)
    im = torch.randint(-1, 2, (B * N * K,), generator=gen, device=dev, dtype=torch.int16)
    A = ((re & 15) | ((im & 15) << 4)).to(torch.uint8)
    del re, im
    Ap = torch.zeros((B, N, Kq), dtype=torch.uint8, device=dev)
    Ap[:, :, :K] = A.view(B, N, K)
    rng = np.random.default_rng(31)
    pos = rng.standard_normal((N, 2)) * 0.015
    ii, jj = np.tril_indices(N)
    uv = torch.from_numpy((pos[ii] - pos[jj]).astype(np.float32)).to(dev)
    freqs = torch.from_numpy(np.linspace(0.8e9, 1.6e9, n_ch + fo)).to(dev)
    print(f"pol_grid ragged N={N} K={K} ch={n_ch}", flush=True)
    grid = torch.zeros((n_ch, Ng, Ng, 2), dtype=torch.float32, device=dev)
    ftdx.herk_pol_grid(g, A, N, K, n_ch, n_pol, uv, freqs, grid, Ng, cell, freq_offset=fo)
    gref = torch.zeros_like(grid)
    ftdx.herk_pol_grid(g, Ap.view(-1), N, Kq, n_ch, n_pol, uv, freqs, gref, Ng, cell, freq_offset=fo)
    torch.cuda.synchronize()
    scale = gref.abs().max().item()
    assert scale > 0
    assert (grid - gref).abs().max().item() <= 1e-5 * scale
    g.close()


@pytest.mark.parametrize("K", [1008, 496])
def test_two_pol_segments_padded_to_stage(oracle, dev, K):
    """Two polarization segments of K a multiple of 16 but not of the 64-sample stage are
    copied zero-padded to the stage (1008 -> 1024, 496 -> 512) so they run the two-segment
    loop without bounds checks; the polarization sum and the gridding stay bit-exact against
    the oracle."""
    import torch
    from ftd_amd import ftdx
    N, n_ch, n_pol, n_ti = 200, 2, 2, 2  # grid cell sums stay below 2^24: exact in any order
    B = n_ch * n_pol * n_ti
    qc = oracle.gen_qc_uniform(K + 5, N * K * B)
    A = torch.from_numpy(qc).to(dev)
    g = ftdx.CutlassComplexGemm()
    g.set_launch_config(5, 1)
    print(f"two-segment padded K={K}", flush=True)
    fused = ftdx.herk_pol_reduce(g, A, N, K, n_ch, n_pol, n_ti)
    torch.cuda.synchronize()
    per = oracle.herk_int64(qc, N, K, B).reshape(n_ch, n_pol, n_ti, -1, 2).sum(axis=1)
    ref = per.reshape(n_ch * n_ti, -1, 2).astype(np.float32)
    assert np.array_equal(_bits(fused.cpu().numpy()), _bits(ref))
    # gridding: channels of n_pol segments (n_time_inner = 1)
    Ng, cell, alpha, fo = 64, 2e-3, 0.5, 0
    n_chg = n_ch * n_ti
    rng = np.random.default_rng(K)
    pos = rng.standard_normal((N, 2)) * 0.015
    ii, jj = np.tril_indices(N)
    uv = (pos[ii] - pos[jj]).astype(np.float32)
    freqs = np.array([0.9e9, 1.0e9, 1.1e9, 1.2e9])
    grid = torch.zeros((n_chg, Ng, Ng, 2), dtype=torch.float32, device=dev)
    ftdx.herk_pol_grid(g, A, N, K, n_chg, n_pol, torch.from_numpy(uv).to(dev), torch.from_numpy(freqs).to(dev),
                       grid, Ng, cell, freq_offset=fo, alpha=alpha)
    torch.cuda.synchronize()
    gref = _grid_ref(oracle, qc, N, K, n_chg, n_pol, uv, freqs, Ng, cell, fo, alpha)
    assert np.abs(gref).sum() > 0
    assert np.array_equal(_bits(grid.cpu().numpy()), _bits(gref))
    g.close()
