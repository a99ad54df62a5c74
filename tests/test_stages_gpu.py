"""GPU tests of the stand-alone stage kernels, the FP16-input HERK, the unfused
(HerkMode::ForceBaseline) path and graph replay, against the oracle
(bit-exact where the arithmetic is exact)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view({4: np.uint32, 2: np.uint16, 8: np.uint64}[a.dtype.itemsize])


@pytest.mark.parametrize("n", [1, 15, 16, 4096, 10007])
def test_promote_qc_to_fp8_pairs(oracle, dev, n):
    import torch
    from ftd_amd import ftdx
    qc = oracle.gen_bytes(n, n)
    out = ftdx.promote_qc_to_fp8(torch.from_numpy(qc).to(dev)).cpu().numpy()
    ref = np.array([oracle.qc_to_fp8_pair(int(b)) for b in qc], dtype=np.uint8).reshape(n, 2)
    assert np.array_equal(out, ref)


def test_promote_qc_to_fp16_planar(oracle, dev):
    import torch
    from ftd_amd import ftdx
    qc = oracle.gen_bytes(3, 5000)
    re, im = ftdx.promote_qc_to_fp16_planar(torch.from_numpy(qc).to(dev))
    tab = np.array([oracle.qc_nibble_fp16(k) for k in range(16)], dtype=np.uint16)
    assert np.array_equal(re.cpu().numpy().view(np.uint16), tab[qc >> 4])
    assert np.array_equal(im.cpu().numpy().view(np.uint16), tab[qc & 15])


@pytest.mark.parametrize("fp64", [False, True])
def test_promote_data_ggp(oracle, dev, fp64):
    import torch
    from ftd_amd import ftdx
    d = oracle.gen_bytes(4, 3001)
    out = ftdx.promote_data(torch.from_numpy(d).to(dev), fp64=fp64).cpu().numpy()
    assert np.array_equal(out.astype(np.float64), oracle.promote_data(d))


def test_promote_qc_to_fp32_and_convert_to_ftd(oracle, dev):
    import torch
    from ftd_amd import ftdx
    qc = oracle.gen_bytes(77, 4096)  # the reference test's shape: 4096 random bytes
    q = torch.from_numpy(qc).to(dev)
    f = ftdx.promote_qc_to_fp32(q).cpu().numpy()
    dec = np.array([oracle.decode_qc(int(b)) for b in qc], dtype=np.float32)
    assert np.array_equal(_bits(f), _bits(dec))  # -0 nibble gives +0
    ftd = ftdx.convert_qc_to_ftd(q).cpu().numpy()
    assert np.array_equal(ftd, oracle.convert_qc_to_ftd(qc))
    # a -0 nibble next to a positive one (0x80 then 0x42): the SWAR two's complement
    # must not carry into the neighbouring byte
    pair = np.array([0x80, 0x42, 0x08, 0x24] * 8, dtype=np.uint8)
    got = ftdx.convert_qc_to_ftd(torch.from_numpy(pair).to(dev)).cpu().numpy()
    assert np.array_equal(got, oracle.convert_qc_to_ftd(pair))
    assert got[:4].tolist() == [0x00, 0x24, 0x00, 0x42]
    # in place, odd length tail
    q2 = torch.from_numpy(qc[:1001].copy()).to(dev)
    ftdx.convert_qc_to_ftd(q2, out=q2)
    assert np.array_equal(q2.cpu().numpy(), oracle.convert_qc_to_ftd(qc[:1001]))


@pytest.mark.parametrize("N,B", [(131, 3), (1100, 2)])  # 1100: rows longer than one 4 x 256 pass
def test_pack_triangle_and_tri_from_herm(oracle, dev, N, B):
    import torch
    from ftd_amd import ftdx
    rng = np.random.default_rng(1)
    full = rng.standard_normal((B, N, N, 2)).astype(np.float32)
    old = rng.standard_normal((B, N * (N + 1) // 2, 2)).astype(np.float32)
    fd = torch.from_numpy(full).to(dev)
    p0 = ftdx.pack_triangle(fd, N, B).cpu().numpy()
    il, jl = np.tril_indices(N)
    assert np.array_equal(p0, full[:, il, jl])
    p1 = ftdx.pack_triangle(fd, N, B, old=torch.from_numpy(old).to(dev), beta=0.5).cpu().numpy()
    ref = (full[:, il, jl].astype(np.float64) + 0.5 * old.astype(np.float64)).astype(np.float32)
    assert np.array_equal(p1, ref)  # one rounding of full + beta*old (exact in float64)
    up = ftdx.tri_from_herm(fd, N, B).cpu().numpy()
    assert np.array_equal(up.astype(np.float64), oracle.tri_from_herm(full.astype(np.float64), N, B))
    up64 = ftdx.tri_from_herm(fd.double(), N, B).cpu().numpy()
    assert np.array_equal(up64, oracle.tri_from_herm(full.astype(np.float64), N, B))


def test_interleave_tri(dev):
    import torch
    from ftd_amd import ftdx
    rng = np.random.default_rng(2)
    re = rng.standard_normal(777).astype(np.float16)
    im = rng.standard_normal(777).astype(np.float16)
    for fp64 in (False, True):
        out = ftdx.interleave_tri(torch.from_numpy(re).to(dev), torch.from_numpy(im).to(dev),
                                  fp64=fp64).cpu().numpy()
        assert np.array_equal(out[:, 0], re.astype(out.dtype))
        assert np.array_equal(out[:, 1], im.astype(out.dtype))


def _fp16_pairs_from_qc(oracle, qc):
    nib = np.array([0, 1, 2, 3, 4, 5, 6, 7, 0, -1, -2, -3, -4, -5, -6, -7], dtype=np.float16)
    return np.stack([nib[qc >> 4], nib[qc & 15]], axis=-1)


@pytest.mark.parametrize("N,K,B", [(64, 128, 2), (130, 100, 1), (257, 33, 2)])
def test_fp16_input_integer_valued_bit_exact(oracle, dev, N, K, B):
    """FP16 input holding integers in [-7, 7] (the reference's herk FP16 tests,
    test_gemm_api.cu:93-110) converts exactly to FP8: bit-exact vs the oracle."""
    import torch
    from ftd_amd import ftdx
    qc = oracle.gen_qc_uniform(40 + N, N * K * B)
    X16 = _fp16_pairs_from_qc(oracle, qc)
    g = ftdx.CutlassComplexGemm()
    C = torch.empty((B, N * (N + 1) // 2, 2), dtype=torch.float32, device=dev)
    assert g.herk(torch.from_numpy(X16).to(dev), C, N, K, B, ftdx.InputPrecision.FP16,
                  ftdx.ComputePrecision.FP8, ftdx.OutputPrecision.FP32, tune=False) == 0
    torch.cuda.synchronize()
    assert np.array_equal(_bits(C.cpu().numpy()), _bits(oracle.herk_f32(qc, N, K, B)))


def test_fp16_input_general_values(oracle, dev):
    """Arbitrary FP16 voltages: FP8 E4M3 quantization (clamp +-448, nearest-even),
    then the HERK of the quantized values within the FP8 MFMA's accumulation."""
    import torch
    from ftd_amd import ftdx
    N, K, B = 96, 200, 2
    rng = np.random.default_rng(8)
    X16 = (rng.standard_normal((B * N, K, 2)) * 20).astype(np.float16)
    X16[0, 0, 0] = 1000.0  # clamps to 448
    g = ftdx.CutlassComplexGemm()
    C = torch.empty((B, N * (N + 1) // 2, 2), dtype=torch.float32, device=dev)
    assert g.herk(torch.from_numpy(X16).to(dev), C, N, K, B, ftdx.InputPrecision.FP16,
                  ftdx.ComputePrecision.FP8, ftdx.OutputPrecision.FP32, tune=False) == 0
    torch.cuda.synchronize()
    q = np.array([oracle.fp8e4m3_to_float(oracle.f32_to_e4m3_clamped(float(v)))
                  for v in X16.astype(np.float32).reshape(-1)]).reshape(B, N, K, 2)
    ref = oracle.herk_c64(q[..., 0] + 1j * q[..., 1], N, K, B)
    got = C.cpu().numpy().astype(np.float64)
    assert np.abs(got - ref).max() <= 1e-4 * np.abs(ref).max()


def test_fp16_input_every_fp16_value(oracle, dev):
    """Every FP16 bit pattern through the FP16-input HERK's E4M3 conversion: N = 2, K = 1,
    row 0 = 1 + 0i, row 1 = (a, b), so output element (1, 0) is exactly (q(a), q(b)) in FP32.
    32768 channels pair the 65536 patterns as (Re, Im), a second call swaps them; q is the
    oracle's clamp-and-round (NaN -> -448, infinities -> +-448)."""
    import torch
    from ftd_amd import ftdx
    bits = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    vals = bits.view(np.float16)
    q = np.array([oracle.fp8e4m3_to_float(oracle.f32_to_e4m3_clamped(float(v)))
                  for v in vals.astype(np.float32)], dtype=np.float32)
    B = 32768
    g = ftdx.CutlassComplexGemm()
    for swap in (False, True):
        X = np.zeros((B, 2, 1, 2), dtype=np.float16)
        X[:, 0, 0, 0] = 1.0
        a, b = (1, 0) if swap else (0, 1)
        X[:, 1, 0, 0] = vals[a::2]
        X[:, 1, 0, 1] = vals[b::2]
        C = torch.empty((B, 3, 2), dtype=torch.float32, device=dev)
        assert g.herk(torch.from_numpy(X.reshape(-1)).to(dev), C, 2, 1, B, ftdx.InputPrecision.FP16,
                      ftdx.ComputePrecision.FP8, ftdx.OutputPrecision.FP32, tune=False) == 0
        torch.cuda.synchronize()
        got = C.cpu().numpy()[:, 1, :]
        want = np.stack([q[a::2], q[b::2]], axis=-1)
        # values, not bits: a -0 product added to the +0 accumulator is +0
        assert np.array_equal(got, want), swap


@pytest.mark.parametrize("N,K,B", [(128, 1024, 4), (100, 70, 2)])
def test_force_baseline_equals_fused(oracle, dev, N, K, B):
    import torch
    from ftd_amd import ftdx
    qc = oracle.gen_bytes(N + K, N * K * B)
    A = torch.from_numpy(qc).to(dev)
    g = ftdx.CutlassComplexGemm()
    g.set_herk_mode(ftdx.HerkMode.ForceBaseline)
    g.init_herk(N, K, B)
    C = torch.empty((B, N * (N + 1) // 2, 2), dtype=torch.float32, device=dev)
    assert g.herk(A, C, N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                  ftdx.OutputPrecision.FP32, tune=False) == 0
    torch.cuda.synchronize()
    assert np.array_equal(_bits(C.cpu().numpy()), _bits(oracle.herk_f32(qc, N, K, B)))
    g.end_herk()


@pytest.mark.parametrize("mode", ["ForceDirect", "ForceBaseline", "ForceDirectBF6"])
def test_graph_replay_lru(oracle, dev, mode):
    """set_direct_herk_graph: the first call of a shape/pointer set is captured,
    later ones replay the graph (cutlass_gemm_api.cu:1738-1807); five distinct
    calls overflow the four-entry cache."""
    import torch
    from ftd_amd import ftdx
    g = ftdx.CutlassComplexGemm()
    g.set_herk_mode(getattr(ftdx.HerkMode, mode.replace("BF6", "")))
    if mode.endswith("BF6"):
        g.set_launch_config(ftdx.TILE_128x128_K64 | ftdx.TILE_BF6, 1)
    g.set_direct_herk_graph(True)
    N, K, B = 96, 128, 2
    g.init_herk(N, K, B)
    cases = []
    for i in range(5):
        qc = oracle.gen_qc_uniform(500 + i, N * K * B)
        A = torch.from_numpy(qc).to(dev)
        C = torch.empty((B, N * (N + 1) // 2, 2), dtype=torch.float32, device=dev)
        cases.append((qc, A, C))
    for rep in range(3):
        for qc, A, C in cases:
            C.fill_(float("nan"))
            torch.cuda.synchronize()
            assert g.herk(A, C, N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                          ftdx.OutputPrecision.FP32, alpha=0.5, tune=False) == 0
            torch.cuda.synchronize()
            ref = oracle.herk_f32(qc, N, K, B, alpha=0.5)
            assert np.array_equal(_bits(C.cpu().numpy()), _bits(ref)), (rep, mode)
    g.end_herk()


def test_graph_replay_follows_launch_config(oracle, dev):
    """A graph captured under one tile configuration is not replayed after the launch
    configuration changes: the call with the new configuration is captured anew, and
    both give the oracle's bytes."""
    import torch
    from ftd_amd import ftdx
    g = ftdx.CutlassComplexGemm()
    g.set_herk_mode(ftdx.HerkMode.ForceDirect)
    g.set_direct_herk_graph(True)
    N, K, B = 130, 192, 2
    qc = oracle.gen_qc_uniform(77, N * K * B)
    A = torch.from_numpy(qc).to(dev)
    C = torch.empty((B, N * (N + 1) // 2, 2), dtype=torch.float32, device=dev)
    ref = oracle.herk_f32(qc, N, K, B)
    seen = []
    for cfg, remap in ((0, 1), (0, 1), (3, 1), (3, 1), (3, 0), (0, 1)):
        g.set_launch_config(cfg, remap)
        C.fill_(float("nan"))
        assert g.herk(A, C, N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                      ftdx.OutputPrecision.FP32, tune=False) == 0
        torch.cuda.synchronize()
        assert np.array_equal(_bits(C.cpu().numpy()), _bits(ref)), (cfg, remap)
        seen.append(g.graph_stats()[1])
    # captures: (0,1) new, replay, (3,1) new, replay, (3,0) new, (0,1) replayed from the cache
    assert seen == [1, 1, 2, 2, 3, 3], seen
    g.end_herk()


@pytest.mark.parametrize("mode", ["ForceBaseline", "ForceDirectBF6"])
def test_graph_replay_of_chunked_precast(dev, mode):
    """A captured call whose pre-cast path runs in several chunks (2048 inputs x 90 channels x
    K=1024: 4 MB of rows per channel, at most 320 MB per chunk) replays the whole sequence
    of conversion and tile launches; init_herk sizes the scratch for one chunk, so the
    captured pointers stay valid.  Every replay equals the fused INT4 output."""
    import torch
    from ftd_amd import ftdx, selfcheck
    N, K, B = 2048, 1024, 90
    A = selfcheck.qc_uniform(N * K * B, 13, dev)
    P = N * (N + 1) // 2
    ref = torch.empty((B, P, 2), dtype=torch.float32, device=dev)
    f = ftdx.CutlassComplexGemm()
    f.set_launch_config(5, 1)
    assert f.herk(A, ref, N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                  ftdx.OutputPrecision.FP32, tune=False) == 0
    f.close()
    g = ftdx.CutlassComplexGemm()
    g.set_herk_mode(getattr(ftdx.HerkMode, mode.replace("BF6", "")))
    g.set_launch_config(5 | (ftdx.TILE_BF6 if mode.endswith("BF6") else 0), 1)
    g.set_direct_herk_graph(True)
    g.init_herk(N, K, B)
    C = torch.empty_like(ref)
    for rep in range(3):
        C.fill_(float("nan"))
        torch.cuda.synchronize()
        assert g.herk(A, C, N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                      ftdx.OutputPrecision.FP32, tune=False) == 0
        torch.cuda.synchronize()
        assert torch.equal(C.view(torch.int32), ref.view(torch.int32)), (mode, rep)
    cached, captures = g.graph_stats()
    assert captures == 1, (cached, captures)
    g.end_herk()
    g.close()
