"""GPU parity of every compiled tile configuration and workgroup order, of the
grouped persistent launch (with and without hipGraph replay) and of the
autotuner path; all bit-exact against the CPU oracle."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint16)


@pytest.mark.parametrize("cfg", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("remap", [0, 1])
@pytest.mark.parametrize("N,K,B", [(1, 1, 2), (63, 70, 3), (129, 64, 2), (257, 100, 1),
                                   (300, 256, 2), (520, 128, 1),
                                   (97, 150, 2), (64, 320, 1)])  # 3 and 5 stages of 64
@pytest.mark.parametrize("mode", ["fused", "precast", "bf6"])
def test_tile_configs_bit_exact(oracle, dev, cfg, remap, N, K, B, mode):
    """mode "precast": FP8 rows converted first (HerkMode.ForceBaseline), whose
    interior tiles are loaded global -> LDS directly; mode "bf6": rows converted to
    E3M2 and the MFMAs run in E3M2 (FTDX_TILE_BF6)."""
    import torch
    from ftd_amd import ftdx
    g = ftdx.CutlassComplexGemm()
    g.set_launch_config(cfg | (ftdx.TILE_BF6 if mode == "bf6" else 0), remap)
    if mode == "precast":
        g.set_herk_mode(ftdx.HerkMode.ForceBaseline)
    A_np = oracle.gen_qc_uniform(31 * N + K + cfg, N * K * B)
    A = torch.from_numpy(A_np).to(dev)
    C = torch.full((B, N * (N + 1) // 2, 2), float("nan"), dtype=torch.float32, device=dev)
    assert g.herk(A, C, N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                  ftdx.OutputPrecision.FP32, tune=False) == 0
    torch.cuda.synchronize()
    ref = oracle.herk_f32(A_np, N, K, B)
    assert np.array_equal(_bits(C.cpu().numpy()), _bits(ref))


@pytest.mark.parametrize("cfg", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("N,K,B", [(200, 96, 2), (300, 256, 2)])  # edge loop; interior tiles
@pytest.mark.parametrize("mode", ["fused", "precast", "bf6"])
def test_tile_configs_fp16_beta(oracle, dev, cfg, N, K, B, mode):
    import torch
    from ftd_amd import ftdx
    g = ftdx.CutlassComplexGemm()
    g.set_launch_config(cfg | (ftdx.TILE_BF6 if mode == "bf6" else 0), 1)
    if mode == "precast":
        g.set_herk_mode(ftdx.HerkMode.ForceBaseline)
    A_np = oracle.gen_qc_uniform(5 + cfg, N * K * B)
    rng = np.random.default_rng(cfg)
    C0 = rng.integers(-100, 100, size=(B, N * (N + 1) // 2, 2)).astype(np.float16)
    A = torch.from_numpy(A_np).to(dev)
    C = torch.from_numpy(C0).to(dev)
    assert g.herk(A, C, N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                  ftdx.OutputPrecision.FP16, alpha=0.5, beta=2.0, tune=False) == 0
    torch.cuda.synchronize()
    ref = oracle.herk_f16(A_np, N, K, B, alpha=0.5, beta=2.0, C=C0)
    assert np.array_equal(_bits(C.cpu().numpy()), ref)


def _grouped_case(oracle, dev, shapes, seed):
    import torch
    probs, refs, inputs = [], [], []
    for i, (N, K) in enumerate(shapes):
        A_np = oracle.gen_qc_uniform(seed * 1000 + i, N * K)
        A = torch.from_numpy(A_np).to(dev)
        C = torch.full((N * (N + 1) // 2, 2), float("nan"), dtype=torch.float32, device=dev)
        probs.append((A, C, N, K, 1.0, 0.0))
        inputs.append(A_np)
        refs.append(oracle.herk_f32(A_np, N, K, 1)[0])
    return probs, refs


# "mixed": some K not a multiple of 16 (both load paths in one kernel); "vec": every
# problem takes 16-byte loads (the 16-byte-only kernel at the tile's own prefetch depth),
# with partial tiles and one-stage problems
GROUPED_SHAPES = {
    "mixed": [(256, 320), (64, 64), (130, 1000), (512, 256), (33, 17), (1, 5), (300, 64),
              (128, 4096), (200, 2048), (96, 100)],
    "vec": [(256, 320), (130, 1024), (512, 256), (300, 64), (128, 4096), (200, 2048), (96, 112),
            (1, 16), (33, 48), (640, 192)],
}


@pytest.mark.parametrize("cfg", [-1, 0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("shapes", ["mixed", "vec"])
def test_grouped_uneven_bit_exact(oracle, dev, cfg, graph, shapes):
    import torch
    from ftd_amd import ftdx
    probs, refs = _grouped_case(oracle, dev, GROUPED_SHAPES[shapes], 7)
    ctx = ftdx.CutlassComplexGemm()
    plan = ftdx.GroupedPlan(ctx, probs, tile_config=cfg, use_graph=graph)
    info = plan.info()
    assert info["items"] > 0 and info["workgroups"] > 0
    for rep in range(3):  # replays give the same bytes
        for p in probs:
            p[1].fill_(float("nan"))
        plan.run()
        torch.cuda.synchronize()
        for (A, C, N, K, _, _), ref in zip(probs, refs):
            assert np.array_equal(_bits(C.cpu().numpy()), _bits(ref)), (rep, N, K)
    plan.close()


def test_grouped_alpha_beta_and_unaligned(oracle, dev):
    import torch
    from ftd_amd import ftdx
    probs, refs = [], []
    rng = np.random.default_rng(2)
    for i, (N, K, alpha, beta, off) in enumerate([(100, 64, 0.5, 1.0, 0), (70, 33, -1.0, 0.25, 3),
                                                  (130, 48, 2.0, 0.0, 1)]):
        A_np = oracle.gen_qc_uniform(900 + i, N * K)
        buf = torch.zeros(N * K + 16, dtype=torch.uint8, device=dev)
        A = buf[off:off + N * K]
        A.copy_(torch.from_numpy(A_np))
        C0 = rng.integers(-50, 50, size=(N * (N + 1) // 2, 2)).astype(np.float32)
        C = torch.from_numpy(C0).to(dev)
        probs.append((A, C, N, K, alpha, beta))
        refs.append(oracle.herk_f32(A_np, N, K, 1, alpha=alpha, beta=beta, C=C0[None])[0])
    ctx = ftdx.CutlassComplexGemm()
    plan = ftdx.GroupedPlan(ctx, probs, tile_config=-1, use_graph=True)
    plan.run()
    torch.cuda.synchronize()
    for (A, C, N, K, a, b), ref in zip(probs, refs):
        assert np.array_equal(_bits(C.cpu().numpy()), _bits(ref)), (N, K, a, b)


@pytest.mark.parametrize("tile_config", [3, 5])
def test_grouped_ragged_rows_padded_by_plan(oracle, dev, tile_config):
    """Problems whose rows 16-byte loads cannot take (K % 16 != 0 or A unaligned, here most
    of them) run on zero-padded copies refreshed by the plan's pad launch: several pad
    workgroups per problem (N * K up to 600 x 700), empty problems between them, and the
    input changed between replays (the copy is taken on every run).  Bit-exact against the
    oracle on every replay."""
    import torch
    from ftd_amd import ftdx
    rng = np.random.default_rng(40 + tile_config)
    shapes = [(int(rng.integers(1, 601)), int(rng.integers(1, 701))) for _ in range(14)]
    shapes += [(0, 30), (50, 0), (600, 700), (257, 16), (33, 1)]
    probs, host = [], []
    for i, (N, K) in enumerate(shapes):
        off = int(rng.integers(0, 16))
        buf = torch.zeros(N * K + 32, dtype=torch.uint8, device=dev)
        A = buf[off:off + N * K]
        C = torch.full((max(N * (N + 1) // 2, 1), 2), float("nan"), dtype=torch.float32, device=dev)
        probs.append((A, C, N, K))
        host.append(None)
    ctx = ftdx.CutlassComplexGemm()
    plan = ftdx.GroupedPlan(ctx, probs, tile_config=tile_config, use_graph=True)
    for rep in range(2):
        for i, (A, C, N, K) in enumerate(probs):
            host[i] = oracle.gen_qc_uniform(7000 + 31 * rep + i, N * K)
            if N * K:
                A.copy_(torch.from_numpy(host[i]))
            C.fill_(float("nan"))
        torch.cuda.synchronize()
        print(f"grouped ragged cfg={tile_config} replay {rep}", flush=True)
        plan.run()
        torch.cuda.synchronize()
        for (A, C, N, K), A_np in zip(probs, host):
            if N == 0:
                continue
            ref = oracle.herk_f32(A_np, N, K, 1)[0]
            assert np.array_equal(_bits(C.cpu().numpy()), _bits(ref)), (rep, N, K)
    plan.close()
    ctx.close()


def test_grouped_matches_plain_loop_bytes(oracle, dev):
    """A reduced batched-uneven set: grouped output bytes == per-problem plain loop."""
    import torch
    from ftd_amd import ftdx
    rng = np.random.default_rng(3)
    Ns = rng.choice([256, 512, 1024], size=24)
    Ks = rng.integers(4, 65, size=24) * 64
    plain, grouped, probs = [], [], []
    ctx = ftdx.CutlassComplexGemm()
    for i, (N, K) in enumerate(zip(Ns.tolist(), Ks.tolist())):
        A = torch.from_numpy(oracle.gen_qc_uniform(5000 + i, N * K)).to(dev)
        Cp = torch.empty((N * (N + 1) // 2, 2), dtype=torch.float32, device=dev)
        Cg = torch.full_like(Cp, float("nan"))
        assert ctx.herk(A, Cp, N, K, 1, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                        ftdx.OutputPrecision.FP32, tune=False) == 0
        plain.append(Cp)
        grouped.append(Cg)
        probs.append((A, Cg, N, K))
    plan = ftdx.GroupedPlan(ctx, probs, use_graph=True)
    plan.run()
    torch.cuda.synchronize()
    for Cp, Cg in zip(plain, grouped):
        assert torch.equal(Cp.view(torch.int32), Cg.view(torch.int32))


@pytest.mark.parametrize("tile_config", ["auto", 4])
def test_grouped_full_size_workload(oracle, dev, tile_config):
    """The bench's batched workload itself: 512 problems (seed 3, N ~ U{256..4096},
    K ~ U{256..4096}) on the bench's input, one grouped hipGraph launch.  Every problem
    is checked by the exact weighted checksum; rows of three N = 4096 problems are
    compared with the oracle."""
    import torch
    from ftd_amd import ftdx, selfcheck
    Ns, Ks, a_off, c_off, oa, oc = selfcheck.uneven_problem_list(512, 3)
    A = selfcheck.qc_uniform(oa, 3, dev)
    C = torch.full((oc,), float("nan"), dtype=torch.float32, device=dev)
    probs = [(A[a:a + N * K], C[c:c + N * (N + 1)], N, K) for a, c, N, K in zip(a_off, c_off, Ns, Ks)]
    ctx = ftdx.CutlassComplexGemm()
    ctx.set_strategy_tune_verbosity(0)
    cfg = ftdx.TILE_AUTO if tile_config == "auto" else tile_config
    plan = ftdx.GroupedPlan(ctx, probs, tile_config=cfg, use_graph=True)
    plan.run()
    plan.run()  # graph replay: the same bytes again
    torch.cuda.synchronize()
    res = selfcheck.check_problem_list(torch, A, C, Ns, Ks, a_off, c_off)
    assert res["ok"], res
    big = [p for p, N in enumerate(Ns) if N == 4096][:3]
    assert len(big) == 3
    rows = [0, 1, 127, 128, 255, 256, 2047, 4095]
    for p in big:
        N, K = Ns[p], Ks[p]
        Ah = A[a_off[p]:a_off[p] + N * K].cpu().numpy()
        ref = np.zeros((1, N * (N + 1) // 2, 2), dtype=np.int64)
        for r in rows:
            oracle.herk_int64_rows(Ah, N, K, 1, r, r + 1, out=ref)
        got = C[c_off[p]:c_off[p] + N * (N + 1)].view(-1, 2)
        for r in rows:
            s, e = r * (r + 1) // 2, (r + 1) * (r + 2) // 2
            assert np.array_equal(got[s:e].cpu().numpy(), ref[0, s:e].astype(np.float32)), (p, r)
    plan.close()


def test_autotune_and_cache_file(oracle, dev, tmp_path):
    import torch
    from ftd_amd import ftdx
    N, K, B = 256, 512, 3
    A_np = oracle.gen_qc_uniform(77, N * K * B)
    A = torch.from_numpy(A_np).to(dev)
    C = torch.empty((B, N * (N + 1) // 2, 2), dtype=torch.float32, device=dev)
    path = str(tmp_path / "tune.txt")
    g = ftdx.CutlassComplexGemm()
    g.set_strategy_tune_verbosity(0)
    g.set_tune_cache_path(path)
    assert g.herk(A, C, N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                  ftdx.OutputPrecision.FP32, tune=True) == 0
    torch.cuda.synchronize()
    assert np.array_equal(_bits(C.cpu().numpy()), _bits(oracle.herk_f32(A_np, N, K, B)))
    text = open(path).read()
    assert text.startswith("# ftdx-tune ") and "\nherk\t256\t512\t3\t0\t1\t" in text
    cfg, remap = g.query_config(N, K, B)
    g2 = ftdx.CutlassComplexGemm()
    n, sk = g2.tune_load(path)
    assert n == 1 and sk == 0
    assert g2.query_config(N, K, B) == (cfg, remap)
    # beta != 0 with tuning: the caller's C is accumulated exactly once
    C0 = np.ones((B, N * (N + 1) // 2, 2), dtype=np.float32)
    C2 = torch.from_numpy(C0).to(dev)
    g3 = ftdx.CutlassComplexGemm()
    g3.set_strategy_tune_verbosity(0)
    assert g3.herk(A, C2, N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                   ftdx.OutputPrecision.FP32, alpha=1.0, beta=1.0, tune=True) == 0
    torch.cuda.synchronize()
    ref = oracle.herk_f32(A_np, N, K, B, alpha=1.0, beta=1.0, C=C0)
    assert np.array_equal(_bits(C2.cpu().numpy()), _bits(ref))


@pytest.mark.parametrize("N,K,B", [(300, 256, 2), (129, 70, 3)])
def test_bf6_signed_zero_and_extremes(oracle, dev, N, K, B):
    """E3M2 operands on data heavy in -0 nibbles (0x8 / 0x80) and +-7: every sign and
    magnitude code, bit-exact against the oracle (FP32 sums stay exact)."""
    import torch
    from ftd_amd import ftdx
    rng = np.random.default_rng(N + K)
    nib = rng.choice(np.array([0x0, 0x8, 0x7, 0xF, 0x1, 0x9, 0x5, 0xD], np.uint8), size=(2, N * K * B))
    A_np = ((nib[0] << 4) | nib[1]).astype(np.uint8)
    g = ftdx.CutlassComplexGemm()
    g.set_launch_config(ftdx.TILE_128x128_K64_P2 | ftdx.TILE_BF6, 1)
    C = torch.full((B, N * (N + 1) // 2, 2), float("nan"), dtype=torch.float32, device=dev)
    assert g.herk(torch.from_numpy(A_np).to(dev), C, N, K, B, ftdx.InputPrecision.INT4,
                  ftdx.ComputePrecision.FP8, ftdx.OutputPrecision.FP32, tune=False) == 0
    torch.cuda.synchronize()
    assert np.array_equal(_bits(C.cpu().numpy()), _bits(oracle.herk_f32(A_np, N, K, B)))
