"""Benchmark of the X-engine HERK on one MI355X.

Headline workload (one "step" = one plain HERK call over all channels):
2048 dual-pol antennas (4096 inputs), 96 channels, K = 2048 time samples;
int4 complex voltages uniform in [-7, 7] (synthetic, seeded), promoted to OCP
FP8 E4M3 inside the kernel, FP32 accumulate, packed lower-triangle complex FP32
output.  Metric: 4M-equivalent TFLOP/s = 8 N^2 K batch / t (the reference's
convention, herk_vs_tcc.sh:141, cutlass_interface/test_herk_int4.cu:97) and its
fraction of the MI355X FP8 dense MFMA peak.

Prints ONE JSON line.  The CPU baseline is the oracle (C, OpenMP, int64
accumulation) timed on a bounded slice of the same workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FP8_DENSE_PEAK_TFLOPS = 5033.0   # 256 CU x 2.4 GHz x 8192 FLOP/CU/clk (AMD published, dense)
HBM_PEAK_GBS = 8000.0
MX_DENSE_PEAK_TFLOPS = 10066.0   # MXFP6/MXFP4 dense (AMD published), 2x the FP8 rate
METRIC = "HERK TFLOP/s (4M-equiv) & % FP8 MFMA peak, 4096 inputs; batched vs plain"
# HBM bytes per launch of the headline kernel (4096 inputs, 96 channels, K = 2048, FP32
# output), per tile configuration, from rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (the
# factors a 4096 MiB device copy gives those counters); summary and commands under profiles/.
MEASURED_TRAFFIC_BYTES = {5: 11.631e9, 4: 15.518e9}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def synth_qc(torch, n: int, seed: int, device) -> "torch.Tensor":
    """QC bytes with Re, Im i.i.d. uniform integers in [-7, 7] (sign-magnitude nibbles);
    the same generator the full-size parity tests use, so a seed names one input."""
    from ftd_amd import selfcheck
    return selfcheck.qc_uniform(n, seed, device)


# Tile geometry per tile configuration id: (TM, TN, WAVES_M, WAVES_N, KC)
TILE_GEOMETRY = {0: (128, 128, 2, 2, 64), 1: (128, 128, 2, 2, 128), 2: (64, 64, 1, 1, 64),
                 3: (256, 128, 4, 2, 64), 4: (128, 128, 2, 2, 64), 5: (256, 128, 4, 2, 64)}


def executed_mfma_flop(N: int, K: int, batch: int, cfg: int) -> float:
    """MFMA FLOP the plain launch executes: every wave of a lower-triangle tile that
    holds an element i >= j runs its full WM x WN block over K rounded up to the stage
    (8 real FLOP per complex MAC, counted the way the MFMA counts them)."""
    TM, TN, WVM, WVN, KC = TILE_GEOMETRY[cfg & 15]
    WM, WN = TM // WVM, TN // WVN
    nbM, nbN, r = -(-N // TM), -(-N // TN), TM // TN
    waves = 0
    for I in range(nbM):
        for J in range(min(nbN, r * (I + 1))):
            for wm in range(WVM):
                for wn in range(WVN):
                    w_r, w_c = I * TM + wm * WM, J * TN + wn * WN
                    if w_r < N and w_r + WM - 1 >= w_c:
                        waves += 1
    return 8.0 * WM * WN * (-(-K // KC) * KC) * waves * batch


def batched_vs_plain(torch, ftdx, dev, count: int, seed: int, reps: int) -> dict:
    """Batched uneven HERK: `count` problems, N ~ U{256..4096 (powers of 2)},
    K ~ U{256, 320, ..., 4096}; one grouped persistent launch (hipGraph replay,
    autotuned tile configuration) against a loop of one plain launch per problem on
    one stream, the loop timed three ways: the library default tile, the grouped
    plan's tile, and each shape's own autotuned tile (the reference's per-call
    tune=true).  Every output is compared byte for byte with the grouped one, and the
    grouped output is checked exactly (weighted checksum of every problem, exact rows
    of the largest problems)."""
    from ftd_amd import selfcheck
    Ns, Ks, a_off, c_off, oa, oc = selfcheck.uneven_problem_list(count, seed)
    A = synth_qc(torch, oa, seed, dev)
    Cp = torch.empty(oc, dtype=torch.float32, device=dev)
    Cg = torch.empty(oc, dtype=torch.float32, device=dev)
    pp = [(A[a:a + N * K], Cp[c:c + N * (N + 1)], N, K) for a, c, N, K in zip(a_off, c_off, Ns, Ks)]
    pg = [(A[a:a + N * K], Cg[c:c + N * (N + 1)], N, K) for a, c, N, K in zip(a_off, c_off, Ns, Ks)]
    flop = sum(8.0 * N * N * K for N, K in zip(Ns, Ks))
    ctx = ftdx.CutlassComplexGemm()
    ctx.set_strategy_tune_verbosity(0)
    arr = ftdx.problem_array(pp)
    t_plan = time.perf_counter()
    plan = ftdx.GroupedPlan(ctx, pg, tile_config=ftdx.TILE_AUTO, use_graph=True)
    t_plan = time.perf_counter() - t_plan
    info = plan.info()
    stream = torch.cuda.current_stream()
    plan.run(stream)
    torch.cuda.synchronize()
    big = [p for p, N in enumerate(Ns) if N == max(Ns)][:3]
    parity = selfcheck.check_problem_list(
        torch, A, Cg, Ns, Ks, a_off, c_off,
        rows_of=[(p, [0, 1, 255, 256, Ns[p] // 2, Ns[p] - 1]) for p in big])

    # plain-loop contexts: default tile (no tuning), the grouped plan's tile, per-shape tuned
    loops = {}
    c_def = ftdx.CutlassComplexGemm()
    c_def.set_strategy_tune_verbosity(0)
    loops["default_tile"] = c_def
    c_grp = ftdx.CutlassComplexGemm()
    c_grp.set_strategy_tune_verbosity(0)
    c_grp.set_launch_config(info["tile_config"], -1)
    loops["grouped_tile"] = c_grp
    c_tun = ftdx.CutlassComplexGemm()
    c_tun.set_strategy_tune_verbosity(0)
    t_tune = time.perf_counter()
    for (a, c, N, K) in {(N, K): q for q in pp for N, K in [(q[2], q[3])]}.values():
        c_tun.herk(a, c, N, K, 1, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                   ftdx.OutputPrecision.FP32, stream=stream, tune=True)
    torch.cuda.synchronize()
    t_tune = time.perf_counter() - t_tune
    loops["per_shape_tuned"] = c_tun
    identical = {}
    for name, c in loops.items():
        Cp.fill_(-1.0)
        ftdx.herk_each(c, pp, stream=stream, arr=arr)
        torch.cuda.synchronize()
        identical[name] = bool(torch.equal(Cp.view(torch.int32), Cg.view(torch.int32)))

    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn):
        e0.record(stream)
        fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1)

    def replays(n):
        for _ in range(n):
            plan.run(stream)

    # grouped: per replay over three back-to-back replays (the plan replayed once per
    # integration, as the headline's steps are), and the first replay after the loops
    tl = {name: [] for name in loops}
    tg, tg1 = [], []
    for _ in range(reps):
        for name, c in loops.items():
            tl[name].append(timed(lambda: ftdx.herk_each(c, pp, stream=stream, arr=arr)))
        tg1.append(timed(lambda: plan.run(stream)))
        tg.append(timed(lambda: replays(3)) / 3)
    med = lambda v: sorted(v)[len(v) // 2]  # noqa: E731
    ms_g = med(tg)
    ms_l = {name: med(v) for name, v in tl.items()}
    plan.close()
    for c in loops.values():
        c.close()
    return {
        "problems": count, "seed": seed, "flop_4m_equiv": flop,
        "grouped_ms": round(ms_g, 4),
        "grouped_tflops": round(flop / (ms_g * 1e-3) / 1e12, 2),
        "grouped_first_replay_ms": round(med(tg1), 4),
        "plain_loop_ms": {k: round(v, 4) for k, v in ms_l.items()},
        "plain_loop_tflops": {k: round(flop / (v * 1e-3) / 1e12, 2) for k, v in ms_l.items()},
        "speedup": {k: round(v / ms_g, 3) for k, v in ms_l.items()},
        "identical_output": identical,
        "grouped_parity": parity,
        "tile_config": info["tile_config"], "workgroups": info["workgroups"],
        "items": info["items"], "schedule_imbalance": round(info["imbalance"], 4),
        "plan_s": round(t_plan, 3), "per_shape_tune_s": round(t_tune, 2), "graph": True,
    }


def mx_study(torch, ftdx, dev, N: int, K: int, B: int, seed: int, reps: int) -> dict:
    """MX block-scaled HERK at 4096 inputs x 32 channels x K=4096: complex
    Gaussian voltages (E|x|^2 = 1) with 1% impulsive 30-sigma outliers of uniform
    phase, as FP16; FP6 E3M2 and FP4 E2M1 with 32-sample E8M0 scales, under the
    reference's element converters (the default) and under OCP nearest-even rounding.
    Error per channel vs the FP64 HERK of the same FP16 voltages (complex128 matmul)."""
    g = torch.Generator(device=dev).manual_seed(seed)
    x = torch.complex(torch.randn((B, N, K), generator=g, device=dev),
                      torch.randn((B, N, K), generator=g, device=dev)) * (0.5 ** 0.5)
    mask = torch.rand((B, N, K), generator=g, device=dev) < 0.01
    phase = (torch.rand((B, N, K), generator=g, device=dev) * 2 - 1) * torch.pi
    x = torch.where(mask, 30.0 * torch.exp(1j * phase), x)
    X16 = torch.view_as_real(x).to(torch.float16).contiguous()
    del x, mask, phase
    ii, jj = torch.tril_indices(N, N, device=dev)
    ctx = ftdx.CutlassComplexGemm()
    flop = 8.0 * N * N * K * B
    out = {"inputs": N, "channels": B, "K": K, "seed": seed,
           "data": "complex Gaussian sigma=1 + 1% 30-sigma outliers, FP16",
           "rounding": "reference converters (default); nearest_even entries: OCP rounding"}
    runs = {}
    for name, comp in (("fp6_e3m2", ftdx.ComputePrecision.FP6), ("fp4_e2m1", ftdx.ComputePrecision.FP4)):
        for rname, rnd in (("reference", ftdx.MxRounding.Reference), ("nearest_even", ftdx.MxRounding.NearestEven)):
            ctx.set_mx_rounding(rnd)
            C = torch.empty((B, N * (N + 1) // 2, 2), dtype=torch.float32, device=dev)
            ftdx.herk_mx(ctx, X16, N, K, B, ftdx.InputPrecision.FP16, comp, output="fp32", C=C)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                ftdx.herk_mx(ctx, X16, N, K, B, ftdx.InputPrecision.FP16, comp, output="fp32", C=C)
            e1.record()
            e1.synchronize()
            runs[(name, rname)] = (e0.elapsed_time(e1) / reps, C, [], [])
    for b in range(B):
        xb = X16[b].double()
        xc = torch.complex(xb[..., 0], xb[..., 1])
        R = (xc @ xc.conj().transpose(0, 1).contiguous())[ii, jj]
        R = torch.view_as_real(R).clone()
        R[ii == jj, 1] = 0.0
        for ms, C, fro, mxe in runs.values():
            d = C[b].double() - R
            fro.append(float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(R)))
            mxe.append(float(d.abs().max() / R.abs().max()))
            del d
        del xb, xc, R
    for (name, rname), (ms, C, fro, mxe) in runs.items():
        rec = {"ms": round(ms, 4), "tflops_4m_equiv": round(flop / (ms * 1e-3) / 1e12, 2),
               "pct_mx_dense_peak": round(100 * flop / (ms * 1e-3) / 1e12 / MX_DENSE_PEAK_TFLOPS, 2),
               "rel_frobenius": {"min": min(fro), "median": sorted(fro)[B // 2], "max": max(fro)},
               "rel_max_element": {"min": min(mxe), "median": sorted(mxe)[B // 2], "max": max(mxe)}}
        if rname == "reference":
            out[name] = rec
        else:
            out[name]["nearest_even"] = rec
    del runs
    torch.cuda.empty_cache()
    return out


def calibration_study(torch, ftdx, dev, N: int, K: int, B: int, seed: int, reps: int) -> dict:
    """Differentiable gain calibration at 1024 inputs x 16 channels x K=2048:
    forward HERK (FP8, exact), gain loss + gradients, HERK adjoint dX = (G+G^H) X
    (FP32 MFMA), timed; errors vs an FP64 reference computed on the device."""
    import numpy as np
    gen = torch.Generator(device=dev).manual_seed(seed)
    v = torch.randint(-7, 8, (2, B * N * K), generator=gen, device=dev, dtype=torch.int16)
    nib = torch.where(v < 0, 8 + (-v), v).to(torch.uint8)
    A = ((nib[0] << 4) | nib[1]).contiguous()
    Xc = torch.complex(v[0].double(), v[1].double()).reshape(B, N, K)
    del v, nib
    ctx = ftdx.CutlassComplexGemm()
    M = ftdx.herk_int4(A, N, K, B)
    ii, jj = torch.tril_indices(N, N, device=dev)
    Mc = torch.view_as_complex(M.double())
    rng = np.random.default_rng(seed)

    def gains():
        return (1 + 0.1 * rng.standard_normal((B, N))) * np.exp(1j * np.pi * rng.uniform(-1, 1, (B, N)))

    gt, g = torch.from_numpy(gains()).to(dev), torch.from_numpy(gains()).to(dev)
    noise = torch.complex(torch.randn(Mc.shape, generator=gen, device=dev, dtype=torch.float64),
                          torch.randn(Mc.shape, generator=gen, device=dev, dtype=torch.float64))
    V = gt[:, ii] * gt[:, jj].conj() * Mc + 50.0 * noise
    V[:, ii == jj] = V[:, ii == jj].real.to(V.dtype)
    V32 = torch.view_as_real(V).float().contiguous()
    g32 = torch.view_as_real(g).float().contiguous()
    V, g = torch.view_as_complex(V32.double()), torch.view_as_complex(g32.double())
    loss, grad, Gamma = ftdx.gain_loss_grad(ctx, V32, M, g32, N, B)
    dX = ftdx.herk_adjoint(ctx, Gamma, A, N, K, B)
    torch.cuda.synchronize()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    e[0].record()
    for _ in range(reps):
        ftdx.herk_int4(A, N, K, B, C=M)
    e[1].record()
    for _ in range(reps):
        ftdx.gain_loss_grad(ctx, V32, M, g32, N, B)
    e[2].record()
    for _ in range(reps):
        ftdx.herk_adjoint(ctx, Gamma, A, N, K, B)
    e[3].record()
    e[3].synchronize()
    t_fwd, t_gain, t_adj = (e[i].elapsed_time(e[i + 1]) / reps for i in range(3))
    R = V - g[:, ii] * g[:, jj].conj() * Mc
    G64 = -2.0 * R * g[:, ii].conj() * g[:, jj]
    acc = torch.zeros((B, N, 2), dtype=torch.float64, device=dev)
    acc.index_add_(1, ii, torch.view_as_real(-2.0 * R * Mc.conj() * g[:, jj]))
    acc.index_add_(1, jj, torch.view_as_real(-2.0 * R.conj() * Mc * g[:, ii]))
    grad64 = torch.view_as_complex(acc)
    H = torch.zeros((B, N, N), dtype=torch.complex128, device=dev)
    H[:, ii, jj] = G64
    H = H + H.conj().transpose(1, 2)
    dX64 = torch.matmul(H, Xc)
    h = 1e-4
    fd_err = 0.0
    for k in (0, N // 2, N - 1):
        for unit in (1.0, 1j):
            def Lg(delta):
                gg = g[0].clone()
                gg[k] += delta
                r = V[0] - gg[ii] * gg[jj].conj() * Mc[0]
                return float((r.abs() ** 2).sum())
            fd = (Lg(h * unit) - Lg(-h * unit)) / (2 * h)
            an = float(grad[0, k, 0] if unit == 1.0 else grad[0, k, 1])
            fd_err = max(fd_err, abs(fd - an) / float(grad64.abs().max()))
    adj_flop = 8.0 * N * N * K * B
    return {
        "inputs": N, "channels": B, "K": K, "seed": seed,
        "forward_herk_ms": round(t_fwd, 4), "gain_loss_grad_ms": round(t_gain, 4),
        "herk_adjoint_ms": round(t_adj, 4),
        "herk_adjoint_tflops": round(adj_flop / (t_adj * 1e-3) / 1e12, 2),
        "loss_rel_err_vs_fp64": float(((loss - (R.abs() ** 2).sum(-1)).abs()
                                       / (R.abs() ** 2).sum(-1)).max()),
        "grad_g_max_rel_err_vs_fp64": float((torch.view_as_complex(grad.double()) - grad64).abs().max()
                                            / grad64.abs().max()),
        "dX_max_rel_err_vs_fp64": float((torch.view_as_complex(dX.double()) - dX64).abs().max()
                                        / dX64.abs().max()),
        "grad_g_max_err_vs_fp64_central_diff_h1e-4": fd_err,
    }


def beamformer_study(torch, ftdx, dev, M: int, N: int, K: int, B: int, seed: int, reps: int) -> dict:
    """Voltage-beamformer product with prepared weights: INT4 QC voltages A [B][M][K]
    times conj-prepared FP16 weights B [B][N][K], fused power detection
    (gemm_prepared_power_int4); 4M-equivalent FLOP 8 M N K per batch."""
    g = torch.Generator(device=dev).manual_seed(seed)
    A = synth_qc(torch, M * K * B, seed, dev)
    Wre = torch.randint(-7, 8, (B, N, K), generator=g, device=dev, dtype=torch.int16).to(torch.float16)
    Wim = torch.randint(-7, 8, (B, N, K), generator=g, device=dev, dtype=torch.int16).to(torch.float16)
    P = torch.empty((B, M, N), dtype=torch.float32, device=dev)
    ctx = ftdx.CutlassComplexGemm()
    ctx.prepare_b(Wre, Wim, N, K, B)
    del Wre, Wim
    for _ in range(2):
        ctx.gemm_prepared_power_int4(A, P, M, N, K, B)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        ctx.gemm_prepared_power_int4(A, P, M, N, K, B)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    tf = 8.0 * M * N * K * B / (ms * 1e-3) / 1e12
    return {"M": M, "N_beams": N, "K_antennas": K, "batch": B, "seed": seed,
            "entry": "gemm_prepared_power_int4", "ms": round(ms, 4),
            "tflops_4m_equiv": round(tf, 2),
            "pct_fp8_peak": round(100.0 * tf / FP8_DENSE_PEAK_TFLOPS, 2)}


def voltage_study(torch, ftdx, dev, n_ant: int, n_beam: int, n_ch: int, n_time: int, n_tps: int, seed: int,
                  reps: int) -> dict:
    """Voltage beamformer end to end (ggp::VoltagePipeline::beamform): QC payload -> uint8
    Stokes-I filterbank.  The pipeline (rows pass + one GEMM whose epilogue detects,
    sums the pols, integrates and quantizes) against the reference's call sequence on
    this library: transpose + pol split, gemm_prepared_power_int4 for pol 0 (beta 0) and
    pol 1 (beta 1) into an FP32 power buffer, then integration + quantization.
    4M-equivalent FLOP 8 * (2 n_time) * n_beam * n_ant per channel."""
    qc = synth_qc(torch, n_ch * 2 * n_ant * n_time, seed, dev)
    g = torch.Generator(device=dev).manual_seed(seed)
    w = torch.rand((n_ch, n_beam, n_ant, 2), generator=g, device=dev) * 2 - 1
    vp = ftdx.VoltagePipeline(n_antennae=n_ant, n_beams=n_beam, n_channels=n_ch, n_time=n_time,
                              n_time_power_sum=n_tps, strategy_tune_verbosity=0)
    vp.set_weights(w)
    out = torch.empty((n_ch, n_beam, n_time // n_tps), dtype=torch.uint8, device=dev)
    scale = 1e-5
    ctx = ftdx.CutlassComplexGemm()
    ctx.set_strategy_tune_verbosity(0)
    w16 = w.to(torch.float16)
    ctx.prepare_b(w16[..., 0].contiguous(), w16[..., 1].contiguous(), n_beam, n_ant, n_ch)
    del w16, w
    P = torch.empty((n_ch, n_time, n_beam), dtype=torch.float32, device=dev)
    ref_out = torch.empty_like(out)
    v = qc.view(n_ch, 2, n_ant, n_time)
    rows = [torch.empty((n_ch, n_time, n_ant), dtype=torch.uint8, device=dev) for _ in range(2)]

    def reference_sequence():
        for p in range(2):
            rows[p].copy_(v[:, p].transpose(1, 2))
            ctx.gemm_prepared_power_int4(rows[p], P, n_time, n_beam, n_ant, n_ch, beta=float(p))
        acc = P.view(n_ch, n_time // n_tps, n_tps, n_beam).sum(dim=2)
        ref_out.copy_(torch.round(acc.transpose(1, 2) * scale).clamp_(0, 255).to(torch.uint8))

    def pipeline():
        vp.beamform(out, qc, scale)

    # FP6 mode (the reference's lower-precision option): QC rows as E3M2 operands against the
    # MX E3M2 weights, one GEMM per payload, then the power / pol-sum / integration pass
    vp6 = ftdx.VoltagePipeline(n_antennae=n_ant, n_beams=n_beam, n_channels=n_ch, n_time=n_time,
                               n_time_power_sum=n_tps, strategy_tune_verbosity=0,
                               compute_mode=ftdx.VOLTAGE_FP6)
    vp6.set_weights((torch.rand((n_ch, n_beam, n_ant, 2), generator=g, device=dev) * 2 - 1))
    out6 = torch.empty_like(out)

    def pipeline_fp6():
        vp6.beamform(out6, qc, scale)

    res = {}
    for name, fn in (("reference_sequence_ms", reference_sequence), ("pipeline_ms", pipeline),
                     ("pipeline_fp6_ms", pipeline_fp6)):
        fn()
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        res[name] = round(e0.elapsed_time(e1) / reps, 4)
    flop = 8.0 * 2 * n_time * n_beam * n_ant * n_ch
    diff = (out.int() - ref_out.int()).abs()
    return {"n_antennae": n_ant, "n_beams": n_beam, "n_channels": n_ch, "n_time": n_time,
            "n_time_power_sum": n_tps, "seed": seed, **res,
            "speedup": round(res["reference_sequence_ms"] / res["pipeline_ms"], 3),
            "tflops_4m_equiv": round(flop / (res["pipeline_ms"] * 1e-3) / 1e12, 2),
            "pct_fp8_peak": round(100.0 * flop / (res["pipeline_ms"] * 1e-3) / 1e12 / FP8_DENSE_PEAK_TFLOPS, 2),
            "fp6_tflops_4m_equiv": round(flop / (res["pipeline_fp6_ms"] * 1e-3) / 1e12, 2),
            "max_abs_diff_vs_reference_sequence": int(diff.max()),
            "mean_output": round(float(out.float().mean()), 2)}


def stage_study(torch, ftdx, dev, N: int, K: int, B: int, seed: int, reps: int, cfg: int, remap: int) -> dict:
    """Stand-alone unpack / pack stages (fused into the HERK on its main path) against the
    HBM peak: QC -> FP8 pairs over the headline input, QC -> FTD byte conversion, the
    N x N -> packed-triangle packs (lower with beta, GGP upper) over 8 channels, and the
    unfused (ForceBaseline: pre-cast rows, then the tile) HERK against the fused one, both on
    the headline's tile configuration.  Bytes are algorithmic: the packs read only the
    triangle of the N x N input."""
    qc = synth_qc(torch, N * K * B, seed, dev)
    n = qc.numel()
    full = torch.randn((8, N, N, 2), device=dev)
    old = torch.randn((8, N * (N + 1) // 2, 2), device=dev)
    res = {}

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    P = N * (N + 1) // 2
    cases = [
        ("promote_qc_to_fp8", lambda: ftdx.promote_qc_to_fp8(qc), 3 * n),
        ("convert_qc_to_ftd", lambda: ftdx.convert_qc_to_ftd(qc), 2 * n),
        ("pack_triangle_beta", lambda: ftdx.pack_triangle(full, N, 8, old=old, beta=0.5), 8 * 3 * P * 8),
        ("tri_from_herm", lambda: ftdx.tri_from_herm(full, N, 8), 8 * 2 * P * 8),
    ]
    for name, fn, nbytes in cases:
        ms = timed(fn)
        res[name] = {"ms": round(ms, 4), "bytes": int(nbytes), "gb_per_s": round(nbytes / ms / 1e6, 1),
                     "frac_hbm_peak": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 3)}
    del full, old
    torch.cuda.empty_cache()
    C = torch.empty((B, P, 2), dtype=torch.float32, device=dev)
    for mode, key in ((ftdx.HerkMode.ForceDirect, "herk_fused_ms"), (ftdx.HerkMode.ForceBaseline, "herk_unfused_ms")):
        ctx = ftdx.CutlassComplexGemm()
        ctx.set_strategy_tune_verbosity(0)
        ctx.set_herk_mode(mode)
        ctx.set_launch_config(cfg, remap)
        res[key] = round(timed(lambda: ctx.herk(qc, C, N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                                                ftdx.OutputPrecision.FP32, tune=False)), 4)
    res["inputs"], res["K"], res["channels"], res["seed"], res["tile_config"] = N, K, B, seed, cfg
    return res


def ragged_k_study(torch, ftdx, dev, N: int, K: int, B: int, seed: int, reps: int, cfg: int, remap: int) -> dict:
    """HERK on rows that 16-byte loads cannot take: K - 8 and K - 1 samples per row (the
    library copies the rows zero-padded to K rounded up to 16, per chunk of channels, then
    runs the fused tile).  Each is timed against the same rows padded beforehand and run
    at the padded K (no copy), and their output bytes are compared."""
    res = {"inputs": N, "channels": B, "tile_config": cfg, "cases": []}
    ctx = ftdx.CutlassComplexGemm()
    ctx.set_strategy_tune_verbosity(0)
    ctx.set_launch_config(cfg, remap)
    P = N * (N + 1) // 2

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    for Kr in (K - 8, K - 1):
        Kq = (Kr + 15) // 16 * 16
        A = synth_qc(torch, N * Kr * B, seed, dev)
        Ap = torch.zeros((B, N, Kq), dtype=torch.uint8, device=dev)
        Ap[:, :, :Kr] = A.view(B, N, Kr)
        C = torch.empty((B, P, 2), dtype=torch.float32, device=dev)
        R = torch.empty_like(C)

        def run(X, k, out):
            return ctx.herk(X, out, N, k, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                            ftdx.OutputPrecision.FP32, tune=False)

        ms = timed(lambda: run(A, Kr, C))
        ms_pre = timed(lambda: run(Ap, Kq, R))
        res["cases"].append({"K": Kr, "ms": round(ms, 4), "K_padded": Kq, "prepadded_ms": round(ms_pre, 4),
                             "ratio": round(ms / ms_pre, 3),
                             "identical_output": bool(torch.equal(C.view(torch.int32), R.view(torch.int32)))})
        del A, Ap, C, R
        torch.cuda.empty_cache()
    ctx.close()
    return res


def pol_reduce_study(torch, ftdx, dev, N: int, K: int, n_ch: int, n_pol: int, n_ti: int, seed: int,
                     reps: int) -> dict:
    """HERK + polarization sum: one fused pass (ftdx_herk_pol_reduce) against the HERK of
    every polarization followed by the stand-alone sum; outputs compared byte for byte."""
    B = n_ch * n_pol * n_ti
    A = synth_qc(torch, N * K * B, seed, dev)
    ctx = ftdx.CutlassComplexGemm()
    ctx.set_strategy_tune_verbosity(0)
    full = torch.empty((B, N * (N + 1) // 2, 2), dtype=torch.float32, device=dev)
    red = torch.empty((n_ch * n_ti, N * (N + 1) // 2, 2), dtype=torch.float32, device=dev)
    fused = torch.empty_like(red)

    def unfused():
        ctx.herk(A, full, N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                 ftdx.OutputPrecision.FP32, tune=False)
        ftdx.pol_reduce_triangle(full, N, n_ch, n_pol, n_ti, out=red)

    def one_pass():
        ftdx.herk_pol_reduce(ctx, A, N, K, n_ch, n_pol, n_ti, C=fused)

    res = {}
    for name, fn in (("unfused_ms", unfused), ("fused_ms", one_pass)):
        fn()
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        res[name] = round(e0.elapsed_time(e1) / reps, 4)
    same = bool(torch.equal(red.view(torch.int32), fused.view(torch.int32)))
    return {"inputs": N, "K": K, "channels": n_ch, "pols": n_pol, "time_inner": n_ti, "seed": seed,
            **res, "speedup": round(res["unfused_ms"] / res["fused_ms"], 3), "identical_output": same}


def grid_study(torch, ftdx, dev, N: int, K: int, n_ch: int, n_pol: int, Ng: int, seed: int, reps: int) -> dict:
    """HERK + polarization sum + pillbox UV gridding: one pass (ftdx_herk_pol_grid, the
    triangle never reaches HBM) against the HERK + pol-sum pass writing the triangle
    followed by the stand-alone scatter.  UV from random antenna positions (autos at the
    grid centre), frequencies 0.7-1.4 GHz.  FP32 atomic cell sums are order-dependent once
    a cell's sum passes 2^24 (the centre cell collects every auto-correlation here), so
    the grids are compared by their largest difference relative to the largest cell."""
    import numpy as np
    A = synth_qc(torch, N * K * n_ch * n_pol, seed, dev)
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1.0, 1.0, (N, 2)) * 600.0  # metres
    ii, jj = np.tril_indices(N)
    uv = torch.from_numpy((pos[ii] - pos[jj]).astype(np.float32)).to(dev)
    freqs = torch.from_numpy(np.linspace(0.7e9, 1.4e9, n_ch)).to(dev)
    # UV cell (wavelengths) so that the longest baseline at the highest frequency lands inside
    cell = float((pos.max(axis=0) - pos.min(axis=0)).max() * 1.4e9 / 299792458.0 * 1.05 / (Ng // 2))
    ctx = ftdx.CutlassComplexGemm()
    ctx.set_strategy_tune_verbosity(0)
    fused = torch.zeros((n_ch, Ng, Ng, 2), dtype=torch.float32, device=dev)
    unfused = torch.zeros_like(fused)
    tri = torch.empty((n_ch, N * (N + 1) // 2, 2), dtype=torch.float32, device=dev)
    n_bl = N * (N + 1) // 2

    def one_pass():
        ftdx.herk_pol_grid(ctx, A, N, K, n_ch, n_pol, uv, freqs, fused, Ng, cell)

    def two_pass():
        ftdx.herk_pol_reduce(ctx, A, N, K, n_ch, n_pol, 1, C=tri)
        ftdx.pillbox_grid(tri, uv, freqs, unfused, n_bl, Ng, n_ch, cell)

    res = {}
    for name, fn in (("unfused_ms", two_pass), ("fused_ms", one_pass)):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        res[name] = round(e0.elapsed_time(e1) / reps, 4)
    fused.zero_()
    unfused.zero_()
    one_pass()
    two_pass()
    torch.cuda.synchronize()
    diff = float((fused - unfused).abs().max() / unfused.abs().max())
    flop = 8.0 * N * N * K * n_pol * n_ch
    return {"inputs": N, "K": K, "channels": n_ch, "pols": n_pol, "grid": Ng, "seed": seed, **res,
            "speedup": round(res["unfused_ms"] / res["fused_ms"], 3),
            "fused_tflops_4m_equiv": round(flop / (res["fused_ms"] * 1e-3) / 1e12, 2),
            "triangle_bytes_avoided": 8 * n_bl * n_ch * 2,
            "max_rel_diff_vs_unfused": diff,
            "identical_output": bool(torch.equal(fused.view(torch.int32), unfused.view(torch.int32)))}


def streaming_study(torch, ftdx, dev, N: int, K: int, B: int, payloads: int, seed: int) -> dict:
    """Host payloads through ftdx_pipeline_* (H2D, HERK, D2H double-buffered): the
    PCIe-inclusive rate next to the HERK alone on resident data."""
    import numpy as np
    rng = np.random.default_rng(seed)
    ins = [rng.integers(0, 256, N * K * B, dtype=np.uint8) for _ in range(payloads)]
    ctx = ftdx.CutlassComplexGemm()
    ctx.set_strategy_tune_verbosity(0)
    pl = ftdx.Pipeline(ctx, N, K, B)
    outs = [np.empty(pl.out_shape(), dtype=np.float32) for _ in range(payloads)]
    pl.correlate_many(ins[:2], outs[:2])  # warm-up (first launches, page faults)
    t0 = time.perf_counter()
    pl.correlate_many(ins, outs)
    wall = time.perf_counter() - t0
    # the same payloads from page-locked host buffers: no staging copies
    keep = [torch.empty(N * K * B, dtype=torch.uint8, pin_memory=True) for _ in range(payloads)]
    pins = []
    for t, x in zip(keep, ins):
        t.numpy()[:] = x
        pins.append(t.numpy())
    pouts_t = [torch.empty(pl.out_shape(), dtype=torch.float32, pin_memory=True) for _ in range(payloads)]
    pouts = [t.numpy() for t in pouts_t]
    pl.correlate_many(pins[:2], pouts[:2])
    t0 = time.perf_counter()
    pl.correlate_many(pins, pouts)
    wall_pinned = time.perf_counter() - t0
    same_pinned = all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(outs, pouts))
    A = torch.from_numpy(ins[0]).to(dev)
    C = torch.empty((B, N * (N + 1) // 2, 2), dtype=torch.float32, device=dev)
    for _ in range(2):
        ctx.herk(A, C, N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                 ftdx.OutputPrecision.FP32, tune=False)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        ctx.herk(A, C, N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                 ftdx.OutputPrecision.FP32, tune=False)
    e1.record()
    torch.cuda.synchronize()
    herk_ms = e0.elapsed_time(e1) / 5
    same = bool(np.array_equal(outs[0].view(np.uint32), C.cpu().numpy().view(np.uint32)))
    pl.close()
    in_b, out_b = N * K * B, B * N * (N + 1) * 4
    return {
        "inputs": N, "channels": B, "K": K, "payloads": payloads, "seed": seed,
        "data": "uniform random QC bytes in pageable host memory",
        "ms_per_payload": round(wall * 1e3 / payloads, 3),
        "herk_only_ms": round(herk_ms, 3),
        "tflops_4m_equiv_pcie_inclusive": round(8.0 * N * N * K * B * payloads / wall / 1e12, 2),
        "host_bytes_per_payload": in_b + out_b,
        "host_gb_per_s": round((in_b + out_b) * payloads / wall / 1e9, 2),
        "first_output_matches_resident_herk": same,
        "page_locked": {"ms_per_payload": round(wall_pinned * 1e3 / payloads, 3),
                        "host_gb_per_s": round((in_b + out_b) * payloads / wall_pinned / 1e9, 2),
                        "outputs_equal_pageable": same_pinned},
    }


def cpu_baseline(N: int, K: int, seed: int, target_s: float) -> dict:
    """Oracle HERK (int64, OpenMP) on rows of one channel; rate in 4M-equiv TFLOP/s."""
    import numpy as np
    from oracle import oracle
    threads = len(os.sched_getaffinity(0))
    env_t = os.environ.get("OMP_NUM_THREADS")
    if env_t and env_t.isdigit():
        threads = min(threads, int(env_t))
    A1 = oracle.gen_qc_uniform(seed, N * K)
    C1 = np.zeros((1, N * (N + 1) // 2, 2), dtype=np.int64)
    tri_macs = K * N * (N + 1) / 2  # complex MACs of one channel's lower triangle

    # calibrate on the heaviest rows of one channel, doubling until >= 0.5 s
    cal_rows = 32
    while True:
        t0 = time.perf_counter()
        oracle.herk_int64_rows(A1, N, K, 1, N - cal_rows, N, threads=threads, out=C1)
        t_cal = time.perf_counter() - t0
        if t_cal >= 0.5 or cal_rows >= N:
            break
        cal_rows = min(N, cal_rows * 4)
    rate = sum(i + 1 for i in range(N - cal_rows, N)) * K / max(t_cal, 1e-6)
    t_channel = tri_macs / rate
    if t_channel >= target_s:
        # part of one channel: the heaviest rows first, until ~target_s of work
        want, r0, macs = rate * target_s, N, 0
        while r0 > 0 and macs < want:
            r0 -= 1
            macs += (r0 + 1) * K
        t0 = time.perf_counter()
        oracle.herk_int64_rows(A1, N, K, 1, r0, N, threads=threads, out=C1)
        t = time.perf_counter() - t0
        channels = macs / tri_macs
        sample = f"rows {r0}..{N - 1} of one channel ({channels * 100:.1f}% of its lower triangle)"
    else:
        nch = max(1, min(96, int(round(target_s / t_channel))))
        A = oracle.gen_qc_uniform(seed, N * K * nch)
        C = np.zeros((nch, N * (N + 1) // 2, 2), dtype=np.int64)
        t0 = time.perf_counter()
        oracle.herk_int64_rows(A, N, K, nch, 0, N, threads=threads, out=C)
        t = time.perf_counter() - t0
        channels = float(nch)
        sample = f"{nch} full channel(s)"
    flops_equiv = 8.0 * N * N * K * channels
    return {
        "value": flops_equiv / t / 1e12,
        "unit": "TFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle int64 HERK (OpenMP, {threads} threads), {sample} of the same "
                   f"workload (N={N}, K={K}), {t:.1f} s; 4M-equiv FLOP = 8 N^2 K per channel"),
        "cpu_model": _cpu_model(),
    }


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--N", type=int, default=4096)
    ap.add_argument("--K", type=int, default=2048)
    ap.add_argument("--batch", type=int, default=96)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-mx", action="store_true")
    ap.add_argument("--no-calib", action="store_true")
    ap.add_argument("--no-stream", action="store_true")
    ap.add_argument("--no-beamformer", action="store_true")
    ap.add_argument("--no-polsum", action="store_true")
    ap.add_argument("--no-grid", action="store_true")
    ap.add_argument("--voltage", action="store_true",
                    help="also time the voltage beamformer pipeline (a caller outside the correlator path)")
    ap.add_argument("--no-stages", action="store_true")
    ap.add_argument("--no-bf6", action="store_true")
    ap.add_argument("--no-ragged", action="store_true")
    ap.add_argument("--batched-problems", type=int, default=512)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the timed steps, one plain device copy of this many MiB "
                         "(reference kernel for scaling rocprof byte counters)")
    ap.add_argument("--tile-config", type=int, default=-1,
                    help="-1: autotune the headline shape; else a fixed FTDX_TILE_* id")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "channels shard across GPUs with no exchange, one process per GPU", file=sys.stderr)
        return 2

    import torch
    from ftd_amd import ftdx

    N, K, B = args.N, args.K, args.batch
    dev = torch.device("cuda:0")
    log(f"bench: plain HERK N={N} K={K} batch={B} steps={args.steps} warmup={args.warmup}")
    A = synth_qc(torch, N * K * B, args.seed, dev)
    C = torch.empty((B, N * (N + 1) // 2, 2), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()
    ctx = ftdx.CutlassComplexGemm()
    ctx.set_strategy_tune_verbosity(0)
    if args.tile_config >= 0:
        ctx.set_launch_config(args.tile_config, 1)

    def step(tune=False):
        rc = ctx.herk(A, C, N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                      ftdx.OutputPrecision.FP32, stream=stream, tune=tune)
        if rc != 0:
            raise RuntimeError(f"herk failed: {rc} {ftdx.last_error()}")

    step(tune=args.tile_config < 0)  # autotune (cached in the context) outside the timed region
    cfg, remap = ctx.query_config(N, K, B)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        starts[s].record(stream)
        step()
        ends[s].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = sorted(starts[s].elapsed_time(ends[s]) for s in range(args.steps))
    ms_total = sum(per)
    ms_per_step = wall * 1e3 / args.steps
    kern_ms = ms_total / args.steps

    flop = 8.0 * N * N * K * B
    value = flop / (ms_per_step * 1e-3) / 1e12
    achieved = flop / (kern_ms * 1e-3) / 1e12
    log(f"bench: {ms_per_step:.3f} ms/step wall, kernel avg {kern_ms:.3f} ms "
        f"(median {per[len(per) // 2]:.3f}), {value:.1f} TFLOP/s 4M-equiv")

    # the output of the last timed step, checked exactly (outside the timed region):
    # weighted checksums of every channel, exact rows of three channels
    from ftd_amd import selfcheck
    t_chk = time.perf_counter()
    rows = sorted({r for r in (0, 1, 31, 32, 127, 128, 129, 255, 256, 1000, N // 2, N - 2, N - 1)
                   if 0 <= r < N})
    parity = selfcheck.check_herk(torch, A, C, N, K, B, rows=rows,
                                  row_channels=sorted({0, B // 2, B - 1}))
    parity["seconds"] = round(time.perf_counter() - t_chk, 2)
    parity["tile_config"] = cfg
    log(f"bench: parity of the timed output: ok={parity['ok']} "
        f"({parity['weighted_checksums_equal']}/{B} channel checksums, "
        f"{parity['exact_rows_checked']} exact rows)")
    exec_flop = executed_mfma_flop(N, K, B, cfg)

    bf6 = None
    if not args.no_bf6:
        # the same tile fed by E3M2 operand rows (one extra pass over the input)
        ctx6 = ftdx.CutlassComplexGemm()
        ctx6.set_launch_config(cfg | ftdx.TILE_BF6, remap)
        C6 = torch.empty_like(C)

        def step6():
            rc = ctx6.herk(A, C6, N, K, B, ftdx.InputPrecision.INT4, ftdx.ComputePrecision.FP8,
                           ftdx.OutputPrecision.FP32, stream=stream, tune=False)
            if rc != 0:
                raise RuntimeError(f"herk (E3M2) failed: {rc} {ftdx.last_error()}")

        for _ in range(2):
            step6()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(5):
            step6()
        e1.record(stream)
        torch.cuda.synchronize()
        ms6 = e0.elapsed_time(e1) / 5
        same = bool(torch.equal(C.view(torch.int32), C6.view(torch.int32)))
        bf6 = {"tile_config": cfg | ftdx.TILE_BF6, "ms": round(ms6, 4),
               "tflops_4m_equiv": round(flop / (ms6 * 1e-3) / 1e12, 2),
               "fp8_headline_ms": round(kern_ms, 4), "identical_output": same,
               "note": "E3M2 conversion pass + E3M2 MFMA tile; opt-in (FTDX_TILE_BF6)"}
        log(f"bench: E3M2 operands {ms6:.3f} ms vs FP8 {kern_ms:.3f} ms, identical={same}")
        del C6, ctx6
        torch.cuda.empty_cache()

    if args.calib_copy_mib > 0:
        n = args.calib_copy_mib * (1 << 20)
        src = torch.ones(n, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"bench: calibration copy of {args.calib_copy_mib} MiB done")
        del src, dst

    batched = None
    if not args.no_batched:
        del C
        torch.cuda.empty_cache()
        log(f"bench: batched uneven HERK, {args.batched_problems} problems ...")
        batched = batched_vs_plain(torch, ftdx, dev, args.batched_problems, 3, 5)
        log(f"bench: grouped {batched['grouped_ms']:.3f} ms per replay (first after the loops "
            f"{batched['grouped_first_replay_ms']:.3f} ms) vs plain loop {batched['plain_loop_ms']} "
            f"-> {batched['speedup']}, identical={batched['identical_output']}, "
            f"parity ok={batched['grouped_parity']['ok']}")

    mx = None
    if not args.no_mx:
        log("bench: MX block-scaled HERK (FP6 / FP4) 4096 x 32 x K=4096 ...")
        mx = mx_study(torch, ftdx, dev, 4096, 4096, 32, 4, 3)
        log(f"bench: MX fp6 {mx['fp6_e3m2']['ms']:.3f} ms rel-fro {mx['fp6_e3m2']['rel_frobenius']['max']:.3e}; "
            f"fp4 {mx['fp4_e2m1']['ms']:.3f} ms rel-fro {mx['fp4_e2m1']['rel_frobenius']['max']:.3e}")

    calib = None
    if not args.no_calib:
        log("bench: gain calibration + HERK adjoint 1024 x 16 x K=2048 ...")
        calib = calibration_study(torch, ftdx, dev, 1024, 2048, 16, 5, 5)
        log(f"bench: calib fwd {calib['forward_herk_ms']:.3f} ms, gain {calib['gain_loss_grad_ms']:.3f} ms, "
            f"adjoint {calib['herk_adjoint_ms']:.3f} ms; dX err {calib['dX_max_rel_err_vs_fp64']:.2e}")
        torch.cuda.empty_cache()

    bf = None
    if not args.no_beamformer:
        log("bench: voltage beamformer GEMM + power, M=256 N=4000 K=1664 batch=128 ...")
        bf = beamformer_study(torch, ftdx, dev, 256, 4000, 1664, 128, 7, 10)
        log(f"bench: beamformer {bf['ms']:.3f} ms, {bf['tflops_4m_equiv']:.0f} TFLOP/s 4M-equiv")
        torch.cuda.empty_cache()

    stages = None
    if not args.no_stages:
        log("bench: stand-alone unpack / pack stages vs HBM peak ...")
        stages = stage_study(torch, ftdx, dev, N, K, B, args.seed, 5, cfg, remap)
        log("bench: " + ", ".join(f"{k} {v['gb_per_s']:.0f} GB/s" for k, v in stages.items() if isinstance(v, dict))
            + f"; HERK fused {stages['herk_fused_ms']:.3f} ms vs unfused {stages['herk_unfused_ms']:.3f} ms")
        torch.cuda.empty_cache()

    ragged = None
    if not args.no_ragged:
        log(f"bench: HERK at K={K - 8} and K={K - 1} (rows padded per chunk) vs pre-padded rows ...")
        ragged = ragged_k_study(torch, ftdx, dev, N, K, B, args.seed, 5, cfg, remap)
        log("bench: " + "; ".join(f"K={c['K']} {c['ms']:.3f} ms vs pre-padded K={c['K_padded']} "
                                  f"{c['prepadded_ms']:.3f} ms, identical={c['identical_output']}"
                                  for c in ragged["cases"]))

    volt = None
    if args.voltage:
        log("bench: voltage pipeline 1664 ant x 4000 beams x 128 ch, 256 times, n_tps 4 ...")
        volt = voltage_study(torch, ftdx, dev, 1664, 4000, 128, 256, 4, 11, 10)
        log(f"bench: voltage pipeline {volt['pipeline_ms']:.3f} ms vs reference sequence "
            f"{volt['reference_sequence_ms']:.3f} ms ({volt['speedup']:.2f}x), "
            f"{volt['tflops_4m_equiv']:.0f} TFLOP/s 4M-equiv; FP6 mode {volt['pipeline_fp6_ms']:.3f} ms")
        torch.cuda.empty_cache()

    polsum = None
    if not args.no_polsum:
        log("bench: HERK + polarization sum, 2048 inputs x 16 ch x 2 pol x 2 ti, K=1024 ...")
        polsum = pol_reduce_study(torch, ftdx, dev, 2048, 1024, 16, 2, 2, 8, 5)
        log(f"bench: pol sum fused {polsum['fused_ms']:.3f} ms vs unfused {polsum['unfused_ms']:.3f} ms "
            f"({polsum['speedup']:.2f}x, identical={polsum['identical_output']})")
        torch.cuda.empty_cache()

    grid = None
    if not args.no_grid:
        log("bench: HERK + pol sum + UV gridding, 4096 inputs x 16 ch x 2 pol, K=1024, 1024^2 grid ...")
        grid = grid_study(torch, ftdx, dev, 4096, 1024, 16, 2, 1024, 12, 5)
        log(f"bench: grid fused {grid['fused_ms']:.3f} ms vs HERK+pol sum then scatter {grid['unfused_ms']:.3f} ms "
            f"({grid['speedup']:.2f}x, max rel diff {grid['max_rel_diff_vs_unfused']:.2e})")
        torch.cuda.empty_cache()

    stream_res = None
    if not args.no_stream:
        log("bench: streaming host payloads 1024 x 16 x K=2048, 8 payloads ...")
        stream_res = streaming_study(torch, ftdx, dev, 1024, 2048, 16, 8, 6)
        log(f"bench: stream {stream_res['ms_per_payload']:.3f} ms/payload "
            f"({stream_res['host_gb_per_s']:.1f} GB/s host traffic; page-locked "
            f"{stream_res['page_locked']['ms_per_payload']:.3f} ms, "
            f"{stream_res['page_locked']['host_gb_per_s']:.1f} GB/s) vs HERK alone "
            f"{stream_res['herk_only_ms']:.3f} ms")

    cpu = None
    if not args.no_cpu:
        log("bench: CPU baseline (oracle) ...")
        cpu = cpu_baseline(N, K, args.seed, args.cpu_seconds)
        log(f"bench: CPU {cpu['value']:.4f} TFLOP/s on {cpu['cores']} threads")

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp8_e4m3 (fp32 accumulate)",
        "data": "synthetic",
        "config": {
            "workload": "plain HERK, int4 QC voltages -> FP8 E4M3 MFMA, packed lower-triangle FP32",
            "inputs": N, "channels": B, "K": K, "seed": args.seed,
            "parallelism": "single GPU",
        },
        "pct_fp8_peak": round(100.0 * value / FP8_DENSE_PEAK_TFLOPS, 2),
        "parity_checked": bool(parity["ok"]),
        "parity": parity,
        "roofline": {
            "bound": "mfma",
            # 4M-equivalent FLOP (8 N^2 K per channel, the metric's convention) per kernel time
            "achieved": round(achieved, 2),
            "peak": FP8_DENSE_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP8_DENSE_PEAK_TFLOPS, 4),
            # MFMA FLOP the launch really executes (lower-triangle tiles only, about half
            # of the 4M-equivalent count) per kernel time
            "executed_flop": exec_flop,
            "executed_tflops": round(exec_flop / (kern_ms * 1e-3) / 1e12, 2),
            "executed_frac": round(exec_flop / (kern_ms * 1e-3) / 1e12 / FP8_DENSE_PEAK_TFLOPS, 4),
            # HBM bytes per launch: not measured in this run -- the rocprofv3 FETCH_SIZE /
            # WRITE_SIZE figures of the committed profile for this shape and tile config
            "traffic": (MEASURED_TRAFFIC_BYTES.get(cfg)
                        if (N, K, B) == (4096, 2048, 96) else None),
            "traffic_source": ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, summary in profiles/ "
                               "(not measured by this run)"),
            "kernel": "herk_int4_fp8_plain_kernel",
            "kernel_ms": round(kern_ms, 4),
            "tile_config": cfg, "xcd_remap": remap,
        },
        "batched_vs_plain": batched,
        "mx_block_scaled": mx,
        "gain_calibration": calib,
        "streaming_host_payloads": stream_res,
        "beamformer_gemm_power": bf,
        "herk_pol_sum_fusion": polsum,
        "herk_pol_grid_fusion": grid,
        "voltage_pipeline": volt,
        "stages_vs_hbm_peak": stages,
        "ragged_k": ragged,
        "e3m2_operands": bf6,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    if ragged and not all(c["identical_output"] for c in ragged["cases"]):
        log("bench: ragged-K output differs from the pre-padded rows' output")
        return 3
    if not parity["ok"] or (batched and not batched["grouped_parity"]["ok"]):
        log("bench: PARITY FAILURE in the timed output")
        return 3
    return 0


if __name__ == "__main__":
    sys.exit(main())
