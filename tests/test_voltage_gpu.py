"""GPU parity of the voltage beamformer (the ggp::VoltagePipeline drop-in) against the
oracle: bit-exact uint8 and FP32 filterbanks (the reference's FP32 operation order on
beams that every summation order gives exactly), the channel-fused / multi-payload
layout of beamform_batched, compute()'s corner turn and quantization, graph replay,
the FP6 path, the reference's own acceptance test against its FP64 CPU beamformer
(tests/voltage_pipeline_test.cpp:135-246, lifecycle :270-300) and its error paths."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu


def _data(oracle, seed, n_ant, n_beam, n_ch, n_time, n_payloads=1):
    payloads = [oracle.gen_bytes(seed * 31 + p, n_ch * 2 * n_ant * n_time) for p in range(n_payloads)]
    w = np.random.default_rng(seed).uniform(-1, 1, (n_ch, n_beam, n_ant, 2)).astype(np.float32)
    return payloads, w


def _pipeline(ftdx, dev, w, **cfg):
    import torch
    vp = ftdx.VoltagePipeline(**cfg)
    vp.set_weights(torch.from_numpy(w).to(dev))
    return vp


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


# n_ant, n_beam, n_ch, n_time, n_tps
CASES = [
    (64, 256, 4, 128, 4),    # the reference's defaults, n_time shortened
    (100, 130, 2, 96, 3),    # ragged antennas and beams; n_tps not dividing 64 (integration pass)
    (64, 64, 2, 200, 8),     # n_time % 16 != 0: byte loads, partial last sample block
    (520, 300, 2, 256, 16),  # several K stages, partial beam tile
    (64, 128, 1, 256, 128),  # n_tps > 64
    (1, 1, 1, 4, 1),         # smallest
]


@pytest.mark.parametrize("n_ant,n_beam,n_ch,n_time,n_tps", CASES)
def test_beamform_and_filterbank_bit_exact(oracle, dev, n_ant, n_beam, n_ch, n_time, n_tps):
    import torch
    from ftd_amd import ftdx
    (qc,), w = _data(oracle, n_ant + n_time, n_ant, n_beam, n_ch, n_time)
    ref_fb = oracle.voltage_beamform([qc], w, n_ant, n_beam, n_ch, n_time, n_tps, out_u8=False)
    scale = float(200.0 / max(ref_fb.max(), 1e-30))
    ref = oracle.voltage_beamform([qc], w, n_ant, n_beam, n_ch, n_time, n_tps, scale=scale)
    print(f"voltage beamform ant={n_ant} beam={n_beam} ch={n_ch} t={n_time} tps={n_tps}", flush=True)
    vp = _pipeline(ftdx, dev, w, n_antennae=n_ant, n_beams=n_beam, n_channels=n_ch, n_time=n_time,
                   n_time_power_sum=n_tps)
    assert vp.uses_direct_path and vp.n_time_out == n_time // n_tps
    d_qc = torch.from_numpy(qc).to(dev)
    out = torch.full((n_ch, n_beam, n_time // n_tps), 77, dtype=torch.uint8, device=dev)
    vp.beamform(out, d_qc, scale)
    vp.compute_filterbank(d_qc)
    fb = vp.filterbank("chanmajor")
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), ref)
    assert np.array_equal(_bits(fb.cpu().numpy().reshape(ref_fb.shape)), _bits(ref_fb))


@pytest.mark.parametrize("cf,npl", [(2, 3), (1, 2), (4, 1), (2, 16)])
def test_beamform_batched_layout_bit_exact(oracle, dev, cf, npl):
    import torch
    from ftd_amd import ftdx
    n_ant, n_beam, n_ch, n_time, n_tps = 64, 96, 4, 64, 4
    payloads, w = _data(oracle, 40 + cf * 7 + npl, n_ant, n_beam, n_ch, n_time, npl)
    scale = 0.02
    ref = oracle.voltage_beamform(payloads, w, n_ant, n_beam, n_ch, n_time, n_tps, scale=scale, ch_fuse=cf)
    print(f"voltage batched cf={cf} payloads={npl}", flush=True)
    vp = _pipeline(ftdx, dev, w, n_antennae=n_ant, n_beams=n_beam, n_channels=n_ch, n_time=n_time,
                   n_time_power_sum=n_tps, ch_fuse=cf, max_payloads=npl)
    d = [torch.from_numpy(p).to(dev) for p in payloads]
    out = torch.zeros(ref.shape, dtype=torch.uint8, device=dev)
    vp.beamform_batched(out, d, scale)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), ref)


def test_compute_corner_turn_and_quantize(oracle, dev):
    import torch
    from ftd_amd import ftdx
    n_ant, n_beam, n_ch, n_time, n_tps = 48, 40, 3, 64, 4
    (qc,), w = _data(oracle, 9, n_ant, n_beam, n_ch, n_time)
    ref_fb = oracle.voltage_beamform([qc], w, n_ant, n_beam, n_ch, n_time, n_tps, out_u8=False)
    scale = float(180.0 / ref_fb.max())
    vp = _pipeline(ftdx, dev, w, n_antennae=n_ant, n_beams=n_beam, n_channels=n_ch, n_time=n_time,
                   n_time_power_sum=n_tps)
    out = torch.zeros((n_ch, n_beam, n_time // n_tps), dtype=torch.uint8, device=dev)
    vp.compute(out, torch.from_numpy(qc).to(dev), None, scale)
    chan = vp.filterbank("chanmajor").cpu().numpy().reshape(ref_fb.shape)
    beam = vp.filterbank("beammajor").cpu().numpy().reshape(n_beam, n_ch, -1)
    assert np.array_equal(_bits(chan), _bits(ref_fb))
    assert np.array_equal(beam, ref_fb.transpose(1, 0, 2))
    # the same rule as beamform(): compute() and beamform() agree byte for byte
    ref = oracle.voltage_beamform([qc], w, n_ant, n_beam, n_ch, n_time, n_tps, scale=scale)
    assert np.array_equal(out.cpu().numpy(), ref)


@pytest.mark.parametrize("Nf,Nb,Nt", [(5, 37, 1), (33, 70, 3), (4, 9, 16), (3, 5, 100)])
def test_corner_turn(dev, Nf, Nb, Nt):
    import torch
    from ftd_amd import ftdx
    x = torch.randn((Nf, Nb, Nt), device=dev)
    y = torch.empty((Nb, Nf, Nt), device=dev)
    ftdx.corner_turn(y, x, Nf, Nb, Nt)
    torch.cuda.synchronize()
    assert torch.equal(y, x.transpose(0, 1).contiguous())


def test_graph_replay_tracks_input_and_recaptures(oracle, dev):
    import torch
    from ftd_amd import ftdx
    n_ant, n_beam, n_ch, n_time, n_tps = 64, 64, 2, 128, 4
    payloads, w = _data(oracle, 12, n_ant, n_beam, n_ch, n_time, 3)
    vp = _pipeline(ftdx, dev, w, n_antennae=n_ant, n_beams=n_beam, n_channels=n_ch, n_time=n_time,
                   n_time_power_sum=n_tps)
    vp.set_graph_enabled(True)
    d_qc = torch.empty(payloads[0].size, dtype=torch.uint8, device=dev)
    out = torch.zeros((n_ch, n_beam, n_time // n_tps), dtype=torch.uint8, device=dev)
    for i, (p, scale) in enumerate([(0, 0.01), (1, 0.01), (2, 0.01), (0, 0.03), (1, 0.03)]):
        d_qc.copy_(torch.from_numpy(payloads[p]))  # eager work between replays
        out.zero_()
        vp.beamform(out, d_qc, scale)
        torch.cuda.synchronize()
        ref = oracle.voltage_beamform([payloads[p]], w, n_ant, n_beam, n_ch, n_time, n_tps, scale=scale)
        assert np.array_equal(out.cpu().numpy(), ref), i


def test_fp6_path_bit_exact(oracle, dev):
    import torch
    from ftd_amd import ftdx
    n_ant, n_beam, n_ch, n_time, n_tps = 64, 72, 3, 96, 4
    (qc,), w = _data(oracle, 21, n_ant, n_beam, n_ch, n_time)
    ref_fb = oracle.voltage_beamform([qc], w, n_ant, n_beam, n_ch, n_time, n_tps, compute="fp6", out_u8=False)
    scale = float(200.0 / ref_fb.max())
    ref = oracle.voltage_beamform([qc], w, n_ant, n_beam, n_ch, n_time, n_tps, scale=scale, compute="fp6")
    vp = _pipeline(ftdx, dev, w, n_antennae=n_ant, n_beams=n_beam, n_channels=n_ch, n_time=n_time,
                   n_time_power_sum=n_tps, compute_mode=ftdx.VOLTAGE_FP6)
    assert not vp.uses_direct_path
    d_qc = torch.from_numpy(qc).to(dev)
    out = torch.zeros(ref.shape, dtype=torch.uint8, device=dev)
    vp.beamform(out, d_qc, scale)
    vp.compute_filterbank(d_qc)
    fb = vp.filterbank("chanmajor").cpu().numpy().reshape(ref_fb.shape)
    assert np.array_equal(out.cpu().numpy(), ref)
    assert np.array_equal(_bits(fb), _bits(ref_fb))


@pytest.mark.parametrize("compute", ["fp8", "fp6"])
def test_reference_acceptance_vs_fp64(oracle, dev, compute):
    """The reference's run_beamformer_test: 128 antennas, 64 beams, 4 channels, 128 times,
    n_tps 4, scale = 200 / max FP64 power, tolerance max(4, 1 + n_ant/32) (FP8) or
    max(15, 1 + n_ant/8) (FP6) against round() of the FP64 power; no element beyond it."""
    import torch
    from ftd_amd import ftdx
    n_ant, n_beam, n_ch, n_time, n_tps = 128, 64, 4, 128, 4
    (qc,), w = _data(oracle, 42, n_ant, n_beam, n_ch, n_time)
    ref64 = oracle.voltage_fp64(qc, w, n_ant, n_beam, n_ch, n_time, n_tps)
    scale = float(np.float32(200.0) / np.float32(ref64.max()))
    mode = ftdx.VOLTAGE_FP8 if compute == "fp8" else ftdx.VOLTAGE_FP6
    vp = _pipeline(ftdx, dev, w, n_antennae=n_ant, n_beams=n_beam, n_channels=n_ch, n_time=n_time,
                   n_time_power_sum=n_tps, compute_mode=mode)
    assert vp.input_size == n_ch * 2 * n_ant * n_time and vp.output_size == n_ch * n_beam * n_time // n_tps
    out = torch.zeros((n_ch, n_beam, n_time // n_tps), dtype=torch.uint8, device=dev)
    vp.beamform(out, torch.from_numpy(qc).to(dev), scale)
    got = out.cpu().numpy().astype(np.int64)
    want = np.clip(np.floor(ref64 * scale + 0.5), 0, 255).astype(np.int64)  # round(): half away from 0
    tol = max(4, 1 + n_ant // 32) if compute == "fp8" else max(15, 1 + n_ant // 8)
    diff = np.abs(got - want)
    print(f"{compute}: max_diff={diff.max()} mean_diff={diff.mean():.2f} tol={tol}", flush=True)
    assert (diff > tol).sum() == 0


def test_lifecycle_sizes_and_errors(dev):
    """Sizes of the reference's lifecycle test and the VP_ERROR cases."""
    import torch
    from ftd_amd import ftdx
    vp = ftdx.VoltagePipeline(n_antennae=64, n_beams=32, n_channels=2, n_time=64, n_time_power_sum=4)
    assert (vp.input_size, vp.output_size, vp.n_time_out) == (2 * 2 * 64 * 64, 2 * 32 * 16, 16)
    qc = torch.zeros(vp.input_size, dtype=torch.uint8, device=dev)
    out = torch.zeros(vp.output_size, dtype=torch.uint8, device=dev)
    with pytest.raises(ftdx.FtdxError, match="set_weights"):
        vp.beamform(out, qc)
    vp.set_weights(torch.zeros((2, 32, 64, 2), device=dev))
    with pytest.raises(ftdx.FtdxError, match="ch_fuse > 1 or n_payloads > 1"):
        vp.beamform_batched(out, [qc])
    out2 = torch.zeros(2 * vp.output_size, dtype=torch.uint8, device=dev)
    with pytest.raises(ftdx.FtdxError, match="n_payloads"):
        vp.beamform_batched(out2, [qc, qc])  # max_payloads = 1
    vp.beamform(out, qc)
    torch.cuda.synchronize()
    assert int(out.max()) == 0
    for bad in (dict(n_polarizations=1), dict(n_time=63), dict(ch_fuse=3), dict(n_channels=2, ch_fuse=4),
                dict(max_payloads=17), dict(n_beams=0)):
        cfg = dict(n_antennae=64, n_beams=32, n_channels=2, n_time=64, n_time_power_sum=4)
        cfg.update(bad)
        with pytest.raises(ftdx.FtdxError):
            ftdx.VoltagePipeline(**cfg)
    fused = ftdx.VoltagePipeline(n_antennae=64, n_beams=32, n_channels=4, n_time=64, n_time_power_sum=4,
                                 ch_fuse=2)
    fused.set_weights(torch.zeros((4, 32, 64, 2), device=dev))
    with pytest.raises(ftdx.FtdxError, match="beamform_batched"):
        fused.beamform(torch.zeros(4 * 32 * 16, dtype=torch.uint8, device=dev),
                       torch.zeros(4 * 2 * 64 * 64, dtype=torch.uint8, device=dev))
    fp6 = ftdx.VoltagePipeline(n_antennae=64, n_beams=32, n_channels=2, n_time=64, n_time_power_sum=4,
                               compute_mode=ftdx.VOLTAGE_FP6, max_payloads=2)
    fp6.set_weights(torch.zeros((2, 32, 64, 2), device=dev))
    with pytest.raises(ftdx.UnsupportedError, match="FP8"):
        fp6.beamform_batched(out2, [qc, qc])


def test_cpp_drop_in_voltage_pipeline(dev):
    """C++ ggp::VoltagePipeline (include/voltage_pipeline.h) through libftdx.so: the
    reference test's sizes (128 antennas, 64 beams, 4 channels, 128 times)."""
    exe = os.path.join(ROOT, "tests", "cpp", "test_voltage_api")
    assert os.path.exists(exe), "tests/cpp/test_voltage_api not built"
    r = subprocess.run([exe, "128", "64", "4", "128", "42"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
