"""CPU pinning of the voltage-beamformer oracle: against the reference's FP64 CPU
beamformer (tests/voltage_pipeline_test.cpp:43-83) on weights every rounding step
keeps exact, the channel-fusion / payload layout against per-channel runs, the
quantization rule at ties, saturation and NaN, and planted mistakes (conjugated
weights, a dropped polarization)."""
import numpy as np
import pytest

# weights that FP16, E4M3 and MX E3M2 (block amax <= 1 -> scale 2^-4 or finer) all hold exactly
EXACT_W = np.array([-1.0, -0.5, -0.25, 0.0, 0.25, 0.5, 1.0], dtype=np.float32)


def _case(oracle, seed, n_ant, n_beam, n_ch, n_time, n_payloads=1, exact=True):
    rng = np.random.default_rng(seed)
    payloads = [oracle.gen_bytes(seed * 31 + p, n_ch * 2 * n_ant * n_time) for p in range(n_payloads)]
    shape = (n_ch, n_beam, n_ant, 2)
    w = rng.choice(EXACT_W, shape) if exact else rng.uniform(-1, 1, shape).astype(np.float32)
    return payloads, w.astype(np.float32)


@pytest.mark.parametrize("compute", ["fp8", "fp6"])
def test_exact_weights_match_fp64_beamformer(oracle, compute):
    # 16 antennas: every power and running sum is a multiple of 1/16 below 2^20, exact in FP32
    n_ant, n_beam, n_ch, n_time, n_tps = 16, 12, 3, 32, 4
    (qc,), w = _case(oracle, 1, n_ant, n_beam, n_ch, n_time)
    fb = oracle.voltage_beamform([qc], w, n_ant, n_beam, n_ch, n_time, n_tps, compute=compute, out_u8=False)
    ref = oracle.voltage_fp64(qc, w, n_ant, n_beam, n_ch, n_time, n_tps)
    assert np.array_equal(fb.astype(np.float64), ref)
    scale = 255.0 / ref.max() * 0.999
    u8 = oracle.voltage_beamform([qc], w, n_ant, n_beam, n_ch, n_time, n_tps, scale=scale, compute=compute)
    want = np.clip(np.rint(ref.astype(np.float32) * np.float32(scale)), 0, 255).astype(np.uint8)
    assert np.array_equal(u8, want)
    assert u8.max() > 200


def test_planted_conjugate_and_dropped_pol_are_caught(oracle):
    n_ant, n_beam, n_ch, n_time, n_tps = 16, 8, 2, 16, 2
    (qc,), w = _case(oracle, 2, n_ant, n_beam, n_ch, n_time)
    ref = oracle.voltage_fp64(qc, w, n_ant, n_beam, n_ch, n_time, n_tps)
    wc = w.copy()
    wc[..., 1] *= -1  # V conj(W) instead of V W
    assert not np.array_equal(oracle.voltage_beamform([qc], wc, n_ant, n_beam, n_ch, n_time, n_tps,
                                                      out_u8=False).astype(np.float64), ref)
    W = oracle.voltage_weights(w, n_ch, n_beam, n_ant, "fp8")
    re, im = oracle.voltage_beams([qc], W, n_ant, n_beam, n_ch, n_time)
    re[:, 1] = 0
    im[:, 1] = 0  # pol 1 dropped
    assert not np.array_equal(oracle.voltage_tail(re, im, n_tps, 0, 1.0, False).astype(np.float64), ref)


def test_fused_layout_matches_per_channel_runs(oracle):
    """beamform_batched with ch_fuse = 2 and 3 payloads: output [n_ch / 2][beam][(p*2 + c)*n_tout
    + t] equals the single-payload output of channel 2b + c of payload p, when the weights of
    fused channels agree."""
    n_ant, n_beam, n_ch, n_time, n_tps, cf, npl = 20, 6, 4, 24, 3, 2, 3
    payloads, w = _case(oracle, 3, n_ant, n_beam, n_ch, n_time, npl, exact=False)
    w[1::2] = w[0::2]  # channel pairs share weights
    # the prepared sets are the first n_ch / ch_fuse (lib/voltage_pipeline.cu:770-776):
    # batch b uses set b, so place pair b's weights there too
    wb = w.copy()
    wb[: n_ch // cf] = w[0::cf]
    fused = oracle.voltage_beamform(payloads, wb, n_ant, n_beam, n_ch, n_time, n_tps, scale=0.05, ch_fuse=cf)
    n_out = n_time // n_tps
    assert fused.shape == (n_ch // cf, n_beam, npl * cf * n_out)
    for p in range(npl):
        single = oracle.voltage_beamform([payloads[p]], w, n_ant, n_beam, n_ch, n_time, n_tps, scale=0.05)
        for b in range(n_ch // cf):
            for c in range(cf):
                got = fused[b, :, (p * cf + c) * n_out:(p * cf + c + 1) * n_out]
                assert np.array_equal(got, single[b * cf + c])


def test_quantization_rule(oracle):
    """rint (nearest-even) of acc * scale, clamped to [0, 255]; NaN -> 0, inf -> 255."""
    re = np.zeros((1, 2, 8, 1), np.float32)
    im = np.zeros_like(re)
    re[0, 0, :, 0] = [1, 1, 1, 1, 1, np.nan, np.inf, 1]
    for scale, want in [(2.5, 2), (3.5, 4), (0.49, 0), (254.5, 254), (255.5, 255), (300, 255), (-3, 0)]:
        out = oracle.voltage_tail(re, im, 1, 0, scale, True)[0, 0]
        assert out[0] == want, (scale, out[0])
    out = oracle.voltage_tail(re, im, 1, 0, 1.0, True)[0, 0]
    assert out[5] == 0 and out[6] == 255


def test_fp6_tail_equals_fp8_tail_on_exact_beams(oracle):
    rng = np.random.default_rng(4)
    re = rng.integers(-20, 20, (2, 2, 12, 5)).astype(np.float32)
    im = rng.integers(-20, 20, (2, 2, 12, 5)).astype(np.float32)
    a = oracle.voltage_tail(re, im, 4, 0, 1.0, False)
    b = oracle.voltage_tail(re, im, 4, 1, 1.0, False)
    assert np.array_equal(a, b)
    want = (re.astype(np.float64) ** 2 + im.astype(np.float64) ** 2).sum(axis=1)  # [b, M, beam]
    want = want.reshape(2, 3, 4, 5).sum(axis=2).transpose(0, 2, 1)
    assert np.array_equal(a.astype(np.float64), want)


def test_exactness_guard_rejects_rounding_cases(oracle):
    """A case whose FP32 partial sums could round is refused, not silently compared."""
    n_ant, n_beam, n_ch, n_time = 64, 2, 1, 16
    (qc,), _ = _case(oracle, 5, n_ant, n_beam, n_ch, n_time)
    w = np.full((n_ch, n_beam, n_ant, 2), 448.0, np.float32)
    w[0, 0, 0, 0] = 2.0 ** -9  # finest E4M3 step next to large weights
    W = oracle.voltage_weights(w, n_ch, n_beam, n_ant, "fp8")
    with pytest.raises(AssertionError, match="not an exact case"):
        oracle.voltage_beams([qc], W, n_ant, n_beam, n_ch, n_time)
