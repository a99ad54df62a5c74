"""GPU parity of the GGP XEngine (ftdx_xengine_*) and the FTD raw-order correlation
(ftdx_correlate_ftd) through the C ABI.  Integer voltages make every sum exact in
FP32 (below 2^24) and FP64, so all comparisons are bit-exact: against the oracle
at small sizes, and against an FP64 torch matmul at the reference test's own size
(tests/xengine_test.cpp:34-57: 2048 antennas, 1024 times, 4 channels, 2 pols,
time_inner 2; the reference checks deviation == 0, xengine_test_gtest.hpp)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _param(ftdx, n_ant, n_chan, n_time, n_ti, n_pol, prec, fmt, n_payload=1, loc=1,
           engine=0):
    return ftdx.new_xengine_param(
        in_location=loc, out_location=loc, packet_format=ftdx.PACKET_DSA2K,
        n_payload=n_payload, n_antennae_per_payload=n_ant, n_channels_per_payload=n_chan,
        n_time_per_payload=n_time, n_time_inner=n_ti, n_polarizations=n_pol,
        compute_prec=prec, output_prec=prec, engine=engine, format=fmt)


def _oracle_payloads(oracle, data, xe, n_payload, herm):
    per = xe.in_payload_bytes
    outs = [oracle.xengine_ggp_i64(data[i * per:(i + 1) * per], xe.N, xe.K, xe.batch, herm)
            for i in range(n_payload)]
    return np.stack(outs)


SHAPES = [(99, 3, 96, 2, 2), (64, 2, 64, 1, 2), (33, 1, 50, 2, 1), (128, 1, 7, 2, 2)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("prec", ["single", "double"])
@pytest.mark.parametrize("fmt", ["tri", "herm"])
def test_xengine_host_buffers_match_oracle(oracle, dev, shape, prec, fmt):
    """The reference test's flow: host input and output, n_payload = 2
    (xengine_test.cpp:343-372, lib/xengine.cpp:116-121, 166-174)."""
    from ftd_amd import ftdx
    n_ant, n_chan, n_time, n_ti, n_pol = shape
    qp = ftdx.QPREC_SINGLE if prec == "single" else ftdx.QPREC_DOUBLE
    mf = ftdx.XENGINE_MAT_TRI if fmt == "tri" else ftdx.XENGINE_MAT_HERM
    xe = ftdx.XEngine(_param(ftdx, n_ant, n_chan, n_time, n_ti, n_pol, qp, mf, n_payload=2))
    data = oracle.gen_bytes(n_ant + n_time, 2 * xe.in_payload_bytes)  # -8 nibbles included
    dt = np.float32 if prec == "single" else np.float64
    out = np.full(2 * xe.out_payload_bytes // np.dtype(dt).itemsize, np.nan, dtype=dt)
    print(f"xengine host {shape} {prec} {fmt}", flush=True)
    xe.compute(out, data)
    ref = _oracle_payloads(oracle, data, xe, 2, fmt == "herm")
    assert np.array_equal(out.reshape(ref.shape), ref.astype(dt))
    xe.close()


@pytest.mark.parametrize("cfg", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("fmt", ["tri", "herm"])
@pytest.mark.parametrize("n_time", [80, 256])  # K = 40 (edge loop only); K = 128 (interior tiles)
def test_xengine_every_tile_config(oracle, dev, cfg, fmt, n_time):
    import torch
    from ftd_amd import ftdx
    n_ant, n_chan, n_ti, n_pol = 300, 2, 2, 1
    mf = ftdx.XENGINE_MAT_TRI if fmt == "tri" else ftdx.XENGINE_MAT_HERM
    xe = ftdx.XEngine(_param(ftdx, n_ant, n_chan, n_time, n_ti, n_pol, ftdx.QPREC_SINGLE, mf,
                             loc=ftdx.LOCATION_DEVICE))
    xe.set_launch_config(cfg, cfg & 1)
    data = oracle.gen_bytes(cfg + 70, xe.in_payload_bytes)
    out = torch.full((xe.out_payload_bytes // 4,), float("nan"), dtype=torch.float32, device=dev)
    print(f"xengine cfg {cfg} {fmt}", flush=True)
    xe.compute(out, torch.from_numpy(data).to(dev))
    ref = _oracle_payloads(oracle, data, xe, 1, fmt == "herm")[0]
    assert np.array_equal(out.cpu().numpy().reshape(ref.shape), ref.astype(np.float32))
    xe.close()


def _torch_gram(data, B, K, N):
    import torch
    d = data.reshape(B, K, N)
    re = ((d << 4).view(torch.int8) >> 4).double()  # sign-extend the low nibble
    im = (d.view(torch.int8) >> 4).double()
    M = torch.complex(re, im)
    MH = M.conj().transpose(1, 2).contiguous()
    return torch.matmul(MH, M)  # FP64: exact for these integer sums


@pytest.mark.parametrize("prec", ["single", "double"])
@pytest.mark.parametrize("fmt", ["tri", "herm"])
def test_xengine_reference_test_size(dev, prec, fmt):
    import torch
    from ftd_amd import ftdx
    n_ant, n_chan, n_time, n_ti, n_pol = 2048, 4, 1024, 2, 2
    qp = ftdx.QPREC_SINGLE if prec == "single" else ftdx.QPREC_DOUBLE
    mf = ftdx.XENGINE_MAT_TRI if fmt == "tri" else ftdx.XENGINE_MAT_HERM
    xe = ftdx.XEngine(_param(ftdx, n_ant, n_chan, n_time, n_ti, n_pol, qp, mf,
                             loc=ftdx.LOCATION_DEVICE))
    assert (xe.N, xe.K, xe.batch) == (2048, 512, 16)
    g = torch.Generator(device=dev).manual_seed(17)
    data = torch.randint(0, 256, (xe.in_payload_bytes,), generator=g, device=dev,
                         dtype=torch.int32).to(torch.uint8)
    dt = torch.float32 if prec == "single" else torch.float64
    out = torch.empty(xe.out_payload_bytes // (4 if prec == "single" else 8), dtype=dt, device=dev)
    print(f"xengine reference size {prec} {fmt}", flush=True)
    xe.compute(out, data)
    C = _torch_gram(data, xe.batch, xe.K, xe.N)
    if fmt == "tri":
        iu, ju = torch.triu_indices(xe.N, xe.N, device=dev)
        C = C[:, iu, ju]
    ref = torch.view_as_real(C).to(dt).reshape(-1)
    assert torch.equal(out, ref)
    xe.close()


def test_xengine_fp64_long_k_is_exact(oracle, dev):
    """K = 200000 > 131072 (the FP32-exact limit for [-8, 7] nibbles): FP64 output
    accumulates K chunks in double and stays exact."""
    import torch
    from ftd_amd import ftdx
    n_ant, n_chan, n_time, n_ti, n_pol = 8, 1, 400000, 2, 1
    xe = ftdx.XEngine(_param(ftdx, n_ant, n_chan, n_time, n_ti, n_pol, ftdx.QPREC_DOUBLE,
                             ftdx.XENGINE_MAT_HERM, loc=ftdx.LOCATION_DEVICE))
    data = np.full(xe.in_payload_bytes, 0x88, np.uint8)  # mostly -8 - 8i ...
    rnd = oracle.gen_bytes(3, xe.in_payload_bytes)
    mask = oracle.gen_bytes(4, xe.in_payload_bytes) < 26  # ... 10% random bytes
    data[mask] = rnd[mask]
    out = torch.empty(xe.out_payload_bytes // 8, dtype=torch.float64, device=dev)
    print("xengine fp64 long K", flush=True)
    xe.compute(out, torch.from_numpy(data).to(dev))
    ref = _oracle_payloads(oracle, data, xe, 1, True)[0]
    assert np.abs(ref).max() > 2 ** 24  # the FP32-inexact regime
    assert np.array_equal(out.cpu().numpy().reshape(ref.shape), ref.astype(np.float64))
    xe.close()


def test_xengine_tcc_engine_ftd_order(oracle, dev):
    """Engine TCC: QC bytes in FTD order [time][ant][chan][pol] -> packed lower
    triangle FP32 per (chan, pol, ti) (lib/xengine.cpp:130-150)."""
    from ftd_amd import ftdx
    n_ant, n_chan, n_time, n_ti, n_pol = 70, 3, 64, 2, 2
    xe = ftdx.XEngine(_param(ftdx, n_ant, n_chan, n_time, n_ti, n_pol, ftdx.QPREC_SINGLE,
                             ftdx.XENGINE_MAT_TRI, engine=ftdx.BLAS_ENGINE_TCC))
    raw = oracle.gen_bytes(9, xe.in_payload_bytes)  # QC bytes, -0 nibbles included
    out = np.empty(xe.out_payload_bytes // 4, dtype=np.float32)
    print("xengine tcc engine", flush=True)
    xe.compute(out, raw)
    ref = oracle.herk_ftd_i64(raw, n_ant, n_chan, n_time, n_ti, n_pol, twos=False)
    assert np.array_equal(out.reshape(ref.shape), ref.astype(np.float32))
    xe.close()


@pytest.mark.parametrize("shape", [(96, 4, 128, 2, 2), (33, 3, 50, 1, 1), (130, 1, 33, 3, 2)])
@pytest.mark.parametrize("twos", [True, False])
def test_correlate_ftd_matches_oracle(oracle, dev, shape, twos):
    import torch
    from ftd_amd import ftdx
    n_ant, n_chan, n_time, n_ti, n_pol = shape
    raw = oracle.gen_bytes(n_ant * 3 + n_time, n_time * n_ant * n_chan * n_pol)
    g = ftdx.CutlassComplexGemm()
    rd = torch.from_numpy(raw).to(dev)
    print(f"correlate_ftd {shape} twos={twos}", flush=True)
    out = ftdx.correlate_ftd(g, rd, n_ant, n_chan, n_time, n_ti, n_pol, twos=twos)
    torch.cuda.synchronize()
    ref = oracle.herk_ftd_i64(raw, n_ant, n_chan, n_time, n_ti, n_pol, twos=twos)
    assert np.array_equal(out.cpu().numpy(), ref.astype(np.float32))
    # the device input is left as it was (the reference converts it in place)
    assert np.array_equal(rd.cpu().numpy(), raw)


def test_xengine_chpc_one_shot(oracle, dev):
    from ftd_amd import ftdx
    p = _param(ftdx, 40, 2, 32, 2, 2, ftdx.QPREC_SINGLE, ftdx.XENGINE_MAT_TRI)
    data = oracle.gen_bytes(21, 40 * 2 * 32 * 2)
    out = np.empty(2 * 2 * 2 * (40 * 41 // 2) * 2, dtype=np.float32)
    ftdx.xengine_chpc(out, data, p)
    ref = oracle.xengine_ggp_i64(data, 40, 16, 8, herm=False)
    assert np.array_equal(out.reshape(ref.shape), ref.astype(np.float32))
