"""GPU parity of the streaming correlation of host payloads (ftdx_pipeline_*), the
counterpart of VisibilityPipeline::correlate / correlate_pipelined
(lib/visibility_pipeline.cu:253-300, 484-545): every payload's output bit-exact
against the oracle, in order, with the double-buffered stages overlapping."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view({4: np.uint32, 2: np.uint16}[a.dtype.itemsize])


@pytest.mark.parametrize("output", ["fp32", "fp16"])
@pytest.mark.parametrize("N,K,B,n", [(128, 256, 4, 5), (33, 40, 3, 2), (200, 64, 2, 1)])
def test_pipeline_payloads_match_oracle(oracle, dev, output, N, K, B, n):
    from ftd_amd import ftdx
    g = ftdx.CutlassComplexGemm()
    pl = ftdx.Pipeline(g, N, K, B, output=output)
    ins = [oracle.gen_bytes(1000 * N + p, N * K * B) for p in range(n)]
    print(f"pipeline {output} N={N} K={K} B={B} payloads={n}", flush=True)
    outs = pl.correlate_many(ins)
    for p in range(n):
        ref = (oracle.herk_f32(ins[p], N, K, B) if output == "fp32"
               else oracle.herk_f16(ins[p], N, K, B))
        assert np.array_equal(_bits(outs[p]), _bits(ref)), p
    # single-payload form, reusing the same buffers afterwards
    one = pl.correlate(ins[-1])
    assert np.array_equal(_bits(one), _bits(outs[-1]))
    pl.close()


def test_pipeline_uses_context_settings(oracle, dev):
    """The context's HERK mode and launch configuration apply inside the pipeline."""
    from ftd_amd import ftdx
    N, K, B = 96, 100, 2
    g = ftdx.CutlassComplexGemm()
    g.set_herk_mode(ftdx.HerkMode.ForceBaseline)
    g.set_launch_config(ftdx.TILE_64x64_K64, 0)
    pl = ftdx.Pipeline(g, N, K, B)
    ins = [oracle.gen_bytes(7 + p, N * K * B) for p in range(3)]
    outs = pl.correlate_many(ins)
    for p in range(3):
        assert np.array_equal(_bits(outs[p]), _bits(oracle.herk_f32(ins[p], N, K, B)))
    assert pl.correlate_many([]) == []
    pl.close()


@pytest.mark.parametrize("pin_in,pin_out", [(True, True), (True, False), (False, True)])
def test_pipeline_page_locked_buffers(oracle, dev, pin_in, pin_out):
    """Page-locked payloads and outputs (torch pin_memory, i.e. hipHostMalloc) skip the
    staging copies; mixed with pageable ones in the same call, every output stays
    bit-exact against the oracle."""
    import torch
    from ftd_amd import ftdx
    N, K, B, n = 128, 256, 4, 5
    g = ftdx.CutlassComplexGemm()
    pl = ftdx.Pipeline(g, N, K, B)
    src = [oracle.gen_bytes(500 + p, N * K * B) for p in range(n)]
    ins, outs, keep = [], [], []
    for p in range(n):
        # every other payload page-locked, so both paths share the double buffers
        pinned = p % 2 == 0
        if pin_in and pinned:
            t = torch.empty(N * K * B, dtype=torch.uint8, pin_memory=True)
            t.numpy()[:] = src[p]
            keep.append(t)
            ins.append(t.numpy())
        else:
            ins.append(src[p])
        if pin_out and pinned:
            t = torch.empty(pl.out_shape(), dtype=torch.float32, pin_memory=True)
            keep.append(t)
            outs.append(t.numpy())
        else:
            outs.append(np.empty(pl.out_shape(), dtype=np.float32))
    print(f"pipeline page-locked in={pin_in} out={pin_out}", flush=True)
    pl.correlate_many(ins, outs)
    for p in range(n):
        assert np.array_equal(_bits(outs[p]), _bits(oracle.herk_f32(src[p], N, K, B))), p
    pl.close()
