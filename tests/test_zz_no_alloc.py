"""GPU test of ftdx_gain_loss_grad_noalloc, the allocation-free gain entry point: its
outputs must equal ftdx_gain_loss_grad's bit for bit (the same kernels, fixed-order
reductions), on two calls in a row, the second with a smaller input, so the fixed
scratch's leftover contents from the first call cannot leak into the second."""
import pytest

pytestmark = pytest.mark.gpu


def _case(torch, dev, N, B, seed):
    gen = torch.Generator(device="cpu").manual_seed(seed)
    P = N * (N + 1) // 2
    V = torch.randn((B, P, 2), generator=gen).to(dev)
    M = torch.randn((B, P, 2), generator=gen).to(dev)
    g = torch.randn((B, N, 2), generator=gen).to(dev)
    return V, M, g


def _outputs(torch, dev, N, B, gamma_like):
    return (torch.full((B,), float("nan"), dtype=torch.float64, device=dev),
            torch.full((B, N, 2), float("nan"), dtype=torch.float32, device=dev),
            torch.full_like(gamma_like, float("nan")))


def test_noalloc_matches_allocating_entry(dev):
    import torch
    from ftd_amd import ftdx
    ctx = ftdx.CutlassComplexGemm()
    for N, B, seed in ((1024, 16, 1), (300, 3, 2)):  # the second call is smaller
        V, M, g = _case(torch, dev, N, B, seed)
        loss_r, grad_r, gam_r = ftdx.gain_loss_grad(ctx, V, M, g, N, B)
        loss, grad, gam = _outputs(torch, dev, N, B, M)
        ftdx.gain_loss_grad_noalloc(V, M, g, N, B, loss, grad, gam)
        torch.cuda.synchronize()
        assert torch.equal(loss.view(torch.int64), loss_r.view(torch.int64)), (N, B)
        assert torch.equal(grad.view(torch.int32), grad_r.view(torch.int32)), (N, B)
        assert torch.equal(gam.view(torch.int32), gam_r.view(torch.int32)), (N, B)
    # without Gamma
    loss, grad, _ = _outputs(torch, dev, N, B, M)
    ftdx.gain_loss_grad_noalloc(V, M, g, N, B, loss, grad, None)
    torch.cuda.synchronize()
    assert torch.equal(grad.view(torch.int32), grad_r.view(torch.int32))
    assert torch.equal(loss.view(torch.int64), loss_r.view(torch.int64))


def test_noalloc_rejects_sizes_past_the_fixed_scratch(dev):
    import torch
    from ftd_amd import ftdx
    N, B = 2048, 16  # 2 * 16 * 64 * 2048 * 8 B = 33.6 MB of partials > 16 MiB
    V = torch.zeros((B, N * (N + 1) // 2, 2), dtype=torch.float32, device=dev)
    g = torch.zeros((B, N, 2), dtype=torch.float32, device=dev)
    loss = torch.zeros(B, dtype=torch.float64, device=dev)
    grad = torch.zeros((B, N, 2), dtype=torch.float32, device=dev)
    with pytest.raises(ftdx.FtdxError, match="fixed"):
        ftdx.gain_loss_grad_noalloc(V, V, g, N, B, loss, grad, None)
