"""GPU tests of differentiable gain calibration (FP32 kernels) against the FP64
oracle and FP64 central finite differences (h = 1e-4).

Forward: M = X X^H from QC voltages (exact FP8 HERK), gains
g_i = (1 + 0.1 N(0,1)) exp(i pi U(-1, 1)), observed V = g_t g_t^H o M + noise;
L = sum_{i>=j} |V_ij - g_i conj(g_j) M_ij|^2.  Backward: grad_g, Gamma = dL/dM,
and the HERK adjoint dX = (Gamma + Gamma^H) X.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gains(rng, shape):
    return (1 + 0.1 * rng.standard_normal(shape)) * np.exp(1j * np.pi * rng.uniform(-1, 1, shape))


def _setup(oracle, dev, N, K, B, seed):
    import torch
    from ftd_amd import ftdx
    A = oracle.gen_qc_uniform(seed, N * K * B)
    Ad = torch.from_numpy(A).to(dev)
    M = ftdx.herk_int4(Ad, N, K, B)  # exact
    torch.cuda.synchronize()
    Mn = M.cpu().numpy().astype(np.float64)
    Mc = Mn[..., 0] + 1j * Mn[..., 1]
    rng = np.random.default_rng(seed)
    il, jl = np.tril_indices(N)
    gt, g = _gains(rng, (B, N)), _gains(rng, (B, N))
    V = gt[:, il] * gt[:, jl].conj() * Mc + 20.0 * (rng.standard_normal(Mc.shape)
                                                    + 1j * rng.standard_normal(Mc.shape))
    V[:, il == jl] = V[:, il == jl].real
    # numpy keeps the operands' memory order ('K'): force C order for the device copies
    V32 = np.ascontiguousarray(np.stack([V.real, V.imag], -1), dtype=np.float32)
    g32 = np.ascontiguousarray(np.stack([g.real, g.imag], -1), dtype=np.float32)
    # the FP64 references use exactly the FP32 values the GPU sees
    V = V32[..., 0].astype(np.float64) + 1j * V32[..., 1]
    g = g32[..., 0].astype(np.float64) + 1j * g32[..., 1]
    X = oracle.decode_qc_array(A).reshape(B, N, K)
    return Ad, M, torch.from_numpy(V32).to(dev), torch.from_numpy(g32).to(dev), X, Mc, V, g


def _c(t):
    a = t.cpu().numpy().astype(np.float64)
    return a[..., 0] + 1j * a[..., 1]


@pytest.mark.parametrize("N,K,B", [(100, 96, 2), (33, 70, 1), (130, 200, 2), (257, 64, 1)])
def test_gain_and_adjoint_vs_fp64_oracle(oracle, dev, N, K, B):
    import torch
    from ftd_amd import ftdx
    Ad, M, Vd, gd, X, Mc, V, g = _setup(oracle, dev, N, K, B, 100 + N)
    ctx = ftdx.CutlassComplexGemm()
    loss, grad, Gamma = ftdx.gain_loss_grad(ctx, Vd, M, gd, N, B)
    dX = ftdx.herk_adjoint(ctx, Gamma, Ad, N, K, B)
    torch.cuda.synchronize()
    l64, g64, G64 = oracle.gain_loss_grad_f64(V, Mc, g, N, B)
    assert np.allclose(loss.cpu().numpy(), l64, rtol=1e-5)
    assert np.abs(_c(grad) - g64).max() <= 1e-4 * np.abs(g64).max()
    assert np.abs(_c(Gamma) - G64).max() <= 1e-5 * np.abs(G64).max()
    d64 = oracle.herk_adjoint_f64(G64, X, N, K, B)
    assert np.abs(_c(dX) - d64).max() <= 1e-4 * np.abs(d64).max()
    # the GEMM alone: the oracle adjoint fed the GPU's own Gamma
    d64g = oracle.herk_adjoint_f64(_c(Gamma), X, N, K, B)
    assert np.abs(_c(dX) - d64g).max() <= 1e-5 * np.abs(d64g).max()


@pytest.mark.parametrize("N,B", [(200, 2), (130, 1)])
def test_adjoint_identity_voltages_exact(dev, N, B):
    """X = I (K = N): dX = Gamma + Gamma^H exactly.  The adjoint splits each FP32 value
    into three bf16 terms whose sum is exact, so every bit of a random 24-bit significand
    must come back; a missing or rounded term fails."""
    import torch
    from ftd_amd import ftdx
    rng = np.random.default_rng(N)
    P = N * (N + 1) // 2
    G = rng.standard_normal((B, P, 2)).astype(np.float32) * np.float32(3.7)
    X = np.zeros((B, N, N), dtype=np.uint8)
    X[:, np.arange(N), np.arange(N)] = 0x10  # Re = +1, Im = 0 (high nibble Re)
    ctx = ftdx.CutlassComplexGemm()
    dX = ftdx.herk_adjoint(ctx, torch.from_numpy(G).to(dev), torch.from_numpy(X.reshape(-1)).to(dev), N, N, B)
    torch.cuda.synchronize()
    il, jl = np.tril_indices(N)
    H = np.zeros((B, N, N, 2), dtype=np.float32)
    H[:, il, jl] = G
    H[:, jl, il, 0] = G[..., 0]
    H[:, jl, il, 1] = -G[..., 1]
    d = np.arange(N)
    H[:, d, d, 0] = 2 * H[:, d, d, 0]
    H[:, d, d, 1] = 0
    assert np.array_equal(dX.cpu().numpy().reshape(B, N, N, 2), H)


def test_gpu_gradients_vs_central_differences(oracle, dev):
    """FP64 central differences of the forward loss at a sample of coordinates."""
    import torch
    from ftd_amd import ftdx
    N, K, B = 48, 40, 1
    Ad, M, Vd, gd, X, Mc, V, g = _setup(oracle, dev, N, K, B, 7)
    ctx = ftdx.CutlassComplexGemm()
    loss, grad, Gamma = ftdx.gain_loss_grad(ctx, Vd, M, gd, N, B)
    dX = ftdx.herk_adjoint(ctx, Gamma, Ad, N, K, B)
    torch.cuda.synchronize()
    il, jl = np.tril_indices(N)

    def L(gg, XX):
        Mx = XX[0] @ XX[0].conj().T
        Mp = Mx[il, jl].copy()
        Mp[il == jl] = Mp[il == jl].real
        return (np.abs(V[0] - gg[0][il] * gg[0][jl].conj() * Mp) ** 2).sum()

    h = 1e-4
    gg, dd = _c(grad), _c(dX)
    gscale, dscale = np.abs(gg).max(), np.abs(dd).max()
    rng = np.random.default_rng(0)
    for k in rng.choice(N, 8, replace=False):
        for unit, part in ((1.0, "re"), (1j, "im")):
            gp, gm = g.copy(), g.copy()
            gp[0, k] += h * unit
            gm[0, k] -= h * unit
            fd = (L(gp, X) - L(gm, X)) / (2 * h)
            an = gg[0, k].real if part == "re" else gg[0, k].imag
            assert abs(fd - an) <= 2e-4 * gscale, (k, part, fd, an)
    for _ in range(8):
        i, k = int(rng.integers(N)), int(rng.integers(K))
        for unit, part in ((1.0, "re"), (1j, "im")):
            Xp, Xm = X.copy(), X.copy()
            Xp[0, i, k] += h * unit
            Xm[0, i, k] -= h * unit
            fd = (L(g, Xp) - L(g, Xm)) / (2 * h)
            an = dd[0, i, k].real if part == "re" else dd[0, i, k].imag
            assert abs(fd - an) <= 2e-4 * dscale, (i, k, part, fd, an)


def test_gain_grad_deterministic(oracle, dev):
    import torch
    from ftd_amd import ftdx
    N, K, B = 300, 64, 2
    Ad, M, Vd, gd, *_ = _setup(oracle, dev, N, K, B, 3)
    ctx = ftdx.CutlassComplexGemm()
    l1, g1, G1 = ftdx.gain_loss_grad(ctx, Vd, M, gd, N, B)
    l2, g2, G2 = ftdx.gain_loss_grad(ctx, Vd, M, gd, N, B)
    torch.cuda.synchronize()
    assert torch.equal(l1, l2) and torch.equal(g1, g2) and torch.equal(G1, G2)


def test_calibration_full_size_vs_fp64(dev):
    """1024 inputs x 16 channels x K=2048: GPU FP32 gradients vs an FP64 torch
    reference on the device, and FP64 central differences at sample coordinates."""
    import torch
    from ftd_amd import ftdx
    N, K, B = 1024, 2048, 16
    gen = torch.Generator(device=dev).manual_seed(5)
    v = torch.randint(-7, 8, (2, B * N * K), generator=gen, device=dev, dtype=torch.int16)
    nib = torch.where(v < 0, 8 + (-v), v).to(torch.uint8)
    Ad = ((nib[0] << 4) | nib[1]).contiguous()
    Xc = torch.complex(v[0].double(), v[1].double()).reshape(B, N, K)
    del v, nib
    M = ftdx.herk_int4(Ad, N, K, B)
    ii, jj = torch.tril_indices(N, N, device=dev)
    diag = ii == jj
    Mc = torch.view_as_complex(M.double())
    rng = np.random.default_rng(5)
    gt = torch.from_numpy(_gains(rng, (B, N))).to(dev)
    g = torch.from_numpy(_gains(rng, (B, N))).to(dev)
    noise = torch.complex(torch.randn(Mc.shape, generator=gen, device=dev, dtype=torch.float64),
                          torch.randn(Mc.shape, generator=gen, device=dev, dtype=torch.float64))
    V = gt[:, ii] * gt[:, jj].conj() * Mc + 50.0 * noise
    V[:, diag] = V[:, diag].real.to(V.dtype)
    V32 = torch.view_as_real(V).float().contiguous()
    g32 = torch.view_as_real(g).float().contiguous()
    V = torch.view_as_complex(V32.double())
    g = torch.view_as_complex(g32.double())
    ctx = ftdx.CutlassComplexGemm()
    loss, grad, Gamma = ftdx.gain_loss_grad(ctx, V32, M, g32, N, B)
    dX = ftdx.herk_adjoint(ctx, Gamma, Ad, N, K, B)
    torch.cuda.synchronize()
    # FP64 reference
    R = V - g[:, ii] * g[:, jj].conj() * Mc
    L64 = (R.abs() ** 2).sum(-1)
    assert torch.allclose(loss, L64, rtol=1e-5)
    G64 = -2.0 * R * g[:, ii].conj() * g[:, jj]
    rowt = -2.0 * R * Mc.conj() * g[:, jj]
    colt = -2.0 * R.conj() * Mc * g[:, ii]
    acc = torch.zeros((B, N, 2), dtype=torch.float64, device=dev)
    acc.index_add_(1, ii, torch.view_as_real(rowt))
    acc.index_add_(1, jj, torch.view_as_real(colt))
    grad64 = torch.view_as_complex(acc)
    gerr = (torch.view_as_complex(grad.double()) - grad64).abs().max() / grad64.abs().max()
    assert gerr < 1e-4, float(gerr)
    H = torch.zeros((B, N, N), dtype=torch.complex128, device=dev)
    H[:, ii, jj] = G64
    H = H + H.conj().transpose(1, 2)
    dX64 = torch.matmul(H, Xc)
    derr = (torch.view_as_complex(dX.double()) - dX64).abs().max() / dX64.abs().max()
    assert derr < 1e-4, float(derr)
    # FP64 central differences of the loss in g for channel 0 at a few coordinates
    h = 1e-4
    for k in (0, 511, 1023):
        for unit in (1.0, 1j):
            def Lg(delta):
                gg = g[0].clone()
                gg[k] += delta
                r = V[0] - gg[ii] * gg[jj].conj() * Mc[0]
                return float((r.abs() ** 2).sum())
            fd = (Lg(h * unit) - Lg(-h * unit)) / (2 * h)
            an = float(grad[0, k, 0] if unit == 1.0 else grad[0, k, 1])
            assert abs(fd - an) <= 1e-4 * float(grad64.abs().max()), (k, unit, fd, an)
