"""CPU pinning of the FP64 gain-calibration oracle by central finite differences.

The oracle's analytic gradients (gain gradient, visibility gradient Gamma, HERK
adjoint dX = (Gamma + Gamma^H) X) are checked against FP64 central differences
(h = 1e-4) of the loss
    L(g, X) = sum_{i>=j} |V_ij - g_i conj(g_j) (X X^H)_ij|^2,
evaluated by a plain numpy forward model (no gradient code involved).
"""
import numpy as np
import pytest


def _packed(G, N):
    il, jl = np.tril_indices(N)
    return G[..., il, jl]


def _forward_loss(V, X, g):
    """V packed [B, P] complex; X [B, N, K]; g [B, N]; plain FP64 numpy forward."""
    B, N, _ = X.shape
    L = np.zeros(B)
    il, jl = np.tril_indices(N)
    for b in range(B):
        M = X[b] @ X[b].conj().T
        Mp = M[il, jl].copy()
        Mp[il == jl] = Mp[il == jl].real  # packed diagonal has Im = 0
        R = V[b] - g[b][il] * g[b][jl].conj() * Mp
        L[b] = (np.abs(R) ** 2).sum()
    return L


def _case(seed, N, K, B):
    rng = np.random.default_rng(seed)
    X = (rng.integers(-7, 8, (B, N, K)) + 1j * rng.integers(-7, 8, (B, N, K))).astype(complex)
    il, jl = np.tril_indices(N)
    M = np.stack([_packed(X[b] @ X[b].conj().T, N) for b in range(B)])
    M[:, il == jl] = M[:, il == jl].real
    gt = (1 + 0.1 * rng.standard_normal((B, N))) * np.exp(1j * np.pi * rng.uniform(-1, 1, (B, N)))
    g = (1 + 0.1 * rng.standard_normal((B, N))) * np.exp(1j * np.pi * rng.uniform(-1, 1, (B, N)))
    V = gt[:, il] * gt[:, jl].conj() * M + 5.0 * (rng.standard_normal(M.shape)
                                                  + 1j * rng.standard_normal(M.shape))
    V[:, il == jl] = V[:, il == jl].real
    return X, M, V, g


def test_gain_gradient_vs_central_differences(oracle):
    N, K, B = 12, 9, 2
    X, M, V, g = _case(1, N, K, B)
    loss, grad, _ = oracle.gain_loss_grad_f64(V, M, g, N, B)
    assert np.allclose(loss, _forward_loss(V, X, g), rtol=1e-12)
    h = 1e-4
    # central differences are O(h^2): the loss is quartic in g with |M| ~ 10^2..10^3,
    # so the truncation term is ~1e-5 of the gradient scale
    scale = np.abs(grad).max()
    for b in range(B):
        for k in range(N):
            for unit in (1.0, 1j):
                gp, gm = g.copy(), g.copy()
                gp[b, k] += h * unit
                gm[b, k] -= h * unit
                fd = (_forward_loss(V, X, gp)[b] - _forward_loss(V, X, gm)[b]) / (2 * h)
                an = grad[b, k].real if unit == 1.0 else grad[b, k].imag
                assert abs(fd - an) <= 2e-5 * scale, (b, k, unit, fd, an)


def test_herk_adjoint_vs_central_differences(oracle):
    N, K, B = 10, 7, 2
    X, M, V, g = _case(2, N, K, B)
    _, _, Gamma = oracle.gain_loss_grad_f64(V, M, g, N, B)
    dX = oracle.herk_adjoint_f64(Gamma, X, N, K, B)
    h = 1e-4
    scale = np.abs(dX).max()
    for b in range(B):
        for i in range(N):
            for k in range(K):
                for unit in (1.0, 1j):
                    Xp, Xm = X.copy(), X.copy()
                    Xp[b, i, k] += h * unit
                    Xm[b, i, k] -= h * unit
                    fd = (_forward_loss(V, Xp, g)[b] - _forward_loss(V, Xm, g)[b]) / (2 * h)
                    an = dX[b, i, k].real if unit == 1.0 else dX[b, i, k].imag
                    assert abs(fd - an) <= 2e-5 * scale, (b, i, k, unit, fd, an)


def test_adjoint_closed_form(oracle):
    """Gamma = identity on the diagonal (Re 1): H = 2I, so dX = 2X."""
    N, K = 6, 5
    rng = np.random.default_rng(3)
    X = rng.standard_normal((1, N, K)) + 1j * rng.standard_normal((1, N, K))
    Gamma = np.zeros((1, N * (N + 1) // 2), dtype=complex)
    d = np.arange(N) * (np.arange(N) + 1) // 2 + np.arange(N)
    Gamma[0, d] = 1.0 + 0.7j  # the imaginary part of the diagonal must not matter
    assert np.allclose(oracle.herk_adjoint_f64(Gamma, X, N, K, 1), 2 * X, atol=1e-14)


@pytest.mark.parametrize("planted", ["swap_conj", "drop_column_term"])
def test_planted_gradient_mistakes_fail_fd(oracle, planted):
    """A plausible mistake in the gradient (conjugate on the wrong factor, a
    dropped column term) disagrees with the finite differences."""
    N, K, B = 8, 6, 1
    X, M, V, g = _case(4, N, K, B)
    _, grad, _ = oracle.gain_loss_grad_f64(V, M, g, N, B)
    bad = grad.conj() if planted == "swap_conj" else grad * 0.5
    h = 1e-4
    worst = 0.0
    for k in range(N):
        gp, gm = g.copy(), g.copy()
        gp[0, k] += h
        gm[0, k] -= h
        fd_r = (_forward_loss(V, X, gp)[0] - _forward_loss(V, X, gm)[0]) / (2 * h)
        gp, gm = g.copy(), g.copy()
        gp[0, k] += 1j * h
        gm[0, k] -= 1j * h
        fd_i = (_forward_loss(V, X, gp)[0] - _forward_loss(V, X, gm)[0]) / (2 * h)
        worst = max(worst, abs(fd_r + 1j * fd_i - bad[0, k]) / max(1.0, abs(bad[0, k])))
    assert worst > 1e-3
