"""CPU pinning of the oracle's pillbox gridding and polarization sum against closed
forms (known UV cells, conjugate placement, auto-correlations, out-of-grid drops),
linearity and a planted u/v swap."""
import numpy as np

C = 299792458.0


def _uv_for_cell(iu, iv, Ng, cell, freq):
    """Baseline in metres whose UV lands exactly on cell (iu, iv) at this frequency."""
    lam = C / freq
    return np.array([(iu - Ng / 2) * cell * lam, (iv - Ng / 2) * cell * lam], dtype=np.float32)


def test_known_cells_and_conjugate(oracle):
    Ng, cell, freq = 32, 1e-3, 1.4e9
    N = 3                                   # baselines (0,0) (1,0) (1,1) (2,0) (2,1) (2,2)
    n_bl = N * (N + 1) // 2
    uv = np.zeros((n_bl, 2), np.float32)
    uv[1] = _uv_for_cell(20, 11, Ng, cell, freq)  # cross (1, 0)
    uv[3] = _uv_for_cell(5, 30, Ng, cell, freq)   # cross (2, 0)
    uv[2] = _uv_for_cell(16, 16, Ng, cell, freq)  # auto (1, 1) at the centre
    uv[0] = _uv_for_cell(40, 3, Ng, cell, freq)   # auto (0, 0) off the grid: dropped
    uv[4] = _uv_for_cell(0, 0, Ng, cell, freq)    # cross (2, 1): conjugate at (32, 32) is off-grid
    uv[5] = _uv_for_cell(17, 16, Ng, cell, freq)  # auto (2, 2)
    vis = np.zeros((1, n_bl, 2), np.float32)
    vis[0, :, 0] = [1, 2, 3, 4, 5, 6]
    vis[0, :, 1] = [0, 20, 0, 40, 50, 0]
    g = oracle.pillbox_grid(vis, uv, np.array([freq]), n_bl, Ng, 1, cell)
    want = np.zeros((1, Ng, Ng, 2), np.float32)
    want[0, 11, 20] += (2, 20)
    want[0, 32 - 11, 32 - 20] += (2, -20)  # (-u, -v) -> Ng - i
    want[0, 30, 5] += (4, 40)
    want[0, 2, 27] += (4, -40)
    want[0, 16, 16] += (3, 0)
    want[0, 0, 0] += (5, 50)
    want[0, 16, 17] += (6, 0)
    assert np.array_equal(g, want)


def test_linearity_and_frequency_scaling(oracle):
    rng = np.random.default_rng(1)
    N, Ng, Nf, cell = 20, 64, 3, 2e-3
    n_bl = N * (N + 1) // 2
    uv = (rng.standard_normal((n_bl, 2)) * 4).astype(np.float32)
    freqs = np.array([1.0e9, 1.2e9, 1.4e9, 1.6e9])
    a = rng.integers(-50, 50, (Nf, n_bl, 2)).astype(np.float32)
    b = rng.integers(-50, 50, (Nf, n_bl, 2)).astype(np.float32)
    ga = oracle.pillbox_grid(a, uv, freqs, n_bl, Ng, Nf, cell, freq_offset=1)
    gb = oracle.pillbox_grid(b, uv, freqs, n_bl, Ng, Nf, cell, freq_offset=1)
    gab = oracle.pillbox_grid(a + b, uv, freqs, n_bl, Ng, Nf, cell, freq_offset=1)
    assert np.array_equal(ga + gb, gab)  # integer sums: exact


def test_cross_only_grid_is_hermitian(oracle):
    """Without auto-correlations every visibility also lands, conjugated, at (-u, -v):
    G[Ng - iv, Ng - iu] = conj(G[iv, iu]) for cells away from the edge (random UVs
    make exact rounding ties, where the mirror breaks, improbable)."""
    rng = np.random.default_rng(3)
    N, Ng, cell, freq = 30, 64, 2e-3, 1.0e9
    n_bl = N * (N + 1) // 2
    uv = (rng.standard_normal((n_bl, 2)) * 0.02).astype(np.float32)
    vis = rng.integers(-9, 10, (1, n_bl, 2)).astype(np.float32)
    il, jl = np.tril_indices(N)
    vis[0, il == jl] = 0
    g = oracle.pillbox_grid(vis, uv, np.array([freq]), n_bl, Ng, 1, cell)[0]
    inner = g[1:, 1:]
    mirror = g[1:, 1:][::-1, ::-1]
    assert np.array_equal(inner[..., 0], mirror[..., 0])
    assert np.array_equal(inner[..., 1], -mirror[..., 1])
    assert np.abs(g).sum() > 0


def test_planted_uv_swap_is_caught(oracle):
    Ng, cell, freq = 16, 1e-3, 1e9
    uv = np.array([[0, 0], _uv_for_cell(3, 12, Ng, cell, freq)], np.float32)
    vis = np.array([[[0, 0], [1, 0]]], np.float32)
    g = oracle.pillbox_grid(vis, uv, np.array([freq]), 2, Ng, 1, cell)
    assert g[0, 12, 3, 0] == 1 and g[0, 3, 12, 0] == 0  # row index is v


def test_pol_reduce(oracle):
    rng = np.random.default_rng(2)
    n_ch, n_pol, n_ti, P = 3, 2, 4, 10
    t = rng.integers(-100, 100, (n_ch * n_pol * n_ti, P, 2)).astype(np.float32)
    r = oracle.pol_reduce(t, n_ch, n_pol, n_ti)
    x = t.reshape(n_ch, n_pol, n_ti, P, 2)
    assert np.array_equal(r.reshape(n_ch, n_ti, P, 2), x[:, 0] + x[:, 1])
