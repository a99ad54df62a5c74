"""Synthetic QC voltages and exact self-checks of a packed HERK result on the device.

Used by bench.py to check the very output it times, and by the full-size parity
tests, so that a seed names the same input in both places.

`qc_uniform` draws Re and Im i.i.d. uniform on the integers [-7, 7] and encodes
them as QC bytes (high nibble Re, low nibble Im, sign-magnitude with +0 for 0),
the reference's INT4 input format (cutlass_interface/test_herk_int4.cu:36-45).
The stream is torch's device generator, so a (seed, device type) pair always
yields the same bytes.

The checks run in float64 on the device.  Every quantity involved is an integer
of magnitude below 2^53, so each sum is exact in any order and the comparisons
are equalities, not tolerances:

* weighted checksum -- for small integer weights w,
  sum_{i,j} w_i conj(w_j) C_ij = sum_k |sum_i w_i a_ik|^2,
  with the packed lower triangle standing for the full Hermitian C
  (C_ji = conj C_ij, Im C_ii = 0);
* exact rows -- packed row i of C against sum_k a_ik conj(a_jk) for j <= i,
  computed directly from the QC bytes.
"""
from __future__ import annotations

_SM_LUT = (0, 1, 2, 3, 4, 5, 6, 7, 0, -1, -2, -3, -4, -5, -6, -7)


def qc_uniform(n: int, seed: int, device):
    """n QC bytes with Re, Im i.i.d. uniform integers in [-7, 7] (uint8 tensor on `device`)."""
    import torch
    g = torch.Generator(device=device).manual_seed(seed)
    v = torch.randint(-7, 8, (2, n), generator=g, device=device, dtype=torch.int16)
    nib = torch.where(v < 0, 8 + (-v), v).to(torch.uint8)
    return (nib[0] << 4) | nib[1]


def uneven_problem_list(count: int, seed: int):
    """Shapes of the batched uneven HERK workload: N ~ U{256, 512, 1024, 2048, 4096},
    K ~ U{256, 320, ..., 4096}; returns (Ns, Ks, A offsets, C offsets, A bytes, C floats),
    each problem's rows starting 256-byte aligned and its triangle 256-byte aligned."""
    import numpy as np
    rng = np.random.default_rng(seed)
    Ns = [int(x) for x in rng.choice([256, 512, 1024, 2048, 4096], size=count)]
    Ks = [int(x) * 64 for x in rng.integers(4, 65, size=count)]
    a_off, c_off, oa, oc = [], [], 0, 0
    for N, K in zip(Ns, Ks):
        a_off.append(oa)
        c_off.append(oc)
        oa += (N * K + 255) // 256 * 256
        oc += (N * (N + 1) + 63) // 64 * 64
    return Ns, Ks, a_off, c_off, oa, oc


def _decode(torch, a):
    """QC bytes -> (re, im) float64 tensors."""
    lut = torch.tensor(_SM_LUT, dtype=torch.float64, device=a.device)
    a = a.long()
    return lut[a >> 4], lut[a & 15]


class PackedIndex:
    """Row / column of every packed lower-triangle element of an N x N problem."""

    def __init__(self, torch, N: int, device):
        ii = torch.arange(N, device=device, dtype=torch.int64)
        self.rows = torch.repeat_interleave(ii, ii + 1)
        self.cols = torch.arange(self.rows.numel(), device=device) - self.rows * (self.rows + 1) // 2
        self.offdiag = self.rows != self.cols


def weighted_checksum(torch, A, C, N: int, K: int, seed: int, index: PackedIndex | None = None):
    """(lhs, rhs) of the weighted-checksum identity for one problem.

    A: N*K QC bytes (device), C: packed lower triangle as N(N+1)/2 [re, im] floats.
    Weights are integers in [-3, 3] from a CPU generator seeded with `seed`.
    """
    dev = A.device
    ar, ai = _decode(torch, A.reshape(N, K))
    g = torch.Generator(device="cpu").manual_seed(seed)
    wr = torch.randint(-3, 4, (N,), generator=g).double().to(dev)
    wi = torch.randint(-3, 4, (N,), generator=g).double().to(dev)
    sr = wr @ ar - wi @ ai          # s_k = sum_i w_i a_ik
    si = wr @ ai + wi @ ar
    rhs = float((sr * sr + si * si).sum())
    if index is None:
        index = PackedIndex(torch, N, dev)
    r, c = index.rows, index.cols
    cc = C.reshape(-1, 2).double()
    # Re(w_i conj(w_j) C_ij); the strictly lower part stands for itself and its mirror
    t = (wr[r] * wr[c] + wi[r] * wi[c]) * cc[:, 0] - (wi[r] * wr[c] - wr[r] * wi[c]) * cc[:, 1]
    lhs = float(2.0 * t[index.offdiag].sum() + t[~index.offdiag].sum())
    return lhs, rhs


def exact_rows(torch, A, C, N: int, K: int, rows) -> list[int]:
    """Rows of the packed triangle that differ from the exact product (empty = all equal)."""
    ar, ai = _decode(torch, A.reshape(N, K))
    cc = C.reshape(-1, 2).double()
    bad = []
    for i in rows:
        # C_ij = sum_k a_ik conj(a_jk), j = 0..i
        re = ar[:i + 1] @ ar[i] + ai[:i + 1] @ ai[i]
        im = ar[:i + 1] @ ai[i] - ai[:i + 1] @ ar[i]
        im[i] = 0.0
        s = i * (i + 1) // 2
        got = cc[s:s + i + 1]
        if not (torch.equal(got[:, 0], re) and torch.equal(got[:, 1], im)):
            bad.append(int(i))
    return bad


def check_herk(torch, A, C, N: int, K: int, batch: int, rows=(), row_channels=(), seed0: int = 100) -> dict:
    """Weighted checksums of every channel plus exact rows of some channels.

    A: batch*N*K QC bytes, C: [batch, N(N+1)/2, 2] FP32, both on the device.
    Returns a summary dict; `ok` is True only if every check is an equality.
    """
    A2 = A.reshape(batch, N * K)
    C2 = C.reshape(batch, -1)
    idx = PackedIndex(torch, N, A.device)
    bad_sum = []
    for b in range(batch):
        lhs, rhs = weighted_checksum(torch, A2[b], C2[b], N, K, seed0 + b, idx)
        if lhs != rhs:
            bad_sum.append(b)
    del idx
    bad_rows = {}
    for b in row_channels:
        bad = exact_rows(torch, A2[b], C2[b], N, K, rows)
        if bad:
            bad_rows[int(b)] = bad
    return {
        "ok": not bad_sum and not bad_rows,
        "weighted_checksums_equal": batch - len(bad_sum),
        "channels": batch,
        "exact_rows_checked": len(rows) * len(row_channels),
        "exact_row_channels": [int(b) for b in row_channels],
        "mismatched_checksum_channels": bad_sum[:8],
        "mismatched_rows": bad_rows,
    }


def check_problem_list(torch, A, C, Ns, Ks, a_off, c_off, seed0: int = 1000, rows_of=()) -> dict:
    """Weighted checksum of every problem of a batched list (flat A / C buffers at the
    given offsets), plus exact rows (`rows_of`: (problem, [rows]) pairs)."""
    idx = {}
    bad = []
    for p, (N, K, ao, co) in enumerate(zip(Ns, Ks, a_off, c_off)):
        if N not in idx:
            idx[N] = PackedIndex(torch, N, A.device)
        lhs, rhs = weighted_checksum(torch, A[ao:ao + N * K], C[co:co + N * (N + 1)], N, K, seed0 + p, idx[N])
        if lhs != rhs:
            bad.append(p)
    idx.clear()
    bad_rows = {}
    n_rows = 0
    for p, rows in rows_of:
        N, K, ao, co = Ns[p], Ks[p], a_off[p], c_off[p]
        n_rows += len(rows)
        b = exact_rows(torch, A[ao:ao + N * K], C[co:co + N * (N + 1)], N, K, rows)
        if b:
            bad_rows[int(p)] = b
    return {
        "ok": not bad and not bad_rows,
        "weighted_checksums_equal": len(Ns) - len(bad),
        "problems": len(Ns),
        "exact_rows_checked": n_rows,
        "mismatched_problems": bad[:8],
        "mismatched_rows": bad_rows,
    }
