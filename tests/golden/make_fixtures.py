"""Writes tests/golden/herk_fixtures.npz: small HERK input/output vectors.

Each case's output comes from the C oracle and is cross-checked here, before it
is written, against an independent numpy complex128 product A A^H (lower
triangle, Im(diag) = 0).  The reference ships no golden vectors
(its tests regenerate data from seeds), so these fixtures are the pinned
vectors of this library.  Regenerate with:  python tests/golden/make_fixtures.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402

# (name, N, K, batch, generator, seed)
CASES = [
    ("tiny_hand", 2, 1, 1, "hand", 0),
    ("small_uniform", 17, 5, 2, "uniform", 7),
    ("raw_bytes_negzero", 32, 32, 4, "bytes", 42),   # shape of CutlassHerkCorrectnessTest
    ("ragged", 45, 77, 1, "uniform", 8),
    ("k_not_aligned", 64, 96, 2, "uniform", 9),       # test_gemm_api herk_k_not_aligned K=96
    ("tile_edge", 130, 64, 1, "uniform", 10),
]

# cfg1 (128 inputs, 16 channels, K=1024, seed 1) is pinned by digests only
DIGEST_CASES = [("cfg1", 128, 1024, 16, "uniform", 1)]


def decode_np(A):
    nib = np.array([0, 1, 2, 3, 4, 5, 6, 7, 0, -1, -2, -3, -4, -5, -6, -7], dtype=np.int64)
    return nib[A >> 4] + 1j * nib[A & 15]


def numpy_herk(A, N, K, B):
    X = decode_np(A.astype(np.int64)).reshape(B, N, K)
    out = np.zeros((B, N * (N + 1) // 2, 2), dtype=np.int64)
    il, jl = np.tril_indices(N)
    for b in range(B):
        G = X[b] @ X[b].conj().T
        v = G[il, jl]
        out[b, :, 0] = np.rint(v.real).astype(np.int64)
        im = np.rint(v.imag).astype(np.int64)
        im[il == jl] = 0
        out[b, :, 1] = im
    return out


def make_input(gen, N, K, B, seed):
    if gen == "hand":
        return np.array([oracle.encode_qc(1, 2), oracle.encode_qc(3, -1)], dtype=np.uint8)
    if gen == "bytes":
        return oracle.gen_bytes(seed, N * K * B)
    return oracle.gen_qc_uniform(seed, N * K * B)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    oracle.build()
    data = {}
    for name, N, K, B, gen, seed in CASES:
        A = make_input(gen, N, K, B, seed)
        C = oracle.herk_int64(A, N, K, B)
        assert np.array_equal(C, numpy_herk(A, N, K, B)), name
        data[f"{name}__A"] = A
        data[f"{name}__C"] = C
        data[f"{name}__meta"] = np.array([N, K, B, seed], dtype=np.int64)
        data[f"{name}__gen"] = np.array(gen)
    for name, N, K, B, gen, seed in DIGEST_CASES:
        A = make_input(gen, N, K, B, seed)
        C = oracle.herk_int64(A, N, K, B)
        assert np.array_equal(C, numpy_herk(A, N, K, B)), name
        F = oracle.herk_f32(A, N, K, B)
        data[f"{name}__meta"] = np.array([N, K, B, seed], dtype=np.int64)
        data[f"{name}__gen"] = np.array(gen)
        data[f"{name}__sha_A"] = np.array(sha(A))
        data[f"{name}__sha_C"] = np.array(sha(C))
        data[f"{name}__sha_F32"] = np.array(sha(F))
    path = os.path.join(HERE, "herk_fixtures.npz")
    np.savez_compressed(path, **data)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
