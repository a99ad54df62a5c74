"""ctypes front end of the CPU oracle (liboracle_herk.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` as the checker.  The product path
(``ftd_amd``, ``libftdx.so``) never imports this module.

The C restatement lives in ``herk_oracle.c``; each function there cites the
reference file:line it follows.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_herk.so")
_lib = None


def build() -> str:
    """Compile liboracle_herk.so in place (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library not built: {_LIB_PATH} (run oracle.build())")
        L = ctypes.CDLL(_LIB_PATH)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        L.ftdo_mt64_nth.restype = ctypes.c_uint64
        L.ftdo_mt64_nth.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.ftdo_encode_qc.restype = ctypes.c_uint8
        L.ftdo_encode_qc.argtypes = [ctypes.c_int, ctypes.c_int]
        L.ftdo_decode_qc.restype = None
        L.ftdo_decode_qc.argtypes = [ctypes.c_uint8, ctypes.POINTER(ctypes.c_int),
                                     ctypes.POINTER(ctypes.c_int)]
        L.ftdo_qc_to_fp8_pair.restype = ctypes.c_uint16
        L.ftdo_qc_to_fp8_pair.argtypes = [ctypes.c_uint8]
        L.ftdo_fp8e4m3_to_float.restype = ctypes.c_float
        L.ftdo_fp8e4m3_to_float.argtypes = [ctypes.c_uint8]
        L.ftdo_gen_qc_uniform.restype = None
        L.ftdo_gen_qc_uniform.argtypes = [ctypes.c_uint64, ctypes.c_int64, u8p]
        L.ftdo_gen_bytes.restype = None
        L.ftdo_gen_bytes.argtypes = [ctypes.c_uint64, ctypes.c_int64, u8p]
        L.ftdo_herk_int64.restype = None
        L.ftdo_herk_int64.argtypes = [u8p, ctypes.POINTER(ctypes.c_int64), ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ftdo_herk_int64_rows.restype = None
        L.ftdo_herk_int64_rows.argtypes = [u8p, ctypes.POINTER(ctypes.c_int64), ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int]
        L.ftdo_herk_f32.restype = None
        L.ftdo_herk_f32.argtypes = [u8p, ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_float, ctypes.c_int]
        L.ftdo_herk_f16.restype = None
        L.ftdo_herk_f16.argtypes = [u8p, ctypes.POINTER(ctypes.c_uint16), ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                    ctypes.c_float, ctypes.c_int]
        L.ftdo_f32_to_f16_rne.restype = ctypes.c_uint16
        L.ftdo_f32_to_f16_rne.argtypes = [ctypes.c_float]
        _lib = L
    return _lib


def _ptr(a: np.ndarray, ctype):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def mt64_nth(seed: int, n: int) -> int:
    return int(lib().ftdo_mt64_nth(seed, n))


def encode_qc(re: int, im: int) -> int:
    return int(lib().ftdo_encode_qc(re, im))


def decode_qc(b: int) -> tuple[int, int]:
    r, i = ctypes.c_int(), ctypes.c_int()
    lib().ftdo_decode_qc(b, ctypes.byref(r), ctypes.byref(i))
    return r.value, i.value


def qc_to_fp8_pair(b: int) -> tuple[int, int]:
    v = int(lib().ftdo_qc_to_fp8_pair(b))
    return v & 0xFF, v >> 8


def fp8e4m3_to_float(b: int) -> float:
    return float(lib().ftdo_fp8e4m3_to_float(b))


def gen_qc_uniform(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint8)
    lib().ftdo_gen_qc_uniform(seed, n, _ptr(out, ctypes.c_uint8))
    return out


def gen_bytes(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint8)
    lib().ftdo_gen_bytes(seed, n, _ptr(out, ctypes.c_uint8))
    return out


def herk_int64(A: np.ndarray, N: int, K: int, batch: int, threads: int = 0) -> np.ndarray:
    """Packed lower triangle, shape [batch, N(N+1)/2, 2] int64."""
    A = np.ascontiguousarray(A, dtype=np.uint8).reshape(-1)
    assert A.size == N * K * batch
    C = np.zeros((batch, N * (N + 1) // 2, 2), dtype=np.int64)
    lib().ftdo_herk_int64(_ptr(A, ctypes.c_uint8), _ptr(C, ctypes.c_int64), N, K, batch, threads)
    return C


def herk_int64_rows(A: np.ndarray, N: int, K: int, batch: int, row0: int, row1: int,
                    threads: int = 0, out: np.ndarray | None = None) -> np.ndarray:
    A = np.ascontiguousarray(A, dtype=np.uint8).reshape(-1)
    assert A.size == N * K * batch
    C = out if out is not None else np.zeros((batch, N * (N + 1) // 2, 2), dtype=np.int64)
    lib().ftdo_herk_int64_rows(_ptr(A, ctypes.c_uint8), _ptr(C, ctypes.c_int64), N, K, batch,
                               row0, row1, threads)
    return C


def herk_f32(A: np.ndarray, N: int, K: int, batch: int, alpha: float = 1.0, beta: float = 0.0,
             C: np.ndarray | None = None, threads: int = 0) -> np.ndarray:
    A = np.ascontiguousarray(A, dtype=np.uint8).reshape(-1)
    assert A.size == N * K * batch
    if C is None:
        C = np.zeros((batch, N * (N + 1) // 2, 2), dtype=np.float32)
    else:
        C = np.ascontiguousarray(C, dtype=np.float32).reshape(batch, N * (N + 1) // 2, 2).copy()
    lib().ftdo_herk_f32(_ptr(A, ctypes.c_uint8), _ptr(C, ctypes.c_float), N, K, batch,
                        alpha, beta, threads)
    return C


def herk_f16(A: np.ndarray, N: int, K: int, batch: int, alpha: float = 1.0, beta: float = 0.0,
             C: np.ndarray | None = None, threads: int = 0) -> np.ndarray:
    """FP16 packed output returned as uint16 bit patterns."""
    A = np.ascontiguousarray(A, dtype=np.uint8).reshape(-1)
    assert A.size == N * K * batch
    if C is None:
        C = np.zeros((batch, N * (N + 1) // 2, 2), dtype=np.uint16)
    else:
        C = np.ascontiguousarray(C).view(np.uint16).reshape(batch, N * (N + 1) // 2, 2).copy()
    lib().ftdo_herk_f16(_ptr(A, ctypes.c_uint8), _ptr(C, ctypes.c_uint16), N, K, batch,
                        alpha, beta, threads)
    return C


def f32_to_f16_rne(x: float) -> int:
    return int(lib().ftdo_f32_to_f16_rne(x))
