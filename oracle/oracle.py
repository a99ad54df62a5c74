"""ctypes front end of the CPU oracle (liboracle_herk.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py``, to check results.  The product path
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
        L.ftdo_mx_encode.restype = ctypes.c_uint8
        L.ftdo_mx_encode.argtypes = [ctypes.c_float, ctypes.c_int]
        L.ftdo_mx_decode.restype = ctypes.c_double
        L.ftdo_mx_decode.argtypes = [ctypes.c_uint8, ctypes.c_int]
        L.ftdo_mx_scale_exponent.restype = ctypes.c_uint8
        L.ftdo_mx_scale_exponent.argtypes = [ctypes.c_float, ctypes.c_int]
        L.ftdo_mx_quantize.restype = None
        L.ftdo_mx_quantize.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, u8p, u8p]
        L.ftdo_mx_encode_ref.restype = ctypes.c_uint8
        L.ftdo_mx_encode_ref.argtypes = [ctypes.c_float, ctypes.c_int]
        L.ftdo_mx_quantize_mode.restype = None
        L.ftdo_mx_quantize_mode.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, u8p, u8p]
        dp = ctypes.POINTER(ctypes.c_double)
        L.ftdo_gain_loss_grad_f64.restype = None
        L.ftdo_gain_loss_grad_f64.argtypes = [dp, dp, dp, ctypes.c_int, ctypes.c_int, dp, dp, dp]
        L.ftdo_herk_adjoint_f64.restype = None
        L.ftdo_herk_adjoint_f64.argtypes = [dp, dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int]
        L.ftdo_promote_data.restype = None
        L.ftdo_promote_data.argtypes = [u8p, ctypes.c_int64, dp]
        L.ftdo_convert_qc_to_ftd.restype = ctypes.c_uint8
        L.ftdo_convert_qc_to_ftd.argtypes = [ctypes.c_uint8]
        L.ftdo_qc_nibble_fp16.restype = ctypes.c_uint16
        L.ftdo_qc_nibble_fp16.argtypes = [ctypes.c_uint]
        L.ftdo_f32_to_e4m3_clamped.restype = ctypes.c_uint8
        L.ftdo_f32_to_e4m3_clamped.argtypes = [ctypes.c_float]
        L.ftdo_tri_from_herm.restype = None
        L.ftdo_tri_from_herm.argtypes = [dp, dp, ctypes.c_int, ctypes.c_int]
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.ftdo_xengine_ggp_i64.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, i64p, ctypes.c_int]
        L.ftdo_xengine_ggp_i64.restype = None
        L.ftdo_herk_ftd_i64.argtypes = [u8p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, i64p, ctypes.c_int]
        L.ftdo_herk_ftd_i64.restype = None
        L.ftdo_gemm_c64.argtypes = [dp, dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int]
        L.ftdo_gemm_c64.restype = None
        fp = ctypes.POINTER(ctypes.c_float)
        L.ftdo_power_f32.argtypes = [fp, fp, fp, ctypes.c_int64, ctypes.c_float, ctypes.c_float, fp]
        L.ftdo_power_f32.restype = None
        L.ftdo_voltage_tail.argtypes = [fp, fp, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_void_p]
        L.ftdo_voltage_tail.restype = None
        L.ftdo_pillbox_grid.argtypes = [fp, fp, dp, fp, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_float]
        L.ftdo_pillbox_grid.restype = None
        L.ftdo_herk_c64.restype = None
        L.ftdo_herk_c64.argtypes = [ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int]
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


# ---- MX block-scaled quantization (fmt: 0 = FP4 E2M1, 1 = FP6 E3M2, 2 = FP6 E2M3) ----
MX_FP4_E2M1, MX_FP6_E3M2, MX_FP6_E2M3 = 0, 1, 2


# element rounding: the reference's converters (fp6_kernels.hpp:27-88,
# fp4_kernels.hpp:7-30) or OCP nearest-even with subnormals
MX_ROUND_REFERENCE, MX_ROUND_NEAREST_EVEN = 0, 1


def mx_encode(x: float, fmt: int) -> int:
    return int(lib().ftdo_mx_encode(x, fmt))


def mx_encode_ref(x: float, fmt: int) -> int:
    return int(lib().ftdo_mx_encode_ref(x, fmt))


def mx_decode(code: int, fmt: int) -> float:
    return float(lib().ftdo_mx_decode(code, fmt))


def mx_scale_exponent(amax: float, fmt: int) -> int:
    return int(lib().ftdo_mx_scale_exponent(amax, fmt))


def mx_quantize(X: np.ndarray, fmt: int, rounding: int = MX_ROUND_NEAREST_EVEN):
    """X: float32 complex pairs [rows, K, 2].  Returns (codes [rows, nblk, 2, 32] uint8,
    scales [rows, nblk, 2] uint8)."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    rows, K = X.shape[0], X.shape[1]
    nblk = (K + 31) // 32
    codes = np.zeros((rows, nblk, 2, 32), dtype=np.uint8)
    scales = np.zeros((rows, nblk, 2), dtype=np.uint8)
    lib().ftdo_mx_quantize_mode(_ptr(X, ctypes.c_float), rows, K, fmt, rounding,
                                _ptr(codes, ctypes.c_uint8), _ptr(scales, ctypes.c_uint8))
    return codes, scales


def mx_dequantize(codes: np.ndarray, scales: np.ndarray, fmt: int, K: int) -> np.ndarray:
    """Back to complex128 [rows, K] from mx_quantize's output."""
    table = np.array([mx_decode(c, fmt) for c in range(64 if fmt else 16)])
    vals = table[codes.astype(np.int64) & (63 if fmt else 15)]
    vals = vals * np.ldexp(1.0, scales.astype(np.int64) - 127)[..., None]
    rows, nblk = scales.shape[0], scales.shape[1]
    re = vals[:, :, 0, :].reshape(rows, nblk * 32)[:, :K]
    im = vals[:, :, 1, :].reshape(rows, nblk * 32)[:, :K]
    return re + 1j * im


def herk_c64(X: np.ndarray, N: int, K: int, batch: int, threads: int = 0) -> np.ndarray:
    """FP64 HERK of complex128 X [batch, N, K] -> packed [batch, N(N+1)/2, 2] float64."""
    Xf = np.ascontiguousarray(np.stack([X.real, X.imag], axis=-1), dtype=np.float64)
    C = np.zeros((batch, N * (N + 1) // 2, 2), dtype=np.float64)
    lib().ftdo_herk_c64(_ptr(Xf, ctypes.c_double), _ptr(C, ctypes.c_double), N, K, batch, threads)
    return C


# ---- gain calibration (FP64) ----
def _c2(a):
    a = np.asarray(a)
    if np.iscomplexobj(a):
        a = np.stack([a.real, a.imag], axis=-1)
    return np.ascontiguousarray(a, dtype=np.float64)


def gain_loss_grad_f64(V, M, g, N: int, batch: int):
    """V, M: packed [batch, P] complex (or [..., 2] float); g: [batch, N] complex.
    Returns (loss [batch], grad [batch, N] complex, Gamma [batch, P] complex)."""
    Vf, Mf, gf = _c2(V), _c2(M), _c2(g)
    P = N * (N + 1) // 2
    assert Vf.size == batch * P * 2 and Mf.size == batch * P * 2 and gf.size == batch * N * 2
    loss = np.zeros(batch)
    grad = np.zeros((batch, N, 2))
    Gam = np.zeros((batch, P, 2))
    lib().ftdo_gain_loss_grad_f64(_ptr(Vf, ctypes.c_double), _ptr(Mf, ctypes.c_double),
                                  _ptr(gf, ctypes.c_double), N, batch,
                                  _ptr(loss, ctypes.c_double), _ptr(grad, ctypes.c_double),
                                  _ptr(Gam, ctypes.c_double))
    return loss, grad[..., 0] + 1j * grad[..., 1], Gam[..., 0] + 1j * Gam[..., 1]


def herk_adjoint_f64(Gamma, X, N: int, K: int, batch: int, threads: int = 0):
    """dX = (Gamma + Gamma^H) X in FP64; Gamma packed [batch, P] complex; X [batch, N, K]
    complex.  Returns complex128 [batch, N, K]."""
    Gf, Xf = _c2(Gamma), _c2(X)
    dX = np.zeros((batch, N, K, 2))
    lib().ftdo_herk_adjoint_f64(_ptr(Gf, ctypes.c_double), _ptr(Xf, ctypes.c_double),
                                _ptr(dX, ctypes.c_double), N, K, batch, threads)
    return dX[..., 0] + 1j * dX[..., 1]


def decode_qc_array(A: np.ndarray) -> np.ndarray:
    """QC bytes -> complex128 values."""
    nib = np.array([0, 1, 2, 3, 4, 5, 6, 7, 0, -1, -2, -3, -4, -5, -6, -7], dtype=np.float64)
    A = np.asarray(A, dtype=np.uint8)
    return nib[A >> 4] + 1j * nib[A & 15]


# ---- stage converters ----
def promote_data(data: np.ndarray) -> np.ndarray:
    d = np.ascontiguousarray(data, dtype=np.uint8).reshape(-1)
    out = np.zeros((d.size, 2))
    lib().ftdo_promote_data(_ptr(d, ctypes.c_uint8), d.size, _ptr(out, ctypes.c_double))
    return out


def convert_qc_to_ftd(data: np.ndarray) -> np.ndarray:
    table = np.array([lib().ftdo_convert_qc_to_ftd(b) for b in range(256)], dtype=np.uint8)
    return table[np.asarray(data, dtype=np.uint8)]


def qc_nibble_fp16(n: int) -> int:
    return int(lib().ftdo_qc_nibble_fp16(n))


def f32_to_e4m3_clamped(x: float) -> int:
    return int(lib().ftdo_f32_to_e4m3_clamped(x))


def tri_from_herm(full: np.ndarray, N: int, batch: int) -> np.ndarray:
    f = np.ascontiguousarray(full, dtype=np.float64).reshape(batch, N, N, 2)
    out = np.zeros((batch, N * (N + 1) // 2, 2))
    lib().ftdo_tri_from_herm(_ptr(f, ctypes.c_double), _ptr(out, ctypes.c_double), N, batch)
    return out


# ---- GGP XEngine and FTD raw-order correlation ----
def xengine_ggp_i64(data: np.ndarray, N: int, K: int, batch: int, herm: bool,
                    threads: int = 0) -> np.ndarray:
    """M^H M per batch of GGP bytes [batch][K][N]: TRI -> [batch, N(N+1)/2, 2] (upper,
    row-major), HERM -> [batch, N, N, 2]; int64."""
    d = np.ascontiguousarray(data, dtype=np.uint8).reshape(-1)
    assert d.size >= N * K * batch
    shape = (batch, N, N, 2) if herm else (batch, N * (N + 1) // 2, 2)
    out = np.zeros(shape, dtype=np.int64)
    lib().ftdo_xengine_ggp_i64(_ptr(d, ctypes.c_uint8), N, K, batch, int(herm),
                               _ptr(out, ctypes.c_int64), threads)
    return out


def herk_ftd_i64(raw: np.ndarray, n_ant: int, n_chan: int, n_time: int, n_ti: int, n_pol: int,
                 twos: bool, threads: int = 0) -> np.ndarray:
    """FTD [time][ant][chan][pol] bytes -> packed lower triangle
    [n_chan*n_pol*n_ti, n_ant(n_ant+1)/2, 2] int64."""
    d = np.ascontiguousarray(raw, dtype=np.uint8).reshape(-1)
    assert d.size >= n_time * n_ant * n_chan * n_pol
    out = np.zeros((n_chan * n_pol * n_ti, n_ant * (n_ant + 1) // 2, 2), dtype=np.int64)
    lib().ftdo_herk_ftd_i64(_ptr(d, ctypes.c_uint8), n_ant, n_chan, n_time, n_ti, n_pol, int(twos),
                            _ptr(out, ctypes.c_int64), threads)
    return out



# ---- complex GEMM (cpu_gemm) and the power epilogue ----
def gemm_c64(A: np.ndarray, B: np.ndarray, threads: int = 0) -> np.ndarray:
    """A complex [batch, M, K], B complex [batch, N, K] -> C = A B^T [batch, M, N] (complex128)."""
    A = np.ascontiguousarray(A, dtype=np.complex128)
    B = np.ascontiguousarray(B, dtype=np.complex128)
    bt, M, K = A.shape
    N = B.shape[1]
    C = np.zeros((bt, M, N), dtype=np.complex128)
    lib().ftdo_gemm_c64(_ptr(A.view(np.float64), ctypes.c_double), _ptr(B.view(np.float64), ctypes.c_double),
                        _ptr(C.view(np.float64), ctypes.c_double), M, N, K, bt, threads)
    return C


def power_f32(re: np.ndarray, im: np.ndarray, alpha: float = 1.0, beta: float = 0.0,
              old: np.ndarray | None = None) -> np.ndarray:
    re = np.ascontiguousarray(re, dtype=np.float32).reshape(-1)
    im = np.ascontiguousarray(im, dtype=np.float32).reshape(-1)
    o = np.zeros_like(re) if old is None else np.ascontiguousarray(old, dtype=np.float32).reshape(-1)
    out = np.empty_like(re)
    fp = ctypes.c_float
    lib().ftdo_power_f32(_ptr(re, fp), _ptr(im, fp), _ptr(o, fp), re.size, alpha, beta, _ptr(out, fp))
    return out



# ---- voltage beamformer (ggp::VoltagePipeline) ----
_E4M3_OF_F16 = None


def fp16_to_e4m3_values(w16: np.ndarray) -> np.ndarray:
    """FP16 values -> the E4M3 values prepare_b turns them into (clamp +-448, nearest-even;
    cast_kernels_common.hpp:30-48), as float64, through a table over all FP16 codes."""
    global _E4M3_OF_F16
    if _E4M3_OF_F16 is None:
        vals = np.arange(65536, dtype=np.uint32).astype(np.uint16).view(np.float16).astype(np.float32)
        _E4M3_OF_F16 = np.array([fp8e4m3_to_float(f32_to_e4m3_clamped(float(v))) for v in vals])
    w16 = np.ascontiguousarray(w16, dtype=np.float16)
    return _E4M3_OF_F16[w16.view(np.uint16)]


def voltage_weights(weights: np.ndarray, n_ch: int, n_beam: int, n_ant: int, compute: str) -> np.ndarray:
    """FP32 [n_ch, n_beam, n_ant, 2] weights as the GEMM sees them (complex128): FP32 ->
    FP16 nearest-even (deinterleave_fp32_to_fp16_kernel, lib/voltage_pipeline.cu:484-498),
    then E4M3 (compute 'fp8') or MX E3M2 with 32-sample scales along antennas ('fp6', the
    reference's element converter)."""
    w16 = np.asarray(weights, dtype=np.float32).reshape(n_ch, n_beam, n_ant, 2).astype(np.float16)
    if compute == "fp8":
        q = fp16_to_e4m3_values(w16)
        return q[..., 0] + 1j * q[..., 1]
    codes, scales = mx_quantize(w16.astype(np.float32).reshape(n_ch * n_beam, n_ant, 2), MX_FP6_E3M2,
                                MX_ROUND_REFERENCE)
    return mx_dequantize(codes, scales, MX_FP6_E3M2, n_ant).reshape(n_ch, n_beam, n_ant)


def voltage_beams(payloads, W: np.ndarray, n_ant: int, n_beam: int, n_ch: int, n_time: int,
                  ch_fuse: int = 1):
    """Beams of the fused GEMM: re, im float32 [batch, 2, M_total, n_beam] with batch b =
    n_ch / ch_fuse using weight set b (the first n_ch / ch_fuse sets are prepared,
    lib/voltage_pipeline.cu:770-776), M_total = n_time * ch_fuse * n_payloads and sample
    m = (payload * ch_fuse + c) * n_time + t of channel b * ch_fuse + c
    (qc_fused_transpose_fp8_kernel, lib/voltage_pipeline.cu:176-236).  The products of
    the rounded weights and the 4-bit voltages are summed in float64; the caller keeps
    them exact (then every FP32 summation order gives the same value)."""
    batch = n_ch // ch_fuse
    V = np.stack([decode_qc_array(np.asarray(p, dtype=np.uint8)).reshape(n_ch, 2, n_ant, n_time)
                  for p in payloads])                                  # [np, ch, pol, ant, t]
    V = V.reshape(len(payloads), batch, ch_fuse, 2, n_ant, n_time)
    V = V.transpose(1, 3, 0, 2, 5, 4).reshape(batch, 2, -1, n_ant)      # [b, pol, m, ant]
    Wb = W[:batch]                                                      # [b, beam, ant]
    C = np.einsum("bpma,bna->bpmn", V, Wb)
    re, im = C.real, C.imag
    # exactness: every product is a multiple of g (the finest power of two among the
    # weight parts, the voltages being integers), so every partial sum in any order is
    # an FP32 value while its magnitude stays below 2^24 g; bound it by sum |v| |w|
    parts = np.abs(np.concatenate([Wb.real.ravel(), Wb.imag.ravel()]))
    parts = parts[parts > 0]
    if parts.size:
        ints = np.ldexp(parts, 80).astype(object)
        g = min(int(x) & -int(x) for x in set(ints.tolist()))
        g = float(np.ldexp(float(g), -80))
        bound = (np.abs(V.real) + np.abs(V.imag)) @ \
            (np.abs(Wb.real) + np.abs(Wb.imag)).transpose(0, 2, 1)[:, None]
        assert bound.max() < 2.0 ** 24 * g, "FP32 partial sums could round: not an exact case"
    return re.astype(np.float32), im.astype(np.float32)


def voltage_tail(re: np.ndarray, im: np.ndarray, nt: int, mode: int, scale: float, out_u8: bool):
    """ftdo_voltage_tail: [batch, 2, M, n_beam] beams -> [batch, n_beam, M / nt] (mode 0 =
    FP8 path: power GEMMs with beta accumulation + time integration; mode 1 = FP6 path:
    power_sum)."""
    re = np.ascontiguousarray(re, dtype=np.float32)
    im = np.ascontiguousarray(im, dtype=np.float32)
    batch, _, M, n_beam = re.shape
    out = np.zeros((batch, n_beam, M // nt), dtype=np.uint8 if out_u8 else np.float32)
    fp = ctypes.c_float
    lib().ftdo_voltage_tail(_ptr(re, fp), _ptr(im, fp), batch, M, n_beam, nt, mode, scale, int(out_u8),
                            out.ctypes.data)
    return out


def voltage_beamform(payloads, weights: np.ndarray, n_ant: int, n_beam: int, n_ch: int, n_time: int,
                     n_tps: int, scale: float = 1.0, ch_fuse: int = 1, compute: str = "fp8",
                     out_u8: bool = True):
    """The pipeline's output for QC payloads (each [n_ch, 2, n_ant, n_time] bytes):
    [n_ch / ch_fuse, n_beam, M_total / n_tps] uint8 (beamform / beamform_batched) or FP32
    (compute_filterbank)."""
    W = voltage_weights(weights, n_ch, n_beam, n_ant, compute)
    re, im = voltage_beams(payloads, W, n_ant, n_beam, n_ch, n_time, ch_fuse)
    return voltage_tail(re, im, n_tps, 0 if compute == "fp8" else 1, scale, out_u8)


def voltage_fp64(payload, weights: np.ndarray, n_ant: int, n_beam: int, n_ch: int, n_time: int, n_tps: int):
    """cpu_beamformer (tests/voltage_pipeline_test.cpp:43-83): FP64 power with the
    unrounded weights, [n_ch, n_beam, n_time / n_tps]."""
    w = np.asarray(weights, dtype=np.float64).reshape(n_ch, n_beam, n_ant, 2)
    Wc = w[..., 0] + 1j * w[..., 1]
    V = decode_qc_array(np.asarray(payload, dtype=np.uint8)).reshape(n_ch, 2, n_ant, n_time)
    beams = np.einsum("cpat,cna->cpnt", V, Wc)                            # [ch, pol, beam, t]
    power = (beams.real ** 2 + beams.imag ** 2).sum(axis=1)               # std::norm; [ch, beam, t]
    return power.reshape(n_ch, n_beam, n_time // n_tps, n_tps).sum(axis=-1)


# ---- visibility consumers ----
def pol_reduce(tri: np.ndarray, n_ch: int, n_pol: int, n_ti: int) -> np.ndarray:
    """[n_ch*n_pol*n_ti, P, 2] -> [n_ch*n_ti, P, 2]: sum over pol in float64, one rounding
    to the input dtype (pol_reduce_triangle_kernel, lib/bench_pipeline.cu:254-302)."""
    t = np.asarray(tri)
    P = t.shape[1]
    s = t.astype(np.float64).reshape(n_ch, n_pol, n_ti, P, 2).sum(axis=1)
    return s.reshape(n_ch * n_ti, P, 2).astype(t.dtype)


def pillbox_grid(vis: np.ndarray, uv_m: np.ndarray, freq_hz: np.ndarray, n_bl: int, Ng: int, Nf: int,
                 cell: float, freq_offset: int = 0, grid: np.ndarray | None = None) -> np.ndarray:
    v = np.ascontiguousarray(vis, dtype=np.float32).reshape(-1)
    u = np.ascontiguousarray(uv_m, dtype=np.float32).reshape(-1)
    fr = np.ascontiguousarray(freq_hz, dtype=np.float64).reshape(-1)
    g = np.zeros((Nf, Ng, Ng, 2), dtype=np.float32) if grid is None else grid
    fp = ctypes.c_float
    lib().ftdo_pillbox_grid(_ptr(v, fp), _ptr(u, fp), _ptr(fr, ctypes.c_double), _ptr(g, fp), n_bl, Ng,
                            Nf, freq_offset, cell)
    return g
