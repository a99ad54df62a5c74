"""Python host side of the X-engine HERK library: a ctypes binding of the C ABI
in ``include/ftdx.h`` (``libftdx.so``) plus a Python mirror of the reference's
``cutlass_gemm_api::CutlassComplexGemm`` HERK surface
(cutlass_interface/cutlass_gemm_api.h:157-789) and GGP glue
(lib/cutlass_interface.cu:150-289).

Device memory and streams come from PyTorch (ROCm); all arithmetic runs in the
HIP kernels of libftdx.so.  There is no CPU fallback: if the library is not
built, every entry point raises.
"""
from __future__ import annotations

import ctypes
import enum
import os
import re
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libftdx.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "ftdx.h")

FTDX_OK = 0
FTDX_ERR_INVALID_ARG = 1
FTDX_ERR_UNSUPPORTED = 2
FTDX_ERR_HIP = 3
FTDX_ERR_NO_MEMORY = 4
FTDX_ERR_IO = 5


class InputPrecision(enum.IntEnum):
    FP16 = 0
    INT4 = 1


class ComputePrecision(enum.IntEnum):
    FP8 = 0
    FP6 = 1
    FP4 = 2


class OutputPrecision(enum.IntEnum):
    FP16 = 0
    FP32 = 1


class MxRounding(enum.IntEnum):
    """Element rounding of the MX (FP6 / FP4) quantizer.  Reference: the reference's
    converters code for code (gemm_complex_sm100/fp6_kernels.hpp:27-58,
    fp4_kernels.hpp:7-30), the default.  NearestEven: OCP MX v1.0 rounding."""
    Reference = 0
    NearestEven = 1


class OutputFormat(enum.IntEnum):
    PackedTriangle = 0


class HerkMode(enum.IntEnum):
    Auto = 0
    ForceDirect = 1
    ForceBaseline = 2


class DirectTileConfig(enum.IntEnum):
    Auto = 0
    K32_B3 = 1
    K64_B2 = 2
    K64_B3 = 3
    K128_B2 = 4


class FtdxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ftdx error {code}: {msg}")
        self.code = code


class UnsupportedError(FtdxError):
    pass


_lib = None
_lock = threading.Lock()

_VP = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_S = ctypes.c_char_p

# name -> (restype, argtypes); must cover every function declared in ftdx.h
_SIGNATURES = {
    "ftdx_last_error": (ctypes.c_char_p, []),
    "ftdx_abi_version": (_I, []),
    "ftdx_create": (_I, [ctypes.POINTER(_VP)]),
    "ftdx_destroy": (_I, [_VP]),
    "ftdx_herk": (_I, [_VP, _VP, _VP, _I, _I, _I, _I, _I, _I, _I, _F, _F, _VP, _I]),
    "ftdx_init_herk": (_I, [_VP, _I, _I, _I, _VP]),
    "ftdx_end_herk": (_I, [_VP]),
    "ftdx_set_herk_mode": (_I, [_VP, _I]),
    "ftdx_set_direct_tile_config": (_I, [_VP, _I]),
    "ftdx_set_direct_herk_graph": (_I, [_VP, _I]),
    "ftdx_set_tune_cache_path": (_I, [_VP, _S]),
    "ftdx_set_strategy_tune_verbosity": (_I, [_VP, _I]),
    "ftdx_set_kernel_tune_verbosity": (_I, [_VP, _I]),
    "ftdx_set_kernel_tune_cache_path": (_I, [_VP, _S]),
    "ftdx_print_build_info": (_I, []),
    "ftdx_herk_int4": (_I, [_VP, _VP, _I, _I, _I, _I, _F, _F, _VP]),
    "ftdx_set_launch_config": (_I, [_VP, _I, _I]),
    "ftdx_herk_graph_stats": (_I, [_VP, ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_longlong)]),
    "ftdx_tune_load": (_I, [_VP, _S, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "ftdx_tune_save": (_I, [_VP, _S]),
    "ftdx_herk_query_config": (_I, [_VP, _I, _I, _I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "ftdx_herk_grouped_plan_create": (_I, [_VP, _VP, _I, _I, _I, _I, ctypes.POINTER(_VP)]),
    "ftdx_herk_grouped_plan_run": (_I, [_VP, _VP]),
    "ftdx_herk_grouped_plan_info": (_I, [_VP, ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_int64),
                                         ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_double)]),
    "ftdx_herk_grouped_plan_destroy": (_I, [_VP]),
    "ftdx_herk_each": (_I, [_VP, _VP, _I, _I, _VP]),
    "ftdx_herk_mx": (_I, [_VP, _VP, _VP, _I, _I, _I, _I, _I, _I, _F, _F, _VP]),
    "ftdx_mx_quantize": (_I, [_VP, _I, _I, _VP, _VP, ctypes.c_int64, _I, _VP]),
    "ftdx_mx_quantize_rounding": (_I, [_VP, _I, _I, _I, _VP, _VP, ctypes.c_int64, _I, _VP]),
    "ftdx_set_mx_rounding": (_I, [_VP, _I]),
    "ftdx_get_mx_rounding": (_I, [_VP]),
    "ftdx_mx_layout": (_I, [_I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "ftdx_gain_loss_grad": (_I, [_VP, _VP, _VP, _VP, _I, _I, _VP, _VP, _VP, _VP]),
    "ftdx_gain_loss_grad_noalloc": (_I, [_VP, _VP, _VP, _I, _I, _VP, _VP, _VP, _VP]),
    "ftdx_herk_adjoint": (_I, [_VP, _VP, _VP, _VP, _I, _I, _I, _I, _VP]),
    "ftdx_promote_qc_to_fp8": (_I, [_VP, _VP, ctypes.c_int64, _VP]),
    "ftdx_promote_qc_to_fp16_planar": (_I, [_VP, _VP, _VP, ctypes.c_int64, _VP]),
    "ftdx_promote_data": (_I, [_VP, _VP, ctypes.c_int64, _I, _VP]),
    "ftdx_pack_triangle": (_I, [_VP, _VP, _VP, _I, _I, _F, _VP]),
    "ftdx_tri_from_herm": (_I, [_VP, _VP, _I, _I, _I, _VP]),
    "ftdx_interleave_tri": (_I, [_VP, _VP, _VP, ctypes.c_int64, _I, _VP]),
    "ftdx_promote_qc_to_fp32": (_I, [_VP, _VP, ctypes.c_int64, _VP]),
    "ftdx_convert_qc_to_ftd": (_I, [_VP, _VP, ctypes.c_int64, _VP]),
    "ftdx_correlate_ftd": (_I, [_VP, _VP, _VP, _I, _I, _I, _I, _I, _I, _VP]),
    "ftdx_xengine_param_init": (_I, [_VP]),
    "ftdx_xengine_create": (_I, [_VP, ctypes.POINTER(_VP)]),
    "ftdx_xengine_compute": (_I, [_VP, _VP, _VP, _VP]),
    "ftdx_xengine_query": (_I, [_VP, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t),
                                ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "ftdx_xengine_set_launch_config": (_I, [_VP, _I, _I]),
    "ftdx_xengine_destroy": (_I, [_VP]),
    "ftdx_xengine_chpc": (_I, [_VP, _VP, _VP]),
    "ftdx_prepare_b": (_I, [_VP, _VP, _VP, _I, _I, _I, _I, _VP]),
    "ftdx_gemm_prepared": (_I, [_VP, _VP, _VP, _VP, _VP, _I, _I, _I, _I, _I, _F, _F, _VP, _I, _I]),
    "ftdx_gemm_prepared_direct": (_I, [_VP, _VP, _VP, _VP, _VP, _I, _I, _I, _I, _F, _F, _VP]),
    "ftdx_gemm_prepared_direct_int4": (_I, [_VP, _VP, _VP, _VP, _I, _I, _I, _I, _F, _F, _VP]),
    "ftdx_gemm_prepared_power": (_I, [_VP, _VP, _VP, _VP, _I, _I, _I, _I, _F, _F, _VP]),
    "ftdx_gemm_prepared_power_fp8": (_I, [_VP, _VP, _VP, _I, _I, _I, _I, _F, _F, _VP]),
    "ftdx_gemm_prepared_power_int4": (_I, [_VP, _VP, _VP, _I, _I, _I, _I, _F, _F, _VP]),
    "ftdx_gemm": (_I, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _I, _I, _I, _I, _I, _I, _F, _F, _VP, _I, _I]),
    "ftdx_pol_reduce_triangle": (_I, [_VP, _VP, _I, _I, _I, _I, _I, _VP]),
    "ftdx_herk_pol_reduce": (_I, [_VP, _VP, _VP, _I, _I, _I, _I, _I, _I, _F, _F, _VP]),
    "ftdx_pillbox_grid": (_I, [_VP, _I, _VP, _VP, _VP, _I, _I, _I, _I, _F, _VP]),
    "ftdx_herk_pol_grid": (_I, [_VP, _VP, _I, _I, _I, _I, _VP, _VP, _VP, _I, _I, _F, _F, _VP]),
    "ftdx_pipeline_create": (_I, [_VP, _I, _I, _I, _I, ctypes.POINTER(_VP)]),
    "ftdx_pipeline_query": (_I, [_VP, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "ftdx_pipeline_correlate": (_I, [_VP, _VP, _VP]),
    "ftdx_pipeline_correlate_many": (_I, [_VP, ctypes.POINTER(_VP), ctypes.POINTER(_VP), _I]),
    "ftdx_pipeline_destroy": (_I, [_VP]),
    "ftdx_voltage_config_init": (_I, [_VP]),
    "ftdx_voltage_create": (_I, [_VP, ctypes.POINTER(_VP)]),
    "ftdx_voltage_destroy": (_I, [_VP]),
    "ftdx_voltage_set_weights": (_I, [_VP, _VP, _VP]),
    "ftdx_voltage_beamform": (_I, [_VP, _VP, _VP, _F, _VP]),
    "ftdx_voltage_beamform_batched": (_I, [_VP, _VP, ctypes.POINTER(_VP), _I, _F, _VP]),
    "ftdx_voltage_compute_filterbank": (_I, [_VP, _VP, _VP]),
    "ftdx_voltage_compute": (_I, [_VP, _VP, _VP, _VP, _F, _VP]),
    "ftdx_voltage_corner_turn": (_I, [_VP, _VP, _VP, _VP]),
    "ftdx_voltage_set_graph_enabled": (_I, [_VP, _I]),
    "ftdx_voltage_query": (_I, [_VP, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(_I), ctypes.POINTER(_I),
                                ctypes.POINTER(_VP), ctypes.POINTER(_VP)]),
    "ftdx_corner_turn": (_I, [_VP, _VP, _I, _I, _I, _VP]),
    "ftdx_gemm_prepared_fused": (_I, [_VP, _VP, _VP, _VP, _VP, _I, _I, _I, _I, _I, _I, _F, _F, _VP]),
    "ftdx_debug_fp6_real_gemm": (_I, [_VP, _I, _I, _I, _VP]),
    "ftdx_set_gemm_setting": (_I, [_VP, _I, _I]),
    "ftdx_get_gemm_setting": (_I, [_VP, _I, ctypes.POINTER(_I)]),
    "ftdx_set_gemm_tune_cache_path": (_I, [_VP, _S]),
    "ftdx_blas_herk_fp32": (_I, [_VP, _VP, _VP, _I, _I, _I, _I]),
    "ftdx_blas_gemm_fp32": (_I, [_VP, _VP, _VP, _VP, _I, _I, _I, _I, _I, _I]),
}

PREC_FP32 = 1
PREC_FP64 = 2

TILE_AUTO = -1
TILE_128x128_K64 = 0
TILE_128x128_K128 = 1
TILE_64x64_K64 = 2
TILE_256x128_K64 = 3
TILE_128x128_K64_P2 = 4
TILE_256x128_K64_P2 = 5
TILE_CONFIGS = (TILE_128x128_K64, TILE_128x128_K128, TILE_64x64_K64, TILE_256x128_K64,
                TILE_128x128_K64_P2, TILE_256x128_K64_P2)
TILE_BF6 = 16  # flag: the tile with pre-cast E3M2 operands (ftdx_herk, INT4 input)

# ftdx_set_gemm_setting / ftdx_get_gemm_setting (the reference's GEMM dispatch setters)
GEMM_SETTING_TILE_CONFIG = 0
GEMM_SETTING_POWER_MODE = 1
GEMM_SETTING_GEMM_MODE = 2
GEMM_SETTING_AUTO_TUNE = 3
GEMM_SETTING_TUNE_VERBOSITY = 4


class HerkProblem(ctypes.Structure):
    """ftdx_herk_problem (include/ftdx.h)."""
    _fields_ = [("A", ctypes.c_void_p), ("C", ctypes.c_void_p), ("N", ctypes.c_int),
                ("K", ctypes.c_int), ("alpha", ctypes.c_float), ("beta", ctypes.c_float)]


def declared_functions(header: str = HEADER_PATH) -> list[str]:
    """Names of the functions declared in include/ftdx.h."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ftdx_[a-z0-9_]+)\s*\(", text)))


def lib():
    """Load libftdx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise RuntimeError(
                        f"libftdx.so not found at {LIB_PATH}: build it first "
                        "(python -c 'import __graft_entry__ as g; g.build()')")
                L = ctypes.CDLL(LIB_PATH)
                for name, (res, args) in _SIGNATURES.items():
                    fn = getattr(L, name)
                    fn.restype = res
                    fn.argtypes = args
                _lib = L
    return _lib


def last_error() -> str:
    return lib().ftdx_last_error().decode(errors="replace")


def _check(rc: int, what: str) -> None:
    if rc == FTDX_OK:
        return
    msg = f"{what}: {last_error()}"
    if rc == FTDX_ERR_UNSUPPORTED:
        raise UnsupportedError(rc, msg)
    raise FtdxError(rc, msg)


def _stream_ptr(stream) -> int | None:
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream or None
    if isinstance(stream, int):
        return stream or None
    return stream.cuda_stream or None


def _ptr(t) -> int | None:
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("ftdx: tensors must be device (HIP) tensors")
    if not t.is_contiguous():
        raise ValueError("ftdx: tensors must be contiguous")
    return t.data_ptr()


def packed_len(N: int) -> int:
    return N * (N + 1) // 2


class CutlassComplexGemm:
    """Mirror of cutlass_gemm_api::CutlassComplexGemm's HERK surface."""

    def __init__(self):
        h = ctypes.c_void_p()
        _check(lib().ftdx_create(ctypes.byref(h)), "ftdx_create")
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().ftdx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def set_herk_mode(self, mode: HerkMode) -> None:
        _check(lib().ftdx_set_herk_mode(self._h, int(mode)), "set_herk_mode")

    def set_direct_tile_config(self, cfg: DirectTileConfig) -> None:
        _check(lib().ftdx_set_direct_tile_config(self._h, int(cfg)), "set_direct_tile_config")

    def set_direct_herk_graph(self, enable: bool) -> None:
        _check(lib().ftdx_set_direct_herk_graph(self._h, int(bool(enable))),
               "set_direct_herk_graph")

    def set_tune_cache_path(self, path: str) -> None:
        _check(lib().ftdx_set_tune_cache_path(self._h, path.encode()), "set_tune_cache_path")

    def set_strategy_tune_verbosity(self, level: int) -> None:
        _check(lib().ftdx_set_strategy_tune_verbosity(self._h, level),
               "set_strategy_tune_verbosity")

    def set_kernel_tune_verbosity(self, level: int) -> None:
        _check(lib().ftdx_set_kernel_tune_verbosity(self._h, level), "set_kernel_tune_verbosity")

    def set_kernel_tune_cache_path(self, path: str) -> None:
        _check(lib().ftdx_set_kernel_tune_cache_path(self._h, path.encode()),
               "set_kernel_tune_cache_path")

    def init_herk(self, N: int, K: int, batch_count: int, stream=None) -> None:
        _check(lib().ftdx_init_herk(self._h, N, K, batch_count, _stream_ptr(stream)), "init_herk")

    def end_herk(self) -> None:
        _check(lib().ftdx_end_herk(self._h), "end_herk")

    def herk(self, A, C, N: int, K: int, batch_count: int,
             input: InputPrecision = InputPrecision.FP16,
             compute: ComputePrecision = ComputePrecision.FP8,
             output: OutputPrecision = OutputPrecision.FP16,
             format: OutputFormat = OutputFormat.PackedTriangle,
             alpha: float = 1.0, beta: float = 0.0, stream=None, tune: bool = True) -> int:
        """C = alpha * A A^H + beta * C (batched), packed lower triangle.

        Raises UnsupportedError for unsupported precision combinations (as the
        reference throws); returns the non-zero status for invalid arguments.
        """
        rc = lib().ftdx_herk(self._h, _ptr(A), _ptr(C), N, K, batch_count, int(input),
                             int(compute), int(output), int(format), alpha, beta,
                             _stream_ptr(stream), int(bool(tune)))
        if rc in (FTDX_ERR_UNSUPPORTED, FTDX_ERR_HIP, FTDX_ERR_NO_MEMORY):
            _check(rc, "herk")
        return rc

    @staticmethod
    def print_build_info() -> None:
        lib().ftdx_print_build_info()

    # ---- complex GEMM with prepared B (cutlass_gemm_api.h:215-470) ----
    def _gemm_rc(self, rc: int, what: str) -> int:
        if rc in (FTDX_ERR_UNSUPPORTED, FTDX_ERR_HIP, FTDX_ERR_NO_MEMORY):
            _check(rc, what)
        return rc

    def prepare_b(self, B_re, B_im, N: int, K: int, batch_count: int,
                  compute: ComputePrecision = ComputePrecision.FP8, stream=None) -> None:
        _check(lib().ftdx_prepare_b(self._h, _ptr(B_re), _ptr(B_im), N, K, batch_count,
                                    int(compute), _stream_ptr(stream)), "prepare_b")

    def gemm_prepared(self, A_re, A_im, C_re, C_im, M: int, N: int, K: int, batch_count: int,
                      output: OutputPrecision = OutputPrecision.FP32, alpha: float = 1.0,
                      beta: float = 0.0, stream=None, tune: bool = True,
                      input: InputPrecision = InputPrecision.FP16) -> int:
        return self._gemm_rc(lib().ftdx_gemm_prepared(
            self._h, _ptr(A_re), _ptr(A_im), _ptr(C_re), _ptr(C_im), M, N, K, batch_count,
            int(output), alpha, beta, _stream_ptr(stream), int(bool(tune)), int(input)),
            "gemm_prepared")

    def gemm_prepared_direct(self, A_re, A_im, C_re, C_im, M, N, K, batch_count, alpha=1.0,
                             beta=0.0, stream=None) -> int:
        return self._gemm_rc(lib().ftdx_gemm_prepared_direct(
            self._h, _ptr(A_re), _ptr(A_im), _ptr(C_re), _ptr(C_im), M, N, K, batch_count,
            alpha, beta, _stream_ptr(stream)), "gemm_prepared_direct")

    def gemm_prepared_direct_int4(self, A_int4, C_re, C_im, M, N, K, batch_count, alpha=1.0,
                                  beta=0.0, stream=None) -> int:
        return self._gemm_rc(lib().ftdx_gemm_prepared_direct_int4(
            self._h, _ptr(A_int4), _ptr(C_re), _ptr(C_im), M, N, K, batch_count, alpha, beta,
            _stream_ptr(stream)), "gemm_prepared_direct_int4")

    def gemm_prepared_power(self, A_re, A_im, C_power, M, N, K, batch_count, alpha=1.0, beta=0.0,
                            stream=None) -> int:
        return self._gemm_rc(lib().ftdx_gemm_prepared_power(
            self._h, _ptr(A_re), _ptr(A_im), _ptr(C_power), M, N, K, batch_count, alpha, beta,
            _stream_ptr(stream)), "gemm_prepared_power")

    def gemm_prepared_power_fp8(self, A_fp8, C_power, M, N, K, batch_count, alpha=1.0, beta=0.0,
                                stream=None) -> int:
        return self._gemm_rc(lib().ftdx_gemm_prepared_power_fp8(
            self._h, _ptr(A_fp8), _ptr(C_power), M, N, K, batch_count, alpha, beta,
            _stream_ptr(stream)), "gemm_prepared_power_fp8")

    def gemm_prepared_fused(self, A_re, A_im, C_re, C_im, M, N, K, batch_count, fuse_factor=0,
                            output: OutputPrecision = OutputPrecision.FP32, alpha=1.0, beta=0.0,
                            stream=None) -> int:
        # every failure throws in the reference (cutlass_gemm_api.cu:982-995)
        _check(lib().ftdx_gemm_prepared_fused(
            self._h, _ptr(A_re), _ptr(A_im), _ptr(C_re), _ptr(C_im), M, N, K, batch_count, fuse_factor,
            int(output), alpha, beta, _stream_ptr(stream)), "gemm_prepared_fused")
        return 0

    def debug_fp6_real_gemm(self, M: int, N: int, K: int, stream=None) -> int:
        rc = lib().ftdx_debug_fp6_real_gemm(self._h, M, N, K, _stream_ptr(stream))
        return 0 if rc == FTDX_OK else (1 if rc == FTDX_ERR_INVALID_ARG else -1)

    def set_gemm_setting(self, setting: int, value: int) -> None:
        _check(lib().ftdx_set_gemm_setting(self._h, setting, int(value)), "set_gemm_setting")

    def get_gemm_setting(self, setting: int) -> int:
        v = ctypes.c_int()
        _check(lib().ftdx_get_gemm_setting(self._h, setting, ctypes.byref(v)), "get_gemm_setting")
        return v.value

    def gemm_prepared_power_int4(self, A_int4, C_power, M, N, K, batch_count, alpha=1.0,
                                 beta=0.0, stream=None) -> int:
        return self._gemm_rc(lib().ftdx_gemm_prepared_power_int4(
            self._h, _ptr(A_int4), _ptr(C_power), M, N, K, batch_count, alpha, beta,
            _stream_ptr(stream)), "gemm_prepared_power_int4")

    def gemm(self, A_re, A_im, B_re, B_im, C_re, C_im, M, N, K, batch_count,
             compute: ComputePrecision = ComputePrecision.FP8,
             output: OutputPrecision = OutputPrecision.FP16, alpha=1.0, beta=0.0, stream=None,
             tune: bool = True, input: InputPrecision = InputPrecision.FP16) -> int:
        return self._gemm_rc(lib().ftdx_gemm(
            self._h, _ptr(A_re), _ptr(A_im), _ptr(B_re), _ptr(B_im), _ptr(C_re), _ptr(C_im), M, N,
            K, batch_count, int(compute), int(output), alpha, beta, _stream_ptr(stream),
            int(bool(tune)), int(input)), "gemm")

    # ---- gfx950 launch configuration and tuning cache ----
    def set_launch_config(self, tile_config: int = -1, xcd_remap: int = -1) -> None:
        _check(lib().ftdx_set_launch_config(self._h, tile_config, xcd_remap), "set_launch_config")

    def set_mx_rounding(self, rounding: MxRounding) -> None:
        """Element rounding of this context's FP6 / FP4 paths (default: the reference's)."""
        _check(lib().ftdx_set_mx_rounding(self._h, int(rounding)), "set_mx_rounding")

    def mx_rounding(self) -> MxRounding:
        return MxRounding(lib().ftdx_get_mx_rounding(self._h))

    def graph_stats(self) -> tuple[int, int]:
        """(graphs cached, graphs captured so far) of the HERK graph replay."""
        n, c = ctypes.c_int(), ctypes.c_longlong()
        _check(lib().ftdx_herk_graph_stats(self._h, ctypes.byref(n), ctypes.byref(c)), "graph_stats")
        return n.value, c.value

    def query_config(self, N: int, K: int, batch: int, output=OutputPrecision.FP32):
        c, r = ctypes.c_int(), ctypes.c_int()
        _check(lib().ftdx_herk_query_config(self._h, N, K, batch, int(output), ctypes.byref(c),
                                            ctypes.byref(r)), "query_config")
        return c.value, r.value

    def tune_load(self, path: str):
        n, sk = ctypes.c_int(), ctypes.c_int()
        _check(lib().ftdx_tune_load(self._h, path.encode(), ctypes.byref(n), ctypes.byref(sk)),
               "tune_load")
        return n.value, sk.value

    def tune_save(self, path: str) -> None:
        _check(lib().ftdx_tune_save(self._h, path.encode()), "tune_save")


def mx_layout(compute: ComputePrecision, K: int) -> tuple[int, int]:
    """(padded 32-sample blocks per row, bytes per 32-code block)."""
    b, c = ctypes.c_int(), ctypes.c_int()
    _check(lib().ftdx_mx_layout(int(compute), K, ctypes.byref(b), ctypes.byref(c)), "mx_layout")
    return b.value, c.value


def mx_quantize(X, input: InputPrecision, compute: ComputePrecision, rows: int, K: int,
                stream=None, rounding: MxRounding = MxRounding.Reference):
    """MX quantization stage on the device: returns (codes, scales) uint8 tensors of
    shapes [rows, blocks, 2, code_bytes] and [rows, blocks, 2]."""
    import torch
    blocks, cb = mx_layout(compute, K)
    codes = torch.zeros((rows, blocks, 2, cb), dtype=torch.uint8, device=X.device)
    scales = torch.zeros((rows, blocks, 2), dtype=torch.uint8, device=X.device)
    _check(lib().ftdx_mx_quantize_rounding(_ptr(X), int(input), int(compute), int(rounding), _ptr(codes),
                                           _ptr(scales), rows, K, _stream_ptr(stream)), "mx_quantize")
    return codes, scales


def herk_mx(ctx: "CutlassComplexGemm", A, N: int, K: int, batch: int,
            input: InputPrecision = InputPrecision.FP16,
            compute: ComputePrecision = ComputePrecision.FP6, output: str = "fp32", C=None,
            alpha: float = 1.0, beta: float = 0.0, stream=None):
    """MX block-scaled HERK (ftdx_herk_mx); returns C [batch, N(N+1)/2, 2]."""
    import torch
    dt = torch.float32 if output == "fp32" else torch.float16
    if C is None:
        C = torch.empty((batch, packed_len(N), 2), dtype=dt, device=A.device)
    out = OutputPrecision.FP32 if output == "fp32" else OutputPrecision.FP16
    _check(lib().ftdx_herk_mx(ctx.handle, _ptr(A), _ptr(C), N, K, batch, int(input),
                              int(compute), int(out), alpha, beta, _stream_ptr(stream)),
           "herk_mx")
    return C


def gain_loss_grad(ctx: "CutlassComplexGemm", V, M, g, N: int, batch: int, want_gamma=True,
                   stream=None):
    """Gain-calibration loss and gradients (ftdx_gain_loss_grad).  V, M: packed
    [batch, N(N+1)/2, 2] float32; g: [batch, N, 2] float32.  Returns
    (loss [batch] float64, grad_g [batch, N, 2] float32, Gamma or None)."""
    import torch
    loss = torch.empty(batch, dtype=torch.float64, device=V.device)
    grad = torch.empty((batch, N, 2), dtype=torch.float32, device=V.device)
    Gamma = torch.empty_like(M) if want_gamma else None
    _check(lib().ftdx_gain_loss_grad(ctx.handle, _ptr(V), _ptr(M), _ptr(g), N, batch, _ptr(loss),
                                     _ptr(grad), _ptr(Gamma) if Gamma is not None else None,
                                     _stream_ptr(stream)), "gain_loss_grad")
    return loss, grad, Gamma


GAIN_NOALLOC_SCRATCH_BYTES = 16 << 20


def gain_loss_grad_noalloc(V, M, g, N: int, batch: int, loss, grad, Gamma=None, stream=None) -> None:
    """ftdx_gain_loss_grad_noalloc: the gain loss and gradients into caller-provided
    outputs (loss [batch] float64, grad [batch, N, 2] float32, Gamma like M or None);
    the library allocates nothing (fixed per-device scratch)."""
    _check(lib().ftdx_gain_loss_grad_noalloc(_ptr(V), _ptr(M), _ptr(g), N, batch, _ptr(loss), _ptr(grad),
                                             _ptr(Gamma) if Gamma is not None else None,
                                             _stream_ptr(stream)), "gain_loss_grad_noalloc")


def herk_adjoint(ctx: "CutlassComplexGemm", Gamma, X, N: int, K: int, batch: int, stream=None):
    """dX = (Gamma + Gamma^H) X (ftdx_herk_adjoint); X QC bytes; returns
    dX [batch, N, K, 2] float32."""
    import torch
    dX = torch.empty((batch, N, K, 2), dtype=torch.float32, device=X.device)
    _check(lib().ftdx_herk_adjoint(ctx.handle, _ptr(Gamma), _ptr(X), _ptr(dX), N, K, batch,
                                   int(InputPrecision.INT4), _stream_ptr(stream)), "herk_adjoint")
    return dX


# ---- stand-alone stage kernels (return new device tensors) ----
def promote_qc_to_fp8(qc, stream=None):
    """QC bytes -> [n, 2] uint8 FP8 E4M3 bytes (Re, Im)."""
    import torch
    out = torch.empty((qc.numel(), 2), dtype=torch.uint8, device=qc.device)
    _check(lib().ftdx_promote_qc_to_fp8(_ptr(qc), _ptr(out), qc.numel(), _stream_ptr(stream)),
           "promote_qc_to_fp8")
    return out


def promote_qc_to_fp16_planar(qc, stream=None):
    import torch
    re = torch.empty(qc.numel(), dtype=torch.float16, device=qc.device)
    im = torch.empty_like(re)
    _check(lib().ftdx_promote_qc_to_fp16_planar(_ptr(qc), _ptr(re), _ptr(im), qc.numel(),
                                                _stream_ptr(stream)), "promote_qc_to_fp16_planar")
    return re, im


def promote_data(data, fp64=False, stream=None):
    """GGP two's-complement (low nibble Re) bytes -> [n, 2] float32/float64."""
    import torch
    out = torch.empty((data.numel(), 2), dtype=torch.float64 if fp64 else torch.float32,
                      device=data.device)
    _check(lib().ftdx_promote_data(_ptr(data), _ptr(out), data.numel(),
                                   PREC_FP64 if fp64 else PREC_FP32, _stream_ptr(stream)),
           "promote_data")
    return out


def pack_triangle(full, N: int, batch: int, old=None, beta: float = 0.0, stream=None):
    """[batch, N, N, 2] float32 -> packed lower [batch, N(N+1)/2, 2] (+ beta old)."""
    import torch
    out = torch.empty((batch, packed_len(N), 2), dtype=torch.float32, device=full.device)
    _check(lib().ftdx_pack_triangle(_ptr(full), _ptr(out), _ptr(old) if old is not None else None,
                                    N, batch, beta, _stream_ptr(stream)), "pack_triangle")
    return out


def tri_from_herm(full, N: int, batch: int, stream=None):
    """[batch, N, N, 2] float32/float64 -> packed upper row-major [batch, N(N+1)/2, 2]."""
    import torch
    out = torch.empty((batch, packed_len(N), 2), dtype=full.dtype, device=full.device)
    prec = PREC_FP64 if full.dtype == torch.float64 else PREC_FP32
    _check(lib().ftdx_tri_from_herm(_ptr(full), _ptr(out), N, batch, prec, _stream_ptr(stream)),
           "tri_from_herm")
    return out


def interleave_tri(re, im, fp64=False, stream=None):
    import torch
    out = torch.empty((re.numel(), 2), dtype=torch.float64 if fp64 else torch.float32,
                      device=re.device)
    _check(lib().ftdx_interleave_tri(_ptr(re), _ptr(im), _ptr(out), re.numel(),
                                     PREC_FP64 if fp64 else PREC_FP32, _stream_ptr(stream)),
           "interleave_tri")
    return out


def promote_qc_to_fp32(qc, stream=None):
    import torch
    out = torch.empty((qc.numel(), 2), dtype=torch.float32, device=qc.device)
    _check(lib().ftdx_promote_qc_to_fp32(_ptr(qc), _ptr(out), qc.numel(), _stream_ptr(stream)),
           "promote_qc_to_fp32")
    return out


def convert_qc_to_ftd(qc, out=None, stream=None):
    import torch
    out = torch.empty_like(qc) if out is None else out
    _check(lib().ftdx_convert_qc_to_ftd(_ptr(qc), _ptr(out), qc.numel(), _stream_ptr(stream)),
           "convert_qc_to_ftd")
    return out


def problem_array(problems):
    """ctypes array of ftdx_herk_problem from (A, C, N, K[, alpha, beta]) tuples."""
    arr = (HerkProblem * max(1, len(problems)))()
    for i, p in enumerate(problems):
        A, C, N, K = p[0], p[1], p[2], p[3]
        alpha = p[4] if len(p) > 4 else 1.0
        beta = p[5] if len(p) > 5 else 0.0
        arr[i] = HerkProblem(_ptr(A) if A is not None else None, _ptr(C), N, K, alpha, beta)
    return arr


def herk_each(ctx: "CutlassComplexGemm", problems, output=None, stream=None, arr=None) -> None:
    """One plain HERK launch per problem, in order (ftdx_herk_each)."""
    output = OutputPrecision.FP32 if output is None else output
    arr = problem_array(problems) if arr is None else arr
    _check(lib().ftdx_herk_each(ctx.handle, arr, len(problems), int(output),
                                _stream_ptr(stream)), "herk_each")


class GroupedPlan:
    """Grouped HERK plan over many (A, C, N, K, alpha, beta) problems
    (ftdx_herk_grouped_plan_*).  Keeps references to the tensors it points at."""

    def __init__(self, ctx: CutlassComplexGemm, problems, output=OutputPrecision.FP32,
                 tile_config: int = TILE_AUTO, use_graph: bool = True):
        arr = problem_array(problems)
        self._keep = [t for p in problems for t in (p[0], p[1])]
        h = ctypes.c_void_p()
        _check(lib().ftdx_herk_grouped_plan_create(ctx.handle, arr, len(problems), int(output),
                                                    tile_config, int(bool(use_graph)),
                                                    ctypes.byref(h)), "grouped_plan_create")
        self._h = h

    def run(self, stream=None) -> None:
        _check(lib().ftdx_herk_grouped_plan_run(self._h, _stream_ptr(stream)), "grouped_plan_run")

    def info(self) -> dict:
        g, n, c, imb = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int(), ctypes.c_double()
        _check(lib().ftdx_herk_grouped_plan_info(self._h, ctypes.byref(g), ctypes.byref(n),
                                                  ctypes.byref(c), ctypes.byref(imb)),
               "grouped_plan_info")
        return {"workgroups": g.value, "items": n.value, "tile_config": c.value,
                "imbalance": imb.value}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().ftdx_herk_grouped_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def herk_int4(A, N: int, K: int, batch: int, C=None, output: str = "fp32",
              alpha: float = 1.0, beta: float = 0.0, stream=None):
    """Plain batched HERK of QC bytes A [batch, N, K] (uint8, device).

    Returns C [batch, N(N+1)/2, 2] (float32 or float16) on the same device.
    """
    import torch
    if A.dtype != torch.uint8:
        raise ValueError("herk_int4: A must be uint8 QC bytes")
    if A.numel() != N * K * batch:
        raise ValueError("herk_int4: A has the wrong number of elements")
    dt = torch.float32 if output == "fp32" else torch.float16
    if C is None:
        C = torch.empty((batch, packed_len(N), 2), dtype=dt, device=A.device)
        if beta != 0.0:
            raise ValueError("herk_int4: beta != 0 needs an input C")
    elif C.dtype != dt or C.numel() != batch * packed_len(N) * 2:
        raise ValueError("herk_int4: C has the wrong dtype or size")
    out = int(OutputPrecision.FP32 if output == "fp32" else OutputPrecision.FP16)
    _check(lib().ftdx_herk_int4(_ptr(A), _ptr(C), N, K, batch, out, alpha, beta,
                                _stream_ptr(stream)), "ftdx_herk_int4")
    return C


# ---- GGP glue (lib/cutlass_interface.cu:150-289) ----
_ggp_api: CutlassComplexGemm | None = None
_ggp_tune = False


def _ggp() -> CutlassComplexGemm:
    global _ggp_api
    if _ggp_api is None:
        _ggp_api = CutlassComplexGemm()
    return _ggp_api


def initCutlassApi(kernel_tune_verbosity: int, kernel_tune_cache_path: str | None,
                   strategy_tune_cache_path: str | None, herk_tune: bool,
                   strategy_tune_verbosity: int = 1) -> None:
    global _ggp_tune
    a = _ggp()
    if kernel_tune_verbosity > 0:
        a.set_kernel_tune_verbosity(kernel_tune_verbosity)
    a.set_strategy_tune_verbosity(strategy_tune_verbosity)
    if kernel_tune_cache_path:
        a.set_kernel_tune_cache_path(kernel_tune_cache_path)
    if strategy_tune_cache_path:
        a.set_tune_cache_path(strategy_tune_cache_path)
    _ggp_tune = bool(herk_tune)


def herkBatchedCutlassQC(qc_data, tri_output, N: int, K: int, batch: int, stream=None) -> int:
    return _ggp().herk(qc_data, tri_output, N, K, batch, InputPrecision.INT4,
                       ComputePrecision.FP8, OutputPrecision.FP32, OutputFormat.PackedTriangle,
                       1.0, 0.0, stream, _ggp_tune)


def herkBatchedCutlassQC_FP16(qc_data, tri_output, N: int, K: int, batch: int,
                              stream=None) -> int:
    return _ggp().herk(qc_data, tri_output, N, K, batch, InputPrecision.INT4,
                       ComputePrecision.FP8, OutputPrecision.FP16, OutputFormat.PackedTriangle,
                       1.0, 0.0, stream, _ggp_tune)


# ---- FTD raw-order correlation (ftdx_correlate_ftd) ----
def correlate_ftd(ctx: "CutlassComplexGemm", raw, n_ant: int, n_chan: int, n_time: int,
                  n_time_inner: int, n_pol: int, twos: bool = True, out=None, stream=None):
    """FTD [time][ant][chan][pol] bytes (device uint8) -> packed lower triangle FP32
    [n_chan*n_pol*n_time_inner, n_ant(n_ant+1)/2, 2] (correlateTCC's contract,
    lib/tcc_interface.cu:251-314)."""
    import torch
    batch = n_chan * n_pol * n_time_inner
    if out is None:
        out = torch.empty((batch, packed_len(n_ant), 2), dtype=torch.float32, device=raw.device)
    _check(lib().ftdx_correlate_ftd(ctx.handle, _ptr(raw), _ptr(out), n_ant, n_chan, n_time,
                                    n_time_inner, n_pol, int(bool(twos)), _stream_ptr(stream)),
           "correlate_ftd")
    return out


# ---- GGP XEngine (lib/xengine.cpp:22-306) ----
LOCATION_HOST = 1
LOCATION_DEVICE = 2
PACKET_DSA110 = 4608
PACKET_CASM = 46080
PACKET_DSA2K = 8192
QPREC_QUARTER, QPREC_HALF, QPREC_SINGLE, QPREC_DOUBLE = 1, 2, 4, 8
BLAS_ENGINE_CUBLAS, BLAS_ENGINE_CUTLASS, BLAS_ENGINE_TCC = 0, 1, 2
XENGINE_MAT_HERM, XENGINE_MAT_TRI = 0, 1


class XEngineParam(ctypes.Structure):
    """ftdx_xengine_param (mirrors XEngineParam, include/ggp.h:101-123)."""
    _fields_ = [("struct_size", ctypes.c_size_t), ("in_location", ctypes.c_int),
                ("out_location", ctypes.c_int), ("packet_format", ctypes.c_int),
                ("n_payload", ctypes.c_size_t), ("n_antennae_per_payload", ctypes.c_size_t),
                ("n_channels_per_payload", ctypes.c_size_t), ("n_time_per_payload", ctypes.c_size_t),
                ("n_time_inner", ctypes.c_size_t), ("n_polarizations", ctypes.c_size_t),
                ("compute_prec", ctypes.c_int), ("output_prec", ctypes.c_int),
                ("engine", ctypes.c_int), ("format", ctypes.c_int)]


def new_xengine_param(**fields) -> XEngineParam:
    """newXEngineParam defaults (lib/check_params.h:104-145), then `fields`; sets
    struct_size."""
    p = XEngineParam()
    _check(lib().ftdx_xengine_param_init(ctypes.byref(p)), "xengine_param_init")
    for k, v in fields.items():
        setattr(p, k, v)
    p.struct_size = ctypes.sizeof(XEngineParam)
    return p


def _host_or_device_ptr(x, location: int, what: str):
    if location == LOCATION_HOST:
        import numpy as np
        if not isinstance(x, np.ndarray) or not x.flags["C_CONTIGUOUS"]:
            raise ValueError(f"XEngine: host {what} must be a C-contiguous numpy array")
        return x.ctypes.data
    return _ptr(x)


class XEngine:
    """ggp::XEngine over ftdx_xengine_* (lib/xengine.cpp:22-224).  Host locations
    take numpy arrays, device locations device tensors."""

    def __init__(self, param: XEngineParam):
        self.param = param
        h = _VP()
        _check(lib().ftdx_xengine_create(ctypes.byref(param), ctypes.byref(h)), "xengine_create")
        self.handle = h.value
        ip, op = ctypes.c_size_t(), ctypes.c_size_t()
        n, k, b = _I(), _I(), _I()
        _check(lib().ftdx_xengine_query(self.handle, ctypes.byref(ip), ctypes.byref(op),
                                        ctypes.byref(n), ctypes.byref(k), ctypes.byref(b)),
               "xengine_query")
        self.in_payload_bytes, self.out_payload_bytes = ip.value, op.value
        self.N, self.K, self.batch = n.value, k.value, b.value

    def set_launch_config(self, tile_config: int = -1, xcd_remap: int = -1) -> None:
        _check(lib().ftdx_xengine_set_launch_config(self.handle, tile_config, xcd_remap),
               "xengine_set_launch_config")

    def compute(self, output, input_data, stream=None) -> None:
        p = self.param
        nin = self.in_payload_bytes * p.n_payload
        nout = self.out_payload_bytes * p.n_payload
        for x, n, what in ((input_data, nin, "input"), (output, nout, "output")):
            nbytes = x.nbytes if hasattr(x, "nbytes") else x.numel() * x.element_size()
            if nbytes < n:
                raise ValueError(f"XEngine: {what} holds {nbytes} bytes, needs {n}")
        sp = None
        if p.in_location == LOCATION_DEVICE or p.out_location == LOCATION_DEVICE:
            sp = _stream_ptr(stream)
        _check(lib().ftdx_xengine_compute(
            self.handle, _host_or_device_ptr(output, p.out_location, "output"),
            _host_or_device_ptr(input_data, p.in_location, "input"), sp), "xengine_compute")

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().ftdx_xengine_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def xengine_chpc(output, input_data, param: XEngineParam) -> None:
    """XEngineCHPC (lib/interface_ggp.cpp:536-555): one-shot create/compute/destroy."""
    _check(lib().ftdx_xengine_chpc(_host_or_device_ptr(output, param.out_location, "output"),
                                   _host_or_device_ptr(input_data, param.in_location, "input"),
                                   ctypes.byref(param)), "xengine_chpc")


# ---- streaming correlation of host payloads (ftdx_pipeline_*) ----
class Pipeline:
    """VisibilityPipeline::correlate / correlate_pipelined
    (lib/visibility_pipeline.cu:253-300, 484-545) over host numpy payloads."""

    def __init__(self, ctx: "CutlassComplexGemm", N: int, K: int, batch: int,
                 output: str = "fp32"):
        self.ctx = ctx  # keeps the context alive
        self.N, self.K, self.batch = N, K, batch
        self.output = output
        h = _VP()
        out = int(OutputPrecision.FP32 if output == "fp32" else OutputPrecision.FP16)
        _check(lib().ftdx_pipeline_create(ctx.handle, N, K, batch, out, ctypes.byref(h)),
               "pipeline_create")
        self.handle = h.value

    def out_shape(self):
        return (self.batch, packed_len(self.N), 2)

    def _out_dtype(self):
        import numpy as np
        return np.float32 if self.output == "fp32" else np.float16

    def correlate_many(self, raw_inputs, outputs=None):
        import numpy as np
        n = len(raw_inputs)
        ins = [np.ascontiguousarray(x, dtype=np.uint8).reshape(-1) for x in raw_inputs]
        for x in ins:
            if x.size != self.N * self.K * self.batch:
                raise ValueError("Pipeline: payload has the wrong number of bytes")
        if outputs is None:
            outputs = [np.empty(self.out_shape(), dtype=self._out_dtype()) for _ in range(n)]
        for o in outputs:
            if not o.flags["C_CONTIGUOUS"] or o.nbytes < self.batch * packed_len(self.N) * 2 * \
                    np.dtype(self._out_dtype()).itemsize:
                raise ValueError("Pipeline: output too small or not contiguous")
        arr_in = (_VP * max(1, n))(*[x.ctypes.data for x in ins])
        arr_out = (_VP * max(1, n))(*[o.ctypes.data for o in outputs])
        _check(lib().ftdx_pipeline_correlate_many(self.handle, arr_out, arr_in, n),
               "pipeline_correlate_many")
        return outputs

    def correlate(self, raw_input, output=None):
        return self.correlate_many([raw_input], None if output is None else [output])[0]

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().ftdx_pipeline_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- voltage beamformer (ftdx_voltage_*) ----
VOLTAGE_FP8 = 0
VOLTAGE_FP6 = 1


class VoltageConfig(ctypes.Structure):
    """ftdx_voltage_config (VoltagePipeline::Config, include/voltage_pipeline.h:121-156)."""
    _fields_ = [("struct_size", ctypes.c_size_t), ("n_antennae", ctypes.c_uint64),
                ("n_beams", ctypes.c_uint64), ("n_channels", ctypes.c_uint64), ("n_time", ctypes.c_uint64),
                ("n_polarizations", ctypes.c_uint64), ("n_time_power_sum", ctypes.c_uint64),
                ("compute_mode", ctypes.c_int), ("kernel_tune_verbosity", ctypes.c_int),
                ("strategy_tune_verbosity", ctypes.c_int), ("gemm_tune", ctypes.c_int),
                ("kernel_tune_cache_path", ctypes.c_char_p), ("strategy_tune_cache_path", ctypes.c_char_p),
                ("n_dm_trials", ctypes.c_int), ("f_min_MHz", ctypes.c_float), ("f_max_MHz", ctypes.c_float),
                ("max_dm", ctypes.c_float), ("total_obs_time_s", ctypes.c_float), ("profile", ctypes.c_int),
                ("ch_fuse", ctypes.c_uint64), ("max_payloads", ctypes.c_uint64)]


def voltage_config(**fields) -> VoltageConfig:
    """Config defaults (include/voltage_pipeline.h:123-156), then `fields`."""
    c = VoltageConfig()
    _check(lib().ftdx_voltage_config_init(ctypes.byref(c)), "voltage_config_init")
    for k, v in fields.items():
        setattr(c, k, v)
    return c


class VoltagePipeline:
    """ggp::VoltagePipeline over ftdx_voltage_* (include/voltage_pipeline.h:49-310).
    Device tensors: QC input uint8 [n_ch, 2, n_ant, n_time], weights float32
    [n_ch, n_beam, n_ant, 2], uint8 output [n_ch, n_beam, n_time_out]."""

    def __init__(self, config: VoltageConfig | None = None, **fields):
        self.config = config if config is not None else voltage_config(**fields)
        h = _VP()
        _check(lib().ftdx_voltage_create(ctypes.byref(self.config), ctypes.byref(h)), "VoltagePipeline")
        self.handle = h.value
        osz, isz = ctypes.c_int64(), ctypes.c_int64()
        nto = ctypes.c_uint64()
        direct, graph = _I(), _I()
        fc, fb = _VP(), _VP()
        _check(lib().ftdx_voltage_query(self.handle, ctypes.byref(osz), ctypes.byref(isz), ctypes.byref(nto),
                                        ctypes.byref(direct), ctypes.byref(graph), ctypes.byref(fc),
                                        ctypes.byref(fb)), "voltage_query")
        self.output_size, self.input_size, self.n_time_out = osz.value, isz.value, nto.value
        self.uses_direct_path = bool(direct.value)
        self._fb_ptrs = (fc.value, fb.value)

    def _need(self, t, nbytes: int, what: str) -> None:
        if t.numel() * t.element_size() < nbytes:
            raise ValueError(f"VoltagePipeline: {what} holds {t.numel() * t.element_size()} bytes, "
                             f"needs {nbytes}")

    def set_weights(self, weights, stream=None) -> None:
        c = self.config
        self._need(weights, 8 * c.n_channels * c.n_beams * c.n_antennae, "weights")
        _check(lib().ftdx_voltage_set_weights(self.handle, _ptr(weights), _stream_ptr(stream)), "set_weights")

    def beamform(self, output, qc_input, scale: float = 1.0, stream=None) -> None:
        self._need(output, self.output_size, "output")
        self._need(qc_input, self.input_size, "input")
        _check(lib().ftdx_voltage_beamform(self.handle, _ptr(output), _ptr(qc_input), scale, _stream_ptr(stream)),
               "beamform")

    def beamform_batched(self, output, qc_inputs, scale: float = 1.0, stream=None) -> None:
        n = len(qc_inputs)
        for x in qc_inputs:
            self._need(x, self.input_size, "input")
        self._need(output, self.output_size * max(n, 1), "output")
        arr = (_VP * max(1, n))(*[_ptr(x) for x in qc_inputs])
        _check(lib().ftdx_voltage_beamform_batched(self.handle, _ptr(output), arr, n, scale, _stream_ptr(stream)),
               "beamform_batched")

    def compute_filterbank(self, qc_input, stream=None) -> None:
        self._need(qc_input, self.input_size, "input")
        _check(lib().ftdx_voltage_compute_filterbank(self.handle, _ptr(qc_input), _stream_ptr(stream)),
               "compute_filterbank")

    def compute(self, output, qc_input, dedisp_out=None, scale: float = 1.0, stream=None) -> None:
        self._need(output, self.output_size, "output")
        self._need(qc_input, self.input_size, "input")
        _check(lib().ftdx_voltage_compute(self.handle, _ptr(output), _ptr(qc_input), _ptr(dedisp_out), scale,
                                          _stream_ptr(stream)), "compute")

    def corner_turn(self, out, inp, stream=None) -> None:
        _check(lib().ftdx_voltage_corner_turn(self.handle, _ptr(out), _ptr(inp), _stream_ptr(stream)),
               "corner_turn")

    def set_graph_enabled(self, enabled: bool) -> None:
        _check(lib().ftdx_voltage_set_graph_enabled(self.handle, int(bool(enabled))), "set_graph_enabled")

    def filterbank(self, which: str = "chanmajor"):
        """Copy of the FP32 filterbank buffer ([n_ch, n_beam, n_time_out] channel-major or
        [n_beam, n_ch, n_time_out] beam-major) as a device tensor (stream-ordered on the
        current stream)."""
        import torch
        c = self.config
        n = c.n_channels * c.n_beams * self.n_time_out
        out = torch.empty(n, dtype=torch.float32, device="cuda")
        src = self._fb_ptrs[0 if which == "chanmajor" else 1]
        # a 1 x 1 corner turn of n values is a plain device copy
        _check(lib().ftdx_corner_turn(_ptr(out), src, 1, 1, n, _stream_ptr(None)), "filterbank copy")
        return out

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().ftdx_voltage_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def corner_turn(out, inp, n_channels: int, n_beams: int, n_time: int, stream=None) -> None:
    """ggp::corner_turn_nf_nb (lib/corner_turn.cu:63-75): [Nf, Nb, Nt] -> [Nb, Nf, Nt] FP32."""
    _check(lib().ftdx_corner_turn(_ptr(out), _ptr(inp), n_channels, n_beams, n_time, _stream_ptr(stream)),
           "corner_turn")


# ---- visibility consumers: polarization sum and pillbox gridding ----
def pol_reduce_triangle(tri, N: int, n_ch: int, n_pol: int, n_time_inner: int, out=None, stream=None):
    """[n_ch*n_pol*n_ti, N(N+1)/2, 2] triangles -> [n_ch*n_ti, N(N+1)/2, 2] summed over pol."""
    import torch
    if out is None:
        out = torch.empty((n_ch * n_time_inner, packed_len(N), 2), dtype=tri.dtype, device=tri.device)
    o = int(OutputPrecision.FP16 if tri.dtype == torch.float16 else OutputPrecision.FP32)
    _check(lib().ftdx_pol_reduce_triangle(_ptr(tri), _ptr(out), N, n_ch, n_pol, n_time_inner, o,
                                          _stream_ptr(stream)), "pol_reduce_triangle")
    return out


def herk_pol_reduce(ctx: "CutlassComplexGemm", A, N: int, K: int, n_ch: int, n_pol: int,
                    n_time_inner: int, output: str = "fp32", C=None, alpha: float = 1.0,
                    beta: float = 0.0, stream=None):
    """Fused HERK + polarization sum (ftdx_herk_pol_reduce)."""
    import torch
    dt = torch.float32 if output == "fp32" else torch.float16
    if C is None:
        C = torch.empty((n_ch * n_time_inner, packed_len(N), 2), dtype=dt, device=A.device)
    o = int(OutputPrecision.FP32 if output == "fp32" else OutputPrecision.FP16)
    _check(lib().ftdx_herk_pol_reduce(ctx.handle, _ptr(A), _ptr(C), N, K, n_ch, n_pol, n_time_inner, o,
                                      alpha, beta, _stream_ptr(stream)), "herk_pol_reduce")
    return C


def pillbox_grid(vis_tri, baseline_uv_m, freq_hz, grid, n_baselines: int, Ng: int, Nf: int,
                 cell_size_rad: float, freq_offset: int = 0, stream=None):
    """Scatter-add Nf packed triangles into grid [Nf, Ng, Ng, 2] float32 (ftdx_pillbox_grid)."""
    import torch
    o = int(OutputPrecision.FP16 if vis_tri.dtype == torch.float16 else OutputPrecision.FP32)
    _check(lib().ftdx_pillbox_grid(_ptr(vis_tri), o, _ptr(baseline_uv_m), _ptr(freq_hz), _ptr(grid),
                                   n_baselines, Ng, Nf, freq_offset, cell_size_rad, _stream_ptr(stream)),
           "pillbox_grid")
    return grid


def herk_pol_grid(ctx: "CutlassComplexGemm", A, N: int, K: int, n_chan: int, n_pol: int, baseline_uv_m,
                  freq_hz, grid, Ng: int, cell_size_rad: float, freq_offset: int = 0, alpha: float = 1.0,
                  stream=None):
    """HERK + polarization sum + pillbox gridding in one pass (ftdx_herk_pol_grid): QC bytes
    A [n_chan, n_pol, N, K] -> adds into grid [n_chan, Ng, Ng, 2] float32 (caller zeroes it)."""
    _check(lib().ftdx_herk_pol_grid(ctx.handle, _ptr(A), N, K, n_chan, n_pol, _ptr(baseline_uv_m),
                                    _ptr(freq_hz), _ptr(grid), Ng, freq_offset, cell_size_rad, alpha,
                                    _stream_ptr(stream)), "herk_pol_grid")
    return grid
