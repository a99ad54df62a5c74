"""CPU checks of the drop-in boundary: the C-ABI library loads and exports
every function include/ftdx.h declares; error paths that need no device; the
public headers compile as C / C++; the C++ wrapper library exports the
reference-named classes and functions.  No kernel is launched here."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    from ftd_amd import ftdx
    L = ftdx.lib()
    declared = ftdx.declared_functions()
    assert "ftdx_herk" in declared and "ftdx_last_error" in declared
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding types every declared function
    assert set(declared) == set(ftdx._SIGNATURES), set(declared) ^ set(ftdx._SIGNATURES)


def test_error_paths_without_device():
    from ftd_amd import ftdx
    L = ftdx.lib()
    assert L.ftdx_abi_version() == 1
    assert L.ftdx_set_herk_mode(None, 0) == ftdx.FTDX_ERR_INVALID_ARG
    assert "NULL context" in ftdx.last_error()
    rc = L.ftdx_herk(None, None, None, 4, 4, 1, 1, 0, 1, 0, 1.0, 0.0, None, 0)
    assert rc == ftdx.FTDX_ERR_INVALID_ARG
    assert L.ftdx_create(None) == ftdx.FTDX_ERR_INVALID_ARG
    assert L.ftdx_destroy(None) == ftdx.FTDX_OK
    assert L.ftdx_set_tune_cache_path(None, b"x") == ftdx.FTDX_ERR_INVALID_ARG


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from ftd_amd import ftdx
    monkeypatch.setattr(ftdx, "_lib", None)
    monkeypatch.setattr(ftdx, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(RuntimeError, match="not found"):
        ftdx.lib()


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no host C compiler")
def test_headers_compile():
    inc = os.path.join(ROOT, "include")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-x", "c",
                    os.path.join(inc, "ftdx.h")], check=True)
    for h in ("cutlass_gemm_api.h", "ggp_cutlass_interface.h", "ggp_xengine.h", "voltage_pipeline.h"):
        subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-x", "c++",
                        "-I", inc, "-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-"],
                       input=f'#include "{h}"\n', text=True, check=True)


def test_cpp_wrapper_exports():
    so = os.path.join(ROOT, "ftd_amd", "libftdx_api.so")
    out = subprocess.run(["nm", "-DC", "--defined-only", so], capture_output=True, text=True,
                         check=True).stdout
    for sym in ["cutlass_gemm_api::CutlassComplexGemm::herk(",
                "cutlass_gemm_api::CutlassComplexGemm::init_herk(",
                "cutlass_gemm_api::CutlassComplexGemm::set_herk_mode(",
                "ggp::initCutlassApi(", "ggp::herkBatchedCutlassQC(",
                "ggp::herkBatchedCutlassQC_FP16("]:
        assert sym in out, sym


def test_tune_cache_file_parsing(tmp_path):
    """Malformed lines and lines of another build fingerprint are skipped
    (the reference's cache parser, strategy_cache.hpp:143-420, skips bad lines)."""
    from ftd_amd import ftdx
    ctx = ftdx.CutlassComplexGemm()  # no device needed for settings / cache I/O
    good = tmp_path / "good.txt"
    ctx.tune_save(str(good))
    header = open(good).read().strip()
    assert header.startswith("# ftdx-tune gfx950-abi1")
    f = tmp_path / "cache.txt"
    f.write_text("\n".join([
        header,
        "herk\t4096\t2048\t96\t0\t1\t3\t1\t7.5",
        "herk\t128\t1024\t16\t0\t1\t0\t0\t0.1",
        "herk\t256\t512\t4\t0\t1\t20\t1\t0.5",         # tile 4, E3M2 operands
        "herk\t1\t2",                                  # too short
        "herk\t10\t10\t1\t0\t1\t9\t0\t1.0",            # unknown tile config
        "gemm\t10\t10\t1\t0\t1\t0\t0\t1.0",            # other kind
        "herk\t10\t10\t1\t0\t1\t0\t0\t1.0\textra",     # trailing field
        "# ftdx-tune some-other-build",
        "herk\t64\t64\t1\t0\t1\t1\t0\t0.01",           # foreign fingerprint
    ]) + "\n")
    n, skipped = ctx.tune_load(str(f))
    assert (n, skipped) == (3, 5)
    assert ctx.query_config(4096, 2048, 96) == (3, 1)
    assert ctx.query_config(128, 1024, 16) == (0, 0)
    assert ctx.query_config(256, 512, 4) == (4 | ftdx.TILE_BF6, 1)
    assert ctx.query_config(64, 64, 1) == (0, 1)       # default
    ctx.set_launch_config(2, 0)
    assert ctx.query_config(4096, 2048, 96) == (2, 0)  # explicit setting wins
    ctx.set_launch_config(-1, -1)
    ctx.set_direct_tile_config(ftdx.DirectTileConfig.K128_B2)
    assert ctx.query_config(4096, 2048, 96)[0] == 1
    out = tmp_path / "round.txt"
    ctx.tune_save(str(out))
    ctx2 = ftdx.CutlassComplexGemm()
    assert ctx2.tune_load(str(out)) == (3, 0)
    with pytest.raises(ftdx.FtdxError):
        ctx2.tune_load(str(tmp_path / "missing.txt"))


def test_xengine_param_validation_without_device():
    """ftdx_xengine_create rejects what the reference XEngine rejects
    (lib/xengine.cpp:34-88, lib/check_params.h:118-119), before touching a device."""
    import ctypes
    from ftd_amd import ftdx
    L = ftdx.lib()
    base = dict(n_payload=1, n_antennae_per_payload=8, n_channels_per_payload=1,
                n_time_per_payload=32, n_time_inner=2, n_polarizations=2,
                compute_prec=ftdx.QPREC_SINGLE, output_prec=ftdx.QPREC_SINGLE,
                format=ftdx.XENGINE_MAT_TRI)

    def create(struct_size=None, **kw):
        p = ftdx.new_xengine_param(**{**base, **kw})
        if struct_size is not None:
            p.struct_size = struct_size
        h = ctypes.c_void_p()
        rc = L.ftdx_xengine_create(ctypes.byref(p), ctypes.byref(h))
        assert h.value is None
        return rc, ftdx.last_error()

    d = ftdx.new_xengine_param()
    assert (d.compute_prec, d.output_prec, d.packet_format, d.format, d.engine) == (
        ftdx.QPREC_HALF, ftdx.QPREC_SINGLE, ftdx.PACKET_DSA2K, ftdx.XENGINE_MAT_HERM,
        ftdx.BLAS_ENGINE_CUBLAS)
    rc, msg = create(compute_prec=ftdx.QPREC_HALF)
    assert rc == ftdx.FTDX_ERR_UNSUPPORTED and "compute precision" in msg
    rc, msg = create(output_prec=ftdx.QPREC_DOUBLE)
    assert rc == ftdx.FTDX_ERR_UNSUPPORTED and "must match" in msg
    rc, msg = create(packet_format=ftdx.PACKET_DSA110)
    assert rc == ftdx.FTDX_ERR_UNSUPPORTED and "DSA110" in msg
    assert create(packet_format=ftdx.PACKET_CASM)[0] == ftdx.FTDX_ERR_INVALID_ARG
    assert create(format=7)[0] == ftdx.FTDX_ERR_INVALID_ARG
    rc, msg = create(engine=ftdx.BLAS_ENGINE_TCC, format=ftdx.XENGINE_MAT_HERM)
    assert rc == ftdx.FTDX_ERR_INVALID_ARG and "triangular" in msg
    assert create(engine=ftdx.BLAS_ENGINE_TCC, compute_prec=ftdx.QPREC_DOUBLE,
                  output_prec=ftdx.QPREC_DOUBLE)[0] == ftdx.FTDX_ERR_UNSUPPORTED
    assert create(engine=9)[0] == ftdx.FTDX_ERR_INVALID_ARG
    assert create(n_payload=0)[0] == ftdx.FTDX_ERR_INVALID_ARG
    assert create(in_location=3)[0] == ftdx.FTDX_ERR_INVALID_ARG
    rc, msg = create(struct_size=3)
    assert rc == ftdx.FTDX_ERR_INVALID_ARG and "struct size" in msg
    assert L.ftdx_xengine_create(None, ctypes.byref(ctypes.c_void_p())) == ftdx.FTDX_ERR_INVALID_ARG
    assert L.ftdx_xengine_destroy(None) == ftdx.FTDX_OK
    assert L.ftdx_correlate_ftd(None, None, None, 4, 1, 4, 1, 2, 1, None) == ftdx.FTDX_ERR_INVALID_ARG


def test_pipeline_validation_without_device():
    import ctypes
    from ftd_amd import ftdx
    L = ftdx.lib()
    h = ctypes.c_void_p()
    assert L.ftdx_pipeline_create(None, 8, 8, 1, 1, ctypes.byref(h)) == ftdx.FTDX_ERR_INVALID_ARG
    assert L.ftdx_pipeline_create(None, 8, 8, 1, 1, None) == ftdx.FTDX_ERR_INVALID_ARG
    assert L.ftdx_pipeline_correlate_many(None, None, None, 1) == ftdx.FTDX_ERR_INVALID_ARG
    assert L.ftdx_pipeline_destroy(None) == ftdx.FTDX_OK


def test_gemm_validation_without_device():
    from ftd_amd import ftdx
    L = ftdx.lib()
    assert L.ftdx_prepare_b(None, None, None, 8, 8, 1, 0, None) == ftdx.FTDX_ERR_INVALID_ARG
    assert L.ftdx_gemm_prepared_direct(None, None, None, None, None, 8, 8, 8, 1, 1.0, 0.0,
                                       None) == ftdx.FTDX_ERR_INVALID_ARG
    assert "NULL context" in ftdx.last_error()
    assert L.ftdx_gemm_prepared_power_int4(None, None, None, 8, 8, 8, 1, 1.0, 0.0,
                                           None) == ftdx.FTDX_ERR_INVALID_ARG
    assert L.ftdx_gemm(None, None, None, None, None, None, None, 8, 8, 8, 1, 0, 1, 1.0, 0.0, None,
                       0, 0) == ftdx.FTDX_ERR_INVALID_ARG


def test_gemm_dispatch_settings_round_trip():
    """The reference's set_gemm_mode / get_gemm_mode, set_power_mode, set_tile_config,
    set_auto_tune (cutlass_gemm_api.h:656-720) keep their values; out-of-range values and
    unknown settings are rejected."""
    from ftd_amd import ftdx
    ctx = ftdx.CutlassComplexGemm()  # settings only: no device needed
    assert ctx.get_gemm_setting(ftdx.GEMM_SETTING_GEMM_MODE) == 0
    ctx.set_gemm_setting(ftdx.GEMM_SETTING_GEMM_MODE, 2)
    assert ctx.get_gemm_setting(ftdx.GEMM_SETTING_GEMM_MODE) == 2
    ctx.set_gemm_setting(ftdx.GEMM_SETTING_TILE_CONFIG, 8)
    ctx.set_gemm_setting(ftdx.GEMM_SETTING_AUTO_TUNE, 0)
    assert ctx.get_gemm_setting(ftdx.GEMM_SETTING_AUTO_TUNE) == 0
    for setting, value in ((ftdx.GEMM_SETTING_GEMM_MODE, 3), (ftdx.GEMM_SETTING_TILE_CONFIG, 9),
                           (ftdx.GEMM_SETTING_POWER_MODE, -1), (7, 0)):
        try:
            ctx.set_gemm_setting(setting, value)
        except ftdx.FtdxError:
            continue
        raise AssertionError(f"setting {setting}={value} accepted")
    L = ftdx.lib()
    assert L.ftdx_gemm_prepared_fused(None, None, None, None, None, 8, 8, 8, 2, 0, 1, 1.0, 0.0,
                                      None) == ftdx.FTDX_ERR_INVALID_ARG
    assert L.ftdx_debug_fp6_real_gemm(None, 8, 8, 8, None) == ftdx.FTDX_ERR_INVALID_ARG
