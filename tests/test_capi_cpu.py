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
    for h in ("cutlass_gemm_api.h", "ggp_cutlass_interface.h"):
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
