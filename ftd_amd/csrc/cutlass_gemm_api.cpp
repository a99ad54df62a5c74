// cutlass_gemm_api.cpp -- C++ drop-in CutlassComplexGemm over the ftdx C ABI.
// Host code only (built with the host compiler); every call goes through ftdx.h.
#include "cutlass_gemm_api.h"

#include <stdexcept>
#include <string>

namespace cutlass_gemm_api {

namespace {

// Reference behaviour: unsupported combinations and device errors throw
// (cutlass_interface/cutlass_gemm_api.cu:1716-1723, shared/config_common.hpp:35-43).
int check(int rc, const char* what) {
    if (rc == FTDX_OK) return 0;
    if (rc == FTDX_ERR_UNSUPPORTED || rc == FTDX_ERR_HIP || rc == FTDX_ERR_NO_MEMORY)
        throw std::runtime_error(std::string(what) + ": " + ftdx_last_error());
    return rc;
}

void check_void(int rc, const char* what) {
    if (rc != FTDX_OK) throw std::runtime_error(std::string(what) + ": " + ftdx_last_error());
}

}  // namespace

CutlassComplexGemm::CutlassComplexGemm() : ctx_(nullptr) {
    check_void(ftdx_create(&ctx_), "CutlassComplexGemm");
}

CutlassComplexGemm::~CutlassComplexGemm() { ftdx_destroy(ctx_); }

void CutlassComplexGemm::set_herk_mode(HerkMode mode) {
    check_void(ftdx_set_herk_mode(ctx_, static_cast<int>(mode)), "set_herk_mode");
}

void CutlassComplexGemm::set_direct_tile_config(DirectTileConfig config) {
    check_void(ftdx_set_direct_tile_config(ctx_, static_cast<int>(config)),
               "set_direct_tile_config");
}

void CutlassComplexGemm::set_direct_herk_graph(bool enable) {
    check_void(ftdx_set_direct_herk_graph(ctx_, enable ? 1 : 0), "set_direct_herk_graph");
}

void CutlassComplexGemm::set_tune_cache_path(const std::string& path) {
    check_void(ftdx_set_tune_cache_path(ctx_, path.c_str()), "set_tune_cache_path");
}

void CutlassComplexGemm::set_strategy_tune_verbosity(int level) {
    check_void(ftdx_set_strategy_tune_verbosity(ctx_, level), "set_strategy_tune_verbosity");
}

void CutlassComplexGemm::set_kernel_tune_verbosity(int level) {
    check_void(ftdx_set_kernel_tune_verbosity(ctx_, level), "set_kernel_tune_verbosity");
}

void CutlassComplexGemm::set_kernel_tune_cache_path(const std::string& path) {
    check_void(ftdx_set_kernel_tune_cache_path(ctx_, path.c_str()), "set_kernel_tune_cache_path");
}

void CutlassComplexGemm::init_herk(int N, int K, int batch_count, hipStream_t stream) {
    check_void(ftdx_init_herk(ctx_, N, K, batch_count, stream), "init_herk");
}

void CutlassComplexGemm::end_herk() { ftdx_end_herk(ctx_); }

int CutlassComplexGemm::herk(const void* A, void* C, int N, int K, int batch_count,
                             InputPrecision input, ComputePrecision compute,
                             OutputPrecision output, OutputFormat format, float alpha,
                             float beta, hipStream_t stream, bool tune) {
    return check(ftdx_herk(ctx_, A, C, N, K, batch_count, static_cast<int>(input),
                           static_cast<int>(compute), static_cast<int>(output),
                           static_cast<int>(format), alpha, beta, stream, tune ? 1 : 0),
                 "herk");
}

int CutlassComplexGemm::gemm(const void* A_re, const void* A_im, const void* B_re, const void* B_im,
                             void* C_re, void* C_im, int M, int N, int K, int batch_count,
                             ComputePrecision compute, OutputPrecision output, float alpha,
                             float beta, hipStream_t stream, bool tune, InputPrecision input) {
    return check(ftdx_gemm(ctx_, A_re, A_im, B_re, B_im, C_re, C_im, M, N, K, batch_count,
                           static_cast<int>(compute), static_cast<int>(output), alpha, beta, stream,
                           tune ? 1 : 0, static_cast<int>(input)),
                 "gemm");
}

void CutlassComplexGemm::prepare_b(const void* B_re, const void* B_im, int N, int K, int batch_count,
                                   ComputePrecision compute, hipStream_t stream) {
    const int rc = ftdx_prepare_b(ctx_, B_re, B_im, N, K, batch_count, static_cast<int>(compute), stream);
    if (rc != FTDX_OK) throw std::runtime_error(std::string("prepare_b: ") + ftdx_last_error());
}

int CutlassComplexGemm::gemm_prepared(const void* A_re, const void* A_im, void* C_re, void* C_im, int M,
                                      int N, int K, int batch_count, OutputPrecision output,
                                      float alpha, float beta, hipStream_t stream, bool tune,
                                      InputPrecision input) {
    return check(ftdx_gemm_prepared(ctx_, A_re, A_im, C_re, C_im, M, N, K, batch_count,
                                    static_cast<int>(output), alpha, beta, stream, tune ? 1 : 0,
                                    static_cast<int>(input)),
                 "gemm_prepared");
}

int CutlassComplexGemm::gemm_prepared_fused(const void* A_re, const void* A_im, void* C_re, void* C_im, int M,
                                            int N, int K, int batch_count, int fuse_factor, OutputPrecision output,
                                            float alpha, float beta, hipStream_t stream) {
    // every failure throws in the reference (cutlass_gemm_api.cu:982-995)
    check_void(ftdx_gemm_prepared_fused(ctx_, A_re, A_im, C_re, C_im, M, N, K, batch_count, fuse_factor,
                                        static_cast<int>(output), alpha, beta, stream),
               "gemm_prepared_fused");
    return 0;
}

int CutlassComplexGemm::debug_fp6_real_gemm(int M, int N, int K, hipStream_t stream) {
    const int rc = ftdx_debug_fp6_real_gemm(ctx_, M, N, K, stream);
    return rc == FTDX_OK ? 0 : (rc == FTDX_ERR_INVALID_ARG ? 1 : -1);
}

void CutlassComplexGemm::set_gemm_tune_verbosity(int level) {
    check_void(ftdx_set_gemm_setting(ctx_, FTDX_GEMM_SETTING_TUNE_VERBOSITY, level < 0 ? 0 : level),
               "set_gemm_tune_verbosity");
}

void CutlassComplexGemm::set_gemm_tune_cache_path(const std::string& path) {
    check_void(ftdx_set_gemm_tune_cache_path(ctx_, path.c_str()), "set_gemm_tune_cache_path");
}

void CutlassComplexGemm::set_tile_config(TileConfig config) {
    check_void(ftdx_set_gemm_setting(ctx_, FTDX_GEMM_SETTING_TILE_CONFIG, static_cast<int>(config)),
               "set_tile_config");
}

void CutlassComplexGemm::set_power_mode(PowerMode mode) {
    check_void(ftdx_set_gemm_setting(ctx_, FTDX_GEMM_SETTING_POWER_MODE, static_cast<int>(mode)), "set_power_mode");
}

void CutlassComplexGemm::set_gemm_mode(GemmMode mode) {
    check_void(ftdx_set_gemm_setting(ctx_, FTDX_GEMM_SETTING_GEMM_MODE, static_cast<int>(mode)), "set_gemm_mode");
}

GemmMode CutlassComplexGemm::get_gemm_mode() const {
    int v = 0;
    check_void(ftdx_get_gemm_setting(ctx_, FTDX_GEMM_SETTING_GEMM_MODE, &v), "get_gemm_mode");
    return static_cast<GemmMode>(v);
}

void CutlassComplexGemm::set_auto_tune(bool enable) {
    check_void(ftdx_set_gemm_setting(ctx_, FTDX_GEMM_SETTING_AUTO_TUNE, enable ? 1 : 0), "set_auto_tune");
}

int CutlassComplexGemm::gemm_prepared_direct(const void* A_re, const void* A_im, float* C_re, float* C_im,
                                             int M, int N, int K, int batch_count, float alpha,
                                             float beta, hipStream_t stream) {
    return check(ftdx_gemm_prepared_direct(ctx_, A_re, A_im, C_re, C_im, M, N, K, batch_count, alpha,
                                           beta, stream),
                 "gemm_prepared_direct");
}

int CutlassComplexGemm::gemm_prepared_power(const void* A_re, const void* A_im, float* C_power, int M,
                                            int N, int K, int batch_count, float alpha, float beta,
                                            hipStream_t stream) {
    return check(ftdx_gemm_prepared_power(ctx_, A_re, A_im, C_power, M, N, K, batch_count, alpha, beta,
                                          stream),
                 "gemm_prepared_power");
}

int CutlassComplexGemm::gemm_prepared_power_fp8(const void* A_fp8_interleaved, float* C_power, int M,
                                                int N, int K, int batch_count, float alpha, float beta,
                                                hipStream_t stream) {
    return check(ftdx_gemm_prepared_power_fp8(ctx_, A_fp8_interleaved, C_power, M, N, K, batch_count,
                                              alpha, beta, stream),
                 "gemm_prepared_power_fp8");
}

int CutlassComplexGemm::gemm_prepared_power_int4(const void* A_int4, float* C_power, int M, int N,
                                                 int K, int batch_count, float alpha, float beta,
                                                 hipStream_t stream) {
    return check(ftdx_gemm_prepared_power_int4(ctx_, A_int4, C_power, M, N, K, batch_count, alpha, beta,
                                               stream),
                 "gemm_prepared_power_int4");
}

int CutlassComplexGemm::gemm_prepared_direct_int4(const void* A_int4, float* C_re, float* C_im, int M,
                                                  int N, int K, int batch_count, float alpha, float beta,
                                                  hipStream_t stream) {
    return check(ftdx_gemm_prepared_direct_int4(ctx_, A_int4, C_re, C_im, M, N, K, batch_count, alpha,
                                                beta, stream),
                 "gemm_prepared_direct_int4");
}

void CutlassComplexGemm::print_build_info() { ftdx_print_build_info(); }

}  // namespace cutlass_gemm_api
