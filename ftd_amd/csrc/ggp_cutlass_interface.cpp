// ggp_cutlass_interface.cpp -- GGP caller glue over the C++ drop-in
// (reference lib/cutlass_interface.cu:38-41, :150-289).  Host code only.
#include "ggp_cutlass_interface.h"

#include "cutlass_gemm_api.h"

namespace ggp {

namespace {

using namespace cutlass_gemm_api;

// Process-wide instance, created on first use (lib/cutlass_interface.cu:38-41).
CutlassComplexGemm& api() {
    static CutlassComplexGemm instance;
    return instance;
}

bool g_herk_tune_enabled = false;

}  // namespace

void initCutlassApi(int kernel_tune_verbosity, const char* kernel_tune_cache_path,
                    const char* strategy_tune_cache_path, bool herk_tune,
                    int strategy_tune_verbosity) {
    CutlassComplexGemm& a = api();
    if (kernel_tune_verbosity > 0) a.set_kernel_tune_verbosity(kernel_tune_verbosity);
    a.set_strategy_tune_verbosity(strategy_tune_verbosity);
    if (kernel_tune_cache_path) a.set_kernel_tune_cache_path(kernel_tune_cache_path);
    if (strategy_tune_cache_path) a.set_tune_cache_path(strategy_tune_cache_path);
    g_herk_tune_enabled = herk_tune;
}

int herkBatchedCutlassQC(const void* qc_data, void* tri_output, int N, int K, int batch,
                         hipStream_t stream) {
    return api().herk(qc_data, tri_output, N, K, batch, InputPrecision::INT4,
                      ComputePrecision::FP8, OutputPrecision::FP32, OutputFormat::PackedTriangle,
                      1.0f, 0.0f, stream, g_herk_tune_enabled);
}

int herkBatchedCutlassQC_FP16(const void* qc_data, void* tri_output, int N, int K, int batch,
                              hipStream_t stream) {
    return api().herk(qc_data, tri_output, N, K, batch, InputPrecision::INT4,
                      ComputePrecision::FP8, OutputPrecision::FP16, OutputFormat::PackedTriangle,
                      1.0f, 0.0f, stream, g_herk_tune_enabled);
}

int herkBatchedCutlassFP16(const void* fp16_data, void* tri_output, int N, int K, int batch, hipStream_t stream) {
    return api().herk(fp16_data, tri_output, N, K, batch, InputPrecision::FP16, ComputePrecision::FP8,
                      OutputPrecision::FP16, OutputFormat::PackedTriangle, 1.0f, 0.0f, stream, g_herk_tune_enabled);
}

ftdx_context* cutlassContext() { return api().context(); }

bool cutlassHerkTune() { return g_herk_tune_enabled; }

}  // namespace ggp
