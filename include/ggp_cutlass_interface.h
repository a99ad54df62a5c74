/**
 * @file ggp_cutlass_interface.h
 * @brief Drop-in GGP caller glue for the X-engine HERK (reference
 *        lib/cutlass_interface.cu:150-289; forward-declared by its callers at
 *        lib/visibility_pipeline.cu:12-22 and lib/visibility_bf.cpp:18).
 *
 * One process-wide CutlassComplexGemm instance serves these functions, as in
 * the reference (lib/cutlass_interface.cu:38-41).  Pointers are device
 * pointers owned by the caller.  Return value 0 = success.
 */
#pragma once
#include <hip/hip_runtime_api.h>

#include <stdexcept>
#include <string>

#include "ftdx.h"

namespace ggp {

/// Configure the shared HERK instance (lib/cutlass_interface.cu:150-174).
void initCutlassApi(int kernel_tune_verbosity, const char* kernel_tune_cache_path,
                    const char* strategy_tune_cache_path, bool herk_tune,
                    int strategy_tune_verbosity = 1);

/// Batched HERK from QC (INT4) input, FP32 packed lower-triangle output
/// (lib/cutlass_interface.cu:178-231).  Any K >= 0 is accepted directly: the
/// gfx950 kernel masks the K tail itself, so no zero-padded copy is made.
int herkBatchedCutlassQC(const void* qc_data, void* tri_output, int N, int K, int batch,
                         hipStream_t stream = 0);

/// Same with FP16 packed output (lib/cutlass_interface.cu:243-289).
/// Overflows to inf once an entry exceeds 65504 (98*K > 65504 at full scale).
int herkBatchedCutlassQC_FP16(const void* qc_data, void* tri_output, int N, int K, int batch,
                              hipStream_t stream = 0);

/// Batched HERK from interleaved FP16 complex input, FP16 packed output
/// (lib/cutlass_interface.cu:291-306).
int herkBatchedCutlassFP16(const void* fp16_data, void* tri_output, int N, int K, int batch,
                           hipStream_t stream = 0);

/// The shared instance's C-ABI context (used by the quda glue below).
ftdx_context* cutlassContext();
/// Whether initCutlassApi enabled HERK tuning.
bool cutlassHerkTune();

}  // namespace ggp

namespace quda {

/// quda::stridedBatchGEMMCutlass (lib/cutlass_interface.cu:316-384, declared at
/// include/blas_lapack.h:59): the HERK case (A == B, trans_a = conjugate transpose
/// (2), trans_b = none (0), m == n, single complex data (2)) narrows A to FP16, runs the
/// HERK (FP8 compute, FP16 packed output) and widens the packed lower triangle into C;
/// returns 8 m n k batch_count, or -1 on failure.  Every other case fails, as in the
/// reference (which aborts through errorQuda): here std::runtime_error.  `Param` is
/// the caller's QudaBLASParam (fields m, n, k, batch_count, trans_a, trans_b,
/// data_type with the reference's enum values); alpha, beta and the leading dimensions
/// are not read, as in the reference.
template <typename Param>
long long stridedBatchGEMMCutlass(void* A_data, void* B_data, void* C_data, const Param& param, int location) {
    (void)location;
    const bool herk = A_data == B_data && (int)param.trans_a == 2 && (int)param.trans_b == 0 && param.m == param.n;
    if (herk && (int)param.data_type == 2) {
        const int rc = ftdx_blas_herk_fp32(ggp::cutlassContext(), A_data, C_data, param.n, param.k,
                                           param.batch_count, ggp::cutlassHerkTune() ? 1 : 0);
        if (rc != FTDX_OK) return -1;
        return (long long)param.batch_count * 8LL * param.m * param.n * param.k;
    }
    if ((int)param.data_type == 2)
        throw std::runtime_error("CUTLASS GEMM API: general complex GEMM not yet implemented via this "
                                 "interface. Use the HERK path (trans_a=CONJ_TRANS, A==B) or cuBLAS.");
    throw std::runtime_error("CUTLASS GEMM API: unsupported data type " + std::to_string((int)param.data_type));
}

/// quda::gemmBatchedCutlass (lib/cutlass_interface.cu:371-459): FP32 interleaved complex
/// A [batch][M][K] and B [batch][N][K] -> C [batch][M][N] through the FP8 complex GEMM,
/// B's imaginary part negated when conjugate_b is false (the reference's flag sense).
/// Returns 8 M N K batch_count, or -1 on failure.
inline long long gemmBatchedCutlass(void* A_data, void* B_data, void* C_data, int M, int N, int K, int batch_count,
                                    bool conjugate_b) {
    const int rc = ftdx_blas_gemm_fp32(ggp::cutlassContext(), A_data, B_data, C_data, M, N, K, batch_count,
                                       conjugate_b ? 1 : 0, 0);
    if (rc != FTDX_OK) return -1;
    return (long long)batch_count * 8LL * M * N * K;
}

}  // namespace quda
