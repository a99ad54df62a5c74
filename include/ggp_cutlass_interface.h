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

}  // namespace ggp
