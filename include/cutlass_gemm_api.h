/**
 * @file cutlass_gemm_api.h
 * @brief Drop-in C++ interface of the MI355X X-engine HERK library.
 *
 * Same namespace, class, method names, enums and argument meaning as the
 * reference's HERK surface (cutlass_interface/cutlass_gemm_api.h:28-119,
 * :157-789), so a caller of CutlassComplexGemm::herk recompiles against this
 * header unchanged except for the stream type, which is hipStream_t.
 * The class is a thin wrapper over the C ABI in ftdx.h (libftdx.so); it is
 * implemented in libftdx_api.so and contains no device code.
 *
 * The HERK methods, the complex GEMM with prepared B (including power detection
 * and channel fusion) and the reference's tuning / dispatch setters are provided.
 *
 * Error behaviour (as the reference): herk() returns 0 on success and non-zero
 * on invalid arguments; unsupported precision combinations and HIP failures
 * throw std::runtime_error.  Not thread-safe: one instance per thread.
 */
#pragma once
#include <hip/hip_runtime_api.h>

#include <string>

#include "ftdx.h"

namespace cutlass_gemm_api {

/// Input data precision (cutlass_gemm_api.h:37-40).
enum class InputPrecision {
    FP16,  ///< Interleaved complex FP16 pairs
    INT4   ///< One QC byte per complex sample: high nibble Re, low nibble Im, sign-magnitude
};

/// Matrix-core compute precision (cutlass_gemm_api.h:43-47).
enum class ComputePrecision {
    FP8,  ///< OCP FP8 E4M3 -- exact for INT4 input
    FP6,  ///< FP6 (block-scaled MX path)
    FP4   ///< FP4 E2M1 (block-scaled MX path, lossy for INT4 input)
};

/// Output data precision (cutlass_gemm_api.h:50-53).
enum class OutputPrecision {
    FP16,  ///< FP16 output (overflows for 98*K > 65504 with full-scale INT4 input)
    FP32   ///< FP32 output
};

/// HERK output format (cutlass_gemm_api.h:56-58).
enum class OutputFormat {
    PackedTriangle  ///< N*(N+1)/2 complex elements per batch (lower triangle)
};

/// HERK dispatch mode (cutlass_gemm_api.h:61-65).
enum class HerkMode {
    Auto,          ///< Library choice (fused path)
    ForceDirect,   ///< Fused single-kernel path: unpack + MFMA HERK + packed store
    ForceBaseline  ///< Unfused path: separate promote stage, then HERK from FP8
};

/// Direct-path tile configuration (cutlass_gemm_api.h:113-119).  On gfx950 the
/// values select among the tile/pipeline configurations of the HERK kernel.
enum class DirectTileConfig { Auto, K32_B3, K64_B2, K64_B3, K128_B2 };

/// The reference's CUTLASS-baseline tile, power and GEMM dispatch choices
/// (cutlass_gemm_api.h:76-136).  This library has one GEMM tile kernel: the values are
/// kept and returned by get_gemm_mode(), results and kernels do not depend on them.
enum class TileConfig {
    Auto, HerkOptimal, GemmOptimal, WideN, Cluster1x2, Cluster2x2, SmallTile, Pingpong, PingpongSmall
};
enum class PowerMode { Auto, ForceDirect, Force4M };
enum class GemmMode { Auto, ForceDirect, Force4M };

class CutlassComplexGemm {
public:
    CutlassComplexGemm();
    ~CutlassComplexGemm();

    void set_herk_mode(HerkMode mode);
    void set_direct_tile_config(DirectTileConfig config);
    void set_direct_herk_graph(bool enable);
    void set_tune_cache_path(const std::string& path);
    void set_strategy_tune_verbosity(int level);
    void set_kernel_tune_verbosity(int level);
    void set_kernel_tune_cache_path(const std::string& path);

    void init_herk(int N, int K, int batch_count, hipStream_t stream = nullptr);
    void end_herk();

    /// C = alpha * A * A^H + beta * C (batched), packed lower triangle.
    int herk(const void* A, void* C, int N, int K, int batch_count,
             InputPrecision input = InputPrecision::FP16,
             ComputePrecision compute = ComputePrecision::FP8,
             OutputPrecision output = OutputPrecision::FP16,
             OutputFormat format = OutputFormat::PackedTriangle,
             float alpha = 1.0f, float beta = 0.0f,
             hipStream_t stream = nullptr,
             bool tune = true);

    // ---- complex GEMM C = alpha A B^T + beta C (cutlass_gemm_api.h:215-470) ----
    // FP16 planar operands are passed as `const __half*` by reference callers; any
    // pointer converts to the `const void*` parameters here.

    /// One call, B converted each time (gemm, :215-225).  FP8, or FP6 / FP4 on the
    /// MX block-scaled path.
    int gemm(const void* A_re, const void* A_im, const void* B_re, const void* B_im, void* C_re,
             void* C_im, int M, int N, int K, int batch_count,
             ComputePrecision compute = ComputePrecision::FP8,
             OutputPrecision output = OutputPrecision::FP16, float alpha = 1.0f, float beta = 0.0f,
             hipStream_t stream = nullptr, bool tune = true,
             InputPrecision input = InputPrecision::FP16);
    /// Converts B once (prepare_b, :251-255): FP8 rows, or MX codes for FP6 / FP4 (the
    /// direct and power variants then throw: they need FP8-prepared B).
    void prepare_b(const void* B_re, const void* B_im, int N, int K, int batch_count,
                   ComputePrecision compute, hipStream_t stream = nullptr);
    int gemm_prepared(const void* A_re, const void* A_im, void* C_re, void* C_im, int M, int N,
                      int K, int batch_count, OutputPrecision output = OutputPrecision::FP32,
                      float alpha = 1.0f, float beta = 0.0f, hipStream_t stream = nullptr,
                      bool tune = true, InputPrecision input = InputPrecision::FP16);
    int gemm_prepared_direct(const void* A_re, const void* A_im, float* C_re, float* C_im, int M,
                             int N, int K, int batch_count, float alpha = 1.0f, float beta = 0.0f,
                             hipStream_t stream = nullptr);
    int gemm_prepared_power(const void* A_re, const void* A_im, float* C_power, int M, int N, int K,
                            int batch_count, float alpha = 1.0f, float beta = 0.0f,
                            hipStream_t stream = nullptr);
    int gemm_prepared_power_fp8(const void* A_fp8_interleaved, float* C_power, int M, int N, int K,
                                int batch_count, float alpha = 1.0f, float beta = 0.0f,
                                hipStream_t stream = nullptr);
    int gemm_prepared_power_int4(const void* A_int4, float* C_power, int M, int N, int K,
                                 int batch_count, float alpha = 1.0f, float beta = 0.0f,
                                 hipStream_t stream = nullptr);
    int gemm_prepared_direct_int4(const void* A_int4, float* C_re, float* C_im, int M, int N, int K,
                                  int batch_count, float alpha = 1.0f, float beta = 0.0f,
                                  hipStream_t stream = nullptr);

    /// fuse_factor batch elements per product along M (gemm_prepared_fused, :440-470);
    /// FP32 output only, FP8-prepared B.
    int gemm_prepared_fused(const void* A_re, const void* A_im, void* C_re, void* C_im, int M, int N, int K,
                            int batch_count, int fuse_factor = 0, OutputPrecision output = OutputPrecision::FP32,
                            float alpha = 1.0f, float beta = 0.0f, hipStream_t stream = nullptr);
    /// Real FP6 GEMM of small integers against the host product (debug_fp6_real_gemm, :477-490).
    int debug_fp6_real_gemm(int M, int N, int K, hipStream_t stream = nullptr);

    void set_gemm_tune_verbosity(int level);
    void set_gemm_tune_cache_path(const std::string& path);
    void set_tile_config(TileConfig config);
    void set_power_mode(PowerMode mode);
    void set_gemm_mode(GemmMode mode);
    GemmMode get_gemm_mode() const;
    void set_auto_tune(bool enable);

    static void print_build_info();

    /// Underlying C-ABI context (for callers mixing the C and C++ interfaces).
    ftdx_context* context() const { return ctx_; }

    CutlassComplexGemm(const CutlassComplexGemm&) = delete;
    CutlassComplexGemm& operator=(const CutlassComplexGemm&) = delete;

private:
    ftdx_context* ctx_;
};

}  // namespace cutlass_gemm_api
