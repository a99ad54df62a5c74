/*
 * ftdx.h -- C ABI of the MI355X (gfx950) X-engine HERK library, libftdx.so.
 *
 * Plain C: pointers, sizes and ints only; no HIP or torch types.  A stream is
 * passed as `void*` holding a hipStream_t (NULL = the default stream).  All
 * data pointers are device pointers owned by the caller.  Every entry point
 * returns FTDX_OK (0) or an FTDX_ERR_* code; ftdx_last_error() gives the
 * message of the last failure on the calling thread.
 *
 * Each entry point replaces one interface of the reference (cpviolator/ftd);
 * the C++ drop-ins in cutlass_gemm_api.h and ggp_cutlass_interface.h are thin
 * wrappers over these functions.
 *
 * Data conventions (kept unchanged from the reference):
 *   INT4 input  : one byte per complex sample, [batch][N][K] (K fastest); high
 *                 nibble = Re, low nibble = Im; sign-magnitude nibbles (bit 3 =
 *                 sign, bits 2:0 = |v|), values in [-7, 7], -0 allowed.
 *                 (cutlass_interface/cutlass_gemm_api.h:185-187, :727-728)
 *   HERK output : C = alpha * A A^H + beta * C, packed lower triangle, element
 *                 (i, j <= i) at i(i+1)/2 + j, interleaved [re, im], batch
 *                 stride N(N+1) scalars (FP32 or FP16); Im(C_ii) = 0 before
 *                 alpha/beta.  (cutlass_interface/shared/herk_kernel_common.hpp:353-424)
 *
 * Thread safety: a context is not thread-safe (one per thread), as the
 * reference class (cutlass_interface/cutlass_gemm_api.h:151-155).
 */
#ifndef FTDX_H
#define FTDX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FTDX_ABI_VERSION 1

/* status codes */
#define FTDX_OK 0
#define FTDX_ERR_INVALID_ARG 1
#define FTDX_ERR_UNSUPPORTED 2
#define FTDX_ERR_HIP 3
#define FTDX_ERR_NO_MEMORY 4
#define FTDX_ERR_IO 5

/* enum values equal the reference's C++ enums (cutlass_interface/cutlass_gemm_api.h:37-119) */
#define FTDX_INPUT_FP16 0  /* InputPrecision::FP16 : interleaved complex FP16 [batch][N][K][2] */
#define FTDX_INPUT_INT4 1  /* InputPrecision::INT4 : QC bytes [batch][N][K] */

#define FTDX_COMPUTE_FP8 0 /* ComputePrecision::FP8 (OCP E4M3) */
#define FTDX_COMPUTE_FP6 1 /* ComputePrecision::FP6 */
#define FTDX_COMPUTE_FP4 2 /* ComputePrecision::FP4 */

#define FTDX_OUTPUT_FP16 0 /* OutputPrecision::FP16 */
#define FTDX_OUTPUT_FP32 1 /* OutputPrecision::FP32 */

#define FTDX_FORMAT_PACKED_TRIANGLE 0 /* OutputFormat::PackedTriangle */

#define FTDX_HERK_AUTO 0           /* HerkMode::Auto */
#define FTDX_HERK_FORCE_DIRECT 1   /* HerkMode::ForceDirect : fused single-kernel path */
#define FTDX_HERK_FORCE_BASELINE 2 /* HerkMode::ForceBaseline : unfused promote + HERK path */

#define FTDX_DIRECT_TILE_AUTO 0    /* DirectTileConfig values (cutlass_gemm_api.h:113-119) */
#define FTDX_DIRECT_TILE_K32_B3 1
#define FTDX_DIRECT_TILE_K64_B2 2
#define FTDX_DIRECT_TILE_K64_B3 3
#define FTDX_DIRECT_TILE_K128_B2 4

typedef struct ftdx_context ftdx_context;

/* Message of the last failure on this thread ("" if none). */
const char* ftdx_last_error(void);

/* ABI version (FTDX_ABI_VERSION). */
int ftdx_abi_version(void);

/* Replaces CutlassComplexGemm::CutlassComplexGemm() / ~CutlassComplexGemm()
 * (cutlass_interface/cutlass_gemm_api.h:165-171; cutlass_gemm_api.cu:322-336).
 * The context binds to the device current at creation. */
int ftdx_create(ftdx_context** out);
int ftdx_destroy(ftdx_context* ctx);

/* Replaces CutlassComplexGemm::herk (cutlass_interface/cutlass_gemm_api.h:750-760;
 * cutlass_gemm_api.cu:1707-1878).  input/compute/output/format take the FTDX_*
 * values above; `tune` != 0 consults (and on a miss fills) the tile-config cache. */
int ftdx_herk(ftdx_context* ctx, const void* A, void* C, int N, int K, int batch_count,
              int input, int compute, int output, int format, float alpha, float beta,
              void* stream, int tune);

/* Replaces CutlassComplexGemm::init_herk / end_herk (cutlass_gemm_api.h:601-617). */
int ftdx_init_herk(ftdx_context* ctx, int N, int K, int batch_count, void* stream);
int ftdx_end_herk(ftdx_context* ctx);

/* Replace the CutlassComplexGemm setters (cutlass_gemm_api.h:620-720). */
int ftdx_set_herk_mode(ftdx_context* ctx, int mode);
int ftdx_set_direct_tile_config(ftdx_context* ctx, int config);
int ftdx_set_direct_herk_graph(ftdx_context* ctx, int enable);
int ftdx_set_tune_cache_path(ftdx_context* ctx, const char* path);
int ftdx_set_strategy_tune_verbosity(ftdx_context* ctx, int level);
int ftdx_set_kernel_tune_verbosity(ftdx_context* ctx, int level);
int ftdx_set_kernel_tune_cache_path(ftdx_context* ctx, const char* path);

/* Replaces CutlassComplexGemm::print_build_info (cutlass_gemm_api.h:773-775). */
int ftdx_print_build_info(void);

/* Stateless convenience over a per-thread default context: INT4 input, FP8
 * compute, packed-triangle output (output = FTDX_OUTPUT_FP32 or _FP16). */
int ftdx_herk_int4(const void* A, void* C, int N, int K, int batch_count, int output,
                   float alpha, float beta, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* FTDX_H */
