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

/* Element rounding of the MX (FP6 / FP4) quantizer; the E8M0 block scale is the same
 * for both.  REFERENCE reproduces the reference's converters code for code
 * (cutlass_interface/gemm_complex_sm100/fp6_kernels.hpp:27-58, fp4_kernels.hpp:7-30:
 * clamp, |x| < 0.25 to signed zero, E3M2 mantissa rounded without carry, E2M1
 * mantissa truncated) and is the default.  NEAREST_EVEN (no reference counterpart)
 * is the OCP MX v1.0 rule: nearest even with subnormals, saturating. */
#define FTDX_MX_ROUND_REFERENCE 0
#define FTDX_MX_ROUND_NEAREST_EVEN 1

#define FTDX_DIRECT_TILE_AUTO 0    /* DirectTileConfig values (cutlass_gemm_api.h:113-119) */
#define FTDX_DIRECT_TILE_K32_B3 1
#define FTDX_DIRECT_TILE_K64_B2 2
#define FTDX_DIRECT_TILE_K64_B3 3
#define FTDX_DIRECT_TILE_K128_B2 4

/* Tile configurations of the gfx950 HERK kernel (rows x cols complex per
 * workgroup tile, samples per LDS stage); FTDX_TILE_AUTO lets the library choose. */
#define FTDX_TILE_AUTO (-1)
#define FTDX_TILE_128x128_K64 0  /* 4 waves, 2 workgroups per CU */
#define FTDX_TILE_128x128_K128 1 /* 4 waves, 1 workgroup per CU */
#define FTDX_TILE_64x64_K64 2    /* 1 wave, 4 workgroups per CU */
#define FTDX_TILE_256x128_K64 3  /* 8 waves, 1 workgroup per CU */
#define FTDX_TILE_128x128_K64_P2 4 /* as 0, global loads two stages ahead */
#define FTDX_TILE_256x128_K64_P2 5 /* as 3, global loads two stages ahead */
/* Flag on the ids above, ftdx_herk with INT4 input only: the same tile fed by operand
 * rows pre-cast to OCP E3M2 (6-bit) by an extra pass per chunk of channels.  Every QC
 * value is exact in E3M2, so results are identical to the FP8 path; the MFMAs run at
 * twice the FP8 rate.
 * Chosen explicitly (ftdx_set_launch_config) or from a tune-cache entry; the autotuner
 * times the FP8 tiles only.  Other entry points use the tile without the flag.  No
 * counterpart in the reference (its INT4 path computes in FP8). */
#define FTDX_TILE_BF6 16

typedef struct ftdx_context ftdx_context;
typedef struct ftdx_grouped_plan ftdx_grouped_plan;

/* One problem of a grouped HERK: C = alpha * A A^H + beta * C with A [N][K] QC
 * bytes and C the packed lower triangle (N(N+1) scalars). */
typedef struct ftdx_herk_problem {
    const void* A;
    void* C;
    int N;
    int K;
    float alpha;
    float beta;
} ftdx_herk_problem;

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
 * values above; `tune` != 0 consults (and on a miss fills) the tile-config cache.
 * INT4 rows of K % 16 != 0, or at an address that is not 16-byte aligned, are copied
 * zero-padded (K rounded up to 16, or to the tile's stage) into the context's scratch (at
 * most 320 MB per chunk of channels) before the fused tile runs on them; the FP16-input, ForceBaseline and E3M2
 * paths stage converted rows there.  The scratch grows on first need unless
 * ftdx_init_herk sized it. */
int ftdx_herk(ftdx_context* ctx, const void* A, void* C, int N, int K, int batch_count,
              int input, int compute, int output, int format, float alpha, float beta,
              void* stream, int tune);

/* Replaces CutlassComplexGemm::init_herk / end_herk (cutlass_gemm_api.h:601-617).
 * init_herk allocates the scratch that calls of this shape need (staged rows for the
 * pre-cast paths and graph mode, padded rows when K % 16 != 0), so that they allocate
 * nothing later; end_herk releases it. */
int ftdx_init_herk(ftdx_context* ctx, int N, int K, int batch_count, void* stream);
int ftdx_end_herk(ftdx_context* ctx);

/* Replace the CutlassComplexGemm setters (cutlass_gemm_api.h:620-720). */
int ftdx_set_herk_mode(ftdx_context* ctx, int mode);
int ftdx_set_direct_tile_config(ftdx_context* ctx, int config);
int ftdx_set_direct_herk_graph(ftdx_context* ctx, int enable);
/* Graph-replay diagnostics (no reference counterpart): graphs held in the context's
 * four-entry cache and graphs captured since it was created.  A call replays a cached
 * graph only if its shape, precisions, scalars, pointers and launch configuration
 * (tile configuration and workgroup order) all match the captured call. */
int ftdx_herk_graph_stats(const ftdx_context* ctx, int* cached, long long* captures);
/* MX element rounding for this context's FP6 / FP4 paths (HERK, prepared and one-shot
 * GEMM): FTDX_MX_ROUND_REFERENCE (default) or FTDX_MX_ROUND_NEAREST_EVEN.  No
 * reference counterpart (the reference has only its own converters).  The getter
 * returns -1 for a NULL context. */
int ftdx_set_mx_rounding(ftdx_context* ctx, int rounding);
int ftdx_get_mx_rounding(const ftdx_context* ctx);
int ftdx_set_tune_cache_path(ftdx_context* ctx, const char* path);
int ftdx_set_strategy_tune_verbosity(ftdx_context* ctx, int level);
int ftdx_set_kernel_tune_verbosity(ftdx_context* ctx, int level);
int ftdx_set_kernel_tune_cache_path(ftdx_context* ctx, const char* path);

/* Explicit launch configuration for ftdx_herk (tile_config = FTDX_TILE_*,
 * xcd_remap = 0/1, either -1 = choose automatically).  Overrides
 * set_direct_tile_config and the tuning cache while set. */
int ftdx_set_launch_config(ftdx_context* ctx, int tile_config, int xcd_remap);

/* Tuning cache I/O (replaces the reference's strategy/kernel cache files,
 * cutlass_interface/strategy_cache.hpp:13-27, tune_cache.hpp:22-37).  Entries
 * carry a build fingerprint; lines of another build or malformed lines are
 * skipped.  `loaded` / `skipped` may be NULL. */
int ftdx_tune_load(ftdx_context* ctx, const char* path, int* loaded, int* skipped);
int ftdx_tune_save(ftdx_context* ctx, const char* path);
/* Launch configuration ftdx_herk would use now for this shape without tuning
 * (explicit setting, else cache entry, else the default). */
int ftdx_herk_query_config(ftdx_context* ctx, int N, int K, int batch_count, int output,
                           int* tile_config, int* xcd_remap);

/* Grouped HERK: many problems of different N and K in one persistent launch.
 * No reference counterpart (the reference loops over uniform batched calls,
 * lib/visibility_pipeline.cu:484-540); it serves sub-arrays and channel groups
 * of uneven size.  A plan uploads the problem list and a static tile schedule;
 * with use_graph != 0 the launch is captured once in a hipGraph and replayed.
 * tile_config = FTDX_TILE_AUTO times the configurations on the problems' own
 * buffers when every beta is 0 (C is overwritten), else uses the default.
 * All outputs share one precision.  The plan keeps the pointers: buffers must
 * outlive it.  Problems whose rows 16-byte loads cannot take (K % 16 != 0, or A
 * not 16-byte aligned) run on zero-padded copies (K rounded up to 16) in device
 * memory the plan owns until it is destroyed; every run refreshes them from A
 * first, so A may change between runs. */
int ftdx_herk_grouped_plan_create(ftdx_context* ctx, const ftdx_herk_problem* problems,
                                  int count, int output, int tile_config, int use_graph,
                                  ftdx_grouped_plan** plan);
int ftdx_herk_grouped_plan_run(ftdx_grouped_plan* plan, void* stream);
int ftdx_herk_grouped_plan_info(const ftdx_grouped_plan* plan, int* num_workgroups,
                                int64_t* num_items, int* tile_config, double* imbalance);
int ftdx_herk_grouped_plan_destroy(ftdx_grouped_plan* plan);

/* The same problem list run the plain way: one ftdx_herk launch per problem, in
 * order, on one stream (the per-sub-array loop of the reference's callers,
 * lib/visibility_pipeline.cu:484-540).  Output bytes equal the grouped plan's. */
int ftdx_herk_each(ftdx_context* ctx, const ftdx_herk_problem* problems, int count, int output,
                   void* stream);

/* MX block-scaled HERK (the reference's FP6/FP4 compute path,
 * cutlass_interface/gemm_complex_sm100/herk_impl.hpp:1067-1340, reached from
 * CutlassComplexGemm::herk with ComputePrecision::FP6/FP4).  A is QC bytes
 * (FTDX_INPUT_INT4) or interleaved complex FP16 [batch][N][K][2]
 * (FTDX_INPUT_FP16).  Re and Im are quantized separately per (row, 32-sample
 * block) with an E8M0 scale 2^ceil(log2(amax / max)) (max 28 for FP6 E3M2,
 * 6 for FP4 E2M1; mxfp_kernels.hpp:147-166) and the context's element rounding
 * (ftdx_set_mx_rounding; the reference's converters by default), then
 * multiplied on scaled FP6/FP4 MFMA with FP32 accumulation.
 * Unlike the reference (FP16 output only) both FP16 and FP32 output work here;
 * ftdx_herk keeps the reference's rule and rejects FP32 output with FP6/FP4.
 * The context owns the quantized scratch (grown on demand, freed by
 * ftdx_end_herk / ftdx_destroy). */
int ftdx_herk_mx(ftdx_context* ctx, const void* A, void* C, int N, int K, int batch_count,
                 int input, int compute, int output, float alpha, float beta, void* stream);

/* The MX quantization stage alone (fused_mxfp_preprocess_kernel_sm100,
 * mxfp_kernels.hpp:265-360), with the reference's element converters.  X as for
 * ftdx_herk_mx with `rows` rows of K samples.  codes: rows * blocks * 2 *
 * code_bytes bytes laid out [row][block][Re|Im][code_bytes], scales: rows *
 * blocks * 2 E8M0 bytes [row][block][Re|Im], with blocks and code_bytes from
 * ftdx_mx_layout. */
int ftdx_mx_quantize(const void* X, int input, int compute, void* codes, void* scales,
                     int64_t rows, int K, void* stream);
/* The same with an explicit element rounding (FTDX_MX_ROUND_*). */
int ftdx_mx_quantize_rounding(const void* X, int input, int compute, int rounding, void* codes,
                              void* scales, int64_t rows, int K, void* stream);
int ftdx_mx_layout(int compute, int K, int* blocks, int* code_bytes);

/* Differentiable gain calibration on packed visibilities (no reference
 * counterpart: the reference has no backward pass).  Per channel b, with M the
 * model visibilities (a HERK output), V the observed ones, both packed lower
 * triangles [batch][N(N+1)/2][2] FP32, and gains g [batch][N][2] FP32:
 *   R_ij = V_ij - g_i conj(g_j) M_ij,   loss_b = sum_{i>=j} |R_ij|^2   (FP64, device)
 *   grad_g [batch][N][2] = dL/dRe g + i dL/dIm g                        (FP32, device)
 *   Gamma (optional, packed like M) = dL/dRe M + i dL/dIm M = -2 R_ij conj(g_i) g_j
 * Sums are reduced in a fixed order (bitwise reproducible). */
int ftdx_gain_loss_grad(ftdx_context* ctx, const void* V, const void* M, const void* g, int N,
                        int batch_count, void* loss, void* grad_g, void* Gamma, void* stream);

/* Allocation-free ftdx_gain_loss_grad.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and results as
 * ftdx_gain_loss_grad, without a context.  The gradient and loss partials live in a
 * fixed device array of FTDX_GAIN_NOALLOC_SCRATCH_BYTES on each device (held until
 * the process ends), so calls are limited to
 *   2 * batch * nb * N * 8 + batch * nb (nb + 1) / 2 * 8 <= 16 MiB,  nb = ceil(N / 32)
 * (N = 1024 with batch 16 needs 8.1 MB); larger calls fail with FTDX_ERR_INVALID_ARG.
 * Every call on a device shares that array, so calls that overlap on one device
 * (two streams, two threads) race.  Nothing in it needs clearing between calls. */
#define FTDX_GAIN_NOALLOC_SCRATCH_BYTES ((size_t)16 << 20)
int ftdx_gain_loss_grad_noalloc(const void* V, const void* M, const void* g, int N, int batch_count,
                                void* loss, void* grad_g, void* Gamma, void* stream);

/* HERK adjoint: for M = X X^H (packed) and an upstream gradient Gamma (packed,
 * the convention above), dX = (Gamma + Gamma^H) X, FP32 [batch][N][K][2].
 * X is QC bytes [batch][N][K] (input = FTDX_INPUT_INT4).  Complex GEMM on bf16
 * MFMA over the exact three-term bf16 split of Gamma + Gamma^H, FP32 accumulation
 * (FP32-GEMM accuracy).  The context owns the scratch: the split planes and the
 * transposed voltages, batch * Np * (12 N + 4 K) bytes with Np = N rounded up to 32. */
int ftdx_herk_adjoint(ftdx_context* ctx, const void* Gamma, const void* X, void* dX, int N, int K,
                      int batch_count, int input, void* stream);

/* ---- stand-alone stage kernels (the fused HERK does these internally) ---- */
#define FTDX_PREC_FP32 1 /* element type selectors for the GGP-convention stages */
#define FTDX_PREC_FP64 2

/* QC bytes -> [fp8 Re, fp8 Im] byte pairs, n complex samples (replaces
 * cast_int4_to_fp8_interleaved, cutlass_interface/cutlass_gemm_api.cu:1553-1623;
 * same bytes as its LUT, -0 nibble -> 0x80). */
int ftdx_promote_qc_to_fp8(const void* qc, void* fp8_pairs, int64_t n, void* stream);
/* QC bytes -> FP16 Re plane and Im plane (replaces cast_int4_to_fp16_planar,
 * cutlass_gemm_api.cu:1627-1698; -0 nibble -> -0.0). */
int ftdx_promote_qc_to_fp16_planar(const void* qc, void* re, void* im, int64_t n, void* stream);
/* GGP convention: (4,4)-bit two's complement, low nibble = Re, high = Im ->
 * interleaved FP32 or FP64 pairs (replaces promoteData, lib/promote_data.cu:54-97). */
int ftdx_promote_data(const void* in, void* out, int64_t n, int prec, void* stream);
/* Full N x N interleaved FP32 matrices [batch][N][N][2] -> packed lower
 * triangle, out = full (+ beta * old when beta != 0 and old != NULL)
 * (replaces pack_scratch_to_triangle_fp32, shared/pack_kernels_common.hpp:602-659). */
int ftdx_pack_triangle(const void* full, void* packed, const void* old, int N, int batch_count,
                       float beta, void* stream);
/* GGP convention: full N x N complex [batch][N][N][2] -> packed UPPER triangle,
 * row-major, (i, j >= i) at i*N - i(i-1)/2 + (j - i) (replaces
 * triangulateFromHerm, lib/tri_from_herm.cu:10-54). */
int ftdx_tri_from_herm(const void* full, void* tri, int N, int batch_count, int prec,
                       void* stream);
/* FP16 Re/Im planes -> interleaved FP32/FP64 pairs (replaces promInterTri,
 * lib/interleave_tri.cu:9-54). */
int ftdx_interleave_tri(const void* re, const void* im, void* out, int64_t n, int prec,
                        void* stream);

/* QC bytes -> interleaved FP32 [re, im] (replaces promoteQcSmToFp32,
 * include/kernels/visibility_bf.cuh:38-56; -0 nibble -> +0). */
int ftdx_promote_qc_to_fp32(const void* qc, void* out, int64_t n, void* stream);
/* QC sign-magnitude (high nibble Re) -> FTD two's-complement nibbles (low nibble
 * Re, high nibble Im), n bytes; qc == out allowed (replaces convertQcSmToFtd,
 * lib/visibility_bf_kernels.cu:422-440). */
int ftdx_convert_qc_to_ftd(const void* qc, void* out, int64_t n, void* stream);

/* ---- FTD raw-order correlation (the TCC engine's contract) ----
 * Input: one byte per complex sample in FTD order [time][ant][chan][pol], time =
 * k * n_time_inner + ti, K = n_time / n_time_inner samples per product.  Output:
 * packed lower triangle FP32 [batch][n_ant(n_ant+1)/2][2], batch = (chan * n_pol +
 * pol) * n_time_inner + ti, element (r2, r1 <= r2) = sum_k x_r2 conj(x_r1), i.e. the
 * HERK convention above.  `twos` = 1: FTD two's-complement nibbles (low nibble Re,
 * values in [-8, 7]) as correlateTCC takes them (lib/tcc_interface.cu:251-314);
 * `twos` = 0: QC sign-magnitude bytes as XEngine's TCC branch receives them before
 * convertQcSmToFtd (lib/xengine.cpp:130-150; CPU form cpu_herk_ftd,
 * tests/visibility_pipeline_test.cpp:54-96).  The input is not modified.  Any
 * n_pol >= 1 (libtcc needs 2); no time-block multiple is required. */
int ftdx_correlate_ftd(ftdx_context* ctx, const void* raw, void* tri_fp32, int n_ant, int n_chan,
                       int n_time, int n_time_inner, int n_pol, int twos, void* stream);

/* ---- GGP XEngine (lib/xengine.cpp:22-224, include/ggp.h:101-123) ----
 * Enum values equal the reference's (include/enum_ggp.h). */
#define FTDX_LOCATION_HOST 1     /* QUDA_CPU_FIELD_LOCATION */
#define FTDX_LOCATION_DEVICE 2   /* QUDA_CUDA_FIELD_LOCATION */
#define FTDX_PACKET_DSA110 4608  /* QUDA_PACKET_FORMAT_DSA110 */
#define FTDX_PACKET_CASM 46080   /* QUDA_PACKET_FORMAT_CASM */
#define FTDX_PACKET_DSA2K 8192   /* QUDA_PACKET_FORMAT_DSA2K */
#define FTDX_QPREC_QUARTER 1     /* QudaPrecision */
#define FTDX_QPREC_HALF 2
#define FTDX_QPREC_SINGLE 4
#define FTDX_QPREC_DOUBLE 8
#define FTDX_BLAS_ENGINE_CUBLAS 0  /* QudaBLASEngine */
#define FTDX_BLAS_ENGINE_CUTLASS 1
#define FTDX_BLAS_ENGINE_TCC 2
#define FTDX_XENGINE_MAT_HERM 0  /* QudaXEngineMatFormat */
#define FTDX_XENGINE_MAT_TRI 1
#define FTDX_INVALID_ENUM (-0x7fffffff - 1) /* QUDA_INVALID_ENUM */

/* Mirrors XEngineParam (include/ggp.h:101-123) minus data_type / data_order /
 * verbosity, which the reference's XEngine does not read. */
typedef struct ftdx_xengine_param {
    size_t struct_size;    /* sizeof(ftdx_xengine_param) or (size_t)FTDX_INVALID_ENUM */
    int in_location;       /* FTDX_LOCATION_* of the input payloads */
    int out_location;      /* FTDX_LOCATION_* of the outputs */
    int packet_format;     /* FTDX_PACKET_DSA2K */
    size_t n_payload;
    size_t n_antennae_per_payload;
    size_t n_channels_per_payload;
    size_t n_time_per_payload;
    size_t n_time_inner;
    size_t n_polarizations;
    int compute_prec;      /* FTDX_QPREC_SINGLE or _DOUBLE */
    int output_prec;       /* must equal compute_prec */
    int engine;            /* FTDX_BLAS_ENGINE_*: CUBLAS / CUTLASS -> GGP layout; TCC -> FTD order */
    int format;            /* FTDX_XENGINE_MAT_HERM or _TRI */
} ftdx_xengine_param;

typedef struct ftdx_xengine ftdx_xengine;

/* newXEngineParam (lib/check_params.h:104-145): DSA2K, compute HALF (which
 * ftdx_xengine_create rejects, as the reference constructor does), output
 * SINGLE, engine CUBLAS, format HERM, sizes invalid; locations host. */
int ftdx_xengine_param_init(ftdx_xengine_param* param);

/* ggp::XEngine::XEngine (lib/xengine.cpp:22-108): validates, sizes and allocates.
 * Layouts (engine CUBLAS or CUTLASS, packet DSA2K):
 *   input per payload: n_time * n_ant * n_chan * n_pol bytes, GGP two's-complement
 *     nibbles (low = Re), batches [chan][pol][time_inner] of [K][n_ant] (ant fastest),
 *     K = n_time / n_time_inner;
 *   output per payload: batch = n_chan * n_pol * n_time_inner matrices of M^H M
 *     (M = K x n_ant), TRI: upper triangle row-major (row q, col p >= q at
 *     q*N + p - q(q+1)/2), HERM: full N x N row-major; FP32 or FP64 [re, im].
 * Engine TCC (packet DSA2K, format TRI, output SINGLE): the ftdx_correlate_ftd
 * layout with QC sign-magnitude input.  Packet DSA110 is FTDX_ERR_UNSUPPORTED. */
int ftdx_xengine_create(const ftdx_xengine_param* param, ftdx_xengine** out);
/* XEngine::compute (lib/xengine.cpp:110-178): for each payload, copy in (host
 * input), convert, HERK, copy out (host output); returns after the stream is
 * idle, as the reference ends with a device synchronize. */
int ftdx_xengine_compute(ftdx_xengine* xe, void* output, const void* input, void* stream);
/* Bytes of one input / output payload and the HERK shape of an engine. */
int ftdx_xengine_query(const ftdx_xengine* xe, size_t* in_payload_bytes, size_t* out_payload_bytes,
                       int* N, int* K, int* batch_count);
/* Launch configuration for the HERK (FTDX_TILE_*, xcd_remap 0/1; -1 = library choice). */
int ftdx_xengine_set_launch_config(ftdx_xengine* xe, int tile_config, int xcd_remap);
int ftdx_xengine_destroy(ftdx_xengine* xe);
/* XEngineCHPC (lib/interface_ggp.cpp:536-555): create, compute on the default
 * stream, destroy. */
int ftdx_xengine_chpc(void* output, const void* input, const ftdx_xengine_param* param);

/* ---- consumers of the visibility triangle ----
 * Polarization sum: input triangles in the HERK batch order [n_ch][n_pol][n_ti]
 * (packed lower, N(N+1) scalars each), output [n_ch][n_ti] = sum over pol, FP32 or
 * FP16 (output = FTDX_OUTPUT_*).  Replaces pol_reduce_triangle_kernel[_fp16]
 * (lib/bench_pipeline.cu:254-302, n_pol = 2 there). */
int ftdx_pol_reduce_triangle(const void* tri_in, void* tri_out, int N, int n_ch, int n_pol, int n_time_inner,
                             int output, void* stream);
/* HERK and polarization sum in one pass: A = QC bytes [n_ch][n_pol][n_ti][N][K], C =
 * [n_ch][n_ti] packed lower triangles of sum_pol X X^H (= the HERK over the n_pol*K
 * samples of both polarizations), with the HERK's alpha / beta / Im(diag) = 0 rules.
 * Equal to ftdx_herk followed by ftdx_pol_reduce_triangle for integer voltages (exact
 * sums), without writing the per-polarization triangles.  Rows of K % 16 != 0 or at an
 * address that is not 16-byte aligned are first copied zero-padded into the context's
 * scratch (grown on first need, as ftdx_herk does). */
int ftdx_herk_pol_reduce(ftdx_context* ctx, const void* A, void* C, int N, int K, int n_ch, int n_pol,
                         int n_time_inner, int output, float alpha, float beta, void* stream);
/* Pillbox UV gridding (PillboxGridScatter, include/kernels/visibility_bf.cuh:259-370;
 * pillbox_grid_scatter_fp32_fp32grid_kernel, include/kernels/imaging_kernels.cuh:219-270):
 * for frequency f < Nf and baseline k (packed lower index, vis_tri [Nf][n_baselines][2]
 * FP32 or FP16 per vis_output), UV in wavelengths = baseline_uv_m[k] * freq_hz[freq_offset
 * + f] / c; cell (roundf(u / cell + Ng/2), roundf(v / cell + Ng/2)); the visibility is
 * added to uv_grid [Nf][Ng][Ng][2] (FP32) there, and its conjugate at (-u, -v) for
 * cross-correlations.  Accumulates: the caller zeroes the grid. */
int ftdx_pillbox_grid(const void* vis_tri, int vis_output, const float* baseline_uv_m,
                      const double* freq_hz, float* uv_grid, int n_baselines, int Ng, int Nf,
                      int freq_offset, float cell_size_rad, void* stream);
/* HERK, polarization sum and pillbox gridding in one pass: the triangle never reaches
 * memory.  Replaces the reference's pol-reduce + grid sequence on the packed triangle
 * (pol_reduce_triangle_kernel, lib/bench_pipeline.cu:254-302; PillboxGridScatter,
 * include/kernels/visibility_bf.cuh:269-370, called at lib/imaging_pipeline.cu:338-356).
 * A = QC bytes [n_chan][n_pol][N][K]; channel f's visibilities alpha * sum_pol X X^H
 * (Im = 0 on the diagonal) go into uv_grid [n_chan][Ng][Ng][2] FP32 by the rules of
 * ftdx_pillbox_grid at frequency freq_hz[freq_offset + f].  Accumulates: the caller
 * zeroes the grid.  Equal to ftdx_herk_pol_reduce (n_time_inner = 1) followed by
 * ftdx_pillbox_grid (Nf = n_chan) whenever the FP32 cell sums are exact (integer
 * visibilities, |partial sums| < 2^24); otherwise equal up to the order of the adds.
 * Rows of K % 16 != 0 or at an unaligned address are padded first, as in
 * ftdx_herk_pol_reduce. */
int ftdx_herk_pol_grid(ftdx_context* ctx, const void* A, int N, int K, int n_chan, int n_pol,
                       const float* baseline_uv_m, const double* freq_hz, float* uv_grid, int Ng,
                       int freq_offset, float cell_size_rad, float alpha, void* stream);

/* ---- complex GEMM with prepared B (the voltage beamformer's product) ----
 * C = alpha * A B^T + beta * C per batch, A [batch][M][K], B [batch][N][K], C
 * [batch][M][N] (cutlass_gemm_api.h:215-470; CPU form cpu_gemm, test_gemm_api.cu:50-67).
 * FP8 E4M3 compute (FP16 inputs clamped to +-448, nearest-even), FP32 accumulate.
 * Planar layouts: FP16 A_re / A_im, FP16 B_re / B_im, C_re / C_im FP32 (or FP16 via
 * ftdx_gemm_prepared's output).  INT4 A: QC bytes [batch][M][K] in A_re (A_im unused).
 * Power: C_power = alpha * (Re^2 + Im^2) + beta * C_power, FP32 [batch][M][N].
 * ComputePrecision FP6 / FP4 return FTDX_ERR_UNSUPPORTED. */
/* CutlassComplexGemm::prepare_b (cutlass_gemm_api.h:251-255): converts B once. */
int ftdx_prepare_b(ftdx_context* ctx, const void* B_re, const void* B_im, int N, int K, int batch_count,
                   int compute, void* stream);
/* gemm_prepared (cutlass_gemm_api.h:290-299); input FTDX_INPUT_FP16 / _INT4,
 * output FTDX_OUTPUT_FP32 / _FP16. */
int ftdx_gemm_prepared(ftdx_context* ctx, const void* A_re, const void* A_im, void* C_re, void* C_im,
                       int M, int N, int K, int batch_count, int output, float alpha, float beta,
                       void* stream, int tune, int input);
/* gemm_prepared_direct (cutlass_gemm_api.h:321-326): FP16 A, FP32 planar C. */
int ftdx_gemm_prepared_direct(ftdx_context* ctx, const void* A_re, const void* A_im, void* C_re,
                              void* C_im, int M, int N, int K, int batch_count, float alpha, float beta,
                              void* stream);
/* gemm_prepared_direct_int4 (cutlass_gemm_api.h:419-424). */
int ftdx_gemm_prepared_direct_int4(ftdx_context* ctx, const void* A_int4, void* C_re, void* C_im, int M,
                                   int N, int K, int batch_count, float alpha, float beta, void* stream);
/* gemm_prepared_power (cutlass_gemm_api.h:348-353). */
int ftdx_gemm_prepared_power(ftdx_context* ctx, const void* A_re, const void* A_im, void* C_power, int M,
                             int N, int K, int batch_count, float alpha, float beta, void* stream);
/* gemm_prepared_power_fp8 (cutlass_gemm_api.h:376-381): A as FP8 [re, im] byte pairs. */
int ftdx_gemm_prepared_power_fp8(ftdx_context* ctx, const void* A_fp8_interleaved, void* C_power, int M,
                                 int N, int K, int batch_count, float alpha, float beta, void* stream);
/* gemm_prepared_power_int4 (cutlass_gemm_api.h:398-401). */
int ftdx_gemm_prepared_power_int4(ftdx_context* ctx, const void* A_int4, void* C_power, int M, int N,
                                  int K, int batch_count, float alpha, float beta, void* stream);
/* gemm_prepared_fused (cutlass_gemm_api.h:440-470; cutlass_gemm_api.cu:951-1007): fuse_factor
 * consecutive batch elements run as one product of M * fuse_factor rows (A is [batch][M][K],
 * so that is the same memory), effective batch e using prepared weight set e, which equals
 * gemm_prepared when the prepared sets are one B repeated.  fuse_factor 0 picks the
 * reference's rule (M_fused in [128, 1024] where the batch allows); 1 is gemm_prepared.
 * FP16 planar A, FP32 output only (FTDX_ERR_UNSUPPORTED otherwise), FP8-prepared B. */
int ftdx_gemm_prepared_fused(ftdx_context* ctx, const void* A_re, const void* A_im, void* C_re, void* C_im,
                             int M, int N, int K, int batch_count, int fuse_factor, int output, float alpha,
                             float beta, void* stream);
/* debug_fp6_real_gemm (cutlass_gemm_api.h:477-490): a real M x N x K product of integers in
 * [-7, 7] through the FP6 complex GEMM (Im = 0), FP16 and FP32 outputs, compared with the
 * host product (exact; E3M2 holds these integers after block scaling); prints to stderr,
 * returns 0 when both match. */
int ftdx_debug_fp6_real_gemm(ftdx_context* ctx, int M, int N, int K, void* stream);
/* GEMM dispatch settings of the reference (set_tile_config, set_power_mode, set_gemm_mode /
 * get_gemm_mode, set_auto_tune, set_gemm_tune_verbosity, set_gemm_tune_cache_path;
 * cutlass_gemm_api.h:540-720).  They choose among the reference's CUTLASS and direct-PTX
 * GEMM kernels; this library has one GEMM tile kernel, so the values are kept (and
 * returned by the getter) without changing results or kernels.  Enum values as the
 * reference's TileConfig (0-8), PowerMode (0-2) and GemmMode (0-2). */
int ftdx_set_gemm_setting(ftdx_context* ctx, int setting, int value);
int ftdx_get_gemm_setting(const ftdx_context* ctx, int setting, int* value);
int ftdx_set_gemm_tune_cache_path(ftdx_context* ctx, const char* path);
#define FTDX_GEMM_SETTING_TILE_CONFIG 0
#define FTDX_GEMM_SETTING_POWER_MODE 1
#define FTDX_GEMM_SETTING_GEMM_MODE 2
#define FTDX_GEMM_SETTING_AUTO_TUNE 3
#define FTDX_GEMM_SETTING_TUNE_VERBOSITY 4
/* gemm (cutlass_gemm_api.h:215-225): one call with B converted each time (does not
 * touch the prepared B). */
int ftdx_gemm(ftdx_context* ctx, const void* A_re, const void* A_im, const void* B_re, const void* B_im,
              void* C_re, void* C_im, int M, int N, int K, int batch_count, int compute, int output,
              float alpha, float beta, void* stream, int tune, int input);

/* ---- the BLAS glue of the GGP library (quda namespace, lib/cutlass_interface.cu:316-459) ----
 * Synchronous, as the reference (it ends with a device synchronize); the scratch is
 * allocated per call.  Pointers are device pointers; FP32 interleaved complex data. */
/* stridedBatchGEMMCutlass's HERK case (A == B, op(A) = A^H, M == N; :316-369): A [batch][N][K]
 * narrowed to FP16 (nearest even), HERK with FP8 compute and FP16 packed output, widened into
 * C, the FP32 packed lower triangle [batch][N(N+1)/2][2]. */
int ftdx_blas_herk_fp32(ftdx_context* ctx, const void* A, void* C, int N, int K, int batch_count, int tune);
/* gemmBatchedCutlass (:371-459): A [batch][M][K], B [batch][N][K] narrowed to FP16 planar, B's
 * imaginary part negated when conjugate_b == 0 (the reference's flag sense), FP8 GEMM
 * C = A B^T with FP32 planar output, interleaved into C [batch][M][N]. */
int ftdx_blas_gemm_fp32(ftdx_context* ctx, const void* A, const void* B, void* C, int M, int N, int K,
                        int batch_count, int conjugate_b, int tune);

/* ---- streaming correlation of host payloads ----
 * VisibilityPipeline::correlate / correlate_pipelined (lib/visibility_pipeline.cu:253-300,
 * 484-545): each payload is QC bytes [batch][N][K] in host memory (K = n_time /
 * n_time_inner, batch = n_chan * n_pol * n_time_inner); the output is the packed
 * lower triangle (FP32 or FP16, output = FTDX_OUTPUT_*) in host memory.  The
 * pipeline owns two device buffer pairs, two pinned staging pairs and three
 * default-priority streams; payload p's H2D, HERK and D2H overlap its neighbours'.
 * The HERK uses ctx's settings (mode, launch configuration, graph replay).  Host
 * buffers may be pageable; a page-locked one (hipHostMalloc) is read or written by
 * the DMA engines directly, without its staging copy.  Returns when every output has
 * been written. */
typedef struct ftdx_pipeline ftdx_pipeline;
int ftdx_pipeline_create(ftdx_context* ctx, int N, int K, int batch_count, int output,
                         ftdx_pipeline** out);
int ftdx_pipeline_query(const ftdx_pipeline* pl, size_t* in_payload_bytes, size_t* out_payload_bytes);
int ftdx_pipeline_correlate(ftdx_pipeline* pl, void* vis_output, const void* raw_input);
int ftdx_pipeline_correlate_many(ftdx_pipeline* pl, void* const* vis_outputs,
                                 const void* const* raw_inputs, int n_payloads);
int ftdx_pipeline_destroy(ftdx_pipeline* pl);

/* ---- voltage beamformer (ggp::VoltagePipeline, include/voltage_pipeline.h:49-310,
 * lib/voltage_pipeline.cu:505-1338) ----
 * QC voltages [n_channels][2 pol][n_antennae][n_time] bytes -> beam power
 * |W V_pol0|^2 + |W V_pol1|^2 summed over n_time_power_sum samples, as uint8
 * rint(power * scale) clamped to [0, 255], layout [n_channels][n_beams][n_time_out],
 * or FP32 (compute_filterbank).  Weights: FP32 [re, im] [n_channels][n_beams][n_antennae],
 * rounded to FP16 then to the compute format (the reference's set_weights).
 * FP8 compute: one pass turns the voltages into GEMM rows; one GEMM over both
 * polarizations detects, sums the polarizations, integrates and quantizes in its
 * epilogue (FP32 power never reaches memory when n_time_power_sum divides 64).
 * FP6 compute: block-scaled E3M2 GEMM to planar FP32 beams, then the power sum.
 * All data pointers are device pointers.  Errors the reference raises through
 * VP_ERROR return FTDX_ERR_INVALID_ARG / _UNSUPPORTED. */
#define FTDX_VOLTAGE_FP8 0 /* VoltageComputeMode::FP8 */
#define FTDX_VOLTAGE_FP6 1 /* VoltageComputeMode::FP6 */

/* VoltagePipeline::Config (include/voltage_pipeline.h:121-156), same fields and defaults. */
typedef struct ftdx_voltage_config {
    size_t struct_size; /* sizeof(ftdx_voltage_config) */
    uint64_t n_antennae;
    uint64_t n_beams;
    uint64_t n_channels;
    uint64_t n_time;
    uint64_t n_polarizations;
    uint64_t n_time_power_sum;
    int compute_mode; /* FTDX_VOLTAGE_* */
    int kernel_tune_verbosity;
    int strategy_tune_verbosity;
    int gemm_tune;
    const char* kernel_tune_cache_path;
    const char* strategy_tune_cache_path;
    int n_dm_trials; /* dedispersion settings: kept for the layout; no dedispersion library here */
    float f_min_MHz;
    float f_max_MHz;
    float max_dm;
    float total_obs_time_s;
    int profile;
    uint64_t ch_fuse;
    uint64_t max_payloads;
} ftdx_voltage_config;

typedef struct ftdx_voltage ftdx_voltage;

/* Config defaults (include/voltage_pipeline.h:123-156). */
int ftdx_voltage_config_init(ftdx_voltage_config* cfg);
/* VoltagePipeline::VoltagePipeline (lib/voltage_pipeline.cu:509-575): validates
 * (n_polarizations == 2, n_time % n_time_power_sum == 0, ch_fuse a power of 2 dividing
 * n_channels, max_payloads in 1..16) and allocates. */
int ftdx_voltage_create(const ftdx_voltage_config* cfg, ftdx_voltage** out);
int ftdx_voltage_destroy(ftdx_voltage* vp);
/* set_weights (lib/voltage_pipeline.cu:737-784); prepares n_channels / ch_fuse weight
 * sets (the first ones), drops a captured graph. */
int ftdx_voltage_set_weights(ftdx_voltage* vp, const float* d_weights, void* stream);
/* beamform (lib/voltage_pipeline.cu:786-989): one payload -> uint8 [n_ch][n_beam][n_time_out].
 * With ch_fuse > 1 the prepared weights do not match n_channels and it fails, as the
 * reference's GEMM check makes it. */
int ftdx_voltage_beamform(ftdx_voltage* vp, uint8_t* d_output, const uint8_t* d_qc_input, float scale,
                          void* stream);
/* beamform_batched (lib/voltage_pipeline.cu:991-1120): d_qc_inputs is a host array of
 * n_payloads device pointers.  FP8 only; needs ch_fuse > 1 or n_payloads > 1.  Output
 * [n_ch / ch_fuse][n_beam][n_time * ch_fuse * n_payloads / n_time_power_sum], the
 * inner index running over (payload, channel in group, time / n_time_power_sum). */
int ftdx_voltage_beamform_batched(ftdx_voltage* vp, uint8_t* d_output, const uint8_t* const* d_qc_inputs,
                                  int n_payloads, float scale, void* stream);
/* compute_filterbank (lib/voltage_pipeline.cu:1122-1226): FP32 [n_ch][n_beam][n_time_out]
 * into the pipeline's channel-major filterbank buffer. */
int ftdx_voltage_compute_filterbank(ftdx_voltage* vp, const uint8_t* d_qc_input, void* stream);
/* compute (lib/voltage_pipeline.cu:1228-1263): filterbank, corner turn to the beam-major
 * buffer, uint8 quantization of the channel-major filterbank into d_output.  With no
 * dedispersion library d_dedisp_out is not written (the reference's dedisperse()
 * returns -1 and compute() goes on). */
int ftdx_voltage_compute(ftdx_voltage* vp, uint8_t* d_output, const uint8_t* d_qc_input, float* d_dedisp_out,
                         float scale, void* stream);
/* corner_turn (lib/voltage_pipeline.cu:1265-1273): [n_ch][n_beam][n_time_out] ->
 * [n_beam][n_ch][n_time_out]. */
int ftdx_voltage_corner_turn(ftdx_voltage* vp, float* d_out, const float* d_in, void* stream);
/* set_graph_enabled (include/voltage_pipeline.h:237): FP8 beamform() calls with the same
 * input, output and scale replay one captured graph. */
int ftdx_voltage_set_graph_enabled(ftdx_voltage* vp, int enabled);
/* output_size, input_size, n_time_out, uses_direct_path, graph_enabled, the filterbank
 * buffers (include/voltage_pipeline.h:208-303); NULL outputs are skipped. */
int ftdx_voltage_query(const ftdx_voltage* vp, int64_t* output_size, int64_t* input_size, uint64_t* n_time_out,
                       int* uses_direct_path, int* graph_enabled, const float** filterbank_chanmajor,
                       const float** filterbank_beammajor);
/* ggp::corner_turn_nf_nb (lib/corner_turn.cu:63-75): [Nf][Nb][Nt] -> [Nb][Nf][Nt] FP32. */
int ftdx_corner_turn(float* d_out, const float* d_in, int n_channels, int n_beams, int n_time, void* stream);

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
