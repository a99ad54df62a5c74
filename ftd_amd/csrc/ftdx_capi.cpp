// ftdx_capi.cpp -- extern "C" entry points of libftdx.so (declared in include/ftdx.h).
//
// Validation, dispatch and per-context settings for the HERK kernels; the
// reference counterpart of each entry point is cited in ftdx.h.
#include "ftdx.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "ftdx_internal.h"
#include "herk_fp8.h"

struct ftdx_context {
    int device = 0;
    int herk_mode = FTDX_HERK_AUTO;
    int direct_tile = FTDX_DIRECT_TILE_AUTO;
    int graph = 0;
    int strategy_verbosity = 1;
    int kernel_verbosity = 0;
    std::string tune_cache_path;
    std::string kernel_cache_path;
    int init_N = 0, init_K = 0, init_batch = 0;
};

namespace ftdx {

namespace {
thread_local std::string g_last_error;
thread_local ftdx_context* g_default_ctx = nullptr;
}  // namespace

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(FTDX_ERR_HIP, std::string(what) + ": " + hipGetErrorName(e) + " (" +
                                  hipGetErrorString(e) + ")");
}

}  // namespace ftdx

using ftdx::fail;
using ftdx::hip_fail;

extern "C" {

const char* ftdx_last_error(void) { return ftdx::g_last_error.c_str(); }

int ftdx_abi_version(void) { return FTDX_ABI_VERSION; }

int ftdx_create(ftdx_context** out) {
    if (!out) return fail(FTDX_ERR_INVALID_ARG, "ftdx_create: out is NULL");
    ftdx_context* c = new (std::nothrow) ftdx_context();
    if (!c) return fail(FTDX_ERR_NO_MEMORY, "ftdx_create: out of host memory");
    hipError_t e = hipGetDevice(&c->device);
    if (e != hipSuccess) {
        delete c;
        return hip_fail(e, "ftdx_create: hipGetDevice");
    }
    *out = c;
    return FTDX_OK;
}

int ftdx_destroy(ftdx_context* ctx) {
    if (!ctx) return FTDX_OK;
    ftdx_end_herk(ctx);
    delete ctx;
    return FTDX_OK;
}

int ftdx_set_herk_mode(ftdx_context* ctx, int mode) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_herk_mode: NULL context");
    if (mode < FTDX_HERK_AUTO || mode > FTDX_HERK_FORCE_BASELINE)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_herk_mode: unknown mode");
    ctx->herk_mode = mode;
    return FTDX_OK;
}

int ftdx_set_direct_tile_config(ftdx_context* ctx, int config) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_direct_tile_config: NULL context");
    if (config < FTDX_DIRECT_TILE_AUTO || config > FTDX_DIRECT_TILE_K128_B2)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_direct_tile_config: unknown config");
    ctx->direct_tile = config;
    return FTDX_OK;
}

int ftdx_set_direct_herk_graph(ftdx_context* ctx, int enable) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_direct_herk_graph: NULL context");
    ctx->graph = enable ? 1 : 0;
    return FTDX_OK;
}

int ftdx_set_tune_cache_path(ftdx_context* ctx, const char* path) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_tune_cache_path: NULL context");
    ctx->tune_cache_path = path ? path : "";
    return FTDX_OK;
}

int ftdx_set_strategy_tune_verbosity(ftdx_context* ctx, int level) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_strategy_tune_verbosity: NULL context");
    ctx->strategy_verbosity = level;
    return FTDX_OK;
}

int ftdx_set_kernel_tune_verbosity(ftdx_context* ctx, int level) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_kernel_tune_verbosity: NULL context");
    ctx->kernel_verbosity = level;
    return FTDX_OK;
}

int ftdx_set_kernel_tune_cache_path(ftdx_context* ctx, const char* path) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_kernel_tune_cache_path: NULL context");
    ctx->kernel_cache_path = path ? path : "";
    return FTDX_OK;
}

int ftdx_init_herk(ftdx_context* ctx, int N, int K, int batch_count, void* stream) {
    (void)stream;
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_init_herk: NULL context");
    if (N < 0 || K < 0 || batch_count < 0)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_init_herk: negative size");
    // The fused path needs no scratch: only remember the reservation.
    ctx->init_N = N;
    ctx->init_K = K;
    ctx->init_batch = batch_count;
    return FTDX_OK;
}

int ftdx_end_herk(ftdx_context* ctx) {
    if (!ctx) return FTDX_OK;
    ctx->init_N = ctx->init_K = ctx->init_batch = 0;
    return FTDX_OK;
}

int ftdx_herk(ftdx_context* ctx, const void* A, void* C, int N, int K, int batch_count,
              int input, int compute, int output, int format, float alpha, float beta,
              void* stream, int tune) {
    (void)tune;
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk: NULL context");
    if (N < 0 || K < 0 || batch_count < 0)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk: negative size");
    if (format != FTDX_FORMAT_PACKED_TRIANGLE)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk: only PackedTriangle output is supported");
    if (output != FTDX_OUTPUT_FP32 && output != FTDX_OUTPUT_FP16)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk: unknown output precision");
    if (compute < FTDX_COMPUTE_FP8 || compute > FTDX_COMPUTE_FP4)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk: unknown compute precision");
    if (input != FTDX_INPUT_INT4 && input != FTDX_INPUT_FP16)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk: unknown input precision");
    // cutlass_interface/cutlass_gemm_api.cu:1721-1723
    if (output == FTDX_OUTPUT_FP32 && compute != FTDX_COMPUTE_FP8)
        return fail(FTDX_ERR_UNSUPPORTED,
                    "FP32 output HERK currently only supports FP8 compute precision");
    if (N == 0 || batch_count == 0) return FTDX_OK;
    if ((!A && K > 0) || !C) return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk: NULL data pointer");
    if (reinterpret_cast<uintptr_t>(C) % (output == FTDX_OUTPUT_FP32 ? 8u : 4u) != 0)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk: C must be aligned to one complex element");
    if ((int64_t)N * (N + 1) / 2 > (int64_t)1 << 40)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk: N too large");
    if (input != FTDX_INPUT_INT4)
        return fail(FTDX_ERR_UNSUPPORTED, "ftdx_herk: FP16 input is not built yet");
    if (compute != FTDX_COMPUTE_FP8)
        return fail(FTDX_ERR_UNSUPPORTED, "ftdx_herk: FP6/FP4 compute is not built yet");

    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "ftdx_herk: hipGetDevice");
    if (dev != ctx->device)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk: context belongs to another device");

    ftdx::HerkPlainArgs a{};
    a.A = static_cast<const uint8_t*>(A);
    a.C = C;
    a.N = N;
    a.K = K;
    a.batch = batch_count;
    a.tiles = ftdx::herk_plain_tiles(N);
    a.alpha = alpha;
    a.beta = beta;
    a.a_bstride = (int64_t)N * K;
    a.c_bstride = (int64_t)N * (N + 1);
    e = ftdx::launch_herk_int4_plain(a, output == FTDX_OUTPUT_FP16,
                                     static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "ftdx_herk: kernel launch");
    return FTDX_OK;
}

int ftdx_herk_int4(const void* A, void* C, int N, int K, int batch_count, int output,
                   float alpha, float beta, void* stream) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "ftdx_herk_int4: hipGetDevice");
    ftdx_context*& d = ftdx::g_default_ctx;
    if (d && d->device != dev) {
        ftdx_destroy(d);
        d = nullptr;
    }
    if (!d) {
        int rc = ftdx_create(&d);
        if (rc != FTDX_OK) return rc;
    }
    return ftdx_herk(d, A, C, N, K, batch_count, FTDX_INPUT_INT4, FTDX_COMPUTE_FP8, output,
                     FTDX_FORMAT_PACKED_TRIANGLE, alpha, beta, stream, 0);
}

int ftdx_print_build_info(void) {
    std::printf("ftdx: MI355X X-engine HERK library, ABI %d\n", FTDX_ABI_VERSION);
    std::printf("  target: gfx950 (CDNA4), FP8 E4M3 MFMA v_mfma_f32_32x32x64_f8f6f4\n");
    int n = 0;
    if (hipGetDeviceCount(&n) == hipSuccess && n > 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) == hipSuccess)
            std::printf("  device %d: %s (%s), %d CUs, %.0f MHz, %.1f GiB\n", dev, p.name,
                        p.gcnArchName, p.multiProcessorCount, p.clockRate / 1e3,
                        p.totalGlobalMem / 1073741824.0);
    } else {
        std::printf("  no HIP device visible\n");
    }
    return FTDX_OK;
}

}  // extern "C"
