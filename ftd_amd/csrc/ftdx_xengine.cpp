// ftdx_xengine.cpp -- the GGP XEngine and the FTD raw-order correlation over the
// gfx950 HERK kernel (C ABI in include/ftdx.h).
//
// Reference: ggp::XEngine (lib/xengine.cpp:22-306) runs, per payload,
//   promoteData (GGP nibbles -> FP32/FP64 interleaved) -> cuBLAS CGEMM/ZGEMM of
//   the full N x N M^H M -> triangulateFromHerm (upper triangle, row-major),
// and its TCC branch runs convertQcSmToFtd -> reorder -> libtcc -> int2-to-float
// (lib/tcc_interface.cu:124-314).  Here both are two launches per payload:
//   1. a transposing converter writes the pre-cast FP8 operand rows of every
//      batch (stages.hip: ggp_to_fp8_rows / ftd_to_fp8_rows);
//   2. the HERK tile kernel computes only the lower-triangle tiles and writes the
//      caller's layout from its epilogue (GGP upper packed, GGP full Hermitian, or
//      the HERK lower packed), in FP32 or FP64.
// Integer inputs make every product exact: FP32 accumulation is exact while the
// sum stays below 2^24 (128 K for [-8, 7] nibbles); FP64 output longer than that
// is accumulated in K chunks with beta = 1 (exact in double).
#include "ftdx.h"

#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <string>

#include "ftdx_internal.h"
#include "herk_fp8.h"
#include "stages.h"

namespace ftdx {

namespace {

// Longest K whose FP32 sums of [-8, 7]^2 products stay exact: 128 K <= 2^24.
constexpr int kExactK = (1 << 24) / 128;

int heuristic_cfg(int N, int batch) {
    const int64_t tiles = (int64_t)herk_tiles(TILE_128x128_K64, N) * batch;
    return (N <= 128 || tiles < 512) ? TILE_64x64_K64 : TILE_128x128_K64;
}

}  // namespace

int xengine_cfg(int forced, int N, int batch) { return forced >= 0 ? forced : heuristic_cfg(N, batch); }

hipError_t run_ggp_payload(const uint8_t* in, void* out, uint8_t* rows, int N, int K, int batch,
                           int fmt, bool fp64, int cfg, int remap, hipStream_t s) {
    hipError_t e = launch_ggp_to_fp8_rows(in, rows, N, K, batch, s);
    if (e != hipSuccess) return e;
    const int Kp = precast_row_samples(K);
    HerkPlainArgs a;
    a.A = rows;
    a.C = out;
    a.N = N;
    a.batch = batch;
    a.alpha = 1.0f;
    a.beta = 0.0f;
    a.a_bstride = (int64_t)N * Kp * 2;
    a.c_bstride = fmt == OUT_HERM_FULL ? (int64_t)N * N * 2 : (int64_t)N * (N + 1);
    a.precast = 1;
    a.lda = (int64_t)Kp * 2;
    if (K <= kExactK || !fp64) {
        a.K = K;
        return launch_herk_ggp(a, cfg, fmt, fp64, remap, s);
    }
    for (int k0 = 0; k0 < K; k0 += kExactK) {  // chunks start on 16-sample boundaries
        a.A = rows + (int64_t)k0 * 2;
        a.K = K - k0 < kExactK ? K - k0 : kExactK;
        a.beta = k0 == 0 ? 0.0f : 1.0f;
        if ((e = launch_herk_ggp(a, cfg, fmt, fp64, remap, s)) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t run_ftd_payload(const uint8_t* raw, float* tri, uint8_t* rows, int n_ant, int n_chan,
                           int n_time, int n_time_inner, int n_pol, bool twos, int cfg, int remap,
                           hipStream_t s) {
    const int K = n_time / n_time_inner;
    const int batch = n_chan * n_pol * n_time_inner;
    hipError_t e = launch_ftd_to_fp8_rows(raw, rows, n_ant, n_chan, n_time, n_time_inner, n_pol, twos, s);
    if (e != hipSuccess) return e;
    const int Kp = precast_row_samples(K);
    HerkPlainArgs a;
    a.A = rows;
    a.C = tri;
    a.N = n_ant;
    a.K = K;
    a.batch = batch;
    a.alpha = 1.0f;
    a.beta = 0.0f;
    a.a_bstride = (int64_t)n_ant * Kp * 2;
    a.c_bstride = (int64_t)n_ant * (n_ant + 1);
    a.precast = 1;
    a.lda = (int64_t)Kp * 2;
    return launch_herk_int4_plain(a, cfg, false, remap, s);
}

size_t rows_bytes(int N, int K, int batch) {
    return (size_t)batch * N * precast_row_samples(K) * 2;
}

}  // namespace ftdx

struct ftdx_xengine {
    ftdx_xengine_param p;
    int N = 0, K = 0, batch = 0;
    bool tcc = false, fp64 = false, herm = false;
    size_t in_payload = 0, out_payload = 0;
    uint8_t* d_in = nullptr;   // device copy of a host input payload
    uint8_t* d_out = nullptr;  // device output of a payload bound for the host
    uint8_t* rows = nullptr;   // pre-cast FP8 operand rows of one payload
    int cfg = -1, remap = -1;
};

using ftdx::fail;
using ftdx::hip_fail;

namespace {

int check_param(const ftdx_xengine_param* p) {
    if (!p) return fail(FTDX_ERR_INVALID_ARG, "xengine: param is NULL");
    if (p->struct_size != (size_t)FTDX_INVALID_ENUM && p->struct_size != sizeof(*p))
        return fail(FTDX_ERR_INVALID_ARG, "xengine: unexpected XEngineParam struct size " +
                                              std::to_string(p->struct_size) + ", expected " +
                                              std::to_string(sizeof(*p)));
    if (p->compute_prec != FTDX_QPREC_SINGLE && p->compute_prec != FTDX_QPREC_DOUBLE)
        return fail(FTDX_ERR_UNSUPPORTED, "xengine: unsupported XEngine compute precision " +
                                              std::to_string(p->compute_prec));
    if (p->output_prec != p->compute_prec)
        return fail(FTDX_ERR_UNSUPPORTED,
                    "xengine: output and compute precisions must match until a copy engine is formulated");
    if (p->format != FTDX_XENGINE_MAT_HERM && p->format != FTDX_XENGINE_MAT_TRI)
        return fail(FTDX_ERR_INVALID_ARG, "xengine: unknown XEngine mat format " + std::to_string(p->format));
    if (p->packet_format == FTDX_PACKET_DSA110)
        return fail(FTDX_ERR_UNSUPPORTED,
                    "xengine: DSA110 packets are not supported (the reference reorder writes "
                    "overlapping addresses, so it defines no result to match)");
    if (p->packet_format != FTDX_PACKET_DSA2K)
        return fail(FTDX_ERR_INVALID_ARG, "xengine: unknown packet format " + std::to_string(p->packet_format));
    if (p->in_location != FTDX_LOCATION_HOST && p->in_location != FTDX_LOCATION_DEVICE)
        return fail(FTDX_ERR_INVALID_ARG, "xengine: in_location must be host (1) or device (2)");
    if (p->out_location != FTDX_LOCATION_HOST && p->out_location != FTDX_LOCATION_DEVICE)
        return fail(FTDX_ERR_INVALID_ARG, "xengine: out_location must be host (1) or device (2)");
    const size_t lim = 0x7FFFFFFF;
    if (p->n_payload == 0 || p->n_antennae_per_payload == 0 || p->n_channels_per_payload == 0 ||
        p->n_time_per_payload == 0 || p->n_time_inner == 0 || p->n_polarizations == 0 ||
        p->n_payload > lim || p->n_antennae_per_payload > lim || p->n_channels_per_payload > lim ||
        p->n_time_per_payload > lim || p->n_time_inner > lim || p->n_polarizations > lim)
        return fail(FTDX_ERR_INVALID_ARG, "xengine: payload sizes must be set (positive, < 2^31)");
    const size_t batch = p->n_channels_per_payload * p->n_polarizations * p->n_time_inner;
    if (batch > lim) return fail(FTDX_ERR_INVALID_ARG, "xengine: n_channels * n_pol * n_time_inner >= 2^31");
    if (p->engine == FTDX_BLAS_ENGINE_TCC) {
        if (p->format != FTDX_XENGINE_MAT_TRI)
            return fail(FTDX_ERR_INVALID_ARG,
                        "xengine: TCC engine only supports triangular output format (QUDA_XENGINE_MAT_TRI)");
        if (p->output_prec != FTDX_QPREC_SINGLE)
            return fail(FTDX_ERR_UNSUPPORTED, "xengine: TCC engine writes FP32 visibilities");
    } else if (p->engine != FTDX_BLAS_ENGINE_CUBLAS && p->engine != FTDX_BLAS_ENGINE_CUTLASS) {
        return fail(FTDX_ERR_INVALID_ARG, "xengine: unknown BLAS engine " + std::to_string(p->engine));
    }
    return FTDX_OK;
}

void release(ftdx_xengine* xe) {
    if (xe->d_in) (void)hipFree(xe->d_in);
    if (xe->d_out) (void)hipFree(xe->d_out);
    if (xe->rows) (void)hipFree(xe->rows);
    xe->d_in = xe->d_out = xe->rows = nullptr;
}

}  // namespace

extern "C" {

int ftdx_xengine_param_init(ftdx_xengine_param* p) {
    if (!p) return fail(FTDX_ERR_INVALID_ARG, "xengine_param_init: param is NULL");
    std::memset(p, 0, sizeof(*p));
    p->struct_size = (size_t)FTDX_INVALID_ENUM;
    p->in_location = FTDX_LOCATION_HOST;
    p->out_location = FTDX_LOCATION_HOST;
    p->packet_format = FTDX_PACKET_DSA2K;
    p->n_payload = p->n_antennae_per_payload = p->n_channels_per_payload = (size_t)FTDX_INVALID_ENUM;
    p->n_time_per_payload = p->n_time_inner = p->n_polarizations = (size_t)FTDX_INVALID_ENUM;
    p->compute_prec = FTDX_QPREC_HALF;
    p->output_prec = FTDX_QPREC_SINGLE;
    p->engine = FTDX_BLAS_ENGINE_CUBLAS;
    p->format = FTDX_XENGINE_MAT_HERM;
    return FTDX_OK;
}

int ftdx_xengine_create(const ftdx_xengine_param* param, ftdx_xengine** out) {
    if (!out) return fail(FTDX_ERR_INVALID_ARG, "xengine_create: out is NULL");
    *out = nullptr;
    int rc = check_param(param);
    if (rc != FTDX_OK) return rc;
    ftdx_xengine* xe = new (std::nothrow) ftdx_xengine();
    if (!xe) return fail(FTDX_ERR_NO_MEMORY, "xengine_create: out of host memory");
    xe->p = *param;
    const ftdx_xengine_param& p = xe->p;
    xe->N = (int)p.n_antennae_per_payload;
    xe->K = (int)(p.n_time_per_payload / p.n_time_inner);
    xe->batch = (int)(p.n_channels_per_payload * p.n_polarizations * p.n_time_inner);
    xe->tcc = p.engine == FTDX_BLAS_ENGINE_TCC;
    xe->fp64 = p.output_prec == FTDX_QPREC_DOUBLE;
    xe->herm = p.format == FTDX_XENGINE_MAT_HERM;
    // lib/xengine.cpp:52-90: input bytes, then output bytes of one payload
    xe->in_payload = p.n_time_per_payload * p.n_antennae_per_payload * p.n_channels_per_payload *
                     p.n_polarizations;
    const size_t elem = xe->fp64 ? 16 : 8;
    const size_t N = p.n_antennae_per_payload;
    const size_t per_batch = xe->herm ? N * N : N * (N + 1) / 2;
    xe->out_payload = elem * per_batch * (size_t)xe->batch;
    hipError_t e;
    if (p.in_location == FTDX_LOCATION_HOST && (e = hipMalloc(&xe->d_in, xe->in_payload)) != hipSuccess) {
        release(xe);
        delete xe;
        return hip_fail(e, "xengine_create: input copy");
    }
    if (p.out_location == FTDX_LOCATION_HOST &&
        (e = hipMalloc(&xe->d_out, xe->out_payload)) != hipSuccess) {
        release(xe);
        delete xe;
        return hip_fail(e, "xengine_create: output copy");
    }
    if ((e = hipMalloc(&xe->rows, ftdx::rows_bytes(xe->N, xe->K, xe->batch) + 16)) != hipSuccess) {
        release(xe);
        delete xe;
        return hip_fail(e, "xengine_create: operand rows");
    }
    *out = xe;
    return FTDX_OK;
}

int ftdx_xengine_set_launch_config(ftdx_xengine* xe, int tile_config, int xcd_remap) {
    if (!xe) return fail(FTDX_ERR_INVALID_ARG, "xengine_set_launch_config: NULL engine");
    if (tile_config < -1 || tile_config >= ftdx::NUM_TILE_CONFIGS || xcd_remap < -1 || xcd_remap > 1)
        return fail(FTDX_ERR_INVALID_ARG, "xengine_set_launch_config: bad tile config or remap");
    xe->cfg = tile_config;
    xe->remap = xcd_remap;
    return FTDX_OK;
}

int ftdx_xengine_query(const ftdx_xengine* xe, size_t* in_payload_bytes, size_t* out_payload_bytes,
                       int* N, int* K, int* batch_count) {
    if (!xe) return fail(FTDX_ERR_INVALID_ARG, "xengine_query: NULL engine");
    if (in_payload_bytes) *in_payload_bytes = xe->in_payload;
    if (out_payload_bytes) *out_payload_bytes = xe->out_payload;
    if (N) *N = xe->N;
    if (K) *K = xe->K;
    if (batch_count) *batch_count = xe->batch;
    return FTDX_OK;
}

int ftdx_xengine_compute(ftdx_xengine* xe, void* output, const void* input, void* stream) {
    if (!xe || !output || !input) return fail(FTDX_ERR_INVALID_ARG, "xengine_compute: NULL argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const ftdx_xengine_param& p = xe->p;
    const int cfg = ftdx::xengine_cfg(xe->cfg, xe->N, xe->batch);
    const int remap = xe->remap >= 0 ? xe->remap : 1;
    hipError_t e;
    for (size_t i = 0; i < p.n_payload; ++i) {
        const uint8_t* src = static_cast<const uint8_t*>(input) + i * xe->in_payload;
        uint8_t* dst = static_cast<uint8_t*>(output) + i * xe->out_payload;
        if (p.in_location == FTDX_LOCATION_HOST) {
            if ((e = hipMemcpyAsync(xe->d_in, src, xe->in_payload, hipMemcpyHostToDevice, s)) != hipSuccess)
                return hip_fail(e, "xengine_compute: input copy");
            src = xe->d_in;
        }
        void* res = p.out_location == FTDX_LOCATION_HOST ? static_cast<void*>(xe->d_out) : dst;
        if (xe->tcc)
            e = ftdx::run_ftd_payload(src, static_cast<float*>(res), xe->rows, xe->N,
                                      (int)p.n_channels_per_payload, (int)p.n_time_per_payload,
                                      (int)p.n_time_inner, (int)p.n_polarizations, false, cfg, remap, s);
        else
            e = ftdx::run_ggp_payload(src, res, xe->rows, xe->N, xe->K, xe->batch,
                                      xe->herm ? ftdx::OUT_HERM_FULL : ftdx::OUT_UPPER_PACKED, xe->fp64, cfg,
                                      remap, s);
        if (e != hipSuccess) return hip_fail(e, "xengine_compute: kernels");
        if (p.out_location == FTDX_LOCATION_HOST &&
            (e = hipMemcpyAsync(dst, xe->d_out, xe->out_payload, hipMemcpyDeviceToHost, s)) != hipSuccess)
            return hip_fail(e, "xengine_compute: output copy");
    }
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "xengine_compute: synchronize");
    return FTDX_OK;
}

int ftdx_xengine_destroy(ftdx_xengine* xe) {
    if (!xe) return FTDX_OK;
    release(xe);
    delete xe;
    return FTDX_OK;
}

int ftdx_xengine_chpc(void* output, const void* input, const ftdx_xengine_param* param) {
    ftdx_xengine* xe = nullptr;
    int rc = ftdx_xengine_create(param, &xe);
    if (rc != FTDX_OK) return rc;
    rc = ftdx_xengine_compute(xe, output, input, nullptr);
    ftdx_xengine_destroy(xe);
    return rc;
}

}  // extern "C"
