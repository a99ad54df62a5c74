// ftdx_pipeline.cpp -- streaming correlation of host payloads (C ABI in ftdx.h).
//
// Reference: VisibilityPipeline::correlate / correlate_pipelined
// (lib/visibility_pipeline.cu:253-300, 484-545): raw QC payloads in host memory ->
// H2D -> herkBatchedCutlassQC[_FP16] -> D2H, ping-ponging two device buffer pairs
// over three streams (H2D, compute, D2H) with one pinned staging buffer per
// direction, synchronizing on the D2H of every payload.
//
// Here the pinned staging is double-buffered as well and the host copy-out of
// payload p-1 runs while payload p is in flight, so host memcpy, both DMA
// directions and the HERK of neighbouring payloads overlap.  Buffer reuse is
// ordered by events:
//   staging_in[b]  reused after H2D of payload p-2 completed      (host waits)
//   d_in[b]        reused after the HERK of payload p-2 completed (H2D stream waits)
//   d_out[b]       reused after D2H of payload p-2 completed      (compute stream waits)
//   staging_out[b] read by the host after D2H of its payload       (host waits)
// The host copies between caller memory (pageable) and the pinned staging are the
// pipeline's bottleneck (one thread copies at 10-15 GB/s, the DMA engines move several
// times that), so large copies are split over up to 8 threads.  A caller buffer that is
// itself page-locked (hipHostMalloc, torch pin_memory) skips its staging copy: the DMA
// engines read or write it directly.
// All streams use the default priority.
#include "ftdx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "ftdx_host.h"
#include "ftdx_internal.h"

struct ftdx_pipeline {
    ftdx_context* ctx = nullptr;
    int N = 0, K = 0, batch = 0, output = FTDX_OUTPUT_FP32;
    size_t in_bytes = 0, out_bytes = 0;
    hipStream_t s_h2d = nullptr, s_comp = nullptr, s_d2h = nullptr;
    uint8_t* d_in[2] = {nullptr, nullptr};
    uint8_t* d_out[2] = {nullptr, nullptr};
    uint8_t* h_in[2] = {nullptr, nullptr};   // pinned staging
    uint8_t* h_out[2] = {nullptr, nullptr};
    hipEvent_t ev_h2d[2] = {nullptr, nullptr};
    hipEvent_t ev_comp[2] = {nullptr, nullptr};
    hipEvent_t ev_d2h[2] = {nullptr, nullptr};
    bool used[2] = {false, false};  // buffer b carried a payload whose events are recorded
};

using ftdx::fail;
using ftdx::hip_fail;
using ftdx::host_copy;

namespace {

// True if p is page-locked host memory that the DMA engines can read or write directly.
// hipPointerGetAttributes rejects pageable pointers; that error is cleared here so the
// next call does not report it.
bool pinned_host(const void* p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

void release(ftdx_pipeline* pl) {
    if (pl->s_comp) (void)hipStreamSynchronize(pl->s_comp);
    if (pl->s_h2d) (void)hipStreamSynchronize(pl->s_h2d);
    if (pl->s_d2h) (void)hipStreamSynchronize(pl->s_d2h);
    for (int b = 0; b < 2; ++b) {
        if (pl->d_in[b]) (void)hipFree(pl->d_in[b]);
        if (pl->d_out[b]) (void)hipFree(pl->d_out[b]);
        if (pl->h_in[b]) (void)hipHostFree(pl->h_in[b]);
        if (pl->h_out[b]) (void)hipHostFree(pl->h_out[b]);
        if (pl->ev_h2d[b]) (void)hipEventDestroy(pl->ev_h2d[b]);
        if (pl->ev_comp[b]) (void)hipEventDestroy(pl->ev_comp[b]);
        if (pl->ev_d2h[b]) (void)hipEventDestroy(pl->ev_d2h[b]);
    }
    if (pl->s_h2d) (void)hipStreamDestroy(pl->s_h2d);
    if (pl->s_comp) (void)hipStreamDestroy(pl->s_comp);
    if (pl->s_d2h) (void)hipStreamDestroy(pl->s_d2h);
}

int herk_on(ftdx_pipeline* pl, const void* d_in, void* d_out, hipStream_t s) {
    return ftdx_herk(pl->ctx, d_in, d_out, pl->N, pl->K, pl->batch, FTDX_INPUT_INT4, FTDX_COMPUTE_FP8,
                     pl->output, FTDX_FORMAT_PACKED_TRIANGLE, 1.0f, 0.0f, s, 0);
}

}  // namespace

extern "C" {

int ftdx_pipeline_create(ftdx_context* ctx, int N, int K, int batch_count, int output,
                         ftdx_pipeline** out) {
    if (!out) return fail(FTDX_ERR_INVALID_ARG, "pipeline_create: out is NULL");
    *out = nullptr;
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "pipeline_create: NULL context");
    if (N <= 0 || K < 0 || batch_count <= 0)
        return fail(FTDX_ERR_INVALID_ARG, "pipeline_create: N and batch must be positive, K >= 0");
    if (output != FTDX_OUTPUT_FP32 && output != FTDX_OUTPUT_FP16)
        return fail(FTDX_ERR_INVALID_ARG, "pipeline_create: output must be FTDX_OUTPUT_FP32 or _FP16");
    ftdx_pipeline* pl = new (std::nothrow) ftdx_pipeline();
    if (!pl) return fail(FTDX_ERR_NO_MEMORY, "pipeline_create: out of host memory");
    pl->ctx = ctx;
    pl->N = N;
    pl->K = K;
    pl->batch = batch_count;
    pl->output = output;
    pl->in_bytes = (size_t)N * K * batch_count;
    pl->out_bytes = (size_t)N * (N + 1) * batch_count * (output == FTDX_OUTPUT_FP16 ? 2 : 4);
    hipError_t e = hipSuccess;
    const size_t ib = pl->in_bytes > 0 ? pl->in_bytes : 1;
    if ((e = hipStreamCreateWithFlags(&pl->s_h2d, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&pl->s_comp, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&pl->s_d2h, hipStreamNonBlocking)) != hipSuccess) {
        release(pl);
        delete pl;
        return hip_fail(e, "pipeline_create: streams");
    }
    for (int b = 0; b < 2 && e == hipSuccess; ++b) {
        if ((e = hipMalloc(&pl->d_in[b], ib)) != hipSuccess) break;
        if ((e = hipMalloc(&pl->d_out[b], pl->out_bytes)) != hipSuccess) break;
        if ((e = hipHostMalloc(&pl->h_in[b], ib, hipHostMallocDefault)) != hipSuccess) break;
        if ((e = hipHostMalloc(&pl->h_out[b], pl->out_bytes, hipHostMallocDefault)) != hipSuccess) break;
        if ((e = hipEventCreateWithFlags(&pl->ev_h2d[b], hipEventDisableTiming)) != hipSuccess) break;
        if ((e = hipEventCreateWithFlags(&pl->ev_comp[b], hipEventDisableTiming)) != hipSuccess) break;
        if ((e = hipEventCreateWithFlags(&pl->ev_d2h[b], hipEventDisableTiming)) != hipSuccess) break;
    }
    if (e != hipSuccess) {
        release(pl);
        delete pl;
        return hip_fail(e, "pipeline_create: buffers");
    }
    *out = pl;
    return FTDX_OK;
}

int ftdx_pipeline_query(const ftdx_pipeline* pl, size_t* in_payload_bytes, size_t* out_payload_bytes) {
    if (!pl) return fail(FTDX_ERR_INVALID_ARG, "pipeline_query: NULL pipeline");
    if (in_payload_bytes) *in_payload_bytes = pl->in_bytes;
    if (out_payload_bytes) *out_payload_bytes = pl->out_bytes;
    return FTDX_OK;
}

int ftdx_pipeline_correlate(ftdx_pipeline* pl, void* vis_output, const void* raw_input) {
    const void* in[1] = {raw_input};
    void* outp[1] = {vis_output};
    return ftdx_pipeline_correlate_many(pl, outp, in, 1);
}

int ftdx_pipeline_correlate_many(ftdx_pipeline* pl, void* const* vis_outputs,
                                 const void* const* raw_inputs, int n_payloads) {
    if (!pl) return fail(FTDX_ERR_INVALID_ARG, "pipeline_correlate: NULL pipeline");
    if (n_payloads < 0 || (n_payloads > 0 && (!vis_outputs || !raw_inputs)))
        return fail(FTDX_ERR_INVALID_ARG, "pipeline_correlate: bad payload arrays");
    for (int p = 0; p < n_payloads; ++p)
        if (!vis_outputs[p] || (!raw_inputs[p] && pl->in_bytes > 0))
            return fail(FTDX_ERR_INVALID_ARG, "pipeline_correlate: NULL payload pointer");
    hipError_t e;
    int rc = FTDX_OK;
    std::vector<char> direct_out(n_payloads > 0 ? n_payloads : 1, 0);
    // copy-out of payload q from its pinned staging to the caller (nothing to copy when
    // the D2H wrote the caller's page-locked buffer itself)
    auto drain = [&](int q) -> int {
        const int b = q & 1;
        hipError_t de = hipEventSynchronize(pl->ev_d2h[b]);
        if (de != hipSuccess) return hip_fail(de, "pipeline_correlate: D2H");
        if (!direct_out[q]) host_copy(vis_outputs[q], pl->h_out[b], pl->out_bytes);
        return FTDX_OK;
    };
    for (int p = 0; p < n_payloads && rc == FTDX_OK; ++p) {
        const int b = p & 1;
        const bool direct_in = pl->in_bytes > 0 && pinned_host(raw_inputs[p]);
        direct_out[p] = pinned_host(vis_outputs[p]) ? 1 : 0;
        if (pl->used[b]) {
            // staging_in[b] was last read by the H2D two payloads back
            if ((e = hipEventSynchronize(pl->ev_h2d[b])) != hipSuccess) {
                rc = hip_fail(e, "pipeline_correlate: H2D");
                break;
            }
        }
        if (pl->in_bytes > 0 && !direct_in) host_copy(pl->h_in[b], raw_inputs[p], pl->in_bytes);
        if (pl->used[b] && (e = hipStreamWaitEvent(pl->s_h2d, pl->ev_comp[b], 0)) != hipSuccess) {
            rc = hip_fail(e, "pipeline_correlate: wait");
            break;
        }
        if (pl->in_bytes > 0 &&
            (e = hipMemcpyAsync(pl->d_in[b], direct_in ? raw_inputs[p] : pl->h_in[b], pl->in_bytes,
                                hipMemcpyHostToDevice, pl->s_h2d)) != hipSuccess) {
            rc = hip_fail(e, "pipeline_correlate: H2D");
            break;
        }
        if ((e = hipEventRecord(pl->ev_h2d[b], pl->s_h2d)) != hipSuccess ||
            (e = hipStreamWaitEvent(pl->s_comp, pl->ev_h2d[b], 0)) != hipSuccess) {
            rc = hip_fail(e, "pipeline_correlate: events");
            break;
        }
        if (pl->used[b] && (e = hipStreamWaitEvent(pl->s_comp, pl->ev_d2h[b], 0)) != hipSuccess) {
            rc = hip_fail(e, "pipeline_correlate: wait");
            break;
        }
        if ((rc = herk_on(pl, pl->d_in[b], pl->d_out[b], pl->s_comp)) != FTDX_OK) break;
        if ((e = hipEventRecord(pl->ev_comp[b], pl->s_comp)) != hipSuccess ||
            (e = hipStreamWaitEvent(pl->s_d2h, pl->ev_comp[b], 0)) != hipSuccess) {
            rc = hip_fail(e, "pipeline_correlate: events");
            break;
        }
        // staging_out[b] is free: its previous payload (p-2) was drained in iteration p-1
        if ((e = hipMemcpyAsync(direct_out[p] ? vis_outputs[p] : pl->h_out[b], pl->d_out[b], pl->out_bytes,
                                hipMemcpyDeviceToHost, pl->s_d2h)) != hipSuccess ||
            (e = hipEventRecord(pl->ev_d2h[b], pl->s_d2h)) != hipSuccess) {
            rc = hip_fail(e, "pipeline_correlate: D2H");
            break;
        }
        pl->used[b] = true;
        if (p > 0) rc = drain(p - 1);
    }
    if (rc == FTDX_OK && n_payloads > 0) rc = drain(n_payloads - 1);
    if (rc != FTDX_OK) {
        // nothing may stay queued against the caller's buffers (page-locked ones are read
        // and written by the copies directly)
        (void)hipStreamSynchronize(pl->s_h2d);
        (void)hipStreamSynchronize(pl->s_comp);
        (void)hipStreamSynchronize(pl->s_d2h);
    }
    return rc;
}

int ftdx_pipeline_destroy(ftdx_pipeline* pl) {
    if (!pl) return FTDX_OK;
    release(pl);
    delete pl;
    return FTDX_OK;
}

}  // extern "C"
