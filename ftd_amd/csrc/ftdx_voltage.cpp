// ftdx_voltage.cpp -- the voltage beamformer pipeline behind the C ABI
// (ggp::VoltagePipeline, include/voltage_pipeline.h:49-310, lib/voltage_pipeline.cu).
//
// FP8 compute (the reference's direct path): per call
//   1. voltage_rows: QC payload(s) -> pre-cast FP8 GEMM rows, the beam tile's order
//      (64 samples of pol 0 then the same 64 of pol 1 per 128-row block);
//   2. one GEMM against the prepared conj(weights) rows whose epilogue forms
//      |pol 0|^2 + |pol 1|^2 per sample, sums n_time_power_sum samples and writes the
//      uint8 (or FP32) filterbank.  When n_time_power_sum does not divide 64 the
//      epilogue writes per-sample Stokes-I power and beam_integrate finishes.
// The reference runs a transpose kernel, two INT4 -> FP8 casts, two power GEMMs
// (pol 1 accumulating into pol 0's FP32 buffer with beta = 1) and an integration
// kernel; the arithmetic per output is the same sequence of FP32 operations.
// FP6 compute (the reference's 4M fallback): QC rows -> block-scaled E3M2 GEMM of
// both pols (planar FP32) -> power_sum.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <string>

#include "ftdx.h"
#include "ftdx_internal.h"
#include "herk_fp8.h"
#include "stages.h"
#include "voltage.h"

struct ftdx_voltage {
    ftdx_voltage_config cfg{};
    int n_ant = 0, n_beam = 0, n_ch = 0, n_time = 0, n_tps = 0, cf = 1, max_pl = 1;
    int n_time_out = 0;
    bool direct = true;
    bool fused_epilogue = true;  // n_tps divides 64
    ftdx_context* ctx = nullptr;
    uint16_t* w_re = nullptr;    // FP16 planar weights [n_ch][n_beam][n_ant]
    uint16_t* w_im = nullptr;
    float* fb_chan = nullptr;    // [n_ch][n_beam][n_time_out]
    float* fb_beam = nullptr;    // [n_beam][n_ch][n_time_out]
    uint8_t* rows = nullptr;     // FP8: beam rows; FP6: QC rows
    float* power = nullptr;      // FP8 with n_tps not dividing 64: per-sample Stokes I
    float* beam_re = nullptr;    // FP6: planar beams [n_ch][2 n_time][n_beam]
    float* beam_im = nullptr;
    bool weights_prepared = false;
    bool graph_enabled = false;
    hipGraphExec_t graph = nullptr;
    hipStream_t cap_stream = nullptr;
    const uint8_t* g_in = nullptr;
    uint8_t* g_out = nullptr;
    float g_scale = 0.0f;
};

namespace {

using ftdx::fail;
using ftdx::hip_fail;

int Kp_of(int n_ant) { return ftdx::precast_row_samples(n_ant); }

int64_t beam_rows_per_batch(int64_t M_total) { return (M_total + 63) / 64 * 128; }

template <typename T>
int dev_alloc(T*& p, size_t count, const char* what) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count > 0 ? count * sizeof(T) : 1);
    if (e != hipSuccess) return hip_fail(e, what);
    p = static_cast<T*>(q);
    return FTDX_OK;
}

template <typename T>
void dev_free(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

void drop_graph(ftdx_voltage* vp) {
    if (vp->graph) (void)hipGraphExecDestroy(vp->graph);
    vp->graph = nullptr;
}

ftdx::VoltageLayout layout(const ftdx_voltage* vp, const uint8_t* const* inputs, int n_payloads, int cf) {
    ftdx::VoltageLayout L;
    std::memset(&L, 0, sizeof(L));
    for (int p = 0; p < n_payloads; ++p) L.payload[p] = inputs[p];
    L.n_payloads = n_payloads;
    L.n_ant = vp->n_ant;
    L.n_time = vp->n_time;
    L.n_ch = vp->n_ch;
    L.ch_fuse = cf;
    return L;
}

// FP8 path body: rows, then the beam GEMM into `out` (uint8 or FP32).
hipError_t direct_body(ftdx_voltage* vp, void* out, bool u8, const uint8_t* const* inputs, int n_payloads,
                       int cf, float scale, hipStream_t s) {
    const int batch = vp->n_ch / cf;
    const int64_t M_total = (int64_t)vp->n_time * cf * n_payloads;
    const int Kp = Kp_of(vp->n_ant);
    const ftdx::VoltageLayout L = layout(vp, inputs, n_payloads, cf);
    hipError_t e = ftdx::launch_voltage_beam_rows(L, vp->rows, s);
    if (e != hipSuccess) return e;
    const uint8_t* brows = ftdx::prepared_fp8_rows(vp->ctx, vp->n_beam, vp->n_ant, batch);
    if (!brows) return hipErrorInvalidValue;  // checked by the callers
    ftdx::GemmArgs g{};
    g.A = vp->rows;
    g.B = brows;
    g.M = (int)beam_rows_per_batch(M_total);
    g.N = vp->n_beam;
    g.K = vp->n_ant;
    g.batch = batch;
    g.alpha = scale;
    g.beta = 0.0f;
    g.lda = g.ldb = (int64_t)2 * Kp;
    g.a_bstride = (int64_t)g.M * 2 * Kp;
    g.b_bstride = (int64_t)vp->n_beam * 2 * Kp;
    g.beam_samples = (int)M_total;
    if (vp->fused_epilogue) {
        g.C = out;
        g.beam_nt = vp->n_tps;
        g.ldc = M_total / vp->n_tps;
        g.c_bstride = (int64_t)vp->n_beam * g.ldc;
        return ftdx::launch_gemm_beam(g, u8, 1, s);
    }
    g.C = vp->power;
    g.beam_nt = 1;
    g.ldc = M_total;
    g.c_bstride = (int64_t)vp->n_beam * M_total;
    if ((e = ftdx::launch_gemm_beam(g, false, 1, s)) != hipSuccess) return e;
    return ftdx::launch_beam_integrate(vp->power, out, u8, (int64_t)batch * vp->n_beam, (int)M_total, vp->n_tps,
                                       scale, s);
}

// FP6 path body (one payload, no channel fusion).
int fp6_body(ftdx_voltage* vp, void* out, bool u8, const uint8_t* input, float scale, hipStream_t s,
             const char* what) {
    const ftdx::VoltageLayout L = layout(vp, &input, 1, 1);
    hipError_t e = ftdx::launch_voltage_qc_rows(L, vp->rows, s);
    if (e != hipSuccess) return hip_fail(e, what);
    int rc = ftdx_gemm_prepared(vp->ctx, vp->rows, nullptr, vp->beam_re, vp->beam_im, 2 * vp->n_time, vp->n_beam,
                                vp->n_ant, vp->n_ch, FTDX_OUTPUT_FP32, 1.0f, 0.0f, s, vp->cfg.gemm_tune,
                                FTDX_INPUT_INT4);
    if (rc != FTDX_OK) return fail(rc, std::string(what) + ": GEMM failed: " + ftdx_last_error());
    e = ftdx::launch_power_sum(vp->beam_re, vp->beam_im, out, u8, vp->n_ch, vp->n_time, vp->n_beam, vp->n_tps,
                               scale, s);
    return e == hipSuccess ? FTDX_OK : hip_fail(e, what);
}

// beamform() / compute_filterbank(): one payload into `out`.
int single_payload(ftdx_voltage* vp, void* out, bool u8, const uint8_t* input, float scale, hipStream_t s,
                   const char* what) {
    // the weights were prepared for n_ch / ch_fuse batches; a per-channel GEMM over
    // n_ch batches fails the prepared-shape check (gemm_complex_fp8.hpp:376)
    if (vp->cf != 1)
        return fail(FTDX_ERR_INVALID_ARG, std::string(what) +
                                              ": GEMM failed: weights are prepared for n_channels / ch_fuse "
                                              "batches (use beamform_batched)");
    if (!vp->direct) return fp6_body(vp, out, u8, input, scale, s, what);
    hipError_t e = direct_body(vp, out, u8, &input, 1, 1, scale, s);
    return e == hipSuccess ? FTDX_OK : hip_fail(e, what);
}

}  // namespace

extern "C" {

int ftdx_voltage_config_init(ftdx_voltage_config* c) {
    if (!c) return fail(FTDX_ERR_INVALID_ARG, "ftdx_voltage_config_init: NULL config");
    std::memset(c, 0, sizeof(*c));
    c->struct_size = sizeof(*c);
    c->n_antennae = 64;
    c->n_beams = 256;
    c->n_channels = 4;
    c->n_time = 1024;
    c->n_polarizations = 2;
    c->n_time_power_sum = 4;
    c->compute_mode = FTDX_VOLTAGE_FP8;
    c->kernel_tune_verbosity = 0;
    c->strategy_tune_verbosity = 1;
    c->gemm_tune = 1;
    c->n_dm_trials = 256;
    c->f_min_MHz = 700.0f;
    c->f_max_MHz = 1500.0f;
    c->max_dm = 1000.0f;
    c->total_obs_time_s = 1.0f;
    c->profile = 0;
    c->ch_fuse = 1;
    c->max_payloads = 1;
    return FTDX_OK;
}

int ftdx_voltage_create(const ftdx_voltage_config* c, ftdx_voltage** out) {
    const char* W = "VoltagePipeline";
    if (!c || !out) return fail(FTDX_ERR_INVALID_ARG, "ftdx_voltage_create: NULL argument");
    *out = nullptr;
    if (c->struct_size != sizeof(ftdx_voltage_config))
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_voltage_create: struct_size mismatch (use ftdx_voltage_config_init)");
    if (c->n_polarizations != 2) return fail(FTDX_ERR_INVALID_ARG, "VoltagePipeline requires n_polarizations == 2");
    if (c->n_time_power_sum < 1 || c->n_time % c->n_time_power_sum != 0)
        return fail(FTDX_ERR_INVALID_ARG, "n_time (" + std::to_string(c->n_time) +
                                              ") must be divisible by n_time_power_sum (" +
                                              std::to_string(c->n_time_power_sum) + ")");
    if (c->n_antennae < 1 || c->n_beams < 1 || c->n_channels < 1 || c->n_time < 1)
        return fail(FTDX_ERR_INVALID_ARG, "All dimensions must be >= 1");
    if (c->ch_fuse < 1 || (c->ch_fuse & (c->ch_fuse - 1)) != 0)
        return fail(FTDX_ERR_INVALID_ARG, "ch_fuse must be a power of 2, got " + std::to_string(c->ch_fuse));
    if (c->n_channels % c->ch_fuse != 0)
        return fail(FTDX_ERR_INVALID_ARG, "n_channels (" + std::to_string(c->n_channels) +
                                              ") must be divisible by ch_fuse (" + std::to_string(c->ch_fuse) + ")");
    if (c->max_payloads < 1 || c->max_payloads > 16)
        return fail(FTDX_ERR_INVALID_ARG, "max_payloads must be 1-16, got " + std::to_string(c->max_payloads));
    if (c->compute_mode != FTDX_VOLTAGE_FP8 && c->compute_mode != FTDX_VOLTAGE_FP6)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_voltage_create: unknown compute_mode");
    const uint64_t M_max = c->n_time * c->ch_fuse * c->max_payloads;
    if (c->n_antennae > (1u << 20) || c->n_beams > (1u << 30) || c->n_channels > (1u << 30) ||
        2 * M_max + 128 > 0x7FFFFFFFu || c->n_channels * c->n_beams * c->n_antennae > (1ull << 40))
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_voltage_create: sizes too large");

    ftdx_voltage* vp = new (std::nothrow) ftdx_voltage();
    if (!vp) return fail(FTDX_ERR_NO_MEMORY, "ftdx_voltage_create: host allocation");
    vp->cfg = *c;
    vp->cfg.kernel_tune_cache_path = nullptr;  // not kept: the caller owns the strings
    vp->cfg.strategy_tune_cache_path = nullptr;
    vp->n_ant = (int)c->n_antennae;
    vp->n_beam = (int)c->n_beams;
    vp->n_ch = (int)c->n_channels;
    vp->n_time = (int)c->n_time;
    vp->n_tps = (int)c->n_time_power_sum;
    vp->cf = (int)c->ch_fuse;
    vp->max_pl = (int)c->max_payloads;
    vp->n_time_out = vp->n_time / vp->n_tps;
    vp->direct = c->compute_mode == FTDX_VOLTAGE_FP8;
    vp->fused_epilogue = 64 % vp->n_tps == 0;

    int rc = ftdx_create(&vp->ctx);
    if (rc == FTDX_OK && c->kernel_tune_verbosity > 0)
        rc = ftdx_set_kernel_tune_verbosity(vp->ctx, c->kernel_tune_verbosity);
    if (rc == FTDX_OK && c->kernel_tune_cache_path)
        rc = ftdx_set_kernel_tune_cache_path(vp->ctx, c->kernel_tune_cache_path);
    if (rc == FTDX_OK) rc = ftdx_set_strategy_tune_verbosity(vp->ctx, c->strategy_tune_verbosity);
    if (rc == FTDX_OK && c->strategy_tune_cache_path)
        rc = ftdx_set_tune_cache_path(vp->ctx, c->strategy_tune_cache_path);

    const size_t n_w = (size_t)vp->n_ch * vp->n_beam * vp->n_ant;
    const size_t n_fb = (size_t)vp->n_ch * vp->n_beam * vp->n_time_out;
    if (rc == FTDX_OK) rc = dev_alloc(vp->w_re, n_w, W);
    if (rc == FTDX_OK) rc = dev_alloc(vp->w_im, n_w, W);
    if (rc == FTDX_OK) rc = dev_alloc(vp->fb_chan, n_fb, W);
    if (rc == FTDX_OK) rc = dev_alloc(vp->fb_beam, n_fb, W);
    if (vp->direct) {
        // rows for the largest call: a single payload over n_ch batches, or the fused
        // batch (n_ch / ch_fuse batches of n_time * ch_fuse * max_payloads samples)
        const int64_t single = (int64_t)vp->n_ch * beam_rows_per_batch(vp->n_time);
        const int64_t fused = (int64_t)(vp->n_ch / vp->cf) * beam_rows_per_batch((int64_t)M_max);
        const int64_t nrows = single > fused ? single : fused;
        if (rc == FTDX_OK) rc = dev_alloc(vp->rows, (size_t)nrows * 2 * Kp_of(vp->n_ant), W);
        if (rc == FTDX_OK && !vp->fused_epilogue) {
            const int64_t ps = (int64_t)vp->n_ch * vp->n_time;
            const int64_t pf = (int64_t)(vp->n_ch / vp->cf) * (int64_t)M_max;
            rc = dev_alloc(vp->power, (size_t)(ps > pf ? ps : pf) * vp->n_beam, W);
        }
    } else {
        const size_t n_rows = (size_t)vp->n_ch * 2 * vp->n_time;
        if (rc == FTDX_OK) rc = dev_alloc(vp->rows, n_rows * vp->n_ant, W);
        if (rc == FTDX_OK) rc = dev_alloc(vp->beam_re, n_rows * vp->n_beam, W);
        if (rc == FTDX_OK) rc = dev_alloc(vp->beam_im, n_rows * vp->n_beam, W);
    }
    if (rc != FTDX_OK) {
        std::string msg = ftdx_last_error();
        ftdx_voltage_destroy(vp);
        return fail(rc, msg);
    }
    *out = vp;
    return FTDX_OK;
}

int ftdx_voltage_destroy(ftdx_voltage* vp) {
    if (!vp) return FTDX_OK;
    drop_graph(vp);
    if (vp->cap_stream) (void)hipStreamDestroy(vp->cap_stream);
    dev_free(vp->w_re);
    dev_free(vp->w_im);
    dev_free(vp->fb_chan);
    dev_free(vp->fb_beam);
    dev_free(vp->rows);
    dev_free(vp->power);
    dev_free(vp->beam_re);
    dev_free(vp->beam_im);
    if (vp->ctx) ftdx_destroy(vp->ctx);
    delete vp;
    return FTDX_OK;
}

int ftdx_voltage_set_weights(ftdx_voltage* vp, const float* d_weights, void* stream) {
    const char* W = "VoltagePipeline::set_weights";
    if (!vp || !d_weights) return fail(FTDX_ERR_INVALID_ARG, std::string(W) + ": NULL argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t n = (int64_t)vp->n_ch * vp->n_beam * vp->n_ant;
    hipError_t e = ftdx::launch_deinterleave_fp16(d_weights, vp->w_re, vp->w_im, n, s);
    if (e != hipSuccess) return hip_fail(e, W);
    // with ch_fuse > 1 the weights are shared by ch_fuse consecutive channels: the first
    // n_ch / ch_fuse weight sets are prepared (lib/voltage_pipeline.cu:770-776)
    int rc = ftdx_prepare_b(vp->ctx, vp->w_re, vp->w_im, vp->n_beam, vp->n_ant, vp->n_ch / vp->cf,
                            vp->direct ? FTDX_COMPUTE_FP8 : FTDX_COMPUTE_FP6, stream);
    if (rc != FTDX_OK) return fail(rc, std::string(W) + ": " + ftdx_last_error());
    vp->weights_prepared = true;
    drop_graph(vp);
    return FTDX_OK;
}

int ftdx_voltage_beamform(ftdx_voltage* vp, uint8_t* d_output, const uint8_t* d_qc_input, float scale,
                          void* stream) {
    const char* W = "VoltagePipeline::beamform";
    if (!vp || !d_output || !d_qc_input) return fail(FTDX_ERR_INVALID_ARG, std::string(W) + ": NULL argument");
    if (!vp->weights_prepared)
        return fail(FTDX_ERR_INVALID_ARG, "set_weights() must be called before beamform()");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (vp->graph_enabled && vp->direct && vp->cf == 1) {
        if (!(vp->graph && vp->g_in == d_qc_input && vp->g_out == d_output && vp->g_scale == scale)) {
            drop_graph(vp);
            hipError_t e;
            if (!vp->cap_stream && (e = hipStreamCreateWithFlags(&vp->cap_stream, hipStreamNonBlocking)) != hipSuccess)
                return hip_fail(e, "beamform: capture stream");
            if ((e = hipStreamBeginCapture(vp->cap_stream, hipStreamCaptureModeRelaxed)) != hipSuccess)
                return hip_fail(e, "beamform: begin capture");
            hipError_t le = direct_body(vp, d_output, true, &d_qc_input, 1, 1, scale, vp->cap_stream);
            hipGraph_t graph = nullptr;
            e = hipStreamEndCapture(vp->cap_stream, &graph);
            if (le != hipSuccess || e != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                return hip_fail(le != hipSuccess ? le : e, "beamform: capture");
            }
            e = hipGraphInstantiate(&vp->graph, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                vp->graph = nullptr;
                return hip_fail(e, "beamform: graph instantiate");
            }
            vp->g_in = d_qc_input;
            vp->g_out = d_output;
            vp->g_scale = scale;
        }
        hipError_t e = hipGraphLaunch(vp->graph, s);
        return e == hipSuccess ? FTDX_OK : hip_fail(e, "beamform: graph launch");
    }
    return single_payload(vp, d_output, true, d_qc_input, scale, s, W);
}

int ftdx_voltage_beamform_batched(ftdx_voltage* vp, uint8_t* d_output, const uint8_t* const* d_qc_inputs,
                                  int n_payloads, float scale, void* stream) {
    const char* W = "VoltagePipeline::beamform_batched";
    if (!vp || !d_output || !d_qc_inputs) return fail(FTDX_ERR_INVALID_ARG, std::string(W) + ": NULL argument");
    if (!vp->weights_prepared)
        return fail(FTDX_ERR_INVALID_ARG, "set_weights() must be called before beamform_batched()");
    if (!vp->direct)
        return fail(FTDX_ERR_UNSUPPORTED, "beamform_batched() requires FP8 compute mode (direct path)");
    if (n_payloads < 1 || n_payloads > vp->max_pl)
        return fail(FTDX_ERR_INVALID_ARG, "n_payloads (" + std::to_string(n_payloads) + ") must be 1-" +
                                              std::to_string(vp->max_pl));
    if (vp->cf == 1 && n_payloads == 1)
        return fail(FTDX_ERR_INVALID_ARG, "beamform_batched() requires ch_fuse > 1 or n_payloads > 1");
    for (int p = 0; p < n_payloads; ++p)
        if (!d_qc_inputs[p]) return fail(FTDX_ERR_INVALID_ARG, std::string(W) + ": NULL payload pointer");
    hipError_t e = direct_body(vp, d_output, true, d_qc_inputs, n_payloads, vp->cf, scale,
                               static_cast<hipStream_t>(stream));
    return e == hipSuccess ? FTDX_OK : hip_fail(e, W);
}

int ftdx_voltage_compute_filterbank(ftdx_voltage* vp, const uint8_t* d_qc_input, void* stream) {
    const char* W = "VoltagePipeline::compute_filterbank";
    if (!vp || !d_qc_input) return fail(FTDX_ERR_INVALID_ARG, std::string(W) + ": NULL argument");
    if (!vp->weights_prepared)
        return fail(FTDX_ERR_INVALID_ARG, "set_weights() must be called before compute_filterbank()");
    return single_payload(vp, vp->fb_chan, false, d_qc_input, 1.0f, static_cast<hipStream_t>(stream), W);
}

int ftdx_voltage_corner_turn(ftdx_voltage* vp, float* d_out, const float* d_in, void* stream) {
    if (!vp) return fail(FTDX_ERR_INVALID_ARG, "VoltagePipeline::corner_turn: NULL pipeline");
    return ftdx_corner_turn(d_out, d_in, vp->n_ch, vp->n_beam, vp->n_time_out, stream);
}

int ftdx_voltage_compute(ftdx_voltage* vp, uint8_t* d_output, const uint8_t* d_qc_input, float* d_dedisp_out,
                         float scale, void* stream) {
    const char* W = "VoltagePipeline::compute";
    (void)d_dedisp_out;  // no dedispersion library: dedisperse() would return -1
    if (!vp || !d_output) return fail(FTDX_ERR_INVALID_ARG, std::string(W) + ": NULL argument");
    int rc = ftdx_voltage_compute_filterbank(vp, d_qc_input, stream);
    if (rc == FTDX_OK) rc = ftdx_voltage_corner_turn(vp, vp->fb_beam, vp->fb_chan, stream);
    if (rc != FTDX_OK) return rc;
    hipError_t e = ftdx::launch_quantize_u8(vp->fb_chan, d_output, scale,
                                            (int64_t)vp->n_ch * vp->n_beam * vp->n_time_out,
                                            static_cast<hipStream_t>(stream));
    return e == hipSuccess ? FTDX_OK : hip_fail(e, W);
}

int ftdx_voltage_set_graph_enabled(ftdx_voltage* vp, int enabled) {
    if (!vp) return fail(FTDX_ERR_INVALID_ARG, "ftdx_voltage_set_graph_enabled: NULL pipeline");
    vp->graph_enabled = enabled != 0;
    return FTDX_OK;
}

int ftdx_voltage_query(const ftdx_voltage* vp, int64_t* output_size, int64_t* input_size, uint64_t* n_time_out,
                       int* uses_direct_path, int* graph_enabled, const float** filterbank_chanmajor,
                       const float** filterbank_beammajor) {
    if (!vp) return fail(FTDX_ERR_INVALID_ARG, "ftdx_voltage_query: NULL pipeline");
    if (output_size) *output_size = (int64_t)vp->n_ch * vp->n_beam * vp->n_time_out;
    if (input_size) *input_size = (int64_t)vp->n_ch * 2 * vp->n_ant * vp->n_time;
    if (n_time_out) *n_time_out = (uint64_t)vp->n_time_out;
    if (uses_direct_path) *uses_direct_path = vp->direct ? 1 : 0;
    if (graph_enabled) *graph_enabled = vp->graph_enabled ? 1 : 0;
    if (filterbank_chanmajor) *filterbank_chanmajor = vp->fb_chan;
    if (filterbank_beammajor) *filterbank_beammajor = vp->fb_beam;
    return FTDX_OK;
}

int ftdx_corner_turn(float* d_out, const float* d_in, int n_channels, int n_beams, int n_time, void* stream) {
    if (n_channels < 0 || n_beams < 0 || n_time < 0)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_corner_turn: negative size");
    if ((int64_t)n_channels * n_beams * n_time == 0) return FTDX_OK;
    if (!d_out || !d_in) return fail(FTDX_ERR_INVALID_ARG, "ftdx_corner_turn: NULL pointer");
    hipError_t e = ftdx::launch_corner_turn(d_out, d_in, n_channels, n_beams, n_time,
                                            static_cast<hipStream_t>(stream));
    return e == hipSuccess ? FTDX_OK : hip_fail(e, "ftdx_corner_turn");
}

}  // extern "C"
