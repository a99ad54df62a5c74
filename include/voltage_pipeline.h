// voltage_pipeline.h -- C++ drop-in for ggp::VoltagePipeline (include/voltage_pipeline.h:49-310,
// lib/voltage_pipeline.cu) over the C ABI in ftdx.h.  Header-only; link libftdx.
//
// Same Config fields and defaults, same methods and layouts.  Streams are hipStream_t
// (passed as void*; nullptr = the default stream).  Errors the reference raises
// through VP_ERROR (errorQuda, or std::runtime_error in its standalone build) are
// thrown as std::runtime_error.  set_profile / print_profile keep the reference's
// standalone behaviour (no per-stage timer); dedisperse() returns -1, the reference's
// result when no dedispersion library is linked.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>

#include "ftdx.h"

namespace ggp {

enum class VoltageComputeMode {
    FP8 = FTDX_VOLTAGE_FP8,
    FP6 = FTDX_VOLTAGE_FP6
};

class VoltagePipeline {
   public:
    struct Config {
        uint64_t n_antennae = 64;
        uint64_t n_beams = 256;
        uint64_t n_channels = 4;
        uint64_t n_time = 1024;
        uint64_t n_polarizations = 2;
        uint64_t n_time_power_sum = 4;
        VoltageComputeMode compute_mode = VoltageComputeMode::FP8;
        int kernel_tune_verbosity = 0;
        int strategy_tune_verbosity = 1;
        bool gemm_tune = true;
        const char* kernel_tune_cache_path = nullptr;
        const char* strategy_tune_cache_path = nullptr;
        int n_dm_trials = 256;
        float f_min_MHz = 700.0f;
        float f_max_MHz = 1500.0f;
        float max_dm = 1000.0f;
        float total_obs_time_s = 1.0f;
        bool profile = false;
        uint64_t ch_fuse = 1;
        uint64_t max_payloads = 1;
    };

    explicit VoltagePipeline(const Config& c) {
        ftdx_voltage_config f;
        ftdx_voltage_config_init(&f);
        f.n_antennae = c.n_antennae;
        f.n_beams = c.n_beams;
        f.n_channels = c.n_channels;
        f.n_time = c.n_time;
        f.n_polarizations = c.n_polarizations;
        f.n_time_power_sum = c.n_time_power_sum;
        f.compute_mode = static_cast<int>(c.compute_mode);
        f.kernel_tune_verbosity = c.kernel_tune_verbosity;
        f.strategy_tune_verbosity = c.strategy_tune_verbosity;
        f.gemm_tune = c.gemm_tune ? 1 : 0;
        f.kernel_tune_cache_path = c.kernel_tune_cache_path;
        f.strategy_tune_cache_path = c.strategy_tune_cache_path;
        f.n_dm_trials = c.n_dm_trials;
        f.f_min_MHz = c.f_min_MHz;
        f.f_max_MHz = c.f_max_MHz;
        f.max_dm = c.max_dm;
        f.total_obs_time_s = c.total_obs_time_s;
        f.profile = c.profile ? 1 : 0;
        f.ch_fuse = c.ch_fuse;
        f.max_payloads = c.max_payloads;
        check(ftdx_voltage_create(&f, &vp_), "VoltagePipeline");
        check(ftdx_voltage_query(vp_, &output_size_, &input_size_, &n_time_out_, &direct_, nullptr,
                                 &fb_chan_, &fb_beam_),
              "VoltagePipeline");
        n_beams_ = c.n_beams;
        n_channels_ = c.n_channels;
        n_time_ = c.n_time;
        ch_fuse_ = c.ch_fuse;
    }
    ~VoltagePipeline() { ftdx_voltage_destroy(vp_); }
    VoltagePipeline(const VoltagePipeline&) = delete;
    VoltagePipeline& operator=(const VoltagePipeline&) = delete;

    void set_weights(const float* d_weights, void* stream = nullptr) {
        check(ftdx_voltage_set_weights(vp_, d_weights, stream), "set_weights");
    }
    void beamform(uint8_t* d_output, const uint8_t* d_qc_input, float scale = 1.0f, void* stream = nullptr) {
        check(ftdx_voltage_beamform(vp_, d_output, d_qc_input, scale, stream), "beamform");
    }
    // d_qc_inputs: host array of n_payloads device pointers
    void beamform_batched(uint8_t* d_output, const uint8_t* const* d_qc_inputs, int n_payloads,
                          float scale = 1.0f, void* stream = nullptr) {
        check(ftdx_voltage_beamform_batched(vp_, d_output, d_qc_inputs, n_payloads, scale, stream),
              "beamform_batched");
    }
    void compute_filterbank(const uint8_t* d_qc_input, void* stream = nullptr) {
        check(ftdx_voltage_compute_filterbank(vp_, d_qc_input, stream), "compute_filterbank");
    }
    void compute(uint8_t* d_output, const uint8_t* d_qc_input, float* d_dedisp_out = nullptr, float scale = 1.0f,
                 void* stream = nullptr) {
        check(ftdx_voltage_compute(vp_, d_output, d_qc_input, d_dedisp_out, scale, stream), "compute");
    }
    void corner_turn(float* d_out, const float* d_in, void* stream = nullptr) {
        check(ftdx_voltage_corner_turn(vp_, d_out, d_in, stream), "corner_turn");
    }
    int dedisperse(float*, const float*, int, void* = nullptr) { return -1; }

    void set_graph_enabled(bool enabled) { check(ftdx_voltage_set_graph_enabled(vp_, enabled ? 1 : 0), "graph"); }
    bool graph_enabled() const {
        int g = 0;
        ftdx_voltage_query(vp_, nullptr, nullptr, nullptr, nullptr, &g, nullptr, nullptr);
        return g != 0;
    }
    bool uses_direct_path() const { return direct_ != 0; }
    void set_profile(bool) {}
    void print_profile() {}

    uint64_t ch_fuse() const { return ch_fuse_; }
    int64_t output_size() const { return output_size_; }
    int64_t input_size() const { return input_size_; }
    uint64_t n_channels() const { return n_channels_; }
    uint64_t n_beams() const { return n_beams_; }
    uint64_t n_time() const { return n_time_; }
    uint64_t n_time_out() const { return n_time_out_; }
    const float* filterbank_chanmajor() const { return fb_chan_; }
    const float* filterbank_beammajor() const { return fb_beam_; }

   private:
    static void check(int rc, const char* what) {
        if (rc != FTDX_OK) throw std::runtime_error(std::string("VoltagePipeline ") + what + ": " + ftdx_last_error());
    }
    ftdx_voltage* vp_ = nullptr;
    int64_t output_size_ = 0, input_size_ = 0;
    uint64_t n_time_out_ = 0;
    int direct_ = 0;
    const float* fb_chan_ = nullptr;
    const float* fb_beam_ = nullptr;
    uint64_t n_beams_ = 0, n_channels_ = 0, n_time_ = 0, ch_fuse_ = 1;
};

// ggp::corner_turn_nf_nb (lib/corner_turn.cu:63-75).
inline void corner_turn_nf_nb(float* output, const float* input, int n_channels, int n_beams, int n_time,
                              void* stream = nullptr) {
    if (ftdx_corner_turn(output, input, n_channels, n_beams, n_time, stream) != FTDX_OK)
        throw std::runtime_error(std::string("corner_turn_nf_nb: ") + ftdx_last_error());
}

}  // namespace ggp
