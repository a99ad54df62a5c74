// test_voltage_api.cpp -- the C++ drop-in ggp::VoltagePipeline on the GPU, in the manner of
// the reference's tests/voltage_pipeline_test.cpp: a lifecycle check of the sizes, then
// FP8 and FP6 beamforming of random QC voltages with uniform [-1, 1) weights, compared
// with an FP64 CPU beamformer on the unrounded weights under the reference's tolerance
// (max(4, 1 + n_ant/32) for FP8, max(15, 1 + n_ant/8) for FP6; scale = 200 / max power),
// and compute() against beamform() byte for byte.
//
// usage: test_voltage_api n_ant n_beam n_ch n_time [seed]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "herk_oracle.h"
#include "voltage_pipeline.h"

#define HIP_OK(x)                                                                       \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) {                                                         \
            std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_),      \
                         __FILE__, __LINE__);                                           \
            std::exit(2);                                                               \
        }                                                                               \
    } while (0)

static double nibble(int x) { return (x & 8) ? -(double)(x & 7) : (double)(x & 7); }

// FP64 power [ch][beam][t_out] of QC [ch][pol][ant][t] against weights [ch][beam][ant][2].
static std::vector<double> cpu_power(const std::vector<uint8_t>& qc, const std::vector<float>& w, int n_ant,
                                     int n_beam, int n_ch, int n_time, int n_tps) {
    const int n_out = n_time / n_tps;
    std::vector<double> out((size_t)n_ch * n_beam * n_out, 0.0);
    for (int ch = 0; ch < n_ch; ++ch)
        for (int beam = 0; beam < n_beam; ++beam)
            for (int t = 0; t < n_time; ++t) {
                double p = 0.0;
                for (int pol = 0; pol < 2; ++pol) {
                    std::complex<double> acc(0.0, 0.0);
                    for (int a = 0; a < n_ant; ++a) {
                        const uint8_t b = qc[(((size_t)ch * 2 + pol) * n_ant + a) * n_time + t];
                        const size_t wi = (((size_t)ch * n_beam + beam) * n_ant + a) * 2;
                        acc += std::complex<double>(w[wi], w[wi + 1]) *
                               std::complex<double>(nibble(b >> 4), nibble(b & 15));
                    }
                    p += std::norm(acc);
                }
                out[((size_t)ch * n_beam + beam) * n_out + t / n_tps] += p;
            }
    return out;
}

static int run(ggp::VoltageComputeMode mode, const char* label, int n_ant, int n_beam, int n_ch, int n_time,
               const std::vector<uint8_t>& qc, const std::vector<float>& w, const std::vector<double>& ref) {
    const int n_tps = 4;
    ggp::VoltagePipeline::Config cfg;
    cfg.n_antennae = n_ant;
    cfg.n_beams = n_beam;
    cfg.n_channels = n_ch;
    cfg.n_time = n_time;
    cfg.n_time_power_sum = n_tps;
    cfg.compute_mode = mode;
    cfg.strategy_tune_verbosity = 0;
    ggp::VoltagePipeline vp(cfg);
    uint8_t *d_qc = nullptr, *d_out = nullptr, *d_out2 = nullptr;
    float* d_w = nullptr;
    HIP_OK(hipMalloc(&d_qc, vp.input_size()));
    HIP_OK(hipMalloc(&d_out, vp.output_size()));
    HIP_OK(hipMalloc(&d_out2, vp.output_size()));
    HIP_OK(hipMalloc(&d_w, w.size() * sizeof(float)));
    HIP_OK(hipMemcpy(d_qc, qc.data(), vp.input_size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_w, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
    vp.set_weights(d_w);
    double max_power = 0.0;
    for (double p : ref) max_power = std::max(max_power, p);
    const float scale = max_power > 0 ? 200.0f / (float)max_power : 1.0f;
    vp.beamform(d_out, d_qc, scale);
    vp.compute(d_out2, d_qc, nullptr, scale);
    HIP_OK(hipDeviceSynchronize());
    std::vector<uint8_t> got(vp.output_size()), got2(vp.output_size());
    HIP_OK(hipMemcpy(got.data(), d_out, got.size(), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(got2.data(), d_out2, got2.size(), hipMemcpyDeviceToHost));
    const int tol = mode == ggp::VoltageComputeMode::FP6 ? std::max(15, 1 + n_ant / 8) : std::max(4, 1 + n_ant / 32);
    int errors = 0, max_diff = 0, mismatch = 0;
    for (size_t i = 0; i < got.size(); ++i) {
        const int want = std::max(0, std::min(255, (int)std::round(ref[i] * scale)));
        const int d = std::abs((int)got[i] - want);
        max_diff = std::max(max_diff, d);
        errors += d > tol;
        mismatch += got[i] != got2[i];
    }
    std::printf("%s: max_diff=%d errors=%d/%zu (tol=%d), compute vs beamform mismatches=%d, direct=%d\n", label,
                max_diff, errors, got.size(), tol, mismatch, (int)vp.uses_direct_path());
    HIP_OK(hipFree(d_qc));
    HIP_OK(hipFree(d_out));
    HIP_OK(hipFree(d_out2));
    HIP_OK(hipFree(d_w));
    return errors + mismatch;
}

int main(int argc, char** argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s n_ant n_beam n_ch n_time [seed]\n", argv[0]);
        return 2;
    }
    const int n_ant = std::atoi(argv[1]), n_beam = std::atoi(argv[2]), n_ch = std::atoi(argv[3]),
              n_time = std::atoi(argv[4]);
    const uint64_t seed = argc > 5 ? std::strtoull(argv[5], nullptr, 10) : 1;
    std::printf("test_voltage_api n_ant=%d n_beam=%d n_ch=%d n_time=%d seed=%llu\n", n_ant, n_beam, n_ch, n_time,
                (unsigned long long)seed);
    std::fflush(stdout);
    int fails = 0;
    {   // lifecycle (the reference's VoltagePipelineLifecycleTest sizes)
        ggp::VoltagePipeline::Config cfg;
        cfg.n_antennae = 64;
        cfg.n_beams = 32;
        cfg.n_channels = 2;
        cfg.n_time = 64;
        cfg.n_time_power_sum = 4;
        ggp::VoltagePipeline vp(cfg);
        const bool ok = vp.input_size() == 2 * 2 * 64 * 64 && vp.output_size() == 2 * 32 * 16 &&
                        vp.n_channels() == 2 && vp.n_beams() == 32 && vp.n_time() == 64 && vp.n_time_out() == 16;
        std::printf("lifecycle: %s\n", ok ? "PASS" : "FAIL");
        fails += !ok;
        bool threw = false;
        try {
            cfg.n_time = 63;
            ggp::VoltagePipeline bad(cfg);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        std::printf("invalid config throws: %s\n", threw ? "PASS" : "FAIL");
        fails += !threw;
    }
    std::vector<uint8_t> qc((size_t)n_ch * 2 * n_ant * n_time);
    ftdo_gen_bytes(seed, (int64_t)qc.size(), qc.data());
    std::vector<float> w((size_t)n_ch * n_beam * n_ant * 2);
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 1;
    for (float& x : w) {  // uniform [-1, 1) from a 64-bit LCG
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        x = (float)((double)(s >> 11) * (2.0 / 9007199254740992.0) - 1.0);
    }
    const std::vector<double> ref = cpu_power(qc, w, n_ant, n_beam, n_ch, n_time, 4);
    fails += run(ggp::VoltageComputeMode::FP8, "FP8", n_ant, n_beam, n_ch, n_time, qc, w, ref);
    fails += run(ggp::VoltageComputeMode::FP6, "FP6", n_ant, n_beam, n_ch, n_time, qc, w, ref);
    std::printf(fails == 0 ? "ALL PASSED\n" : "FAILED\n");
    return fails == 0 ? 0 : 1;
}
