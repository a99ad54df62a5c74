// test_herk_api.cpp -- end-to-end test of the C++ drop-in HERK API on the GPU,
// in the manner of the reference's cutlass_interface/test_herk_int4.cu:290-460:
// generate int4 voltages, call CutlassComplexGemm::herk (INT4 -> FP32 and FP16)
// and ggp::herkBatchedCutlassQC, compare with the CPU oracle.  The reference test
// allows max_err < max(1, 0.02K); this one requires bit-exact FP32 output.
//
// usage: test_herk_api N K batch [seed]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "cutlass_gemm_api.h"
#include "ggp_cutlass_interface.h"
#include "ggp_xengine.h"
#include "herk_oracle.h"

#define HIP_OK(x)                                                                       \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) {                                                         \
            std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_),      \
                         __FILE__, __LINE__);                                           \
            std::exit(2);                                                               \
        }                                                                               \
    } while (0)

using namespace cutlass_gemm_api;

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s N K batch [seed]\n", argv[0]);
        return 2;
    }
    const int N = std::atoi(argv[1]), K = std::atoi(argv[2]), batch = std::atoi(argv[3]);
    const uint64_t seed = argc > 4 ? std::strtoull(argv[4], nullptr, 10) : 1;
    const int64_t nA = (int64_t)N * K * batch;
    const int64_t nC = (int64_t)N * (N + 1) * batch;  // scalars
    std::printf("test_herk_api N=%d K=%d batch=%d seed=%llu\n", N, K, batch,
                (unsigned long long)seed);
    std::fflush(stdout);

    std::vector<uint8_t> hA(nA);
    ftdo_gen_qc_uniform(seed, nA, hA.data());
    std::vector<float> ref(nC, 0.0f);
    ftdo_herk_f32(hA.data(), ref.data(), N, K, batch, 1.0f, 0.0f, 0);
    std::vector<uint16_t> ref16(nC, 0);
    ftdo_herk_f16(hA.data(), ref16.data(), N, K, batch, 1.0f, 0.0f, 0);

    uint8_t* dA = nullptr;
    float* dC = nullptr;
    uint16_t* dC16 = nullptr;
    HIP_OK(hipMalloc(&dA, nA > 0 ? nA : 1));
    HIP_OK(hipMalloc(&dC, nC * sizeof(float)));
    HIP_OK(hipMalloc(&dC16, nC * sizeof(uint16_t)));
    HIP_OK(hipMemcpy(dA, hA.data(), nA, hipMemcpyHostToDevice));

    int failures = 0;
    CutlassComplexGemm api;
    api.set_strategy_tune_verbosity(0);

    auto compare32 = [&](const char* label) {
        std::vector<float> got(nC);
        HIP_OK(hipMemcpy(got.data(), dC, nC * sizeof(float), hipMemcpyDeviceToHost));
        int64_t bad = 0;
        for (int64_t i = 0; i < nC; ++i)
            if (std::memcmp(&got[i], &ref[i], 4) != 0) {
                if (bad < 5)
                    std::printf("    [%lld] got %.1f ref %.1f\n", (long long)i, got[i], ref[i]);
                ++bad;
            }
        std::printf("  %-40s %s (%lld mismatches)\n", label, bad ? "FAIL" : "PASS",
                    (long long)bad);
        failures += bad ? 1 : 0;
    };

    // Test 1: herk INT4 -> FP32 via the class
    HIP_OK(hipMemset(dC, 0xFF, nC * sizeof(float)));
    int rc = api.herk(dA, dC, N, K, batch, InputPrecision::INT4, ComputePrecision::FP8,
                      OutputPrecision::FP32, OutputFormat::PackedTriangle, 1.0f, 0.0f, nullptr,
                      false);
    HIP_OK(hipDeviceSynchronize());
    if (rc != 0) {
        std::printf("  herk returned %d\n", rc);
        ++failures;
    } else {
        compare32("CutlassComplexGemm::herk INT4->FP32");
    }

    // Test 2: herk INT4 -> FP16 via the class
    HIP_OK(hipMemset(dC16, 0xFF, nC * sizeof(uint16_t)));
    rc = api.herk(dA, dC16, N, K, batch, InputPrecision::INT4, ComputePrecision::FP8,
                  OutputPrecision::FP16, OutputFormat::PackedTriangle, 1.0f, 0.0f, nullptr, false);
    HIP_OK(hipDeviceSynchronize());
    {
        std::vector<uint16_t> got(nC);
        HIP_OK(hipMemcpy(got.data(), dC16, nC * 2, hipMemcpyDeviceToHost));
        int64_t bad = 0;
        for (int64_t i = 0; i < nC; ++i) bad += got[i] != ref16[i];
        std::printf("  %-40s %s (%lld mismatches)\n", "CutlassComplexGemm::herk INT4->FP16",
                    (rc || bad) ? "FAIL" : "PASS", (long long)bad);
        failures += (rc || bad) ? 1 : 0;
    }

    // Test 3: GGP glue
    ggp::initCutlassApi(0, nullptr, nullptr, false, 0);
    HIP_OK(hipMemset(dC, 0xFF, nC * sizeof(float)));
    rc = ggp::herkBatchedCutlassQC(dA, dC, N, K, batch, nullptr);
    HIP_OK(hipDeviceSynchronize());
    if (rc != 0) {
        std::printf("  herkBatchedCutlassQC returned %d\n", rc);
        ++failures;
    } else {
        compare32("ggp::herkBatchedCutlassQC");
    }

    // Test 4: unsupported combination throws, as the reference
    bool threw = false;
    try {
        api.herk(dA, dC, N, K, batch, InputPrecision::INT4, ComputePrecision::FP4,
                 OutputPrecision::FP32);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    std::printf("  %-40s %s\n", "FP4 + FP32 output throws", threw ? "PASS" : "FAIL");
    failures += threw ? 0 : 1;

    // Test 5: ggp::XEngine (TRI, FP32, host buffers) against the oracle's M^H M
    {
        ftdx_xengine_param xp;
        ftdx_xengine_param_init(&xp);
        xp.struct_size = sizeof(xp);
        xp.n_payload = 1;
        xp.n_antennae_per_payload = 48;
        xp.n_channels_per_payload = 2;
        xp.n_time_per_payload = 64;
        xp.n_time_inner = 2;
        xp.n_polarizations = 2;
        xp.compute_prec = xp.output_prec = FTDX_QPREC_SINGLE;
        xp.format = FTDX_XENGINE_MAT_TRI;
        const int xN = 48, xK = 32, xB = 8;
        std::vector<uint8_t> raw((size_t)xN * 64 * 2 * 2);
        ftdo_gen_bytes(seed + 1, (int64_t)raw.size(), raw.data());
        const size_t P = (size_t)xN * (xN + 1) / 2;
        std::vector<float> out(P * xB * 2, -1.0f);
        std::vector<int64_t> ref(P * xB * 2);
        ftdo_xengine_ggp_i64(raw.data(), xN, xK, xB, 0, ref.data(), 0);
        bool ok = true;
        try {
            ggp::XEngine xe(xp);
            xe.compute(out.data(), raw.data());
            for (size_t i = 0; i < out.size(); ++i) ok = ok && out[i] == (float)ref[i];
        } catch (const std::runtime_error& e) {
            std::printf("  XEngine threw: %s\n", e.what());
            ok = false;
        }
        std::printf("  %-40s %s\n", "ggp::XEngine TRI FP32 (host payload)", ok ? "PASS" : "FAIL");
        failures += ok ? 0 : 1;
        xp.compute_prec = FTDX_QPREC_HALF;  // rejected, as the reference constructor does
        bool threw2 = false;
        try {
            ggp::XEngine bad(xp);
        } catch (const std::runtime_error&) {
            threw2 = true;
        }
        std::printf("  %-40s %s\n", "XEngine HALF compute throws", threw2 ? "PASS" : "FAIL");
        failures += threw2 ? 0 : 1;
    }

    // Test 6: prepared-B complex GEMM with fused power (INT4 voltages) vs the oracle
    {
        const int gM = 64, gN = 96, gK = 80, gB = 2;
        std::vector<uint8_t> qc((size_t)gM * gK * gB);
        ftdo_gen_qc_uniform(seed + 7, (int64_t)qc.size(), qc.data());
        std::vector<uint16_t> bre((size_t)gN * gK * gB), bim(bre.size());
        std::vector<double> A((size_t)gM * gK * gB * 2), Bm((size_t)gN * gK * gB * 2);
        for (size_t i = 0; i < qc.size(); ++i) {
            int r, m;
            ftdo_decode_qc(qc[i], &r, &m);
            A[2 * i] = r;
            A[2 * i + 1] = m;
        }
        for (size_t i = 0; i < bre.size(); ++i) {
            const int r = (int)(i * 7 % 15) - 7, m = (int)(i * 11 % 15) - 7;
            bre[i] = ftdo_f32_to_f16_rne((float)r);
            bim[i] = ftdo_f32_to_f16_rne((float)m);
            Bm[2 * i] = r;
            Bm[2 * i + 1] = m;
        }
        std::vector<double> Cref((size_t)gM * gN * gB * 2);
        ftdo_gemm_c64(A.data(), Bm.data(), Cref.data(), gM, gN, gK, gB, 0);
        std::vector<float> re(Cref.size() / 2), im(re.size()), want(re.size()), got(re.size());
        for (size_t i = 0; i < re.size(); ++i) {
            re[i] = (float)Cref[2 * i];
            im[i] = (float)Cref[2 * i + 1];
        }
        ftdo_power_f32(re.data(), im.data(), re.data() /* unused: beta = 0 */, (int64_t)re.size(),
                       1.0f, 0.0f, want.data());
        uint8_t* dq;
        uint16_t *dbr, *dbi;
        float* dp;
        HIP_OK(hipMalloc(&dq, qc.size()));
        HIP_OK(hipMalloc(&dbr, bre.size() * 2));
        HIP_OK(hipMalloc(&dbi, bim.size() * 2));
        HIP_OK(hipMalloc(&dp, got.size() * 4));
        HIP_OK(hipMemcpy(dq, qc.data(), qc.size(), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dbr, bre.data(), bre.size() * 2, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dbi, bim.data(), bim.size() * 2, hipMemcpyHostToDevice));
        bool ok = true;
        try {
            api.prepare_b(dbr, dbi, gN, gK, gB, ComputePrecision::FP8);
            ok = api.gemm_prepared_power_int4(dq, dp, gM, gN, gK, gB) == 0;
            HIP_OK(hipDeviceSynchronize());
            HIP_OK(hipMemcpy(got.data(), dp, got.size() * 4, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < got.size(); ++i) ok = ok && std::memcmp(&got[i], &want[i], 4) == 0;
        } catch (const std::runtime_error& e) {
            std::printf("  GEMM threw: %s\n", e.what());
            ok = false;
        }
        std::printf("  %-40s %s\n", "prepare_b + gemm_prepared_power_int4", ok ? "PASS" : "FAIL");
        failures += ok ? 0 : 1;
        HIP_OK(hipFree(dq));
        HIP_OK(hipFree(dbr));
        HIP_OK(hipFree(dbi));
        HIP_OK(hipFree(dp));
    }

    {   // quda::stridedBatchGEMMCutlass, HERK case (lib/cutlass_interface.cu:316-369), with a
        // parameter struct of the reference's field names and enum values
        struct BlasParam {
            int trans_a = 2, trans_b = 0, m, n, k, batch_count, data_type = 2;
        } bp;
        const int hN = 40, hK = 12, hB = 2;  // |C| <= 98 * 12: the FP16 triangle stays exact
        bp.m = bp.n = hN;
        bp.k = hK;
        bp.batch_count = hB;
        std::vector<uint8_t> qc((size_t)hN * hK * hB);
        ftdo_gen_qc_uniform(seed + 7, (int64_t)qc.size(), qc.data());
        std::vector<float> x(qc.size() * 2);
        for (size_t i = 0; i < qc.size(); ++i) {
            int r, m;
            ftdo_decode_qc(qc[i], &r, &m);
            x[2 * i] = (float)r;
            x[2 * i + 1] = (float)m;
        }
        std::vector<float> want((size_t)hN * (hN + 1) * hB), got(want.size());
        ftdo_herk_f32(qc.data(), want.data(), hN, hK, hB, 1.0f, 0.0f, 0);
        float *dx, *dtri;
        HIP_OK(hipMalloc(&dx, x.size() * 4));
        HIP_OK(hipMalloc(&dtri, got.size() * 4));
        HIP_OK(hipMemcpy(dx, x.data(), x.size() * 4, hipMemcpyHostToDevice));
        const long long flops = quda::stridedBatchGEMMCutlass(dx, dx, dtri, bp, 2);
        HIP_OK(hipMemcpy(got.data(), dtri, got.size() * 4, hipMemcpyDeviceToHost));
        bool ok = flops == 8LL * hN * hN * hK * hB;
        for (size_t i = 0; i < got.size(); ++i) ok = ok && std::memcmp(&got[i], &want[i], 4) == 0;
        bool threw = false;
        try {
            bp.trans_a = 0;  // not the HERK pattern: the reference aborts
            (void)quda::stridedBatchGEMMCutlass(dx, dx, dtri, bp, 2);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        ok = ok && threw;
        std::printf("  %-40s %s\n", "quda::stridedBatchGEMMCutlass (HERK)", ok ? "PASS" : "FAIL");
        failures += ok ? 0 : 1;
        HIP_OK(hipFree(dx));
        HIP_OK(hipFree(dtri));
    }

    HIP_OK(hipFree(dA));
    HIP_OK(hipFree(dC));
    HIP_OK(hipFree(dC16));
    std::printf("%s\n", failures ? "FAILED" : "ALL PASSED");
    return failures ? 1 : 0;
}
