// voltage.h -- stage kernels of the voltage beamformer (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ftdx {

constexpr int kVoltageMaxPayloads = 16;

// QC voltage payloads, each [n_ch][2][n_ant][n_time] bytes (time fastest), seen as
// batch b = 0 .. n_ch / ch_fuse - 1 of M_total = n_time * ch_fuse * n_payloads
// samples m = (payload * ch_fuse + c) * n_time + t from channel b * ch_fuse + c
// (lib/voltage_pipeline.cu:176-236).
struct VoltageLayout {
    const uint8_t* payload[kVoltageMaxPayloads];
    int n_payloads, n_ant, n_time, n_ch, ch_fuse;
};

// -> pre-cast FP8 operand rows for the beam GEMM: per batch 128 * ceil(M_total / 64)
// rows of pitch 2 * Kp bytes (Kp = n_ant rounded up to 16), row block c holding
// samples 64c .. 64c + 63 of pol 0 then the same samples of pol 1; rows past M_total
// and samples past n_ant are zero.
hipError_t launch_voltage_beam_rows(const VoltageLayout& L, uint8_t* rows, hipStream_t s);
// -> QC bytes [batch][2 * M_total][n_ant]: pol 0's samples then pol 1's.
hipError_t launch_voltage_qc_rows(const VoltageLayout& L, uint8_t* rows, hipStream_t s);

// P [batch][n_beam][M] FP32 -> out [batch][n_beam][M / nt]: acc = sum of nt
// consecutive values in order; uint8 rint(acc * scale) clamped to [0, 255] or FP32.
hipError_t launch_beam_integrate(const float* P, void* out, bool out_u8, int64_t rows, int M, int nt,
                                 float scale, hipStream_t s);

// Planar beams C_re / C_im [n_ch][2 * n_time][n_beam] (pol 0 rows, then pol 1) ->
// out [n_ch][n_beam][n_time / nt], power += |pol 0|^2 + |pol 1|^2 per sample
// (lib/voltage_pipeline.cu:327-370, 419-456), uint8 or FP32.
hipError_t launch_power_sum(const float* re, const float* im, void* out, bool out_u8, int n_ch,
                            int n_time, int n_beam, int nt, float scale, hipStream_t s);

// [Nf][Nb][Nt] -> [Nb][Nf][Nt] (lib/corner_turn.cu:23-75).
hipError_t launch_corner_turn(float* out, const float* in, int Nf, int Nb, int Nt, hipStream_t s);

// out[i] = rint(in[i] * scale) clamped to [0, 255] (lib/voltage_pipeline.cu:464-478).
hipError_t launch_quantize_u8(const float* in, uint8_t* out, float scale, int64_t n, hipStream_t s);

// FP32 [re, im] pairs -> FP16 Re / Im planes, nearest-even (lib/voltage_pipeline.cu:484-498).
hipError_t launch_deinterleave_fp16(const float* in, uint16_t* re, uint16_t* im, int64_t n,
                                    hipStream_t s);

}  // namespace ftdx
