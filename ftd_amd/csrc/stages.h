// stages.h -- host-side interface of the stand-alone stage kernels (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ftdx {

// Samples per pre-cast operand row (K rounded up to 16); row pitch = 2 * that bytes.
int precast_row_samples(int K);

// QC rows [rows][K] (any alignment) -> zero-padded rows [rows][Kp] (Kp >= K, a multiple of
// 16; out 16-byte aligned).
hipError_t launch_pad_qc_rows(const uint8_t* in, uint8_t* out, int64_t rows, int K, int Kp, hipStream_t s);

// One row block of a multi-block pad launch: src [rows][K] -> dst [rows][Kp] (Kp >= K, a
// multiple of 16; dst 16-byte aligned).
struct PadRowsJob {
    const uint8_t* src;
    uint8_t* dst;
    int64_t rows;
    int K, Kp;
};
// One workgroup of that launch: kPadBlockSegments 16-byte output segments of job `job`
// from segment `seg` (fewer at the job's end).  A job of rows * Kp / 16 segments takes
// ceil(that / kPadBlockSegments) workgroups.
constexpr int kPadBlockSegments = 2048;
struct PadBlock {
    int job;
    int64_t seg;
};
// jobs, blocks: device arrays; one workgroup per entry of blocks (nblocks < 2^31).
hipError_t launch_pad_qc_rows_multi(const PadRowsJob* jobs, const PadBlock* blocks, int nblocks, hipStream_t s);
hipError_t launch_promote_qc_fp8_interleaved(const uint8_t* in, uint8_t* out, int64_t n,
                                             hipStream_t s);
hipError_t launch_promote_qc_fp8_planar(const uint8_t* in, uint8_t* out, int64_t rows, int K,
                                        hipStream_t s);
// QC rows -> E3M2 ("BF6") operand rows for LOAD_BF6: 32 bytes per 16 samples, pitch 2*Kp.
hipError_t launch_promote_qc_bf6_planar(const uint8_t* in, uint8_t* out, int64_t rows, int K,
                                        hipStream_t s);
hipError_t launch_promote_fp16_fp8_planar(const uint16_t* in, uint8_t* out, int64_t rows, int K,
                                          hipStream_t s);
hipError_t launch_promote_qc_fp16_planar(const uint8_t* in, uint16_t* re, uint16_t* im, int64_t n,
                                         hipStream_t s);
hipError_t launch_promote_data(const uint8_t* in, void* out, int64_t n, bool fp64, hipStream_t s);
hipError_t launch_pack_triangle(const float2* full, float2* packed, const float2* old, int N,
                                int batch, float beta, hipStream_t s);
hipError_t launch_tri_from_herm(const void* full, void* tri, int N, int batch, bool fp64,
                                hipStream_t s);
hipError_t launch_promote_qc_fp32(const uint8_t* in, float2* out, int64_t n, hipStream_t s);
hipError_t launch_convert_qc_to_ftd(const uint8_t* in, uint8_t* out, int64_t n, hipStream_t s);
// GGP [batch][K][N] two's-complement bytes -> pre-cast rows [batch][N][Kp] (pitch 2*Kp B).
hipError_t launch_ggp_to_fp8_rows(const uint8_t* in, uint8_t* out, int N, int K, int batch,
                                  hipStream_t s);
// FTD raw [time][ant][chan][pol] bytes (time = k*n_time_inner + ti) -> pre-cast rows
// [batch][ant][Kp], batch = (chan*n_pol + pol)*n_time_inner + ti, K = n_time / n_time_inner.
// twos: two's complement nibbles, low = Re; else QC sign-magnitude, high = Re.
hipError_t launch_ftd_to_fp8_rows(const uint8_t* in, uint8_t* out, int n_ant, int n_chan, int n_time,
                                  int n_time_inner, int n_pol, bool twos, hipStream_t s);
// Planar FP16 [rows][K] Re and Im -> pre-cast rows; conj: negate Im (FP8 sign bit).
hipError_t launch_promote_fp16_planar_rows(const uint16_t* re, const uint16_t* im, uint8_t* out,
                                           int64_t rows, int K, bool conj, hipStream_t s);
// FP8 E4M3 [rows][K][2] interleaved pairs -> pre-cast rows.
hipError_t launch_promote_fp8_interleaved_rows(const uint8_t* in, uint8_t* out, int64_t rows, int K,
                                               hipStream_t s);
hipError_t launch_interleave_tri(const uint16_t* re, const uint16_t* im, void* out, int64_t n,
                                 bool fp64, hipStream_t s);

// FP32 -> FP16 nearest-even and FP16 -> FP32, n scalars (the BLAS glue's narrow / widen,
// lib/cutlass_interface.cu:63-96).
hipError_t launch_narrow_fp32(const float* in, uint16_t* out, int64_t n, hipStream_t s);
hipError_t launch_widen_fp16(const uint16_t* in, float* out, int64_t n, hipStream_t s);
// FP32 [re, im] -> FP16 Re / Im planes, Im negated when asked (lib/cutlass_interface.cu:98-113);
// FP32 Re / Im planes -> [re, im] (:116-130).  n complex values.
hipError_t launch_deinterleave_narrow(const float* in, uint16_t* re, uint16_t* im, int64_t n, bool negate_im,
                                      hipStream_t s);
hipError_t launch_interleave_fp32(const float* re, const float* im, float* out, int64_t n, hipStream_t s);

}  // namespace ftdx
