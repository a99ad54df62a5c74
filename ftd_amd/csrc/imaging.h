// imaging.h -- host-side interface of the visibility-consumer kernels (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ftdx {

// [n_ch][n_pol][n_ti] packed triangles of N -> [n_ch][n_ti], summed over pol (FP32 or FP16).
hipError_t launch_pol_reduce(const void* in, void* out, int N, int n_ch, int n_pol, int n_ti, bool fp16,
                             hipStream_t s);
// Pillbox scatter of Nf packed triangles (n_bl baselines each, FP32 or FP16 [re, im])
// into an FP32 [Nf][Ng][Ng][2] grid (accumulates; the caller zeroes it).
hipError_t launch_pillbox_grid(const void* vis, bool fp16, const float* uv_m, const double* freq_hz,
                               float* grid, int64_t n_bl, int Ng, int Nf, int freq_offset, float cell,
                               hipStream_t s);

}  // namespace ftdx
