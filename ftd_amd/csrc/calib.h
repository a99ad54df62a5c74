// calib.h -- host-side interface of the gain-calibration and HERK-adjoint kernels (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ftdx {

// Triangle tiles (32x32) per channel and 32-blocks per dimension of the gain kernel.
int gain_tiles(int N);
int gain_blocks(int N);

// V, M: packed lower triangles [batch][N(N+1)/2] complex; g: [batch][N].
// Scratch: rowpart, colpart [batch][gain_blocks(N)][N] complex; losspart
// [batch][gain_tiles(N)] double.  Outputs: loss [batch] (double), grad [batch][N],
// Gamma (packed, may be null).
hipError_t launch_gain_loss_grad(const float2* V, const float2* M, const float2* g, int N, int batch,
                                 float2* Gamma, float2* rowpart, float2* colpart, double* losspart,
                                 double* loss, float2* grad, hipStream_t s);

// Bytes of scratch launch_gain_loss_grad needs: rowpart, colpart, then losspart, each
// starting at a multiple of 8 bytes.
size_t gain_scratch_bytes(int N, int batch);

// Size of the fixed device array behind the allocation-free gain entry point, and its
// device address on the current device (looked up once per device, then kept).
constexpr size_t kGainStaticScratchBytes = (size_t)16 << 20;
hipError_t gain_static_scratch(void** out);

// dX [batch][N][K] complex = (Gamma + Gamma^H) X with X QC bytes [batch][N][K];
// scratch holds adjoint_scratch_bytes(N, K, batch) bytes (16-byte aligned): the bf16
// split planes of H and the transposed bf16 voltages.
size_t adjoint_scratch_bytes(int N, int K, int batch);
hipError_t launch_herk_adjoint(const float2* Gamma, const uint8_t* X, void* scratch, float2* dX, int N,
                               int K, int batch, hipStream_t s);

}  // namespace ftdx
