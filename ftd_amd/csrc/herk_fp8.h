// herk_fp8.h -- host-side launch interface of the plain FP8 HERK kernels (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ftdx {

struct HerkPlainArgs {
    const uint8_t* A;   // [batch][N][K] QC bytes
    void* C;            // [batch][N(N+1)] scalars (float or half), packed lower triangle
    int N, K, batch;
    int tiles;          // lower-triangle tiles per batch
    float alpha, beta;
    int64_t a_bstride;  // bytes of A per batch (N*K)
    int64_t c_bstride;  // scalars of C per batch (N*(N+1))
};

int herk_plain_tiles(int N);
hipError_t launch_herk_int4_plain(const HerkPlainArgs& a, bool out_fp16, hipStream_t stream);

}  // namespace ftdx
