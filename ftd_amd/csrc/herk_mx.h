// herk_mx.h -- host-side interface of the MX block-scaled HERK kernels (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ftdx {

enum MxFormat { MX_FP4_E2M1 = 0, MX_FP6_E3M2 = 1 };
enum MxInput { MX_IN_INT4 = 0, MX_IN_FP16 = 1, MX_IN_FP16_PLANAR = 2 };

// Bytes of one 32-code block (FP4: 16, FP6: 24).
int mx_code_bytes(int fmt);
// 32-sample blocks per row, padded to whole LDS stages.
int mx_padded_blocks(int K);

// X: [rows][K] QC bytes (MX_IN_INT4) or [rows][K][2] FP16 (MX_IN_FP16).
// codes: rows * nblkp * 2 * mx_code_bytes(fmt) bytes; scales: rows * nblkp * 2 bytes.
hipError_t launch_mx_quantize(const void* X, int input, int fmt, uint8_t* codes, uint8_t* scales,
                              int64_t rows, int K, hipStream_t s);
// MX_IN_FP16_PLANAR: X = Re, X2 = Im ([rows][K] halves each).  conj: store the codes of
// conj(X) (Im sign bits flipped).
hipError_t launch_mx_quantize_ex(const void* X, const void* X2, int input, int fmt, bool conj,
                                 uint8_t* codes, uint8_t* scales, int64_t rows, int K, hipStream_t s);

// C = alpha A B^T + beta C on MX codes of A ([batch][M] rows) and of conj(B) ([batch][N]
// rows), both quantized with the same K; planar outputs Cre / Cim [batch][M][ldc].
struct MxGemmArgs {
    const uint8_t* acodes;
    const uint8_t* ascales;
    const uint8_t* bcodes;
    const uint8_t* bscales;
    void* Cre;
    void* Cim;
    int M, N, batch, nblkp;
    float alpha, beta;
    int64_t ldc, c_bstride;
};

hipError_t launch_gemm_mx(const MxGemmArgs& a, int fmt, bool out_fp16, hipStream_t s);

struct MxHerkArgs {
    const uint8_t* codes;   // [batch][N][nblkp][2][PB]
    const uint8_t* scales;  // [batch][N][nblkp][2]
    void* C;
    int N, batch, nblkp;
    float alpha, beta;
    int64_t c_bstride;      // scalars of C per batch
};

hipError_t launch_herk_mx(const MxHerkArgs& a, int fmt, bool out_fp16, hipStream_t s);

}  // namespace ftdx
