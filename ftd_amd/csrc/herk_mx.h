// herk_mx.h -- host-side interface of the MX block-scaled quantizer (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ftdx {

enum MxFormat { MX_FP4_E2M1 = 0, MX_FP6_E3M2 = 1 };
enum MxInput { MX_IN_INT4 = 0, MX_IN_FP16 = 1, MX_IN_FP16_PLANAR = 2 };
// Element conversion of the scaled values: the reference's converters (clamp, flush below
// 0.25, mantissa rounded without carry / E2M1 truncated) or OCP nearest-even with
// subnormals.  The E8M0 scale rule is the same for both.
enum MxRounding { MX_ROUND_REFERENCE = 0, MX_ROUND_NEAREST_EVEN = 1 };

// Bytes of one 32-code block (FP4: 16, FP6: 24).
int mx_code_bytes(int fmt);
// 32-sample blocks per row, padded to whole LDS stages.
int mx_padded_blocks(int K);

// X: [rows][K] QC bytes (MX_IN_INT4) or [rows][K][2] FP16 (MX_IN_FP16).
// codes: rows * nblkp * 2 * mx_code_bytes(fmt) bytes; scales: rows * nblkp * 2 bytes.
hipError_t launch_mx_quantize(const void* X, int input, int fmt, int rounding, uint8_t* codes, uint8_t* scales,
                              int64_t rows, int K, hipStream_t s);
// MX_IN_FP16_PLANAR: X = Re, X2 = Im ([rows][K] halves each).  conj: store the codes of
// conj(X) (Im sign bits flipped).
hipError_t launch_mx_quantize_ex(const void* X, const void* X2, int input, int fmt, int rounding, bool conj,
                                 uint8_t* codes, uint8_t* scales, int64_t rows, int K, hipStream_t s);

// The operand-row layout of the HERK / GEMM tile (LOAD_MX6 / LOAD_MX4): rows_out [rows]
// [nblkp][2][32] bytes, each 32-byte part = the 32 codes, the E8M0 scale byte, zeros; row
// pitch nblkp * 64 bytes (the tile's pre-cast geometry for nblkp * 32 samples).
hipError_t launch_mx_quantize_rows(const void* X, const void* X2, int input, int fmt, int rounding, bool conj,
                                   uint8_t* rows_out, int64_t rows, int K, hipStream_t s);

// Dense E2M1 operand rows (LOAD_MX4D): rows_out [rows][nblkp][2][16] code bytes (Re part,
// Im part), row pitch nblkp * 32 bytes; scales [rows / nrow][nblkp / 4][nrow][8], the
// (Re, Im) E8M0 bytes of each 4-block stage.  nblkp = mx_dense_blocks(K), a multiple of 4.
int mx_dense_blocks(int K);
hipError_t launch_mx_quantize_dense(const void* X, int input, int rounding, uint8_t* rows_out, uint8_t* scales,
                                    int64_t rows, int64_t nrow, int K, hipStream_t s);

}  // namespace ftdx
