/*
 * herk_oracle.h -- CPU oracle for the X-engine HERK (test infrastructure only).
 *
 * This library checks results: only tests, the smoke check and the CPU-baseline
 * leg of bench.py load it.  The product path (libftdx.so) never links or calls it.
 *
 * Conventions restated from the reference (cutlass_gemm_api side):
 *   - INT4 "QC" byte per complex sample: high nibble = Re, low nibble = Im,
 *     each nibble sign-magnitude (bit 3 = sign, bits 2:0 = |v|), range [-7, 7],
 *     -0 legal.  (cutlass_interface/test_herk_int4.cu:36-45,
 *     tests/visibility_pipeline_test.cpp:42-53)
 *   - Input layout [batch][N][K], K fastest.
 *   - C[i][j] = sum_k A[i][k] * conj(A[j][k]) for j <= i; Im(C[i][i]) forced 0;
 *     packed lower triangle, tri = i(i+1)/2 + j, interleaved [re, im];
 *     batch stride N(N+1) scalars.  (cutlass_interface/test_herk_int4.cu:51-70,
 *     cutlass_interface/shared/herk_kernel_common.hpp:353-424)
 *   - Epilogue: v = acc; if diag && imag: v = 0; v *= alpha;
 *     if beta != 0: v = fma(beta, old, v).  (herk_kernel_common.hpp:405-420)
 */
#ifndef FTD_HERK_ORACLE_H
#define FTD_HERK_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- portable generator: mt19937_64 (64-bit Mersenne twister) ---- */
typedef struct { uint64_t mt[312]; int idx; } ftdo_mt64;
void     ftdo_mt64_seed(ftdo_mt64* s, uint64_t seed);
uint64_t ftdo_mt64_next(ftdo_mt64* s);
/* n-th (1-based) output of mt19937_64 seeded with `seed` */
uint64_t ftdo_mt64_nth(uint64_t seed, uint64_t n);

/* ---- QC sign-magnitude codec ---- */
uint8_t ftdo_encode_qc(int re, int im);
void    ftdo_decode_qc(uint8_t b, int* re, int* im);
/* QC byte -> [fp8(Re), fp8(Im)] as (Re byte) | (Im byte << 8), OCP E4M3 */
uint16_t ftdo_qc_to_fp8_pair(uint8_t b);
/* OCP E4M3 byte -> float (independent decoder, all 256 codes) */
float   ftdo_fp8e4m3_to_float(uint8_t b);

/* Uniform [-7,7] complex int4 samples, encoded QC, from 4-bit fields of the
 * mt19937_64 stream (fields LSB first, value 15 rejected, v = u - 7; Re then Im). */
void ftdo_gen_qc_uniform(uint64_t seed, int64_t n, uint8_t* out);
/* Raw random bytes (includes -0 nibbles): 8 bytes per word, LSB first. */
void ftdo_gen_bytes(uint64_t seed, int64_t n, uint8_t* out);

/* ---- HERK oracles (plain path, cutlass_gemm_api convention) ---- */
/* C: [batch][N(N+1)/2][2] int64 (re, im), Im(diag) = 0. threads<=0: OpenMP default */
void ftdo_herk_int64(const uint8_t* A, int64_t* C, int N, int K, int batch, int threads);
/* Same, rows [row0, row1) of each batch only (bounded CPU-baseline sample). */
void ftdo_herk_int64_rows(const uint8_t* A, int64_t* C, int N, int K, int batch,
                          int row0, int row1, int threads);
/* FP32 output with the kernel epilogue (alpha, beta; C read when beta != 0). */
void ftdo_herk_f32(const uint8_t* A, float* C, int N, int K, int batch,
                   float alpha, float beta, int threads);
/* FP16 output (IEEE binary16 bits, round-to-nearest-even from the FP32 value). */
void ftdo_herk_f16(const uint8_t* A, uint16_t* C, int N, int K, int batch,
                   float alpha, float beta, int threads);

uint16_t ftdo_f32_to_f16_rne(float f);
float    ftdo_f16_to_f32(uint16_t h);

/* ---- MX block-scaled quantization (FP6 E3M2 / FP4 E2M1, 32-sample blocks) ----
 * fmt: 0 = FP4 E2M1 (max 6), 1 = FP6 E3M2 (max 28), 2 = FP6 E2M3 (max 7.5).
 * Scale per (row, 32-sample block, Re|Im part): E8M0 biased exponent
 * e = clamp(ceil(log2(amax / max)) + 127, 0, 254), 0 when amax == 0
 * (reference rule: cutlass_interface/gemm_complex_sm100/mxfp_kernels.hpp:147-166).
 * Elements: x * 2^(127-e) rounded to the format, nearest-even, with subnormals,
 * saturating at +-max (OCP Microscaling Formats v1.0 conversion).
 * Codes are returned one per byte (sign in the format's top bit). */
uint8_t ftdo_mx_encode(float x, int fmt);
/* The reference's own converters (fp6_kernels.hpp:27-88, fp4_kernels.hpp:7-30):
 * clamp to +-max (NaN -> -max), |x| below 0.25 (E2M3: 0.0625) -> signed zero,
 * exponent clamped to the normal range, mantissa rint(frac * 2^mbits) clamped
 * to 2^mbits - 1 without carry (E2M1: one bit, frac >= 0.5). */
uint8_t ftdo_mx_encode_ref(float x, int fmt);
double  ftdo_mx_decode(uint8_t code, int fmt);
uint8_t ftdo_mx_scale_exponent(float amax, int fmt);
/* X: complex samples as float pairs [rows][K][2].  codes: [rows][nblk][2][32],
 * scales: [rows][nblk][2] with nblk = ceil(K/32) (tail samples quantize as 0). */
void ftdo_mx_quantize(const float* X, int rows, int K, int fmt, uint8_t* codes, uint8_t* scales);
/* rounding: 0 = the reference's converters (ftdo_mx_encode_ref), 1 = nearest even
 * with subnormals (ftdo_mx_encode, what ftdo_mx_quantize uses). */
void ftdo_mx_quantize_mode(const float* X, int rows, int K, int fmt, int rounding, uint8_t* codes,
                           uint8_t* scales);
/* FP64 HERK of complex double input [batch][N][K][2], packed lower triangle
 * [batch][N(N+1)/2][2], Im(diag) = 0. */
void ftdo_herk_c64(const double* X, double* C, int N, int K, int batch, int threads);

/* ---- gain calibration (FP64) ----
 * V, M: packed lower triangles [batch][N(N+1)/2][2]; g: [batch][N][2].
 * R_ij = V_ij - g_i conj(g_j) M_ij (i >= j), loss_b = sum |R_ij|^2,
 * grad[b][k] = dL/dRe g_k + i dL/dIm g_k, Gamma_ij = dL/dRe M_ij + i dL/dIm M_ij. */
void ftdo_gain_loss_grad_f64(const double* V, const double* M, const double* g, int N, int batch,
                             double* loss, double* grad, double* Gamma);
/* dX = (Gamma + Gamma^H) X; Gamma packed lower; X, dX: [batch][N][K][2]. */
void ftdo_herk_adjoint_f64(const double* Gamma, const double* X, double* dX, int N, int K,
                           int batch, int threads);

/* ---- stage converters ---- */
/* GGP (4,4)-bit two's complement, low nibble = Re -> doubles [re, im]
 * (include/kernels/promote_data.cuh:80-81). */
void ftdo_promote_data(const uint8_t* in, int64_t n, double* out);
/* QC sign-magnitude -> FTD two's-complement byte (low = Re, high = Im)
 * (tests/visibility_pipeline_test.cpp:105-117). */
uint8_t ftdo_convert_qc_to_ftd(uint8_t b);
/* QC nibble -> FP16 bits (cutlass_gemm_api.cu:1548-1551 table: -0 -> 0x8000). */
uint16_t ftdo_qc_nibble_fp16(unsigned nibble);
/* float -> OCP E4M3 after clamping to +-448, nearest-even
 * (cast_kernels_common.hpp:30-48: clamp, then __nv_fp8_e4m3). */
uint8_t ftdo_f32_to_e4m3_clamped(float x);
/* full [batch][N][N][2] -> packed upper row-major (include/kernels/tri_from_herm.cuh:28-73). */
void ftdo_tri_from_herm(const double* full, double* tri, int N, int batch);

/* GGP XEngine M^H M per batch ([batch][K][N] two's-complement input); herm = 0:
 * upper triangle row-major [batch][N(N+1)/2][2], herm = 1: full [batch][N][N][2]. */
void ftdo_xengine_ggp_i64(const uint8_t* in, int N, int K, int batch, int herm, int64_t* out,
                          int threads);
/* FTD raw-order HERK, packed lower triangle [n_chan*n_pol*n_ti][n_ant(n_ant+1)/2][2]. */
void ftdo_herk_ftd_i64(const uint8_t* raw, int n_ant, int n_chan, int n_time, int n_ti, int n_pol,
                       int twos, int64_t* out, int threads);

/* Complex GEMM sum_k A[i][k] B[j][k] (cpu_gemm), interleaved FP64. */
void ftdo_gemm_c64(const double* A, const double* B, double* C, int M, int N, int K, int batch,
                   int threads);
/* The GEMM epilogue's power detection in float arithmetic. */
/* Voltage beamformer tail (power, polarization sum, time integration, quantization). */
void ftdo_voltage_tail(const float* re, const float* im, int64_t batch, int M, int n_beam, int nt, int mode,
                       float scale, int out_u8, void* out);
void ftdo_power_f32(const float* re, const float* im, const float* old, int64_t n, float alpha,
                    float beta, float* out);

/* Pillbox UV gridding of Nf packed FP32 triangles into an FP32 grid (accumulates). */
void ftdo_pillbox_grid(const float* vis, const float* uv_m, const double* freq_hz, float* grid,
                       int64_t n_bl, int Ng, int Nf, int freq_offset, float cell);

#ifdef __cplusplus
}
#endif
#endif
