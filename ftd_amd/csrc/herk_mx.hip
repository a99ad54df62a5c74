// herk_mx.hip -- MX block-scaled quantization on gfx950 (FP4 E2M1 / FP6 E3M2 codes,
// E8M0 scale per 32 samples) for the block-scaled HERK and complex GEMM.
//
// Reference behaviour restated: the block-scaled HERK of
// cutlass_interface/gemm_complex_sm100/herk_impl.hpp:1067-1340 quantizes Re and
// Im separately with one E8M0 scale per (row, 32-sample block), scale
// 2^ceil(log2(amax / narrow_max)) (mxfp_kernels.hpp:147-166), then runs three
// block-scaled sub-GEMMs.  Here one scaled MFMA family does the whole product:
//
//   quantize:  X (INT4 QC bytes, FP16 interleaved or FP16 planar complex) -> codes +
//              scales, either split (codes [row][blk][part][PB], scales [row][blk][part],
//              the ftdx_mx_quantize layout) or as operand rows for the tile kernel
//              (per block and part 32 bytes: codes, scale byte, zeros).  blk is padded
//              to an even count (one 64-sample tile stage = 2 blocks).
//   HERK / GEMM: the tile runner of herk_tile.h (LOAD_MX6 / LOAD_MX4), per 32-sample
//              block one v_mfma_scale_f32_32x32x64_f8f6f4 step; lanes 0-31 carry the Re
//              part, lanes 32-63 the Im part of a row:
//                 Re C_ij = sum (re_i re_j + im_i im_j)     A = [re | im],  B = [re | im]
//                 Im C_ij = sum (im_i re_j - re_i im_j)     A = [im | -re], B = [re | im]
//              each lane's 32 values share one scale, which the MFMA applies.
// Element conversion of the values times the block's inverse scale, two rules:
//   MX_ROUND_REFERENCE     the reference's converters (fp6_kernels.hpp:27-58 for E3M2,
//                          fp4_kernels.hpp:7-30 for E2M1): clamp to +-max, magnitudes
//                          below 0.25 to a zero with the sign of x < 0, exponent clamped
//                          to the normal range, E3M2 mantissa rounded to nearest even and
//                          clamped to 3 (no carry into the exponent), E2M1 mantissa bit
//                          truncated; computed here on the float's bit fields.
//   MX_ROUND_NEAREST_EVEN  nearest-even with subnormals, saturating at the format maximum
//                          (OCP Microscaling Formats v1.0), by the gfx950 converters
//                          (v_cvt_scalef32_pk_fp4_f32, v_cvt_scalef32_2xpk16_bf6_f32).
// Both are checked against the CPU restatement's codes.  Codes per value are packed LSB
// first (FP4: nibble i at bits 4i; FP6: bits 6i..6i+5 of the lane's dwords).
#include <hip/hip_runtime.h>

#include "herk_common.h"
#include "herk_mx.h"
#include "herk_tile.h"

namespace ftdx {

typedef int v6i __attribute__((ext_vector_type(6)));

namespace {

template <int FMT>
struct MxFmt;

template <>
struct MxFmt<MX_FP4_E2M1> {
    static constexpr int PB = 16;        // bytes per 32 codes
    static constexpr int DW = 4;         // dwords per lane operand
    static constexpr float MAXV = 6.0f;
    static __device__ __forceinline__ uint32_t neg(int d) { (void)d; return 0x88888888u; }
};

template <>
struct MxFmt<MX_FP6_E3M2> {
    static constexpr int PB = 24;
    static constexpr int DW = 6;
    static constexpr float MAXV = 28.0f;
    // sign bits at 6i+5 of the 192-bit lane operand, per dword
    static __device__ __forceinline__ uint32_t neg(int d) {
        return (d % 3 == 0) ? 0x20820820u : (d % 3 == 1) ? 0x08208208u : 0x82082082u;
    }
};

template <int FMT>
__device__ __forceinline__ uint32_t mx_scale_exponent(float amax) {
    if (amax == 0.0f) return 0u;
    const float ratio = amax / MxFmt<FMT>::MAXV;
    if (ratio == 0.0f) return 0u;  // denormal amax underflowing the ratio: smallest scale
    int e;
    const float m = __builtin_frexpf(ratio, &e);
    int p = (m == 0.5f) ? e - 1 : e;
    int b = p + 127;
    b = b < 0 ? 0 : (b > 254 ? 254 : b);
    return (uint32_t)b;
}

// The reference's element converter on a scaled value (fp6_kernels.hpp:27-58,
// fp4_kernels.hpp:7-30), from the float's bit fields: a >= 0.25 is a normal float whose
// exponent e gives the E3M2 biased exponent e + 3 (1..7 for a in [0.25, 28]) or the E2M1
// one max(e + 1, 1); the E3M2 mantissa is the top 23 - 21 bits rounded to nearest even
// (rint(4 * (frac - 1)) of the reference) and clamped to 3, the E2M1 bit is bit 22.
template <int FMT>
__device__ __forceinline__ uint32_t ref_element_code(float x) {
    constexpr float MAXV = MxFmt<FMT>::MAXV;
    constexpr int S = FMT == MX_FP4_E2M1 ? 3 : 5;
    float v = __builtin_fminf(__builtin_fmaxf(x, -MAXV), MAXV);
    if (x != x) v = -MAXV;  // fmaxf(NaN, -max) = -max in the reference
    const uint32_t sign = v < 0.0f ? 1u : 0u;
    const float a = __builtin_fabsf(v);
    if (a < 0.25f) return sign << S;
    const uint32_t bits = __float_as_uint(a);
    const int e = (int)(bits >> 23) - 127;
    const uint32_t man = bits & 0x7FFFFFu;
    if constexpr (FMT == MX_FP4_E2M1) {
        const int eb = e + 1 < 1 ? 1 : e + 1;
        return (sign << 3) | ((uint32_t)eb << 1) | (man >> 22);
    } else {
        uint32_t m = (man + 0xFFFFFu + ((man >> 21) & 1u)) >> 21;
        m = m > 3u ? 3u : m;
        return (sign << 5) | ((uint32_t)(e + 3) << 2) | m;
    }
}

// Sign-magnitude nibble -> value; the "-0" nibble gives +0 (integer negate first).
__device__ __forceinline__ float qc_nibble_value(uint32_t n) {
    const int m = (int)(n & 7u);
    return (float)((n & 8u) ? -m : m);
}

// One thread per (row, 32-sample block): reads 32 complex samples, writes the
// Re and Im code blocks and their two scale bytes.  Blocks past K (and the
// padding block) quantize as zeros with scale 0.
// MX_IN_FP16_PLANAR: X holds the Re halves, X2 the Im halves, both [rows][K].  conj:
// the Im codes are stored negated (sign bit flipped after encoding, which is exact
// since the encoding is symmetric), i.e. the codes of conj(X).
// ROWS: the operand-row layout of the HERK tile instead (codes = rows, scales unused):
// per (row, block) the Re part then the Im part, 32 bytes each: codes, scale byte, zeros.
// DENSE (E2M1, with ROWS): codes = rows [rows][nblkp][2][16] (the 16 code bytes of the Re
// part, then of the Im part), scales = [rows / nrow][nblkp / 4][nrow][8]: per stage of 4
// blocks, each row's (Re, Im) scale bytes of those blocks.
template <int FMT, int IN, int RND, bool ROWS = false, bool DENSE = false>
__global__ void __launch_bounds__(256) mx_quantize_kernel(const void* __restrict__ X,
                                                          const void* __restrict__ X2,
                                                          uint8_t* __restrict__ codes,
                                                          uint8_t* __restrict__ scales,
                                                          int64_t rows, int K, int nblkp, int conj, int vec,
                                                          int64_t nrow = 0) {
    using F = MxFmt<FMT>;
    const int64_t total = rows * nblkp;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float v[2][32];
    if (t >= total) return;
    const int64_t row = t / nblkp;
    const int blk = (int)(t - row * nblkp);
    if (vec && blk * 32 + 32 <= K) {
        // whole block, 16-byte aligned rows: wide loads of the block's contiguous bytes
        if constexpr (IN == MX_IN_INT4) {
            const uint4* src = reinterpret_cast<const uint4*>(static_cast<const uint8_t*>(X) + row * K + blk * 32);
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const uint4 u = src[q];
                const uint32_t w4[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const uint32_t b = (w4[i >> 2] >> (8 * (i & 3))) & 0xFFu;
                    v[0][q * 16 + i] = qc_nibble_value(b >> 4);
                    v[1][q * 16 + i] = qc_nibble_value(b & 15u);
                }
            }
        } else if constexpr (IN == MX_IN_FP16_PLANAR) {
#pragma unroll
            for (int part = 0; part < 2; ++part) {
                const uint4* src = reinterpret_cast<const uint4*>(
                    static_cast<const _Float16*>(part ? X2 : X) + row * K + blk * 32);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint4 u = src[q];
                    const uint32_t w4[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        const uint16_t hb = (uint16_t)(w4[i >> 1] >> (16 * (i & 1)));
                        _Float16 hv;
                        __builtin_memcpy(&hv, &hb, 2);
                        v[part][q * 8 + i] = (float)hv;
                    }
                }
            }
        } else {
            const uint4* src =
                reinterpret_cast<const uint4*>(static_cast<const _Float16*>(X) + (row * K + blk * 32) * 2);
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const uint4 u = src[q];
                const uint32_t w4[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint16_t lo = (uint16_t)(w4[i] & 0xFFFFu), hi = (uint16_t)(w4[i] >> 16);
                    _Float16 re, im;
                    __builtin_memcpy(&re, &lo, 2);
                    __builtin_memcpy(&im, &hi, 2);
                    v[0][q * 4 + i] = (float)re;
                    v[1][q * 4 + i] = (float)im;
                }
            }
        }
    } else {
#pragma unroll
    for (int i = 0; i < 32; ++i) {
        const int k = blk * 32 + i;
        float re = 0.0f, im = 0.0f;
        if (k < K) {
            if (IN == MX_IN_INT4) {
                const uint32_t b = static_cast<const uint8_t*>(X)[row * K + k];
                re = qc_nibble_value(b >> 4);
                im = qc_nibble_value(b & 15u);
            } else if (IN == MX_IN_FP16_PLANAR) {
                re = (float)static_cast<const _Float16*>(X)[row * K + k];
                im = (float)static_cast<const _Float16*>(X2)[row * K + k];
            } else {
                const _Float16* h = static_cast<const _Float16*>(X) + (row * K + k) * 2;
                re = (float)h[0];
                im = (float)h[1];
            }
        }
        v[0][i] = re;
        v[1][i] = im;
    }
    }
#pragma unroll
    for (int part = 0; part < 2; ++part) {
        float amax = 0.0f;
#pragma unroll
        for (int i = 0; i < 32; ++i) amax = __builtin_fmaxf(amax, __builtin_fabsf(v[part][i]));
        const uint32_t se = mx_scale_exponent<FMT>(amax);
        const float inv = __builtin_ldexpf(1.0f, 127 - (int)se);
        // v * inv is exact (a power of two within range); then either the reference's
        // converter per element, or the gfx950 converters, which round to nearest even
        // with subnormals and saturate (the OCP MX element rule)
        uint32_t w[6] = {0u, 0u, 0u, 0u, 0u, 0u};
        if constexpr (RND == MX_ROUND_REFERENCE) {
#pragma unroll
            for (int i = 0; i < 32; ++i) {
                const uint32_t c = ref_element_code<FMT>(v[part][i] * inv);
                if constexpr (FMT == MX_FP4_E2M1) {
                    w[i >> 3] |= c << (4 * (i & 7));
                } else {
                    const int p = 6 * i;
                    w[p >> 5] |= c << (p & 31);
                    if ((p & 31) > 26) w[(p >> 5) + 1] |= c >> (32 - (p & 31));
                }
            }
        } else if constexpr (FMT == MX_FP4_E2M1) {
#pragma unroll
            for (int i = 0; i < 32; i += 2) {
                const uint32_t b = __builtin_amdgcn_cvt_scalef32_pk_fp4_f32(0u, v[part][i] * inv, v[part][i + 1] * inv,
                                                                            1.0f, 0) & 0xFFu;
                w[i >> 3] |= b << (4 * (i & 7));
            }
        } else {
            // the 2 x 16 form interleaves its sources: code 2i from the first, 2i + 1 from
            // the second, code k at bits 6k .. 6k + 5
            v16f ev, od;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                ev[i] = v[part][2 * i] * inv;
                od[i] = v[part][2 * i + 1] * inv;
            }
            const v6i c = __builtin_amdgcn_cvt_scalef32_2xpk16_bf6_f32(ev, od, 1.0f);
#pragma unroll
            for (int d = 0; d < 6; ++d) w[d] = (uint32_t)c[d];
        }
        if constexpr (DENSE) {
            static_assert(FMT == MX_FP4_E2M1 && ROWS, "dense rows are E2M1 operand rows");
            uint32_t o[4];
#pragma unroll
            for (int d = 0; d < 4; ++d) o[d] = (part == 1 && conj) ? w[d] ^ F::neg(d) : w[d];
            *reinterpret_cast<uint4*>(codes + ((row * nblkp + blk) * 2 + part) * 16) = make_uint4(o[0], o[1], o[2], o[3]);
            const int64_t pb = row / nrow, n = row - pb * nrow;
            scales[((pb * (nblkp / 4) + blk / 4) * nrow + n) * 8 + (blk & 3) * 2 + part] = (uint8_t)se;
        } else if constexpr (ROWS) {
            uint32_t o[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
#pragma unroll
            for (int d = 0; d < F::DW; ++d) o[d] = (part == 1 && conj) ? w[d] ^ F::neg(d) : w[d];
            o[F::DW] = se;
            uint4* dst = reinterpret_cast<uint4*>(codes + ((row * nblkp + blk) * 2 + part) * 32);
            dst[0] = make_uint4(o[0], o[1], o[2], o[3]);
            dst[1] = make_uint4(o[4], o[5], o[6], o[7]);
        } else {
            uint32_t* dst = reinterpret_cast<uint32_t*>(codes + ((row * nblkp + blk) * 2 + part) * F::PB);
#pragma unroll
            for (int d = 0; d < F::DW; ++d) dst[d] = (part == 1 && conj) ? w[d] ^ F::neg(d) : w[d];
            scales[(row * nblkp + blk) * 2 + part] = (uint8_t)se;
        }
    }
}

constexpr int MX_BLK_PER_STAGE = 2;  // 32-sample blocks per 64-sample tile stage

template <int FMT, int RND, bool ROWS>
hipError_t quantize_fmt_rnd(const void* X, const void* X2, int input, bool conj, uint8_t* codes,
                            uint8_t* scales, int64_t rows, int K, int nblkp, hipStream_t s) {
    const int64_t total = rows * nblkp;
    const unsigned blocks = (unsigned)((total + 255) / 256);
    const int cj = conj ? 1 : 0;
    // whole blocks may be read 16 bytes at a time when every row starts 16-byte aligned
    const int esz = input == MX_IN_INT4 ? 1 : input == MX_IN_FP16_PLANAR ? 2 : 4;
    const int vec = ((int64_t)K * esz % 16 == 0 &&
                     ((reinterpret_cast<uintptr_t>(X) | reinterpret_cast<uintptr_t>(X2)) & 15u) == 0) ? 1 : 0;
    if (input == MX_IN_INT4)
        hipLaunchKernelGGL((mx_quantize_kernel<FMT, MX_IN_INT4, RND, ROWS>), dim3(blocks), dim3(256), 0, s, X, X2,
                           codes, scales, rows, K, nblkp, cj, vec);
    else if (input == MX_IN_FP16_PLANAR)
        hipLaunchKernelGGL((mx_quantize_kernel<FMT, MX_IN_FP16_PLANAR, RND, ROWS>), dim3(blocks), dim3(256), 0,
                           s, X, X2, codes, scales, rows, K, nblkp, cj, vec);
    else
        hipLaunchKernelGGL((mx_quantize_kernel<FMT, MX_IN_FP16, RND, ROWS>), dim3(blocks), dim3(256), 0, s, X, X2,
                           codes, scales, rows, K, nblkp, cj, vec);
    return hipGetLastError();
}

template <int FMT, bool ROWS = false>
hipError_t quantize_fmt(const void* X, const void* X2, int input, int rounding, bool conj, uint8_t* codes,
                        uint8_t* scales, int64_t rows, int K, int nblkp, hipStream_t s) {
    if (rounding == MX_ROUND_REFERENCE)
        return quantize_fmt_rnd<FMT, MX_ROUND_REFERENCE, ROWS>(X, X2, input, conj, codes, scales, rows, K, nblkp, s);
    if (rounding == MX_ROUND_NEAREST_EVEN)
        return quantize_fmt_rnd<FMT, MX_ROUND_NEAREST_EVEN, ROWS>(X, X2, input, conj, codes, scales, rows, K, nblkp,
                                                                  s);
    return hipErrorInvalidValue;
}

}  // namespace

int mx_code_bytes(int fmt) { return fmt == MX_FP4_E2M1 ? MxFmt<MX_FP4_E2M1>::PB : MxFmt<MX_FP6_E3M2>::PB; }

int mx_padded_blocks(int K) {
    const int nblk = (K + 31) / 32;
    return (nblk + MX_BLK_PER_STAGE - 1) / MX_BLK_PER_STAGE * MX_BLK_PER_STAGE;
}

hipError_t launch_mx_quantize_ex(const void* X, const void* X2, int input, int fmt, int rounding, bool conj,
                                 uint8_t* codes, uint8_t* scales, int64_t rows, int K, hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    const int nblkp = mx_padded_blocks(K);
    if (nblkp == 0) return hipSuccess;
    if (fmt == MX_FP4_E2M1)
        return quantize_fmt<MX_FP4_E2M1>(X, X2, input, rounding, conj, codes, scales, rows, K, nblkp, s);
    if (fmt == MX_FP6_E3M2)
        return quantize_fmt<MX_FP6_E3M2>(X, X2, input, rounding, conj, codes, scales, rows, K, nblkp, s);
    return hipErrorInvalidValue;
}

hipError_t launch_mx_quantize_rows(const void* X, const void* X2, int input, int fmt, int rounding, bool conj,
                                   uint8_t* rows_out, int64_t rows, int K, hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    const int nblkp = mx_padded_blocks(K);
    if (nblkp == 0) return hipSuccess;
    if (fmt == MX_FP4_E2M1)
        return quantize_fmt<MX_FP4_E2M1, true>(X, X2, input, rounding, conj, rows_out, nullptr, rows, K, nblkp, s);
    if (fmt == MX_FP6_E3M2)
        return quantize_fmt<MX_FP6_E3M2, true>(X, X2, input, rounding, conj, rows_out, nullptr, rows, K, nblkp, s);
    return hipErrorInvalidValue;
}

int mx_dense_blocks(int K) { return ((K + 31) / 32 + 3) / 4 * 4; }

template <int RND>
hipError_t quantize_dense_rnd(const void* X, int input, uint8_t* rows_out, uint8_t* scales, int64_t rows,
                              int64_t nrow, int K, int nblkp, hipStream_t s) {
    const int64_t total = rows * nblkp;
    const unsigned blocks = (unsigned)((total + 255) / 256);
    const int esz = input == MX_IN_INT4 ? 1 : 4;
    const int vec = ((int64_t)K * esz % 16 == 0 && (reinterpret_cast<uintptr_t>(X) & 15u) == 0) ? 1 : 0;
    if (input == MX_IN_INT4)
        hipLaunchKernelGGL((mx_quantize_kernel<MX_FP4_E2M1, MX_IN_INT4, RND, true, true>), dim3(blocks), dim3(256), 0,
                           s, X, nullptr, rows_out, scales, rows, K, nblkp, 0, vec, nrow);
    else if (input == MX_IN_FP16)
        hipLaunchKernelGGL((mx_quantize_kernel<MX_FP4_E2M1, MX_IN_FP16, RND, true, true>), dim3(blocks), dim3(256), 0,
                           s, X, nullptr, rows_out, scales, rows, K, nblkp, 0, vec, nrow);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_mx_quantize_dense(const void* X, int input, int rounding, uint8_t* rows_out, uint8_t* scales,
                                    int64_t rows, int64_t nrow, int K, hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    const int nblkp = mx_dense_blocks(K);
    if (nblkp == 0 || nrow <= 0 || rows % nrow != 0) return hipErrorInvalidValue;
    if (rounding == MX_ROUND_REFERENCE)
        return quantize_dense_rnd<MX_ROUND_REFERENCE>(X, input, rows_out, scales, rows, nrow, K, nblkp, s);
    if (rounding == MX_ROUND_NEAREST_EVEN)
        return quantize_dense_rnd<MX_ROUND_NEAREST_EVEN>(X, input, rows_out, scales, rows, nrow, K, nblkp, s);
    return hipErrorInvalidValue;
}

hipError_t launch_mx_quantize(const void* X, int input, int fmt, int rounding, uint8_t* codes, uint8_t* scales,
                              int64_t rows, int K, hipStream_t s) {
    return launch_mx_quantize_ex(X, nullptr, input, fmt, rounding, false, codes, scales, rows, K, s);
}

}  // namespace ftdx
