// herk_common.h -- device helpers shared by the gfx950 HERK kernels.
//
// Data conventions (cutlass_gemm_api side of the reference):
//   * input: one "QC" byte per complex sample, high nibble = Re, low nibble = Im,
//     each sign-magnitude (bit 3 sign, bits 2:0 magnitude), layout [batch][N][K]
//     (cutlass_interface/cutlass_gemm_api.h:185-187, :727-728);
//   * output: packed lower triangle, tri(i,j) = i(i+1)/2 + j, interleaved [re, im],
//     batch stride N(N+1) scalars (cutlass_interface/shared/herk_kernel_common.hpp:353-424).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ftdx {

typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef float v4f __attribute__((ext_vector_type(4)));

// 4 QC bytes (4 complex samples) -> 8 OCP E4M3 bytes [re0, im0, re1, im1 | re2, im2, re3, im3].
// The magnitude table {0,1,..,7} -> {0x00,0x38,0x40,0x44,0x48,0x4A,0x4C,0x4E} is two
// dwords indexed by v_perm_b32; the sign bit of each nibble moves to bit 7 of its byte.
// A "-0" nibble becomes 0x80 (OCP E4M3 negative zero, a valid zero on gfx950).
__device__ __forceinline__ void qc4_to_fp8x8(uint32_t x, uint32_t& o0, uint32_t& o1) {
    const uint32_t TLO = 0x44403800u;  // fp8(0..3) in bytes 0..3
    const uint32_t THI = 0x4E4C4A48u;  // fp8(4..7) in bytes 4..7
    uint32_t m_im = x & 0x07070707u;
    uint32_t m_re = (x >> 4) & 0x07070707u;
    uint32_t f_im = __builtin_amdgcn_perm(THI, TLO, m_im);
    uint32_t f_re = __builtin_amdgcn_perm(THI, TLO, m_re);
    f_im |= (x << 4) & 0x80808080u;
    f_re |= x & 0x80808080u;
    // bytes 0..3 of the 8-byte pool are f_re, bytes 4..7 are f_im
    o0 = __builtin_amdgcn_perm(f_im, f_re, 0x05010400u);
    o1 = __builtin_amdgcn_perm(f_im, f_re, 0x07030602u);
}

// [re, im, re', im'] (fp8) -> [im, -re, im', -re'].  Row vector v_i with
// v_i . u_j = Im(a_i conj(a_j)) when u_j is the plain interleaved row.
__device__ __forceinline__ uint32_t conj_swap_fp8(uint32_t v) {
    return __builtin_amdgcn_perm(v, v, 0x02030001u) ^ 0x80008000u;
}

__device__ __forceinline__ v8i conj_swap_fp8x32(v8i a) {
    v8i r;
#pragma unroll
    for (int i = 0; i < 8; ++i) r[i] = (int)conj_swap_fp8((uint32_t)a[i]);
    return r;
}

// D = A(32x64 fp8) * B(64x32 fp8) + C, FP32 accumulate (plain, unscaled form).
__device__ __forceinline__ v16f mfma_32x32x64_fp8(v8i a, v8i b, v16f c) {
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, 0, 0, 0, 0, 0, 0);
}

// Operand formats of the unscaled f8f6f4 MFMA (cbsz / blgp codes): OCP E4M3, and E3M2
// ("BF6": sign, 3 exponent bits, bias 3, 2 mantissa bits), 32 six-bit values in the low
// 192 bits of the 8-VGPR operand.  E3M2 holds every integer in [-8, 8] exactly, so integer
// QC products sum exactly in FP32 either way; the 6-bit format runs the MFMA at twice
// the FP8 rate.
// OPF_MX6 / OPF_MX4: MX block-scaled E3M2 / E2M1 operand rows.  Per 32-sample block a row
// holds two 32-byte parts, Re then Im: the part's 32 codes (24 / 16 bytes), its E8M0
// scale byte right after them, zeros to 32 bytes.  A lane's operand is one part, so its
// scale is dword opf_scale_dword of the operand itself.
enum OperandFormat { OPF_FP8 = 0, OPF_BF6 = 3, OPF_MX6 = 16 + 3, OPF_MX4 = 16 + 4, OPF_MX4D = 48 + 4 };
// OPF_MX4D: dense E2M1 rows -- per 32-sample block 16 code bytes of Re then 16 of Im, the
// E8M0 scales in a separate stage-major array -- on the direct-to-LDS path only.

constexpr bool opf_mx(int opf) { return opf >= 16; }
constexpr int opf_scale_dword(int opf) { return opf == OPF_MX6 ? 6 : 4; }

template <int OPF>
__device__ __forceinline__ v16f mfma_32x32x64(v8i a, v8i b, v16f c) {
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, OPF, OPF, 0, 0, 0, 0);
}

// Scaled form for the MX operand rows: each lane's 32 values times its own scale.
template <int OPF>
__device__ __forceinline__ v16f mfma_32x32x64_mx(v8i a, v8i b, v16f c) {
    constexpr int code = OPF & 15;
    constexpr int sd = opf_scale_dword(OPF);
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, code, code, 0, a[sd], 0, b[sd]);
}

// The partner operand of the Im product, read from the other part of the same block:
// lanes 0-31 (Re lanes) take the Im part as is, lanes 32-63 take the Re part negated
// (sign bit of every code: bit 5 of each 6-bit E3M2 code, bit 3 of each E2M1 nibble).
template <int OPF>
__device__ __forceinline__ v8i mx_partner(v8i x, bool negate) {
    if constexpr (OPF == OPF_MX6) {
        const uint32_t m0 = negate ? 0x20820820u : 0u, m1 = negate ? 0x08208208u : 0u,
                       m2 = negate ? 0x82082082u : 0u;
#pragma unroll
        for (int d = 0; d < 6; ++d) x[d] ^= (int)(d % 3 == 0 ? m0 : d % 3 == 1 ? m1 : m2);
    } else {
        const uint32_t m = negate ? 0x88888888u : 0u;
#pragma unroll
        for (int d = 0; d < 4; ++d) x[d] ^= (int)m;
    }
    return x;
}

// logical workgroup id for hardware id w when consecutive ids go round-robin over
// 8 XCDs: XCD x gets one contiguous range of logical ids.
__device__ __forceinline__ int xcd_logical(int w, int total) {
    const int x = w & 7, local = w >> 3;
    const int q = total >> 3, r = total & 7;
    return x * q + (x < r ? x : r) + local;
}

// lower-triangle tile index t -> (bi, bj), bj <= bi, t = bi(bi+1)/2 + bj
__device__ __forceinline__ void tri_decode(int t, int& bi, int& bj) {
    int r = (int)((__builtin_sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
    while (r * (r + 1) / 2 > t) --r;
    while ((r + 1) * (r + 2) / 2 <= t) ++r;
    bi = r;
    bj = t - r * (r + 1) / 2;
}

__device__ __forceinline__ int64_t tri_index(int64_t i, int64_t j) { return i * (i + 1) / 2 + j; }

}  // namespace ftdx
