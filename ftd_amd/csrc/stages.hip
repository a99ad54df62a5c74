// stages.hip -- stand-alone (unfused) stage kernels of the X-engine path on gfx950.
//
// The fused HERK kernel does all of these inside its load stage / epilogue; the
// stand-alone forms serve callers that use the stages directly (the reference's
// GGP element-wise kernels and CUTLASS-path casts), the unfused HerkMode path,
// and FP16 input.  All are HBM-bound: 16-byte vector loads/stores where the
// layout allows, one grid-stride pass.
//
//   promote_qc_fp8_interleaved   QC -> [fp8 Re, fp8 Im] bytes     (cutlass_gemm_api.cu:1566-1623)
//   promote_qc_fp8_planar        QC -> pre-cast operand rows (dword-planar FP8)
//   pad_qc_rows[_multi]          QC rows of any K and alignment -> zero-padded rows of K
//                                rounded up to 16 (the fused HERK's 16-byte loads); the
//                                multi form copies many row blocks in one launch
//   promote_fp16_fp8_planar      FP16 [re, im] -> pre-cast rows, clamp +-448 then
//                                nearest-even E4M3 (cast_kernels_common.hpp:30-48)
//   promote_qc_fp16_planar       QC -> FP16 Re plane + Im plane   (cutlass_gemm_api.cu:1639-1698)
//   promote_data                 GGP (4,4)-bit two's complement, low nibble = Re ->
//                                FP32/FP64 [re, im]               (include/kernels/promote_data.cuh:70-88)
//   pack_triangle                N x N [re, im] FP32 -> packed lower triangle (+ beta old)
//                                                                 (shared/pack_kernels_common.hpp:602-659)
//   tri_from_herm                N x N complex -> packed upper row-major  (include/kernels/tri_from_herm.cuh:28-73)
//   interleave_tri               FP16 Re/Im planes -> [re, im] FP32/FP64  (include/kernels/interleave_tri.cuh:27-42)
//   promote_fp16_planar_rows     FP16 Re / Im planes -> pre-cast rows, optionally conjugated
//                                (prepare_b's conj(B), cutlass_gemm_api.h:236-255)
//   promote_fp8_interleaved_rows FP8 [re, im] pairs -> pre-cast rows (gemm_prepared_power_fp8's A)
//   ggp_to_fp8_rows              GGP XEngine input [batch][K][N] (two's complement, low
//                                nibble = Re; lib/xengine.cpp:150-160, promote_data.cuh:80-81)
//                                -> pre-cast operand rows [batch][N][Kp], transposed
//   ftd_to_fp8_rows              FTD raw order [time][ant][chan][pol] (QC or two's
//                                complement) -> pre-cast rows [batch][ant][Kp] with
//                                batch = (chan*n_pol + pol)*n_ti + ti (lib/tcc_interface.cu:124-178,
//                                visibility_pipeline_test.cpp:54-79)
#include <hip/hip_runtime.h>

#include "herk_common.h"
#include "herk_tile.h"
#include "stages.h"

namespace ftdx {

namespace {

constexpr int ST_THREADS = 256;

unsigned grid_for(int64_t work_items) {
    int64_t g = (work_items + ST_THREADS - 1) / ST_THREADS;
    if (g > 65536) g = 65536;  // grid-stride beyond this
    return (unsigned)(g > 0 ? g : 1);
}

// 4 QC bytes -> [re0, im0, re1, im1 | re2, im2, re3, im3] FP8 bytes (reference layout)
__device__ __forceinline__ void qc4_to_fp8_interleaved(uint32_t x, uint32_t& o0, uint32_t& o1) {
    uint32_t re, im;
    qc4_to_fp8_planar(x, re, im);
    o0 = __builtin_amdgcn_perm(im, re, 0x05010400u);
    o1 = __builtin_amdgcn_perm(im, re, 0x07030602u);
}

__global__ void __launch_bounds__(ST_THREADS)
promote_qc_fp8_interleaved_kernel(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                  int64_t n) {
    const int64_t n16 = n / 16;
    const int64_t stride = (int64_t)gridDim.x * ST_THREADS;
    for (int64_t t = (int64_t)blockIdx.x * ST_THREADS + threadIdx.x; t < n16; t += stride) {
        const uint4 v = reinterpret_cast<const uint4*>(in)[t];
        uint4 a, b;
        qc4_to_fp8_interleaved(v.x, a.x, a.y);
        qc4_to_fp8_interleaved(v.y, a.z, a.w);
        qc4_to_fp8_interleaved(v.z, b.x, b.y);
        qc4_to_fp8_interleaved(v.w, b.z, b.w);
        reinterpret_cast<uint4*>(out)[2 * t] = a;
        reinterpret_cast<uint4*>(out)[2 * t + 1] = b;
    }
    // tail (< 16 samples), one thread per sample
    const int64_t t = (int64_t)blockIdx.x * ST_THREADS + threadIdx.x;
    if (t < n - n16 * 16) {
        const int64_t i = n16 * 16 + t;
        uint32_t re, im;
        qc4_to_fp8_planar(in[i], re, im);
        out[2 * i] = (uint8_t)(re & 0xFF);
        out[2 * i + 1] = (uint8_t)(im & 0xFF);
    }
}

// QC rows [rows][K] -> pre-cast rows [rows][Kp] (Kp = K rounded up to 16 samples),
// 16 samples (32 output bytes) per thread.
template <bool VEC>
__global__ void __launch_bounds__(ST_THREADS)
promote_qc_fp8_planar_kernel(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                             int64_t rows, int K, int Kp) {
    const int segs = Kp / 16;
    const int64_t total = rows * segs;
    const int64_t stride = (int64_t)gridDim.x * ST_THREADS;
    for (int64_t t = (int64_t)blockIdx.x * ST_THREADS + threadIdx.x; t < total; t += stride) {
        const int64_t r = t / segs;
        const int s = (int)(t - r * segs);
        const int k0 = s * 16;
        uint32_t w[4] = {0u, 0u, 0u, 0u};
        const uint8_t* src = in + r * K + k0;
        if (VEC) {
            const uint4 v = *reinterpret_cast<const uint4*>(src);
            w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
        } else {
            for (int e = 0; e < 16; ++e)
                if (k0 + e < K) w[e >> 2] |= (uint32_t)src[e] << (8 * (e & 3));
        }
        uint4 a, b;
        qc4_to_fp8_planar(w[0], a.x, a.y);
        qc4_to_fp8_planar(w[1], a.z, a.w);
        qc4_to_fp8_planar(w[2], b.x, b.y);
        qc4_to_fp8_planar(w[3], b.z, b.w);
        uint4* dst = reinterpret_cast<uint4*>(out + r * (int64_t)Kp * 2 + (int64_t)k0 * 2);
        dst[0] = a;
        dst[1] = b;
    }
}

// 4 QC bytes -> E3M2 codes of their Re and Im nibbles, one code per byte (bits 5:0).
// Magnitudes {0..7} -> {0x00, 0x0C, 0x10, 0x12, 0x14, 0x15, 0x16, 0x17} (two table dwords
// indexed by v_perm_b32); the nibble's sign bit becomes bit 5.
__device__ __forceinline__ void qc4_to_bf6_bytes(uint32_t x, uint32_t& re, uint32_t& im) {
    const uint32_t TLO = 0x12100C00u;  // e3m2(0..3)
    const uint32_t THI = 0x17161514u;  // e3m2(4..7)
    re = __builtin_amdgcn_perm(THI, TLO, (x >> 4) & 0x07070707u) | ((x >> 2) & 0x20202020u);
    im = __builtin_amdgcn_perm(THI, TLO, x & 0x07070707u) | ((x << 2) & 0x20202020u);
}

// four byte-wide codes -> 24 bits, code k at bits 6k..6k+5
__device__ __forceinline__ uint32_t pack4_bf6(uint32_t c) {
    return (c & 0x3Fu) | ((c >> 2) & 0xFC0u) | ((c >> 4) & 0x3F000u) | ((c >> 6) & 0xFC0000u);
}

// 16 byte-wide codes (c0: codes 0-3, ..., c3: codes 12-15) -> 96 bits in d0..d2
__device__ __forceinline__ void pack16_bf6(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                           uint32_t& d0, uint32_t& d1, uint32_t& d2) {
    const uint32_t p0 = pack4_bf6(c0), p1 = pack4_bf6(c1), p2 = pack4_bf6(c2), p3 = pack4_bf6(c3);
    d0 = p0 | (p1 << 24);
    d1 = (p1 >> 8) | (p2 << 16);
    d2 = (p2 >> 16) | (p3 << 8);
}

// QC rows -> E3M2 operand rows, one 32-byte slot per 16 samples: dwords 0-2 the Re
// codes of samples 0..15, dwords 3-5 the Im codes, dwords 6-7 zero.  Same row pitch
// (2 * Kp bytes) and zero padding as the FP8 planar rows.
template <bool VEC>
__global__ void __launch_bounds__(ST_THREADS)
promote_qc_bf6_planar_kernel(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                             int64_t rows, int K, int Kp) {
    const int segs = Kp / 16;
    const int64_t total = rows * segs;
    const int64_t stride = (int64_t)gridDim.x * ST_THREADS;
    for (int64_t t = (int64_t)blockIdx.x * ST_THREADS + threadIdx.x; t < total; t += stride) {
        const int64_t r = t / segs;
        const int s = (int)(t - r * segs);
        const int k0 = s * 16;
        uint32_t w[4] = {0u, 0u, 0u, 0u};
        const uint8_t* src = in + r * K + k0;
        if (VEC) {
            const uint4 v = *reinterpret_cast<const uint4*>(src);
            w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
        } else {
            for (int e = 0; e < 16; ++e)
                if (k0 + e < K) w[e >> 2] |= (uint32_t)src[e] << (8 * (e & 3));
        }
        uint32_t re[4], im[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) qc4_to_bf6_bytes(w[q], re[q], im[q]);
        uint4 a, b;
        pack16_bf6(re[0], re[1], re[2], re[3], a.x, a.y, a.z);
        pack16_bf6(im[0], im[1], im[2], im[3], a.w, b.x, b.y);
        b.z = 0u;
        b.w = 0u;
        uint4* dst = reinterpret_cast<uint4*>(out + r * (int64_t)Kp * 2 + (int64_t)k0 * 2);
        dst[0] = a;
        dst[1] = b;
    }
}

__device__ __forceinline__ float half_bits_to_float(uint32_t h) {
    _Float16 v;
    const uint16_t b = (uint16_t)h;
    __builtin_memcpy(&v, &b, 2);
    return (float)v;
}

// Four FP16 values (two dwords of bit pairs) -> four OCP E4M3 bytes: clamped to +-448 first
// (a NaN becomes -448, as fmaxf(NaN, -448) in cast_kernels_common.hpp:36-37), then gfx950's
// v_cvt_pk_fp8_f32 rounds to nearest even, subnormals included.  Checked on every FP16 bit
// pattern against the CPU restatement.
__device__ __forceinline__ uint32_t fp16x4_to_e4m3(uint32_t lo, uint32_t hi) {
    float v[4] = {half_bits_to_float(lo & 0xFFFFu), half_bits_to_float(lo >> 16), half_bits_to_float(hi & 0xFFFFu),
                  half_bits_to_float(hi >> 16)};
#pragma unroll
    for (int e = 0; e < 4; ++e) v[e] = __builtin_fminf(__builtin_fmaxf(v[e], -448.0f), 448.0f);
    int w = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], 0, false);
    w = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], w, true);
    return (uint32_t)w;
}

// FP16 [rows][K][2] -> pre-cast rows; 8 samples (32 input bytes, 16 output bytes) per
// thread, 16-byte loads when the row's 8 samples are whole and 16-byte aligned.
__global__ void __launch_bounds__(ST_THREADS)
promote_fp16_fp8_planar_kernel(const uint16_t* __restrict__ in, uint8_t* __restrict__ out,
                               int64_t rows, int K, int Kp, int vec) {
    const int groups = Kp / 8;
    const int64_t total = rows * groups;
    const int64_t stride = (int64_t)gridDim.x * ST_THREADS;
    for (int64_t t = (int64_t)blockIdx.x * ST_THREADS + threadIdx.x; t < total; t += stride) {
        const int64_t r = t / groups;
        const int gi = (int)(t - r * groups);
        const int k0 = gi * 8;
        uint32_t d[8];  // FP16 pairs of samples k0 .. k0+7: (re | im << 16) per sample
        if (vec && k0 + 8 <= K) {
            const uint4* src = reinterpret_cast<const uint4*>(in + (r * K + k0) * 2);
            const uint4 a = src[0], b = src[1];
            d[0] = a.x; d[1] = a.y; d[2] = a.z; d[3] = a.w;
            d[4] = b.x; d[5] = b.y; d[6] = b.z; d[7] = b.w;
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = k0 + e;
                d[e] = k < K ? reinterpret_cast<const uint32_t*>(in)[r * K + k] : 0u;
            }
        }
        uint32_t re[2], im[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t* q = d + 4 * h;
            // gather the Re halves and the Im halves of four samples
            re[h] = fp16x4_to_e4m3(__builtin_amdgcn_perm(q[1], q[0], 0x05040100u),
                                   __builtin_amdgcn_perm(q[3], q[2], 0x05040100u));
            im[h] = fp16x4_to_e4m3(__builtin_amdgcn_perm(q[1], q[0], 0x07060302u),
                                   __builtin_amdgcn_perm(q[3], q[2], 0x07060302u));
        }
        uint4* dst = reinterpret_cast<uint4*>(out + r * (int64_t)Kp * 2 + (int64_t)k0 * 2);
        *dst = make_uint4(re[0], im[0], re[1], im[1]);
    }
}

// Planar FP16 Re / Im [rows][K] -> pre-cast rows; conj negates Im (the prepared B
// operand of the GEMM holds conj(B)).  8 samples per thread, 16-byte loads when the row's
// 8 samples are whole and 16-byte aligned.
__global__ void __launch_bounds__(ST_THREADS)
promote_fp16_planar_rows_kernel(const uint16_t* __restrict__ re_in, const uint16_t* __restrict__ im_in,
                                uint8_t* __restrict__ out, int64_t rows, int K, int Kp, uint32_t im_sign, int vec) {
    const int groups = Kp / 8;
    const int64_t total = rows * groups;
    const int64_t stride = (int64_t)gridDim.x * ST_THREADS;
    for (int64_t t = (int64_t)blockIdx.x * ST_THREADS + threadIdx.x; t < total; t += stride) {
        const int64_t r = t / groups;
        const int gi = (int)(t - r * groups);
        const int k0 = gi * 8;
        uint32_t rr[4], ii[4];   // the 8 samples' FP16 bits, two per dword
        uint32_t neg[2] = {0u, 0u};  // conj: sign bytes of the samples that exist
        if (vec && k0 + 8 <= K) {
            const uint4 a = *reinterpret_cast<const uint4*>(re_in + r * K + k0);
            const uint4 b = *reinterpret_cast<const uint4*>(im_in + r * K + k0);
            rr[0] = a.x; rr[1] = a.y; rr[2] = a.z; rr[3] = a.w;
            ii[0] = b.x; ii[1] = b.y; ii[2] = b.z; ii[3] = b.w;
            neg[0] = neg[1] = im_sign * 0x01010101u;
        } else {
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                const int k = k0 + 2 * d;
                const uint32_t r0 = k < K ? re_in[r * K + k] : 0u, r1 = k + 1 < K ? re_in[r * K + k + 1] : 0u;
                const uint32_t i0 = k < K ? im_in[r * K + k] : 0u, i1 = k + 1 < K ? im_in[r * K + k + 1] : 0u;
                rr[d] = r0 | (r1 << 16);
                ii[d] = i0 | (i1 << 16);
            }
#pragma unroll
            for (int e = 0; e < 8; ++e)
                if (k0 + e < K) neg[e >> 2] |= im_sign << (8 * (e & 3));
        }
        const uint32_t re0 = fp16x4_to_e4m3(rr[0], rr[1]), re1 = fp16x4_to_e4m3(rr[2], rr[3]);
        const uint32_t im0 = fp16x4_to_e4m3(ii[0], ii[1]) ^ neg[0], im1 = fp16x4_to_e4m3(ii[2], ii[3]) ^ neg[1];
        uint4* dst = reinterpret_cast<uint4*>(out + r * (int64_t)Kp * 2 + (int64_t)k0 * 2);
        *dst = make_uint4(re0, im0, re1, im1);
    }
}

// FP8 E4M3 interleaved [re, im] byte pairs [rows][K][2] -> pre-cast rows: per 4
// samples, bytes re0 im0 re1 im1 | re2 im2 re3 im3 -> re0..re3 | im0..im3.
__global__ void __launch_bounds__(ST_THREADS)
promote_fp8_interleaved_rows_kernel(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                    int64_t rows, int K, int Kp) {
    const int groups = Kp / 4;
    const int64_t total = rows * groups;
    const int64_t stride = (int64_t)gridDim.x * ST_THREADS;
    for (int64_t t = (int64_t)blockIdx.x * ST_THREADS + threadIdx.x; t < total; t += stride) {
        const int64_t r = t / groups;
        const int gi = (int)(t - r * groups);
        const int k0 = gi * 4;
        const uint8_t* src = in + (r * K + k0) * 2;
        uint32_t lo = 0u, hi = 0u;
        if (k0 + 4 <= K && ((reinterpret_cast<uintptr_t>(src) & 7u) == 0)) {
            const uint2 v = *reinterpret_cast<const uint2*>(src);
            lo = v.x;
            hi = v.y;
        } else {
            for (int e = 0; e < 8; ++e)
                if (k0 + e / 2 < K) (e < 4 ? lo : hi) |= (uint32_t)src[e] << (8 * (e & 3));
        }
        const uint32_t re = __builtin_amdgcn_perm(hi, lo, 0x06040200u);  // bytes 0, 2, 4, 6
        const uint32_t im = __builtin_amdgcn_perm(hi, lo, 0x07050301u);  // bytes 1, 3, 5, 7
        uint2* dst = reinterpret_cast<uint2*>(out + r * (int64_t)Kp * 2 + (int64_t)gi * 8);
        *dst = make_uint2(re, im);
    }
}

__device__ __forceinline__ uint16_t qc_nibble_fp16(uint32_t n) {
    // sign-magnitude nibble -> FP16 bits (the -0 nibble gives -0.0, as the reference LUT)
    const uint32_t m = n & 7u;
    uint32_t bits = 0u;
    if (m) {
        const int e = 31 - __builtin_clz(m);  // floor(log2 m)
        bits = ((uint32_t)(e + 15) << 10) | ((m << (10 - e)) & 0x3FFu);
    }
    return (uint16_t)(bits | ((n & 8u) << 12));
}

__global__ void __launch_bounds__(ST_THREADS)
promote_qc_fp16_planar_kernel(const uint8_t* __restrict__ in, uint16_t* __restrict__ re,
                              uint16_t* __restrict__ im, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * ST_THREADS;
    for (int64_t i = (int64_t)blockIdx.x * ST_THREADS + threadIdx.x; i < n; i += stride) {
        const uint32_t b = in[i];
        re[i] = qc_nibble_fp16(b >> 4);
        im[i] = qc_nibble_fp16(b & 15u);
    }
}

template <typename T>
__global__ void __launch_bounds__(ST_THREADS)
promote_data_kernel(const uint8_t* __restrict__ in, T* __restrict__ out, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * ST_THREADS;
    for (int64_t i = (int64_t)blockIdx.x * ST_THREADS + threadIdx.x; i < n; i += stride) {
        const int b = in[i];
        const int lo = ((b & 15) ^ 8) - 8;  // two's complement nibble
        const int hi = (((b >> 4) & 15) ^ 8) - 8;
        out[2 * i] = (T)lo;
        out[2 * i + 1] = (T)hi;
    }
}

// one row per block (grid N x batch), as the reference's row-per-block pack
__global__ void __launch_bounds__(ST_THREADS)
pack_triangle_kernel(const float2* __restrict__ full, float2* __restrict__ packed,
                     const float2* __restrict__ old, int N, float beta) {
    const int row = (int)blockIdx.x, b = (int)blockIdx.y;
    const float2* src = full + ((int64_t)b * N + row) * N;
    const int64_t base = (int64_t)b * N * (N + 1) / 2 + tri_index(row, 0);
    for (int col = threadIdx.x; col <= row; col += ST_THREADS) {
        float2 v = src[col];
        if (beta != 0.0f && old) {
            const float2 o = old[base + col];
            v.x = __builtin_fmaf(beta, o.x, v.x);
            v.y = __builtin_fmaf(beta, o.y, v.y);
        }
        packed[base + col] = v;
    }
}

template <typename T>
__global__ void __launch_bounds__(ST_THREADS)
tri_from_herm_kernel(const T* __restrict__ full, T* __restrict__ tri, int N) {
    const int row = (int)blockIdx.x, b = (int)blockIdx.y;
    const T* src = full + (((int64_t)b * N + row) * N) * 2;
    // upper triangle, row-major: (row, col >= row) at row*N - row(row-1)/2 + col - row
    T* dst = tri + ((int64_t)b * N * (N + 1) / 2 + (int64_t)row * N - (int64_t)row * (row - 1) / 2 - row) * 2;
    for (int col = row + threadIdx.x; col < N; col += ST_THREADS) {
        dst[2 * col] = src[2 * col];
        dst[2 * col + 1] = src[2 * col + 1];
    }
}

template <typename T>
__global__ void __launch_bounds__(ST_THREADS)
interleave_tri_kernel(const uint16_t* __restrict__ re, const uint16_t* __restrict__ im,
                      T* __restrict__ out, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * ST_THREADS;
    for (int64_t i = (int64_t)blockIdx.x * ST_THREADS + threadIdx.x; i < n; i += stride) {
        out[2 * i] = (T)half_bits_to_float(re[i]);
        out[2 * i + 1] = (T)half_bits_to_float(im[i]);
    }
}

// QC sign-magnitude (high nibble Re) -> interleaved FP32 [re, im]; the -0 nibble
// gives +0 (integer product first, as PromoteQcSm, include/kernels/visibility_bf.cuh:38-56)
__global__ void __launch_bounds__(ST_THREADS)
promote_qc_fp32_kernel(const uint8_t* __restrict__ in, float2* __restrict__ out, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * ST_THREADS;
    for (int64_t i = (int64_t)blockIdx.x * ST_THREADS + threadIdx.x; i < n; i += stride) {
        const int b = in[i];
        const int re = ((b >> 4) & 8) ? -((b >> 4) & 7) : ((b >> 4) & 7);
        const int im = (b & 8) ? -(b & 7) : (b & 7);
        out[i] = make_float2((float)re, (float)im);
    }
}

// QC sign-magnitude (high = Re) -> FTD two's complement nibbles (low = Re, high = Im),
// 16 bytes per thread; in == out allowed.
__device__ __forceinline__ uint32_t qc_to_ftd4(uint32_t x) {
    // per byte: re = sm(hi nibble), im = sm(lo nibble); out = (im & 15) << 4 | (re & 15)
    const uint32_t mag_hi = (x >> 4) & 0x07070707u, neg_hi = (x >> 7) & 0x01010101u;
    const uint32_t mag_lo = x & 0x07070707u, neg_lo = (x >> 3) & 0x01010101u;
    // two's complement of a 4-bit field: (m ^ 0xF) + 1 for negative fields.  m <= 7
    // gives at most 0x10 per byte (the -0 case, m = 0), so no carry reaches the next
    // byte and the final mask drops bit 4.
    const uint32_t re = ((mag_hi ^ (neg_hi * 0x0Fu)) + neg_hi) & 0x0F0F0F0Fu;
    const uint32_t im = ((mag_lo ^ (neg_lo * 0x0Fu)) + neg_lo) & 0x0F0F0F0Fu;
    return re | (im << 4);
}

__global__ void __launch_bounds__(ST_THREADS)
convert_qc_to_ftd_kernel(const uint8_t* in, uint8_t* out, int64_t n) {
    const int64_t n16 = n / 16;
    const int64_t stride = (int64_t)gridDim.x * ST_THREADS;
    for (int64_t t = (int64_t)blockIdx.x * ST_THREADS + threadIdx.x; t < n16; t += stride) {
        uint4 v = reinterpret_cast<const uint4*>(in)[t];
        v.x = qc_to_ftd4(v.x);
        v.y = qc_to_ftd4(v.y);
        v.z = qc_to_ftd4(v.z);
        v.w = qc_to_ftd4(v.w);
        reinterpret_cast<uint4*>(out)[t] = v;
    }
    const int64_t t = (int64_t)blockIdx.x * ST_THREADS + threadIdx.x;
    if (t < n - n16 * 16) {
        const int64_t i = n16 * 16 + t;
        out[i] = (uint8_t)(qc_to_ftd4(in[i]) & 0xFFu);
    }
}

// ---- GGP XEngine input: [batch][K][N] two's-complement bytes -> pre-cast rows ----

// 4 GGP bytes (4 samples; low nibble = Re, high = Im, two's complement in [-8, 7])
// -> u dwords re = fp8(Re0..Re3), im = fp8(Im0..Im3).  A 16-entry byte table over
// two v_perm lookups (nibbles 0-7 and 8-15) and a per-byte select on bit 3.
__device__ __forceinline__ uint32_t ggp_nib4_to_fp8(uint32_t nib) {
    const uint32_t P0 = 0x44403800u, P1 = 0x4E4C4A48u;  // fp8(0, 1, 2, 3), fp8(4 .. 7)
    const uint32_t M0 = 0xCACCCED0u, M1 = 0xB8C0C4C8u;  // fp8(-8, -7, -6, -5), fp8(-4 .. -1)
    const uint32_t sel = nib & 0x07070707u;
    const uint32_t pos = __builtin_amdgcn_perm(P1, P0, sel);
    const uint32_t neg = __builtin_amdgcn_perm(M1, M0, sel);
    const uint32_t m = ((nib >> 3) & 0x01010101u) * 0xFFu;
    return (pos & ~m) | (neg & m);
}

// d0..d3: one dword per sample holding 4 antennas' bytes -> one dword per antenna
// holding its 4 samples' bytes (byte a of dword s goes to byte s of output a).
__device__ __forceinline__ void transpose4x4_bytes(uint32_t d0, uint32_t d1, uint32_t d2, uint32_t d3,
                                                   uint32_t (&o)[4]) {
    const uint32_t l01 = __builtin_amdgcn_perm(d1, d0, 0x05010400u);  // d0.b0 d1.b0 d0.b1 d1.b1
    const uint32_t h01 = __builtin_amdgcn_perm(d1, d0, 0x07030602u);  // d0.b2 d1.b2 d0.b3 d1.b3
    const uint32_t l23 = __builtin_amdgcn_perm(d3, d2, 0x05010400u);
    const uint32_t h23 = __builtin_amdgcn_perm(d3, d2, 0x07030602u);
    o[0] = __builtin_amdgcn_perm(l23, l01, 0x05040100u);
    o[1] = __builtin_amdgcn_perm(l23, l01, 0x07060302u);
    o[2] = __builtin_amdgcn_perm(h23, h01, 0x05040100u);
    o[3] = __builtin_amdgcn_perm(h23, h01, 0x07060302u);
}

// Transposing converters: the input holds, for every sample k, a run of R
// contiguous bytes (index r) that belong to R different operand rows; the output
// is pre-cast operand rows [row][Kp] (pitch 2*Kp bytes, samples K .. Kp-1 zero).
//   GGP XEngine  (outer o = batch b):        byte at b*K*N + k*N + r,
//                r = antenna, row = b*N + r
//   FTD raw      (outer o = ti*n_ant + ant): byte at (k*n_ti + ti)*n_ant*C + ant*C + r,
//                C = n_chan*n_pol, r = chan*n_pol + pol,
//                row = (r*n_ti + ti)*n_ant + ant   (batch = (chan*n_pol + pol)*n_ti + ti)
struct RowsLayout {
    int R;            // contiguous run length (row-index extent)
    int K, Kp;
    int outers;       // number of outer indices
    int64_t k_stride; // bytes between samples k and k+1
    int ftd;          // 0: GGP layout, 1: FTD raw layout
    int n_ti, n_ant;  // FTD only
    __device__ int64_t in_base(int o) const {
        if (!ftd) return (int64_t)o * K * R;
        const int ti = o / n_ant, ant = o - ti * n_ant;
        return (int64_t)ti * n_ant * R + (int64_t)ant * R;
    }
    __device__ int64_t out_row(int o, int r) const {
        if (!ftd) return (int64_t)o * R + r;
        const int ti = o / n_ant, ant = o - ti * n_ant;
        return ((int64_t)r * n_ti + ti) * n_ant + ant;
    }
};

// Thread = 4 row indices x 16 samples.  Within a wave 16 lanes run along r (64-byte
// reads of each sample's run) and 4 along samples (128-byte runs of each output
// row).  Block: 64 quads of r x 64 samples; grid.z strides over the outer index.
// TWOS: two's complement nibbles with low = Re (GGP / FTD); else QC sign-magnitude
// with high = Re.
template <bool ALIGNED, bool TWOS>
__global__ void __launch_bounds__(ST_THREADS)
transpose_to_fp8_rows_kernel(const uint8_t* __restrict__ in, uint8_t* __restrict__ out, RowsLayout L) {
    const int tid = threadIdx.x;
    const int rq = (tid & 15) + 16 * (tid >> 6);
    const int kg = (tid >> 4) & 3;
    const int r0 = ((int)blockIdx.y * 64 + rq) * 4;
    const int k0 = ((int)blockIdx.x * 4 + kg) * 16;
    if (r0 >= L.R || k0 >= L.Kp) return;
    for (int o = blockIdx.z; o < L.outers; o += gridDim.z) {
        const uint8_t* src = in + L.in_base(o) + r0;
        uint32_t d[16];
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            const int k = k0 + s;
            uint32_t v = 0u;
            if (k < L.K) {
                const uint8_t* run = src + (int64_t)k * L.k_stride;
                if (ALIGNED) {
                    v = *reinterpret_cast<const uint32_t*>(run);
                } else {
#pragma unroll
                    for (int a = 0; a < 4; ++a)
                        if (r0 + a < L.R) v |= (uint32_t)run[a] << (8 * a);
                }
            }
            d[s] = v;
        }
        uint32_t t[4][4];  // [sample group][row index]
#pragma unroll
        for (int g = 0; g < 4; ++g) transpose4x4_bytes(d[4 * g], d[4 * g + 1], d[4 * g + 2], d[4 * g + 3], t[g]);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            if (r0 + a >= L.R) break;
            uint32_t w[8];
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                if (TWOS) {
                    w[2 * g] = ggp_nib4_to_fp8(t[g][a] & 0x0F0F0F0Fu);             // Re: low nibbles
                    w[2 * g + 1] = ggp_nib4_to_fp8((t[g][a] >> 4) & 0x0F0F0F0Fu);  // Im: high nibbles
                } else {
                    qc4_to_fp8_planar(t[g][a], w[2 * g], w[2 * g + 1]);
                }
            }
            // samples >= K were read as 0x00 bytes, which both codes map to fp8 +0
            uint4* dst = reinterpret_cast<uint4*>(out + L.out_row(o, r0 + a) * (int64_t)L.Kp * 2 +
                                                  (int64_t)k0 * 2);
            dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
            dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
        }
    }
}

template <bool TWOS>
hipError_t launch_transpose_rows(const uint8_t* in, uint8_t* out, const RowsLayout& L, hipStream_t s) {
    dim3 grid((unsigned)((L.Kp + 63) / 64), (unsigned)((L.R + 255) / 256),
              (unsigned)(L.outers < 65535 ? L.outers : 65535));
    const bool aligned = (L.R % 4 == 0) && (L.k_stride % 4 == 0) &&
                         ((reinterpret_cast<uintptr_t>(in) & 3u) == 0);
    if (aligned)
        hipLaunchKernelGGL((transpose_to_fp8_rows_kernel<true, TWOS>), grid, dim3(ST_THREADS), 0, s, in,
                           out, L);
    else
        hipLaunchKernelGGL((transpose_to_fp8_rows_kernel<false, TWOS>), grid, dim3(ST_THREADS), 0, s, in,
                           out, L);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_ggp_to_fp8_rows(const uint8_t* in, uint8_t* out, int N, int K, int batch,
                                  hipStream_t s) {
    if (N <= 0 || K <= 0 || batch <= 0) return hipSuccess;
    RowsLayout L;
    L.R = N;
    L.K = K;
    L.Kp = precast_row_samples(K);
    L.outers = batch;
    L.k_stride = N;
    L.ftd = 0;
    L.n_ti = 1;
    L.n_ant = 1;
    return launch_transpose_rows<true>(in, out, L, s);
}

hipError_t launch_ftd_to_fp8_rows(const uint8_t* in, uint8_t* out, int n_ant, int n_chan, int n_time,
                                  int n_time_inner, int n_pol, bool twos, hipStream_t s) {
    const int K = n_time / n_time_inner;
    if (n_ant <= 0 || n_chan <= 0 || n_pol <= 0 || K <= 0) return hipSuccess;
    RowsLayout L;
    L.R = n_chan * n_pol;
    L.K = K;
    L.Kp = precast_row_samples(K);
    L.outers = n_time_inner * n_ant;
    L.k_stride = (int64_t)n_time_inner * n_ant * n_chan * n_pol;
    L.ftd = 1;
    L.n_ti = n_time_inner;
    L.n_ant = n_ant;
    return twos ? launch_transpose_rows<true>(in, out, L, s) : launch_transpose_rows<false>(in, out, L, s);
}

hipError_t launch_promote_qc_fp32(const uint8_t* in, float2* out, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(promote_qc_fp32_kernel, dim3(grid_for(n)), dim3(ST_THREADS), 0, s, in, out, n);
    return hipGetLastError();
}

hipError_t launch_convert_qc_to_ftd(const uint8_t* in, uint8_t* out, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(convert_qc_to_ftd_kernel, dim3(grid_for(n / 16 + 16)), dim3(ST_THREADS), 0, s,
                       in, out, n);
    return hipGetLastError();
}

namespace {

// FP32 -> FP16 (nearest even) and back; FP32 [re, im] -> FP16 planar (optionally -im);
// FP32 planar -> [re, im].  Grid-stride element loops.
__global__ void __launch_bounds__(ST_THREADS) narrow_fp32_kernel(const float* __restrict__ in,
                                                                _Float16* __restrict__ out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * ST_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * ST_THREADS)
        out[i] = (_Float16)in[i];
}

__global__ void __launch_bounds__(ST_THREADS) widen_fp16_kernel(const _Float16* __restrict__ in,
                                                               float* __restrict__ out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * ST_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * ST_THREADS)
        out[i] = (float)in[i];
}

__global__ void __launch_bounds__(ST_THREADS) deinterleave_narrow_kernel(const float2* __restrict__ in,
                                                                        _Float16* __restrict__ re,
                                                                        _Float16* __restrict__ im, int64_t n,
                                                                        int negate_im) {
    for (int64_t i = (int64_t)blockIdx.x * ST_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * ST_THREADS) {
        const float2 v = in[i];
        re[i] = (_Float16)v.x;
        im[i] = (_Float16)(negate_im ? -v.y : v.y);
    }
}

__global__ void __launch_bounds__(ST_THREADS) interleave_fp32_kernel(const float* __restrict__ re,
                                                                    const float* __restrict__ im,
                                                                    float2* __restrict__ out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * ST_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * ST_THREADS)
        out[i] = make_float2(re[i], im[i]);
}

// QC rows of K bytes at any alignment -> rows of Kp >= K bytes (a multiple of 16), zero
// padded, so the fused HERK can take its 16-byte loads (zero samples add nothing to the
// product).  One thread per 16 output bytes (segment t of rows * Kp / 16): the (one or
// two) aligned 16-byte words that hold its input bytes are loaded and funnel-shifted into
// place, and bytes past the row's end are cleared.  Only words that lie wholly inside
// [in, in + rows * K) are loaded; a segment whose words reach outside (the first or last
// few of the buffer) gathers bytes.
__device__ __forceinline__ void pad_segment(const uint8_t* __restrict__ in, uint8_t* __restrict__ out,
                                            int64_t rows, int K, int Kp, int64_t t) {
    const int segs = Kp / 16;
    const uintptr_t lo = reinterpret_cast<uintptr_t>(in), hi = lo + (uintptr_t)(rows * K);
    const int64_t r = t / segs;
    const int k0 = (int)(t - r * segs) * 16;
    const int v = K - k0 < 16 ? K - k0 : 16;  // input bytes of this segment
    if (v <= 0) {  // wholly past the row's end
        *reinterpret_cast<uint4*>(out + r * Kp + k0) = make_uint4(0u, 0u, 0u, 0u);
        return;
    }
    const uintptr_t addr = lo + (uintptr_t)(r * K + k0);
    const uint32_t off = (uint32_t)(addr & 15u);
    const uintptr_t base = addr - off;
    const bool two = off + (uint32_t)v > 16u;
    uint32_t w[4];
    if (base >= lo && base + (two ? 32u : 16u) <= hi) {
        const uint4 q0 = *reinterpret_cast<const uint4*>(base);
        const uint4 q1 = two ? *reinterpret_cast<const uint4*>(base + 16) : make_uint4(0u, 0u, 0u, 0u);
        const uint32_t d[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
        const uint32_t sw = off >> 2, sb = 8u * (off & 3u);
        uint32_t e[5];
#pragma unroll
        for (int j = 0; j < 5; ++j)  // e[j] = d[j + sw], with constant register indices
            e[j] = sw == 0 ? d[j] : sw == 1 ? d[j + 1] : sw == 2 ? d[j + 2] : d[j + 3];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint64_t pair = ((uint64_t)e[i + 1] << 32) | e[i];
            const int lim = v - 4 * i;  // bytes of this dword inside the row
            const uint32_t keep = lim >= 4 ? 0xFFFFFFFFu : lim <= 0 ? 0u : (1u << (8 * lim)) - 1u;
            w[i] = (uint32_t)(pair >> sb) & keep;
        }
    } else {
        const uint8_t* src = in + r * K + k0;
#pragma unroll
        for (int i = 0; i < 4; ++i) w[i] = 0u;
#pragma unroll
        for (int b = 0; b < 16; ++b)
            if (b < v) w[b >> 2] |= (uint32_t)src[b] << (8 * (b & 3));
    }
    *reinterpret_cast<uint4*>(out + r * Kp + k0) = make_uint4(w[0], w[1], w[2], w[3]);
}

__global__ void __launch_bounds__(ST_THREADS)
pad_qc_rows_kernel(const uint8_t* __restrict__ in, uint8_t* __restrict__ out, int64_t rows, int K, int Kp) {
    const int64_t total = rows * (Kp / 16);
    const int64_t stride = (int64_t)gridDim.x * ST_THREADS;
    for (int64_t t = (int64_t)blockIdx.x * ST_THREADS + threadIdx.x; t < total; t += stride)
        pad_segment(in, out, rows, K, Kp, t);
}

// Several row blocks in one launch (the grouped HERK's problems): workgroup w copies
// segments [blocks[w].seg, + kPadBlockSegments) of job blocks[w].job, consecutive threads
// on consecutive segments.
static_assert(kPadBlockSegments % ST_THREADS == 0, "whole segment rounds per workgroup");
__global__ void __launch_bounds__(ST_THREADS)
pad_qc_rows_multi_kernel(const PadRowsJob* __restrict__ jobs, const PadBlock* __restrict__ blocks) {
    const PadBlock bk = blocks[blockIdx.x];
    const PadRowsJob jb = jobs[bk.job];
    const int64_t total = jb.rows * (jb.Kp / 16);
#pragma unroll 2
    for (int i = 0; i < kPadBlockSegments / ST_THREADS; ++i) {
        const int64_t t = bk.seg + i * ST_THREADS + threadIdx.x;
        if (t < total) pad_segment(jb.src, jb.dst, jb.rows, jb.K, jb.Kp, t);
    }
}

}  // namespace

hipError_t launch_narrow_fp32(const float* in, uint16_t* out, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(narrow_fp32_kernel, dim3(grid_for(n)), dim3(ST_THREADS), 0, s, in,
                       reinterpret_cast<_Float16*>(out), n);
    return hipGetLastError();
}

hipError_t launch_widen_fp16(const uint16_t* in, float* out, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(widen_fp16_kernel, dim3(grid_for(n)), dim3(ST_THREADS), 0, s,
                       reinterpret_cast<const _Float16*>(in), out, n);
    return hipGetLastError();
}

hipError_t launch_deinterleave_narrow(const float* in, uint16_t* re, uint16_t* im, int64_t n, bool negate_im,
                                      hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(deinterleave_narrow_kernel, dim3(grid_for(n)), dim3(ST_THREADS), 0, s,
                       reinterpret_cast<const float2*>(in), reinterpret_cast<_Float16*>(re),
                       reinterpret_cast<_Float16*>(im), n, negate_im ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_interleave_fp32(const float* re, const float* im, float* out, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(interleave_fp32_kernel, dim3(grid_for(n)), dim3(ST_THREADS), 0, s, re, im,
                       reinterpret_cast<float2*>(out), n);
    return hipGetLastError();
}

int precast_row_samples(int K) { return (K + 15) / 16 * 16; }

hipError_t launch_pad_qc_rows_multi(const PadRowsJob* jobs, const PadBlock* blocks, int nblocks, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(pad_qc_rows_multi_kernel, dim3((unsigned)nblocks), dim3(ST_THREADS), 0, s, jobs, blocks);
    return hipGetLastError();
}

hipError_t launch_pad_qc_rows(const uint8_t* in, uint8_t* out, int64_t rows, int K, int Kp, hipStream_t s) {
    if (rows <= 0 || K <= 0) return hipSuccess;
    if (Kp < K || Kp % 16 != 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pad_qc_rows_kernel, dim3(grid_for(rows * (Kp / 16))), dim3(ST_THREADS), 0, s, in, out,
                       rows, K, Kp);
    return hipGetLastError();
}

hipError_t launch_promote_qc_fp8_interleaved(const uint8_t* in, uint8_t* out, int64_t n,
                                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(promote_qc_fp8_interleaved_kernel, dim3(grid_for(n / 16 + 16)),
                       dim3(ST_THREADS), 0, s, in, out, n);
    return hipGetLastError();
}

hipError_t launch_promote_qc_fp8_planar(const uint8_t* in, uint8_t* out, int64_t rows, int K,
                                        hipStream_t s) {
    if (rows <= 0 || K <= 0) return hipSuccess;
    const int Kp = precast_row_samples(K);
    const int64_t work = rows * (Kp / 16);
    const bool vec = (K % 16 == 0) && ((reinterpret_cast<uintptr_t>(in) & 15u) == 0);
    if (vec)
        hipLaunchKernelGGL((promote_qc_fp8_planar_kernel<true>), dim3(grid_for(work)), dim3(ST_THREADS),
                           0, s, in, out, rows, K, Kp);
    else
        hipLaunchKernelGGL((promote_qc_fp8_planar_kernel<false>), dim3(grid_for(work)), dim3(ST_THREADS),
                           0, s, in, out, rows, K, Kp);
    return hipGetLastError();
}

hipError_t launch_promote_qc_bf6_planar(const uint8_t* in, uint8_t* out, int64_t rows, int K,
                                        hipStream_t s) {
    if (rows <= 0 || K <= 0) return hipSuccess;
    const int Kp = precast_row_samples(K);
    const int64_t work = rows * (Kp / 16);
    const bool vec = (K % 16 == 0) && ((reinterpret_cast<uintptr_t>(in) & 15u) == 0);
    if (vec)
        hipLaunchKernelGGL((promote_qc_bf6_planar_kernel<true>), dim3(grid_for(work)), dim3(ST_THREADS),
                           0, s, in, out, rows, K, Kp);
    else
        hipLaunchKernelGGL((promote_qc_bf6_planar_kernel<false>), dim3(grid_for(work)), dim3(ST_THREADS),
                           0, s, in, out, rows, K, Kp);
    return hipGetLastError();
}

hipError_t launch_promote_fp16_fp8_planar(const uint16_t* in, uint8_t* out, int64_t rows, int K,
                                          hipStream_t s) {
    if (rows <= 0 || K <= 0) return hipSuccess;
    const int Kp = precast_row_samples(K);
    // 16-byte loads need every row (K samples of 4 bytes) to start 16-byte aligned
    const int vec = (K % 4 == 0 && (reinterpret_cast<uintptr_t>(in) & 15u) == 0) ? 1 : 0;
    hipLaunchKernelGGL(promote_fp16_fp8_planar_kernel, dim3(grid_for(rows * (Kp / 8))),
                       dim3(ST_THREADS), 0, s, in, out, rows, K, Kp, vec);
    return hipGetLastError();
}

hipError_t launch_promote_fp16_planar_rows(const uint16_t* re, const uint16_t* im, uint8_t* out,
                                           int64_t rows, int K, bool conj, hipStream_t s) {
    if (rows <= 0 || K <= 0) return hipSuccess;
    const int Kp = precast_row_samples(K);
    // 16-byte loads need every row (K samples of 2 bytes) to start 16-byte aligned
    const int vec = (K % 8 == 0 && ((reinterpret_cast<uintptr_t>(re) | reinterpret_cast<uintptr_t>(im)) & 15u) == 0)
                        ? 1 : 0;
    hipLaunchKernelGGL(promote_fp16_planar_rows_kernel, dim3(grid_for(rows * (Kp / 8))),
                       dim3(ST_THREADS), 0, s, re, im, out, rows, K, Kp, conj ? 0x80u : 0u, vec);
    return hipGetLastError();
}

hipError_t launch_promote_fp8_interleaved_rows(const uint8_t* in, uint8_t* out, int64_t rows, int K,
                                               hipStream_t s) {
    if (rows <= 0 || K <= 0) return hipSuccess;
    const int Kp = precast_row_samples(K);
    hipLaunchKernelGGL(promote_fp8_interleaved_rows_kernel, dim3(grid_for(rows * (Kp / 4))),
                       dim3(ST_THREADS), 0, s, in, out, rows, K, Kp);
    return hipGetLastError();
}

hipError_t launch_promote_qc_fp16_planar(const uint8_t* in, uint16_t* re, uint16_t* im, int64_t n,
                                         hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(promote_qc_fp16_planar_kernel, dim3(grid_for(n)), dim3(ST_THREADS), 0, s, in,
                       re, im, n);
    return hipGetLastError();
}

hipError_t launch_promote_data(const uint8_t* in, void* out, int64_t n, bool fp64, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (fp64)
        hipLaunchKernelGGL((promote_data_kernel<double>), dim3(grid_for(n)), dim3(ST_THREADS), 0, s,
                           in, static_cast<double*>(out), n);
    else
        hipLaunchKernelGGL((promote_data_kernel<float>), dim3(grid_for(n)), dim3(ST_THREADS), 0, s,
                           in, static_cast<float*>(out), n);
    return hipGetLastError();
}

hipError_t launch_pack_triangle(const float2* full, float2* packed, const float2* old, int N,
                                int batch, float beta, hipStream_t s) {
    if (N <= 0 || batch <= 0) return hipSuccess;
    hipLaunchKernelGGL(pack_triangle_kernel, dim3((unsigned)N, (unsigned)batch), dim3(ST_THREADS), 0,
                       s, full, packed, old, N, beta);
    return hipGetLastError();
}

hipError_t launch_tri_from_herm(const void* full, void* tri, int N, int batch, bool fp64,
                                hipStream_t s) {
    if (N <= 0 || batch <= 0) return hipSuccess;
    if (fp64)
        hipLaunchKernelGGL((tri_from_herm_kernel<double>), dim3((unsigned)N, (unsigned)batch),
                           dim3(ST_THREADS), 0, s, static_cast<const double*>(full),
                           static_cast<double*>(tri), N);
    else
        hipLaunchKernelGGL((tri_from_herm_kernel<float>), dim3((unsigned)N, (unsigned)batch),
                           dim3(ST_THREADS), 0, s, static_cast<const float*>(full),
                           static_cast<float*>(tri), N);
    return hipGetLastError();
}

hipError_t launch_interleave_tri(const uint16_t* re, const uint16_t* im, void* out, int64_t n,
                                 bool fp64, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (fp64)
        hipLaunchKernelGGL((interleave_tri_kernel<double>), dim3(grid_for(n)), dim3(ST_THREADS), 0, s,
                           re, im, static_cast<double*>(out), n);
    else
        hipLaunchKernelGGL((interleave_tri_kernel<float>), dim3(grid_for(n)), dim3(ST_THREADS), 0, s,
                           re, im, static_cast<float*>(out), n);
    return hipGetLastError();
}

}  // namespace ftdx
