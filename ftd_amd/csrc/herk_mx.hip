// herk_mx.hip -- MX block-scaled HERK on gfx950 (FP4 E2M1 / FP6 E3M2 operands,
// E8M0 scale per 32 samples, FP32 accumulate).
//
// Reference behaviour restated: the block-scaled HERK of
// cutlass_interface/gemm_complex_sm100/herk_impl.hpp:1067-1340 quantizes Re and
// Im separately with one E8M0 scale per (row, 32-sample block), scale
// 2^ceil(log2(amax / narrow_max)) (mxfp_kernels.hpp:147-166), then runs three
// block-scaled sub-GEMMs.  Here one scaled MFMA family does the whole product:
//
//   quantize:  X (INT4 QC bytes or FP16 interleaved complex) -> codes + scales
//              scratch layout  codes  [row][blk][part][PB]  (part 0 = Re, 1 = Im)
//                              scales [row][blk][part]      (E8M0 bytes)
//              blk padded to an even count (one LDS stage = 2 blocks).
//   HERK / GEMM: per 32-sample block one v_mfma_scale_f32_32x32x64_f8f6f4 step,
//              lanes 0-31 carry the Re block, lanes 32-63 the Im block of a row:
//                 Re C_ij = sum (re_i re_j + im_i im_j)     A = [re | im],  B = [re | im]
//                 Im C_ij = sum (im_i re_j - re_i im_j)     A = [im | -re], B = [re | im]
//              each lane's 32 values share one scale, which the MFMA applies.
// Element conversion: nearest-even with subnormals, saturating at the format
// maximum (OCP Microscaling Formats v1.0); codes per value are packed LSB
// first (FP4: nibble i at bits 4i; FP6: bits 6i..6i+5 of the lane's dwords).
#include <hip/hip_runtime.h>

#include "herk_common.h"
#include "herk_mx.h"
#include "herk_tile.h"

namespace ftdx {

namespace {

template <int FMT>
struct MxFmt;

template <>
struct MxFmt<MX_FP4_E2M1> {
    static constexpr int PB = 16;        // bytes per 32 codes
    static constexpr int DW = 4;         // dwords per lane operand
    static constexpr int MFMA_FMT = 4;   // cbsz/blgp code: FP4 E2M1
    static constexpr int MBITS = 1, BIAS = 1, SIGN_SHIFT = 3, MAXCODE = 7;
    static constexpr float MAXV = 6.0f;
    static __device__ __forceinline__ uint32_t neg(int d) { (void)d; return 0x88888888u; }
};

template <>
struct MxFmt<MX_FP6_E3M2> {
    static constexpr int PB = 24;
    static constexpr int DW = 6;
    static constexpr int MFMA_FMT = 3;   // BF6 = E3M2
    static constexpr int MBITS = 2, BIAS = 3, SIGN_SHIFT = 5, MAXCODE = 31;
    static constexpr float MAXV = 28.0f;
    // sign bits at 6i+5 of the 192-bit lane operand, per dword
    static __device__ __forceinline__ uint32_t neg(int d) {
        return (d % 3 == 0) ? 0x20820820u : (d % 3 == 1) ? 0x08208208u : 0x82082082u;
    }
};

// Value -> code, nearest-even with subnormals, saturating (same steps as the
// CPU restatement, so codes are bit-identical).
template <int FMT>
__device__ __forceinline__ uint32_t mx_encode(float x) {
    using F = MxFmt<FMT>;
    const uint32_t sign = __builtin_signbitf(x) ? 1u : 0u;
    const float a = __builtin_fabsf(x);
    uint32_t code;
    if (a >= F::MAXV) {
        code = F::MAXCODE;
    } else {
        int e;
        __builtin_frexpf(a, &e);
        e -= 1;
        const int emin = 1 - F::BIAS;
        if (a == 0.0f || e < emin) {
            code = (uint32_t)__builtin_rintf(__builtin_ldexpf(a, F::MBITS - emin));
        } else {
            uint32_t q = (uint32_t)__builtin_rintf(__builtin_ldexpf(a, F::MBITS - e));
            if (q == (2u << F::MBITS)) {
                q >>= 1;
                ++e;
            }
            code = ((uint32_t)(e + F::BIAS) << F::MBITS) | (q - (1u << F::MBITS));
        }
        if (code > (uint32_t)F::MAXCODE) code = F::MAXCODE;
    }
    return (sign << F::SIGN_SHIFT) | code;
}

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
template <int FMT, int IN>
__global__ void __launch_bounds__(256) mx_quantize_kernel(const void* __restrict__ X,
                                                          const void* __restrict__ X2,
                                                          uint8_t* __restrict__ codes,
                                                          uint8_t* __restrict__ scales,
                                                          int64_t rows, int K, int nblkp, int conj) {
    using F = MxFmt<FMT>;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows * nblkp) return;
    const int64_t row = t / nblkp;
    const int blk = (int)(t - row * nblkp);
    float v[2][32];
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
#pragma unroll
    for (int part = 0; part < 2; ++part) {
        float amax = 0.0f;
#pragma unroll
        for (int i = 0; i < 32; ++i) amax = __builtin_fmaxf(amax, __builtin_fabsf(v[part][i]));
        const uint32_t se = mx_scale_exponent<FMT>(amax);
        const float inv = __builtin_ldexpf(1.0f, 127 - (int)se);
        uint32_t w[6] = {0u, 0u, 0u, 0u, 0u, 0u};
#pragma unroll
        for (int i = 0; i < 32; ++i) {
            const uint32_t c = mx_encode<FMT>(v[part][i] * inv);
            if (FMT == MX_FP4_E2M1) {
                w[i >> 3] |= c << (4 * (i & 7));
            } else {
                const int bit = 6 * i, d = bit >> 5, o = bit & 31;
                w[d] |= c << o;
                if (o > 26) w[d + 1] |= c >> (32 - o);
            }
        }
        uint32_t* dst = reinterpret_cast<uint32_t*>(codes + ((row * nblkp + blk) * 2 + part) * F::PB);
#pragma unroll
        for (int d = 0; d < F::DW; ++d) dst[d] = (part == 1 && conj) ? w[d] ^ F::neg(d) : w[d];
        scales[(row * nblkp + blk) * 2 + part] = (uint8_t)se;
    }
}

// ---------------------------------------------------------------- HERK
constexpr int MX_TILE = 128;
constexpr int MX_THREADS = 256;
constexpr int MX_BLK_PER_STAGE = 2;

template <int FMT>
struct MxCfg {
    using F = MxFmt<FMT>;
    static constexpr int ROW_DATA = MX_BLK_PER_STAGE * 2 * F::PB;  // code bytes per row per stage
    static constexpr int PITCH = ROW_DATA + 16;
    static constexpr int ROWS = 2 * MX_TILE;                        // row block + column block
    static constexpr int CODE_BYTES = ROWS * PITCH;
    static constexpr int SCALE_BYTES = ROWS * 4;                    // 2 blocks x 2 parts
    static constexpr int STAGE = CODE_BYTES + SCALE_BYTES;
    static constexpr int LDS_BYTES = 2 * STAGE;
    static constexpr int SEGS_PER_ROW = ROW_DATA / 16;
    static constexpr int SEGS_PER_THREAD = ROWS * SEGS_PER_ROW / MX_THREADS;
    static_assert(ROWS * SEGS_PER_ROW % MX_THREADS == 0, "segments must divide over threads");
    static_assert(ROWS == MX_THREADS, "one scale dword per thread");
};

template <int FMT>
__device__ __forceinline__ v8i mx_read_operand(const uint8_t* p) {
    v8i r;
    if (FMT == MX_FP4_E2M1) {
        uint4 a = *reinterpret_cast<const uint4*>(p);
        r[0] = (int)a.x; r[1] = (int)a.y; r[2] = (int)a.z; r[3] = (int)a.w;
        r[4] = r[5] = r[6] = r[7] = 0;
    } else {
        uint2 a = *reinterpret_cast<const uint2*>(p);
        uint2 b = *reinterpret_cast<const uint2*>(p + 8);
        uint2 c = *reinterpret_cast<const uint2*>(p + 16);
        r[0] = (int)a.x; r[1] = (int)a.y; r[2] = (int)b.x; r[3] = (int)b.y;
        r[4] = (int)c.x; r[5] = (int)c.y; r[6] = r[7] = 0;
    }
    return r;
}

template <int FMT>
__device__ __forceinline__ v16f mx_mfma(v8i a, v8i b, v16f c, int sa, int sb) {
    constexpr int code = MxFmt<FMT>::MFMA_FMT;
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, c, code, code, 0, sa, 0, sb);
}

// FP6 operands need more registers than two waves per SIMD allow without spills.
template <int FMT, typename OutT>
__global__ void __launch_bounds__(MX_THREADS, FMT == MX_FP4_E2M1 ? 2 : 1)
herk_mx_kernel(const uint8_t* __restrict__ codes, const uint8_t* __restrict__ scales, void* C,
               int N, int nblkp, int tiles, float alpha, float beta, int64_t c_bstride) {
    using Cfg = MxCfg<FMT>;
    using F = MxFmt<FMT>;
    __shared__ __attribute__((aligned(16))) uint8_t lds[Cfg::LDS_BYTES];

    const int w = (int)blockIdx.x;
    const int b = w / tiles;
    int I, J;
    tri_decode_r(w - b * tiles, 1, I, J);
    const int row0 = I * MX_TILE, col0 = J * MX_TILE;
    const bool diag = (I == J);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int wrow0 = row0 + wm * 64, wcol0 = col0 + wn * 64;
    const bool active = (wrow0 < N) && (wrow0 + 63 >= wcol0);

    const int64_t row_bytes = (int64_t)nblkp * 2 * F::PB;
    const uint8_t* cb = codes + (int64_t)b * N * row_bytes;
    const uint8_t* sb = scales + (int64_t)b * N * nblkp * 2;
    OutT* Cb = static_cast<OutT*>(C) + (int64_t)b * c_bstride;

    v16f acc_re[2][2], acc_im[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                acc_re[x][y][r] = 0.0f;
                acc_im[x][y][r] = 0.0f;
            }

    const int nst = nblkp / MX_BLK_PER_STAGE;
    uint4 g[Cfg::SEGS_PER_THREAD];
    uint32_t gs;
    auto load = [&](int s) {
#pragma unroll
        for (int q = 0; q < Cfg::SEGS_PER_THREAD; ++q) {
            const int seg = tid + MX_THREADS * q;
            const int lrow = seg / Cfg::SEGS_PER_ROW, sg = seg % Cfg::SEGS_PER_ROW;
            const bool is_col = lrow >= MX_TILE;
            const int grow = is_col ? col0 + (lrow - MX_TILE) : row0 + lrow;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (!(is_col && diag) && grow < N)
                v = *reinterpret_cast<const uint4*>(cb + grow * row_bytes +
                                                     (int64_t)s * Cfg::ROW_DATA + sg * 16);
            g[q] = v;
        }
        {
            const bool is_col = tid >= MX_TILE;
            const int grow = is_col ? col0 + (tid - MX_TILE) : row0 + tid;
            gs = 0u;
            if (!(is_col && diag) && grow < N)
                gs = *reinterpret_cast<const uint32_t*>(sb + (int64_t)grow * nblkp * 2 + s * 4);
        }
    };
    auto store = [&](uint8_t* st) {
#pragma unroll
        for (int q = 0; q < Cfg::SEGS_PER_THREAD; ++q) {
            const int seg = tid + MX_THREADS * q;
            const int lrow = seg / Cfg::SEGS_PER_ROW, sg = seg % Cfg::SEGS_PER_ROW;
            if (lrow >= MX_TILE && diag) continue;
            *reinterpret_cast<uint4*>(st + lrow * Cfg::PITCH + sg * 16) = g[q];
        }
        if (!(tid >= MX_TILE && diag))
            *reinterpret_cast<uint32_t*>(st + Cfg::CODE_BYTES + tid * 4) = gs;
    };

    if (nst > 0) {
        load(0);
        store(lds);
    }
    __syncthreads();

    const int lr = lane & 31, h = lane >> 5;
    const int a_row = wm * 64 + lr;
    const int b_row = (diag ? 0 : MX_TILE) + wn * 64 + lr;

    for (int s = 0; s < nst; ++s) {
        const bool more = s + 1 < nst;
        if (more) load(s + 1);
        const uint8_t* st = lds + (s & 1) * Cfg::STAGE;
        const uint8_t* ss = st + Cfg::CODE_BYTES;
        if (active) {
#pragma unroll
            for (int blk = 0; blk < MX_BLK_PER_STAGE; ++blk) {
                v8i bu[2];
                int sbu[2];
#pragma unroll
                for (int y = 0; y < 2; ++y) {
                    const int c = b_row + y * 32;
                    bu[y] = mx_read_operand<FMT>(st + c * Cfg::PITCH + (blk * 2 + h) * F::PB);
                    sbu[y] = ss[c * 4 + blk * 2 + h];
                }
#pragma unroll
                for (int x = 0; x < 2; ++x) {
                    const int r = a_row + x * 32;
                    const v8i au = mx_read_operand<FMT>(st + r * Cfg::PITCH + (blk * 2 + h) * F::PB);
                    v8i av = mx_read_operand<FMT>(st + r * Cfg::PITCH + (blk * 2 + 1 - h) * F::PB);
                    const int sau = ss[r * 4 + blk * 2 + h];
                    const int sav = ss[r * 4 + blk * 2 + 1 - h];
                    if (h) {
#pragma unroll
                        for (int d = 0; d < F::DW; ++d) av[d] ^= (int)F::neg(d);
                    }
#pragma unroll
                    for (int y = 0; y < 2; ++y) {
                        acc_re[x][y] = mx_mfma<FMT>(au, bu[y], acc_re[x][y], sau, sbu[y]);
                        acc_im[x][y] = mx_mfma<FMT>(av, bu[y], acc_im[x][y], sav, sbu[y]);
                    }
                }
            }
        }
        if (more) store(lds + ((s + 1) & 1) * Cfg::STAGE);
        __syncthreads();
    }

    if (!active) return;
    const int n = lane & 31, g4 = (lane >> 5) * 4;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int gj = wcol0 + y * 32 + n;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int gi = wrow0 + x * 32 + (r >> 2) * 8 + g4 + (r & 3);
                if (gi < N && gj <= gi) {
                    const int64_t idx = tri_index(gi, gj);
                    float re = acc_re[x][y][r];
                    float im = (gi == gj) ? 0.0f : acc_im[x][y][r];
                    re = re * alpha;
                    im = im * alpha;
                    if (beta != 0.0f) {
                        float ore, oim;
                        OutIO<OutT>::load(Cb, idx, ore, oim);
                        re = __builtin_fmaf(beta, ore, re);
                        im = __builtin_fmaf(beta, oim, im);
                    }
                    OutIO<OutT>::store(Cb, idx, re, im);
                }
            }
        }
}

// ---------------------------------------------------------------- complex GEMM
// C = alpha A B^T + beta C (planar Re / Im outputs, row pitch ldc) on MX codes: the row
// block from A's codes, the column block from the codes of conj(B), so the HERK
// operand algebra above gives Re = sum(ar br - ai bi), Im = sum(ai br + ar bi).  Full
// grid of 128x128 tiles, column-major tile order within a batch.
template <int FMT, typename OutT>
__global__ void __launch_bounds__(MX_THREADS, FMT == MX_FP4_E2M1 ? 2 : 1)
gemm_mx_kernel(MxGemmArgs p, int tiles_m, int tiles) {
    using Cfg = MxCfg<FMT>;
    using F = MxFmt<FMT>;
    __shared__ __attribute__((aligned(16))) uint8_t lds[Cfg::LDS_BYTES];

    const int w = (int)blockIdx.x;
    const int b = w / tiles;
    const int t = w - b * tiles;
    const int J = t / tiles_m, I = t - J * tiles_m;
    const int row0 = I * MX_TILE, col0 = J * MX_TILE;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int wrow0 = row0 + wm * 64, wcol0 = col0 + wn * 64;
    const bool active = wrow0 < p.M && wcol0 < p.N;
    const int nblkp = p.nblkp;

    const int64_t row_bytes = (int64_t)nblkp * 2 * F::PB;
    const uint8_t* ac = p.acodes + (int64_t)b * p.M * row_bytes;
    const uint8_t* as = p.ascales + (int64_t)b * p.M * nblkp * 2;
    const uint8_t* bc = p.bcodes + (int64_t)b * p.N * row_bytes;
    const uint8_t* bs = p.bscales + (int64_t)b * p.N * nblkp * 2;

    v16f acc_re[2][2], acc_im[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                acc_re[x][y][r] = 0.0f;
                acc_im[x][y][r] = 0.0f;
            }

    const int nst = nblkp / MX_BLK_PER_STAGE;
    uint4 g[Cfg::SEGS_PER_THREAD];
    uint32_t gs;
    auto load = [&](int s) {
#pragma unroll
        for (int q = 0; q < Cfg::SEGS_PER_THREAD; ++q) {
            const int seg = tid + MX_THREADS * q;
            const int lrow = seg / Cfg::SEGS_PER_ROW, sg = seg % Cfg::SEGS_PER_ROW;
            const bool is_col = lrow >= MX_TILE;
            const int grow = is_col ? col0 + (lrow - MX_TILE) : row0 + lrow;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (grow < (is_col ? p.N : p.M))
                v = *reinterpret_cast<const uint4*>((is_col ? bc : ac) + grow * row_bytes +
                                                     (int64_t)s * Cfg::ROW_DATA + sg * 16);
            g[q] = v;
        }
        {
            const bool is_col = tid >= MX_TILE;
            const int grow = is_col ? col0 + (tid - MX_TILE) : row0 + tid;
            gs = 0u;
            if (grow < (is_col ? p.N : p.M))
                gs = *reinterpret_cast<const uint32_t*>((is_col ? bs : as) + (int64_t)grow * nblkp * 2 + s * 4);
        }
    };
    auto store = [&](uint8_t* st) {
#pragma unroll
        for (int q = 0; q < Cfg::SEGS_PER_THREAD; ++q) {
            const int seg = tid + MX_THREADS * q;
            const int lrow = seg / Cfg::SEGS_PER_ROW, sg = seg % Cfg::SEGS_PER_ROW;
            *reinterpret_cast<uint4*>(st + lrow * Cfg::PITCH + sg * 16) = g[q];
        }
        *reinterpret_cast<uint32_t*>(st + Cfg::CODE_BYTES + tid * 4) = gs;
    };

    if (nst > 0) {
        load(0);
        store(lds);
    }
    __syncthreads();

    const int lr = lane & 31, h = lane >> 5;
    const int a_row = wm * 64 + lr;
    const int b_row = MX_TILE + wn * 64 + lr;
    for (int s = 0; s < nst; ++s) {
        const bool more = s + 1 < nst;
        if (more) load(s + 1);
        const uint8_t* st = lds + (s & 1) * Cfg::STAGE;
        const uint8_t* ss = st + Cfg::CODE_BYTES;
        if (active) {
#pragma unroll
            for (int blk = 0; blk < MX_BLK_PER_STAGE; ++blk) {
                v8i bu[2];
                int sbu[2];
#pragma unroll
                for (int y = 0; y < 2; ++y) {
                    const int c = b_row + y * 32;
                    bu[y] = mx_read_operand<FMT>(st + c * Cfg::PITCH + (blk * 2 + h) * F::PB);
                    sbu[y] = ss[c * 4 + blk * 2 + h];
                }
#pragma unroll
                for (int x = 0; x < 2; ++x) {
                    const int r = a_row + x * 32;
                    const v8i au = mx_read_operand<FMT>(st + r * Cfg::PITCH + (blk * 2 + h) * F::PB);
                    v8i av = mx_read_operand<FMT>(st + r * Cfg::PITCH + (blk * 2 + 1 - h) * F::PB);
                    const int sau = ss[r * 4 + blk * 2 + h];
                    const int sav = ss[r * 4 + blk * 2 + 1 - h];
                    if (h) {
#pragma unroll
                        for (int d = 0; d < F::DW; ++d) av[d] ^= (int)F::neg(d);
                    }
#pragma unroll
                    for (int y = 0; y < 2; ++y) {
                        acc_re[x][y] = mx_mfma<FMT>(au, bu[y], acc_re[x][y], sau, sbu[y]);
                        acc_im[x][y] = mx_mfma<FMT>(av, bu[y], acc_im[x][y], sav, sbu[y]);
                    }
                }
            }
        }
        if (more) store(lds + ((s + 1) & 1) * Cfg::STAGE);
        __syncthreads();
    }

    if (!active) return;
    OutT* Cr = static_cast<OutT*>(p.Cre) + (int64_t)b * p.c_bstride;
    OutT* Ci = static_cast<OutT*>(p.Cim) + (int64_t)b * p.c_bstride;
    const int n = lane & 31, g4 = (lane >> 5) * 4;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int gj = wcol0 + y * 32 + n;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int gi = wrow0 + x * 32 + (r >> 2) * 8 + g4 + (r & 3);
                if (gi < p.M && gj < p.N) {
                    const int64_t idx = (int64_t)gi * p.ldc + gj;
                    float re = acc_re[x][y][r] * p.alpha, im = acc_im[x][y][r] * p.alpha;
                    if (p.beta != 0.0f) {
                        re = __builtin_fmaf(p.beta, (float)Cr[idx], re);
                        im = __builtin_fmaf(p.beta, (float)Ci[idx], im);
                    }
                    __builtin_nontemporal_store((OutT)re, Cr + idx);
                    __builtin_nontemporal_store((OutT)im, Ci + idx);
                }
            }
        }
}

template <int FMT>
hipError_t quantize_fmt(const void* X, const void* X2, int input, bool conj, uint8_t* codes,
                        uint8_t* scales, int64_t rows, int K, int nblkp, hipStream_t s) {
    const int64_t total = rows * nblkp;
    const unsigned blocks = (unsigned)((total + 255) / 256);
    const int cj = conj ? 1 : 0;
    if (input == MX_IN_INT4)
        hipLaunchKernelGGL((mx_quantize_kernel<FMT, MX_IN_INT4>), dim3(blocks), dim3(256), 0, s, X, X2,
                           codes, scales, rows, K, nblkp, cj);
    else if (input == MX_IN_FP16_PLANAR)
        hipLaunchKernelGGL((mx_quantize_kernel<FMT, MX_IN_FP16_PLANAR>), dim3(blocks), dim3(256), 0, s,
                           X, X2, codes, scales, rows, K, nblkp, cj);
    else
        hipLaunchKernelGGL((mx_quantize_kernel<FMT, MX_IN_FP16>), dim3(blocks), dim3(256), 0, s, X, X2,
                           codes, scales, rows, K, nblkp, cj);
    return hipGetLastError();
}

template <int FMT>
hipError_t gemm_fmt(const MxGemmArgs& a, bool out_fp16, hipStream_t s) {
    const int tiles_m = (a.M + MX_TILE - 1) / MX_TILE, tiles_n = (a.N + MX_TILE - 1) / MX_TILE;
    const int64_t tiles = (int64_t)tiles_m * tiles_n;
    const int64_t blocks = tiles * a.batch;
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    if (out_fp16)
        hipLaunchKernelGGL((gemm_mx_kernel<FMT, _Float16>), dim3((unsigned)blocks), dim3(MX_THREADS), 0,
                           s, a, tiles_m, (int)tiles);
    else
        hipLaunchKernelGGL((gemm_mx_kernel<FMT, float>), dim3((unsigned)blocks), dim3(MX_THREADS), 0, s,
                           a, tiles_m, (int)tiles);
    return hipGetLastError();
}

template <int FMT>
hipError_t herk_fmt(const MxHerkArgs& a, bool out_fp16, hipStream_t s) {
    const int nb = (a.N + MX_TILE - 1) / MX_TILE;
    const int tiles = nb * (nb + 1) / 2;
    const int64_t blocks = (int64_t)tiles * a.batch;
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    if (out_fp16)
        hipLaunchKernelGGL((herk_mx_kernel<FMT, _Float16>), dim3((unsigned)blocks), dim3(MX_THREADS),
                           0, s, a.codes, a.scales, a.C, a.N, a.nblkp, tiles, a.alpha, a.beta,
                           a.c_bstride);
    else
        hipLaunchKernelGGL((herk_mx_kernel<FMT, float>), dim3((unsigned)blocks), dim3(MX_THREADS), 0,
                           s, a.codes, a.scales, a.C, a.N, a.nblkp, tiles, a.alpha, a.beta,
                           a.c_bstride);
    return hipGetLastError();
}

}  // namespace

int mx_code_bytes(int fmt) { return fmt == MX_FP4_E2M1 ? MxFmt<MX_FP4_E2M1>::PB : MxFmt<MX_FP6_E3M2>::PB; }

int mx_padded_blocks(int K) {
    const int nblk = (K + 31) / 32;
    return (nblk + MX_BLK_PER_STAGE - 1) / MX_BLK_PER_STAGE * MX_BLK_PER_STAGE;
}

hipError_t launch_mx_quantize_ex(const void* X, const void* X2, int input, int fmt, bool conj,
                                 uint8_t* codes, uint8_t* scales, int64_t rows, int K, hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    const int nblkp = mx_padded_blocks(K);
    if (nblkp == 0) return hipSuccess;
    if (fmt == MX_FP4_E2M1)
        return quantize_fmt<MX_FP4_E2M1>(X, X2, input, conj, codes, scales, rows, K, nblkp, s);
    if (fmt == MX_FP6_E3M2)
        return quantize_fmt<MX_FP6_E3M2>(X, X2, input, conj, codes, scales, rows, K, nblkp, s);
    return hipErrorInvalidValue;
}

hipError_t launch_mx_quantize(const void* X, int input, int fmt, uint8_t* codes, uint8_t* scales,
                              int64_t rows, int K, hipStream_t s) {
    return launch_mx_quantize_ex(X, nullptr, input, fmt, false, codes, scales, rows, K, s);
}

hipError_t launch_gemm_mx(const MxGemmArgs& a, int fmt, bool out_fp16, hipStream_t s) {
    if (a.M <= 0 || a.N <= 0 || a.batch <= 0) return hipSuccess;
    if (fmt == MX_FP4_E2M1) return gemm_fmt<MX_FP4_E2M1>(a, out_fp16, s);
    if (fmt == MX_FP6_E3M2) return gemm_fmt<MX_FP6_E3M2>(a, out_fp16, s);
    return hipErrorInvalidValue;
}

hipError_t launch_herk_mx(const MxHerkArgs& a, int fmt, bool out_fp16, hipStream_t s) {
    if (a.N <= 0 || a.batch <= 0) return hipSuccess;
    if (fmt == MX_FP4_E2M1) return herk_fmt<MX_FP4_E2M1>(a, out_fp16, s);
    if (fmt == MX_FP6_E3M2) return herk_fmt<MX_FP6_E3M2>(a, out_fp16, s);
    return hipErrorInvalidValue;
}

}  // namespace ftdx
