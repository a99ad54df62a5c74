// herk_tile.h -- one lower-triangle tile of C = alpha * A A^H + beta * C on gfx950.
//
// Shared by the plain batched kernel and the grouped persistent kernel.
//
// Operand formulation (4 real products per complex MAC, the "4M" count the
// reference's direct kernel also executes, herk_kernel_common.hpp:29-32):
//   staged row a_i over 4 samples = [fp8 re0..re3 | fp8 im0..im3]  (u_i, two dwords)
//   conjugate partner             = [im0..im3 | -re0..-re3]        (v_i: dword swap + sign)
//   Re C_ij = u_i . u_j,   Im C_ij = v_i . u_j
// The MFMA sums over all 64 bytes of a 32x32x64 step, so any fixed byte order
// inside a lane is valid as long as A- and B-operands use the same one; the
// dword-planar order makes v_i four XORs per operand instead of byte permutes.
//
// Data flow per stage of KC complex samples:
//   global QC bytes (16 B / thread / segment) -> VGPR -> unpack to OCP E4M3 ->
//   LDS (double-buffered, one barrier per stage) -> ds_read_b128 -> MFMA.
#pragma once
#include "herk_common.h"
#include "herk_fp8.h"

namespace ftdx {

// Tile configuration: TM x TN complex tile, WAVES_M x WAVES_N waves, each wave a
// (TM/WAVES_M) x (TN/WAVES_N) block of 32x32 MFMA sub-tiles; KC samples per stage.
// WAVES_PER_EU is the occupancy target given to the compiler (caps VGPRs at
// 512 / WAVES_PER_EU); WG_PER_CU is how many workgroups fit one CU (LDS- and
// register-limited), used to size persistent grids.
template <int TM_, int TN_, int WAVES_M_, int WAVES_N_, int KC_, int WAVES_PER_EU_, int WG_PER_CU_,
          int DEPTH_ = 1>
struct TileCfg {
    static constexpr int TM = TM_, TN = TN_;
    // Global-load prefetch distance in stages: 1 = the next stage's loads are in
    // flight while this stage computes; 2 = two register sets, loads issued two
    // stages ahead (covers twice the latency for SEGS_PER_THREAD more VGPRs).
    static constexpr int DEPTH = DEPTH_;
    static constexpr int WAVES_M = WAVES_M_, WAVES_N = WAVES_N_;
    static constexpr int KC = KC_;
    static constexpr int WAVES_PER_EU = WAVES_PER_EU_;
    static constexpr int WG_PER_CU = WG_PER_CU_;
    static constexpr int WM = TM / WAVES_M, WN = TN / WAVES_N;
    static constexpr int FM = WM / 32, FN = WN / 32;
    static constexpr int THREADS = 64 * WAVES_M * WAVES_N;
    static constexpr int PITCH = 2 * KC + 16;         // LDS bytes per staged row (+16 B pad)
    static constexpr int STAGE = (TM + TN) * PITCH;   // row block then column block
    static constexpr int LDS_BYTES = 2 * STAGE;
    // Staged epilogue: each wave gathers 16 or 32 rows x WN columns of FP32 [re, im]
    // in its own LDS region.
    static constexpr bool EPI_STAGED =
        WN % 64 == 0 && WAVES_M * WAVES_N * 16 * (WN * 8 + 16) <= LDS_BYTES;
    static constexpr int SEGS_PER_ROW = KC / 16;      // 16-byte global segments per row
    static constexpr int SEGS = (TM + TN) * SEGS_PER_ROW;
    static constexpr int SEGS_PER_THREAD = SEGS / THREADS;
    static constexpr int SEGS8_PER_ROW = KC / 8;       // 16-byte segments of pre-cast FP8 rows
    static constexpr int SEGS8_PER_THREAD = (TM + TN) * SEGS8_PER_ROW / THREADS;
    static_assert(((TM + TN) * SEGS8_PER_ROW) % THREADS == 0, "FP8 segments must divide over threads");
    static_assert(WM % 32 == 0 && WN % 32 == 0, "wave block must be a multiple of 32");
    static_assert(KC % 32 == 0, "stage must hold whole 32-sample MFMA steps");
    static_assert(SEGS % THREADS == 0, "segments must divide evenly over threads");
    static_assert(TM % TN == 0, "row tile must be a multiple of the column tile");
    static_assert(LDS_BYTES * WG_PER_CU <= 160 * 1024, "LDS of resident workgroups exceeds 160 KB");
};

// s_waitcnt vmcnt(N) (other counters unconstrained), gfx9 encoding
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 0xF) | (0x7 << 4) | (0xF << 8) | (((N >> 4) & 0x3) << 14));
}

// LOAD_VEC16 / LOAD_BYTES: A holds QC bytes [N][K] (row pitch K), unpacked in
// the load stage.  LOAD_FP8: A holds pre-cast operand rows (row pitch lda bytes,
// 8 bytes per 4 samples in the dword-planar order), copied to LDS as they are.
// LOAD_BF6: pre-cast E3M2 rows with the same geometry (one 32-byte slot per 16
// samples: 12 bytes of Re codes, 12 of Im codes, 8 of padding), MFMAs in E3M2.
// LOAD_MX6 / LOAD_MX4: MX block-scaled operand rows (OPF_MX6 / OPF_MX4 in herk_common.h),
// the same 64 bytes per 32 samples, MFMAs scaled per lane.
// LOAD_MX4D: dense E2M1 rows (OPF_MX4D), interior tiles of 64-sample-stage configurations
// only: each 128-byte stage row holds 4 blocks of 32 samples (16 code bytes of Re, then Im),
// and TileJob::S holds the scales, [stage][row][8]: (Re, Im) bytes of the stage's 4 blocks.
enum LoadMode { LOAD_VEC16 = 0, LOAD_BYTES = 1, LOAD_FP8 = 2, LOAD_BF6 = 3, LOAD_MX6 = 4, LOAD_MX4 = 5,
                LOAD_MX4D = 6 };

struct TileJob {
    const uint8_t* A;  // this problem's operand rows
    void* C;           // this problem's packed triangle (scalars)
    int N, K;          // N: rows (HERK: also columns)
    int row0, col0;    // first row / column of the tile
    float alpha, beta;
    int64_t lda;       // row pitch of A in bytes (LOAD_FP8 only)
    // K segments (QC load modes): the product runs over nseg * K samples, segment s
    // being the rows at A + s * seg_stride (row pitch K).  nseg = 1: one matrix.
    int nseg;
    int64_t seg_stride;
    // GEMM layouts only (C = alpha A B^T + beta C, full M x Ncol grid): the column
    // operand's pre-cast rows and pitch, the column count, the imaginary output
    // plane (planar layout) and the output row pitch in elements.
    const uint8_t* B;
    int64_t ldb;
    int Ncol;
    void* C2;
    int64_t ldc;
    // OUT_GEMM_BEAM: samples per output and valid samples of this problem
    int beam_nt, beam_samples;
    // LOAD_MX4D: this problem's scales (row 0 of stage 0) and the bytes between stages
    const uint8_t* S;
    int64_t s_stage;
    // OUT_GRID: baseline UV (metres, packed lower index), wavelength (metres), grid size
    // and cell (radians); C is this batch's grid plane [Ng][Ng][2]
    const float* uv;
    double lambda;
    int Ng;
    float cell;
};

// 4 QC bytes (4 samples) -> u dwords: re = fp8(Re0..Re3), im = fp8(Im0..Im3).
__device__ __forceinline__ void qc4_to_fp8_planar(uint32_t x, uint32_t& re, uint32_t& im) {
    const uint32_t TLO = 0x44403800u;  // fp8(0..3)
    const uint32_t THI = 0x4E4C4A48u;  // fp8(4..7)
    uint32_t m_im = x & 0x07070707u;
    uint32_t m_re = (x >> 4) & 0x07070707u;
    im = __builtin_amdgcn_perm(THI, TLO, m_im) | ((x << 4) & 0x80808080u);
    re = __builtin_amdgcn_perm(THI, TLO, m_re) | (x & 0x80808080u);
}

// v from u for one 32-byte operand: [re | im] -> [im | -re] per dword pair.
__device__ __forceinline__ v8i conj_planar(v8i a) {
    v8i r;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        r[2 * q] = a[2 * q + 1];
        r[2 * q + 1] = a[2 * q] ^ (int)0x80808080u;
    }
    return r;
}

// The same for an E3M2 operand slot: dwords 0-2 hold the 16 six-bit Re codes, 3-5 the
// 16 Im codes (code k at bits 6k..6k+5 of the 96-bit half), 6-7 padding.  v = [im | -re]:
// swap the halves and flip bit 5 of every Re code.
__device__ __forceinline__ v8i conj_planar_bf6(v8i a) {
    v8i r;
    r[0] = a[3];
    r[1] = a[4];
    r[2] = a[5];
    r[3] = a[0] ^ (int)0x20820820u;
    r[4] = a[1] ^ (int)0x08208208u;
    r[5] = a[2] ^ (int)0x82082082u;
    r[6] = 0;
    r[7] = 0;
    return r;
}

template <int OPF>
__device__ __forceinline__ v8i conj_op(v8i a) {
    if constexpr (OPF == OPF_BF6) return conj_planar_bf6(a);
    else return conj_planar(a);
}

// Output layouts of one problem's Gram matrix (element (i, j) below is
// C_ij = sum_k a_ik conj(a_jk), the value the tile computes for j <= i):
//   OUT_LOWER_PACKED  C_ij at i(i+1)/2 + j                      (cutlass_gemm_api herk)
//   OUT_UPPER_PACKED  row q = j, column p = i >= q at q*N + p - q(q+1)/2, value C_ij:
//                     GGP's upper triangle of M^H M, row-major (tri_from_herm.cuh:28-73,
//                     xengine_test.cpp:282-287)
//   OUT_HERM_FULL     the full N x N matrix M^H M row-major: C_ij at j*N + i and
//                     conj(C_ij) at i*N + j (xengine_test.cpp:288-292)
// The GGP layouts keep their stores coalesced by swapping the MFMA operands: the
// accumulator then holds the transposed tile, so consecutive lanes own consecutive
// i, which is the contiguous index of both GGP layouts.  (enum OutFormat: herk_fp8.h)

template <typename OutT>
struct OutIO;

typedef float v2f __attribute__((ext_vector_type(2)));
typedef float v4f_a8 __attribute__((ext_vector_type(4), aligned(8)));
typedef uint32_t v2u_a4 __attribute__((ext_vector_type(2), aligned(4)));

// Output is written once and never re-read by the kernel: stores are
// non-temporal so the L2 keeps the re-read input rows instead.
template <>
struct OutIO<float> {
    static __device__ __forceinline__ void load(const void* C, int64_t idx, float& re, float& im) {
        v2f v = reinterpret_cast<const v2f*>(C)[idx];
        re = v[0];
        im = v[1];
    }
    static __device__ __forceinline__ void store(void* C, int64_t idx, float re, float im) {
        v2f v;
        v[0] = re;
        v[1] = im;
        __builtin_nontemporal_store(v, reinterpret_cast<v2f*>(C) + idx);
    }
    // elements idx and idx + 1 in one 16-byte access (8-byte aligned)
    static __device__ __forceinline__ void load2(const void* C, int64_t idx, float& r0, float& i0,
                                                 float& r1, float& i1) {
        const v4f_a8 v = *reinterpret_cast<const v4f_a8*>(reinterpret_cast<const v2f*>(C) + idx);
        r0 = v[0]; i0 = v[1]; r1 = v[2]; i1 = v[3];
    }
    static __device__ __forceinline__ void store2(void* C, int64_t idx, float r0, float i0, float r1,
                                                  float i1) {
        v4f_a8 v;
        v[0] = r0; v[1] = i0; v[2] = r1; v[3] = i1;
        __builtin_nontemporal_store(v, reinterpret_cast<v4f_a8*>(reinterpret_cast<v2f*>(C) + idx));
    }
};

template <>
struct OutIO<_Float16> {
    static __device__ __forceinline__ void load(const void* C, int64_t idx, float& re, float& im) {
        uint32_t w = reinterpret_cast<const uint32_t*>(C)[idx];
        uint16_t lo = (uint16_t)(w & 0xFFFFu), hi = (uint16_t)(w >> 16);
        _Float16 h0, h1;
        __builtin_memcpy(&h0, &lo, 2);
        __builtin_memcpy(&h1, &hi, 2);
        re = (float)h0;
        im = (float)h1;
    }
    static __device__ __forceinline__ void store(void* C, int64_t idx, float re, float im) {
        _Float16 h0 = (_Float16)re, h1 = (_Float16)im;  // round to nearest even
        uint16_t lo, hi;
        __builtin_memcpy(&lo, &h0, 2);
        __builtin_memcpy(&hi, &h1, 2);
        __builtin_nontemporal_store((uint32_t)lo | ((uint32_t)hi << 16),
                                    reinterpret_cast<uint32_t*>(C) + idx);
    }
    static __device__ __forceinline__ void load2(const void* C, int64_t idx, float& r0, float& i0,
                                                 float& r1, float& i1) {
        load(C, idx, r0, i0);
        load(C, idx + 1, r1, i1);
    }
    static __device__ __forceinline__ void store2(void* C, int64_t idx, float r0, float i0, float r1,
                                                  float i1) {
        _Float16 h[4] = {(_Float16)r0, (_Float16)i0, (_Float16)r1, (_Float16)i1};
        uint16_t u[4];
        __builtin_memcpy(u, h, 8);
        v2u_a4 v;
        v[0] = (uint32_t)u[0] | ((uint32_t)u[1] << 16);
        v[1] = (uint32_t)u[2] | ((uint32_t)u[3] << 16);
        __builtin_nontemporal_store(v, reinterpret_cast<v2u_a4*>(reinterpret_cast<uint32_t*>(C) + idx));
    }
};

typedef double v2d __attribute__((ext_vector_type(2)));

// FP64 output: the FP32 accumulators hold exact integers (|sum| < 2^24), so the
// widening is exact; alpha and beta are applied in double.
template <>
struct OutIO<double> {
    static __device__ __forceinline__ void load(const void* C, int64_t idx, double& re, double& im) {
        v2d v = reinterpret_cast<const v2d*>(C)[idx];
        re = v[0];
        im = v[1];
    }
    static __device__ __forceinline__ void store(void* C, int64_t idx, double re, double im) {
        v2d v;
        v[0] = re;
        v[1] = im;
        __builtin_nontemporal_store(v, reinterpret_cast<v2d*>(C) + idx);
    }
    static __device__ __forceinline__ void load2(const void* C, int64_t idx, double& r0, double& i0,
                                                 double& r1, double& i1) {
        load(C, idx, r0, i0);
        load(C, idx + 1, r1, i1);
    }
    static __device__ __forceinline__ void store2(void* C, int64_t idx, double r0, double i0,
                                                  double r1, double i1) {
        store(C, idx, r0, i0);
        store(C, idx + 1, r1, i1);
    }
};

__device__ __forceinline__ float fma_e(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_e(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename OutT>
struct EpiType { typedef float type; };
template <>
struct EpiType<double> { typedef double type; };

template <class Cfg>
struct TileRunner {
    // Staged-row source for segment slot q of thread tid: global row, LDS row.
    static __device__ __forceinline__ void seg_coords(int q, int tid, int& lrow, int& seg) {
        const int s = tid + Cfg::THREADS * q;
        lrow = s / Cfg::SEGS_PER_ROW;  // 0 .. TM+TN-1 (rows first, then columns)
        seg = s % Cfg::SEGS_PER_ROW;
    }

    // Pre-cast FP8 rows: segment slot q of thread tid covers 8 samples of one row.
    // GEMM: column rows come from j.B (pitch ldb, Ncol rows) instead of j.A.
    template <bool GEMM = false, bool BF6 = false>
    static __device__ __forceinline__ void load_stage_fp8(const TileJob& j, int k0, bool share_cols,
                                                          int tid, uint4 (&g)[Cfg::SEGS8_PER_THREAD]) {
        // the operand fields as values: selecting between j.A and j.B (or their pitches)
        // directly lets the compiler select between the fields' addresses instead, which
        // keeps the job in private memory
        const uint8_t* const A = j.A;
        const uint8_t* const B = GEMM ? j.B : j.A;
        const int64_t lda = j.lda, ldb = GEMM ? j.ldb : j.lda;
        const int nrow = j.N, ncol = GEMM ? j.Ncol : j.N;
#pragma unroll
        for (int q = 0; q < Cfg::SEGS8_PER_THREAD; ++q) {
            const int s = tid + Cfg::THREADS * q;
            const int lrow = s / Cfg::SEGS8_PER_ROW, seg = s % Cfg::SEGS8_PER_ROW;
            const bool is_col = lrow >= Cfg::TM;
            const int grow = is_col ? j.col0 + (lrow - Cfg::TM) : j.row0 + lrow;
            const int kk = k0 + seg * 8;
            const bool from_b = GEMM && is_col;
            const uint8_t* base = from_b ? B : A;
            const int64_t pitch = from_b ? ldb : lda;
            const int lim = from_b ? ncol : nrow;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            // pre-cast rows are zero-padded to a multiple of 16 samples; an E3M2 slot's
            // two 16-byte halves both carry all 16 of its samples
            const int kslot = BF6 ? (kk & ~15) : kk;
            if (!(is_col && share_cols) && grow < lim && kslot < j.K)
                v = *reinterpret_cast<const uint4*>(base + (int64_t)grow * pitch + (int64_t)kk * 2);
            g[q] = v;
        }
    }

    static __device__ __forceinline__ void store_stage_fp8(uint8_t* st, bool share_cols, int tid,
                                                           const uint4 (&g)[Cfg::SEGS8_PER_THREAD]) {
#pragma unroll
        for (int q = 0; q < Cfg::SEGS8_PER_THREAD; ++q) {
            const int s = tid + Cfg::THREADS * q;
            const int lrow = s / Cfg::SEGS8_PER_ROW, seg = s % Cfg::SEGS8_PER_ROW;
            if (lrow >= Cfg::TM && share_cols) continue;
            *reinterpret_cast<uint4*>(st + lrow * Cfg::PITCH + seg * 16) = g[q];
        }
    }

    template <int MODE>
    static __device__ __forceinline__ void load_stage(const TileJob& j, int k0, bool share_cols,
                                                      int tid, uint4 (&g)[Cfg::SEGS_PER_THREAD]) {
#pragma unroll
        for (int q = 0; q < Cfg::SEGS_PER_THREAD; ++q) {
            int lrow, seg;
            seg_coords(q, tid, lrow, seg);
            const bool is_col = lrow >= Cfg::TM;
            const int grow = is_col ? j.col0 + (lrow - Cfg::TM) : j.row0 + lrow;
            const int kk = k0 + seg * 16;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (!(is_col && share_cols) && grow < j.N) {
                if (j.nseg <= 1) {
                    const uint8_t* src = j.A + (int64_t)grow * j.K + kk;
                    if (MODE == LOAD_VEC16) {
                        if (kk < j.K) v = *reinterpret_cast<const uint4*>(src);
                    } else {
                        uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
                        for (int e = 0; e < 16; ++e)
                            if (kk + e < j.K) w[e >> 2] |= (uint32_t)src[e] << (8 * (e & 3));
                        v = make_uint4(w[0], w[1], w[2], w[3]);
                    }
                } else if (MODE == LOAD_VEC16) {
                    // K % 16 == 0 here, so a 16-byte segment lies inside one K segment
                    const int ps = kk / j.K, kin = kk - ps * j.K;
                    if (ps < j.nseg)
                        v = *reinterpret_cast<const uint4*>(j.A + ps * j.seg_stride +
                                                            (int64_t)grow * j.K + kin);
                } else {
                    uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
                    for (int e = 0; e < 16; ++e) {
                        const int kv = kk + e, ps = kv / j.K, kin = kv - ps * j.K;
                        if (ps < j.nseg)
                            w[e >> 2] |= (uint32_t)j.A[ps * j.seg_stride + (int64_t)grow * j.K + kin]
                                         << (8 * (e & 3));
                    }
                    v = make_uint4(w[0], w[1], w[2], w[3]);
                }
            }
            g[q] = v;
        }
    }

    static __device__ __forceinline__ void store_stage(uint8_t* st, bool share_cols, int tid,
                                                       const uint4 (&g)[Cfg::SEGS_PER_THREAD]) {
#pragma unroll
        for (int q = 0; q < Cfg::SEGS_PER_THREAD; ++q) {
            int lrow, seg;
            seg_coords(q, tid, lrow, seg);
            if (lrow >= Cfg::TM && share_cols) continue;
            uint4 lo, hi;
            qc4_to_fp8_planar(g[q].x, lo.x, lo.y);
            qc4_to_fp8_planar(g[q].y, lo.z, lo.w);
            qc4_to_fp8_planar(g[q].z, hi.x, hi.y);
            qc4_to_fp8_planar(g[q].w, hi.z, hi.w);
            uint8_t* dst = st + lrow * Cfg::PITCH + seg * 32;
            *reinterpret_cast<uint4*>(dst) = lo;
            *reinterpret_cast<uint4*>(dst + 16) = hi;
        }
    }

    static __device__ __forceinline__ v8i lds_read32(const uint8_t* p) {
        uint4 a = *reinterpret_cast<const uint4*>(p);
        uint4 b = *reinterpret_cast<const uint4*>(p + 16);
        v8i r;
        r[0] = (int)a.x; r[1] = (int)a.y; r[2] = (int)a.z; r[3] = (int)a.w;
        r[4] = (int)b.x; r[5] = (int)b.y; r[6] = (int)b.z; r[7] = (int)b.w;
        return r;
    }

    // Runs one tile; all threads of the workgroup must call it (it uses barriers).
    // Ends with a barrier, so the caller may reuse `lds` right away.
    template <typename OutT>
    static __device__ __forceinline__ void put(const TileJob& j, int64_t idx, float are, float aim) {
        typedef typename EpiType<OutT>::type E;
        const E alpha = (E)j.alpha, beta = (E)j.beta;
        E re = (E)are * alpha, im = (E)aim * alpha;
        if (beta != (E)0) {
            E ore, oim;
            OutIO<OutT>::load(j.C, idx, ore, oim);
            re = fma_e(beta, ore, re);  // one rounding, as the oracle's fma(beta, old, .)
            im = fma_e(beta, oim, im);
        }
        OutIO<OutT>::store(j.C, idx, re, im);
    }

    // put() of elements idx and idx + 1 with one load / store each.
    template <typename OutT>
    static __device__ __forceinline__ void put2(const TileJob& j, int64_t idx, float r0, float i0,
                                                float r1, float i1) {
        typedef typename EpiType<OutT>::type E;
        const E alpha = (E)j.alpha, beta = (E)j.beta;
        E a0 = (E)r0 * alpha, b0 = (E)i0 * alpha, a1 = (E)r1 * alpha, b1 = (E)i1 * alpha;
        if (beta != (E)0) {
            E o0, p0, o1, p1;
            OutIO<OutT>::load2(j.C, idx, o0, p0, o1, p1);
            a0 = fma_e(beta, o0, a0);
            b0 = fma_e(beta, p0, b0);
            a1 = fma_e(beta, o1, a1);
            b1 = fma_e(beta, p1, b1);
        }
        OutIO<OutT>::store2(j.C, idx, a0, b0, a1, b1);
    }

    // Lower-packed epilogue through LDS.  Packed-triangle rows start at arbitrary
    // 8-byte offsets, so the accumulator layout's stores (two 256-byte runs of two
    // rows per instruction) leave partial cache lines at every run end.  Here each
    // wave gathers RP rows of its own block in a private LDS region (the stage
    // buffers are free after the main loop) and writes each row's WN-column run as
    // one 512-byte instruction.  No cross-wave barrier until the final one, which
    // lets the caller reuse `lds`.
    template <typename OutT>
    static __device__ __forceinline__ void epilogue_staged(
        const TileJob& j, uint8_t* __restrict__ lds, const v16f (&acc_re)[Cfg::FM][Cfg::FN],
        const v16f (&acc_im)[Cfg::FM][Cfg::FN], bool active, int wrow0, int wcol0, int lane,
        int wave) {
        constexpr int NW = Cfg::WAVES_M * Cfg::WAVES_N;
        constexpr int EPW = Cfg::WN * 8 + 16;
        constexpr int HS = (NW * 32 * EPW <= Cfg::LDS_BYTES) ? 1 : 2;  // passes per 32-row block
        constexpr int RP = 32 / HS;
        static_assert(NW * RP * EPW <= Cfg::LDS_BYTES, "wave epilogue regions exceed LDS");
        static_assert(Cfg::WN % 64 == 0, "row runs are written 64 columns at a time");
        uint8_t* region = lds + wave * (RP * EPW);
        const int n = lane & 31, g4 = (lane >> 5) * 4;
        if (active) {
#pragma unroll
            for (int a = 0; a < Cfg::FM; ++a)
#pragma unroll
                for (int h = 0; h < HS; ++h) {
#pragma unroll
                    for (int b = 0; b < Cfg::FN; ++b)
#pragma unroll
                        for (int r = h * (16 / HS); r < (h + 1) * (16 / HS); ++r) {
                            const int lrow = (r >> 2) * 8 + g4 + (r & 3) - h * RP;
                            v2f v;
                            v[0] = acc_re[a][b][r];
                            v[1] = acc_im[a][b][r];
                            *reinterpret_cast<v2f*>(region + lrow * EPW + (b * 32 + n) * 8) = v;
                        }
                    __builtin_amdgcn_wave_barrier();
                    // two rows per instruction: lanes 0-31 row lr0, 32-63 row lr0 + 1, each
                    // lane two consecutive columns (one 16-byte LDS read, one store)
                    const int hl = lane >> 5, cp = (lane & 31) * 2;
#pragma unroll 1
                    for (int lr0 = 0; lr0 < RP; lr0 += 2) {
                        const int lr = lr0 + hl;
                        const int gi = wrow0 + a * 32 + h * RP + lr;
                        const int64_t rbase = tri_index(gi, 0);
#pragma unroll
                        for (int c0 = 0; c0 < Cfg::WN; c0 += 64) {
                            const int c = c0 + cp;
                            const int gj = wcol0 + c;
                            if (gi < j.N && gj <= gi) {
                                const v4f_a8 v = *reinterpret_cast<const v4f_a8*>(region + lr * EPW + c * 8);
                                if (gj < gi)  // (gi, gj) off the diagonal; (gi, gj + 1) may be on it
                                    put2<OutT>(j, rbase + gj, v[0], v[1], v[2], gj + 1 == gi ? 0.0f : v[3]);
                                else
                                    put<OutT>(j, rbase + gj, v[0], 0.0f);
                            }
                        }
                    }
                    __builtin_amdgcn_wave_barrier();
                }
        }
        __syncthreads();
    }

    // GEMM-layout epilogue: element (i, j) of C = alpha A B^T + beta C at i*ldc + j.
    template <typename OutT, int FMT>
    static __device__ __forceinline__ void put_gemm(const TileJob& j, int64_t idx, float re, float im) {
        if constexpr (FMT == OUT_GEMM_POWER) {
            // power = alpha (re^2 + im^2) [+ beta old]: the reference's expression
            // (cutlass_gemm_api.cu:425-437), evaluated as fma(re, re, im * im)
            float p = j.alpha * __builtin_fmaf(re, re, im * im);
            float* P = static_cast<float*>(j.C);
            if (j.beta != 0.0f) p = __builtin_fmaf(j.beta, P[idx], p);
            __builtin_nontemporal_store(p, P + idx);
        } else {
            OutT* Cr = static_cast<OutT*>(j.C);
            OutT* Ci = static_cast<OutT*>(j.C2);
            float r = re * j.alpha, m = im * j.alpha;
            if (j.beta != 0.0f) {
                r = __builtin_fmaf(j.beta, (float)Cr[idx], r);
                m = __builtin_fmaf(j.beta, (float)Ci[idx], m);
            }
            __builtin_nontemporal_store((OutT)r, Cr + idx);
            __builtin_nontemporal_store((OutT)m, Ci + idx);
        }
    }

    // Beamformer epilogue (OUT_GEMM_BEAM).  Tile rows 0-63 are samples m0 .. m0 + 63 of
    // polarization 0 and rows 64-127 the same samples of polarization 1 (m0 = 64 *
    // row block); columns are beams.  Each element's power fma(re, re, im * im) goes to
    // LDS; then each output (beam, group of nt samples) is
    //     acc = 0; for dt: acc += (P0[m] + P1[m]),   m = m0 + g * nt + dt
    // in that order, the reference's two power GEMMs with beta = 1 for pol 1 followed by
    // its time integration (lib/voltage_pipeline.cu:283-318, 934-967), and is written
    // to out[beam][m / nt] as FP32 acc or as the uint8 rint(acc * scale) clamped to
    // [0, 255] (NaN -> 0, as __float2int_rn).  Groups lie inside one tile (nt divides 64).
    template <typename OutT>
    static __device__ __forceinline__ void epilogue_beam(
        const TileJob& j, uint8_t* __restrict__ lds, const v16f (&acc_re)[Cfg::FM][Cfg::FN],
        const v16f (&acc_im)[Cfg::FM][Cfg::FN], bool active, int wrow0, int wcol0, int lane, int tid) {
        static_assert(Cfg::TM == 128, "two 64-sample polarization halves per row block");
        constexpr int LP = Cfg::TN + 8;  // floats per LDS row: rows r and r + 4 on disjoint banks
        static_assert(Cfg::TM * LP * 4 <= Cfg::LDS_BYTES, "power tile exceeds LDS");
        float* P = reinterpret_cast<float*>(lds);
        __syncthreads();  // every wave is done with the stage buffers
        if (active) {
            const int n = lane & 31, g4 = (lane >> 5) * 4;
#pragma unroll
            for (int a = 0; a < Cfg::FM; ++a)
#pragma unroll
                for (int b = 0; b < Cfg::FN; ++b) {
                    const int lc = wcol0 - j.col0 + b * 32 + n;
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int lr = wrow0 - j.row0 + a * 32 + (r >> 2) * 8 + g4 + (r & 3);
                        const float re = acc_re[a][b][r], im = acc_im[a][b][r];
                        P[lr * LP + lc] = __builtin_fmaf(re, re, im * im);
                    }
                }
        }
        __syncthreads();
        const int nt = j.beam_nt, groups = 64 / nt;
        const int m0 = (j.row0 >> 7) * 64;
        const int col = tid % Cfg::TN;
        const int beam = j.col0 + col;
        if (beam >= j.Ncol) return;
        OutT* out = static_cast<OutT*>(j.C) + (int64_t)beam * j.ldc + m0 / nt;
        for (int g = tid / Cfg::TN; g < groups; g += Cfg::THREADS / Cfg::TN) {
            if (m0 + g * nt >= j.beam_samples) break;
            const float* p = P + g * nt * LP + col;
            float acc = 0.0f;
            for (int dt = 0; dt < nt; ++dt) acc += p[dt * LP] + p[(dt + 64) * LP];
            if constexpr (sizeof(OutT) == 1) {
                const float q = __builtin_fminf(__builtin_fmaxf(__builtin_rintf(acc * j.alpha), 0.0f), 255.0f);
                out[g] = (OutT)(int)q;
            } else {
                out[g] = acc;
            }
        }
    }

    // Gridding epilogue (OUT_GRID): the pillbox scatter of the reference
    // (include/kernels/visibility_bf.cuh:269-370) applied to each element straight from
    // the accumulators, so the triangle never reaches HBM.  Element (i, j <= i) is
    // baseline i(i+1)/2 + j; UV in wavelengths = float(uv_m / lambda) (double division);
    // cell = roundf(u / cell + Ng/2); the visibility alpha * C_ij (Im = 0 on the diagonal)
    // is added there and, for i != j, its conjugate at (-u, -v).
    static __device__ __forceinline__ void epilogue_grid(const TileJob& j, const v16f (&acc_re)[Cfg::FM][Cfg::FN],
                                                         const v16f (&acc_im)[Cfg::FM][Cfg::FN], int wrow0,
                                                         int wcol0, int lane) {
        float* grid = static_cast<float*>(j.C);
        const float half_ng = 0.5f * j.Ng;
        const int n = lane & 31, g4 = (lane >> 5) * 4;
#pragma unroll
        for (int a = 0; a < Cfg::FM; ++a)
#pragma unroll
            for (int b = 0; b < Cfg::FN; ++b) {
                const int gj = wcol0 + b * 32 + n;
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int gi = wrow0 + a * 32 + (r >> 2) * 8 + g4 + (r & 3);
                    if (gi >= j.N || gj > gi) continue;
                    const float vre = acc_re[a][b][r] * j.alpha;
                    const float vim = gi == gj ? 0.0f : acc_im[a][b][r] * j.alpha;
                    const int64_t k = tri_index(gi, gj);
                    const float u_lam = (float)((double)j.uv[2 * k] / j.lambda);
                    const float v_lam = (float)((double)j.uv[2 * k + 1] / j.lambda);
                    const int iu = (int)__builtin_roundf(u_lam / j.cell + half_ng);
                    const int iv = (int)__builtin_roundf(v_lam / j.cell + half_ng);
                    if (iu >= 0 && iu < j.Ng && iv >= 0 && iv < j.Ng) {
                        float* g = grid + 2 * ((int64_t)iv * j.Ng + iu);
                        unsafeAtomicAdd(g, vre);
                        unsafeAtomicAdd(g + 1, vim);
                    }
                    if (gi != gj) {
                        const int iuc = (int)__builtin_roundf(-u_lam / j.cell + half_ng);
                        const int ivc = (int)__builtin_roundf(-v_lam / j.cell + half_ng);
                        if (iuc >= 0 && iuc < j.Ng && ivc >= 0 && ivc < j.Ng) {
                            float* g = grid + 2 * ((int64_t)ivc * j.Ng + iuc);
                            unsafeAtomicAdd(g, vre);
                            unsafeAtomicAdd(g + 1, -vim);
                        }
                    }
                }
            }
    }

    // One stage of MFMAs on the staged rows in `st` (KC samples, all 32-sample steps).
    template <int FMT, int OPF = OPF_FP8>
    static __device__ __forceinline__ void mfma_stage(const uint8_t* st, int a_off, int b_off,
                                                      v16f (&acc_re)[Cfg::FM][Cfg::FN],
                                                      v16f (&acc_im)[Cfg::FM][Cfg::FN]) {
#pragma unroll
        for (int ks = 0; ks < Cfg::KC / 32; ++ks) {
            v8i au[Cfg::FM], bu[Cfg::FN];
#pragma unroll
            for (int a = 0; a < Cfg::FM; ++a)
                au[a] = lds_read32(st + a_off + a * 32 * Cfg::PITCH + ks * 64);
#pragma unroll
            for (int b = 0; b < Cfg::FN; ++b)
                bu[b] = lds_read32(st + b_off + b * 32 * Cfg::PITCH + ks * 64);
            if constexpr (opf_mx(OPF)) {
                static_assert(FMT == OUT_LOWER_PACKED || gemm_layout(FMT), "MX rows: row-major layouts");
                // the Im partner is the block's other part, 32 bytes away
                const bool hi = (threadIdx.x & 32) != 0;
                const int dv = hi ? -32 : 32;
#pragma unroll
                for (int a = 0; a < Cfg::FM; ++a) {
                    const v8i av =
                        mx_partner<OPF>(lds_read32(st + a_off + dv + a * 32 * Cfg::PITCH + ks * 64), hi);
#pragma unroll
                    for (int b = 0; b < Cfg::FN; ++b) {
                        acc_re[a][b] = mfma_32x32x64_mx<OPF>(au[a], bu[b], acc_re[a][b]);
                        acc_im[a][b] = mfma_32x32x64_mx<OPF>(av, bu[b], acc_im[a][b]);
                    }
                }
                continue;
            }
#pragma unroll
            for (int a = 0; a < Cfg::FM; ++a) {
                const v8i av = conj_op<OPF>(au[a]);
#pragma unroll
                for (int b = 0; b < Cfg::FN; ++b) {
                    if constexpr (FMT == OUT_LOWER_PACKED || FMT == OUT_GRID || gemm_layout(FMT)) {
                        acc_re[a][b] = mfma_32x32x64<OPF>(au[a], bu[b], acc_re[a][b]);
                        acc_im[a][b] = mfma_32x32x64<OPF>(av, bu[b], acc_im[a][b]);
                    } else {  // transposed accumulator, same values
                        acc_re[a][b] = mfma_32x32x64<OPF>(bu[b], au[a], acc_re[a][b]);
                        acc_im[a][b] = mfma_32x32x64<OPF>(bu[b], av, acc_im[a][b]);
                    }
                }
            }
        }
    }

    // ---- Branch-free tiles of the fused QC path ----
    // K is a whole number of stages and the product runs over one or two K segments (a
    // polarization pair: a stage never straddles them), or one segment has a partial last
    // stage (K a multiple of 16, full_tile_loop's TAIL).  The loads then need no bounds
    // checks, so each stage is one basic block -- global loads for a later stage, this
    // stage's ds_read + MFMA, the next stage's unpack + ds_write -- which the scheduler
    // can interleave.  Edge tiles (rows past N) run it too: full_offsets reads row N - 1
    // in place of each missing row, whose products land only in elements (i >= N or
    // j > i) that no epilogue stores.  A diagonal tile loads its column block as well (the
    // same rows as part of its row block) to keep the body uniform; the MFMAs read the
    // row copy.
    // TAILS: the launch's K is not a whole number of stages (kernels built for that case)
    template <bool TAILS>
    static __device__ __forceinline__ bool full_tile(const TileJob& j) {
        const bool stages = TAILS ? (j.nseg == 1 && j.K % 16 == 0 && j.K % Cfg::KC != 0)
                                  : (j.nseg >= 1 && j.nseg <= 2 && j.K % Cfg::KC == 0);
        return stages && j.K > 0 && j.N > 0 && (int64_t)j.N * j.K <= 0xFFFFFFFFll;
    }

    // Byte offset of each 16-byte segment of this thread inside one problem's rows (a row
    // past N reads row N - 1).
    static __device__ __forceinline__ void full_offsets(const TileJob& j, int tid,
                                                        uint32_t (&off)[Cfg::SEGS_PER_THREAD]) {
#pragma unroll
        for (int q = 0; q < Cfg::SEGS_PER_THREAD; ++q) {
            int lrow, seg;
            seg_coords(q, tid, lrow, seg);
            const int want = lrow >= Cfg::TM ? j.col0 + (lrow - Cfg::TM) : j.row0 + lrow;
            const int grow = want < j.N ? want : j.N - 1;
            off[q] = (uint32_t)grow * (uint32_t)j.K + (uint32_t)(seg * 16);
        }
    }

    static __device__ __forceinline__ void full_load(const uint8_t* base,
                                                     const uint32_t (&off)[Cfg::SEGS_PER_THREAD],
                                                     uint4 (&g)[Cfg::SEGS_PER_THREAD]) {
#pragma unroll
        for (int q = 0; q < Cfg::SEGS_PER_THREAD; ++q)
            g[q] = *reinterpret_cast<const uint4*>(base + off[q]);
    }

    static __device__ __forceinline__ void full_store(uint8_t* st, int tid,
                                                      const uint4 (&g)[Cfg::SEGS_PER_THREAD]) {
#pragma unroll
        for (int q = 0; q < Cfg::SEGS_PER_THREAD; ++q) {
            int lrow, seg;
            seg_coords(q, tid, lrow, seg);
            uint4 lo, hi;
            qc4_to_fp8_planar(g[q].x, lo.x, lo.y);
            qc4_to_fp8_planar(g[q].y, lo.z, lo.w);
            qc4_to_fp8_planar(g[q].z, hi.x, hi.y);
            qc4_to_fp8_planar(g[q].w, hi.z, hi.w);
            uint8_t* dst = st + lrow * Cfg::PITCH + seg * 32;
            *reinterpret_cast<uint4*>(dst) = lo;
            *reinterpret_cast<uint4*>(dst + 16) = hi;
        }
    }

    // TAIL (one K segment, K a multiple of 16 but not of KC): the last stage holds
    // K - (nst - 1) * KC samples.  Its 16-byte segments past K load the same row's first
    // segment of that stage instead (in bounds) and are zeroed before the unpack, so they
    // add nothing; both choices are selects on a uniform condition, and every stage stays
    // one basic block.
    template <int FMT, bool ACTIVE, int DEPTH, bool TAIL>
    static __device__ __forceinline__ void full_tile_loop(const TileJob& j, uint8_t* __restrict__ lds,
                                                          int tid, int nst, int a_off, int b_off,
                                                          v16f (&acc_re)[Cfg::FM][Cfg::FN],
                                                          v16f (&acc_im)[Cfg::FM][Cfg::FN]) {
        constexpr int NG = Cfg::SEGS_PER_THREAD;
        uint32_t off[NG];
        full_offsets(j, tid, off);
        // stage s's first sample; past the end the last stage is loaded again (in
        // bounds, never stored where it is read).  With two K segments, stages from sps
        // on are in the second one, seg_stride bytes after the first.
        const int sps = j.nseg > 1 ? j.K / Cfg::KC : nst;
        const int64_t jump = j.seg_stride - (int64_t)j.K;
        auto src = [&](int s) {
            const int t = s < nst ? s : nst - 1;
            return j.A + (int64_t)t * Cfg::KC + (t >= sps ? jump : 0);
        };
        // A thread's segments all sit at the same sample of their rows (THREADS is a
        // multiple of the segments per row): (tid % SEGS_PER_ROW) * 16.  In the last stage
        // a segment past K moves back by that much and is zeroed.
        static_assert(Cfg::THREADS % Cfg::SEGS_PER_ROW == 0, "segment position must not depend on q");
        uint32_t tback = 0u, tkeep = 0xFFFFFFFFu;
        if constexpr (TAIL) {
            const int kt = j.K - (nst - 1) * Cfg::KC;  // samples of the last stage
            const int seg = tid % Cfg::SEGS_PER_ROW;
            const bool v = seg * 16 < kt;
            tback = v ? 0u : (uint32_t)(seg * 16);
            tkeep = v ? 0xFFFFFFFFu : 0u;
        }
        auto load = [&](int s, uint4 (&g)[NG]) {
            if constexpr (TAIL) {
                const uint32_t back = s >= nst - 1 ? tback : 0u;
                uint32_t o[NG];
#pragma unroll
                for (int q = 0; q < NG; ++q) o[q] = off[q] - back;
                full_load(src(s), o, g);
            } else {
                full_load(src(s), off, g);
            }
        };
        // unpack + write stage s (held in g)
        auto store = [&](uint8_t* st, uint4 (&g)[NG], int s) {
            if constexpr (TAIL) {
                const uint32_t m = s >= nst - 1 ? tkeep : 0xFFFFFFFFu;
#pragma unroll
                for (int q = 0; q < NG; ++q) {
                    g[q].x &= m;
                    g[q].y &= m;
                    g[q].z &= m;
                    g[q].w &= m;
                }
            }
            full_store(st, tid, g);
        };
        if constexpr (DEPTH == 2) {
            uint4 ga[NG], gb[NG];
            load(0, ga);
            store(lds, ga, 0);
            load(1, ga);
            __syncthreads();
            for (int s = 0; s < nst; s += 2) {
                load(s + 2, gb);
                if constexpr (ACTIVE) mfma_stage<FMT>(lds, a_off, b_off, acc_re, acc_im);
                store(lds + Cfg::STAGE, ga, s + 1);
                __syncthreads();
                if (s + 1 >= nst) break;
                load(s + 3, ga);
                if constexpr (ACTIVE) mfma_stage<FMT>(lds + Cfg::STAGE, a_off, b_off, acc_re, acc_im);
                store(lds, gb, s + 2);
                __syncthreads();
            }
        } else {
            uint4 g[NG];
            load(0, g);
            store(lds, g, 0);
            __syncthreads();
            for (int s = 0; s < nst; ++s) {
                load(s + 1, g);
                if constexpr (ACTIVE)
                    mfma_stage<FMT>(lds + (s & 1) * Cfg::STAGE, a_off, b_off, acc_re, acc_im);
                store(lds + ((s + 1) & 1) * Cfg::STAGE, g, s + 1);
                __syncthreads();
            }
        }
    }

    // ---- Interior tiles of pre-cast rows: global -> LDS without registers ----
    // Stage layout: TM + TN rows of 2*KC bytes, unpadded; 16-byte chunk c of row r is
    // stored at chunk position c ^ f(r).  The LDS serves a ds_read_b128 16 lanes (16
    // rows of one operand) at a time from 64 four-byte banks, i.e. 16 chunk slots per
    // 256 bytes; f(r) = (r / RPP) mod DCPR, RPP rows sharing one 256-byte bank period,
    // gives those 16 rows 16 distinct slots.  One global_load_lds_dwordx4 per wave fills 1 KB of
    // consecutive LDS (lane L -> byte 16 L); each lane fetches the global chunk that
    // belongs at its position.  No VGPR staging, no ds_write, no unpack VALU.
    static constexpr int DROW = 2 * Cfg::KC;                   // bytes per staged row
    static constexpr int DCPR = DROW / 16;                     // 16-byte chunks per row
    static constexpr int RPP = DROW < 256 ? 256 / DROW : 1;    // rows per 256-byte bank period
    static __device__ __forceinline__ int swz(int row) { return (row / RPP) & (DCPR - 1); }
    static constexpr int DSTAGE = (Cfg::TM + Cfg::TN) * DROW;  // bytes per stage
    static constexpr int NWAVES = Cfg::WAVES_M * Cfg::WAVES_N;
    static constexpr int DOPS = DSTAGE / (1024 * NWAVES);      // loads per wave per stage
    static constexpr bool DMA_OK =
        DSTAGE % (1024 * NWAVES) == 0 && 2 * DSTAGE <= Cfg::LDS_BYTES && DCPR * RPP == 16;

    // Column rows: HERK reads them from A (pitch lda), the GEMM layouts from B (pitch ldb,
    // Ncol rows).  Every load instruction covers 1024 / DROW whole rows of one block.
    // Edge tiles (rows or columns past the matrix) take the direct loads too: dma_offsets
    // reads the block's last row in place of each missing one, and those rows feed only
    // outputs that every epilogue leaves unstored (rows >= N, columns >= Ncol or j > i).
    // EDGES = false keeps to interior tiles (dense E2M1 rows read their scales per block).
    template <bool GEMM, bool EDGES = true>
    static __device__ __forceinline__ bool dma_tile(const TileJob& j) {
        const uint8_t* B = GEMM ? j.B : j.A;
        const int64_t ldb = GEMM ? j.ldb : j.lda;
        const int ncol = GEMM ? j.Ncol : j.N;
        const bool inside = j.row0 + Cfg::TM <= j.N && j.col0 + Cfg::TN <= ncol;
        return j.K > 0 && j.K % Cfg::KC == 0 && (EDGES || inside) && j.N > 0 && ncol > 0 &&
               ((reinterpret_cast<uintptr_t>(j.A) | reinterpret_cast<uintptr_t>(B)) & 15u) == 0 &&
               j.lda % 16 == 0 && ldb % 16 == 0 && (int64_t)j.N * j.lda <= 0xFFFFFFFFll &&
               (int64_t)ncol * ldb <= 0xFFFFFFFFll;
    }
    static_assert(Cfg::TM % (1024 / DROW) == 0, "a load instruction must not straddle the blocks");
    static __device__ __forceinline__ bool dma_op_is_col(int q, int wave) {
        return (q * NWAVES + wave) * (1024 / DROW) >= Cfg::TM;
    }

    // Global byte offset (inside its block's rows, stage 0) of each load of this lane; a
    // row past the matrix (edge tiles) reads the last row instead.
    template <bool GEMM>
    static __device__ __forceinline__ void dma_offsets(const TileJob& j, int lane, int wave,
                                                       uint32_t (&off)[DOPS]) {
        const int64_t ldb = GEMM ? j.ldb : j.lda;
        const int ncol = GEMM ? j.Ncol : j.N;
#pragma unroll
        for (int q = 0; q < DOPS; ++q) {
            const int byte = (q * NWAVES + wave) * 1024 + lane * 16;  // position in the stage
            const int row = byte / DROW, pc = (byte % DROW) / 16;
            const int c = pc ^ swz(row);                               // logical chunk held there
            const int ra = j.row0 + row < j.N ? j.row0 + row : j.N - 1;
            const int rb = j.col0 + row - Cfg::TM < ncol ? j.col0 + row - Cfg::TM : ncol - 1;
            off[q] = row < Cfg::TM ? (uint32_t)ra * (uint32_t)j.lda + (uint32_t)(c * 16)
                                   : (uint32_t)rb * (uint32_t)ldb + (uint32_t)(c * 16);
        }
    }

    // baseA / baseB: this stage's first byte of the row / column operand rows
    static __device__ __forceinline__ void dma_stage(const uint8_t* baseA, const uint8_t* baseB,
                                                     uint8_t* st, int wave, const uint32_t (&off)[DOPS]) {
#pragma unroll
        for (int q = 0; q < DOPS; ++q)
            __builtin_amdgcn_global_load_lds(
                (const void*)((dma_op_is_col(q, wave) ? baseB : baseA) + off[q]),
                (void __attribute__((address_space(3)))*)(st + (q * NWAVES + wave) * 1024), 16, 0, 0);
    }

    // One 32-byte part of an MX block on the swizzled layout (first 16 bytes at p0, rest at
    // p1): only the dwords the scaled MFMA reads -- the codes (4 for E2M1, 6 for E3M2) and
    // the scale (dword 4 / 6) -- and not the zero padding behind them.
    template <int OPF>
    static __device__ __forceinline__ v8i lds_mx_part(const uint8_t* r, int p0, int p1) {
        const uint4 x = *reinterpret_cast<const uint4*>(r + p0);
        v8i v;
        v[0] = (int)x.x; v[1] = (int)x.y; v[2] = (int)x.z; v[3] = (int)x.w;
        v[5] = 0; v[6] = 0; v[7] = 0;
        if constexpr (OPF == OPF_MX6) {
            const uint2 y = *reinterpret_cast<const uint2*>(r + p1);
            v[4] = (int)y.x; v[5] = (int)y.y;
            v[6] = (int)*reinterpret_cast<const uint32_t*>(r + p1 + 8);
        } else {
            v[4] = (int)*reinterpret_cast<const uint32_t*>(r + p1);
        }
        return v;
    }

    // ---- Dense E2M1 rows: the stage's scales, 8 bytes per row, beside the codes ----
    static constexpr int SC_ROW = 8;                               // scale bytes per row per stage
    static constexpr int SC_STAGE = (Cfg::TM + Cfg::TN) * SC_ROW;  // row block, then column block
    static constexpr int SC_OPS = SC_STAGE / 1024;                 // 1 KB loads per stage
    static constexpr bool MX4D_OK = Cfg::KC == 64 && (Cfg::TM * SC_ROW) % 1024 == 0 &&
                                    (Cfg::TN * SC_ROW) % 1024 == 0 && SC_OPS <= NWAVES &&
                                    2 * DSTAGE + 2 * SC_STAGE <= Cfg::LDS_BYTES;

    // Waves 0 .. SC_OPS-1 each load 1 KB of the stage's scales (rows' 8 bytes are
    // contiguous per stage in S, so the row block and the column block are two runs).
    static __device__ __forceinline__ void dma_scales(const TileJob& j, int stage, uint8_t* sc, int wave,
                                                      int lane) {
        if (wave < SC_OPS) {
            const int byte = wave * 1024 + lane * 16;
            const uint8_t* base = j.S + (int64_t)stage * j.s_stage;
            const uint8_t* src = byte < Cfg::TM * SC_ROW ? base + (int64_t)j.row0 * SC_ROW + byte
                                                         : base + (int64_t)j.col0 * SC_ROW + (byte - Cfg::TM * SC_ROW);
            __builtin_amdgcn_global_load_lds((const void*)src,
                                             (void __attribute__((address_space(3)))*)(sc + wave * 1024), 16, 0,
                                             0);
        }
    }

    // E8M0 byte `idx` (0 .. 7) of a row's 8 scale bytes, as the scale operand's low byte
    static __device__ __forceinline__ int scale_byte(uint2 v, int idx) {
        return (int)(((idx < 4 ? v.x : v.y) >> (8 * (idx & 3))) & 0xFFu);
    }

    // One stage of dense E2M1 MFMAs: 4 blocks; lanes 0-31 hold the Re part of a block and
    // lanes 32-63 the Im part, the Im product's partner is the other part (16 bytes away
    // before the swizzle) with Re negated, as on the MX rows.
    template <int FMT>
    static __device__ __forceinline__ void mfma_stage_mx4d(const uint8_t* st, const uint8_t* sc, int a_row,
                                                           int b_row, int lane, v16f (&acc_re)[Cfg::FM][Cfg::FN],
                                                           v16f (&acc_im)[Cfg::FM][Cfg::FN]) {
        static_assert(FMT == OUT_LOWER_PACKED, "dense E2M1 rows: packed HERK output");
        const int sw = swz(lane & 31), half = lane >> 5;
        uint2 sa[Cfg::FM], sb[Cfg::FN];
#pragma unroll
        for (int a = 0; a < Cfg::FM; ++a) sa[a] = *reinterpret_cast<const uint2*>(sc + (a_row + a * 32) * SC_ROW);
#pragma unroll
        for (int b = 0; b < Cfg::FN; ++b) sb[b] = *reinterpret_cast<const uint2*>(sc + (b_row + b * 32) * SC_ROW);
#pragma unroll
        for (int blk = 0; blk < 4; ++blk) {
            const int c = blk * 2 + half;  // this lane's part of the block
            const int po = (c ^ sw) * 16, pp = ((c ^ 1) ^ sw) * 16;
            v8i bu[Cfg::FN];
            int bs[Cfg::FN];
#pragma unroll
            for (int b = 0; b < Cfg::FN; ++b) {
                const uint4 x = *reinterpret_cast<const uint4*>(st + (b_row + b * 32) * DROW + po);
                bu[b][0] = (int)x.x; bu[b][1] = (int)x.y; bu[b][2] = (int)x.z; bu[b][3] = (int)x.w;
                bu[b][4] = 0; bu[b][5] = 0; bu[b][6] = 0; bu[b][7] = 0;
                bs[b] = scale_byte(sb[b], blk * 2 + half);
            }
#pragma unroll
            for (int a = 0; a < Cfg::FM; ++a) {
                const uint8_t* r = st + (a_row + a * 32) * DROW;
                const uint4 x = *reinterpret_cast<const uint4*>(r + po);
                const uint4 y = *reinterpret_cast<const uint4*>(r + pp);
                v8i au, av;
                au[0] = (int)x.x; au[1] = (int)x.y; au[2] = (int)x.z; au[3] = (int)x.w;
                av[0] = (int)y.x; av[1] = (int)y.y; av[2] = (int)y.z; av[3] = (int)y.w;
#pragma unroll
                for (int d = 4; d < 8; ++d) {
                    au[d] = 0;
                    av[d] = 0;
                }
                av = mx_partner<OPF_MX4>(av, half != 0);
                const int as = scale_byte(sa[a], blk * 2 + half), ps = scale_byte(sa[a], blk * 2 + 1 - half);
#pragma unroll
                for (int b = 0; b < Cfg::FN; ++b) {
                    acc_re[a][b] =
                        __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(au, bu[b], acc_re[a][b], 4, 4, 0, as, 0, bs[b]);
                    acc_im[a][b] =
                        __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bu[b], acc_im[a][b], 4, 4, 0, ps, 0, bs[b]);
                }
            }
        }
    }

    // One stage of MFMAs on the swizzled layout.
    template <int FMT, int OPF>
    static __device__ __forceinline__ void mfma_stage_dma(const uint8_t* st, int a_row, int b_row,
                                                          int lane, v16f (&acc_re)[Cfg::FM][Cfg::FN],
                                                          v16f (&acc_im)[Cfg::FM][Cfg::FN]) {
        // operand rows are (multiple of 32) + (lane & 31): one swizzle for all of them
        const int sw = swz(lane & 31), half = (lane >> 5) * 2;
#pragma unroll
        for (int ks = 0; ks < Cfg::KC / 32; ++ks) {
            const int p0 = ((ks * 4 + half) ^ sw) * 16, p1 = p0 ^ 16;
            v8i au[Cfg::FM], bu[Cfg::FN];
            if constexpr (opf_mx(OPF)) {
#pragma unroll
                for (int a = 0; a < Cfg::FM; ++a) au[a] = lds_mx_part<OPF>(st + (a_row + a * 32) * DROW, p0, p1);
#pragma unroll
                for (int b = 0; b < Cfg::FN; ++b) bu[b] = lds_mx_part<OPF>(st + (b_row + b * 32) * DROW, p0, p1);
            } else {
#pragma unroll
                for (int a = 0; a < Cfg::FM; ++a) {
                    const uint8_t* r = st + (a_row + a * 32) * DROW;
                    const uint4 x = *reinterpret_cast<const uint4*>(r + p0);
                    const uint4 y = *reinterpret_cast<const uint4*>(r + p1);
                    au[a][0] = (int)x.x; au[a][1] = (int)x.y; au[a][2] = (int)x.z; au[a][3] = (int)x.w;
                    au[a][4] = (int)y.x; au[a][5] = (int)y.y; au[a][6] = (int)y.z; au[a][7] = (int)y.w;
                }
#pragma unroll
                for (int b = 0; b < Cfg::FN; ++b) {
                    const uint8_t* r = st + (b_row + b * 32) * DROW;
                    const uint4 x = *reinterpret_cast<const uint4*>(r + p0);
                    const uint4 y = *reinterpret_cast<const uint4*>(r + p1);
                    bu[b][0] = (int)x.x; bu[b][1] = (int)x.y; bu[b][2] = (int)x.z; bu[b][3] = (int)x.w;
                    bu[b][4] = (int)y.x; bu[b][5] = (int)y.y; bu[b][6] = (int)y.z; bu[b][7] = (int)y.w;
                }
            }
            if constexpr (opf_mx(OPF)) {
                static_assert(FMT == OUT_LOWER_PACKED || gemm_layout(FMT), "MX rows: row-major layouts");
                // the block's other part: logical chunk ^ 2, so byte offset ^ 32 under the swizzle
                const bool hi = half != 0;
#pragma unroll
                for (int a = 0; a < Cfg::FM; ++a) {
                    const v8i av = mx_partner<OPF>(
                        lds_mx_part<OPF>(st + (a_row + a * 32) * DROW, p0 ^ 32, p1 ^ 32), hi);
#pragma unroll
                    for (int b = 0; b < Cfg::FN; ++b) {
                        acc_re[a][b] = mfma_32x32x64_mx<OPF>(au[a], bu[b], acc_re[a][b]);
                        acc_im[a][b] = mfma_32x32x64_mx<OPF>(av, bu[b], acc_im[a][b]);
                    }
                }
                continue;
            }
#pragma unroll
            for (int a = 0; a < Cfg::FM; ++a) {
                const v8i av = conj_op<OPF>(au[a]);
#pragma unroll
                for (int b = 0; b < Cfg::FN; ++b) {
                    if constexpr (FMT == OUT_LOWER_PACKED || FMT == OUT_GRID || gemm_layout(FMT)) {
                        acc_re[a][b] = mfma_32x32x64<OPF>(au[a], bu[b], acc_re[a][b]);
                        acc_im[a][b] = mfma_32x32x64<OPF>(av, bu[b], acc_im[a][b]);
                    } else {
                        acc_re[a][b] = mfma_32x32x64<OPF>(bu[b], au[a], acc_re[a][b]);
                        acc_im[a][b] = mfma_32x32x64<OPF>(bu[b], av, acc_im[a][b]);
                    }
                }
            }
        }
    }

    // One stage of the direct-load loop: loads of the next stage into `nxt` (and its
    // scales into `snx`), MFMAs on `cur`.  The buffers are distinct, and saying so
    // (__restrict__) lets the compiler see that this stage's LDS reads cannot alias the
    // loads in flight: it then waits for those loads only at the next barrier instead of
    // before the first read of `cur`, so the loads overlap the MFMAs.
    template <int FMT, bool ACTIVE, int OPF>
    static __device__ __forceinline__ void dma_stage_body(const TileJob& j, const uint8_t* __restrict__ cur,
                                                          uint8_t* __restrict__ nxt, const uint8_t* __restrict__ scur,
                                                          uint8_t* __restrict__ snx, const uint8_t* B,
                                                          const uint32_t (&off)[DOPS], int s, int nst, int lane,
                                                          int wave, int a_row, int b_row,
                                                          v16f (&acc_re)[Cfg::FM][Cfg::FN],
                                                          v16f (&acc_im)[Cfg::FM][Cfg::FN]) {
        constexpr bool DENSE = OPF == OPF_MX4D;
        if (s + 1 < nst) {
            const int64_t k1 = (int64_t)(s + 1) * 2 * Cfg::KC;
            dma_stage(j.A + k1, B + k1, nxt, wave, off);
            if constexpr (DENSE) dma_scales(j, s + 1, snx, wave, lane);
        }
        if constexpr (ACTIVE) {
            if constexpr (DENSE)
                mfma_stage_mx4d<FMT>(cur, scur, a_row, b_row, lane, acc_re, acc_im);
            else
                mfma_stage_dma<FMT, OPF>(cur, a_row, b_row, lane, acc_re, acc_im);
        }
    }

    // Two stage buffers, loads one stage ahead, one barrier per stage: the barrier
    // that publishes stage s also retires every wave's reads of stage s - 1, whose
    // buffer the loads of stage s + 1 then overwrite.  Only these loads are in flight
    // in the loop, so vmcnt(0) waits for exactly them.  Ends with a barrier.
    template <int FMT, bool ACTIVE, int OPF>
    static __device__ __forceinline__ void dma_tile_loop(const TileJob& j, uint8_t* __restrict__ lds,
                                                         int lane, int wave, int nst, int a_row,
                                                         int b_row, v16f (&acc_re)[Cfg::FM][Cfg::FN],
                                                         v16f (&acc_im)[Cfg::FM][Cfg::FN]) {
        constexpr bool GEMM = gemm_layout(FMT);
        const uint8_t* B = GEMM ? j.B : j.A;
        uint32_t off[DOPS];
        dma_offsets<GEMM>(j, lane, wave, off);
        constexpr bool DENSE = OPF == OPF_MX4D;
        uint8_t* const scl = lds + 2 * DSTAGE;  // dense E2M1: two scale buffers after the codes
        dma_stage(j.A, B, lds, wave, off);
        if constexpr (DENSE) dma_scales(j, 0, scl, wave, lane);
        for (int s = 0; s < nst; ++s) {
            wait_vmcnt<0>();  // this wave's loads of stage s landed
            __syncthreads();
            const int c = s & 1, n = c ^ 1;
            dma_stage_body<FMT, ACTIVE, OPF>(j, lds + c * DSTAGE, lds + n * DSTAGE, scl + c * SC_STAGE,
                                             scl + n * SC_STAGE, B, off, s, nst, lane, wave, a_row, b_row, acc_re,
                                             acc_im);
        }
        __syncthreads();
    }

    // ---- Three stage buffers, loads two stages ahead (GEMM layouts) ----
    // The GEMM streams its column operand from HBM once per tile, so its loads need more
    // time in flight than one stage.  With a third buffer the loads of stage s + 2 are
    // issued right after the barrier that opens stage s, into the buffer of stage s - 1,
    // which every wave finished reading before that barrier.  The barrier is written out
    // (wait for this wave's loads of stage s + 1 only, then s_barrier): the release fence
    // of __syncthreads() would also wait for the loads of stage s + 2.
    static constexpr int DMA3_LDS = 3 * DSTAGE;
    static constexpr bool DMA3_OK = DMA_OK && Cfg::WG_PER_CU == 1 && DMA3_LDS <= 160 * 1024;

    template <int VM>
    static __device__ __forceinline__ void stage_barrier() {
        wait_vmcnt<VM>();
        __builtin_amdgcn_s_waitcnt(0xC07F & ~0x0F00);  // lgkmcnt(0): this wave's LDS reads are done
        __asm__ __volatile__("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        __asm__ __volatile__("" ::: "memory");
    }

    template <int FMT, bool ACTIVE>
    static __device__ __forceinline__ void dma3_stage_body(const TileJob& j, const uint8_t* __restrict__ cur,
                                                           uint8_t* __restrict__ fre, const uint8_t* B,
                                                           const uint32_t (&off)[DOPS], int s, int nst, int lane,
                                                           int wave, int a_row, int b_row,
                                                           v16f (&acc_re)[Cfg::FM][Cfg::FN],
                                                           v16f (&acc_im)[Cfg::FM][Cfg::FN]) {
        if (s + 2 < nst) {
            const int64_t k2 = (int64_t)(s + 2) * 2 * Cfg::KC;
            dma_stage(j.A + k2, B + k2, fre, wave, off);
        }
        if constexpr (ACTIVE) mfma_stage_dma<FMT, OPF_FP8>(cur, a_row, b_row, lane, acc_re, acc_im);
    }

    template <int FMT, bool ACTIVE>
    static __device__ __forceinline__ void dma3_tile_loop(const TileJob& j, uint8_t* __restrict__ lds, int lane,
                                                          int wave, int nst, int a_row, int b_row,
                                                          v16f (&acc_re)[Cfg::FM][Cfg::FN],
                                                          v16f (&acc_im)[Cfg::FM][Cfg::FN]) {
        constexpr bool GEMM = gemm_layout(FMT);
        const uint8_t* B = GEMM ? j.B : j.A;
        uint32_t off[DOPS];
        dma_offsets<GEMM>(j, lane, wave, off);
        dma_stage(j.A, B, lds, wave, off);
        if (nst > 1) dma_stage(j.A + 2 * Cfg::KC, B + 2 * Cfg::KC, lds + DSTAGE, wave, off);
        int c = 0;  // buffer of stage s
        for (int s = 0; s < nst; ++s) {
            if (s + 1 < nst)
                stage_barrier<DOPS>();  // stage s landed; stage s + 1 may still be in flight
            else
                stage_barrier<0>();
            const int f = c == 0 ? 2 : c - 1;  // buffer of stage s - 1 = buffer of stage s + 2
            dma3_stage_body<FMT, ACTIVE>(j, lds + c * DSTAGE, lds + f * DSTAGE, B, off, s, nst, lane, wave, a_row,
                                         b_row, acc_re, acc_im);
            c = c == 2 ? 0 : c + 1;
        }
        wait_vmcnt<0>();
        __syncthreads();
    }

    // SHALLOW forces prefetch depth 1 (the persistent grouped kernel keeps more state
    // live and would spill the second register set).
    template <typename OutT, int MODE, int FMT = OUT_LOWER_PACKED, bool SHALLOW = false, bool TAILS = false>
    static __device__ __forceinline__ void run(const TileJob& j, uint8_t* __restrict__ lds) {
        constexpr bool GEMM = gemm_layout(FMT);
        static_assert(!GEMM || MODE == LOAD_FP8 || MODE == LOAD_MX6 || MODE == LOAD_MX4,
                      "GEMM layouts read pre-cast rows");
        static_assert(MODE != LOAD_MX4D || (MX4D_OK && DMA_OK && FMT == OUT_LOWER_PACKED),
                      "dense E2M1 rows: direct-load HERK tiles of 64-sample stages");
        constexpr bool PRE = MODE == LOAD_FP8 || MODE == LOAD_BF6 || MODE == LOAD_MX6 ||
                             MODE == LOAD_MX4 || MODE == LOAD_MX4D;  // pre-cast operand rows
        constexpr int OPF = MODE == LOAD_BF6    ? OPF_BF6
                            : MODE == LOAD_MX6  ? OPF_MX6
                            : MODE == LOAD_MX4  ? OPF_MX4
                            : MODE == LOAD_MX4D ? OPF_MX4D
                                                : OPF_FP8;
        const int tid = threadIdx.x;
        // wave index in an SGPR: the per-wave conditions below are then scalar branches
        const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
        const int wm = wave / Cfg::WAVES_N, wn = wave % Cfg::WAVES_N;
        const int wrow0 = j.row0 + wm * Cfg::WM;  // first row of this wave's block
        const int wcol0 = j.col0 + wn * Cfg::WN;
        // The column block lies inside the row block (diagonal tiles): read it from the
        // row region instead of loading it twice.
        const bool share_cols =
            !GEMM && (j.col0 >= j.row0) && (j.col0 + Cfg::TN <= j.row0 + Cfg::TM);
        // HERK: wave has output iff some (i >= j) with i < N in its block; GEMM: any
        // in-range element.
        const bool active = GEMM ? (wrow0 < j.N && wcol0 < j.Ncol)
                                 : (wrow0 < j.N) && (wrow0 + Cfg::WM - 1 >= wcol0);

        v16f acc_re[Cfg::FM][Cfg::FN], acc_im[Cfg::FM][Cfg::FN];
#pragma unroll
        for (int a = 0; a < Cfg::FM; ++a)
#pragma unroll
            for (int b = 0; b < Cfg::FN; ++b)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    acc_re[a][b][r] = 0.0f;
                    acc_im[a][b][r] = 0.0f;
                }

        const int Kv = j.nseg > 1 ? j.K * j.nseg : j.K;  // samples the product runs over
        const int nst = (Kv + Cfg::KC - 1) / Cfg::KC;
        constexpr int NG = PRE ? Cfg::SEGS8_PER_THREAD : Cfg::SEGS_PER_THREAD;
        auto load = [&](int k0, uint4 (&g)[NG]) {
            if constexpr (PRE) load_stage_fp8<GEMM, MODE == LOAD_BF6>(j, k0, share_cols, tid, g);
            else load_stage<MODE>(j, k0, share_cols, tid, g);
        };
        auto store = [&](uint8_t* st, const uint4 (&g)[NG]) {
            if constexpr (PRE) store_stage_fp8(st, share_cols, tid, g);
            else store_stage(st, share_cols, tid, g);
        };

        const int lr = lane & 31, lh = (lane >> 5) * 32;
        const int a_off = (wm * Cfg::WM + lr) * Cfg::PITCH + lh;
        const int b_off = (share_cols ? (j.col0 - j.row0) : Cfg::TM) * Cfg::PITCH +
                          (wn * Cfg::WN + lr) * Cfg::PITCH + lh;

        auto compute = [&](const uint8_t* st) {
            if (active) mfma_stage<FMT, OPF>(st, a_off, b_off, acc_re, acc_im);
        };

        // Depth 2 only where its second register set fits without spilling: the
        // fused 16-byte load mode (pre-cast stages are twice as wide).
        constexpr int DEPTH = (MODE == LOAD_VEC16 && !SHALLOW) ? Cfg::DEPTH : 1;
        bool looped = false;
        if constexpr (MODE == LOAD_VEC16 && !GEMM) {
            if (full_tile<TAILS>(j)) {
                if (active)
                    full_tile_loop<FMT, true, DEPTH, TAILS>(j, lds, tid, nst, a_off, b_off, acc_re, acc_im);
                else
                    full_tile_loop<FMT, false, DEPTH, TAILS>(j, lds, tid, nst, a_off, b_off, acc_re, acc_im);
                looped = true;
            }
        }
        if constexpr (MODE == LOAD_FP8 && GEMM && DMA3_OK) {
            if (dma_tile<GEMM>(j)) {
                const int a_row = wm * Cfg::WM + (lane & 31);
                const int b_row = Cfg::TM + wn * Cfg::WN + (lane & 31);
                if (active)
                    dma3_tile_loop<FMT, true>(j, lds, lane, wave, nst, a_row, b_row, acc_re, acc_im);
                else
                    dma3_tile_loop<FMT, false>(j, lds, lane, wave, nst, a_row, b_row, acc_re, acc_im);
                looped = true;
            }
        } else if constexpr (PRE && DMA_OK) {
            if (dma_tile<GEMM, MODE != LOAD_MX4D>(j)) {
                const int a_row = wm * Cfg::WM + (lane & 31);
                const int b_row = (share_cols ? (j.col0 - j.row0) : Cfg::TM) + wn * Cfg::WN + (lane & 31);
                if (active)
                    dma_tile_loop<FMT, true, OPF>(j, lds, lane, wave, nst, a_row, b_row, acc_re, acc_im);
                else
                    dma_tile_loop<FMT, false, OPF>(j, lds, lane, wave, nst, a_row, b_row, acc_re, acc_im);
                looped = true;
            }
        }
        if (looped) {
        } else if constexpr (MODE == LOAD_MX4D) {
            // not reached: the launch takes dense rows only when every tile is an interior
            // tile on the direct path (the accumulators stay zero otherwise)
        } else if constexpr (DEPTH == 1) {
            uint4 g[NG];
            if (nst > 0) {
                load(0, g);
                store(lds, g);
            }
            __syncthreads();
            for (int s = 0; s < nst; ++s) {
                const bool more = (s + 1 < nst);
                if (more) load((s + 1) * Cfg::KC, g);
                compute(lds + (s & 1) * Cfg::STAGE);
                if (more) store(lds + ((s + 1) & 1) * Cfg::STAGE, g);
                __syncthreads();
            }
        } else {
            // Two register sets: ga holds the stage after the one being computed
            // while gb is loading the one after that; roles swap every stage, so the
            // loop runs in pairs (stage s in LDS buffer 0, s + 1 in buffer 1).
            uint4 ga[NG], gb[NG];
            if (nst > 0) {
                load(0, ga);
                store(lds, ga);
            }
            if (nst > 1) load(Cfg::KC, ga);
            __syncthreads();
            for (int s = 0; s < nst; s += 2) {
                if (s + 2 < nst) load((s + 2) * Cfg::KC, gb);
                compute(lds);
                if (s + 1 < nst) store(lds + Cfg::STAGE, ga);
                __syncthreads();
                if (s + 1 >= nst) break;
                if (s + 3 < nst) load((s + 3) * Cfg::KC, ga);
                compute(lds + Cfg::STAGE);
                if (s + 2 < nst) store(lds, gb);
                __syncthreads();
            }
        }

        if constexpr (FMT == OUT_GEMM_BEAM) {
            epilogue_beam<OutT>(j, lds, acc_re, acc_im, active, wrow0, wcol0, lane, tid);
            return;
        } else if constexpr (FMT == OUT_GRID) {
            if (active) epilogue_grid(j, acc_re, acc_im, wrow0, wcol0, lane);
            return;
        } else if constexpr (GEMM) {
            // B rows hold conj(B), so the HERK product A conj(B)^H = A B^T; lane =
            // column in the accumulator layout: consecutive lanes store consecutive j.
            if (!active) return;
            const int n = lane & 31, g4 = (lane >> 5) * 4;
#pragma unroll
            for (int a = 0; a < Cfg::FM; ++a)
#pragma unroll
                for (int b = 0; b < Cfg::FN; ++b) {
                    const int gj = wcol0 + b * 32 + n;
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int gi = wrow0 + a * 32 + (r >> 2) * 8 + g4 + (r & 3);
                        if (gi < j.N && gj < j.Ncol)
                            put_gemm<OutT, FMT>(j, (int64_t)gi * j.ldc + gj, acc_re[a][b][r],
                                                acc_im[a][b][r]);
                    }
                }
            return;
        } else if constexpr (FMT == OUT_LOWER_PACKED && Cfg::EPI_STAGED) {
            epilogue_staged<OutT>(j, lds, acc_re, acc_im, active, wrow0, wcol0, lane, wave);
            return;
        }
        if (!active) return;

        // Epilogue. 32x32 accumulator layout: lane holds column (lane & 31); register
        // r holds row 8*(r/4) + 4*(lane/32) + (r%4) -- of (i, j) for the lower-packed
        // layout, of (j, i) for the transposed GGP layouts.  Re and Im of one element
        // sit in the same lane and register, so each lane writes interleaved [re, im].
        const int n = lane & 31, g4 = (lane >> 5) * 4;
#pragma unroll
        for (int a = 0; a < Cfg::FM; ++a)
#pragma unroll
            for (int b = 0; b < Cfg::FN; ++b) {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int rr = (r >> 2) * 8 + g4 + (r & 3);
                    const int gi = wrow0 + a * 32 + (FMT == OUT_LOWER_PACKED ? rr : n);
                    const int gj = wcol0 + b * 32 + (FMT == OUT_LOWER_PACKED ? n : rr);
                    if (gi < j.N && gj <= gi) {
                        const float re = acc_re[a][b][r];
                        const float im = (gi == gj) ? 0.0f : acc_im[a][b][r];
                        if constexpr (FMT == OUT_LOWER_PACKED) {
                            put<OutT>(j, tri_index(gi, gj), re, im);
                        } else if constexpr (FMT == OUT_UPPER_PACKED) {
                            put<OutT>(j, (int64_t)gj * j.N + gi - ((int64_t)gj * (gj + 1) >> 1), re, im);
                        } else if constexpr (FMT == OUT_HERM_FULL) {
                            put<OutT>(j, (int64_t)gj * j.N + gi, re, im);
                            if (gi != gj) put<OutT>(j, (int64_t)gi * j.N + gj, re, -im);
                        }
                    }
                }
            }
    }
};

// Lower-triangle tiles of an N x N problem for TM x TN tiles (TM = r * TN):
// row block I holds column blocks J with J*TN <= (I+1)*TM - 1, capped at nbN.
__host__ __device__ inline int tri_tiles(int N, int TM, int TN) {
    const int nbM = (N + TM - 1) / TM, nbN = (N + TN - 1) / TN, r = TM / TN;
    // sum over I of min(nbN, r*(I+1))
    int64_t t = 0;
    for (int I = 0; I < nbM; ++I) t += (r * (I + 1) < nbN) ? r * (I + 1) : nbN;
    return (int)t;
}

// tile index t -> (I, J) for TM = r * TN (all row blocks but the last hold r*(I+1)).
__device__ __forceinline__ void tri_decode_r(int t, int r, int& I, int& J) {
    // uncapped count before row block I: r * I (I + 1) / 2
    const float q = (float)t / (float)r;
    int i = (int)((__builtin_sqrtf(8.0f * q + 1.0f) - 1.0f) * 0.5f);
    while (i > 0 && r * i * (i + 1) / 2 > t) --i;
    while (r * (i + 1) * (i + 2) / 2 <= t) ++i;
    I = i;
    J = t - r * i * (i + 1) / 2;
}

}  // namespace ftdx
