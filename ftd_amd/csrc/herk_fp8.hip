// herk_fp8.hip -- batched HERK on gfx950 FP8 MFMA with INT4 (QC) input.
//
//   C_b = alpha * A_b A_b^H + beta * C_b,  A_b in C^{N x K}, packed lower triangle.
//
// One kernel replaces the reference's three-stage direct path
//   cast_int4_to_fp8_interleaved (cutlass_interface/cutlass_gemm_api.cu:1566-1623)
//   herk_direct_fp8_precast_kernel (cutlass_interface/shared/herk_kernel_common.hpp:175-425)
//   packed-triangle store           (herk_kernel_common.hpp:353-424)
// The per-tile work is in herk_tile.h.  Two launch shapes:
//   * plain:   one workgroup per (tile, batch), optional XCD-aware order so the
//              workgroups one XCD runs together share input rows in its L2;
//   * grouped: persistent workgroups walking a host-built static schedule over the
//              tiles of many problems of different N and K (one launch).
// FP32 accumulation is exact while |sum| < 2^24, i.e. 98*K < 2^24 for [-7, 7] input.
#include "herk_fp8.h"
#include "herk_tile.h"

namespace ftdx {

namespace {

using Cfg0 = TileCfg<128, 128, 2, 2, 64, 2, 2>;
using Cfg1 = TileCfg<128, 128, 2, 2, 128, 1, 1>;
using Cfg2 = TileCfg<64, 64, 1, 1, 64, 1, 4>;
using Cfg3 = TileCfg<256, 128, 4, 2, 64, 2, 1>;
using Cfg4 = TileCfg<128, 128, 2, 2, 64, 2, 2, 2>;  // Cfg0 with loads two stages ahead
using Cfg5 = TileCfg<256, 128, 4, 2, 64, 2, 1, 2>;  // Cfg3 with loads two stages ahead

// Calls f(CfgN{}) for tile configuration id `cfg`; hipErrorInvalidValue for unknown ids.
template <typename F>
hipError_t with_cfg(int cfg, F&& f) {
    switch (cfg) {
        case TILE_128x128_K64: return f(Cfg0{});
        case TILE_128x128_K128: return f(Cfg1{});
        case TILE_64x64_K64: return f(Cfg2{});
        case TILE_256x128_K64: return f(Cfg3{});
        case TILE_128x128_K64_P2: return f(Cfg4{});
        case TILE_256x128_K64_P2: return f(Cfg5{});
        default: return hipErrorInvalidValue;
    }
}

template <typename F>
int cfg_value(int cfg, F&& f, int fallback) {
    int v = fallback;
    (void)with_cfg(cfg, [&](auto c) {
        v = f(c);
        return hipSuccess;
    });
    return v;
}

// TAILS: K is not a whole number of stages (16-byte loads, one K segment); those launches
// get the branch-free loop's variant for a partial last stage, built apart so the other
// kernels' register allocation does not carry it.
template <class Cfg, typename OutT, int MODE, int FMT = OUT_LOWER_PACKED, bool TAILS = false>
__global__ void __launch_bounds__(Cfg::THREADS, Cfg::WAVES_PER_EU)
herk_int4_fp8_plain_kernel(HerkPlainArgs p, int tiles, int xcd_remap) {
    __shared__ __attribute__((aligned(16))) uint8_t lds[Cfg::LDS_BYTES];
    const int total = tiles * p.batch;
    const int w = xcd_remap ? xcd_logical((int)blockIdx.x, total) : (int)blockIdx.x;
    const int b = w / tiles;
    const int t = w - b * tiles;
    int I, J;
    tri_decode_r(t, Cfg::TM / Cfg::TN, I, J);
    TileJob j = {};
    // output batch b = (outer, inner) of seg_group inners; its K segments are the
    // input batches (outer * nseg + s) * seg_group + inner, s = 0 .. nseg-1
    const int grp = p.seg_group, ob = b / grp, ib = b - ob * grp;
    j.A = p.A + ((int64_t)ob * p.nseg * grp + ib) * p.a_bstride;
    j.nseg = p.nseg;
    j.seg_stride = (int64_t)grp * p.a_bstride;
    j.C = static_cast<OutT*>(p.C) + (int64_t)b * p.c_bstride;
    j.N = p.N;
    j.K = p.K;
    j.row0 = I * Cfg::TM;
    j.col0 = J * Cfg::TN;
    j.alpha = p.alpha;
    j.beta = p.beta;
    j.lda = p.lda;
    if constexpr (MODE == LOAD_MX4D) {
        j.S = p.S + (int64_t)b * p.s_bstride;
        j.s_stage = p.s_stage;
    }
    if constexpr (FMT == OUT_GRID) {
        j.uv = p.uv_m;
        j.lambda = 299792458.0 / p.freq_hz[p.freq_offset + b];  // speed of light / frequency
        j.Ng = p.Ng;
        j.cell = p.cell;
    }
    TileRunner<Cfg>::template run<OutT, MODE, FMT, false, TAILS>(j, lds);
}

// VEC_ONLY: every problem of the plan takes 16-byte loads, so only that load path is
// compiled in and the tile keeps its own prefetch depth; otherwise both load paths are
// inlined and loads run one stage ahead (the second register set would spill).
template <class Cfg, typename OutT, bool VEC_ONLY>
__global__ void __launch_bounds__(Cfg::THREADS, Cfg::WAVES_PER_EU)
herk_int4_fp8_grouped_kernel(const GroupProblemDev* __restrict__ probs,
                             const GroupItemDev* __restrict__ items,
                             const int* __restrict__ wg_begin) {
    __shared__ __attribute__((aligned(16))) uint8_t lds[Cfg::LDS_BYTES];
    const int beg = wg_begin[blockIdx.x], end = wg_begin[blockIdx.x + 1];
    for (int it = beg; it < end; ++it) {
        const GroupItemDev item = items[it];
        const GroupProblemDev pr = probs[item.prob];
        int I, J;
        tri_decode_r(item.tile, Cfg::TM / Cfg::TN, I, J);
        TileJob j = {};
        j.A = pr.A;
        j.C = pr.C;
        j.N = pr.N;
        j.K = pr.K;
        j.row0 = I * Cfg::TM;
        j.col0 = J * Cfg::TN;
        j.alpha = pr.alpha;
        j.beta = pr.beta;
        j.lda = pr.K;
        j.nseg = 1;
        if constexpr (VEC_ONLY)
            TileRunner<Cfg>::template run<OutT, LOAD_VEC16, OUT_LOWER_PACKED, false>(j, lds);
        else if (pr.vec)
            TileRunner<Cfg>::template run<OutT, LOAD_VEC16, OUT_LOWER_PACKED, true>(j, lds);
        else
            TileRunner<Cfg>::template run<OutT, LOAD_BYTES, OUT_LOWER_PACKED, true>(j, lds);
    }
}

template <class Cfg>
hipError_t launch_plain_cfg(const HerkPlainArgs& a, bool out_fp16, int xcd_remap,
                            hipStream_t stream) {
    const int tiles = tri_tiles(a.N, Cfg::TM, Cfg::TN);
    const int64_t blocks = (int64_t)tiles * a.batch;
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    const bool vec = (a.K % 16 == 0) && ((reinterpret_cast<uintptr_t>(a.A) & 15u) == 0);
    dim3 grid((unsigned)blocks), block(Cfg::THREADS);
    if (a.precast == 5) {
        if constexpr (TileRunner<Cfg>::MX4D_OK && TileRunner<Cfg>::DMA_OK) {
            if (out_fp16)
                hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, _Float16, LOAD_MX4D>), grid, block, 0, stream,
                                   a, tiles, xcd_remap);
            else
                hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, float, LOAD_MX4D>), grid, block, 0, stream, a,
                                   tiles, xcd_remap);
            return hipGetLastError();
        } else {
            return hipErrorInvalidValue;
        }
    }
    if (a.precast == 3 || a.precast == 4) {
        if (out_fp16) {
            if (a.precast == 3)
                hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, _Float16, LOAD_MX6>), grid, block, 0,
                                   stream, a, tiles, xcd_remap);
            else
                hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, _Float16, LOAD_MX4>), grid, block, 0,
                                   stream, a, tiles, xcd_remap);
        } else {
            if (a.precast == 3)
                hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, float, LOAD_MX6>), grid, block, 0,
                                   stream, a, tiles, xcd_remap);
            else
                hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, float, LOAD_MX4>), grid, block, 0,
                                   stream, a, tiles, xcd_remap);
        }
        return hipGetLastError();
    }
    if (a.precast == 2) {
        if (out_fp16)
            hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, _Float16, LOAD_BF6>), grid, block, 0,
                               stream, a, tiles, xcd_remap);
        else
            hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, float, LOAD_BF6>), grid, block, 0,
                               stream, a, tiles, xcd_remap);
        return hipGetLastError();
    }
    if (a.precast) {
        if (out_fp16)
            hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, _Float16, LOAD_FP8>), grid, block, 0,
                               stream, a, tiles, xcd_remap);
        else
            hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, float, LOAD_FP8>), grid, block, 0,
                               stream, a, tiles, xcd_remap);
        return hipGetLastError();
    }
    const bool tails = vec && a.nseg == 1 && a.K % Cfg::KC != 0;
    if (out_fp16) {
        if (tails)
            hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, _Float16, LOAD_VEC16, OUT_LOWER_PACKED, true>), grid,
                               block, 0, stream, a, tiles, xcd_remap);
        else if (vec)
            hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, _Float16, LOAD_VEC16>), grid, block,
                               0, stream, a, tiles, xcd_remap);
        else
            hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, _Float16, LOAD_BYTES>), grid, block,
                               0, stream, a, tiles, xcd_remap);
    } else {
        if (tails)
            hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, float, LOAD_VEC16, OUT_LOWER_PACKED, true>), grid,
                               block, 0, stream, a, tiles, xcd_remap);
        else if (vec)
            hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, float, LOAD_VEC16>), grid, block, 0,
                               stream, a, tiles, xcd_remap);
        else
            hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, float, LOAD_BYTES>), grid, block, 0,
                               stream, a, tiles, xcd_remap);
    }
    return hipGetLastError();
}

// Complex GEMM C = alpha A B^T + beta C over pre-cast rows of A and conj(B): full
// tile grid in column-major tile order within a batch (consecutive workgroups share
// the column block of B, the large operand in the beamformer shapes, while the few
// row blocks of A stay in L2), XCD-aware order when asked.
template <class Cfg, typename OutT, int FMT, int MODE = LOAD_FP8>
__global__ void __launch_bounds__(Cfg::THREADS, Cfg::WAVES_PER_EU)
gemm_fp8_kernel(GemmArgs p, int tiles_m, int tiles, int xcd_remap) {
    // three direct-load stage buffers where the configuration runs one workgroup per CU
    constexpr int LDS = (MODE == LOAD_FP8 && TileRunner<Cfg>::DMA3_OK && TileRunner<Cfg>::DMA3_LDS > Cfg::LDS_BYTES)
                            ? TileRunner<Cfg>::DMA3_LDS
                            : Cfg::LDS_BYTES;
    __shared__ __attribute__((aligned(16))) uint8_t lds[LDS];
    const int total = tiles * p.batch;
    const int w = xcd_remap ? xcd_logical((int)blockIdx.x, total) : (int)blockIdx.x;
    const int b = w / tiles;
    const int t = w - b * tiles;
    const int J = t / tiles_m, I = t - J * tiles_m;
    TileJob j = {};
    j.A = p.A + (int64_t)b * p.a_bstride;
    j.B = p.B + (int64_t)b * p.b_bstride;
    j.C = static_cast<OutT*>(p.C) + (int64_t)b * p.c_bstride;
    j.C2 = p.C2 ? static_cast<void*>(static_cast<OutT*>(p.C2) + (int64_t)b * p.c_bstride) : nullptr;
    j.N = p.M;
    j.Ncol = p.N;
    j.K = p.K;
    j.row0 = I * Cfg::TM;
    j.col0 = J * Cfg::TN;
    j.alpha = p.alpha;
    j.beta = p.beta;
    j.lda = p.lda;
    j.ldb = p.ldb;
    j.ldc = p.ldc;
    j.nseg = 1;
    j.beam_nt = p.beam_nt;
    j.beam_samples = p.beam_samples;
    TileRunner<Cfg>::template run<OutT, MODE, FMT>(j, lds);
}

template <class Cfg>
hipError_t launch_gemm_cfg(const GemmArgs& a, int fmt, bool out_fp16, int xcd_remap, hipStream_t stream) {
    const int64_t tiles_m = (a.M + Cfg::TM - 1) / Cfg::TM, tiles_n = (a.N + Cfg::TN - 1) / Cfg::TN;
    const int64_t blocks = tiles_m * tiles_n * a.batch;
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    dim3 grid((unsigned)blocks), block(Cfg::THREADS);
    const int tm = (int)tiles_m, tiles = (int)(tiles_m * tiles_n);
    if (a.operand != 0) {  // MX operand rows: planar output on the 128 x 128 tile
        if constexpr (Cfg::TM == 128 && Cfg::TN == 128 && Cfg::KC == 64 && Cfg::DEPTH == 1) {
            if (fmt != OUT_GEMM_PLANAR) return hipErrorInvalidValue;
            if (a.operand == 3 && out_fp16)
                hipLaunchKernelGGL((gemm_fp8_kernel<Cfg, _Float16, OUT_GEMM_PLANAR, LOAD_MX6>), grid, block, 0,
                                   stream, a, tm, tiles, xcd_remap);
            else if (a.operand == 3)
                hipLaunchKernelGGL((gemm_fp8_kernel<Cfg, float, OUT_GEMM_PLANAR, LOAD_MX6>), grid, block, 0,
                                   stream, a, tm, tiles, xcd_remap);
            else if (out_fp16)
                hipLaunchKernelGGL((gemm_fp8_kernel<Cfg, _Float16, OUT_GEMM_PLANAR, LOAD_MX4>), grid, block, 0,
                                   stream, a, tm, tiles, xcd_remap);
            else
                hipLaunchKernelGGL((gemm_fp8_kernel<Cfg, float, OUT_GEMM_PLANAR, LOAD_MX4>), grid, block, 0,
                                   stream, a, tm, tiles, xcd_remap);
            return hipGetLastError();
        } else {
            return hipErrorInvalidValue;
        }
    }
    if (fmt == OUT_GEMM_POWER)
        hipLaunchKernelGGL((gemm_fp8_kernel<Cfg, float, OUT_GEMM_POWER>), grid, block, 0, stream, a, tm,
                           tiles, xcd_remap);
    else if (fmt == OUT_GEMM_PLANAR && out_fp16)
        hipLaunchKernelGGL((gemm_fp8_kernel<Cfg, _Float16, OUT_GEMM_PLANAR>), grid, block, 0, stream, a,
                           tm, tiles, xcd_remap);
    else if (fmt == OUT_GEMM_PLANAR)
        hipLaunchKernelGGL((gemm_fp8_kernel<Cfg, float, OUT_GEMM_PLANAR>), grid, block, 0, stream, a, tm,
                           tiles, xcd_remap);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_gemm_beam_cfg0(const GemmArgs& a, bool out_u8, int xcd_remap, hipStream_t stream) {
    using Cfg = Cfg0;
    if (a.beam_nt <= 0 || 64 % a.beam_nt != 0 || a.M != (a.beam_samples + 63) / 64 * Cfg::TM)
        return hipErrorInvalidValue;
    const int64_t tiles_m = a.M / Cfg::TM, tiles_n = (a.N + Cfg::TN - 1) / Cfg::TN;
    const int64_t blocks = tiles_m * tiles_n * a.batch;
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    if (blocks == 0) return hipSuccess;
    dim3 grid((unsigned)blocks), block(Cfg::THREADS);
    const int tm = (int)tiles_m, tiles = (int)(tiles_m * tiles_n);
    if (out_u8)
        hipLaunchKernelGGL((gemm_fp8_kernel<Cfg, uint8_t, OUT_GEMM_BEAM>), grid, block, 0, stream, a, tm,
                           tiles, xcd_remap);
    else
        hipLaunchKernelGGL((gemm_fp8_kernel<Cfg, float, OUT_GEMM_BEAM>), grid, block, 0, stream, a, tm,
                           tiles, xcd_remap);
    return hipGetLastError();
}

// Gridding epilogue on QC input (K segments allowed); FP32 grid.
template <class Cfg>
hipError_t launch_grid_cfg(const HerkPlainArgs& a, int xcd_remap, hipStream_t stream) {
    const int tiles = tri_tiles(a.N, Cfg::TM, Cfg::TN);
    const int64_t blocks = (int64_t)tiles * a.batch;
    if (blocks > 0x7FFFFFFF || a.precast || !a.uv_m || !a.freq_hz || a.Ng <= 0) return hipErrorInvalidValue;
    const bool vec = (a.K % 16 == 0) && ((reinterpret_cast<uintptr_t>(a.A) & 15u) == 0);
    dim3 grid((unsigned)blocks), block(Cfg::THREADS);
    if (vec)
        hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, float, LOAD_VEC16, OUT_GRID>), grid, block, 0, stream, a,
                           tiles, xcd_remap);
    else
        hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, float, LOAD_BYTES, OUT_GRID>), grid, block, 0, stream, a,
                           tiles, xcd_remap);
    return hipGetLastError();
}

// GGP layouts: pre-cast FP8 operand rows only.
template <class Cfg>
hipError_t launch_ggp_cfg(const HerkPlainArgs& a, int fmt, bool out_fp64, int xcd_remap,
                          hipStream_t stream) {
    const int tiles = tri_tiles(a.N, Cfg::TM, Cfg::TN);
    const int64_t blocks = (int64_t)tiles * a.batch;
    if (blocks > 0x7FFFFFFF || !a.precast) return hipErrorInvalidValue;
    dim3 grid((unsigned)blocks), block(Cfg::THREADS);
    if (fmt == OUT_UPPER_PACKED) {
        if (out_fp64)
            hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, double, LOAD_FP8, OUT_UPPER_PACKED>),
                               grid, block, 0, stream, a, tiles, xcd_remap);
        else
            hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, float, LOAD_FP8, OUT_UPPER_PACKED>),
                               grid, block, 0, stream, a, tiles, xcd_remap);
    } else if (fmt == OUT_HERM_FULL) {
        if (out_fp64)
            hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, double, LOAD_FP8, OUT_HERM_FULL>),
                               grid, block, 0, stream, a, tiles, xcd_remap);
        else
            hipLaunchKernelGGL((herk_int4_fp8_plain_kernel<Cfg, float, LOAD_FP8, OUT_HERM_FULL>),
                               grid, block, 0, stream, a, tiles, xcd_remap);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <class Cfg>
hipError_t launch_grouped_cfg(const GroupProblemDev* probs, const GroupItemDev* items,
                              const int* wg_begin, int num_wg, bool out_fp16, bool all_vec, hipStream_t stream) {
    dim3 grid((unsigned)num_wg), block(Cfg::THREADS);
    if (all_vec) {
        if (out_fp16)
            hipLaunchKernelGGL((herk_int4_fp8_grouped_kernel<Cfg, _Float16, true>), grid, block, 0, stream,
                               probs, items, wg_begin);
        else
            hipLaunchKernelGGL((herk_int4_fp8_grouped_kernel<Cfg, float, true>), grid, block, 0, stream,
                               probs, items, wg_begin);
    } else {
        if (out_fp16)
            hipLaunchKernelGGL((herk_int4_fp8_grouped_kernel<Cfg, _Float16, false>), grid, block, 0, stream,
                               probs, items, wg_begin);
        else
            hipLaunchKernelGGL((herk_int4_fp8_grouped_kernel<Cfg, float, false>), grid, block, 0, stream,
                               probs, items, wg_begin);
    }
    return hipGetLastError();
}

}  // namespace

int tile_config_rows(int cfg) {
    return cfg_value(cfg, [](auto c) { return decltype(c)::TM; }, 0);
}

int tile_config_cols(int cfg) {
    return cfg_value(cfg, [](auto c) { return decltype(c)::TN; }, 0);
}

int tile_config_wg_per_cu(int cfg) {
    return cfg_value(cfg, [](auto c) { return decltype(c)::WG_PER_CU; }, 1);
}

int herk_tiles(int cfg, int N) { return tri_tiles(N, tile_config_rows(cfg), tile_config_cols(cfg)); }

bool mx4d_dense_ok(int cfg, int N, int64_t lda) {
    return cfg_value(
               cfg,
               [&](auto c) {
                   using Cfg = decltype(c);
                   return (TileRunner<Cfg>::MX4D_OK && TileRunner<Cfg>::DMA_OK && N > 0 && N % Cfg::TM == 0 &&
                           N % Cfg::TN == 0 && lda > 0 && lda % 16 == 0 && (int64_t)N * lda <= 0xFFFFFFFFll)
                              ? 1
                              : 0;
               },
               0) != 0;
}

hipError_t launch_herk_int4_plain(const HerkPlainArgs& a, int cfg, bool out_fp16, int xcd_remap,
                                  hipStream_t stream) {
    if (a.N <= 0 || a.batch <= 0) return hipSuccess;
    return with_cfg(cfg, [&](auto c) {
        return launch_plain_cfg<decltype(c)>(a, out_fp16, xcd_remap, stream);
    });
}

hipError_t launch_herk_ggp(const HerkPlainArgs& a, int cfg, int fmt, bool out_fp64, int xcd_remap,
                           hipStream_t stream) {
    if (a.N <= 0 || a.batch <= 0) return hipSuccess;
    return with_cfg(cfg, [&](auto c) {
        return launch_ggp_cfg<decltype(c)>(a, fmt, out_fp64, xcd_remap, stream);
    });
}

hipError_t launch_herk_grid(const HerkPlainArgs& a, int cfg, int xcd_remap, hipStream_t stream) {
    if (a.N <= 0 || a.batch <= 0) return hipSuccess;
    return with_cfg(cfg, [&](auto c) { return launch_grid_cfg<decltype(c)>(a, xcd_remap, stream); });
}

hipError_t launch_gemm_beam(const GemmArgs& a, bool out_u8, int xcd_remap, hipStream_t stream) {
    return launch_gemm_beam_cfg0(a, out_u8, xcd_remap, stream);
}

hipError_t launch_gemm_fp8(const GemmArgs& a, int cfg, int fmt, bool out_fp16, int xcd_remap,
                           hipStream_t stream) {
    if (a.M <= 0 || a.N <= 0 || a.batch <= 0) return hipSuccess;
    return with_cfg(cfg, [&](auto c) {
        return launch_gemm_cfg<decltype(c)>(a, fmt, out_fp16, xcd_remap, stream);
    });
}

hipError_t launch_herk_int4_grouped(const GroupProblemDev* probs, const GroupItemDev* items,
                                    const int* wg_begin, int num_wg, int cfg, bool out_fp16, bool all_vec,
                                    hipStream_t stream) {
    if (num_wg <= 0) return hipSuccess;
    return with_cfg(cfg, [&](auto c) {
        return launch_grouped_cfg<decltype(c)>(probs, items, wg_begin, num_wg, out_fp16, all_vec, stream);
    });
}

}  // namespace ftdx
