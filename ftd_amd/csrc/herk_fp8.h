// herk_fp8.h -- host-side launch interface of the FP8 HERK kernels (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ftdx {

// Tile configurations compiled into the library (ids are stable: they appear in
// the tuning cache file).
enum TileConfigId {
    TILE_128x128_K64 = 0,   // 4 waves (2x2, 64x64 each), 2 workgroups / CU
    TILE_128x128_K128 = 1,  // 4 waves, 128-sample stages, 1 workgroup / CU
    TILE_64x64_K64 = 2,     // 1 wave, 4 workgroups / CU: small N, fine-grained balance
    TILE_256x128_K64 = 3,   // 8 waves (4x2, 64x64 each), 1 workgroup / CU
    TILE_128x128_K64_P2 = 4,  // as 0, global loads issued two stages ahead
    TILE_256x128_K64_P2 = 5,  // as 3, global loads issued two stages ahead
    NUM_TILE_CONFIGS = 6
};

// Flag on a tile configuration id (ftdx_herk, INT4 input): the same tile fed by
// pre-cast E3M2 operand rows.  Exact for QC data, MFMAs at twice the FP8 rate, one
// extra pass over the input (per chunk of batches, ftdx_herk).
constexpr int TILE_BF6 = 16;
inline bool valid_tile_config(int c) { return c >= 0 && (c & ~TILE_BF6) < NUM_TILE_CONFIGS; }
inline int tile_base(int c) { return c & ~TILE_BF6; }


const char* tile_config_name(int cfg);
int tile_config_rows(int cfg);   // TM
int tile_config_cols(int cfg);   // TN
int tile_config_wg_per_cu(int cfg);

// Output layout of a launch (element (i, j <= i) below is sum_k a_ik conj(a_jk)):
//   OUT_LOWER_PACKED  (i, j) at i(i+1)/2 + j                   (the HERK layout)
//   OUT_UPPER_PACKED  GGP: row q = j, column p = i at q*N + p - q(q+1)/2
//   OUT_HERM_FULL     GGP: full M^H M row-major, (i, j) at j*N + i, conj at i*N + j
//   OUT_GEMM_PLANAR   C = alpha A B^T + beta C, Re and Im planes [M][ldc] (B pre-cast as conj)
//   OUT_GEMM_POWER    P = alpha |A B^T|^2 + beta P, FP32 [M][ldc]
//   OUT_GEMM_BEAM     beamformer detect + integrate: 128-row tile = 64 samples of pol 0
//                     then the same 64 of pol 1; out[col][m / nt] = sum over the nt
//                     samples of (|pol 0|^2 + |pol 1|^2), quantized (uint8) or FP32
//   OUT_GRID          pillbox UV gridding of each element (i, j <= i) instead of a store:
//                     alpha * C_ij added into the FP32 [Ng][Ng][2] grid cell of baseline
//                     i(i+1)/2 + j at this batch's frequency, and its conjugate at (-u, -v)
//                     for i != j (PillboxGridScatter, include/kernels/visibility_bf.cuh:269-370)
enum OutFormat {
    OUT_LOWER_PACKED = 0,
    OUT_UPPER_PACKED = 1,
    OUT_HERM_FULL = 2,
    OUT_GEMM_PLANAR = 3,
    OUT_GEMM_POWER = 4,
    OUT_GEMM_BEAM = 5,
    OUT_GRID = 6
};

// Full M x N GEMM grid with B from its own rows (not the lower triangle of A A^H).
constexpr bool gemm_layout(int fmt) {
    return fmt == OUT_GEMM_PLANAR || fmt == OUT_GEMM_POWER || fmt == OUT_GEMM_BEAM;
}

struct HerkPlainArgs {
    const uint8_t* A;   // [batch][N][K] QC bytes, or pre-cast FP8 rows when precast != 0
    void* C;            // [batch][N(N+1)] scalars (float or half), packed lower triangle
                        // (GGP launches: the layout and type launch_herk_ggp names)
    int N, K, batch;
    float alpha, beta;
    int64_t a_bstride;  // bytes of A per batch (N*K)
    int64_t c_bstride;  // scalars of C per batch (N*(N+1))
    int precast = 0;    // 1: A is pre-cast FP8 (dword-planar) rows, 2: pre-cast E3M2 rows
                        // (32 bytes per 16 samples), 3 / 4: MX E3M2 / E2M1 rows (64 bytes
                        // per 32-sample block, scales inline), 5: dense E2M1 rows (32 bytes
                        // per block, scales in S; K counts 2-byte units, a multiple of 64,
                        // and every tile must be interior: mx4d_dense_ok); row pitch lda bytes
    int64_t lda = 0;
    // precast 5: scales [batch][stage][N][8], the (Re, Im) E8M0 bytes of each stage's 4 blocks
    const uint8_t* S = nullptr;
    int64_t s_stage = 0, s_bstride = 0;
    // Sum over K segments (QC input only): output batch b = ob * seg_group + ib sums
    // the HERKs of input batches (ob * nseg + s) * seg_group + ib, s < nseg, as one
    // product over nseg * K samples (polarization sum: nseg = n_pol, seg_group =
    // n_time_inner).  Defaults: one input batch per output batch.
    int nseg = 1;
    int seg_group = 1;
    // OUT_GRID (launch_herk_grid): C is the grid [batch][Ng][Ng][2] (c_bstride = 2 Ng^2
    // floats); baseline UV in metres [N(N+1)/2][2] by packed lower index; output batch b
    // uses frequency freq_hz[freq_offset + b] (Hz) and cell size `cell` (radians).
    const float* uv_m = nullptr;
    const double* freq_hz = nullptr;
    int freq_offset = 0;
    int Ng = 0;
    float cell = 0.0f;
};

// Lower-triangle tiles per problem for a configuration.
int herk_tiles(int cfg, int N);

// Whether dense E2M1 rows (precast 5) run on configuration cfg for N rows of lda bytes:
// a 64-sample-stage configuration whose tiles are all interior.
bool mx4d_dense_ok(int cfg, int N, int64_t lda);

// Plain batched launch: grid = tiles x batch, XCD-aware workgroup order when
// xcd_remap != 0.
hipError_t launch_herk_int4_plain(const HerkPlainArgs& a, int cfg, bool out_fp16, int xcd_remap,
                                  hipStream_t stream);

// HERK (QC input, K segments as in HerkPlainArgs::nseg) whose epilogue grids the
// visibilities (OUT_GRID) instead of storing them; FP32 grid, accumulated atomically.
hipError_t launch_herk_grid(const HerkPlainArgs& a, int cfg, int xcd_remap, hipStream_t stream);

// GGP-convention outputs of a pre-cast (precast = 1) batch: fmt 1 = upper triangle
// of M^H M packed row-major, fmt 2 = full N x N Hermitian row-major; element type
// float or (out_fp64) double.  C batch stride c_bstride counts scalars.
hipError_t launch_herk_ggp(const HerkPlainArgs& a, int cfg, int fmt, bool out_fp64, int xcd_remap,
                           hipStream_t stream);

// ---- complex GEMM over pre-cast rows (the prepared-B beamformer path) ----
struct GemmArgs {
    const uint8_t* A;   // pre-cast rows of A [batch][M] (pitch lda bytes)
    const uint8_t* B;   // pre-cast rows of conj(B) [batch][N] (pitch ldb bytes)
    void* C;            // OUT_GEMM_PLANAR: Re plane; OUT_GEMM_POWER: power (float)
    void* C2;           // OUT_GEMM_PLANAR: Im plane
    int M, N, K, batch;
    float alpha, beta;
    int64_t lda, ldb, a_bstride, b_bstride;  // bytes
    int64_t ldc, c_bstride;                  // elements
    // OUT_GEMM_BEAM: samples per output (divides 64), valid samples per batch (rows
    // of pol 0); alpha is the quantization scale.
    int beam_nt = 0, beam_samples = 0;
    // operand rows: 0 pre-cast FP8, 3 / 4 MX E3M2 / E2M1 rows (K = padded samples,
    // OUT_GEMM_PLANAR only)
    int operand = 0;
};

// fmt = OUT_GEMM_PLANAR (float or, out_fp16, half planes) or OUT_GEMM_POWER (float).
hipError_t launch_gemm_fp8(const GemmArgs& a, int cfg, int fmt, bool out_fp16, int xcd_remap,
                           hipStream_t stream);
// OUT_GEMM_BEAM on the 128 x 128 tile: uint8 (out_u8, rint(acc * alpha) clamped to
// [0, 255]) or FP32 acc; M = 128 * ceil(beam_samples / 64) rows per batch.
hipError_t launch_gemm_beam(const GemmArgs& a, bool out_u8, int xcd_remap, hipStream_t stream);

// ---- grouped persistent launch ----
struct GroupProblemDev {
    const uint8_t* A;
    void* C;
    int N, K;
    float alpha, beta;
    int vec;     // 1: 16-byte loads legal (K % 16 == 0, A 16-byte aligned)
    int tiles;
};

struct GroupItemDev {
    int prob;
    int tile;
};

// Persistent kernel: workgroup g processes items[wg_begin[g] .. wg_begin[g+1]).  all_vec:
// every problem has vec set (the kernel then compiles only the 16-byte load path).
hipError_t launch_herk_int4_grouped(const GroupProblemDev* probs, const GroupItemDev* items,
                                    const int* wg_begin, int num_wg, int cfg, bool out_fp16, bool all_vec,
                                    hipStream_t stream);

}  // namespace ftdx
