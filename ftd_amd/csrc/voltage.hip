// voltage.hip -- stage kernels of the voltage beamformer (gfx950).
//
//   voltage_rows     QC payloads [n_ch][pol][ant][time] -> GEMM operand rows with time as
//                    the row index: pre-cast FP8 rows in the beam tile's pol-interleaved
//                    order (FP8 path) or QC bytes (FP6 path).  One pass over the input
//                    replaces the reference's transpose + pol-split kernel and the GEMM
//                    API's INT4 -> FP8 cast (lib/voltage_pipeline.cu:176-276).
//   beam_integrate   per-sample Stokes-I power -> integrated / quantized filterbank
//   power_sum        planar beams of both pols -> filterbank (the FP6 path)
//   corner_turn      [Nf][Nb][Nt] -> [Nb][Nf][Nt]
//   quantize_u8, deinterleave_fp16
#include <hip/hip_runtime.h>

#include "herk_common.h"
#include "herk_tile.h"
#include "voltage.h"

namespace ftdx {

namespace {

constexpr int VT = 256;
// LDS bytes per antenna row of the 128-sample tile: 33 dwords, so the gathers of four
// 16-antenna segments by one wave land on disjoint banks
constexpr int VLP = 132;

unsigned grid_of(int64_t n, int64_t cap = 8192) {
    int64_t g = (n + VT - 1) / VT;
    return (unsigned)(g < 1 ? 1 : (g > cap ? cap : g));
}

__device__ __forceinline__ const uint8_t* payload_base(const VoltageLayout& L, int p) {
    const uint8_t* base = L.payload[0];
#pragma unroll
    for (int q = 1; q < kVoltageMaxPayloads; ++q)
        if (p == q) base = L.payload[q];
    return base;
}

// Address of sample m (fused index) of antenna `ant`, polarization `pol`, batch b.
__device__ __forceinline__ const uint8_t* sample_ptr(const VoltageLayout& L, int b, int pol, int ant, int m) {
    const int per_payload = L.n_time * L.ch_fuse;
    const int p = m / per_payload;
    const int r = m - p * per_payload;
    const int c = r / L.n_time;
    const int t = r - c * L.n_time;
    const int64_t chan = (int64_t)b * L.ch_fuse + c;
    return payload_base(L, p) + ((chan * 2 + pol) * L.n_ant + ant) * (int64_t)L.n_time + t;
}

// One workgroup: batch b, samples 128q .. 128q + 127 (two 64-sample row blocks, so
// every 128-byte line of an antenna's time series is read by one workgroup) of both
// polarizations, antennas 64at .. 64at + 63.  Loads run along time (the input's fast
// axis) into an LDS tile [pol][ant][sample]; stores run along antennas (the operand
// rows' fast axis).  MODE 0: pre-cast FP8 rows in beam order (row block c = samples
// 64c .. 64c + 63, pol 0 then pol 1); MODE 1: QC bytes [b][pol][M_total][n_ant].
template <int MODE, bool VEC>
__global__ void __launch_bounds__(VT) voltage_rows_kernel(VoltageLayout L, uint8_t* __restrict__ rows,
                                                         int chunks, int ant_tiles, int Kp) {
    __shared__ __attribute__((aligned(16))) uint8_t tile[2 * 64 * VLP];
    const int M_total = L.n_time * L.ch_fuse * L.n_payloads;
    const int pairs = (chunks + 1) / 2;
    int w = blockIdx.x;
    const int at = w % ant_tiles;
    w /= ant_tiles;
    const int q = w % pairs;
    const int b = w / pairs;
    const int tid = threadIdx.x;
    for (int it = tid; it < 1024; it += VT) {
        const int seg = it & 7, al = (it >> 3) & 63, pol = it >> 9;
        const int ant = at * 64 + al;
        const int m0 = q * 128 + seg * 16;
        uint32_t x[4] = {0u, 0u, 0u, 0u};
        if (ant < L.n_ant && m0 < M_total) {
            if constexpr (VEC) {  // the 16 samples are one aligned run of one channel
                const uint4 v = *reinterpret_cast<const uint4*>(sample_ptr(L, b, pol, ant, m0));
                x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
            } else {
                for (int e = 0; e < 16; ++e)
                    if (m0 + e < M_total)
                        x[e >> 2] |= (uint32_t)*sample_ptr(L, b, pol, ant, m0 + e) << (8 * (e & 3));
            }
        }
        uint32_t* d = reinterpret_cast<uint32_t*>(tile + (pol * 64 + al) * VLP + seg * 16);
        d[0] = x[0]; d[1] = x[1]; d[2] = x[2]; d[3] = x[3];
    }
    __syncthreads();
    for (int it = tid; it < 1024; it += VT) {
        const int seg = it & 3, mm = (it >> 2) & 127, pol = it >> 9;
        const int a0 = at * 64 + seg * 16;
        const int c = q * 2 + (mm >> 6);
        if (c >= chunks) continue;
        uint32_t x[4] = {0u, 0u, 0u, 0u};
        const uint8_t* src = tile + (pol * 64 + seg * 16) * VLP + mm;
#pragma unroll
        for (int e = 0; e < 16; ++e) x[e >> 2] |= (uint32_t)src[e * VLP] << (8 * (e & 3));
        if constexpr (MODE == 0) {
            if (a0 >= Kp) continue;
            const int64_t row = (int64_t)b * chunks * 128 + c * 128 + pol * 64 + (mm & 63);
            uint8_t* dst = rows + row * (2 * Kp) + 2 * a0;
            uint4 lo, hi;
            qc4_to_fp8_planar(x[0], lo.x, lo.y);
            qc4_to_fp8_planar(x[1], lo.z, lo.w);
            qc4_to_fp8_planar(x[2], hi.x, hi.y);
            qc4_to_fp8_planar(x[3], hi.z, hi.w);
            *reinterpret_cast<uint4*>(dst) = lo;
            *reinterpret_cast<uint4*>(dst + 16) = hi;
        } else {
            const int m = q * 128 + mm;
            if (m >= M_total || a0 >= L.n_ant) continue;
            uint8_t* dst = rows + (((int64_t)b * 2 + pol) * M_total + m) * L.n_ant + a0;
            if (VEC && a0 + 16 <= L.n_ant && (L.n_ant & 15) == 0) {
                *reinterpret_cast<uint4*>(dst) = make_uint4(x[0], x[1], x[2], x[3]);
            } else {
                for (int e = 0; e < 16 && a0 + e < L.n_ant; ++e) dst[e] = (uint8_t)(x[e >> 2] >> (8 * (e & 3)));
            }
        }
    }
}

__device__ __forceinline__ uint8_t quantize(float v, float scale) {
    // rint (nearest-even) of v * scale clamped to [0, 255]; NaN -> 0 (__float2int_rn
    // gives 0 for NaN, saturates +-inf)
    const float q = __builtin_fminf(__builtin_fmaxf(__builtin_rintf(v * scale), 0.0f), 255.0f);
    return (uint8_t)(int)q;
}

template <bool U8>
__global__ void __launch_bounds__(VT) beam_integrate_kernel(const float* __restrict__ P, void* __restrict__ out,
                                                           int64_t rows, int M, int nt, float scale) {
    const int n_out = M / nt;
    const int64_t total = rows * n_out;
    for (int64_t i = (int64_t)blockIdx.x * VT + threadIdx.x; i < total; i += (int64_t)gridDim.x * VT) {
        const int64_t r = i / n_out;
        const int t = (int)(i - r * n_out);
        const float* p = P + r * M + (int64_t)t * nt;
        float acc = 0.0f;
        for (int dt = 0; dt < nt; ++dt) acc += p[dt];
        if constexpr (U8) static_cast<uint8_t*>(out)[i] = quantize(acc, scale);
        else static_cast<float*>(out)[i] = acc;
    }
}

// Thread per (channel, output sample, beam), beam fastest (coalesced plane reads).
template <bool U8>
__global__ void __launch_bounds__(VT) power_sum_kernel(const float* __restrict__ re, const float* __restrict__ im,
                                                      void* __restrict__ out, int n_ch, int n_time, int n_beam,
                                                      int nt, float scale) {
    const int n_out = n_time / nt;
    const int64_t total = (int64_t)n_ch * n_out * n_beam;
    for (int64_t i = (int64_t)blockIdx.x * VT + threadIdx.x; i < total; i += (int64_t)gridDim.x * VT) {
        const int beam = (int)(i % n_beam);
        const int64_t q = i / n_beam;
        const int t_out = (int)(q % n_out);
        const int64_t ch = q / n_out;
        const int64_t base = ch * 2 * n_time * (int64_t)n_beam + beam;
        float power = 0.0f;
        for (int dt = 0; dt < nt; ++dt) {
            const int64_t i0 = base + (int64_t)(t_out * nt + dt) * n_beam;
            const int64_t i1 = i0 + (int64_t)n_time * n_beam;
            const float ra = re[i0], ia = im[i0], rb = re[i1], ib = im[i1];
            // ra^2 + ia^2 + rb^2 + ib^2 with the reference's contraction fixed
            power += __builtin_fmaf(ib, ib, __builtin_fmaf(rb, rb, __builtin_fmaf(ia, ia, ra * ra)));
        }
        const int64_t o = (ch * n_beam + beam) * n_out + t_out;
        if constexpr (U8) static_cast<uint8_t*>(out)[o] = quantize(power, scale);
        else static_cast<float*>(out)[o] = power;
    }
}

// Corner turn with Nt >= 16: each (f, b) row of Nt values moves as a unit; one
// workgroup per destination row, coalesced on both sides.
__global__ void __launch_bounds__(VT) corner_turn_rows_kernel(float* __restrict__ out, const float* __restrict__ in,
                                                             int Nf, int Nb, int Nt) {
    const int64_t nrows = (int64_t)Nf * Nb;
    for (int64_t r = blockIdx.x; r < nrows; r += gridDim.x) {
        const int64_t bb = r / Nf;
        const int64_t f = r - bb * Nf;
        const float* src = in + (f * Nb + bb) * Nt;
        float* dst = out + r * Nt;
        for (int t = threadIdx.x; t < Nt; t += VT) dst[t] = src[t];
    }
}

// Corner turn with small Nt: 32 x 32 (f, b) tiles through LDS for each t.
__global__ void __launch_bounds__(VT) corner_turn_tile_kernel(float* __restrict__ out, const float* __restrict__ in,
                                                             int Nf, int Nb, int Nt, int tiles_b, int tiles_f) {
    __shared__ float tile[32][33];
    const int64_t per_t = (int64_t)tiles_b * tiles_f;
    for (int64_t w = blockIdx.x; w < per_t * Nt; w += gridDim.x) {
        const int t = (int)(w / per_t);
        const int64_t q = w - (int64_t)t * per_t;
        const int fb0 = (int)(q / tiles_b) * 32, bb0 = (int)(q % tiles_b) * 32;
        const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
        for (int i = ty; i < 32; i += VT / 32) {
            const int f = fb0 + i, bb = bb0 + tx;
            if (f < Nf && bb < Nb) tile[i][tx] = in[((int64_t)f * Nb + bb) * Nt + t];
        }
        __syncthreads();
        for (int i = ty; i < 32; i += VT / 32) {
            const int bb = bb0 + i, f = fb0 + tx;
            if (f < Nf && bb < Nb) out[((int64_t)bb * Nf + f) * Nt + t] = tile[tx][i];
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(VT) quantize_u8_kernel(const float* __restrict__ in, uint8_t* __restrict__ out,
                                                        float scale, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * VT + threadIdx.x; i < n; i += (int64_t)gridDim.x * VT)
        out[i] = quantize(in[i], scale);
}

__global__ void __launch_bounds__(VT) deinterleave_fp16_kernel(const float2* __restrict__ in, _Float16* __restrict__ re,
                                                              _Float16* __restrict__ im, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * VT + threadIdx.x; i < n; i += (int64_t)gridDim.x * VT) {
        const float2 v = in[i];
        re[i] = (_Float16)v.x;
        im[i] = (_Float16)v.y;
    }
}

bool layout_vec(const VoltageLayout& L) {
    if (L.n_time % 16 != 0) return false;
    for (int p = 0; p < L.n_payloads; ++p)
        if (reinterpret_cast<uintptr_t>(L.payload[p]) & 15u) return false;
    return true;
}

template <int MODE>
hipError_t launch_rows(const VoltageLayout& L, uint8_t* rows, hipStream_t s) {
    if (L.n_payloads < 1 || L.n_payloads > kVoltageMaxPayloads || L.ch_fuse < 1 || L.n_ch % L.ch_fuse != 0)
        return hipErrorInvalidValue;
    const int64_t M_total = (int64_t)L.n_time * L.ch_fuse * L.n_payloads;
    const int batch = L.n_ch / L.ch_fuse;
    if (M_total <= 0 || L.n_ant <= 0 || batch <= 0) return hipSuccess;
    const int Kp = (L.n_ant + 15) / 16 * 16;
    const int chunks = (int)((M_total + 63) / 64);
    const int ant_tiles = (Kp + 63) / 64;
    const int64_t blocks = (int64_t)batch * ((chunks + 1) / 2) * ant_tiles;
    if (blocks > 0x7FFFFFFF || M_total > 0x7FFFFFFF) return hipErrorInvalidValue;
    if (layout_vec(L))
        hipLaunchKernelGGL((voltage_rows_kernel<MODE, true>), dim3((unsigned)blocks), dim3(VT), 0, s, L, rows,
                           chunks, ant_tiles, Kp);
    else
        hipLaunchKernelGGL((voltage_rows_kernel<MODE, false>), dim3((unsigned)blocks), dim3(VT), 0, s, L, rows,
                           chunks, ant_tiles, Kp);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_voltage_beam_rows(const VoltageLayout& L, uint8_t* rows, hipStream_t s) {
    return launch_rows<0>(L, rows, s);
}

hipError_t launch_voltage_qc_rows(const VoltageLayout& L, uint8_t* rows, hipStream_t s) {
    return launch_rows<1>(L, rows, s);
}

hipError_t launch_beam_integrate(const float* P, void* out, bool out_u8, int64_t rows, int M, int nt,
                                 float scale, hipStream_t s) {
    if (rows <= 0 || M <= 0 || nt <= 0) return hipSuccess;
    const unsigned g = grid_of(rows * (M / nt));
    if (out_u8)
        hipLaunchKernelGGL(beam_integrate_kernel<true>, dim3(g), dim3(VT), 0, s, P, out, rows, M, nt, scale);
    else
        hipLaunchKernelGGL(beam_integrate_kernel<false>, dim3(g), dim3(VT), 0, s, P, out, rows, M, nt, scale);
    return hipGetLastError();
}

hipError_t launch_power_sum(const float* re, const float* im, void* out, bool out_u8, int n_ch, int n_time,
                            int n_beam, int nt, float scale, hipStream_t s) {
    if (n_ch <= 0 || n_time <= 0 || n_beam <= 0 || nt <= 0) return hipSuccess;
    const unsigned g = grid_of((int64_t)n_ch * (n_time / nt) * n_beam);
    if (out_u8)
        hipLaunchKernelGGL(power_sum_kernel<true>, dim3(g), dim3(VT), 0, s, re, im, out, n_ch, n_time, n_beam, nt,
                           scale);
    else
        hipLaunchKernelGGL(power_sum_kernel<false>, dim3(g), dim3(VT), 0, s, re, im, out, n_ch, n_time, n_beam, nt,
                           scale);
    return hipGetLastError();
}

hipError_t launch_corner_turn(float* out, const float* in, int Nf, int Nb, int Nt, hipStream_t s) {
    if (Nf <= 0 || Nb <= 0 || Nt <= 0) return hipSuccess;
    if (Nt >= 16) {
        const int64_t rows = (int64_t)Nf * Nb;
        hipLaunchKernelGGL(corner_turn_rows_kernel, dim3((unsigned)(rows < 65536 ? rows : 65536)), dim3(VT), 0, s,
                           out, in, Nf, Nb, Nt);
    } else {
        const int tb = (Nb + 31) / 32, tf = (Nf + 31) / 32;
        const int64_t work = (int64_t)tb * tf * Nt;
        hipLaunchKernelGGL(corner_turn_tile_kernel, dim3((unsigned)(work < 65536 ? work : 65536)), dim3(VT), 0, s,
                           out, in, Nf, Nb, Nt, tb, tf);
    }
    return hipGetLastError();
}

hipError_t launch_quantize_u8(const float* in, uint8_t* out, float scale, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(quantize_u8_kernel, dim3(grid_of(n)), dim3(VT), 0, s, in, out, scale, n);
    return hipGetLastError();
}

hipError_t launch_deinterleave_fp16(const float* in, uint16_t* re, uint16_t* im, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(deinterleave_fp16_kernel, dim3(grid_of(n)), dim3(VT), 0, s,
                       reinterpret_cast<const float2*>(in), reinterpret_cast<_Float16*>(re),
                       reinterpret_cast<_Float16*>(im), n);
    return hipGetLastError();
}

}  // namespace ftdx
