// imaging.hip -- consumers of the packed visibility triangle on gfx950:
// polarization reduction and pillbox UV gridding.
//
//   pol_reduce_triangle  [ch][pol][ti] triangles -> [ch][ti] sum over pol
//                        (pol_reduce_triangle_kernel[_fp16], lib/bench_pipeline.cu:254-302,
//                        there for n_pol = 2).  The fused alternative is the HERK over
//                        K segments (HerkPlainArgs::nseg), which never writes the
//                        per-polarization triangles.
//   pillbox_grid         (frequency, baseline) -> atomic add into the UV cell of the
//                        baseline's UV at that frequency, and the conjugate into
//                        (-u, -v) for cross-correlations (PillboxGridScatter,
//                        include/kernels/visibility_bf.cuh:259-370;
//                        pillbox_grid_scatter_fp32_fp32grid_kernel,
//                        include/kernels/imaging_kernels.cuh:219-270).
//
// Both are HBM/atomic-bound, one grid-stride pass.  The UV cell arithmetic keeps the
// reference's float / double steps (wavelength in double, u_lambda rounded to float,
// roundf of u_lambda / cell + Ng/2), so the cells agree.  The baseline index is decoded
// exactly in integers (the reference decodes it in float, which loses the diagonal
// once 8k + 1 exceeds 2^24).  FP32 atomic adds of integer-valued visibilities are exact
// and therefore order-independent while a cell's sum stays below 2^24.
#include <hip/hip_runtime.h>
#include <math.h>

#include "imaging.h"

namespace ftdx {

namespace {

constexpr int IM_THREADS = 256;

unsigned grid_for(int64_t n) {
    int64_t g = (n + IM_THREADS - 1) / IM_THREADS;
    if (g > 65536) g = 65536;
    return (unsigned)(g > 0 ? g : 1);
}

template <typename T>
__device__ __forceinline__ float to_f(T v) { return (float)v; }

template <typename T>
__global__ void __launch_bounds__(IM_THREADS)
pol_reduce_kernel(const T* __restrict__ in, T* __restrict__ out, int64_t per, int n_ch, int n_pol,
                  int n_ti) {
    const int64_t total = (int64_t)n_ch * n_ti * per;
    const int64_t stride = (int64_t)gridDim.x * IM_THREADS;
    for (int64_t i = (int64_t)blockIdx.x * IM_THREADS + threadIdx.x; i < total; i += stride) {
        const int64_t e = i % per;
        const int64_t b = i / per;  // ch * n_ti + ti
        const int64_t ch = b / n_ti, ti = b - ch * n_ti;
        const int64_t base = (ch * n_pol * n_ti + ti) * per + e;
        float s = to_f(in[base]);
        for (int p = 1; p < n_pol; ++p) s += to_f(in[base + (int64_t)p * n_ti * per]);
        out[i] = (T)s;  // one rounding (exact for FP32 integer sums; RNE for FP16)
    }
}

__device__ __forceinline__ void tri_row_col(int64_t k, int& row, int& col) {
    int r = (int)((__builtin_sqrt(8.0 * (double)k + 1.0) - 1.0) * 0.5);
    while ((int64_t)r * (r + 1) / 2 > k) --r;
    while ((int64_t)(r + 1) * (r + 2) / 2 <= k) ++r;
    row = r;
    col = (int)(k - (int64_t)r * (r + 1) / 2);
}

template <typename T>
__global__ void __launch_bounds__(IM_THREADS)
pillbox_grid_kernel(const T* __restrict__ vis, const float* __restrict__ uv_m,
                    const double* __restrict__ freq_hz, float* __restrict__ grid, int64_t n_bl,
                    int Ng, int Nf, int freq_offset, float cell) {
    const double C_LIGHT = 299792458.0;
    const int64_t total = (int64_t)Nf * n_bl;
    const int64_t plane = (int64_t)Ng * Ng;
    const float half_ng = 0.5f * Ng;
    const int64_t stride = (int64_t)gridDim.x * IM_THREADS;
    for (int64_t i = (int64_t)blockIdx.x * IM_THREADS + threadIdx.x; i < total; i += stride) {
        const int64_t f = i / n_bl, k = i - f * n_bl;
        int row, col;
        tri_row_col(k, row, col);
        const float vre = to_f(vis[2 * i]), vim = to_f(vis[2 * i + 1]);
        const double lambda = C_LIGHT / freq_hz[freq_offset + f];
        const float u_lam = (float)((double)uv_m[2 * k] / lambda);
        const float v_lam = (float)((double)uv_m[2 * k + 1] / lambda);
        const int iu = (int)roundf(u_lam / cell + half_ng);
        const int iv = (int)roundf(v_lam / cell + half_ng);
        if (iu >= 0 && iu < Ng && iv >= 0 && iv < Ng) {
            float* g = grid + 2 * (f * plane + (int64_t)iv * Ng + iu);
            unsafeAtomicAdd(g, vre);
            unsafeAtomicAdd(g + 1, vim);
        }
        if (row != col) {
            const int iuc = (int)roundf(-u_lam / cell + half_ng);
            const int ivc = (int)roundf(-v_lam / cell + half_ng);
            if (iuc >= 0 && iuc < Ng && ivc >= 0 && ivc < Ng) {
                float* g = grid + 2 * (f * plane + (int64_t)ivc * Ng + iuc);
                unsafeAtomicAdd(g, vre);
                unsafeAtomicAdd(g + 1, -vim);
            }
        }
    }
}

}  // namespace

hipError_t launch_pol_reduce(const void* in, void* out, int N, int n_ch, int n_pol, int n_ti, bool fp16,
                             hipStream_t s) {
    const int64_t per = (int64_t)N * (N + 1);  // scalars per triangle
    const int64_t total = (int64_t)n_ch * n_ti * per;
    if (total <= 0) return hipSuccess;
    if (fp16)
        hipLaunchKernelGGL((pol_reduce_kernel<_Float16>), dim3(grid_for(total)), dim3(IM_THREADS), 0, s,
                           static_cast<const _Float16*>(in), static_cast<_Float16*>(out), per, n_ch,
                           n_pol, n_ti);
    else
        hipLaunchKernelGGL((pol_reduce_kernel<float>), dim3(grid_for(total)), dim3(IM_THREADS), 0, s,
                           static_cast<const float*>(in), static_cast<float*>(out), per, n_ch, n_pol,
                           n_ti);
    return hipGetLastError();
}

hipError_t launch_pillbox_grid(const void* vis, bool fp16, const float* uv_m, const double* freq_hz,
                               float* grid, int64_t n_bl, int Ng, int Nf, int freq_offset, float cell,
                               hipStream_t s) {
    const int64_t total = (int64_t)Nf * n_bl;
    if (total <= 0 || Ng <= 0) return hipSuccess;
    if (fp16)
        hipLaunchKernelGGL((pillbox_grid_kernel<_Float16>), dim3(grid_for(total)), dim3(IM_THREADS), 0, s,
                           static_cast<const _Float16*>(vis), uv_m, freq_hz, grid, n_bl, Ng, Nf,
                           freq_offset, cell);
    else
        hipLaunchKernelGGL((pillbox_grid_kernel<float>), dim3(grid_for(total)), dim3(IM_THREADS), 0, s,
                           static_cast<const float*>(vis), uv_m, freq_hz, grid, n_bl, Ng, Nf, freq_offset,
                           cell);
    return hipGetLastError();
}

}  // namespace ftdx
