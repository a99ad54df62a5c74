// calib.hip -- differentiable gain calibration on packed visibilities and the
// HERK adjoint, on gfx950.
//
// Forward model (per channel b): model visibilities M = X X^H (the HERK output,
// packed lower triangle), complex gains g, observed V.  For i >= j:
//     R_ij = V_ij - g_i conj(g_j) M_ij,        L = sum_{i>=j} |R_ij|^2
// Gradients in the "d/dRe + i d/dIm" convention:
//     grad_g_k = -2 sum_{j<=k} R_kj conj(M_kj) g_j  -  2 sum_{i>=k} conj(R_ik) M_ik g_i
//     Gamma_ij = dL/dRe M_ij + i dL/dIm M_ij = -2 R_ij conj(g_i) g_j     (i >= j)
// and, through M = X X^H, the HERK adjoint
//     dX = (Gamma + Gamma^H) X
// (with Gamma lower-triangular; its diagonal enters as 2 Re Gamma_ii).
//
// Kernels:
//   gain_tile_kernel    32x32 tiles of the triangle: residual, Gamma, loss and
//                       per-tile row / column partial sums of the gain gradient;
//   gain_reduce_kernel  fixed-order reduction of the partials (deterministic);
//   hermitian_expand_kernel  Gamma (packed) -> H = Gamma + Gamma^H (dense N x N) as three
//                       bf16 split planes, LDS-staged so both the block and its conjugate
//                       transpose are written with coalesced rows;
//   adjoint_x_rows_kernel  QC voltages -> bf16 (Re, Im) rows per time sample (transposed);
//   adjoint_gemm_kernel dX = H X: complex GEMM on bf16 MFMA (v_mfma_f32_32x32x16_bf16)
//                       over the three exact split terms of H; FP32 accumulation.
#include <hip/hip_runtime.h>

#include <mutex>

#include "calib.h"
#include "herk_common.h"

// Fixed device scratch of the allocation-free gain entry point
// (ftdx_gain_loss_grad_noalloc): the row / column gradient partials and the loss
// partials live here instead of in a context buffer.  Each device the library runs on
// holds its own copy until the process ends.  Host code never passes this name to a
// kernel: it uses the device address that gain_static_scratch() looked up.
static __device__ __align__(256) uint8_t scratch[ftdx::kGainStaticScratchBytes];

namespace ftdx {

namespace {

constexpr int GT = 32;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

__global__ void __launch_bounds__(256)
gain_tile_kernel(const float2* __restrict__ V, const float2* __restrict__ M,
                 const float2* __restrict__ g, int N, int nb, int tiles,
                 float2* __restrict__ Gamma, float2* __restrict__ rowpart,
                 float2* __restrict__ colpart, double* __restrict__ losspart) {
    __shared__ float2 rowc[GT][GT + 1];
    __shared__ float2 colc[8][GT];
    __shared__ double lred[256];
    const int w = (int)blockIdx.x;
    const int b = w / tiles, t = w - b * tiles;
    int I, J;
    tri_decode(t, I, J);
    const int tid = threadIdx.x, c = tid & 31, r0 = tid >> 5;
    const int64_t P = (int64_t)N * (N + 1) / 2;
    const int j = J * GT + c;
    const float2 gj = (j < N) ? g[(int64_t)b * N + j] : make_float2(0.f, 0.f);
    float2 colsum = make_float2(0.f, 0.f);
    double lsum = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int rr = r0 + 8 * q;
        const int i = I * GT + rr;
        float2 rc = make_float2(0.f, 0.f);
        if (i < N && j <= i) {
            const int64_t idx = (int64_t)b * P + tri_index(i, j);
            const float2 v = V[idx], m = M[idx];
            const float2 gi = g[(int64_t)b * N + i];
            const float2 p = cmul(gi, cconj(gj));  // g_i conj(g_j)
            const float2 z = cmul(p, m);
            const float2 R = make_float2(v.x - z.x, v.y - z.y);
            lsum += (double)R.x * R.x + (double)R.y * R.y;
            if (Gamma) {
                const float2 gm = cmul(R, cconj(p));  // R conj(g_i) g_j
                Gamma[idx] = make_float2(-2.f * gm.x, -2.f * gm.y);
            }
            const float2 r1 = cmul(cmul(R, cconj(m)), gj);  // row i term
            const float2 c1 = cmul(cmul(cconj(R), m), gi);  // column j term
            rc = make_float2(-2.f * r1.x, -2.f * r1.y);
            colsum.x += -2.f * c1.x;
            colsum.y += -2.f * c1.y;
        }
        rowc[rr][c] = rc;
    }
    colc[r0][c] = colsum;
    lred[tid] = lsum;
    __syncthreads();
    if (tid < GT) {
        const int i = I * GT + tid;
        float2 s = make_float2(0.f, 0.f);
        for (int cc = 0; cc < GT; ++cc) {
            s.x += rowc[tid][cc].x;
            s.y += rowc[tid][cc].y;
        }
        if (i < N) rowpart[((int64_t)b * nb + J) * N + i] = s;
    } else if (tid < 2 * GT) {
        const int cc = tid - GT, jj = J * GT + cc;
        float2 s = make_float2(0.f, 0.f);
        for (int r = 0; r < 8; ++r) {
            s.x += colc[r][cc].x;
            s.y += colc[r][cc].y;
        }
        if (jj < N) colpart[((int64_t)b * nb + I) * N + jj] = s;
    }
    for (int st = 128; st > 0; st >>= 1) {
        __syncthreads();
        if (tid < st) lred[tid] += lred[tid + st];
    }
    if (tid == 0) losspart[(int64_t)b * tiles + t] = lred[0];
}

// grid: (ceil(N/256), batch) for the gradient; one extra block row for the loss.
__global__ void __launch_bounds__(256)
gain_reduce_kernel(const float2* __restrict__ rowpart, const float2* __restrict__ colpart,
                   const double* __restrict__ losspart, int N, int nb, int tiles,
                   float2* __restrict__ grad, double* __restrict__ loss) {
    const int b = (int)blockIdx.y;
    if ((int)blockIdx.x == (int)gridDim.x - 1) {
        // loss: fixed-order tree over this channel's tile partials
        __shared__ double red[256];
        double s = 0.0;
        for (int t = threadIdx.x; t < tiles; t += 256) s += losspart[(int64_t)b * tiles + t];
        red[threadIdx.x] = s;
        for (int st = 128; st > 0; st >>= 1) {
            __syncthreads();
            if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        }
        if (threadIdx.x == 0) loss[b] = red[0];
        return;
    }
    const int k = (int)blockIdx.x * 256 + threadIdx.x;
    if (k >= N) return;
    const int kb = k / GT;
    float2 s = make_float2(0.f, 0.f);
    for (int Jb = 0; Jb <= kb; ++Jb) {
        const float2 v = rowpart[((int64_t)b * nb + Jb) * N + k];
        s.x += v.x;
        s.y += v.y;
    }
    for (int Ib = kb; Ib < nb; ++Ib) {
        const float2 v = colpart[((int64_t)b * nb + Ib) * N + k];
        s.x += v.x;
        s.y += v.y;
    }
    grad[(int64_t)b * N + k] = s;
}

// (a, b) -> three bf16 pairs, term t in s[t] (a in the low half, b in the high half): the
// top 8 significant bits of each value (truncated), then the next 8 of the remainder, then
// the last 8, so a = sum of its three terms exactly; each remainder a - trunc(a) is exact.
__device__ __forceinline__ void split3(float a, float b, uint32_t (&s)[3]) {
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const uint32_t ha = __float_as_uint(a) & 0xFFFF0000u, hb = __float_as_uint(b) & 0xFFFF0000u;
        s[t] = (ha >> 16) | hb;
        a -= __uint_as_float(ha);
        b -= __uint_as_float(hb);
    }
}

// H = Gamma + Gamma^H from the packed lower triangle, written as the adjoint GEMM's row
// operand: plane t of [3][N][Np] holds term t of the bf16 split of (Re H_ij, -Im H_ij);
// columns N .. Np-1 are zero.  One 32x32 tile pair per block.
__global__ void __launch_bounds__(256)
hermitian_expand_kernel(const float2* __restrict__ Gamma, int N, int Np, int tiles,
                        uint32_t* __restrict__ H3) {
    __shared__ float2 blk[GT][GT + 1];
    const int w = (int)blockIdx.x;
    const int b = w / tiles, t = w - b * tiles;
    int I, J;
    tri_decode(t, I, J);
    const int tid = threadIdx.x, c = tid & 31, r0 = tid >> 5;
    const int64_t P = (int64_t)N * (N + 1) / 2;
    const int64_t plane = (int64_t)N * Np;
    const float2* Gb = Gamma + (int64_t)b * P;
    uint32_t* Hb = H3 + (int64_t)b * 3 * plane;
    auto put = [&](int i, int j, float2 v) {
        uint32_t sp[3];
        split3(v.x, -v.y, sp);
#pragma unroll
        for (int k = 0; k < 3; ++k) Hb[k * plane + (int64_t)i * Np + j] = sp[k];
    };
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int rr = r0 + 8 * q, i = I * GT + rr, j = J * GT + c;
        float2 v = make_float2(0.f, 0.f);
        if (i < N && j <= i) v = Gb[tri_index(i, j)];
        blk[rr][c] = v;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int rr = r0 + 8 * q;
        // block (I, J): H_ij for i in I-block, j in J-block (zero where j >= N)
        {
            const int i = I * GT + rr, j = J * GT + c;
            if (i < N && j < Np) {
                float2 v;
                if (i > j) v = blk[rr][c];
                else if (i < j) v = cconj(blk[c][rr]);       // diagonal block only
                else v = make_float2(2.f * blk[rr][c].x, 0.f);
                put(i, j, v);
            }
        }
        // mirrored block (J, I) for off-diagonal tiles: H_ji = conj(Gamma_ij)
        if (I != J) {
            const int i = J * GT + rr, j = I * GT + c;  // row in J-block, column in I-block
            if (i < N && j < Np) put(i, j, cconj(blk[c][rr]));
        }
    }
}

__device__ __forceinline__ float qc_val(uint32_t n) {
    const int m = (int)(n & 7u);
    return (float)((n & 8u) ? -m : m);
}

// X QC bytes [batch][N][K] -> the adjoint GEMM's column operand XT [batch][K][Np]: bf16
// pairs (Re, Im) (small integers: the top half of the FP32 pattern, exactly), zero for
// N <= j < Np.  Tiles of 32 rows j x 128 columns k through LDS, each thread reading 4
// consecutive bytes of a row (one dword when X rows are 4-byte aligned); grid
// (ceil(K/128), ceil(Np/32), batch).
constexpr int XR_K = 128;

__device__ __forceinline__ uint32_t qc_bf16_pair(uint32_t byte) {
    return (__float_as_uint(qc_val(byte >> 4)) >> 16) | (__float_as_uint(qc_val(byte & 15u)) & 0xFFFF0000u);
}

__global__ void __launch_bounds__(256)
adjoint_x_rows_kernel(const uint8_t* __restrict__ X, int N, int K, int Np, int vec, uint32_t* __restrict__ XT) {
    __shared__ uint32_t t[GT][XR_K + 1];
    const int k0 = (int)blockIdx.x * XR_K, j0 = (int)blockIdx.y * GT, b = (int)blockIdx.z;
    const int tid = threadIdx.x, c4 = (tid & 31) * 4, r0 = tid >> 5;
    const uint8_t* Xb = X + (int64_t)b * N * K;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int jr = r0 + 8 * q, j = j0 + jr, k = k0 + c4;
        uint32_t w = 0u;
        if (j < N) {
            const uint8_t* src = Xb + (int64_t)j * K + k;
            if (vec && k + 4 <= K) {
                w = *reinterpret_cast<const uint32_t*>(src);
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (k + e < K) w |= (uint32_t)src[e] << (8 * e);
            }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) t[jr][c4 + e] = j < N ? qc_bf16_pair((w >> (8 * e)) & 0xFFu) : 0u;
    }
    __syncthreads();
    uint32_t* XTb = XT + (int64_t)b * K * Np;
    const int jc = tid & 31;
#pragma unroll 4
    for (int q = 0; q < XR_K / 8; ++q) {
        const int kr = r0 + 8 * q, k = k0 + kr, j = j0 + jc;
        if (k < K && j < Np) XTb[(int64_t)k * Np + j] = t[jc][kr];
    }
}

// ---- dX = H X on bf16 MFMA, H split into three bf16 terms ----
// H3 and XT come from the two passes above.  QC voltages are integers in [-7, 7], exact in
// bf16, so every product of a split term and a voltage is exact in FP32 and the MFMA
// accumulates them in FP32: the accuracy of an FP32 GEMM at 3/16 of its MFMA time.
// Complex values run along K as (re, im) bf16 pairs: the row operand holds (Hr, -Hi), the
// column operand (Xr, Xi), giving Re; the column pair swapped with one sign, (Xi, -Xr),
// gives Im.
constexpr int AT_M = 128, AT_N = 256, AT_J = 32, AT_THREADS = 512;  // 8 waves: 2 x 4 of 64 x 64
constexpr int AT_ROW = AT_J * 4;                // staged row bytes: 32 complex as bf16 pairs
constexpr int AT_SPLIT = AT_M * AT_ROW;         // one split plane of H rows
constexpr int AT_STAGE = 3 * AT_SPLIT + AT_N * AT_ROW;

typedef __bf16 v8bf16 __attribute__((ext_vector_type(8)));

// 16-byte chunk c (4 complex, 8 per row) of staged row r sits at chunk c ^ ((r / 2) mod 8):
// the 16 rows one ds_read_b128 pass serves land in 16 distinct groups of 4 banks.
__device__ __forceinline__ int at_chunk(int r, int c) { return c ^ ((r >> 1) & 7); }

__device__ __forceinline__ v8bf16 as_bf16x8(uint4 u) {
    v8bf16 v;
    __builtin_memcpy(&v, &u, 16);
    return v;
}

__global__ void __launch_bounds__(AT_THREADS)
adjoint_gemm_kernel(const uint32_t* __restrict__ H3, const uint32_t* __restrict__ XT,
                    float2* __restrict__ dX, int N, int K, int Np, int tiles_m, int tiles_n) {
    __shared__ __attribute__((aligned(16))) uint8_t lds[2 * AT_STAGE];
    const int per_b = tiles_m * tiles_n;
    // the column tiles of one row block run on one XCD and share its rows in that L2
    const int w = xcd_logical((int)blockIdx.x, (int)gridDim.x);
    const int b = w / per_b, rem = w - b * per_b;
    const int tm = rem / tiles_n, tn = rem - tm * tiles_n;
    const int i0 = tm * AT_M, k0 = tn * AT_N;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 2, wn = wave & 3;
    const int64_t plane = (int64_t)N * Np;

    v16f acc_re[2][2], acc_im[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                acc_re[x][y][r] = 0.f;
                acc_im[x][y][r] = 0.f;
            }

    // Items e = tid + 512 q: row operand row e / 8 (q = 0, 1), column operand column e / 8
    // (q = 0 .. 3), each 4 complex (16 bytes) at j0 + 4 (e % 8).  Rows past N and columns
    // past K load zeros; j never passes Np.
    constexpr int HQ = AT_M * 8 / AT_THREADS, XQ = AT_N * 8 / AT_THREADS;
    static_assert(HQ * AT_THREADS == AT_M * 8 && XQ * AT_THREADS == AT_N * 8, "whole items per thread");
    const uint32_t* hsrc[HQ];
    const uint32_t* xsrc[XQ];
    bool hok[HQ], xok[XQ];
#pragma unroll
    for (int q = 0; q < HQ; ++q) {
        const int e = tid + AT_THREADS * q, r = e >> 3, jq = e & 7;
        hok[q] = i0 + r < N;
        hsrc[q] = H3 + (int64_t)b * 3 * plane + (int64_t)(hok[q] ? i0 + r : 0) * Np + jq * 4;
    }
#pragma unroll
    for (int q = 0; q < XQ; ++q) {
        const int e = tid + AT_THREADS * q, r = e >> 3, jq = e & 7;
        xok[q] = k0 + r < K;
        xsrc[q] = XT + ((int64_t)b * K + (xok[q] ? k0 + r : 0)) * Np + jq * 4;
    }
    uint4 hreg[HQ][3], xreg[XQ];
    auto load = [&](int j0) {
#pragma unroll
        for (int q = 0; q < HQ; ++q)
#pragma unroll
            for (int t = 0; t < 3; ++t)
                hreg[q][t] = hok[q] ? *reinterpret_cast<const uint4*>(hsrc[q] + t * plane + j0)
                                    : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
        for (int q = 0; q < XQ; ++q)
            xreg[q] = xok[q] ? *reinterpret_cast<const uint4*>(xsrc[q] + j0) : make_uint4(0u, 0u, 0u, 0u);
    };
    auto store = [&](int buf) {
        uint8_t* st = lds + buf * AT_STAGE;
#pragma unroll
        for (int q = 0; q < HQ; ++q) {
            const int e = tid + AT_THREADS * q, r = e >> 3, jq = e & 7;
            const int off = r * AT_ROW + at_chunk(r, jq) * 16;
#pragma unroll
            for (int t = 0; t < 3; ++t) *reinterpret_cast<uint4*>(st + t * AT_SPLIT + off) = hreg[q][t];
        }
#pragma unroll
        for (int q = 0; q < XQ; ++q) {
            const int e = tid + AT_THREADS * q, r = e >> 3, jq = e & 7;
            *reinterpret_cast<uint4*>(st + 3 * AT_SPLIT + r * AT_ROW + at_chunk(r, jq) * 16) = xreg[q];
        }
    };

    const int nst = Np / AT_J;
    load(0);
    store(0);
    __syncthreads();
    const int lr = lane & 31, half = lane >> 5;
    for (int s = 0; s < nst; ++s) {
        const bool more = s + 1 < nst;
        if (more) load((s + 1) * AT_J);
        const uint8_t* st = lds + (s & 1) * AT_STAGE;
        const uint8_t* sb = st + 3 * AT_SPLIT;
#pragma unroll
        for (int ks = 0; ks < AT_J / 8; ++ks) {
            // this lane's 8 K values: chunk ks * 2 + half, the same complex indices in
            // the row and the column operand
            const int c = ks * 2 + half;
            v8bf16 bre[2], bim[2];
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                const int kk = wn * 64 + y * 32 + lr;
                const uint4 u = *reinterpret_cast<const uint4*>(sb + kk * AT_ROW + at_chunk(kk, c) * 16);
                // (Xr, Xi) -> (Xi, -Xr): swap the halves, negate the new high half
                const uint4 v = make_uint4(((u.x >> 16) | (u.x << 16)) ^ 0x80000000u,
                                           ((u.y >> 16) | (u.y << 16)) ^ 0x80000000u,
                                           ((u.z >> 16) | (u.z << 16)) ^ 0x80000000u,
                                           ((u.w >> 16) | (u.w << 16)) ^ 0x80000000u);
                bre[y] = as_bf16x8(u);
                bim[y] = as_bf16x8(v);
            }
#pragma unroll
            for (int t = 0; t < 3; ++t)
#pragma unroll
                for (int x = 0; x < 2; ++x) {
                    const int r = wm * 64 + x * 32 + lr;
                    const v8bf16 a = as_bf16x8(*reinterpret_cast<const uint4*>(st + t * AT_SPLIT + r * AT_ROW +
                                                                              at_chunk(r, c) * 16));
#pragma unroll
                    for (int y = 0; y < 2; ++y) {
                        acc_re[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bre[y], acc_re[x][y], 0, 0, 0);
                        acc_im[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bim[y], acc_im[x][y], 0, 0, 0);
                    }
                }
        }
        if (more) store((s & 1) ^ 1);
        __syncthreads();
    }

    float2* dXb = dX + (int64_t)b * N * K;
    const int n = lane & 31, g4 = (lane >> 5) * 4;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int k = k0 + wn * 64 + y * 32 + n;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = i0 + wm * 64 + x * 32 + (r >> 2) * 8 + g4 + (r & 3);
                if (i < N && k < K) dXb[(int64_t)i * K + k] = make_float2(acc_re[x][y][r], acc_im[x][y][r]);
            }
        }
}

}  // namespace

int gain_tiles(int N) {
    const int nb = (N + GT - 1) / GT;
    return nb * (nb + 1) / 2;
}

int gain_blocks(int N) { return (N + GT - 1) / GT; }

hipError_t launch_gain_loss_grad(const float2* V, const float2* M, const float2* g, int N, int batch,
                                 float2* Gamma, float2* rowpart, float2* colpart, double* losspart,
                                 double* loss, float2* grad, hipStream_t s) {
    if (N <= 0 || batch <= 0) return hipSuccess;
    const int nb = gain_blocks(N), tiles = gain_tiles(N);
    const int64_t blocks = (int64_t)tiles * batch;
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    hipLaunchKernelGGL(gain_tile_kernel, dim3((unsigned)blocks), dim3(256), 0, s, V, M, g, N, nb,
                       tiles, Gamma, rowpart, colpart, losspart);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    dim3 grid((unsigned)((N + 255) / 256 + 1), (unsigned)batch);
    hipLaunchKernelGGL(gain_reduce_kernel, grid, dim3(256), 0, s, rowpart, colpart, losspart, N, nb,
                       tiles, grad, loss);
    return hipGetLastError();
}

int adjoint_padded_n(int N) { return (N + AT_J - 1) / AT_J * AT_J; }

size_t adjoint_scratch_bytes(int N, int K, int batch) {
    const size_t Np = (size_t)adjoint_padded_n(N);
    return (size_t)batch * Np * (12 * (size_t)N + 4 * (size_t)K);
}

hipError_t launch_herk_adjoint(const float2* Gamma, const uint8_t* X, void* scratch, float2* dX, int N,
                               int K, int batch, hipStream_t s) {
    if (N <= 0 || batch <= 0) return hipSuccess;
    const int Np = adjoint_padded_n(N);
    uint32_t* H3 = static_cast<uint32_t*>(scratch);
    uint32_t* XT = H3 + (size_t)batch * 3 * N * Np;
    const int tiles = gain_tiles(N);
    int64_t blocks = (int64_t)tiles * batch;
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    hipLaunchKernelGGL(hermitian_expand_kernel, dim3((unsigned)blocks), dim3(256), 0, s, Gamma, N, Np,
                       tiles, H3);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || K <= 0) return e;
    if (batch > 65535) return hipErrorInvalidValue;
    const int vec = (K % 4 == 0 && (reinterpret_cast<uintptr_t>(X) & 3u) == 0) ? 1 : 0;
    hipLaunchKernelGGL(adjoint_x_rows_kernel,
                       dim3((unsigned)((K + XR_K - 1) / XR_K), (unsigned)((Np + GT - 1) / GT), (unsigned)batch),
                       dim3(256), 0, s, X, N, K, Np, vec, XT);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int tm = (N + AT_M - 1) / AT_M, tn = (K + AT_N - 1) / AT_N;
    blocks = (int64_t)tm * tn * batch;
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    hipLaunchKernelGGL(adjoint_gemm_kernel, dim3((unsigned)blocks), dim3(AT_THREADS), 0, s, H3, XT, dX,
                       N, K, Np, tm, tn);
    return hipGetLastError();
}

size_t gain_scratch_bytes(int N, int batch) {
    if (N <= 0 || batch <= 0) return 0;
    const size_t part = (size_t)batch * gain_blocks(N) * N * sizeof(float2);
    return 2 * part + (size_t)batch * gain_tiles(N) * sizeof(double);
}

// Device address of `scratch` on the current device, looked up once per device (with
// that device current) and kept; later calls return the kept address.
hipError_t gain_static_scratch(void** out) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* addr[kMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        addr[dev] = p;
    }
    *out = addr[dev];
    return hipSuccess;
}

}  // namespace ftdx
