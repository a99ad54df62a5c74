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
//   hermitian_expand_kernel  Gamma (packed) -> H = Gamma + Gamma^H (dense N x N),
//                       LDS-staged so both the block and its conjugate transpose
//                       are written with coalesced rows;
//   adjoint_gemm_kernel dX = H X: FP32 MFMA (v_mfma_f32_32x32x2_f32) complex GEMM,
//                       QC voltages decoded to FP32 in the load stage; the Re/Im
//                       partner operand is a lane-half swap with one sign.
#include <hip/hip_runtime.h>

#include "calib.h"
#include "herk_common.h"

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

// H = Gamma + Gamma^H from the packed lower triangle; one 32x32 tile pair per block.
__global__ void __launch_bounds__(256)
hermitian_expand_kernel(const float2* __restrict__ Gamma, int N, int tiles,
                        float2* __restrict__ H) {
    __shared__ float2 blk[GT][GT + 1];
    const int w = (int)blockIdx.x;
    const int b = w / tiles, t = w - b * tiles;
    int I, J;
    tri_decode(t, I, J);
    const int tid = threadIdx.x, c = tid & 31, r0 = tid >> 5;
    const int64_t P = (int64_t)N * (N + 1) / 2;
    const float2* Gb = Gamma + (int64_t)b * P;
    float2* Hb = H + (int64_t)b * N * N;
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
        // block (I, J): H_ij for i in I-block, j in J-block
        {
            const int i = I * GT + rr, j = J * GT + c;
            if (i < N && j < N) {
                float2 v;
                if (i > j) v = blk[rr][c];
                else if (i < j) v = cconj(blk[c][rr]);       // diagonal block only
                else v = make_float2(2.f * blk[rr][c].x, 0.f);
                Hb[(int64_t)i * N + j] = v;
            }
        }
        // mirrored block (J, I) for off-diagonal tiles: H_ji = conj(Gamma_ij)
        if (I != J) {
            const int i = J * GT + rr, j = I * GT + c;  // row in J-block, column in I-block
            if (i < N && j < N) Hb[(int64_t)i * N + j] = cconj(blk[c][rr]);
        }
    }
}

// ---- dX = H X with FP32 MFMA ----
constexpr int AT_M = 128, AT_N = 128, AT_J = 16, AT_THREADS = 256;
constexpr int AT_HP = AT_J + 1;   // float2 pitch of a staged H row
constexpr int AT_XP = AT_N + 2;   // float2 pitch of a staged X row

__device__ __forceinline__ float qc_val(uint32_t n) {
    const int m = (int)(n & 7u);
    return (float)((n & 8u) ? -m : m);
}

__global__ void __launch_bounds__(AT_THREADS, 2)
adjoint_gemm_kernel(const float2* __restrict__ H, const uint8_t* __restrict__ X,
                    float2* __restrict__ dX, int N, int K, int tiles_m, int tiles_n) {
    __shared__ float2 hs[2][AT_M][AT_HP];
    __shared__ float2 xs[2][AT_J][AT_XP];
    const int w = (int)blockIdx.x;
    const int per_b = tiles_m * tiles_n;
    const int b = w / per_b, rem = w - b * per_b;
    const int tm = rem / tiles_n, tn = rem - tm * tiles_n;
    const int i0 = tm * AT_M, k0 = tn * AT_N;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const float2* Hb = H + (int64_t)b * N * N;
    const uint8_t* Xb = X + (int64_t)b * N * K;

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

    // per stage: H rows i0..i0+127, cols j0..j0+15 (8 float2 per thread);
    //            X rows j0..j0+15, cols k0..k0+127 (8 bytes per thread)
    float2 hreg[8];
    uint32_t xreg[8];
    auto load = [&](int j0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = tid + AT_THREADS * q;  // 0..2047
            const int r = e / AT_J, cj = e % AT_J;
            const int i = i0 + r, j = j0 + cj;
            hreg[q] = (i < N && j < N) ? Hb[(int64_t)i * N + j] : make_float2(0.f, 0.f);
        }
        const int r = tid / 16, c8 = (tid % 16) * 8;  // 16 rows x 16 groups of 8 columns
        const int j = j0 + r;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = k0 + c8 + e;
            xreg[e] = (j < N && k < K) ? (uint32_t)Xb[(int64_t)j * K + k] : 0u;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = tid + AT_THREADS * q;
            hs[buf][e / AT_J][e % AT_J] = hreg[q];
        }
        const int r = tid / 16, c8 = (tid % 16) * 8;
#pragma unroll
        for (int e = 0; e < 8; ++e)
            xs[buf][r][c8 + e] = make_float2(qc_val(xreg[e] >> 4), qc_val(xreg[e] & 15u));
    };

    const int nst = (N + AT_J - 1) / AT_J;
    load(0);
    store(0);
    __syncthreads();
    const int lr = lane & 31, kh = lane >> 5;  // kh: MFMA k index (0: real part, 1: imag part)
    for (int s = 0; s < nst; ++s) {
        const bool more = s + 1 < nst;
        if (more) load((s + 1) * AT_J);
        const int buf = s & 1;
#pragma unroll 4
        for (int jj = 0; jj < AT_J; ++jj) {
            float are[2], aim[2], bv[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                const float2 h = hs[buf][wm * 64 + x * 32 + lr][jj];
                // Re: sum Hr Xr - Hi Xi ; Im: sum Hi Xr + Hr Xi
                are[x] = kh ? -h.y : h.x;
                aim[x] = kh ? h.x : h.y;
            }
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                const float2 xv = xs[buf][jj][wn * 64 + y * 32 + lr];
                bv[y] = kh ? xv.y : xv.x;
            }
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) {
                    acc_re[x][y] = __builtin_amdgcn_mfma_f32_32x32x2f32(are[x], bv[y], acc_re[x][y], 0, 0, 0);
                    acc_im[x][y] = __builtin_amdgcn_mfma_f32_32x32x2f32(aim[x], bv[y], acc_im[x][y], 0, 0, 0);
                }
        }
        if (more) store(buf ^ 1);
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

hipError_t launch_herk_adjoint(const float2* Gamma, const uint8_t* X, float2* H, float2* dX, int N,
                               int K, int batch, hipStream_t s) {
    if (N <= 0 || batch <= 0) return hipSuccess;
    const int tiles = gain_tiles(N);
    int64_t blocks = (int64_t)tiles * batch;
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    hipLaunchKernelGGL(hermitian_expand_kernel, dim3((unsigned)blocks), dim3(256), 0, s, Gamma, N,
                       tiles, H);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || K <= 0) return e;
    const int tm = (N + AT_M - 1) / AT_M, tn = (K + AT_N - 1) / AT_N;
    blocks = (int64_t)tm * tn * batch;
    if (blocks > 0x7FFFFFFF) return hipErrorInvalidValue;
    hipLaunchKernelGGL(adjoint_gemm_kernel, dim3((unsigned)blocks), dim3(AT_THREADS), 0, s, H, X, dX,
                       N, K, tm, tn);
    return hipGetLastError();
}

}  // namespace ftdx
