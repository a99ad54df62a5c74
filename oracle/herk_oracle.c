/*
 * herk_oracle.c -- CPU oracle for the X-engine HERK.  TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, the smoke check and bench.py's CPU-baseline leg to check
 * results; never linked into or called by the product library.
 *
 * Restates, in plain C:
 *   encode_int4           cutlass_interface/test_herk_int4.cu:39-45
 *   decode_qc_byte        tests/visibility_pipeline_test.cpp:43-53
 *   qc_to_fp8_lut_256     cutlass_interface/cutlass_gemm_api.cu:1436-1536
 *   cpu_herk              cutlass_interface/test_herk_int4.cu:51-70 (int64 instead
 *                         of complex<double>: exact, same values)
 *   cpu_herk_ab           cutlass_interface/test_gemm_api.cu:726-751
 *   FP32/FP16 epilogue    cutlass_interface/shared/herk_kernel_common.hpp:405-420
 *   MX scale / codes      gemm_complex_sm100/mxfp_kernels.hpp:122-176 (scale rule);
 *                         elements by the reference's converters (fp6_kernels.hpp:27-88,
 *                         fp4_kernels.hpp:7-30) or by OCP MX v1.0 nearest-even rounding
 *   gain loss / adjoint   FP64 closed forms of the calibration gradients
 *   promote_to_float      tests/xengine_test.cpp:162-167 (GGP two's complement)
 *   cpu_convert_qc_sm_to_ftd  tests/visibility_pipeline_test.cpp:101-112
 *   triangulateFromHerm   include/kernels/tri_from_herm.cuh:28-73
 *   verifyXEngine GEMM    tests/xengine_test.cpp:213-321 (M^H M, TRI / HERM)
 *   cpu_herk_ftd          tests/visibility_pipeline_test.cpp:54-96 (int64)
 *   cpu_gemm              cutlass_interface/test_gemm_api.cu:50-67 (complex GEMM, B [N][K])
 *   power_reduce          cutlass_interface/cutlass_gemm_api.cu:425-437 (float formula)
 *   PillboxGridScatter    include/kernels/visibility_bf.cuh:259-370 (UV cells, conjugate)
 * Parallelism follows the reference's OpenMP "parallel for schedule(dynamic)"
 * over (batch, row) (test_herk_int4.cu:56, :323).
 *
 * Build: -O3 -fopenmp -ffp-contract=off (the epilogue's fma is explicit).
 */
#include "herk_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* mt19937_64 (Matsumoto & Nishimura 2000; parameters as in C++11 std::mt19937_64) */
#define MT_N 312
#define MT_M 156
#define MT_UM 0xFFFFFFFF80000000ULL
#define MT_LM 0x000000007FFFFFFFULL

void ftdo_mt64_seed(ftdo_mt64* s, uint64_t seed) {
    s->mt[0] = seed;
    for (int i = 1; i < MT_N; ++i)
        s->mt[i] = 6364136223846793005ULL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 62)) + (uint64_t)i;
    s->idx = MT_N;
}

static void mt64_twist(ftdo_mt64* s) {
    for (int i = 0; i < MT_N; ++i) {
        uint64_t x = (s->mt[i] & MT_UM) | (s->mt[(i + 1) % MT_N] & MT_LM);
        uint64_t xa = x >> 1;
        if (x & 1ULL) xa ^= 0xB5026F5AA96619E9ULL;
        s->mt[i] = s->mt[(i + MT_M) % MT_N] ^ xa;
    }
    s->idx = 0;
}

uint64_t ftdo_mt64_next(ftdo_mt64* s) {
    if (s->idx >= MT_N) mt64_twist(s);
    uint64_t y = s->mt[s->idx++];
    y ^= (y >> 29) & 0x5555555555555555ULL;
    y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
    y ^= (y << 37) & 0xFFF7EEE000000000ULL;
    y ^= (y >> 43);
    return y;
}

uint64_t ftdo_mt64_nth(uint64_t seed, uint64_t n) {
    ftdo_mt64 s;
    ftdo_mt64_seed(&s, seed);
    uint64_t v = 0;
    for (uint64_t i = 0; i < n; ++i) v = ftdo_mt64_next(&s);
    return v;
}

/* ------------------------------------------------------------------ */
/* QC codec */
uint8_t ftdo_encode_qc(int re, int im) {
    uint8_t rn = (re < 0) ? (uint8_t)(0x8 | ((-re) & 0x7)) : (uint8_t)(re & 0x7);
    uint8_t in = (im < 0) ? (uint8_t)(0x8 | ((-im) & 0x7)) : (uint8_t)(im & 0x7);
    return (uint8_t)((rn << 4) | in);
}

static int decode_nibble(unsigned n) {
    int mag = (int)(n & 0x7);
    return (n & 0x8) ? -mag : mag;
}

void ftdo_decode_qc(uint8_t b, int* re, int* im) {
    *re = decode_nibble((b >> 4) & 0xF);
    *im = decode_nibble(b & 0xF);
}

/* OCP E4M3 encoding of a sign-magnitude nibble, derived from the format
 * (bias 7, 3 mantissa bits): |v| = 2^(e-7) * (1 + m/8). */
static uint8_t nibble_to_fp8(unsigned n) {
    unsigned mag = n & 0x7, sign = (n & 0x8) ? 0x80u : 0u;
    if (mag == 0) return (uint8_t)sign;
    unsigned e = 0;
    while ((mag >> (e + 1)) != 0) ++e;          /* floor(log2 mag) */
    unsigned frac = (mag << 3 >> e) & 0x7;      /* 3-bit mantissa (exact for mag <= 7) */
    return (uint8_t)(sign | ((e + 7) << 3) | frac);
}

uint16_t ftdo_qc_to_fp8_pair(uint8_t b) {
    return (uint16_t)(nibble_to_fp8((b >> 4) & 0xF) | ((unsigned)nibble_to_fp8(b & 0xF) << 8));
}

float ftdo_fp8e4m3_to_float(uint8_t b) {
    int sign = (b & 0x80) ? -1 : 1;
    int e = (b >> 3) & 0xF, m = b & 0x7;
    if (e == 0xF && m == 0x7) return NAN;        /* OCP E4M3: only S.1111.111 is NaN */
    float v = (e == 0) ? ldexpf((float)m / 8.0f, -6) : ldexpf(1.0f + (float)m / 8.0f, e - 7);
    return sign * v;
}

/* ------------------------------------------------------------------ */
/* generators */
void ftdo_gen_qc_uniform(uint64_t seed, int64_t n, uint8_t* out) {
    ftdo_mt64 s;
    ftdo_mt64_seed(&s, seed);
    uint64_t word = 0;
    int left = 0;
    for (int64_t i = 0; i < n; ++i) {
        int v[2];
        for (int c = 0; c < 2; ++c) {
            for (;;) {
                if (left == 0) { word = ftdo_mt64_next(&s); left = 16; }
                unsigned u = (unsigned)(word & 0xF);
                word >>= 4;
                --left;
                if (u != 15) { v[c] = (int)u - 7; break; }
            }
        }
        out[i] = ftdo_encode_qc(v[0], v[1]);
    }
}

void ftdo_gen_bytes(uint64_t seed, int64_t n, uint8_t* out) {
    ftdo_mt64 s;
    ftdo_mt64_seed(&s, seed);
    uint64_t word = 0;
    int left = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (left == 0) { word = ftdo_mt64_next(&s); left = 8; }
        out[i] = (uint8_t)(word & 0xFF);
        word >>= 8;
        --left;
    }
}

/* ------------------------------------------------------------------ */
/* HERK */
static void decode_planes(const uint8_t* A, int16_t* re, int16_t* im, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        int r, m;
        ftdo_decode_qc(A[i], &r, &m);
        re[i] = (int16_t)r;
        im[i] = (int16_t)m;
    }
}

/* One output row i of one batch: C_row[j] for j <= i. int32 partial sums over
 * blocks of 2^20 samples (|term| <= 98) then int64, so any K is exact. */
static void herk_row(const int16_t* re, const int16_t* im, int K, int i, int64_t* Crow) {
    const int16_t* ar = re + (int64_t)i * K;
    const int16_t* ai = im + (int64_t)i * K;
    for (int j = 0; j <= i; ++j) {
        const int16_t* br = re + (int64_t)j * K;
        const int16_t* bi = im + (int64_t)j * K;
        int64_t sr = 0, si = 0;
        for (int k0 = 0; k0 < K; k0 += (1 << 20)) {
            int k1 = (K - k0 > (1 << 20)) ? k0 + (1 << 20) : K;
            int32_t pr = 0, pi = 0;
            for (int k = k0; k < k1; ++k) {
                /* a * conj(b) = (ar br + ai bi) + i (ai br - ar bi) */
                pr += (int32_t)ar[k] * br[k] + (int32_t)ai[k] * bi[k];
                pi += (int32_t)ai[k] * br[k] - (int32_t)ar[k] * bi[k];
            }
            sr += pr;
            si += pi;
        }
        Crow[2 * j] = sr;
        Crow[2 * j + 1] = (i == j) ? 0 : si;
    }
}

void ftdo_herk_int64_rows(const uint8_t* A, int64_t* C, int N, int K, int batch,
                          int row0, int row1, int threads) {
    if (N <= 0 || K < 0 || batch <= 0) return;
    if (row0 < 0) row0 = 0;
    if (row1 > N) row1 = N;
    int64_t nk = (int64_t)N * K;
    int16_t* re = (int16_t*)malloc(sizeof(int16_t) * (size_t)(nk * batch > 0 ? nk * batch : 1));
    int16_t* im = (int16_t*)malloc(sizeof(int16_t) * (size_t)(nk * batch > 0 ? nk * batch : 1));
    decode_planes(A, re, im, nk * batch);
    int64_t packed2 = (int64_t)N * (N + 1); /* scalars per batch */
    int64_t rows = (int64_t)(row1 - row0);
    int64_t total = rows * batch;
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nt)
#endif
    for (int64_t t = 0; t < total; ++t) {
        int b = (int)(t / rows);
        /* heavy rows first: better dynamic balance */
        int i = row1 - 1 - (int)(t % rows);
        int64_t* Crow = C + b * packed2 + (int64_t)i * (i + 1);
        herk_row(re + b * nk, im + b * nk, K, i, Crow);
    }
    (void)threads;
    free(re);
    free(im);
}

void ftdo_herk_int64(const uint8_t* A, int64_t* C, int N, int K, int batch, int threads) {
    ftdo_herk_int64_rows(A, C, N, K, batch, 0, N, threads);
}

static float epilogue(int64_t acc, int diag_im, float alpha, float beta, float old) {
    float v = diag_im ? 0.0f : (float)acc;
    v = v * alpha;
    if (beta != 0.0f) v = fmaf(beta, old, v);
    return v;
}

void ftdo_herk_f32(const uint8_t* A, float* C, int N, int K, int batch,
                   float alpha, float beta, int threads) {
    int64_t packed2 = (int64_t)N * (N + 1);
    int64_t* acc = (int64_t*)malloc(sizeof(int64_t) * (size_t)(packed2 * batch));
    ftdo_herk_int64(A, acc, N, K, batch, threads);
    for (int b = 0; b < batch; ++b)
        for (int i = 0; i < N; ++i)
            for (int j = 0; j <= i; ++j)
                for (int c = 0; c < 2; ++c) {
                    int64_t o = b * packed2 + ((int64_t)i * (i + 1) / 2 + j) * 2 + c;
                    C[o] = epilogue(acc[o], i == j && c == 1, alpha, beta, C[o]);
                }
    free(acc);
}

/* IEEE binary32 -> binary16, round to nearest even, overflow to inf */
uint16_t ftdo_f32_to_f16_rne(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    uint16_t sign = (uint16_t)((x >> 16) & 0x8000u);
    uint32_t ax = x & 0x7FFFFFFFu;
    if (ax >= 0x7F800000u) /* inf / nan */
        return (uint16_t)(sign | 0x7C00u | (ax > 0x7F800000u ? 0x200u : 0u));
    if (ax >= 0x477FF000u) /* >= 65520 rounds to inf */
        return (uint16_t)(sign | 0x7C00u);
    if (ax < 0x33000001u) /* < 2^-25 (or exactly): rounds to 0 */
        return sign;
    int e = (int)(ax >> 23) - 127;
    uint32_t mant = (ax & 0x7FFFFFu) | 0x800000u;
    if (e < -14) { /* subnormal half */
        int shift = -e - 14 + 13;  /* total right shift of the 24-bit significand */
        uint32_t q = mant >> shift;
        uint32_t rem = mant & ((1u << shift) - 1u);
        uint32_t half = 1u << (shift - 1);
        if (rem > half || (rem == half && (q & 1u))) ++q;
        return (uint16_t)(sign | q);
    }
    uint32_t q = mant >> 13, rem = mant & 0x1FFFu;
    uint32_t h = ((uint32_t)(e + 15) << 10) | (q & 0x3FFu);
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) ++h; /* may carry into exponent: correct */
    return (uint16_t)(sign | h);
}

float ftdo_f16_to_f32(uint16_t h) {
    int sign = (h & 0x8000) ? -1 : 1;
    int e = (h >> 10) & 0x1F, m = h & 0x3FF;
    if (e == 0x1F) return m ? NAN : sign * INFINITY;
    float v = (e == 0) ? ldexpf((float)m, -24) : ldexpf((float)(m | 0x400), e - 25);
    return sign * v;
}

void ftdo_herk_f16(const uint8_t* A, uint16_t* C, int N, int K, int batch,
                   float alpha, float beta, int threads) {
    int64_t packed2 = (int64_t)N * (N + 1);
    int64_t* acc = (int64_t*)malloc(sizeof(int64_t) * (size_t)(packed2 * batch));
    ftdo_herk_int64(A, acc, N, K, batch, threads);
    for (int b = 0; b < batch; ++b)
        for (int i = 0; i < N; ++i)
            for (int j = 0; j <= i; ++j)
                for (int c = 0; c < 2; ++c) {
                    int64_t o = b * packed2 + ((int64_t)i * (i + 1) / 2 + j) * 2 + c;
                    float old = ftdo_f16_to_f32(C[o]);
                    C[o] = ftdo_f32_to_f16_rne(epilogue(acc[o], i == j && c == 1, alpha, beta, old));
                }
    free(acc);
}

/* ------------------------------------------------------------------ */
/* MX block-scaled quantization */
static void mx_params(int fmt, int* mbits, int* bias, float* maxv) {
    switch (fmt) {
        case 0: *mbits = 1; *bias = 1; *maxv = 6.0f; break;   /* FP4 E2M1 */
        case 1: *mbits = 2; *bias = 3; *maxv = 28.0f; break;  /* FP6 E3M2 */
        default: *mbits = 3; *bias = 1; *maxv = 7.5f; break;  /* FP6 E2M3 */
    }
}

static int mx_sign_shift(int fmt) { return fmt == 0 ? 3 : 5; }

uint8_t ftdo_mx_encode(float x, int fmt) {
    int mbits, bias;
    float maxv;
    mx_params(fmt, &mbits, &bias, &maxv);
    const unsigned sign = signbit(x) ? 1u : 0u;
    const float a = fabsf(x);
    const int ebits = (fmt == 1) ? 3 : 2;
    const unsigned maxcode = (((1u << ebits) - 1u) << mbits) | ((1u << mbits) - 1u);
    unsigned code;
    if (a >= maxv) {
        code = maxcode;
    } else {
        int e;
        frexpf(a, &e);
        e -= 1; /* a in [2^e, 2^(e+1)) */
        const int emin = 1 - bias;
        if (a == 0.0f || e < emin) {
            /* subnormal grid step 2^(emin - mbits); q may reach 2^mbits = min normal */
            code = (unsigned)rintf(ldexpf(a, mbits - emin));
        } else {
            unsigned q = (unsigned)rintf(ldexpf(a, mbits - e)); /* in [2^m, 2^(m+1)] */
            if (q == (2u << mbits)) {
                q >>= 1;
                ++e;
            }
            code = ((unsigned)(e + bias) << mbits) | (q - (1u << mbits));
        }
        if (code > maxcode) code = maxcode;
    }
    return (uint8_t)((sign << mx_sign_shift(fmt)) | code);
}

/* The reference's element converters, step for step:
 *   FP6 E3M2  cutlass_interface/gemm_complex_sm100/fp6_kernels.hpp:27-58
 *   FP6 E2M3  fp6_kernels.hpp:64-88
 *   FP4 E2M1  cutlass_interface/gemm_complex_sm100/fp4_kernels.hpp:7-30
 * Clamp to +-max with fminf/fmaxf (a NaN becomes -max), magnitudes below 0.25
 * (E2M3: 0.0625) become a zero carrying the sign of x < 0, otherwise the
 * frexp exponent is clamped to the normal range and the mantissa is
 * rint(frac * 2^mbits) clamped to 2^mbits - 1 (no carry into the exponent);
 * E2M1 takes its one mantissa bit as frac >= 0.5 (truncation). */
uint8_t ftdo_mx_encode_ref(float x, int fmt) {
    int mbits, bias;
    float maxv;
    mx_params(fmt, &mbits, &bias, &maxv);
    const int s = mx_sign_shift(fmt);
    const int emax_biased = (fmt == 1) ? 7 : 3;
    const float flush = (fmt == 2) ? 0.0625f : 0.25f;
    const float v = fminf(fmaxf(x, -maxv), maxv);
    const unsigned sign = (v < 0.0f) ? 1u : 0u;
    const float a = fabsf(v);
    if (a < flush) return (uint8_t)(sign << s);
    int e = 0;
    float frac = frexpf(a, &e);
    frac *= 2.0f;
    e -= 1;
    int eb = e + bias;
    if (eb < 1) eb = 1;
    if (eb > emax_biased) eb = emax_biased;
    const float mf = frac - 1.0f;
    int m;
    if (fmt == 0) {
        m = (mf >= 0.5f) ? 1 : 0;
    } else {
        m = (int)rintf(mf * (float)(1 << mbits));
        if (m > (1 << mbits) - 1) m = (1 << mbits) - 1;
        if (m < 0) m = 0;
    }
    return (uint8_t)((sign << s) | ((unsigned)eb << mbits) | (unsigned)m);
}

double ftdo_mx_decode(uint8_t code, int fmt) {
    int mbits, bias;
    float maxv;
    mx_params(fmt, &mbits, &bias, &maxv);
    const int s = mx_sign_shift(fmt);
    const unsigned mag = code & ((1u << s) - 1u);
    const unsigned e = mag >> mbits, m = mag & ((1u << mbits) - 1u);
    double v = (e == 0) ? ldexp((double)m, 1 - bias - mbits)
                        : ldexp((double)(m | (1u << mbits)), (int)e - bias - mbits);
    return ((code >> s) & 1u) ? -v : v;
}

uint8_t ftdo_mx_scale_exponent(float amax, int fmt) {
    int mbits, bias;
    float maxv;
    mx_params(fmt, &mbits, &bias, &maxv);
    if (amax == 0.0f) return 0;
    const float ratio = amax / maxv;
    if (ratio == 0.0f) return 0; /* denormal amax underflowing the ratio: smallest scale */
    int e;
    const float m = frexpf(ratio, &e); /* ratio = m * 2^e, m in [0.5, 1) */
    int p = (m == 0.5f) ? e - 1 : e;   /* ceil(log2(ratio)) */
    int b = p + 127;
    if (b < 0) b = 0;
    if (b > 254) b = 254;
    return (uint8_t)b;
}

void ftdo_mx_quantize(const float* X, int rows, int K, int fmt, uint8_t* codes, uint8_t* scales) {
    ftdo_mx_quantize_mode(X, rows, K, fmt, 1, codes, scales);
}

void ftdo_mx_quantize_mode(const float* X, int rows, int K, int fmt, int rounding, uint8_t* codes,
                           uint8_t* scales) {
    const int nblk = (K + 31) / 32;
    for (int r = 0; r < rows; ++r)
        for (int b = 0; b < nblk; ++b)
            for (int part = 0; part < 2; ++part) {
                float amax = 0.0f;
                for (int i = 0; i < 32; ++i) {
                    const int k = b * 32 + i;
                    if (k < K) {
                        const float v = fabsf(X[((int64_t)r * K + k) * 2 + part]);
                        if (v > amax) amax = v;
                    }
                }
                const uint8_t se = ftdo_mx_scale_exponent(amax, fmt);
                scales[((int64_t)r * nblk + b) * 2 + part] = se;
                const float inv = ldexpf(1.0f, 127 - (int)se);
                for (int i = 0; i < 32; ++i) {
                    const int k = b * 32 + i;
                    const float v = (k < K) ? X[((int64_t)r * K + k) * 2 + part] : 0.0f;
                    codes[(((int64_t)r * nblk + b) * 2 + part) * 32 + i] =
                        rounding == 0 ? ftdo_mx_encode_ref(v * inv, fmt) : ftdo_mx_encode(v * inv, fmt);
                }
            }
}

void ftdo_herk_c64(const double* X, double* C, int N, int K, int batch, int threads) {
    const int64_t total = (int64_t)N * batch;
    const int64_t p2 = (int64_t)N * (N + 1);
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nt)
#endif
    for (int64_t t = 0; t < total; ++t) {
        const int b = (int)(t / N), i = N - 1 - (int)(t % N);
        const double* xi = X + ((int64_t)b * N + i) * K * 2;
        double* Crow = C + b * p2 + (int64_t)i * (i + 1);
        for (int j = 0; j <= i; ++j) {
            const double* xj = X + ((int64_t)b * N + j) * K * 2;
            double sr = 0.0, si = 0.0;
            for (int k = 0; k < K; ++k) {
                const double ar = xi[2 * k], ai = xi[2 * k + 1];
                const double br = xj[2 * k], bi = xj[2 * k + 1];
                sr += ar * br + ai * bi;
                si += ai * br - ar * bi;
            }
            Crow[2 * j] = sr;
            Crow[2 * j + 1] = (i == j) ? 0.0 : si;
        }
    }
    (void)threads;
}

/* ------------------------------------------------------------------ */
/* Gain calibration and HERK adjoint (FP64).  Derivation: with z_ij =
 * g_i conj(g_j) M_ij and R = V - z, dL = -2 Re sum conj(R) dz; collecting the
 * coefficients of dg_k and conj(dg_k) gives the gradient below; dL/dM gives
 * Gamma; through M = X X^H, dX = (Gamma + Gamma^H) X. */
void ftdo_gain_loss_grad_f64(const double* V, const double* M, const double* g, int N, int batch,
                             double* loss, double* grad, double* Gamma) {
    const int64_t P = (int64_t)N * (N + 1) / 2;
    for (int b = 0; b < batch; ++b) {
        double L = 0.0;
        double* gr = grad + (int64_t)b * N * 2;
        for (int k = 0; k < 2 * N; ++k) gr[k] = 0.0;
        const double* gb = g + (int64_t)b * N * 2;
        for (int i = 0; i < N; ++i)
            for (int j = 0; j <= i; ++j) {
                const int64_t o = ((int64_t)b * P + (int64_t)i * (i + 1) / 2 + j) * 2;
                const double gir = gb[2 * i], gii = gb[2 * i + 1];
                const double gjr = gb[2 * j], gji = gb[2 * j + 1];
                /* p = g_i conj(g_j) */
                const double pr = gir * gjr + gii * gji, pi = gii * gjr - gir * gji;
                const double mr = M[o], mi = M[o + 1];
                const double zr = pr * mr - pi * mi, zi = pr * mi + pi * mr;
                const double rr = V[o] - zr, ri = V[o + 1] - zi;
                L += rr * rr + ri * ri;
                if (Gamma) { /* -2 R conj(p) */
                    Gamma[o] = -2.0 * (rr * pr + ri * pi);
                    Gamma[o + 1] = -2.0 * (ri * pr - rr * pi);
                }
                /* row i term: -2 R conj(M) g_j */
                const double ar = rr * mr + ri * mi, ai = ri * mr - rr * mi;
                gr[2 * i] += -2.0 * (ar * gjr - ai * gji);
                gr[2 * i + 1] += -2.0 * (ar * gji + ai * gjr);
                /* column j term: -2 conj(R) M g_i */
                const double br = rr * mr + ri * mi, bi = rr * mi - ri * mr;
                gr[2 * j] += -2.0 * (br * gir - bi * gii);
                gr[2 * j + 1] += -2.0 * (br * gii + bi * gir);
            }
        loss[b] = L;
    }
}

void ftdo_herk_adjoint_f64(const double* Gamma, const double* X, double* dX, int N, int K,
                           int batch, int threads) {
    const int64_t P = (int64_t)N * (N + 1) / 2;
    const int64_t total = (int64_t)batch * N;
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nt)
#endif
    for (int64_t t = 0; t < total; ++t) {
        const int b = (int)(t / N), i = (int)(t % N);
        double* out = dX + ((int64_t)b * N + i) * K * 2;
        for (int k = 0; k < 2 * K; ++k) out[k] = 0.0;
        for (int j = 0; j < N; ++j) {
            double hr, hi;
            if (i > j) {
                const int64_t o = ((int64_t)b * P + (int64_t)i * (i + 1) / 2 + j) * 2;
                hr = Gamma[o];
                hi = Gamma[o + 1];
            } else if (i < j) {
                const int64_t o = ((int64_t)b * P + (int64_t)j * (j + 1) / 2 + i) * 2;
                hr = Gamma[o];
                hi = -Gamma[o + 1];
            } else {
                const int64_t o = ((int64_t)b * P + (int64_t)i * (i + 1) / 2 + i) * 2;
                hr = 2.0 * Gamma[o];
                hi = 0.0;
            }
            const double* xj = X + ((int64_t)b * N + j) * K * 2;
            for (int k = 0; k < K; ++k) {
                out[2 * k] += hr * xj[2 * k] - hi * xj[2 * k + 1];
                out[2 * k + 1] += hr * xj[2 * k + 1] + hi * xj[2 * k];
            }
        }
    }
    (void)threads;
}

/* ------------------------------------------------------------------ */
/* stage converters */
static int twos_nibble(unsigned n) { return (int)((n & 0xF) ^ 0x8) - 8; }

void ftdo_promote_data(const uint8_t* in, int64_t n, double* out) {
    for (int64_t i = 0; i < n; ++i) {
        out[2 * i] = (double)twos_nibble(in[i] & 0xF);
        out[2 * i + 1] = (double)twos_nibble(in[i] >> 4);
    }
}

uint8_t ftdo_convert_qc_to_ftd(uint8_t b) {
    int re, im;
    ftdo_decode_qc(b, &re, &im);
    return (uint8_t)((((unsigned)im & 0xF) << 4) | ((unsigned)re & 0xF));
}

uint16_t ftdo_qc_nibble_fp16(unsigned nibble) {
    const unsigned m = nibble & 7u;
    const uint16_t sign = (nibble & 8u) ? 0x8000u : 0u;
    return (uint16_t)(sign | ftdo_f32_to_f16_rne((float)m));
}

uint8_t ftdo_f32_to_e4m3_clamped(float x) {
    /* fminf(fmaxf(x, -448), 448) as cast_kernels_common.hpp:36-37: a NaN becomes -448 */
    x = fminf(fmaxf(x, -448.0f), 448.0f);
    const unsigned sign = signbit(x) ? 0x80u : 0u;
    const float a = fabsf(x);
    unsigned code;
    int e;
    frexpf(a, &e);
    e -= 1;
    if (a == 0.0f || e < -6) {
        code = (unsigned)rintf(ldexpf(a, 9)); /* subnormal step 2^-9 */
    } else {
        unsigned q = (unsigned)rintf(ldexpf(a, 3 - e));
        if (q == 16u) {
            q = 8u;
            ++e;
        }
        code = ((unsigned)(e + 7) << 3) | (q - 8u);
    }
    if (code > 0x7Eu) code = 0x7Eu;
    return (uint8_t)(sign | code);
}

void ftdo_tri_from_herm(const double* full, double* tri, int N, int batch) {
    const int64_t P = (int64_t)N * (N + 1) / 2;
    for (int b = 0; b < batch; ++b)
        for (int i = 0; i < N; ++i)
            for (int j = i; j < N; ++j) {
                const int64_t src = (((int64_t)b * N + i) * N + j) * 2;
                const int64_t dst = ((int64_t)b * P + (int64_t)i * N + j - (int64_t)i * (i + 1) / 2) * 2;
                tri[dst] = full[src];
                tri[dst + 1] = full[src + 1];
            }
}

/* ------------------------------------------------------------------ */
/* GGP XEngine: per batch b, M_b = K x N row-major (antenna fastest) of
 * promote_to_float values (tests/xengine_test.cpp:162-167: two's-complement
 * nibbles, low = Re), C_b = M_b^H M_b (xengine_test.cpp:254-279).  TRI: upper
 * triangle row-major, (q, p >= q) at q*N + p - q(q+1)/2 (xengine_test.cpp:282-287,
 * tri_from_herm.cuh:28-73); HERM: the full N x N matrix row-major.  int64 exact. */
void ftdo_xengine_ggp_i64(const uint8_t* in, int N, int K, int batch, int herm, int64_t* out,
                          int threads) {
    if (N <= 0 || batch <= 0) return;
    const int64_t nk = (int64_t)N * (K > 0 ? K : 0);
    int16_t* re = (int16_t*)malloc(sizeof(int16_t) * (size_t)(nk * batch > 0 ? nk * batch : 1));
    int16_t* im = (int16_t*)malloc(sizeof(int16_t) * (size_t)(nk * batch > 0 ? nk * batch : 1));
    /* transpose to [b][antenna][time] planes */
    for (int b = 0; b < batch; ++b)
        for (int t = 0; t < K; ++t)
            for (int a = 0; a < N; ++a) {
                const uint8_t v = in[(int64_t)b * nk + (int64_t)t * N + a];
                re[(int64_t)b * nk + (int64_t)a * K + t] = (int16_t)twos_nibble(v & 0xF);
                im[(int64_t)b * nk + (int64_t)a * K + t] = (int16_t)twos_nibble(v >> 4);
            }
    const int64_t per = herm ? (int64_t)N * N : (int64_t)N * (N + 1) / 2;
    const int64_t total = (int64_t)batch * N;
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nt)
#endif
    for (int64_t t = 0; t < total; ++t) {
        const int b = (int)(t / N), q = (int)(t % N);
        const int16_t* qr = re + b * nk + (int64_t)q * K;
        const int16_t* qi = im + b * nk + (int64_t)q * K;
        int64_t* o = out + b * per * 2;
        for (int p = q; p < N; ++p) {
            const int16_t* pr = re + b * nk + (int64_t)p * K;
            const int16_t* pi = im + b * nk + (int64_t)p * K;
            int64_t sr = 0, si = 0;
            for (int k0 = 0; k0 < K; k0 += (1 << 20)) {
                const int k1 = (K - k0 > (1 << 20)) ? k0 + (1 << 20) : K;
                int32_t ar = 0, ai = 0;
                for (int k = k0; k < k1; ++k) {
                    /* conj(x_q) x_p = (qr pr + qi pi) + i (qr pi - qi pr) */
                    ar += (int32_t)qr[k] * pr[k] + (int32_t)qi[k] * pi[k];
                    ai += (int32_t)qr[k] * pi[k] - (int32_t)qi[k] * pr[k];
                }
                sr += ar;
                si += ai;
            }
            if (herm) {
                o[((int64_t)q * N + p) * 2] = sr;
                o[((int64_t)q * N + p) * 2 + 1] = si;
                o[((int64_t)p * N + q) * 2] = sr;
                o[((int64_t)p * N + q) * 2 + 1] = -si;
            } else {
                const int64_t idx = (int64_t)q * N + p - (int64_t)q * (q + 1) / 2;
                o[idx * 2] = sr;
                o[idx * 2 + 1] = si;
            }
        }
    }
    (void)threads;
    free(re);
    free(im);
}

/* FTD raw-order HERK (cpu_herk_ftd, tests/visibility_pipeline_test.cpp:54-96), int64:
 * batch b = (chan * n_pol + pol) * n_ti + ti; x_a[k] is the byte at
 * ((k * n_ti + ti) * n_ant + a) * n_chan * n_pol + chan * n_pol + pol; output the
 * packed lower triangle of sum_k x_i conj(x_j) per batch.  twos != 0: two's-complement
 * nibbles, low = Re (FTD, lib/tcc_interface.cu:124-133); else QC sign-magnitude. */
void ftdo_herk_ftd_i64(const uint8_t* raw, int n_ant, int n_chan, int n_time, int n_ti, int n_pol,
                       int twos, int64_t* out, int threads) {
    if (n_ant <= 0 || n_chan <= 0 || n_pol <= 0 || n_ti <= 0) return;
    const int K = n_time / n_ti;
    const int batch = n_chan * n_pol * n_ti;
    const int64_t nk = (int64_t)n_ant * K;
    int16_t* re = (int16_t*)malloc(sizeof(int16_t) * (size_t)(nk * batch > 0 ? nk * batch : 1));
    int16_t* im = (int16_t*)malloc(sizeof(int16_t) * (size_t)(nk * batch > 0 ? nk * batch : 1));
    const int64_t C = (int64_t)n_chan * n_pol;
    for (int b = 0; b < batch; ++b) {
        const int ti = b % n_ti, pol = (b / n_ti) % n_pol, chan = b / (n_pol * n_ti);
        for (int a = 0; a < n_ant; ++a)
            for (int k = 0; k < K; ++k) {
                const int64_t t = (int64_t)k * n_ti + ti;
                const uint8_t v = raw[(t * n_ant + a) * C + (int64_t)chan * n_pol + pol];
                int r, m;
                if (twos) {
                    r = twos_nibble(v & 0xF);
                    m = twos_nibble(v >> 4);
                } else {
                    ftdo_decode_qc(v, &r, &m);
                }
                re[(int64_t)b * nk + (int64_t)a * K + k] = (int16_t)r;
                im[(int64_t)b * nk + (int64_t)a * K + k] = (int16_t)m;
            }
    }
    const int64_t packed2 = (int64_t)n_ant * (n_ant + 1);
    const int64_t total = (int64_t)batch * n_ant;
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nt)
#endif
    for (int64_t t = 0; t < total; ++t) {
        const int b = (int)(t / n_ant);
        const int i = n_ant - 1 - (int)(t % n_ant);
        herk_row(re + b * nk, im + b * nk, K, i, out + b * packed2 + (int64_t)i * (i + 1));
    }
    (void)threads;
    free(re);
    free(im);
}

/* ------------------------------------------------------------------ */
/* Complex GEMM (cpu_gemm, cutlass_interface/test_gemm_api.cu:50-67): per batch
 * C[i][j] = sum_k A[i][k] B[j][k] (B stored [N][K], no conjugate), FP64, complex
 * interleaved.  Exact for integer inputs while sums stay below 2^53. */
void ftdo_gemm_c64(const double* A, const double* B, double* C, int M, int N, int K, int batch,
                   int threads) {
    if (M <= 0 || N <= 0 || batch <= 0) return;
    const int64_t total = (int64_t)batch * M;
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nt)
#endif
    for (int64_t t = 0; t < total; ++t) {
        const int64_t b = t / M, i = t % M;
        const double* a = A + ((b * M + i) * (int64_t)K) * 2;
        for (int j = 0; j < N; ++j) {
            const double* bb = B + ((b * N + j) * (int64_t)K) * 2;
            double sr = 0.0, si = 0.0;
            for (int k = 0; k < K; ++k) {
                sr += a[2 * k] * bb[2 * k] - a[2 * k + 1] * bb[2 * k + 1];
                si += a[2 * k] * bb[2 * k + 1] + a[2 * k + 1] * bb[2 * k];
            }
            C[((b * M + i) * (int64_t)N + j) * 2] = sr;
            C[((b * M + i) * (int64_t)N + j) * 2 + 1] = si;
        }
    }
    (void)threads;
}

/* Fused power detection of the GEMM epilogue, in the float operations the library
 * evaluates: p = alpha * fmaf(re, re, im * im); with beta != 0, p = fmaf(beta, old, p)
 * (the reference's alpha * (re^2 + im^2) + beta * old, cutlass_gemm_api.cu:425-437,
 * with its contraction fixed). */
void ftdo_power_f32(const float* re, const float* im, const float* old, int64_t n, float alpha,
                    float beta, float* out) {
    for (int64_t i = 0; i < n; ++i) {
        float p = alpha * fmaf(re[i], re[i], im[i] * im[i]);
        if (beta != 0.0f) p = fmaf(beta, old[i], p);
        out[i] = p;
    }
}

/* ------------------------------------------------------------------ */
/* Voltage beamformer tail: from exact beams re / im [batch][2 pol][M][n_beam] (FP32) to
 * the filterbank out [batch][n_beam][M / nt], uint8 (out_u8) or FP32.
 * mode 0, the FP8 path (lib/voltage_pipeline.cu:886-910, 934-967): the power GEMM of
 *   pol 0 with beta = 0, of pol 1 with beta = 1 into the same buffer (power epilogue
 *   p = alpha * fma(re, re, im * im), then fma(beta, old, p); cutlass_gemm_api.cu:425-437
 *   with its contraction fixed), then time_integrate_quantize_kernel
 *   (lib/voltage_pipeline.cu:283-318): acc = 0; acc += P[t] for nt samples in order.
 * mode 1, the FP6 path (power_sum_quantize_kernel, lib/voltage_pipeline.cu:327-370):
 *   power += ra*ra + ia*ia + rb*rb + ib*ib, contracted as
 *   fma(ib, ib, fma(rb, rb, fma(ia, ia, ra * ra))).
 * uint8: __float2int_rn(acc * scale) clamped to [0, 255] (nearest-even; NaN -> 0). */
static uint8_t vq_u8(float v, float scale) {
    float q = rintf(v * scale);
    q = fmaxf(q, 0.0f);  /* NaN -> 0 */
    q = fminf(q, 255.0f);
    return (uint8_t)(int)q;
}

void ftdo_voltage_tail(const float* re, const float* im, int64_t batch, int M, int n_beam, int nt, int mode,
                       float scale, int out_u8, void* out) {
    const int n_out = M / nt;
    for (int64_t b = 0; b < batch; ++b)
        for (int beam = 0; beam < n_beam; ++beam)
            for (int t = 0; t < n_out; ++t) {
                float acc = 0.0f;
                for (int dt = 0; dt < nt; ++dt) {
                    const int64_t i0 = ((b * 2 + 0) * M + (int64_t)t * nt + dt) * n_beam + beam;
                    const int64_t i1 = ((b * 2 + 1) * M + (int64_t)t * nt + dt) * n_beam + beam;
                    if (mode == 0) {
                        const float p0 = 1.0f * fmaf(re[i0], re[i0], im[i0] * im[i0]);
                        float p = 1.0f * fmaf(re[i1], re[i1], im[i1] * im[i1]);
                        p = fmaf(1.0f, p0, p);
                        acc += p;
                    } else {
                        const float ra = re[i0], ia = im[i0], rb = re[i1], ib = im[i1];
                        acc += fmaf(ib, ib, fmaf(rb, rb, fmaf(ia, ia, ra * ra)));
                    }
                }
                const int64_t o = (b * n_beam + beam) * n_out + t;
                if (out_u8) ((uint8_t*)out)[o] = vq_u8(acc, scale);
                else ((float*)out)[o] = acc;
            }
}

/* ------------------------------------------------------------------ */
/* Pillbox UV gridding (PillboxGridScatter, include/kernels/visibility_bf.cuh:259-370;
 * uv_cell, include/kernels/imaging_kernels.cuh:33-45): sequential scatter-add of
 * Nf packed triangles vis [Nf][n_bl][2] into grid [Nf][Ng][Ng][2]; the conjugate goes to
 * (-u, -v) for cross-correlations.  The baseline index is decoded exactly. */
void ftdo_pillbox_grid(const float* vis, const float* uv_m, const double* freq_hz, float* grid,
                       int64_t n_bl, int Ng, int Nf, int freq_offset, float cell) {
    const double C_LIGHT = 299792458.0;
    const float half_ng = 0.5f * (float)Ng;
    const int64_t plane = (int64_t)Ng * Ng;
    for (int f = 0; f < Nf; ++f) {
        const double lambda = C_LIGHT / freq_hz[freq_offset + f];
        int64_t k = 0;
        for (int row = 0; k < n_bl; ++row)
            for (int col = 0; col <= row && k < n_bl; ++col, ++k) {
                const float vre = vis[2 * ((int64_t)f * n_bl + k)];
                const float vim = vis[2 * ((int64_t)f * n_bl + k) + 1];
                const float u_lam = (float)((double)uv_m[2 * k] / lambda);
                const float v_lam = (float)((double)uv_m[2 * k + 1] / lambda);
                const int iu = (int)roundf(u_lam / cell + half_ng);
                const int iv = (int)roundf(v_lam / cell + half_ng);
                if (iu >= 0 && iu < Ng && iv >= 0 && iv < Ng) {
                    float* g = grid + 2 * ((int64_t)f * plane + (int64_t)iv * Ng + iu);
                    g[0] += vre;
                    g[1] += vim;
                }
                if (row != col) {
                    const int iuc = (int)roundf(-u_lam / cell + half_ng);
                    const int ivc = (int)roundf(-v_lam / cell + half_ng);
                    if (iuc >= 0 && iuc < Ng && ivc >= 0 && ivc < Ng) {
                        float* g = grid + 2 * ((int64_t)f * plane + (int64_t)ivc * Ng + iuc);
                        g[0] += vre;
                        g[1] += -vim;
                    }
                }
            }
    }
}
