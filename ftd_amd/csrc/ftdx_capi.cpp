// ftdx_capi.cpp -- extern "C" entry points of libftdx.so (declared in include/ftdx.h).
//
// Validation, launch-configuration choice (explicit setting, tuning cache or the
// autotuner), and dispatch of the plain HERK; the reference counterpart of each
// entry point is cited in ftdx.h.
#include "ftdx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "ftdx_internal.h"
#include "ftdx_tune.h"
#include "herk_fp8.h"
#include "herk_mx.h"
#include "calib.h"
#include "stages.h"
#include "imaging.h"

struct ftdx_context {
    int device = 0;
    int herk_mode = FTDX_HERK_AUTO;
    int direct_tile = FTDX_DIRECT_TILE_AUTO;
    int forced_cfg = -1;     // ftdx_set_launch_config
    int forced_remap = -1;
    int graph = 0;
    int strategy_verbosity = 1;
    int kernel_verbosity = 0;
    std::string tune_cache_path;
    std::string kernel_cache_path;
    bool tune_loaded = false;
    ftdx::TuneCache tune;
    int init_N = 0, init_K = 0, init_batch = 0;
    // MX element rounding (ftdx_set_mx_rounding): the reference's converters by default
    int mx_rounding = FTDX_MX_ROUND_REFERENCE;
    // MX path scratch (operand rows: codes with their scales inline), grown on demand
    uint8_t* mx_codes = nullptr;
    size_t mx_codes_cap = 0;
    // calibration scratch: gradient partials, loss partials, dense Hermitian matrix
    uint8_t* cal_part = nullptr;
    size_t cal_part_cap = 0;
    uint8_t* cal_herm = nullptr;
    size_t cal_herm_cap = 0;
    // pre-cast FP8 operand rows for the unfused / FP16-input paths
    uint8_t* precast = nullptr;
    size_t precast_cap = 0;
    // plain-path graph replay (set_direct_herk_graph): small LRU of captured calls
    struct GraphEntry {
        bool used = false;
        int N = 0, K = 0, batch = 0, input = 0, output = 0, mode = 0;
        int cfg = -1, remap = -1;  // launch configuration the graph was captured with
        float alpha = 0.f, beta = 0.f;
        const void* A = nullptr;
        void* C = nullptr;
        hipGraphExec_t exec = nullptr;
        unsigned long long stamp = 0;
    };
    GraphEntry graphs[4];
    unsigned long long graph_clock = 0;
    long long graph_captures = 0;  // graphs captured so far (ftdx_herk_graph_stats)
    hipStream_t cap_stream = nullptr;
    // complex GEMM: prepared conj(B) rows (prepare_b) and a separate buffer for the
    // one-shot gemm(), so it never clobbers a prepared B
    // (FP8 compute: pre-cast rows; FP6 / FP4 compute: MX codes in *_b, scales in *_bs)
    uint8_t* prep_b = nullptr;
    size_t prep_b_cap = 0;
    int prep_N = -1, prep_K = -1, prep_batch = -1;
    int prep_fmt = -1;  // -1: FP8 rows, else the MxFormat of the codes
    // the reference's GEMM dispatch settings, kept for its getters (one kernel family here)
    int gemm_settings[5] = {0, 0, 0, 1, 1};
    std::string gemm_tune_cache_path;
    uint8_t* gemm_b = nullptr;
    size_t gemm_b_cap = 0;
};

namespace ftdx {

namespace {
thread_local std::string g_last_error;
thread_local ftdx_context* g_default_ctx = nullptr;

// Default launch configuration when nothing is tuned or set.
constexpr int kDefaultCfg = TILE_128x128_K64;
constexpr int kDefaultRemap = 1;
// MX HERK (block-scaled operand rows) default tile
constexpr int kMxDefaultCfg = TILE_256x128_K64;

// DirectTileConfig (reference names, cutlass_gemm_api.h:113-119) -> gfx950 tile
// configuration: the reference enum picks K chunk x pipeline depth; here it
// picks among the compiled tile shapes.
int direct_tile_to_cfg(int d) {
    switch (d) {
        case FTDX_DIRECT_TILE_K32_B3: return TILE_64x64_K64;
        case FTDX_DIRECT_TILE_K64_B2: return TILE_128x128_K64;
        case FTDX_DIRECT_TILE_K64_B3: return TILE_256x128_K64;
        case FTDX_DIRECT_TILE_K128_B2: return TILE_128x128_K128;
        default: return -1;
    }
}

std::string default_tune_path() {
    const char* e = std::getenv("FTDX_TUNE_CACHE_PATH");
    return e ? std::string(e) : std::string();
}

}  // namespace

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(FTDX_ERR_HIP, std::string(what) + ": " + hipGetErrorName(e) + " (" +
                                  hipGetErrorString(e) + ")");
}

const uint8_t* prepared_fp8_rows(const ftdx_context* ctx, int N, int K, int batch) {
    if (!ctx || !ctx->prep_b || ctx->prep_fmt >= 0 || ctx->prep_N != N || ctx->prep_K != K ||
        ctx->prep_batch != batch)
        return nullptr;
    return ctx->prep_b;
}

namespace {

void ensure_tune_loaded(ftdx_context* ctx) {
    if (ctx->tune_loaded) return;
    ctx->tune_loaded = true;
    if (ctx->tune_cache_path.empty()) ctx->tune_cache_path = default_tune_path();
    if (!ctx->tune_cache_path.empty()) {
        int skipped = 0;
        int n = ctx->tune.load(ctx->tune_cache_path, &skipped);
        if (ctx->strategy_verbosity >= 2 && n >= 0)
            std::printf("ftdx tune: loaded %d entries from %s (%d lines skipped)\n", n,
                        ctx->tune_cache_path.c_str(), skipped);
    }
}

// Explicit choice (launch config / direct tile), else cache, else default.
void pick_config(ftdx_context* ctx, const TuneKey& key, int& cfg, int& remap, bool& from_cache) {
    from_cache = false;
    cfg = ctx->forced_cfg >= 0 ? ctx->forced_cfg : direct_tile_to_cfg(ctx->direct_tile);
    remap = ctx->forced_remap;
    TuneEntry e;
    if (cfg < 0 && ctx->tune.find(key, e)) {
        cfg = e.cfg;
        if (remap < 0) remap = e.xcd_remap;
        from_cache = true;
    }
    if (cfg < 0) cfg = kDefaultCfg;
    if (remap < 0) remap = kDefaultRemap;
}

// Compute units of the current device, cached per device.
int device_cu_count() {
    static std::atomic<int> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return 0;
    if (dev < 64) {
        const int c = cache[dev].load(std::memory_order_relaxed);
        if (c > 0) return c;
    }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (dev < 64) cache[dev].store(cus, std::memory_order_relaxed);
    return cus;
}

// Pre-cast paths (FP16 input, ForceBaseline, E3M2 operands) convert and multiply a chunk
// of batches at a time: the chunk's operand rows (at most kPrecastChunkBytes) are then
// still in the memory-side cache when its tiles read them, and the scratch holds one
// chunk.  Among chunk sizes up to that cap, the one whose launches leave the fewest idle
// workgroup slots in their last rounds (plus a small cost per launch) wins.
constexpr size_t kPrecastChunkBytes = size_t(320) << 20;

// batches of `per` scratch bytes each that fit one chunk
int chunk_cap_bytes(size_t per, int batch) {
    const size_t cap = per > 0 ? kPrecastChunkBytes / per : (size_t)batch;
    return (int)std::max<size_t>(1, std::min<size_t>(cap, (size_t)std::max(batch, 1)));
}

int precast_chunk_cap(int N, int K, int batch) {
    return chunk_cap_bytes((size_t)N * 2 * precast_row_samples(K), batch);
}

// units_out: output batches per unit of `batch` (a unit is one group of polarization
// segments on the polarization-sum path)
int pick_chunk(int cfg, int N, int batch, int cmax, int units_out = 1) {
    if (cmax >= batch) return std::max(batch, 1);
    const int64_t slots = (int64_t)std::max(device_cu_count(), 1) * tile_config_wg_per_cu(cfg);
    const int64_t tiles = herk_tiles(cfg, N) * units_out;
    double best = 1e30;
    int pick = cmax;
    for (int c = cmax; c >= 1; --c) {
        int64_t idle = 0;
        for (int b0 = 0; b0 < batch; b0 += c) {
            const int64_t t = tiles * std::min(c, batch - b0);
            idle += (t + slots - 1) / slots * slots - t;
        }
        const double cost = (double)idle / (double)slots + 0.1 * ((batch + c - 1) / c);
        if (cost < best) {
            best = cost;
            pick = c;
        }
    }
    return pick;
}

// The pre-cast path of one call: per chunk of batches, convert A's rows into the
// context's scratch (FP8 planar rows, or E3M2 rows when bf6), then run the tile over
// them.  The scratch must hold N * precast_chunk_cap(...) rows of 2 * Kp bytes.
hipError_t run_precast(ftdx_context* ctx, const HerkPlainArgs& a, int input, bool bf6, int cfg, bool fp16,
                       int remap, hipStream_t st) {
    const int Kp = precast_row_samples(a.K);
    const int chunk = pick_chunk(cfg, a.N, a.batch, precast_chunk_cap(a.N, a.K, a.batch));
    const int64_t in_bytes = (int64_t)a.N * a.K * (input == FTDX_INPUT_FP16 ? 4 : 1);
    const int64_t out_bytes = a.c_bstride * (fp16 ? 2 : 4);
    for (int b0 = 0; b0 < a.batch; b0 += chunk) {
        const int nb = std::min(chunk, a.batch - b0);
        const uint8_t* src = a.A + (int64_t)b0 * in_bytes;
        const int64_t rows = (int64_t)a.N * nb;
        hipError_t e = input == FTDX_INPUT_FP16
            ? launch_promote_fp16_fp8_planar(reinterpret_cast<const uint16_t*>(src), ctx->precast, rows, a.K, st)
            : bf6 ? launch_promote_qc_bf6_planar(src, ctx->precast, rows, a.K, st)
                  : launch_promote_qc_fp8_planar(src, ctx->precast, rows, a.K, st);
        if (e != hipSuccess) return e;
        HerkPlainArgs b = a;
        b.A = ctx->precast;
        b.C = static_cast<uint8_t*>(a.C) + (int64_t)b0 * out_bytes;
        b.batch = nb;
        b.precast = bf6 ? 2 : 1;
        b.lda = (int64_t)2 * Kp;
        b.a_bstride = (int64_t)a.N * 2 * Kp;
        if ((e = launch_herk_int4_plain(b, cfg, fp16, remap, st)) != hipSuccess) return e;
    }
    return hipSuccess;
}

// Samples per padded row for tile configuration cfg: K rounded up to 16, or to the tile's
// stage when that adds at most 1/16 more samples (the copy then runs the loop without a
// partial last stage, which costs more than those zero samples).
int padded_row_samples(int K, int cfg) {
    const int k16 = precast_row_samples(K);
    const int st = cfg == TILE_128x128_K128 ? 128 : 64;
    const int kst = (K + st - 1) / st * st;
    return kst - k16 <= k16 / 16 ? kst : k16;
}

// the longest padded row of any configuration (scratch sizing)
int padded_row_cap(int K) {
    return std::max(padded_row_samples(K, TILE_128x128_K64), padded_row_samples(K, TILE_128x128_K128));
}

// The fused path for QC rows that 16-byte loads cannot take (K not a multiple of 16, or
// A not 16-byte aligned): per chunk of batches, copy the rows zero-padded to
// padded_row_samples(K, cfg) into the context's scratch (zero samples add nothing), then
// run the fused tile on them.  With K segments (polarization sum) a chunk holds whole groups of nseg * seg_group
// input batches, the seg_group output batches they sum into.  launch(b) runs one chunk;
// out_elem: bytes per scalar of C.  The scratch must hold padded_scratch_bytes(a).
int padded_chunk_units(const HerkPlainArgs& a) {
    const int grp = std::max(a.seg_group, 1);
    return chunk_cap_bytes((size_t)a.N * padded_row_cap(a.K) * a.nseg * grp, a.batch / grp);
}

size_t padded_scratch_bytes(const HerkPlainArgs& a) {
    return (size_t)a.N * padded_row_cap(a.K) * a.nseg * std::max(a.seg_group, 1) * padded_chunk_units(a);
}

// QC rows the fused tile reads with 16-byte loads
bool rows_take_vector_loads(const void* A, int K) {
    return K % 16 == 0 && (reinterpret_cast<uintptr_t>(A) & 15u) == 0;
}

// Polarization-sum launches: pad rows 16-byte loads cannot take, and also two K segments of
// K % stage != 0 when padding to the stage is cheap -- the two-segment loop without bounds
// checks needs whole stages, and the bounds-checked one is slower than the copy (2048
// inputs x 16 channels x 2 x 2, K = 1008: 0.60 ms; K = 1000 padded to 1024: 0.50 ms).
bool pol_rows_padded(const void* A, int K, int nseg, int cfg) {
    if (K <= 0) return false;
    if (!rows_take_vector_loads(A, K)) return true;
    const int st = cfg == TILE_128x128_K128 ? 128 : 64;
    return nseg == 2 && K % st != 0 && padded_row_samples(K, cfg) % st == 0;
}

// the plain HERK's case: one input batch per output batch
int padded_chunk_cap(int N, int K, int batch) {
    return chunk_cap_bytes((size_t)N * padded_row_cap(K), batch);
}

template <class Launch>
hipError_t run_padded(ftdx_context* ctx, const HerkPlainArgs& a, int cfg, int out_elem, hipStream_t st,
                      Launch&& launch) {
    const int Kq = padded_row_samples(a.K, cfg);
    const int grp = std::max(a.seg_group, 1), in_per = a.nseg * grp;
    const int units = a.batch / grp;
    const int chunk = pick_chunk(cfg, a.N, units, padded_chunk_units(a), grp);
    for (int u0 = 0; u0 < units; u0 += chunk) {
        const int nu = std::min(chunk, units - u0);
        hipError_t e = launch_pad_qc_rows(a.A + (int64_t)u0 * in_per * a.a_bstride, ctx->precast,
                                          (int64_t)a.N * nu * in_per, a.K, Kq, st);
        if (e != hipSuccess) return e;
        HerkPlainArgs b = a;
        b.A = ctx->precast;
        b.C = static_cast<uint8_t*>(a.C) + (int64_t)u0 * grp * a.c_bstride * out_elem;
        b.batch = nu * grp;
        b.K = Kq;
        b.a_bstride = (int64_t)a.N * Kq;
        b.freq_offset = a.freq_offset + u0 * grp;
        if ((e = launch(b)) != hipSuccess) return e;
    }
    return hipSuccess;
}

// Times every (tile configuration, workgroup order) on this call's buffers and
// stores the fastest (cutlass_interface/strategy_cache.hpp:504-800: warm-up, then
// event-timed repeats, failures skipped).  With beta != 0 the timing runs use a
// scratch output so the caller's C is not accumulated into.
// flag = TILE_BF6: the E3M2 variants, each timed with its conversion pass into the
// context's pre-cast rows (already sized by the caller).
int autotune(ftdx_context* ctx, const HerkPlainArgs& a0, bool out_fp16, hipStream_t s,
             const TuneKey& key, TuneEntry& best, int flag) {
    HerkPlainArgs a = a0;
    void* scratch = nullptr;
    hipError_t e;
    if (a.beta != 0.0f) {
        size_t bytes = (size_t)a.c_bstride * a.batch * (out_fp16 ? 2 : 4);
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || bytes > free_b * 8 / 10) {
            if (ctx->strategy_verbosity >= 1)
                std::printf("ftdx tune: skipped (scratch %.1f GB exceeds 80%% of free memory)\n",
                            bytes / 1e9);
            return FTDX_ERR_NO_MEMORY;
        }
        if ((e = hipMalloc(&scratch, bytes)) != hipSuccess) return hip_fail(e, "autotune: scratch");
        a.C = scratch;
    }
    hipEvent_t t0, t1;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    auto launch = [&](int cfg, int remap) -> hipError_t {
        if (!(cfg & TILE_BF6)) return launch_herk_int4_plain(a, cfg, out_fp16, remap, s);
        return run_precast(ctx, a, FTDX_INPUT_INT4, true, tile_base(cfg), out_fp16, remap, s);
    };
    // mean ms of `reps` launches, or a negative value if the launch failed
    auto time_cand = [&](int cfg, int remap, int reps) -> float {
        (void)hipEventRecord(t0, s);
        for (int r = 0; r < reps; ++r)
            if (launch(cfg, remap) != hipSuccess) {
                (void)hipGetLastError();
                return -1.0f;
            }
        (void)hipEventRecord(t1, s);
        if (hipEventSynchronize(t1) != hipSuccess) {
            (void)hipGetLastError();
            return -1.0f;
        }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, t0, t1);
        return ms / reps;
    };
    // first pass: every (configuration, workgroup order), one warm-up launch each
    struct Cand {
        int cfg, remap;
        float ms;
    };
    std::vector<Cand> cands;
    for (int base = 0; base < NUM_TILE_CONFIGS; ++base)
        for (int remap = 0; remap <= 1; ++remap) {
            const int cfg = base | flag;
            if (time_cand(cfg, remap, 1) < 0.0f) continue;
            const float ms = time_cand(cfg, remap, 3);
            if (ms < 0.0f) continue;
            if (ctx->strategy_verbosity >= 2)
                std::printf("ftdx tune:   %-16s xcd_remap=%d  %.4f ms\n", tile_config_name(cfg), remap, ms);
            cands.push_back({cfg, remap, ms});
        }
    best.ms = 1e30f;
    best.cfg = kDefaultCfg | flag;
    best.xcd_remap = kDefaultRemap;
    if (!cands.empty()) {
        // second pass: the three fastest again in mirrored order (A B C C B A), so a
        // clock that drifts under sustained MFMA load weighs on every finalist alike
        std::sort(cands.begin(), cands.end(), [](const Cand& x, const Cand& y) { return x.ms < y.ms; });
        const size_t nf = cands.size() < 3 ? cands.size() : 3;
        std::vector<float> sum(nf, 0.0f);
        for (int pass = 0; pass < 2; ++pass)
            for (size_t q = 0; q < nf; ++q) {
                const size_t i = pass == 0 ? q : nf - 1 - q;
                const float ms = time_cand(cands[i].cfg, cands[i].remap, 3);
                sum[i] += ms < 0.0f ? 1e30f : ms;
            }
        for (size_t i = 0; i < nf; ++i)
            if (sum[i] / 2 < best.ms) {
                best.ms = sum[i] / 2;
                best.cfg = cands[i].cfg;
                best.xcd_remap = cands[i].remap;
            }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (scratch) (void)hipFree(scratch);
    ctx->tune.put(key, best);
    if (ctx->strategy_verbosity >= 1)
        std::printf("ftdx tune: N=%d K=%d batch=%d -> %s xcd_remap=%d (%.4f ms)\n", key.N, key.K,
                    key.batch, tile_config_name(best.cfg), best.xcd_remap, best.ms);
    if (!ctx->tune_cache_path.empty() && !ctx->tune.save(ctx->tune_cache_path) &&
        ctx->strategy_verbosity >= 1)
        std::printf("ftdx tune: could not write %s\n", ctx->tune_cache_path.c_str());
    return FTDX_OK;
}

// Grow-only device scratch; the old buffer is released after the device is idle.
int ensure_scratch(uint8_t*& p, size_t& cap, size_t need, const char* what) {
    if (need <= cap && p) return FTDX_OK;
    hipError_t e;
    if (p) {
        if ((e = hipDeviceSynchronize()) != hipSuccess) return hip_fail(e, what);
        (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    if ((e = hipMalloc(&p, need > 0 ? need : 1)) != hipSuccess) return hip_fail(e, what);
    cap = need;
    return FTDX_OK;
}

// Grows the pre-cast operand buffer; captured graphs of the pre-cast path point at
// the old buffer, so they are dropped when it moves.
int grow_precast(ftdx_context* ctx, size_t need, const char* what) {
    const uint8_t* before = ctx->precast;
    int rc = ensure_scratch(ctx->precast, ctx->precast_cap, need, what);
    if (rc == FTDX_OK && before && before != ctx->precast) {
        for (auto& gph : ctx->graphs)
            if (gph.used && gph.mode != 0) {
                (void)hipGraphExecDestroy(gph.exec);
                gph = ftdx_context::GraphEntry();
            }
    }
    return rc;
}

// A operand formats of the complex GEMM
enum GemmInput { GEMM_A_FP16_PLANAR, GEMM_A_INT4, GEMM_A_FP8_INTERLEAVED };

// Converts A to pre-cast rows (context scratch) and runs the GEMM against the rows
// of conj(B) at `brows`.  fmt: OUT_GEMM_PLANAR (c0 = Re, c1 = Im) or OUT_GEMM_POWER.
int gemm_run(ftdx_context* ctx, const uint8_t* brows, GemmInput kind, const void* a0, const void* a1,
             void* c0, void* c1, int M, int N, int K, int batch, int fmt, bool out_fp16, float alpha,
             float beta, void* stream, const char* what) {
    const std::string w(what);
    if (M < 0) return fail(FTDX_ERR_INVALID_ARG, w + ": M must be >= 0");
    if (M == 0 || N == 0 || batch == 0) return FTDX_OK;
    if (!c0 || (fmt == OUT_GEMM_PLANAR && !c1))
        return fail(FTDX_ERR_INVALID_ARG, w + ": NULL output pointer");
    if (K > 0 && (!a0 || (kind == GEMM_A_FP16_PLANAR && !a1)))
        return fail(FTDX_ERR_INVALID_ARG, w + ": NULL A pointer");
    if ((int64_t)M * N > ((int64_t)1 << 40) || (int64_t)M * K > ((int64_t)1 << 40))
        return fail(FTDX_ERR_INVALID_ARG, w + ": sizes too large");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int Kp = precast_row_samples(K);
    const int64_t rows = (int64_t)M * batch;
    int rc = grow_precast(ctx, (size_t)rows * Kp * 2 + 16, "gemm: A rows");
    if (rc != FTDX_OK) return rc;
    hipError_t e = hipSuccess;
    if (K > 0) {
        if (kind == GEMM_A_FP16_PLANAR)
            e = launch_promote_fp16_planar_rows(static_cast<const uint16_t*>(a0),
                                                static_cast<const uint16_t*>(a1), ctx->precast, rows, K,
                                                false, s);
        else if (kind == GEMM_A_INT4)
            e = launch_promote_qc_fp8_planar(static_cast<const uint8_t*>(a0), ctx->precast, rows, K, s);
        else
            e = launch_promote_fp8_interleaved_rows(static_cast<const uint8_t*>(a0), ctx->precast, rows,
                                                    K, s);
        if (e != hipSuccess) return hip_fail(e, what);
    }
    GemmArgs g;
    g.A = ctx->precast;
    g.B = brows;
    g.C = c0;
    g.C2 = fmt == OUT_GEMM_PLANAR ? c1 : nullptr;
    g.M = M;
    g.N = N;
    g.K = K;
    g.batch = batch;
    g.alpha = alpha;
    g.beta = beta;
    g.lda = g.ldb = (int64_t)Kp * 2;
    g.a_bstride = (int64_t)M * Kp * 2;
    g.b_bstride = (int64_t)N * Kp * 2;
    g.ldc = N;
    g.c_bstride = (int64_t)M * N;
    // 256x128 tiles, one workgroup per CU, whose direct loads run two stages ahead in three
    // LDS buffers: fastest at the beamformer shape (M = 256, N = 4000, K = 1664: 0.90 ms
    // against 0.91-0.96 ms for the 128x128 tiles, two workgroups per CU)
    const int cfg = ctx->forced_cfg >= 0 ? tile_base(ctx->forced_cfg) : TILE_256x128_K64;
    const int remap = ctx->forced_remap >= 0 ? ctx->forced_remap : kDefaultRemap;
    e = launch_gemm_fp8(g, cfg, fmt, out_fp16, remap, s);
    return e == hipSuccess ? FTDX_OK : hip_fail(e, what);
}

// Argument checks shared by the two gain entry points (sizes, pointers, alignment).
int gain_check_args(const char* what, const void* V, const void* M, const void* g, int N, int batch,
                    const void* loss, const void* grad_g, const void* Gamma) {
    const std::string w(what);
    if (N < 0 || batch < 0) return fail(FTDX_ERR_INVALID_ARG, w + ": negative size");
    if (N == 0 || batch == 0) return FTDX_OK;
    if (!V || !M || !g || !loss || !grad_g) return fail(FTDX_ERR_INVALID_ARG, w + ": NULL pointer");
    const uintptr_t al = reinterpret_cast<uintptr_t>(V) | reinterpret_cast<uintptr_t>(M) |
                         reinterpret_cast<uintptr_t>(g) | reinterpret_cast<uintptr_t>(loss) |
                         reinterpret_cast<uintptr_t>(grad_g) | reinterpret_cast<uintptr_t>(Gamma);
    if (al & 7u) return fail(FTDX_ERR_INVALID_ARG, w + ": buffers must be 8-byte aligned");
    return FTDX_OK;
}

// Launches the gain kernels with their partials in `scr` (gain_scratch_bytes(N, batch)).
int gain_launch(const char* what, uint8_t* scr, const void* V, const void* M, const void* g, int N, int batch,
                void* loss, void* grad_g, void* Gamma, void* stream) {
    const size_t part = (size_t)batch * gain_blocks(N) * N * sizeof(float2);
    hipError_t e = launch_gain_loss_grad(
        static_cast<const float2*>(V), static_cast<const float2*>(M), static_cast<const float2*>(g), N, batch,
        static_cast<float2*>(Gamma), reinterpret_cast<float2*>(scr), reinterpret_cast<float2*>(scr + part),
        reinterpret_cast<double*>(scr + 2 * part), static_cast<double*>(loss), static_cast<float2*>(grad_g),
        static_cast<hipStream_t>(stream));
    return e == hipSuccess ? FTDX_OK : hip_fail(e, (std::string(what) + ": launch").c_str());
}

int mx_format(int compute) {
    return compute == FTDX_COMPUTE_FP4 ? MX_FP4_E2M1 : MX_FP6_E3M2;
}

}  // namespace

}  // namespace ftdx

using ftdx::fail;
using ftdx::hip_fail;

extern "C" {

const char* ftdx_last_error(void) { return ftdx::g_last_error.c_str(); }

int ftdx_abi_version(void) { return FTDX_ABI_VERSION; }

int ftdx_create(ftdx_context** out) {
    if (!out) return fail(FTDX_ERR_INVALID_ARG, "ftdx_create: out is NULL");
    ftdx_context* c = new (std::nothrow) ftdx_context();
    if (!c) return fail(FTDX_ERR_NO_MEMORY, "ftdx_create: out of host memory");
    // Without a visible device the context still holds settings and the tuning
    // cache; calls that launch work then fail with a device error.
    if (hipGetDevice(&c->device) != hipSuccess) {
        (void)hipGetLastError();
        c->device = -1;
    }
    *out = c;
    return FTDX_OK;
}

int ftdx_destroy(ftdx_context* ctx) {
    if (!ctx) return FTDX_OK;
    ftdx_end_herk(ctx);
    if (ctx->prep_b || ctx->gemm_b) {
        (void)hipDeviceSynchronize();
        if (ctx->prep_b) (void)hipFree(ctx->prep_b);
        if (ctx->gemm_b) (void)hipFree(ctx->gemm_b);
    }
    delete ctx;
    return FTDX_OK;
}

int ftdx_set_herk_mode(ftdx_context* ctx, int mode) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_herk_mode: NULL context");
    if (mode < FTDX_HERK_AUTO || mode > FTDX_HERK_FORCE_BASELINE)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_herk_mode: unknown mode");
    ctx->herk_mode = mode;
    return FTDX_OK;
}

int ftdx_set_direct_tile_config(ftdx_context* ctx, int config) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_direct_tile_config: NULL context");
    if (config < FTDX_DIRECT_TILE_AUTO || config > FTDX_DIRECT_TILE_K128_B2)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_direct_tile_config: unknown config");
    ctx->direct_tile = config;
    return FTDX_OK;
}

int ftdx_set_launch_config(ftdx_context* ctx, int tile_config, int xcd_remap) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_launch_config: NULL context");
    if ((tile_config != -1 && !ftdx::valid_tile_config(tile_config)) || xcd_remap < -1 || xcd_remap > 1)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_launch_config: out of range");
    ctx->forced_cfg = tile_config;
    ctx->forced_remap = xcd_remap;
    return FTDX_OK;
}

int ftdx_set_mx_rounding(ftdx_context* ctx, int rounding) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_mx_rounding: NULL context");
    if (rounding != FTDX_MX_ROUND_REFERENCE && rounding != FTDX_MX_ROUND_NEAREST_EVEN)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_mx_rounding: unknown rounding");
    ctx->mx_rounding = rounding;
    return FTDX_OK;
}

int ftdx_get_mx_rounding(const ftdx_context* ctx) { return ctx ? ctx->mx_rounding : -1; }

int ftdx_set_direct_herk_graph(ftdx_context* ctx, int enable) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_direct_herk_graph: NULL context");
    ctx->graph = enable ? 1 : 0;
    return FTDX_OK;
}

int ftdx_herk_graph_stats(const ftdx_context* ctx, int* cached, long long* captures) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_graph_stats: NULL context");
    int n = 0;
    for (const auto& gph : ctx->graphs) n += gph.used ? 1 : 0;
    if (cached) *cached = n;
    if (captures) *captures = ctx->graph_captures;
    return FTDX_OK;
}

int ftdx_set_tune_cache_path(ftdx_context* ctx, const char* path) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_tune_cache_path: NULL context");
    ctx->tune_cache_path = path ? path : "";
    ctx->tune_loaded = false;
    return FTDX_OK;
}

int ftdx_set_strategy_tune_verbosity(ftdx_context* ctx, int level) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_strategy_tune_verbosity: NULL context");
    ctx->strategy_verbosity = level;
    return FTDX_OK;
}

int ftdx_set_kernel_tune_verbosity(ftdx_context* ctx, int level) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_kernel_tune_verbosity: NULL context");
    ctx->kernel_verbosity = level;
    return FTDX_OK;
}

int ftdx_set_kernel_tune_cache_path(ftdx_context* ctx, const char* path) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_kernel_tune_cache_path: NULL context");
    ctx->kernel_cache_path = path ? path : "";
    return FTDX_OK;
}

int ftdx_tune_load(ftdx_context* ctx, const char* path, int* loaded, int* skipped) {
    if (!ctx || !path) return fail(FTDX_ERR_INVALID_ARG, "ftdx_tune_load: NULL argument");
    int sk = 0;
    int n = ctx->tune.load(path, &sk);
    ctx->tune_loaded = true;
    if (loaded) *loaded = n < 0 ? 0 : n;
    if (skipped) *skipped = sk;
    if (n < 0) return fail(FTDX_ERR_IO, std::string("ftdx_tune_load: cannot read ") + path);
    return FTDX_OK;
}

int ftdx_tune_save(ftdx_context* ctx, const char* path) {
    if (!ctx || !path) return fail(FTDX_ERR_INVALID_ARG, "ftdx_tune_save: NULL argument");
    if (!ctx->tune.save(path)) return fail(FTDX_ERR_IO, std::string("ftdx_tune_save: cannot write ") + path);
    return FTDX_OK;
}

int ftdx_herk_query_config(ftdx_context* ctx, int N, int K, int batch_count, int output,
                           int* tile_config, int* xcd_remap) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_query_config: NULL context");
    ftdx::ensure_tune_loaded(ctx);
    ftdx::TuneKey key{N, K, batch_count, output == FTDX_OUTPUT_FP16 ? 1 : 0, K % 16 == 0 ? 1 : 0};
    int cfg, remap;
    bool cached;
    ftdx::pick_config(ctx, key, cfg, remap, cached);
    if (tile_config) *tile_config = cfg;
    if (xcd_remap) *xcd_remap = remap;
    return FTDX_OK;
}

int ftdx_init_herk(ftdx_context* ctx, int N, int K, int batch_count, void* stream) {
    (void)stream;
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_init_herk: NULL context");
    if (N < 0 || K < 0 || batch_count < 0)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_init_herk: negative size");
    // The fused path on 16-byte-aligned rows of K % 16 == 0 needs no scratch.  The
    // unfused (ForceBaseline) and FP16-input paths stage pre-cast FP8 rows, and the fused
    // path on rows of K % 16 != 0 zero-padded copies, one chunk of channels at a time:
    // allocate that now so that later calls (and graph capture) allocate nothing
    // (cutlass_gemm_api.h:601-617).  Rows at an address that is not 16-byte aligned are
    // padded too; their scratch is allocated by the first such call unless graph mode
    // or ForceBaseline already sized it.
    ctx->init_N = N;
    ctx->init_K = K;
    ctx->init_batch = batch_count;
    ftdx::ensure_tune_loaded(ctx);
    const int Kp = ftdx::precast_row_samples(K);
    size_t need = 0;
    if (ctx->herk_mode == FTDX_HERK_FORCE_BASELINE || ctx->graph)
        need = (size_t)N * ftdx::precast_chunk_cap(N, K, batch_count) * 2 * Kp;
    if (K % 16 != 0 || ctx->graph)
        need = std::max(need, (size_t)N * ftdx::padded_chunk_cap(N, K, batch_count) * ftdx::padded_row_cap(K));
    if (need > 0) {
        int rc = ftdx::grow_precast(ctx, need, "ftdx_init_herk: scratch");
        if (rc != FTDX_OK) return rc;
    }
    return FTDX_OK;
}

int ftdx_end_herk(ftdx_context* ctx) {
    if (!ctx) return FTDX_OK;
    ctx->init_N = ctx->init_K = ctx->init_batch = 0;
    bool any_graph = false;
    for (auto& gph : ctx->graphs) any_graph = any_graph || gph.used;
    if (ctx->mx_codes || ctx->cal_part || ctx->cal_herm || ctx->precast ||
        any_graph) {
        (void)hipDeviceSynchronize();
        if (ctx->mx_codes) (void)hipFree(ctx->mx_codes);
        if (ctx->cal_part) (void)hipFree(ctx->cal_part);
        if (ctx->cal_herm) (void)hipFree(ctx->cal_herm);
        if (ctx->precast) (void)hipFree(ctx->precast);
        for (auto& gph : ctx->graphs) {
            if (gph.used && gph.exec) (void)hipGraphExecDestroy(gph.exec);
            gph = ftdx_context::GraphEntry();
        }
    }
    if (ctx->cap_stream) {
        (void)hipStreamDestroy(ctx->cap_stream);
        ctx->cap_stream = nullptr;
    }
    ctx->mx_codes = nullptr;
    ctx->mx_codes_cap = 0;
    ctx->cal_part = ctx->cal_herm = nullptr;
    ctx->cal_part_cap = ctx->cal_herm_cap = 0;
    ctx->precast = nullptr;
    ctx->precast_cap = 0;
    return FTDX_OK;
}

int ftdx_gain_loss_grad(ftdx_context* ctx, const void* V, const void* M, const void* g, int N,
                        int batch_count, void* loss, void* grad_g, void* Gamma, void* stream) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_gain_loss_grad: NULL context");
    int rc = ftdx::gain_check_args("ftdx_gain_loss_grad", V, M, g, N, batch_count, loss, grad_g, Gamma);
    if (rc != FTDX_OK || N == 0 || batch_count == 0) return rc;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "ftdx_gain_loss_grad: hipGetDevice");
    if (dev != ctx->device)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_gain_loss_grad: context belongs to another device (or none)");
    rc = ftdx::ensure_scratch(ctx->cal_part, ctx->cal_part_cap, ftdx::gain_scratch_bytes(N, batch_count),
                              "ftdx_gain_loss_grad: scratch");
    if (rc != FTDX_OK) return rc;
    return ftdx::gain_launch("ftdx_gain_loss_grad", ctx->cal_part, V, M, g, N, batch_count, loss, grad_g,
                             Gamma, stream);
}

static_assert(ftdx::kGainStaticScratchBytes == FTDX_GAIN_NOALLOC_SCRATCH_BYTES,
              "the fixed gain scratch must match its public size");

int ftdx_gain_loss_grad_noalloc(const void* V, const void* M, const void* g, int N, int batch_count,
                                void* loss, void* grad_g, void* Gamma, void* stream) {
    int rc = ftdx::gain_check_args("ftdx_gain_loss_grad_noalloc", V, M, g, N, batch_count, loss, grad_g,
                                   Gamma);
    if (rc != FTDX_OK || N == 0 || batch_count == 0) return rc;
    const size_t need = ftdx::gain_scratch_bytes(N, batch_count);
    if (need > ftdx::kGainStaticScratchBytes)
        return fail(FTDX_ERR_INVALID_ARG,
                    "ftdx_gain_loss_grad_noalloc: N=" + std::to_string(N) + ", batch=" +
                        std::to_string(batch_count) + " needs " + std::to_string(need) +
                        " bytes of scratch, above the fixed " + std::to_string(ftdx::kGainStaticScratchBytes) +
                        " (FTDX_GAIN_NOALLOC_SCRATCH_BYTES); use ftdx_gain_loss_grad");
    void* scr = nullptr;
    hipError_t e = ftdx::gain_static_scratch(&scr);
    if (e != hipSuccess) return hip_fail(e, "ftdx_gain_loss_grad_noalloc: scratch address");
    return ftdx::gain_launch("ftdx_gain_loss_grad_noalloc", static_cast<uint8_t*>(scr), V, M, g, N,
                             batch_count, loss, grad_g, Gamma, stream);
}

int ftdx_herk_adjoint(ftdx_context* ctx, const void* Gamma, const void* X, void* dX, int N, int K,
                      int batch_count, int input, void* stream) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_adjoint: NULL context");
    if (input != FTDX_INPUT_INT4)
        return fail(FTDX_ERR_UNSUPPORTED, "ftdx_herk_adjoint: only INT4 (QC) voltages are supported");
    if (N < 0 || K < 0 || batch_count < 0)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_adjoint: negative size");
    if (N == 0 || K == 0 || batch_count == 0) return FTDX_OK;
    if (!Gamma || !X || !dX) return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_adjoint: NULL pointer");
    if ((reinterpret_cast<uintptr_t>(Gamma) | reinterpret_cast<uintptr_t>(dX)) & 7u)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_adjoint: Gamma and dX must be 8-byte aligned");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "ftdx_herk_adjoint: hipGetDevice");
    if (dev != ctx->device)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_adjoint: context belongs to another device (or none)");
    const size_t hb = ftdx::adjoint_scratch_bytes(N, K, batch_count);
    int rc = ftdx::ensure_scratch(ctx->cal_herm, ctx->cal_herm_cap, hb, "ftdx_herk_adjoint: scratch");
    if (rc != FTDX_OK) return rc;
    e = ftdx::launch_herk_adjoint(static_cast<const float2*>(Gamma), static_cast<const uint8_t*>(X),
                                  ctx->cal_herm, static_cast<float2*>(dX),
                                  N, K, batch_count, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "ftdx_herk_adjoint: launch");
    return FTDX_OK;
}

int ftdx_mx_layout(int compute, int K, int* blocks, int* code_bytes) {
    if (compute != FTDX_COMPUTE_FP6 && compute != FTDX_COMPUTE_FP4)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_mx_layout: compute must be FP6 or FP4");
    if (K < 0) return fail(FTDX_ERR_INVALID_ARG, "ftdx_mx_layout: negative K");
    if (blocks) *blocks = ftdx::mx_padded_blocks(K);
    if (code_bytes) *code_bytes = ftdx::mx_code_bytes(ftdx::mx_format(compute));
    return FTDX_OK;
}

int ftdx_mx_quantize(const void* X, int input, int compute, void* codes, void* scales,
                     int64_t rows, int K, void* stream) {
    return ftdx_mx_quantize_rounding(X, input, compute, FTDX_MX_ROUND_REFERENCE, codes, scales, rows, K, stream);
}

int ftdx_mx_quantize_rounding(const void* X, int input, int compute, int rounding, void* codes, void* scales,
                              int64_t rows, int K, void* stream) {
    if (compute != FTDX_COMPUTE_FP6 && compute != FTDX_COMPUTE_FP4)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_mx_quantize: compute must be FP6 or FP4");
    if (rounding != FTDX_MX_ROUND_REFERENCE && rounding != FTDX_MX_ROUND_NEAREST_EVEN)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_mx_quantize: unknown rounding");
    if (input != FTDX_INPUT_INT4 && input != FTDX_INPUT_FP16)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_mx_quantize: unknown input precision");
    if (rows < 0 || K < 0) return fail(FTDX_ERR_INVALID_ARG, "ftdx_mx_quantize: negative size");
    if (rows == 0 || K == 0) return FTDX_OK;
    if (!X || !codes || !scales) return fail(FTDX_ERR_INVALID_ARG, "ftdx_mx_quantize: NULL pointer");
    if ((reinterpret_cast<uintptr_t>(codes) & 7u) || (input == FTDX_INPUT_FP16 && (reinterpret_cast<uintptr_t>(X) & 3u)))
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_mx_quantize: misaligned buffer");
    hipError_t e = ftdx::launch_mx_quantize(
        X, input == FTDX_INPUT_INT4 ? ftdx::MX_IN_INT4 : ftdx::MX_IN_FP16, ftdx::mx_format(compute), rounding,
        static_cast<uint8_t*>(codes), static_cast<uint8_t*>(scales), rows, K,
        static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "ftdx_mx_quantize: launch");
    return FTDX_OK;
}

int ftdx_herk_mx(ftdx_context* ctx, const void* A, void* C, int N, int K, int batch_count,
                 int input, int compute, int output, float alpha, float beta, void* stream) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_mx: NULL context");
    if (N < 0 || K < 0 || batch_count < 0)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_mx: negative size");
    if (compute != FTDX_COMPUTE_FP6 && compute != FTDX_COMPUTE_FP4)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_mx: compute must be FP6 or FP4");
    if (input != FTDX_INPUT_INT4 && input != FTDX_INPUT_FP16)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_mx: unknown input precision");
    if (output != FTDX_OUTPUT_FP32 && output != FTDX_OUTPUT_FP16)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_mx: unknown output precision");
    if (N == 0 || batch_count == 0) return FTDX_OK;
    if ((!A && K > 0) || !C) return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_mx: NULL data pointer");
    if (reinterpret_cast<uintptr_t>(C) % (output == FTDX_OUTPUT_FP32 ? 8u : 4u) != 0 ||
        (input == FTDX_INPUT_FP16 && (reinterpret_cast<uintptr_t>(A) & 3u)))
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_mx: misaligned buffer");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "ftdx_herk_mx: hipGetDevice");
    if (dev != ctx->device)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_mx: context belongs to another device (or none)");

    const int cfg = ctx->forced_cfg >= 0 ? ftdx::tile_base(ctx->forced_cfg) : ftdx::kMxDefaultCfg;
    const int remap = ctx->forced_remap >= 0 ? ctx->forced_remap : ftdx::kDefaultRemap;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int fmt = ftdx::mx_format(compute);
    // E2M1 on a configuration whose tiles are all interior: dense rows (32 bytes per
    // block, scales apart), twice the samples per stage of the inline-scale rows.
    const int nblkd = ftdx::mx_dense_blocks(K);
    if (fmt == ftdx::MX_FP4_E2M1 && nblkd > 0 && ftdx::mx4d_dense_ok(cfg, N, (int64_t)nblkd * 32)) {
        const int64_t rows = (int64_t)N * batch_count;
        const size_t code_bytes = (size_t)rows * nblkd * 32, scale_bytes = (size_t)rows * (nblkd / 4) * 8;
        int rc = ftdx::ensure_scratch(ctx->mx_codes, ctx->mx_codes_cap, code_bytes + scale_bytes + 16,
                                      "ftdx_herk_mx: scratch");
        if (rc != FTDX_OK) return rc;
        uint8_t* scales = ctx->mx_codes + code_bytes;  // code_bytes is a multiple of 128
        e = ftdx::launch_mx_quantize_dense(A, input == FTDX_INPUT_INT4 ? ftdx::MX_IN_INT4 : ftdx::MX_IN_FP16,
                                           ctx->mx_rounding, ctx->mx_codes, scales, rows, N, K, s);
        if (e != hipSuccess) return hip_fail(e, "ftdx_herk_mx: quantize launch");
        ftdx::HerkPlainArgs a{};
        a.A = ctx->mx_codes;
        a.C = C;
        a.N = N;
        a.K = nblkd * 16;  // 2-byte units: 64 per 4-block stage
        a.batch = batch_count;
        a.alpha = alpha;
        a.beta = beta;
        a.a_bstride = (int64_t)N * nblkd * 32;
        a.c_bstride = (int64_t)N * (N + 1);
        a.precast = 5;
        a.lda = (int64_t)nblkd * 32;
        a.S = scales;
        a.s_stage = (int64_t)N * 8;
        a.s_bstride = (int64_t)(nblkd / 4) * N * 8;
        e = ftdx::launch_herk_int4_plain(a, cfg, output == FTDX_OUTPUT_FP16, remap, s);
        if (e != hipSuccess) return hip_fail(e, "ftdx_herk_mx: kernel launch");
        return FTDX_OK;
    }

    // Quantize into MX operand rows (codes and scale per 32-sample part, the tile's
    // 64-bytes-per-32-samples geometry), then the HERK tile kernel with scaled MFMAs.
    const int nblkp = ftdx::mx_padded_blocks(K);
    const int64_t rows = (int64_t)N * batch_count;
    const size_t row_bytes = (size_t)nblkp * 64;
    int rc = ftdx::ensure_scratch(ctx->mx_codes, ctx->mx_codes_cap, (size_t)rows * row_bytes + 16,
                                  "ftdx_herk_mx: scratch");
    if (rc != FTDX_OK) return rc;
    if (nblkp > 0) {
        e = ftdx::launch_mx_quantize_rows(A, nullptr, input == FTDX_INPUT_INT4 ? ftdx::MX_IN_INT4 : ftdx::MX_IN_FP16,
                                          fmt, ctx->mx_rounding, false, ctx->mx_codes, rows, K, s);
        if (e != hipSuccess) return hip_fail(e, "ftdx_herk_mx: quantize launch");
    }
    ftdx::HerkPlainArgs a{};
    a.A = ctx->mx_codes;
    a.C = C;
    a.N = N;
    a.K = nblkp * 32;
    a.batch = batch_count;
    a.alpha = alpha;
    a.beta = beta;
    a.a_bstride = (int64_t)N * row_bytes;
    a.c_bstride = (int64_t)N * (N + 1);
    a.precast = fmt == ftdx::MX_FP6_E3M2 ? 3 : 4;
    a.lda = (int64_t)row_bytes;
    e = ftdx::launch_herk_int4_plain(a, cfg, output == FTDX_OUTPUT_FP16, remap, s);
    if (e != hipSuccess) return hip_fail(e, "ftdx_herk_mx: kernel launch");
    return FTDX_OK;
}

int ftdx_herk(ftdx_context* ctx, const void* A, void* C, int N, int K, int batch_count,
              int input, int compute, int output, int format, float alpha, float beta,
              void* stream, int tune) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk: NULL context");
    if (N < 0 || K < 0 || batch_count < 0)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk: negative size");
    if (format != FTDX_FORMAT_PACKED_TRIANGLE)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk: only PackedTriangle output is supported");
    if (output != FTDX_OUTPUT_FP32 && output != FTDX_OUTPUT_FP16)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk: unknown output precision");
    if (compute < FTDX_COMPUTE_FP8 || compute > FTDX_COMPUTE_FP4)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk: unknown compute precision");
    if (input != FTDX_INPUT_INT4 && input != FTDX_INPUT_FP16)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk: unknown input precision");
    // cutlass_interface/cutlass_gemm_api.cu:1721-1723
    if (output == FTDX_OUTPUT_FP32 && compute != FTDX_COMPUTE_FP8)
        return fail(FTDX_ERR_UNSUPPORTED,
                    "FP32 output HERK currently only supports FP8 compute precision");
    if (N == 0 || batch_count == 0) return FTDX_OK;
    if ((!A && K > 0) || !C) return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk: NULL data pointer");
    if (reinterpret_cast<uintptr_t>(C) % (output == FTDX_OUTPUT_FP32 ? 8u : 4u) != 0)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk: C must be aligned to one complex element");
    if ((int64_t)N * (N + 1) / 2 > (int64_t)1 << 40)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk: N too large");
    // FP6 compute on INT4 input is lossless (cutlass_gemm_api.h:45: E3M2 holds every QC
    // value): it runs the E3M2 operand tiles.  Other FP6 / FP4 cases take the block-scaled
    // MX path (FP16 output, as the reference).
    const bool e3m2 = compute == FTDX_COMPUTE_FP6 && input == FTDX_INPUT_INT4;
    if (compute != FTDX_COMPUTE_FP8 && !e3m2)
        return ftdx_herk_mx(ctx, A, C, N, K, batch_count, input, compute, output, alpha, beta,
                            stream);
    if (input == FTDX_INPUT_FP16 && (reinterpret_cast<uintptr_t>(A) & 3u))
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk: FP16 input must be 4-byte aligned");

    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "ftdx_herk: hipGetDevice");
    if (dev != ctx->device)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk: context belongs to another device (or none)");

    ftdx::HerkPlainArgs a{};
    a.A = static_cast<const uint8_t*>(A);
    a.C = C;
    a.N = N;
    a.K = K;
    a.batch = batch_count;
    a.alpha = alpha;
    a.beta = beta;
    a.a_bstride = (int64_t)N * K;
    a.c_bstride = (int64_t)N * (N + 1);
    const bool fp16 = output == FTDX_OUTPUT_FP16;
    hipStream_t s = static_cast<hipStream_t>(stream);

    ftdx::ensure_tune_loaded(ctx);
    // Pre-cast FP8 (unfused) path: FP16 input always; INT4 input under ForceBaseline.
    // E3M2 compute always converts first (its own tune-cache entries: vec = 3).
    const bool precast = !e3m2 && ((input == FTDX_INPUT_FP16) || (ctx->herk_mode == FTDX_HERK_FORCE_BASELINE));
    // Fused path on QC rows that 16-byte loads cannot take (K % 16 != 0 or A unaligned):
    // the rows are copied zero-padded first (tune-cache entries: vec = 4).
    const bool padded = !e3m2 && !precast && !ftdx::rows_take_vector_loads(A, K);
    const int vec = e3m2 ? 3 : precast ? 2 : padded ? 4 : 1;
    ftdx::TuneKey key{N, K, batch_count, fp16 ? 1 : 0, vec};
    int cfg, remap;
    bool cached;
    ftdx::pick_config(ctx, key, cfg, remap, cached);
    const bool explicit_cfg = ctx->forced_cfg >= 0 || ctx->direct_tile != FTDX_DIRECT_TILE_AUTO;
    const int Kp = ftdx::precast_row_samples(K);
    const size_t rows_need = (size_t)N * ftdx::precast_chunk_cap(N, K, batch_count) * 2 * Kp;
    if (precast) {
        int rc = ftdx::grow_precast(ctx, rows_need, "ftdx_herk: scratch");
        if (rc != FTDX_OK) return rc;
    }
    if (e3m2) {
        int rc = ftdx::grow_precast(ctx, rows_need, "ftdx_herk: scratch");
        if (rc != FTDX_OK) return rc;
    }
    const size_t pad_need = (size_t)N * ftdx::padded_chunk_cap(N, K, batch_count) * ftdx::padded_row_cap(K);
    if (padded) {
        int rc = ftdx::grow_precast(ctx, pad_need, "ftdx_herk: scratch");
        if (rc != FTDX_OK) return rc;
    }
    if (tune && !cached && !explicit_cfg && !precast) {
        ftdx::TuneEntry best;
        ftdx::HerkPlainArgs t = a;
        if (padded) {
            // time the tiles on the first chunk of padded rows (the 64-sample stages' padding)
            const int Kt = ftdx::padded_row_samples(K, ftdx::TILE_128x128_K64);
            t.batch = ftdx::padded_chunk_cap(N, K, batch_count);
            if ((e = ftdx::launch_pad_qc_rows(a.A, ctx->precast, (int64_t)N * t.batch, K, Kt, s)) != hipSuccess)
                return hip_fail(e, "ftdx_herk: pad rows");
            t.A = ctx->precast;
            t.K = Kt;
            t.a_bstride = (int64_t)N * Kt;
        }
        if (ftdx::autotune(ctx, t, fp16, s, key, best, e3m2 ? ftdx::TILE_BF6 : 0) == FTDX_OK) {
            cfg = best.cfg;
            if (ctx->forced_remap < 0) remap = best.xcd_remap;
        }
    }
    // E3M2 operand tiles: FP6 compute on INT4 input, or an INT4 FP8 call whose tile
    // configuration carries the FTDX_TILE_BF6 flag (explicit choice or cache entry)
    const bool bf6 = e3m2 || (!precast && (cfg & ftdx::TILE_BF6));
    cfg = ftdx::tile_base(cfg);
    if (bf6) {
        int rc = ftdx::grow_precast(ctx, rows_need, "ftdx_herk: scratch");
        if (rc != FTDX_OK) return rc;
    }
    auto run = [&](hipStream_t st) -> hipError_t {
        if (precast || bf6) return ftdx::run_precast(ctx, a, input, bf6, cfg, fp16, remap, st);
        if (padded)
            return ftdx::run_padded(ctx, a, cfg, fp16 ? 2 : 4, st, [&](const ftdx::HerkPlainArgs& b) {
                return ftdx::launch_herk_int4_plain(b, cfg, fp16, remap, st);
            });
        return ftdx::launch_herk_int4_plain(a, cfg, fp16, remap, st);
    };
    // Graph replay of identical calls (cutlass_gemm_api.cu:1738-1807): capture the
    // launch sequence once per (shape, scalars, pointers), then replay it.
    if (ctx->graph && ctx->herk_mode != FTDX_HERK_AUTO) {
        const int mode = bf6 ? 2 : precast ? 1 : padded ? 3 : 0;
        ftdx_context::GraphEntry* hit = nullptr;
        for (auto& gph : ctx->graphs)
            if (gph.used && gph.N == N && gph.K == K && gph.batch == batch_count &&
                gph.input == input && gph.output == output && gph.mode == mode && gph.cfg == cfg &&
                gph.remap == remap && gph.alpha == alpha && gph.beta == beta && gph.A == A && gph.C == C)
                hit = &gph;
        if (!hit) {
            if (!ctx->cap_stream &&
                (e = hipStreamCreateWithFlags(&ctx->cap_stream, hipStreamNonBlocking)) != hipSuccess)
                return hip_fail(e, "ftdx_herk: capture stream");
            if ((e = hipStreamBeginCapture(ctx->cap_stream, hipStreamCaptureModeRelaxed)) != hipSuccess)
                return hip_fail(e, "ftdx_herk: begin capture");
            hipError_t le = run(ctx->cap_stream);
            hipGraph_t graph = nullptr;
            e = hipStreamEndCapture(ctx->cap_stream, &graph);
            if (le != hipSuccess || e != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                return hip_fail(le != hipSuccess ? le : e, "ftdx_herk: capture");
            }
            hipGraphExec_t exec = nullptr;
            e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) return hip_fail(e, "ftdx_herk: graph instantiate");
            ftdx_context::GraphEntry* slot = &ctx->graphs[0];
            for (auto& gph : ctx->graphs) {
                if (!gph.used) { slot = &gph; break; }
                if (gph.stamp < slot->stamp) slot = &gph;
            }
            if (slot->used && slot->exec) (void)hipGraphExecDestroy(slot->exec);
            slot->used = true;
            slot->N = N; slot->K = K; slot->batch = batch_count;
            slot->input = input; slot->output = output; slot->mode = mode;
            slot->cfg = cfg; slot->remap = remap;
            ++ctx->graph_captures;
            slot->alpha = alpha; slot->beta = beta; slot->A = A; slot->C = C;
            slot->exec = exec;
            hit = slot;
        }
        hit->stamp = ++ctx->graph_clock;
        if ((e = hipGraphLaunch(hit->exec, s)) != hipSuccess) return hip_fail(e, "ftdx_herk: graph launch");
        return FTDX_OK;
    }
    e = run(s);
    if (e != hipSuccess) return hip_fail(e, "ftdx_herk: kernel launch");
    return FTDX_OK;
}

int ftdx_promote_qc_to_fp8(const void* qc, void* fp8_pairs, int64_t n, void* stream) {
    if (n < 0) return fail(FTDX_ERR_INVALID_ARG, "ftdx_promote_qc_to_fp8: negative size");
    if (n == 0) return FTDX_OK;
    if (!qc || !fp8_pairs) return fail(FTDX_ERR_INVALID_ARG, "ftdx_promote_qc_to_fp8: NULL pointer");
    if ((reinterpret_cast<uintptr_t>(qc) & 15u) || (reinterpret_cast<uintptr_t>(fp8_pairs) & 15u))
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_promote_qc_to_fp8: buffers must be 16-byte aligned");
    hipError_t e = ftdx::launch_promote_qc_fp8_interleaved(static_cast<const uint8_t*>(qc),
                                                           static_cast<uint8_t*>(fp8_pairs), n,
                                                           static_cast<hipStream_t>(stream));
    return e == hipSuccess ? FTDX_OK : hip_fail(e, "ftdx_promote_qc_to_fp8: launch");
}

int ftdx_promote_qc_to_fp16_planar(const void* qc, void* re, void* im, int64_t n, void* stream) {
    if (n < 0) return fail(FTDX_ERR_INVALID_ARG, "ftdx_promote_qc_to_fp16_planar: negative size");
    if (n == 0) return FTDX_OK;
    if (!qc || !re || !im) return fail(FTDX_ERR_INVALID_ARG, "ftdx_promote_qc_to_fp16_planar: NULL pointer");
    if ((reinterpret_cast<uintptr_t>(re) | reinterpret_cast<uintptr_t>(im)) & 1u)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_promote_qc_to_fp16_planar: misaligned output");
    hipError_t e = ftdx::launch_promote_qc_fp16_planar(static_cast<const uint8_t*>(qc),
                                                       static_cast<uint16_t*>(re),
                                                       static_cast<uint16_t*>(im), n,
                                                       static_cast<hipStream_t>(stream));
    return e == hipSuccess ? FTDX_OK : hip_fail(e, "ftdx_promote_qc_to_fp16_planar: launch");
}

int ftdx_promote_data(const void* in, void* out, int64_t n, int prec, void* stream) {
    if (prec != FTDX_PREC_FP32 && prec != FTDX_PREC_FP64)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_promote_data: prec must be FTDX_PREC_FP32 or _FP64");
    if (n < 0) return fail(FTDX_ERR_INVALID_ARG, "ftdx_promote_data: negative size");
    if (n == 0) return FTDX_OK;
    if (!in || !out) return fail(FTDX_ERR_INVALID_ARG, "ftdx_promote_data: NULL pointer");
    if (reinterpret_cast<uintptr_t>(out) & (prec == FTDX_PREC_FP64 ? 7u : 3u))
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_promote_data: misaligned output");
    hipError_t e = ftdx::launch_promote_data(static_cast<const uint8_t*>(in), out, n,
                                             prec == FTDX_PREC_FP64, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? FTDX_OK : hip_fail(e, "ftdx_promote_data: launch");
}

int ftdx_pack_triangle(const void* full, void* packed, const void* old, int N, int batch_count,
                       float beta, void* stream) {
    if (N < 0 || batch_count < 0) return fail(FTDX_ERR_INVALID_ARG, "ftdx_pack_triangle: negative size");
    if (N == 0 || batch_count == 0) return FTDX_OK;
    if (!full || !packed) return fail(FTDX_ERR_INVALID_ARG, "ftdx_pack_triangle: NULL pointer");
    if ((reinterpret_cast<uintptr_t>(full) | reinterpret_cast<uintptr_t>(packed) |
         reinterpret_cast<uintptr_t>(old)) & 7u)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_pack_triangle: buffers must be 8-byte aligned");
    hipError_t e = ftdx::launch_pack_triangle(static_cast<const float2*>(full), static_cast<float2*>(packed),
                                              static_cast<const float2*>(old), N, batch_count, beta,
                                              static_cast<hipStream_t>(stream));
    return e == hipSuccess ? FTDX_OK : hip_fail(e, "ftdx_pack_triangle: launch");
}

int ftdx_tri_from_herm(const void* full, void* tri, int N, int batch_count, int prec, void* stream) {
    if (prec != FTDX_PREC_FP32 && prec != FTDX_PREC_FP64)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_tri_from_herm: prec must be FTDX_PREC_FP32 or _FP64");
    if (N < 0 || batch_count < 0) return fail(FTDX_ERR_INVALID_ARG, "ftdx_tri_from_herm: negative size");
    if (N == 0 || batch_count == 0) return FTDX_OK;
    if (!full || !tri) return fail(FTDX_ERR_INVALID_ARG, "ftdx_tri_from_herm: NULL pointer");
    hipError_t e = ftdx::launch_tri_from_herm(full, tri, N, batch_count, prec == FTDX_PREC_FP64,
                                              static_cast<hipStream_t>(stream));
    return e == hipSuccess ? FTDX_OK : hip_fail(e, "ftdx_tri_from_herm: launch");
}

int ftdx_promote_qc_to_fp32(const void* qc, void* out, int64_t n, void* stream) {
    if (n < 0) return fail(FTDX_ERR_INVALID_ARG, "ftdx_promote_qc_to_fp32: negative size");
    if (n == 0) return FTDX_OK;
    if (!qc || !out) return fail(FTDX_ERR_INVALID_ARG, "ftdx_promote_qc_to_fp32: NULL pointer");
    if (reinterpret_cast<uintptr_t>(out) & 7u)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_promote_qc_to_fp32: output must be 8-byte aligned");
    hipError_t e = ftdx::launch_promote_qc_fp32(static_cast<const uint8_t*>(qc), static_cast<float2*>(out),
                                                n, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? FTDX_OK : hip_fail(e, "ftdx_promote_qc_to_fp32: launch");
}

int ftdx_convert_qc_to_ftd(const void* qc, void* out, int64_t n, void* stream) {
    if (n < 0) return fail(FTDX_ERR_INVALID_ARG, "ftdx_convert_qc_to_ftd: negative size");
    if (n == 0) return FTDX_OK;
    if (!qc || !out) return fail(FTDX_ERR_INVALID_ARG, "ftdx_convert_qc_to_ftd: NULL pointer");
    if ((reinterpret_cast<uintptr_t>(qc) | reinterpret_cast<uintptr_t>(out)) & 15u)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_convert_qc_to_ftd: buffers must be 16-byte aligned");
    hipError_t e = ftdx::launch_convert_qc_to_ftd(static_cast<const uint8_t*>(qc),
                                                  static_cast<uint8_t*>(out), n,
                                                  static_cast<hipStream_t>(stream));
    return e == hipSuccess ? FTDX_OK : hip_fail(e, "ftdx_convert_qc_to_ftd: launch");
}

int ftdx_correlate_ftd(ftdx_context* ctx, const void* raw, void* tri_fp32, int n_ant, int n_chan,
                       int n_time, int n_time_inner, int n_pol, int twos, void* stream) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_correlate_ftd: NULL context");
    if (n_ant < 0 || n_chan < 0 || n_time < 0 || n_time_inner <= 0 || n_pol < 0)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_correlate_ftd: negative size or n_time_inner <= 0");
    const int64_t batch = (int64_t)n_chan * n_pol * n_time_inner;
    if (batch > 0x7FFFFFFF || (int64_t)n_time * n_ant * n_chan * n_pol > ((int64_t)1 << 40))
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_correlate_ftd: sizes too large");
    const int K = n_time / n_time_inner;
    if (n_ant == 0 || batch == 0) return FTDX_OK;
    if (!tri_fp32 || (!raw && K > 0)) return fail(FTDX_ERR_INVALID_ARG, "ftdx_correlate_ftd: NULL pointer");
    int rc = ftdx::grow_precast(ctx, ftdx::rows_bytes(n_ant, K, (int)batch) + 16, "ftdx_correlate_ftd: rows");
    if (rc != FTDX_OK) return rc;
    const int cfg = ftdx::xengine_cfg(ctx->forced_cfg >= 0 ? ftdx::tile_base(ctx->forced_cfg)
                                                           : ftdx::direct_tile_to_cfg(ctx->direct_tile),
                                      n_ant, (int)batch);
    const int remap = ctx->forced_remap >= 0 ? ctx->forced_remap : 1;
    hipError_t e = ftdx::run_ftd_payload(static_cast<const uint8_t*>(raw), static_cast<float*>(tri_fp32),
                                         ctx->precast, n_ant, n_chan, n_time, n_time_inner, n_pol,
                                         twos != 0, cfg, remap, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? FTDX_OK : hip_fail(e, "ftdx_correlate_ftd: kernels");
}

// ---- triangle consumers: polarization sum and pillbox gridding ----
int ftdx_pol_reduce_triangle(const void* tri_in, void* tri_out, int N, int n_ch, int n_pol, int n_time_inner,
                             int output, void* stream) {
    if (N < 0 || n_ch < 0 || n_pol <= 0 || n_time_inner <= 0)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_pol_reduce_triangle: bad sizes");
    if (output != FTDX_OUTPUT_FP32 && output != FTDX_OUTPUT_FP16)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_pol_reduce_triangle: output must be FP32 or FP16");
    if ((int64_t)N * (N + 1) * n_ch * n_time_inner == 0) return FTDX_OK;
    if (!tri_in || !tri_out) return fail(FTDX_ERR_INVALID_ARG, "ftdx_pol_reduce_triangle: NULL pointer");
    hipError_t e = ftdx::launch_pol_reduce(tri_in, tri_out, N, n_ch, n_pol, n_time_inner,
                                           output == FTDX_OUTPUT_FP16, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? FTDX_OK : hip_fail(e, "ftdx_pol_reduce_triangle");
}

int ftdx_herk_pol_reduce(ftdx_context* ctx, const void* A, void* C, int N, int K, int n_ch, int n_pol,
                         int n_time_inner, int output, float alpha, float beta, void* stream) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_pol_reduce: NULL context");
    if (N < 0 || K < 0 || n_ch < 0 || n_pol <= 0 || n_time_inner <= 0)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_pol_reduce: bad sizes");
    if (output != FTDX_OUTPUT_FP32 && output != FTDX_OUTPUT_FP16)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_pol_reduce: output must be FP32 or FP16");
    const int64_t out_batch = (int64_t)n_ch * n_time_inner;
    if ((int64_t)K * n_pol > 0x7FFFFFFF || out_batch * n_pol > 0x7FFFFFFF)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_pol_reduce: sizes too large");
    if (N == 0 || out_batch == 0) return FTDX_OK;
    if (!C || (!A && K > 0)) return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_pol_reduce: NULL pointer");
    ftdx::HerkPlainArgs a{};
    a.A = static_cast<const uint8_t*>(A);
    a.C = C;
    a.N = N;
    a.K = K;
    a.batch = (int)out_batch;
    a.alpha = alpha;
    a.beta = beta;
    a.a_bstride = (int64_t)N * K;
    a.c_bstride = (int64_t)N * (N + 1);
    a.nseg = n_pol;
    a.seg_group = n_time_inner;
    const int cfg = ctx->forced_cfg >= 0 ? ftdx::tile_base(ctx->forced_cfg)
                                         : (ftdx::direct_tile_to_cfg(ctx->direct_tile) >= 0
                                                ? ftdx::direct_tile_to_cfg(ctx->direct_tile)
                                                : ftdx::kDefaultCfg);
    const int remap = ctx->forced_remap >= 0 ? ctx->forced_remap : ftdx::kDefaultRemap;
    const bool fp16 = output == FTDX_OUTPUT_FP16;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e;
    if (ftdx::pol_rows_padded(A, K, n_pol, cfg)) {
        // rows padded per chunk of polarization groups, as ftdx_herk does
        int rc = ftdx::grow_precast(ctx, ftdx::padded_scratch_bytes(a), "ftdx_herk_pol_reduce: scratch");
        if (rc != FTDX_OK) return rc;
        e = ftdx::run_padded(ctx, a, cfg, fp16 ? 2 : 4, s, [&](const ftdx::HerkPlainArgs& b) {
            return ftdx::launch_herk_int4_plain(b, cfg, fp16, remap, s);
        });
    } else {
        e = ftdx::launch_herk_int4_plain(a, cfg, fp16, remap, s);
    }
    return e == hipSuccess ? FTDX_OK : hip_fail(e, "ftdx_herk_pol_reduce");
}

int ftdx_herk_pol_grid(ftdx_context* ctx, const void* A, int N, int K, int n_chan, int n_pol,
                       const float* baseline_uv_m, const double* freq_hz, float* uv_grid, int Ng,
                       int freq_offset, float cell_size_rad, float alpha, void* stream) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_pol_grid: NULL context");
    if (N < 0 || K < 0 || n_chan < 0 || n_pol <= 0 || Ng < 0 || freq_offset < 0)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_pol_grid: bad sizes");
    if ((int64_t)K * n_pol > 0x7FFFFFFF || (int64_t)n_chan * n_pol > 0x7FFFFFFF ||
        (int64_t)Ng * Ng > ((int64_t)1 << 40))
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_pol_grid: sizes too large");
    if (N == 0 || n_chan == 0 || Ng == 0) return FTDX_OK;
    if (!uv_grid || !baseline_uv_m || !freq_hz || (!A && K > 0))
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_pol_grid: NULL pointer");
    ftdx::HerkPlainArgs a{};
    a.A = static_cast<const uint8_t*>(A);
    a.C = uv_grid;
    a.N = N;
    a.K = K;
    a.batch = n_chan;
    a.alpha = alpha;
    a.beta = 0.0f;
    a.a_bstride = (int64_t)N * K;
    a.c_bstride = (int64_t)2 * Ng * Ng;
    a.nseg = n_pol;
    a.seg_group = 1;
    a.uv_m = baseline_uv_m;
    a.freq_hz = freq_hz;
    a.freq_offset = freq_offset;
    a.Ng = Ng;
    a.cell = cell_size_rad;
    const int cfg = ctx->forced_cfg >= 0 ? ftdx::tile_base(ctx->forced_cfg)
                                         : (ftdx::direct_tile_to_cfg(ctx->direct_tile) >= 0
                                                ? ftdx::direct_tile_to_cfg(ctx->direct_tile)
                                                : ftdx::kDefaultCfg);
    const int remap = ctx->forced_remap >= 0 ? ctx->forced_remap : ftdx::kDefaultRemap;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e;
    if (ftdx::pol_rows_padded(A, K, n_pol, cfg)) {
        // rows padded per chunk of channels, as ftdx_herk does
        int rc = ftdx::grow_precast(ctx, ftdx::padded_scratch_bytes(a), "ftdx_herk_pol_grid: scratch");
        if (rc != FTDX_OK) return rc;
        e = ftdx::run_padded(ctx, a, cfg, 4, s, [&](const ftdx::HerkPlainArgs& b) {
            return ftdx::launch_herk_grid(b, cfg, remap, s);
        });
    } else {
        e = ftdx::launch_herk_grid(a, cfg, remap, s);
    }
    return e == hipSuccess ? FTDX_OK : hip_fail(e, "ftdx_herk_pol_grid");
}

int ftdx_pillbox_grid(const void* vis_tri, int vis_output, const float* baseline_uv_m,
                      const double* freq_hz, float* uv_grid, int n_baselines, int Ng, int Nf,
                      int freq_offset, float cell_size_rad, void* stream) {
    if (n_baselines < 0 || Ng < 0 || Nf < 0 || freq_offset < 0)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_pillbox_grid: bad sizes");
    if (vis_output != FTDX_OUTPUT_FP32 && vis_output != FTDX_OUTPUT_FP16)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_pillbox_grid: visibilities must be FP32 or FP16");
    if ((int64_t)n_baselines * Nf == 0 || Ng == 0) return FTDX_OK;
    if (!vis_tri || !baseline_uv_m || !freq_hz || !uv_grid)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_pillbox_grid: NULL pointer");
    hipError_t e = ftdx::launch_pillbox_grid(vis_tri, vis_output == FTDX_OUTPUT_FP16, baseline_uv_m, freq_hz,
                                             uv_grid, n_baselines, Ng, Nf, freq_offset, cell_size_rad,
                                             static_cast<hipStream_t>(stream));
    return e == hipSuccess ? FTDX_OK : hip_fail(e, "ftdx_pillbox_grid");
}

// ---- complex GEMM with prepared B ----
namespace {

// FP8: conj(B) as pre-cast rows (fmt = -1).  FP6 / FP4: MX operand rows of conj(B), codes
// and scales (fmt = the MxFormat), the block-scaled path of cutlass_gemm_api.h:40-47.
int prepare_rows(ftdx_context* ctx, uint8_t*& buf, size_t& cap,
                 const void* B_re, const void* B_im, int N, int K, int batch, int compute, void* stream,
                 const char* what, int& fmt) {
    const std::string w(what);
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, w + ": NULL context");
    if (N < 0 || K < 0 || batch < 0) return fail(FTDX_ERR_INVALID_ARG, w + ": negative size");
    if (compute < FTDX_COMPUTE_FP8 || compute > FTDX_COMPUTE_FP4)
        return fail(FTDX_ERR_INVALID_ARG, w + ": unknown compute precision");
    if (K > 0 && (int64_t)N * batch > 0 && (!B_re || !B_im))
        return fail(FTDX_ERR_INVALID_ARG, w + ": NULL B pointer");
    if (compute != FTDX_COMPUTE_FP8) {
        fmt = ftdx::mx_format(compute);
        const int nblkp = ftdx::mx_padded_blocks(K);
        const int64_t rows = (int64_t)N * batch;
        int rc = ftdx::ensure_scratch(buf, cap, (size_t)rows * nblkp * 64 + 16, what);
        if (rc != FTDX_OK) return rc;
        hipError_t e = ftdx::launch_mx_quantize_rows(B_re, B_im, ftdx::MX_IN_FP16_PLANAR, fmt, ctx->mx_rounding,
                                                     true, buf, rows, K,
                                                     static_cast<hipStream_t>(stream));
        return e == hipSuccess ? FTDX_OK : hip_fail(e, what);
    }
    fmt = -1;
    const int Kp = ftdx::precast_row_samples(K);
    const size_t need = (size_t)N * batch * Kp * 2 + 16;
    int rc = ftdx::ensure_scratch(buf, cap, need, what);
    if (rc != FTDX_OK) return rc;
    hipError_t e = ftdx::launch_promote_fp16_planar_rows(static_cast<const uint16_t*>(B_re),
                                                        static_cast<const uint16_t*>(B_im), buf,
                                                        (int64_t)N * batch, K, true,
                                                        static_cast<hipStream_t>(stream));
    return e == hipSuccess ? FTDX_OK : hip_fail(e, what);
}

int check_prepared(ftdx_context* ctx, int N, int K, int batch, const char* what, bool fp8_only = false) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, std::string(what) + ": NULL context");
    if (!ctx->prep_b || ctx->prep_N != N || ctx->prep_K != K || ctx->prep_batch != batch)
        return fail(FTDX_ERR_INVALID_ARG,
                    std::string(what) + ": prepare_b() must have been called with matching N, K, batch");
    // the direct and power variants need FP8-prepared B (cutlass_gemm_api.h:319, 346, 374, 396)
    if (fp8_only && ctx->prep_fmt >= 0)
        return fail(FTDX_ERR_UNSUPPORTED,
                    std::string(what) + ": prepare_b() must have been called with ComputePrecision::FP8");
    return FTDX_OK;
}

// Block-scaled complex GEMM: quantize A (FP16 planar or INT4) into the MX scratch, then
// C = alpha A B^T + beta C against the MX codes of conj(B).
int gemm_mx_run(ftdx_context* ctx, const uint8_t* brows, int fmt, bool int4, const void* a0, const void* a1,
                void* c0, void* c1, int M, int N, int K, int batch, bool out_fp16, float alpha, float beta,
                void* stream, const char* what) {
    const std::string w(what);
    if (M < 0) return fail(FTDX_ERR_INVALID_ARG, w + ": M must be >= 0");
    if (M == 0 || N == 0 || batch == 0) return FTDX_OK;
    if (!c0 || !c1) return fail(FTDX_ERR_INVALID_ARG, w + ": NULL output pointer");
    if (K > 0 && (!a0 || (!int4 && !a1))) return fail(FTDX_ERR_INVALID_ARG, w + ": NULL A pointer");
    if ((int64_t)M * N > ((int64_t)1 << 40) || (int64_t)M * K > ((int64_t)1 << 40))
        return fail(FTDX_ERR_INVALID_ARG, w + ": sizes too large");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int nblkp = ftdx::mx_padded_blocks(K);
    const int64_t rows = (int64_t)M * batch, pitch = (int64_t)nblkp * 64;
    int rc = ftdx::ensure_scratch(ctx->mx_codes, ctx->mx_codes_cap, (size_t)(rows * pitch) + 16, what);
    if (rc != FTDX_OK) return rc;
    hipError_t e = ftdx::launch_mx_quantize_rows(a0, a1, int4 ? ftdx::MX_IN_INT4 : ftdx::MX_IN_FP16_PLANAR, fmt,
                                                 ctx->mx_rounding, false, ctx->mx_codes, rows, K, s);
    if (e != hipSuccess) return hip_fail(e, what);
    ftdx::GemmArgs g;
    g.A = ctx->mx_codes;
    g.B = brows;
    g.C = c0;
    g.C2 = c1;
    g.M = M;
    g.N = N;
    g.K = nblkp * 32;
    g.batch = batch;
    g.alpha = alpha;
    g.beta = beta;
    g.lda = g.ldb = pitch;
    g.a_bstride = (int64_t)M * pitch;
    g.b_bstride = (int64_t)N * pitch;
    g.ldc = N;
    g.c_bstride = (int64_t)M * N;
    g.operand = fmt == ftdx::MX_FP6_E3M2 ? 3 : 4;
    const int remap = ctx->forced_remap >= 0 ? ctx->forced_remap : ftdx::kDefaultRemap;
    e = ftdx::launch_gemm_fp8(g, ftdx::TILE_128x128_K64, ftdx::OUT_GEMM_PLANAR, out_fp16, remap, s);
    return e == hipSuccess ? FTDX_OK : hip_fail(e, what);
}

}  // namespace

int ftdx_prepare_b(ftdx_context* ctx, const void* B_re, const void* B_im, int N, int K, int batch_count,
                   int compute, void* stream) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_prepare_b: NULL context");
    int fmt = -1;
    int rc = prepare_rows(ctx, ctx->prep_b, ctx->prep_b_cap, B_re, B_im, N,
                          K, batch_count, compute, stream, "ftdx_prepare_b", fmt);
    if (rc != FTDX_OK) {
        ctx->prep_N = ctx->prep_K = ctx->prep_batch = -1;
        return rc;
    }
    ctx->prep_fmt = fmt;
    ctx->prep_N = N;
    ctx->prep_K = K;
    ctx->prep_batch = batch_count;
    return FTDX_OK;
}

int ftdx_gemm_prepared(ftdx_context* ctx, const void* A_re, const void* A_im, void* C_re, void* C_im,
                       int M, int N, int K, int batch_count, int output, float alpha, float beta,
                       void* stream, int tune, int input) {
    (void)tune;
    int rc = check_prepared(ctx, N, K, batch_count, "ftdx_gemm_prepared");
    if (rc != FTDX_OK) return rc;
    if (output != FTDX_OUTPUT_FP32 && output != FTDX_OUTPUT_FP16)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_gemm_prepared: output must be FP32 or FP16");
    if (input != FTDX_INPUT_FP16 && input != FTDX_INPUT_INT4)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_gemm_prepared: input must be FP16 or INT4");
    if (ctx->prep_fmt >= 0)
        return gemm_mx_run(ctx, ctx->prep_b, ctx->prep_fmt, input == FTDX_INPUT_INT4, A_re, A_im,
                           C_re, C_im, M, N, K, batch_count, output == FTDX_OUTPUT_FP16, alpha, beta, stream,
                           "ftdx_gemm_prepared");
    return ftdx::gemm_run(ctx, ctx->prep_b, input == FTDX_INPUT_INT4 ? ftdx::GEMM_A_INT4 : ftdx::GEMM_A_FP16_PLANAR,
                          A_re, A_im, C_re, C_im, M, N, K, batch_count, ftdx::OUT_GEMM_PLANAR,
                          output == FTDX_OUTPUT_FP16, alpha, beta, stream, "ftdx_gemm_prepared");
}

int ftdx_gemm_prepared_direct(ftdx_context* ctx, const void* A_re, const void* A_im, void* C_re,
                              void* C_im, int M, int N, int K, int batch_count, float alpha, float beta,
                              void* stream) {
    int rc = check_prepared(ctx, N, K, batch_count, "ftdx_gemm_prepared_direct", true);
    if (rc != FTDX_OK) return rc;
    return ftdx::gemm_run(ctx, ctx->prep_b, ftdx::GEMM_A_FP16_PLANAR, A_re, A_im, C_re, C_im, M, N, K,
                          batch_count, ftdx::OUT_GEMM_PLANAR, false, alpha, beta, stream,
                          "ftdx_gemm_prepared_direct");
}

int ftdx_gemm_prepared_direct_int4(ftdx_context* ctx, const void* A_int4, void* C_re, void* C_im, int M,
                                   int N, int K, int batch_count, float alpha, float beta, void* stream) {
    int rc = check_prepared(ctx, N, K, batch_count, "ftdx_gemm_prepared_direct_int4", true);
    if (rc != FTDX_OK) return rc;
    return ftdx::gemm_run(ctx, ctx->prep_b, ftdx::GEMM_A_INT4, A_int4, nullptr, C_re, C_im, M, N, K,
                          batch_count, ftdx::OUT_GEMM_PLANAR, false, alpha, beta, stream,
                          "ftdx_gemm_prepared_direct_int4");
}

int ftdx_gemm_prepared_power(ftdx_context* ctx, const void* A_re, const void* A_im, void* C_power, int M,
                             int N, int K, int batch_count, float alpha, float beta, void* stream) {
    int rc = check_prepared(ctx, N, K, batch_count, "ftdx_gemm_prepared_power", true);
    if (rc != FTDX_OK) return rc;
    return ftdx::gemm_run(ctx, ctx->prep_b, ftdx::GEMM_A_FP16_PLANAR, A_re, A_im, C_power, nullptr, M, N,
                          K, batch_count, ftdx::OUT_GEMM_POWER, false, alpha, beta, stream,
                          "ftdx_gemm_prepared_power");
}

int ftdx_gemm_prepared_power_fp8(ftdx_context* ctx, const void* A_fp8_interleaved, void* C_power, int M,
                                 int N, int K, int batch_count, float alpha, float beta, void* stream) {
    int rc = check_prepared(ctx, N, K, batch_count, "ftdx_gemm_prepared_power_fp8", true);
    if (rc != FTDX_OK) return rc;
    return ftdx::gemm_run(ctx, ctx->prep_b, ftdx::GEMM_A_FP8_INTERLEAVED, A_fp8_interleaved, nullptr,
                          C_power, nullptr, M, N, K, batch_count, ftdx::OUT_GEMM_POWER, false, alpha, beta,
                          stream, "ftdx_gemm_prepared_power_fp8");
}

int ftdx_gemm_prepared_power_int4(ftdx_context* ctx, const void* A_int4, void* C_power, int M, int N,
                                  int K, int batch_count, float alpha, float beta, void* stream) {
    int rc = check_prepared(ctx, N, K, batch_count, "ftdx_gemm_prepared_power_int4", true);
    if (rc != FTDX_OK) return rc;
    return ftdx::gemm_run(ctx, ctx->prep_b, ftdx::GEMM_A_INT4, A_int4, nullptr, C_power, nullptr, M, N, K,
                          batch_count, ftdx::OUT_GEMM_POWER, false, alpha, beta, stream,
                          "ftdx_gemm_prepared_power_int4");
}

namespace {

// fuse factor 0: the reference's choice (cutlass_gemm_api.cu:951-972): M already >= 256
// stays unfused; else the largest power of two up to 32 that divides the batch with
// M_fused <= 1024, then, if M_fused < 128, the smallest divisor reaching 128 (up to 128).
int auto_fuse(int M, int batch) {
    if (M <= 0 || batch <= 0 || M >= 256) return 1;
    int ff = 1;
    for (int f = 2; f <= 32 && M * f <= 1024; f *= 2)
        if (batch % f == 0) ff = f;
    if (M * ff < 128)
        for (int f = ff + 1; f <= 128; ++f)
            if (batch % f == 0 && M * f >= 128) return f;
    return ff;
}

}  // namespace

int ftdx_gemm_prepared_fused(ftdx_context* ctx, const void* A_re, const void* A_im, void* C_re, void* C_im,
                             int M, int N, int K, int batch_count, int fuse_factor, int output, float alpha,
                             float beta, void* stream) {
    const char* W = "ftdx_gemm_prepared_fused";
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, std::string(W) + ": NULL context");
    if (output != FTDX_OUTPUT_FP32)
        return fail(FTDX_ERR_UNSUPPORTED, "gemm_prepared_fused: only FP32 output is currently supported");
    if (M < 0 || batch_count < 0) return fail(FTDX_ERR_INVALID_ARG, std::string(W) + ": negative size");
    const int ff = fuse_factor <= 0 ? auto_fuse(M, batch_count) : fuse_factor;
    if (batch_count > 0 && batch_count % ff != 0)
        return fail(FTDX_ERR_INVALID_ARG, "gemm_prepared_fused: fuse_factor must divide batch_count evenly");
    if (ff == 1)
        return ftdx_gemm_prepared(ctx, A_re, A_im, C_re, C_im, M, N, K, batch_count, output, alpha, beta, stream, 0,
                                  FTDX_INPUT_FP16);
    if (!ctx->prep_b || ctx->prep_N != N || ctx->prep_K != K)
        return fail(FTDX_ERR_INVALID_ARG, std::string(W) + ": prepare_b() must have been called with matching N, K");
    if (ctx->prep_fmt >= 0)
        return fail(FTDX_ERR_UNSUPPORTED, std::string(W) + ": prepare_b() must have been called with FP8 compute");
    const int eb = batch_count / ff;
    if (eb > ctx->prep_batch)
        return fail(FTDX_ERR_INVALID_ARG, std::string(W) + ": more fused batches than prepared weight sets");
    if ((int64_t)M * ff > 0x7FFFFFFF) return fail(FTDX_ERR_INVALID_ARG, std::string(W) + ": sizes too large");
    return ftdx::gemm_run(ctx, ctx->prep_b, ftdx::GEMM_A_FP16_PLANAR, A_re, A_im, C_re, C_im, M * ff, N, K, eb,
                          ftdx::OUT_GEMM_PLANAR, false, alpha, beta, stream, W);
}

int ftdx_set_gemm_setting(ftdx_context* ctx, int setting, int value) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_gemm_setting: NULL context");
    static const int limits[5] = {8, 2, 2, 1, 1 << 30};
    if (setting < 0 || setting > 4) return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_gemm_setting: unknown setting");
    if (value < 0 || value > limits[setting]) return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_gemm_setting: bad value");
    ctx->gemm_settings[setting] = value;
    return FTDX_OK;
}

int ftdx_get_gemm_setting(const ftdx_context* ctx, int setting, int* value) {
    if (!ctx || !value) return fail(FTDX_ERR_INVALID_ARG, "ftdx_get_gemm_setting: NULL argument");
    if (setting < 0 || setting > 4) return fail(FTDX_ERR_INVALID_ARG, "ftdx_get_gemm_setting: unknown setting");
    *value = ctx->gemm_settings[setting];
    return FTDX_OK;
}

int ftdx_set_gemm_tune_cache_path(ftdx_context* ctx, const char* path) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_set_gemm_tune_cache_path: NULL context");
    ctx->gemm_tune_cache_path = path ? path : "";
    return FTDX_OK;
}

int ftdx_debug_fp6_real_gemm(ftdx_context* ctx, int M, int N, int K, void* stream) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_debug_fp6_real_gemm: NULL context");
    if (M <= 0 || N <= 0 || K <= 0 || (int64_t)M * K > (1 << 26) || (int64_t)N * K > (1 << 26) ||
        (int64_t)M * N > (1 << 26))
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_debug_fp6_real_gemm: sizes must be positive and modest");
    std::fprintf(stderr, "\n=== debug_fp6_real_gemm M=%d N=%d K=%d ===\n", M, N, K);
    const size_t sA = (size_t)M * K, sB = (size_t)N * K, sC = (size_t)M * N;
    // integers in [-7, 7] from a fixed linear congruential stream
    std::vector<uint16_t> hA(sA), hB(sB), hz(std::max(sA, sB), 0);
    std::vector<double> ref(sC, 0.0);
    std::vector<int> a(sA), b(sB);
    uint64_t st = 42;
    auto next = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return (int)((st >> 33) % 15) - 7;
    };
    auto f16 = [](int v) {  // exact FP16 bits of a small integer
        if (v == 0) return (uint16_t)0;
        const uint16_t sgn = v < 0 ? 0x8000 : 0;
        int m = v < 0 ? -v : v, e = 0;
        while ((m >> (e + 1)) != 0) ++e;
        return (uint16_t)(sgn | ((e + 15) << 10) | ((m << (10 - e)) & 0x3FF));
    };
    for (size_t i = 0; i < sA; ++i) hA[i] = f16(a[i] = next());
    for (size_t i = 0; i < sB; ++i) hB[i] = f16(b[i] = next());
    for (int m = 0; m < M; ++m)
        for (int n = 0; n < N; ++n) {
            double acc = 0.0;
            for (int k = 0; k < K; ++k) acc += (double)a[(size_t)m * K + k] * b[(size_t)n * K + k];
            ref[(size_t)m * N + n] = acc;
        }
    hipStream_t s = static_cast<hipStream_t>(stream);
    void *dA = nullptr, *dB = nullptr, *dz = nullptr, *dC = nullptr, *dCi = nullptr;
    hipError_t e = hipMalloc(&dA, sA * 2);
    if (e == hipSuccess) e = hipMalloc(&dB, sB * 2);
    if (e == hipSuccess) e = hipMalloc(&dz, hz.size() * 2);
    if (e == hipSuccess) e = hipMalloc(&dC, sC * 4);
    if (e == hipSuccess) e = hipMalloc(&dCi, sC * 4);
    if (e == hipSuccess) e = hipMemcpy(dA, hA.data(), sA * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dB, hB.data(), sB * 2, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dz, hz.data(), hz.size() * 2, hipMemcpyHostToDevice);
    int result = e == hipSuccess ? 0 : -1;
    if (result == 0) {
        for (int pass = 0; pass < 2 && result == 0; ++pass) {
            const int out = pass == 0 ? FTDX_OUTPUT_FP16 : FTDX_OUTPUT_FP32;
            int rc = ftdx_gemm(ctx, dA, dz, dB, dz, dC, dCi, M, N, K, 1, FTDX_COMPUTE_FP6, out, 1.0f, 0.0f, s, 0,
                               FTDX_INPUT_FP16);
            if (rc == FTDX_OK) e = hipStreamSynchronize(s);
            if (rc != FTDX_OK || e != hipSuccess) {
                result = -1;
                break;
            }
            double max_ref = 0.0, max_err = 0.0;
            if (pass == 0) {
                std::vector<uint16_t> h(sC);
                if (hipMemcpy(h.data(), dC, sC * 2, hipMemcpyDeviceToHost) != hipSuccess) { result = -1; break; }
                for (size_t i = 0; i < sC; ++i) {
                    _Float16 v;
                    std::memcpy(&v, &h[i], 2);
                    max_ref = std::max(max_ref, std::fabs(ref[i]));
                    max_err = std::max(max_err, std::fabs((double)(float)v - ref[i]));
                }
            } else {
                std::vector<float> h(sC);
                if (hipMemcpy(h.data(), dC, sC * 4, hipMemcpyDeviceToHost) != hipSuccess) { result = -1; break; }
                for (size_t i = 0; i < sC; ++i) {
                    max_ref = std::max(max_ref, std::fabs(ref[i]));
                    max_err = std::max(max_err, std::fabs((double)h[i] - ref[i]));
                }
            }
            // FP16 output rounds sums beyond 2048 to its 11-bit precision
            const double tol = pass == 0 ? max_ref * 1e-3 : 0.0;
            const bool ok = max_err <= tol;
            std::fprintf(stderr, "  %s output: max_ref=%.1f max_err=%.2f  %s\n", pass == 0 ? "FP16" : "FP32",
                         max_ref, max_err, ok ? "PASS" : "FAIL");
            if (!ok) result = 1;
        }
    }
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dz);
    (void)hipFree(dC);
    (void)hipFree(dCi);
    if (result < 0) return hip_fail(e == hipSuccess ? hipErrorUnknown : e, "ftdx_debug_fp6_real_gemm");
    return result == 0 ? FTDX_OK : fail(FTDX_ERR_INVALID_ARG, "debug_fp6_real_gemm: results differ");
}

namespace {

// Scratch owned by one BLAS-glue call (the reference allocates and frees per call).
struct CallScratch {
    std::vector<void*> p;
    ~CallScratch() {
        for (void* q : p) (void)hipFree(q);
    }
    void* get(size_t bytes, hipError_t& e) {
        void* q = nullptr;
        if (e == hipSuccess) e = hipMalloc(&q, bytes > 0 ? bytes : 1);
        if (e == hipSuccess) p.push_back(q);
        return q;
    }
};

}  // namespace

int ftdx_blas_herk_fp32(ftdx_context* ctx, const void* A, void* C, int N, int K, int batch_count, int tune) {
    const char* W = "ftdx_blas_herk_fp32";
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, std::string(W) + ": NULL context");
    if (N < 0 || K < 0 || batch_count < 0) return fail(FTDX_ERR_INVALID_ARG, std::string(W) + ": negative size");
    if ((int64_t)N * batch_count == 0) return FTDX_OK;
    if (!C || (!A && K > 0)) return fail(FTDX_ERR_INVALID_ARG, std::string(W) + ": NULL pointer");
    const int64_t nA = (int64_t)N * K * 2 * batch_count, nC = (int64_t)N * (N + 1) * batch_count;
    CallScratch sc;
    hipError_t e = hipSuccess;
    uint16_t* a16 = static_cast<uint16_t*>(sc.get((size_t)nA * 2, e));
    uint16_t* c16 = static_cast<uint16_t*>(sc.get((size_t)nC * 2, e));
    if (e == hipSuccess) e = ftdx::launch_narrow_fp32(static_cast<const float*>(A), a16, nA, nullptr);
    if (e != hipSuccess) return hip_fail(e, W);
    int rc = ftdx_herk(ctx, a16, c16, N, K, batch_count, FTDX_INPUT_FP16, FTDX_COMPUTE_FP8, FTDX_OUTPUT_FP16,
                       FTDX_FORMAT_PACKED_TRIANGLE, 1.0f, 0.0f, nullptr, tune);
    if (rc != FTDX_OK) return rc;
    e = ftdx::launch_widen_fp16(c16, static_cast<float*>(C), nC, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return e == hipSuccess ? FTDX_OK : hip_fail(e, W);
}

int ftdx_blas_gemm_fp32(ftdx_context* ctx, const void* A, const void* B, void* C, int M, int N, int K,
                        int batch_count, int conjugate_b, int tune) {
    const char* W = "ftdx_blas_gemm_fp32";
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, std::string(W) + ": NULL context");
    if (M < 0 || N < 0 || K < 0 || batch_count < 0) return fail(FTDX_ERR_INVALID_ARG, std::string(W) + ": negative size");
    if ((int64_t)M * N * batch_count == 0) return FTDX_OK;
    if (!C || (K > 0 && (!A || !B))) return fail(FTDX_ERR_INVALID_ARG, std::string(W) + ": NULL pointer");
    const int64_t nA = (int64_t)M * K * batch_count, nB = (int64_t)N * K * batch_count,
                  nC = (int64_t)M * N * batch_count;
    CallScratch sc;
    hipError_t e = hipSuccess;
    uint16_t* are = static_cast<uint16_t*>(sc.get((size_t)nA * 2, e));
    uint16_t* aim = static_cast<uint16_t*>(sc.get((size_t)nA * 2, e));
    uint16_t* bre = static_cast<uint16_t*>(sc.get((size_t)nB * 2, e));
    uint16_t* bim = static_cast<uint16_t*>(sc.get((size_t)nB * 2, e));
    float* cre = static_cast<float*>(sc.get((size_t)nC * 4, e));
    float* cim = static_cast<float*>(sc.get((size_t)nC * 4, e));
    if (e == hipSuccess) e = ftdx::launch_deinterleave_narrow(static_cast<const float*>(A), are, aim, nA, false, nullptr);
    if (e == hipSuccess)
        e = ftdx::launch_deinterleave_narrow(static_cast<const float*>(B), bre, bim, nB, conjugate_b == 0, nullptr);
    if (e != hipSuccess) return hip_fail(e, W);
    int rc = ftdx_gemm(ctx, are, aim, bre, bim, cre, cim, M, N, K, batch_count, FTDX_COMPUTE_FP8, FTDX_OUTPUT_FP32,
                       1.0f, 0.0f, nullptr, tune, FTDX_INPUT_FP16);
    if (rc != FTDX_OK) return rc;
    e = ftdx::launch_interleave_fp32(cre, cim, static_cast<float*>(C), nC, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return e == hipSuccess ? FTDX_OK : hip_fail(e, W);
}

int ftdx_gemm(ftdx_context* ctx, const void* A_re, const void* A_im, const void* B_re, const void* B_im,
              void* C_re, void* C_im, int M, int N, int K, int batch_count, int compute, int output,
              float alpha, float beta, void* stream, int tune, int input) {
    (void)tune;
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_gemm: NULL context");
    if (output != FTDX_OUTPUT_FP32 && output != FTDX_OUTPUT_FP16)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_gemm: output must be FP32 or FP16");
    if (input != FTDX_INPUT_FP16 && input != FTDX_INPUT_INT4)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_gemm: input must be FP16 or INT4");
    int fmt = -1;
    int rc = prepare_rows(ctx, ctx->gemm_b, ctx->gemm_b_cap, B_re, B_im, N, K,
                          batch_count, compute, stream, "ftdx_gemm", fmt);
    if (rc != FTDX_OK) return rc;
    if (fmt >= 0)
        return gemm_mx_run(ctx, ctx->gemm_b, fmt, input == FTDX_INPUT_INT4, A_re, A_im, C_re, C_im,
                           M, N, K, batch_count, output == FTDX_OUTPUT_FP16, alpha, beta, stream, "ftdx_gemm");
    return ftdx::gemm_run(ctx, ctx->gemm_b, input == FTDX_INPUT_INT4 ? ftdx::GEMM_A_INT4 : ftdx::GEMM_A_FP16_PLANAR,
                          A_re, A_im, C_re, C_im, M, N, K, batch_count, ftdx::OUT_GEMM_PLANAR,
                          output == FTDX_OUTPUT_FP16, alpha, beta, stream, "ftdx_gemm");
}

int ftdx_interleave_tri(const void* re, const void* im, void* out, int64_t n, int prec, void* stream) {
    if (prec != FTDX_PREC_FP32 && prec != FTDX_PREC_FP64)
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_interleave_tri: prec must be FTDX_PREC_FP32 or _FP64");
    if (n < 0) return fail(FTDX_ERR_INVALID_ARG, "ftdx_interleave_tri: negative size");
    if (n == 0) return FTDX_OK;
    if (!re || !im || !out) return fail(FTDX_ERR_INVALID_ARG, "ftdx_interleave_tri: NULL pointer");
    hipError_t e = ftdx::launch_interleave_tri(static_cast<const uint16_t*>(re),
                                               static_cast<const uint16_t*>(im), out, n,
                                               prec == FTDX_PREC_FP64, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? FTDX_OK : hip_fail(e, "ftdx_interleave_tri: launch");
}

int ftdx_herk_each(ftdx_context* ctx, const ftdx_herk_problem* problems, int count, int output,
                   void* stream) {
    if (!ctx) return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_each: NULL context");
    if (count < 0 || (count > 0 && !problems))
        return fail(FTDX_ERR_INVALID_ARG, "ftdx_herk_each: bad problem list");
    for (int p = 0; p < count; ++p) {
        const ftdx_herk_problem& q = problems[p];
        int rc = ftdx_herk(ctx, q.A, q.C, q.N, q.K, 1, FTDX_INPUT_INT4, FTDX_COMPUTE_FP8, output,
                           FTDX_FORMAT_PACKED_TRIANGLE, q.alpha, q.beta, stream, 0);
        if (rc != FTDX_OK) return rc;
    }
    return FTDX_OK;
}

int ftdx_herk_int4(const void* A, void* C, int N, int K, int batch_count, int output,
                   float alpha, float beta, void* stream) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "ftdx_herk_int4: hipGetDevice");
    ftdx_context*& d = ftdx::g_default_ctx;
    if (d && d->device != dev) {
        ftdx_destroy(d);
        d = nullptr;
    }
    if (!d) {
        int rc = ftdx_create(&d);
        if (rc != FTDX_OK) return rc;
    }
    return ftdx_herk(d, A, C, N, K, batch_count, FTDX_INPUT_INT4, FTDX_COMPUTE_FP8, output,
                     FTDX_FORMAT_PACKED_TRIANGLE, alpha, beta, stream, 0);
}

int ftdx_print_build_info(void) {
    std::printf("ftdx: MI355X X-engine HERK library, ABI %d\n", FTDX_ABI_VERSION);
    std::printf("  target: gfx950 (CDNA4), FP8 E4M3 MFMA v_mfma_f32_32x32x64_f8f6f4\n");
    std::printf("  tile configurations:");
    for (int c = 0; c < ftdx::NUM_TILE_CONFIGS; ++c) std::printf(" %s", ftdx::tile_config_name(c));
    std::printf("\n  tune fingerprint: %s\n", ftdx::TuneCache::fingerprint().c_str());
    int n = 0;
    if (hipGetDeviceCount(&n) == hipSuccess && n > 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) == hipSuccess)
            std::printf("  device %d: %s (%s), %d CUs, %.0f MHz, %.1f GiB\n", dev, p.name,
                        p.gcnArchName, p.multiProcessorCount, p.clockRate / 1e3,
                        p.totalGlobalMem / 1073741824.0);
    } else {
        std::printf("  no HIP device visible\n");
    }
    return FTDX_OK;
}

}  // extern "C"
