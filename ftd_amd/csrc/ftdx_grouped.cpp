// ftdx_grouped.cpp -- grouped HERK over many problems of uneven N and K in one
// persistent launch (optionally replayed from a captured hipGraph).
//
// The reference runs one uniform (N, K, batch) HERK per call
// (cutlass_interface/cutlass_gemm_api.cu:1707-1878); its callers loop over
// sub-arrays.  Here a plan holds, in device memory, the problem descriptors and
// a static schedule: every lower-triangle tile of every problem is one item;
// items are sorted by cost (K rounded to the stage, less for diagonal tiles)
// and assigned longest-first to the least-loaded persistent workgroup.  Ties go
// to workgroups of the same XCD first, so items that share input rows run on
// one XCD and meet in its L2.
//
// Problems whose rows 16-byte loads cannot take (K % 16 != 0, or A not 16-byte aligned)
// are run on zero-padded copies of their rows (K rounded up to 16) in a buffer the plan
// owns, refreshed by one pad launch before the HERK launch on every run (captured in the
// same graph), so the HERK kernel is built with the 16-byte load path only.  If that
// buffer cannot be allocated, the plan runs the kernel that holds both load paths.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "ftdx.h"
#include "ftdx_host.h"
#include "ftdx_internal.h"
#include "herk_fp8.h"
#include "stages.h"

struct ftdx_grouped_plan {
    int device = 0;
    int cfg = 0;
    int out_fp16 = 0;
    int num_wg = 0;
    int count = 0;
    bool all_vec = false;  // every problem takes 16-byte loads
    int64_t items = 0;
    double max_load = 0, mean_load = 0;
    ftdx::GroupProblemDev* d_probs = nullptr;
    ftdx::GroupItemDev* d_items = nullptr;
    int* d_wg_begin = nullptr;
    uint8_t* d_pad = nullptr;             // padded rows of the problems that need them
    ftdx::PadRowsJob* d_pad_jobs = nullptr;
    ftdx::PadBlock* d_pad_blocks = nullptr;
    int pad_blocks = 0;
    bool use_graph = false;
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace ftdx {

namespace {

int stage_samples(int cfg) { return cfg == TILE_128x128_K128 ? 128 : 64; }

}  // namespace

int grouped_build(ftdx_grouped_plan* P, const ftdx_herk_problem* problems, int count, int cfg,
                  int out_fp16, int cu_count) {
    std::vector<int> Ns(count), Ks(count);
    for (int p = 0; p < count; ++p) {
        Ns[p] = problems[p].N;
        Ks[p] = problems[p].K;
    }
    GroupSchedule sch;
    const int rc = build_group_schedule(Ns.data(), Ks.data(), count, tile_config_rows(cfg), tile_config_cols(cfg),
                                        stage_samples(cfg), tile_config_wg_per_cu(cfg), cu_count, sch);
    if (rc == -2) return fail(FTDX_ERR_INVALID_ARG, "grouped: too many tiles");
    if (rc != 0) return fail(FTDX_ERR_INVALID_ARG, "grouped: invalid problem list");
    std::vector<GroupProblemDev> probs(count);
    bool all_vec = true;
    size_t pad_bytes = 0;
    for (int p = 0; p < count; ++p) {
        const ftdx_herk_problem& q = problems[p];
        GroupProblemDev& d = probs[p];
        d.A = static_cast<const uint8_t*>(q.A);
        d.C = q.C;
        d.N = q.N;
        d.K = q.K;
        d.alpha = q.alpha;
        d.beta = q.beta;
        // (N = 0 or K = 0: no loads at all)
        d.vec = q.N <= 0 || q.K <= 0 || ((q.K % 16 == 0) && ((reinterpret_cast<uintptr_t>(q.A) & 15u) == 0));
        d.tiles = sch.tiles[p];
        all_vec = all_vec && d.vec;
        if (!d.vec) pad_bytes += (size_t)q.N * precast_row_samples(q.K);
    }
    hipError_t e;
    if (pad_bytes > 0 && hipMalloc(&P->d_pad, pad_bytes) == hipSuccess) {
        // every problem that needs it reads its padded copy; the plan then takes 16-byte loads only
        std::vector<PadRowsJob> jobs;
        std::vector<PadBlock> blocks;
        size_t off = 0;
        for (int p = 0; p < count; ++p) {
            GroupProblemDev& d = probs[p];
            if (d.vec) continue;
            const int Kq = precast_row_samples(d.K);
            const int64_t segs = (int64_t)d.N * (Kq / 16);
            for (int64_t g = 0; g < segs; g += kPadBlockSegments) blocks.push_back({(int)jobs.size(), g});
            jobs.push_back({d.A, P->d_pad + off, (int64_t)d.N, d.K, Kq});
            d.A = P->d_pad + off;
            d.K = Kq;
            d.vec = 1;
            off += (size_t)d.N * Kq;
        }
        if ((e = hipMalloc(&P->d_pad_jobs, sizeof(PadRowsJob) * jobs.size())) != hipSuccess)
            return hip_fail(e, "grouped plan: hipMalloc");
        if ((e = hipMemcpy(P->d_pad_jobs, jobs.data(), sizeof(PadRowsJob) * jobs.size(),
                           hipMemcpyHostToDevice)) != hipSuccess)
            return hip_fail(e, "grouped plan: upload");
        if (blocks.size() > 0x7FFFFFFF) return fail(FTDX_ERR_INVALID_ARG, "grouped plan: too many rows to pad");
        if ((e = hipMalloc(&P->d_pad_blocks, sizeof(PadBlock) * blocks.size())) != hipSuccess)
            return hip_fail(e, "grouped plan: hipMalloc");
        if ((e = hipMemcpy(P->d_pad_blocks, blocks.data(), sizeof(PadBlock) * blocks.size(),
                           hipMemcpyHostToDevice)) != hipSuccess)
            return hip_fail(e, "grouped plan: upload");
        P->pad_blocks = (int)blocks.size();
        all_vec = true;
    } else if (pad_bytes > 0) {
        (void)hipGetLastError();  // out of memory for the copies: both load paths instead
    }
    static_assert(sizeof(GroupItemDev) == sizeof(ScheduleItem), "schedule items are uploaded as they are");
    const std::vector<ScheduleItem>& flat = sch.items;
    const std::vector<int>& begin = sch.begin;
    const int G = sch.num_wg;
    P->items = (int64_t)flat.size();
    P->max_load = sch.max_load;
    P->mean_load = sch.mean_load;

    if ((e = hipMalloc(&P->d_probs, sizeof(GroupProblemDev) * std::max(count, 1))) != hipSuccess)
        return hip_fail(e, "grouped plan: hipMalloc");
    if ((e = hipMalloc(&P->d_items, sizeof(GroupItemDev) * std::max<size_t>(flat.size(), 1))) !=
        hipSuccess)
        return hip_fail(e, "grouped plan: hipMalloc");
    if ((e = hipMalloc(&P->d_wg_begin, sizeof(int) * (G + 1))) != hipSuccess)
        return hip_fail(e, "grouped plan: hipMalloc");
    if ((e = hipMemcpy(P->d_probs, probs.data(), sizeof(GroupProblemDev) * count,
                       hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "grouped plan: upload");
    if (!flat.empty() &&
        (e = hipMemcpy(P->d_items, flat.data(), sizeof(GroupItemDev) * flat.size(),
                       hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "grouped plan: upload");
    if ((e = hipMemcpy(P->d_wg_begin, begin.data(), sizeof(int) * (G + 1),
                       hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "grouped plan: upload");
    P->num_wg = G;
    P->all_vec = all_vec;
    P->cfg = cfg;
    P->out_fp16 = out_fp16;
    P->count = count;
    return FTDX_OK;
}

hipError_t grouped_launch(ftdx_grouped_plan* P, hipStream_t s) {
    if (P->pad_blocks > 0) {
        hipError_t e = launch_pad_qc_rows_multi(P->d_pad_jobs, P->d_pad_blocks, P->pad_blocks, s);
        if (e != hipSuccess) return e;
    }
    return launch_herk_int4_grouped(P->d_probs, P->d_items, P->d_wg_begin, P->num_wg, P->cfg,
                                    P->out_fp16 != 0, P->all_vec, s);
}

void grouped_release(ftdx_grouped_plan* P) {
    if (!P) return;
    if (P->exec) (void)hipGraphExecDestroy(P->exec);
    if (P->cap_stream) (void)hipStreamDestroy(P->cap_stream);
    if (P->d_probs) (void)hipFree(P->d_probs);
    if (P->d_items) (void)hipFree(P->d_items);
    if (P->d_wg_begin) (void)hipFree(P->d_wg_begin);
    if (P->d_pad) (void)hipFree(P->d_pad);
    if (P->d_pad_jobs) (void)hipFree(P->d_pad_jobs);
    if (P->d_pad_blocks) (void)hipFree(P->d_pad_blocks);
    P->d_pad = nullptr;
    P->d_pad_jobs = nullptr;
    P->d_pad_blocks = nullptr;
    P->pad_blocks = 0;
    P->exec = nullptr;
    P->cap_stream = nullptr;
    P->d_probs = nullptr;
    P->d_items = nullptr;
    P->d_wg_begin = nullptr;
}

int grouped_capture(ftdx_grouped_plan* P) {
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&P->cap_stream, hipStreamNonBlocking)) != hipSuccess)
        return hip_fail(e, "grouped plan: hipStreamCreate");
    if ((e = hipStreamBeginCapture(P->cap_stream, hipStreamCaptureModeRelaxed)) != hipSuccess)
        return hip_fail(e, "grouped plan: hipStreamBeginCapture");
    hipError_t le = grouped_launch(P, P->cap_stream);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(P->cap_stream, &graph);
    if (le != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return hip_fail(le, "grouped plan: captured launch");
    }
    if (e != hipSuccess) return hip_fail(e, "grouped plan: hipStreamEndCapture");
    e = hipGraphInstantiate(&P->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_fail(e, "grouped plan: hipGraphInstantiate");
    P->use_graph = true;
    return FTDX_OK;
}

// Times each tile configuration's plan on the problems' own buffers (outputs are
// overwritten: only used when every beta is 0) and returns the fastest.
int grouped_autotune(const ftdx_herk_problem* problems, int count, int out16, int cus, int& best,
                     float& best_ms) {
    hipEvent_t t0, t1;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    best_ms = 1e30f;
    for (int cfg = 0; cfg < NUM_TILE_CONFIGS; ++cfg) {
        ftdx_grouped_plan P;
        int rc = grouped_build(&P, problems, count, cfg, out16, cus);
        if (rc != FTDX_OK) {
            grouped_release(&P);
            continue;
        }
        float ms = 1e30f;
        if (grouped_launch(&P, nullptr) == hipSuccess) {
            (void)hipEventRecord(t0, nullptr);
            const int reps = 2;
            for (int r = 0; r < reps; ++r) (void)grouped_launch(&P, nullptr);
            (void)hipEventRecord(t1, nullptr);
            if (hipEventSynchronize(t1) == hipSuccess) {
                (void)hipEventElapsedTime(&ms, t0, t1);
                ms /= reps;
            }
        }
        (void)hipGetLastError();
        grouped_release(&P);
        if (ms < best_ms) {
            best_ms = ms;
            best = cfg;
        }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (best_ms >= 1e30f) return fail(FTDX_ERR_HIP, "grouped autotune: no configuration ran");
    return FTDX_OK;
}

}  // namespace ftdx

using ftdx::fail;
using ftdx::hip_fail;

extern "C" {

int ftdx_herk_grouped_plan_create(ftdx_context* ctx, const ftdx_herk_problem* problems, int count,
                                  int output, int tile_config, int use_graph,
                                  ftdx_grouped_plan** plan) {
    if (!ctx || !plan) return fail(FTDX_ERR_INVALID_ARG, "grouped plan: NULL context or plan");
    *plan = nullptr;
    if (count < 0 || (count > 0 && !problems))
        return fail(FTDX_ERR_INVALID_ARG, "grouped plan: bad problem list");
    if (output != FTDX_OUTPUT_FP32 && output != FTDX_OUTPUT_FP16)
        return fail(FTDX_ERR_INVALID_ARG, "grouped plan: unknown output precision");
    if (tile_config < -1 || tile_config >= ftdx::NUM_TILE_CONFIGS)
        return fail(FTDX_ERR_INVALID_ARG, "grouped plan: unknown tile configuration");
    for (int p = 0; p < count; ++p) {
        const ftdx_herk_problem& q = problems[p];
        if (q.N < 0 || q.K < 0) return fail(FTDX_ERR_INVALID_ARG, "grouped plan: negative size");
        if (q.N > 0 && ((!q.A && q.K > 0) || !q.C))
            return fail(FTDX_ERR_INVALID_ARG, "grouped plan: NULL data pointer");
        if (q.C && reinterpret_cast<uintptr_t>(q.C) % (output == FTDX_OUTPUT_FP32 ? 8u : 4u))
            return fail(FTDX_ERR_INVALID_ARG, "grouped plan: C must be aligned to one element");
    }
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "grouped plan: hipGetDevice");
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
        return hip_fail(e, "grouped plan: device attribute");
    const int out16 = output == FTDX_OUTPUT_FP16 ? 1 : 0;
    int cfg = tile_config >= 0 ? tile_config : ftdx::TILE_128x128_K64;
    bool all_beta0 = true;
    for (int p = 0; p < count; ++p) all_beta0 = all_beta0 && problems[p].beta == 0.0f;
    if (tile_config < 0 && all_beta0 && count > 0) {
        float best_ms = 0.0f;
        int rc = ftdx::grouped_autotune(problems, count, out16, cus, cfg, best_ms);
        if (rc != FTDX_OK) return rc;
    }
    auto* P = new ftdx_grouped_plan();
    P->device = dev;
    int rc = ftdx::grouped_build(P, problems, count, cfg, out16, cus);
    if (rc == FTDX_OK && use_graph && P->items > 0) rc = ftdx::grouped_capture(P);
    if (rc != FTDX_OK) {
        ftdx::grouped_release(P);
        delete P;
        return rc;
    }
    *plan = P;
    return FTDX_OK;
}

int ftdx_herk_grouped_plan_run(ftdx_grouped_plan* plan, void* stream) {
    if (!plan) return fail(FTDX_ERR_INVALID_ARG, "grouped run: NULL plan");
    if (plan->items == 0) return FTDX_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipError_t e = plan->use_graph ? hipGraphLaunch(plan->exec, s) : ftdx::grouped_launch(plan, s);
    if (e != hipSuccess) return hip_fail(e, "grouped run: launch");
    return FTDX_OK;
}

int ftdx_herk_grouped_plan_info(const ftdx_grouped_plan* plan, int* num_workgroups,
                                int64_t* num_items, int* tile_config, double* imbalance) {
    if (!plan) return fail(FTDX_ERR_INVALID_ARG, "grouped info: NULL plan");
    if (num_workgroups) *num_workgroups = plan->num_wg;
    if (num_items) *num_items = plan->items;
    if (tile_config) *tile_config = plan->cfg;
    if (imbalance) *imbalance = plan->mean_load > 0 ? plan->max_load / plan->mean_load : 1.0;
    return FTDX_OK;
}

int ftdx_herk_grouped_plan_destroy(ftdx_grouped_plan* plan) {
    if (!plan) return FTDX_OK;
    ftdx::grouped_release(plan);
    delete plan;
    return FTDX_OK;
}

}  // extern "C"
