// ftdx_host.cpp -- grouped-launch schedule and threaded host copy (see ftdx_host.h),
// plus the tile configuration names; no HIP API in this file.
#include "ftdx_host.h"

#include <algorithm>
#include <cstring>
#include <functional>
#include <numeric>
#include <queue>
#include <thread>
#include <utility>

#include "herk_fp8.h"

namespace ftdx {

const char* tile_config_name(int cfg) {
    static const char* const names[2][NUM_TILE_CONFIGS] = {
        {"t128x128_k64", "t128x128_k128", "t64x64_k64", "t256x128_k64", "t128x128_k64_p2",
         "t256x128_k64_p2"},
        {"t128x128_k64_bf6", "t128x128_k128_bf6", "t64x64_k64_bf6", "t256x128_k64_bf6",
         "t128x128_k64_p2_bf6", "t256x128_k64_p2_bf6"}};
    if (!valid_tile_config(cfg)) return "unknown";
    return names[(cfg & TILE_BF6) ? 1 : 0][tile_base(cfg)];
}

int64_t host_tri_tiles(int N, int TM, int TN) {
    if (N <= 0 || TM <= 0 || TN <= 0 || TM % TN != 0) return 0;
    const int64_t nbM = (N + TM - 1) / TM, nbN = (N + TN - 1) / TN, r = TM / TN;
    int64_t t = 0;
    for (int64_t I = 0; I < nbM; ++I) t += std::min(nbN, r * (I + 1));
    return t;
}

namespace {

double tile_cost(int kc, int K, bool diag) {
    const double k = (double)((K + kc - 1) / kc) * kc + 192.0;
    return diag ? 0.75 * k : k;
}

}  // namespace

int build_group_schedule(const int* N, const int* K, int count, int TM, int TN, int kc, int wg_per_cu,
                         int cu_count, GroupSchedule& out) {
    out = GroupSchedule();
    if (count < 0 || (count > 0 && (!N || !K)) || TM <= 0 || TN <= 0 || TM % TN != 0 || kc <= 0 ||
        wg_per_cu <= 0)
        return -1;
    const int r = TM / TN;
    struct Item {
        double cost;
        int prob, tile;
    };
    std::vector<Item> items;
    int64_t total = 0;
    for (int p = 0; p < count; ++p) {
        if (N[p] < 0 || K[p] < 0) return -1;
        total += host_tri_tiles(N[p], TM, TN);
        if (total > 0x7FFFFFFF) return -2;
    }
    items.reserve((size_t)total);
    out.tiles.resize(count);
    for (int p = 0; p < count; ++p) {
        int t = 0;
        const int nbM = (N[p] + TM - 1) / TM, nbN = (N[p] + TN - 1) / TN;
        for (int I = 0; I < nbM; ++I) {
            const int cnt = std::min(nbN, r * (I + 1));
            for (int J = 0; J < cnt; ++J, ++t) {
                const bool diag = ((int64_t)J * TN >= (int64_t)I * TM);
                items.push_back({tile_cost(kc, K[p], diag), p, t});
            }
        }
        out.tiles[p] = t;
    }
    std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) { return a.cost > b.cost; });

    int G = (int)std::min<int64_t>((int64_t)std::max(cu_count, 1) * wg_per_cu, (int64_t)items.size());
    if (G < 1) G = 1;
    // rank: workgroups of XCD 0 first, then XCD 1, ...
    std::vector<int> rank(G);
    {
        const int q = G / 8, rem = G % 8;
        for (int g = 0; g < G; ++g) {
            const int x = g % 8, local = g / 8;
            rank[g] = x * q + std::min(x, rem) + local;
        }
    }
    using Bin = std::pair<double, int>;  // (load, rank)
    std::priority_queue<Bin, std::vector<Bin>, std::greater<Bin>> heap;
    std::vector<int> by_rank(G);
    for (int g = 0; g < G; ++g) {
        heap.push({0.0, rank[g]});
        by_rank[rank[g]] = g;
    }
    std::vector<std::vector<ScheduleItem>> lists(G);
    std::vector<double> load(G, 0.0);
    for (const Item& it : items) {
        Bin b = heap.top();
        heap.pop();
        const int g = by_rank[b.second];
        lists[g].push_back({it.prob, it.tile});
        load[g] = b.first + it.cost;
        heap.push({load[g], b.second});
    }
    out.begin.assign(G + 1, 0);
    out.items.reserve(items.size());
    for (int g = 0; g < G; ++g) {
        out.begin[g] = (int)out.items.size();
        out.items.insert(out.items.end(), lists[g].begin(), lists[g].end());
    }
    out.begin[G] = (int)out.items.size();
    out.num_wg = G;
    out.max_load = *std::max_element(load.begin(), load.end());
    out.mean_load = std::accumulate(load.begin(), load.end(), 0.0) / G;
    return 0;
}

void host_copy(void* dst, const void* src, size_t n, unsigned max_threads) {
    constexpr size_t kPiece = (size_t)4 << 20;
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t want = std::min<size_t>({(size_t)std::max(max_threads, 1u), (size_t)(hw ? hw : 1),
                                          (n + kPiece - 1) / kPiece});
    if (want <= 1) {
        if (n) std::memcpy(dst, src, n);
        return;
    }
    const size_t per = ((n + want - 1) / want + 4095) & ~(size_t)4095;
    auto piece = [=](size_t t) {
        const size_t off = t * per;
        if (off < n)
            std::memcpy(static_cast<uint8_t*>(dst) + off, static_cast<const uint8_t*>(src) + off,
                        std::min(per, n - off));
    };
    std::vector<std::thread> workers;
    size_t t = 1;
    try {
        workers.reserve(want - 1);
        for (; t < want; ++t) workers.emplace_back(piece, t);
    } catch (...) {  // no thread available: copy the remaining pieces here
        for (; t < want; ++t) piece(t);
    }
    piece(0);
    for (auto& w : workers) w.join();
}

}  // namespace ftdx
