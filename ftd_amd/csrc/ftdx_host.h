// ftdx_host.h -- host-only logic of libftdx.so that touches no HIP API (internal):
// the grouped launch's static schedule and the multi-threaded host copy.  Kept apart
// so it builds with plain g++ under the address / undefined-behaviour sanitizers.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace ftdx {

// Lower-triangle tiles of an N x N problem for TM x TN tiles, TM a multiple of TN:
// row block I holds column blocks J with J * TN <= (I + 1) * TM - 1, capped at N / TN.
int64_t host_tri_tiles(int N, int TM, int TN);

// Static schedule of a grouped launch: every lower-triangle tile of every problem is
// one item; items go longest-first (K rounded up to the stage `kc`, plus an epilogue
// share, 0.75x on the diagonal) to the least-loaded of G persistent workgroups, ties
// to workgroups of one XCD first (hardware ids go round-robin over 8 XCDs).
// Workgroup g runs items[begin[g] .. begin[g + 1]).
struct ScheduleItem {
    int prob;
    int tile;
};
struct GroupSchedule {
    int num_wg = 0;
    std::vector<int> tiles;          // per problem
    std::vector<ScheduleItem> items;  // grouped by workgroup
    std::vector<int> begin;          // num_wg + 1 entries
    double max_load = 0.0, mean_load = 0.0;
};
// Returns 0, or -1 for invalid arguments (count < 0, a negative N or K, a bad tile
// shape) or -2 when the items do not fit an int index.
int build_group_schedule(const int* N, const int* K, int count, int TM, int TN, int kc, int wg_per_cu,
                         int cu_count, GroupSchedule& out);

// memcpy of n bytes split over up to max_threads threads in pieces of at least 4 MiB
// (page-aligned piece starts); falls back to the calling thread when no thread starts.
void host_copy(void* dst, const void* src, size_t n, unsigned max_threads = 8);

}  // namespace ftdx
