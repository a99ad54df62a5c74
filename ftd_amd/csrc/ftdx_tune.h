// ftdx_tune.h -- launch-configuration autotuner and its cache (internal).
//
// Replaces the reference's strategy cache + kernel tune cache
// (cutlass_interface/strategy_cache.hpp:127-800, tune_cache.hpp:215-918):
// per problem key, time every compiled tile configuration and workgroup order
// with HIP events, keep the fastest, and persist it in a tab-separated file
// whose header carries a build fingerprint (entries of another build are ignored).
#pragma once
#include <map>
#include <string>

namespace ftdx {

struct TuneKey {
    int N, K, batch, out_fp16, vec;
    bool operator<(const TuneKey& o) const;
};

struct TuneEntry {
    int cfg = 0;
    int xcd_remap = 1;
    float ms = 0.0f;
};

class TuneCache {
public:
    bool find(const TuneKey& k, TuneEntry& e) const;
    void put(const TuneKey& k, const TuneEntry& e);
    // Loads entries of this library's fingerprint; malformed or foreign lines are
    // skipped and counted.  Returns the number of entries read (or -1: no file).
    int load(const std::string& path, int* skipped = nullptr);
    bool save(const std::string& path) const;
    size_t size() const { return map_.size(); }
    static std::string fingerprint();

private:
    std::map<TuneKey, TuneEntry> map_;
};

}  // namespace ftdx
