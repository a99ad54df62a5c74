// ftdx_tune.cpp -- tuning cache file I/O (format: see ftdx_tune.h).
//
//   # ftdx-tune <fingerprint>
//   herk <N> <K> <batch> <out_fp16> <vec> <cfg> <xcd_remap> <ms>     (tab-separated)
#include "ftdx_tune.h"

#include <unistd.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <thread>
#include <tuple>
#include <vector>

#include "herk_fp8.h"

namespace ftdx {

bool TuneKey::operator<(const TuneKey& o) const {
    return std::tie(N, K, batch, out_fp16, vec) < std::tie(o.N, o.K, o.batch, o.out_fp16, o.vec);
}

bool TuneCache::find(const TuneKey& k, TuneEntry& e) const {
    auto it = map_.find(k);
    if (it == map_.end()) return false;
    e = it->second;
    return true;
}

void TuneCache::put(const TuneKey& k, const TuneEntry& e) { map_[k] = e; }

std::string TuneCache::fingerprint() {
    std::string f = "gfx950-abi1";
    for (int flag : {0, TILE_BF6})
        for (int c = 0; c < NUM_TILE_CONFIGS; ++c) {
            f += '-';
            f += tile_config_name(c | flag);
        }
    return f;
}

int TuneCache::load(const std::string& path, int* skipped) {
    std::ifstream in(path);
    if (!in) return -1;
    std::string line;
    int n = 0, bad = 0;
    bool ours = false;
    const std::string fp = fingerprint();
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        if (line[0] == '#') {
            ours = (line == "# ftdx-tune " + fp);
            continue;
        }
        if (!ours) {
            ++bad;
            continue;
        }
        std::istringstream ss(line);
        std::string kind;
        TuneKey k{};
        TuneEntry e{};
        if (!(ss >> kind >> k.N >> k.K >> k.batch >> k.out_fp16 >> k.vec >> e.cfg >> e.xcd_remap >>
              e.ms) ||
            kind != "herk" || !valid_tile_config(e.cfg) || k.N < 0 || k.K < 0 ||
            k.batch < 0) {
            ++bad;
            continue;
        }
        std::string extra;
        if (ss >> extra) {
            ++bad;
            continue;
        }
        map_[k] = e;
        ++n;
    }
    if (skipped) *skipped = bad;
    return n;
}

bool TuneCache::save(const std::string& path) const {
    // A temporary name of this process and call, renamed over `path` once complete:
    // processes (one per GPU) and threads sharing one cache path never write into
    // each other's file, and readers see either the old or the new cache whole.
    static std::atomic<unsigned long> serial{0};
    std::ostringstream tn;
    tn << path << ".tmp." << getpid() << '.' << std::hash<std::thread::id>()(std::this_thread::get_id()) << '.'
       << serial.fetch_add(1);
    const std::string tmp = tn.str();
    {
        std::ofstream out(tmp, std::ios::trunc);
        if (!out) return false;
        out << "# ftdx-tune " << fingerprint() << "\n";
        for (const auto& kv : map_) {
            const TuneKey& k = kv.first;
            const TuneEntry& e = kv.second;
            out << "herk\t" << k.N << '\t' << k.K << '\t' << k.batch << '\t' << k.out_fp16 << '\t'
                << k.vec << '\t' << e.cfg << '\t' << e.xcd_remap << '\t' << e.ms << "\n";
        }
        if (!out) {
            out.close();
            std::remove(tmp.c_str());
            return false;
        }
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::remove(tmp.c_str());
        return false;
    }
    return true;
}

}  // namespace ftdx
