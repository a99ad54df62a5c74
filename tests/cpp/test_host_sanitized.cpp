// Host-side logic of libftdx.so under AddressSanitizer + UndefinedBehaviorSanitizer
// (built with g++, no GPU): the tuning-cache parser and writer, the grouped launch's
// schedule builder and the threaded host copy, driven with malformed cache files,
// concurrent writers and extreme problem lists.  The reference builds its own sanitizer
// configuration (CMakeLists.txt:273-274, -fsanitize=address,undefined).
// Usage: test_host_sanitized <scratch directory>; prints "ALL PASSED" on success.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <set>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "ftdx_host.h"
#include "ftdx_tune.h"
#include "herk_fp8.h"

static int g_fail = 0;
#define CHECK(c)                                                              \
    do {                                                                      \
        if (!(c)) {                                                           \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);          \
            ++g_fail;                                                         \
        }                                                                     \
    } while (0)

static void write_file(const std::string& path, const std::string& text) {
    std::ofstream o(path, std::ios::binary | std::ios::trunc);
    o << text;
}

static std::string header() { return "# ftdx-tune " + ftdx::TuneCache::fingerprint() + "\n"; }

static void test_tune_cache(const std::string& dir) {
    using ftdx::TuneCache;
    using ftdx::TuneEntry;
    using ftdx::TuneKey;
    const std::string p = dir + "/tune.tsv";
    // round trip
    {
        TuneCache c;
        for (int i = 0; i < 50; ++i) {
            TuneEntry e;
            e.cfg = i % ftdx::NUM_TILE_CONFIGS;
            e.xcd_remap = i & 1;
            e.ms = 0.5f * i;
            c.put(TuneKey{64 * (i + 1), 32 * i, i % 7, i & 1, 1}, e);
        }
        CHECK(c.save(p));
        TuneCache d;
        int sk = -1;
        CHECK(d.load(p, &sk) == 50);
        CHECK(sk == 0);
        TuneEntry e;
        CHECK(d.find(TuneKey{64 * 8, 32 * 7, 0, 1, 1}, e) && e.cfg == 7 % ftdx::NUM_TILE_CONFIGS);
    }
    // missing file
    {
        TuneCache c;
        CHECK(c.load(dir + "/does-not-exist.tsv") == -1);
    }
    // malformed content: each case -> (entries read, lines skipped)
    const std::string good = "herk\t128\t64\t1\t0\t1\t3\t1\t0.25\n";
    std::string longline(1 << 20, '7');
    std::string binary;
    for (int b = 0; b < 256; ++b) binary.push_back((char)b);
    struct Case {
        std::string text;
        int n, skipped;
    };
    std::vector<Case> cases = {
        {"", 0, 0},
        {header(), 0, 0},
        {header() + good, 1, 0},
        {header() + good.substr(0, good.size() - 1), 1, 0},                 // no final newline
        {"# ftdx-tune another-build\n" + good, 0, 1},                       // foreign fingerprint
        {good, 0, 1},                                                       // no header
        {header() + "herk\t128\t64\n", 0, 1},                               // too few fields
        {header() + "herk\t128\t64\t1\t0\t1\t3\t1\t0.25\textra\n", 0, 1},   // too many
        {header() + "gemm\t128\t64\t1\t0\t1\t3\t1\t0.25\n", 0, 1},          // unknown kind
        {header() + "herk\t-5\t64\t1\t0\t1\t3\t1\t0.25\n", 0, 1},           // negative N
        {header() + "herk\t128\t64\t1\t0\t1\t99\t1\t0.25\n", 0, 1},         // unknown config
        {header() + "herk\t99999999999999999999\t64\t1\t0\t1\t3\t1\t0.25\n", 0, 1},  // overflow
        {header() + "herk\tabc\t64\t1\t0\t1\t3\t1\t0.25\n", 0, 1},          // not a number
        {header() + "herk\t128\t64\t1\t0\t1\t3\t1\tnan\n", 0, 1},           // NaN time rejected
        {header() + longline + "\n" + good, 1, 1},
        {header() + binary + "\n" + good, 1, 2},                          // byte 10 splits the line
        {header() + std::string("herk\t1\t2\0\t3", 12) + "\n" + good, 1, 1},  // embedded NUL
        {header() + good + header() + "# ftdx-tune x\n" + good, 1, 1},      // header switch
        {header() + "herk 128 64 1 0 1 3 1 0.25\r\n", 1, 0},                // CRLF and spaces accepted
    };
    for (size_t i = 0; i < cases.size(); ++i) {
        write_file(p, cases[i].text);
        TuneCache c;
        int sk = -1;
        const int n = c.load(p, &sk);
        if (n != cases[i].n || sk != cases[i].skipped)
            std::printf("  case %zu: read %d skipped %d (expected %d, %d)\n", i, n, sk, cases[i].n,
                        cases[i].skipped);
        CHECK(n == cases[i].n && sk == cases[i].skipped);
    }
    // random garbage never crashes the parser
    std::mt19937_64 rng(7);
    for (int t = 0; t < 200; ++t) {
        std::string txt = header();
        const int len = (int)(rng() % 4096);
        for (int i = 0; i < len; ++i) {
            const char alphabet[] = "herk\t\n 0123456789-.e#x";
            txt.push_back((rng() % 8 == 0) ? (char)(rng() & 0xFF) : alphabet[rng() % (sizeof(alphabet) - 1)]);
        }
        write_file(p, txt);
        TuneCache c;
        int sk = 0;
        CHECK(c.load(p, &sk) >= 0);
    }
    // concurrent writers sharing one path: every save replaces the file whole
    {
        const std::string q = dir + "/shared.tsv";
        std::vector<std::thread> th;
        for (int w = 0; w < 4; ++w)
            th.emplace_back([&, w] {
                TuneCache c;
                for (int i = 0; i < 64; ++i) {
                    TuneEntry e;
                    e.cfg = (w + i) % ftdx::NUM_TILE_CONFIGS;
                    c.put(TuneKey{w * 1000 + i, 64, 1, 0, 1}, e);
                    (void)c.save(q);
                }
            });
        for (auto& t : th) t.join();
        TuneCache d;
        int sk = -1;
        const int n = d.load(q, &sk);
        CHECK(n == 64 && sk == 0);  // one writer's complete final cache
    }
}

static void test_schedule() {
    using ftdx::GroupSchedule;
    std::mt19937_64 rng(11);
    struct Shape {
        int TM, TN, kc, wg;
    };
    const Shape shapes[] = {{128, 128, 64, 2}, {128, 128, 128, 1}, {64, 64, 64, 4}, {256, 128, 64, 1}};
    auto verify = [&](const std::vector<int>& N, const std::vector<int>& K, const Shape& s, int cus) {
        GroupSchedule g;
        const int rc = ftdx::build_group_schedule(N.data(), K.data(), (int)N.size(), s.TM, s.TN, s.kc, s.wg, cus, g);
        CHECK(rc == 0);
        if (rc != 0) return;
        CHECK(g.num_wg >= 1 && (int)g.begin.size() == g.num_wg + 1);
        CHECK(g.begin.front() == 0 && g.begin.back() == (int)g.items.size());
        for (int w = 0; w < g.num_wg; ++w) CHECK(g.begin[w] <= g.begin[w + 1]);
        int64_t want = 0;
        for (size_t p = 0; p < N.size(); ++p) {
            CHECK(g.tiles[p] == ftdx::host_tri_tiles(N[p], s.TM, s.TN));
            want += g.tiles[p];
        }
        CHECK((int64_t)g.items.size() == want);
        std::set<std::pair<int, int>> seen;
        for (const auto& it : g.items) {
            CHECK(it.prob >= 0 && it.prob < (int)N.size());
            if (it.prob < 0 || it.prob >= (int)N.size()) continue;
            CHECK(it.tile >= 0 && it.tile < g.tiles[it.prob]);
            CHECK(seen.insert({it.prob, it.tile}).second);  // each tile exactly once
        }
        CHECK(g.max_load >= g.mean_load);
    };
    for (const Shape& s : shapes) {
        verify({}, {}, s, 256);                        // no problems
        verify({1}, {0}, s, 256);                      // one element, K = 0
        verify({1, 2, 3}, {1, 1, 1}, s, 1);            // one CU
        verify({4096}, {2048}, s, 256);                // the headline problem
        verify({20000}, {64}, s, 256);                 // one large N
        for (int t = 0; t < 20; ++t) {
            const int count = 1 + (int)(rng() % 600);
            std::vector<int> N(count), K(count);
            for (int p = 0; p < count; ++p) {
                N[p] = (int)(rng() % 4200);
                K[p] = (int)(rng() % 5000);
            }
            verify(N, K, s, 1 + (int)(rng() % 512));
        }
    }
    GroupSchedule g;
    const int bn[] = {-1}, bk[] = {64};
    CHECK(ftdx::build_group_schedule(bn, bk, 1, 256, 128, 64, 1, 256, g) == -1);   // negative N
    CHECK(ftdx::build_group_schedule(bk, bn, 1, 256, 128, 64, 1, 256, g) == -1);   // negative K
    CHECK(ftdx::build_group_schedule(bk, bk, 1, 100, 64, 64, 1, 256, g) == -1);    // TM not a multiple of TN
    CHECK(ftdx::build_group_schedule(nullptr, nullptr, 3, 256, 128, 64, 1, 256, g) == -1);
    const int huge[] = {1 << 30};
    CHECK(ftdx::build_group_schedule(huge, bk, 1, 64, 64, 64, 1, 256, g) == -2);  // tiles overflow int
}

static void test_host_copy() {
    std::mt19937_64 rng(3);
    const size_t sizes[] = {0, 1, 7, 4095, 4096, (4u << 20) - 1, 4u << 20, (4u << 20) + 1, (33u << 20) + 7};
    for (size_t n : sizes)
        for (unsigned threads : {1u, 2u, 8u, 64u})
            for (size_t so : {0, 3})
                for (size_t dof : {0, 5}) {
                    std::vector<uint8_t> src(n + so), dst(n + dof + 16, 0xAB);
                    for (auto& b : src) b = (uint8_t)rng();
                    ftdx::host_copy(dst.data() + dof, src.data() + so, n, threads);
                    CHECK(n == 0 || std::memcmp(dst.data() + dof, src.data() + so, n) == 0);
                    bool guard = true;
                    for (size_t i = 0; i < dof; ++i) guard = guard && dst[i] == 0xAB;
                    for (size_t i = n + dof; i < dst.size(); ++i) guard = guard && dst[i] == 0xAB;
                    CHECK(guard);
                }
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    test_tune_cache(dir);
    test_schedule();
    test_host_copy();
    if (g_fail) {
        std::printf("%d FAILED\n", g_fail);
        return 1;
    }
    std::printf("ALL PASSED\n");
    return 0;
}
