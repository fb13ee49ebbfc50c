// Checks of RangeCarver (anchorwave_amd/csrc/ksw2_carve.h), the engine
// pool's bookkeeping: after every operation of a long random sequence of
// allocations, early part releases and frees, the free runs and the parts
// still owned tile [0, size) exactly, and the owned bytes equal what the
// live blocks hold minus what they gave back.  One fixed case first: a block
// carved from another block's released part survives the first block's free.
#include <stdio.h>

#include <random>
#include <vector>

#include "../../anchorwave_amd/csrc/ksw2_carve.h"

using kswhip::RangeCarver;

static int check_tiling(const RangeCarver &c, uint64_t size, uint64_t expect_owned)
{
    std::map<uint64_t, uint64_t> all;
    uint64_t owned = 0;
    for (const auto &f : c.free_runs()) all[f.first] = f.second;
    for (const auto &u : c.parts()) {
        if (all.count(u.first)) return 1;
        all[u.first] = u.second.len, owned += u.second.len;
    }
    uint64_t at = 0;
    for (const auto &r : all) {
        if (r.first != at) return 2;  // gap or overlap
        at += r.second;
    }
    if (at != size) return 3;
    if (owned != expect_owned) return 4;
    return 0;
}

int main()
{
    {
        RangeCarver c;
        c.reset(1 << 20);
        uint64_t a, b;
        if (!c.alloc(64 * 1024, &a) || a != 0) return printf("fixed: first alloc\n"), 1;
        if (c.release(a + 4096, 32768) != 32768) return printf("fixed: release\n"), 1;
        if (!c.alloc(16384, &b) || b != 4096) return printf("fixed: carve from the released part\n"), 1;
        c.free(a);  // must leave b's bytes owned
        if (check_tiling(c, 1 << 20, 16384)) return printf("fixed: free took a later block's part\n"), 1;
        c.free(b);
        if (check_tiling(c, 1 << 20, 0) || c.largest_free() != (1u << 20)) return printf("fixed: coalesce\n"), 1;
    }
    std::mt19937_64 rng(12345);
    for (int trial = 0; trial < 100; ++trial) {
        const uint64_t size = 1ull << 30;
        RangeCarver c;
        c.reset(size);
        struct Blk {
            uint64_t off, len;
            std::vector<std::pair<uint64_t, uint64_t>> parts;
            std::vector<char> gone;
        };
        std::vector<Blk> live;
        for (int step = 0; step < 2000; ++step) {
            const int op = (int)(rng() % 3);
            if (op == 0 || live.empty()) {
                // a plan-like block: a head, several 256-aligned parts, a tail
                Blk b;
                uint64_t at = 256 * (1 + rng() % 20);
                const int np = 1 + (int)(rng() % 8);
                for (int k = 0; k < np; ++k) {
                    const uint64_t l = 256 * (1 + rng() % 200);
                    b.parts.push_back({at, l});
                    at += l;
                }
                const uint64_t len = at + 1 + rng() % 3000;
                if (!c.alloc(len, &b.off)) continue;
                b.len = (len + 255) / 256 * 256;
                b.gone.assign(np, 0);
                live.push_back(b);
            } else if (op == 1) {
                Blk &b = live[rng() % live.size()];
                const int k = (int)(rng() % b.parts.size());
                if (b.gone[k]) continue;
                if (c.release(b.off + b.parts[k].first, b.parts[k].second) != b.parts[k].second)
                    return printf("trial %d step %d: release refused\n", trial, step), 1;
                b.gone[k] = 1;
            } else {
                const size_t i = rng() % live.size();
                c.free(live[i].off);
                live.erase(live.begin() + (long)i);
            }
            uint64_t expect = 0;
            for (const Blk &b : live) {
                expect += b.len;
                for (size_t k = 0; k < b.parts.size(); ++k)
                    if (b.gone[k]) expect -= b.parts[k].second;
            }
            if (const int e = check_tiling(c, size, expect))
                return printf("trial %d step %d op %d: invariant %d broken\n", trial, step, op, e), 1;
        }
    }
    printf("carve ok\n");
    return 0;
}
