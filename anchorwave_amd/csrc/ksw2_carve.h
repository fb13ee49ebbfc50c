// ksw2_carve.h -- first-fit carving of one address range (host bookkeeping
// only, no device calls): the engine's device pool keeps its free list here.
//
// A block is allocated whole and may give back parts of itself early
// (release: one pair's direction bytes once its walk has finished); a later
// block may then be carved from those parts, inside the first block's extent.
// Freeing a block gives back only the parts it still owns.  Offsets and
// sizes are whole units of kAlign bytes.  Not thread-safe: the owner locks.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <iterator>
#include <map>

namespace kswhip {

class RangeCarver {
public:
    static constexpr uint64_t kAlign = 256;

    void reset(uint64_t bytes)
    {
        free_.clear(), used_.clear(), blocks_.clear();
        if (bytes) free_[0] = bytes / kAlign * kAlign;
    }
    // First fit of n bytes (rounded up to kAlign); false when no free run is large enough.
    bool alloc(uint64_t n, uint64_t *off)
    {
        n = up(std::max<uint64_t>(n, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            const uint64_t rest = it->second - n, at = it->first + n;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used_[*off] = Part{n, *off};
            blocks_[*off] = n;
            return true;
        }
        return false;
    }
    // Frees what block `off` still owns; unknown offsets are ignored.
    void free(uint64_t off)
    {
        auto b = blocks_.find(off);
        if (b == blocks_.end()) return;
        const uint64_t end = off + b->second;
        blocks_.erase(b);
        for (auto u = used_.lower_bound(off); u != used_.end() && u->first < end;) {
            if (u->second.owner != off) {  // carved later from a part this block gave back
                ++u;
                continue;
            }
            give_back(u->first, u->second.len);
            u = used_.erase(u);
        }
    }
    // Gives back [off, off + n), shrunk to whole kAlign units, when it lies
    // inside one part a block still owns; otherwise does nothing.  Returns the
    // bytes given back.
    uint64_t release(uint64_t off, uint64_t n)
    {
        const uint64_t a = up(off), e = (off + n) / kAlign * kAlign;
        if (e <= a) return 0;
        auto u = used_.upper_bound(a);
        if (u == used_.begin()) return 0;
        --u;
        const uint64_t s = u->first, ue = u->first + u->second.len, owner = u->second.owner;
        if (e > ue) return 0;
        used_.erase(u);
        if (a > s) used_[s] = Part{a - s, owner};
        if (ue > e) used_[e] = Part{ue - e, owner};
        give_back(a, e - a);
        return e - a;
    }
    uint64_t largest_free() const
    {
        uint64_t m = 0;
        for (const auto &f : free_) m = std::max(m, f.second);
        return m;
    }
    // Views for checks: free runs (offset -> bytes) and owned parts.
    struct Part {
        uint64_t len, owner;  // bytes, and the offset of the block it belongs to
    };
    const std::map<uint64_t, uint64_t> &free_runs() const { return free_; }
    const std::map<uint64_t, Part> &parts() const { return used_; }

private:
    static uint64_t up(uint64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }
    // coalescing insert into the free list
    void give_back(uint64_t start, uint64_t len)
    {
        auto next = free_.lower_bound(start);
        if (next != free_.end() && next->first == start + len) len += next->second, next = free_.erase(next);
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == start) start = prev->first, len += prev->second, free_.erase(prev);
        }
        free_[start] = len;
    }
    std::map<uint64_t, uint64_t> free_;
    std::map<uint64_t, Part> used_;        // parts blocks still own, by offset
    std::map<uint64_t, uint64_t> blocks_;  // blocks as allocated: offset -> bytes
};

}  // namespace kswhip
