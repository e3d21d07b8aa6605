// rf_arena.h -- best-fit sub-allocator over one pre-allocated device block.
//
// Host bookkeeping only (offsets into the block), so it is plain C++ and is
// unit-tested on the CPU.  Free ranges are kept coalesced; allocations are
// rounded up to the alignment.  Not thread-safe: the caller serialises.
#pragma once
#include <cstddef>
#include <cstdint>
#include <iterator>
#include <map>

class rf_arena {
public:
    void init(size_t capacity, size_t align = 256) {
        align_ = align ? align : 1;
        cap_ = capacity / align_ * align_;
        free_.clear();
        used_.clear();
        inUse_ = 0;
        peak_ = 0;
        if (cap_) free_[0] = cap_;
    }

    // Smallest free range that fits (lowest offset on ties); false if none.
    bool alloc(size_t bytes, size_t* offset) {
        const size_t need = round(bytes ? bytes : 1);
        auto best = free_.end();
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= need && (best == free_.end() || it->second < best->second)) best = it;
        if (best == free_.end()) return false;
        const size_t off = best->first, rest = best->second - need;
        free_.erase(best);
        if (rest) free_[off + need] = rest;
        used_[off] = need;
        inUse_ += need;
        if (inUse_ > peak_) peak_ = inUse_;
        *offset = off;
        return true;
    }

    // Returns the range of an allocation to the free set; false if offset is
    // not the start of a live allocation.
    bool release(size_t offset) {
        auto u = used_.find(offset);
        if (u == used_.end()) return false;
        size_t off = u->first, len = u->second;
        used_.erase(u);
        inUse_ -= len;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + len == next->first) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += len;
                return true;
            }
        }
        free_[off] = len;
        return true;
    }

    size_t round(size_t b) const { return (b + align_ - 1) / align_ * align_; }
    size_t capacity() const { return cap_; }
    size_t used() const { return inUse_; }
    size_t peak() const { return peak_; }
    size_t free_ranges() const { return free_.size(); }
    size_t largest_free() const {
        size_t m = 0;
        for (auto& f : free_) m = f.second > m ? f.second : m;
        return m;
    }
    size_t live() const { return used_.size(); }

private:
    size_t align_ = 256, cap_ = 0, inUse_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;   // offset -> length, coalesced
    std::map<size_t, size_t> used_;   // offset -> length
};
