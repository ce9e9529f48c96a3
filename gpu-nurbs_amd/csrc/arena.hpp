// arena.hpp -- first-fit range allocator over one device allocation (host-only bookkeeping).
//
// Used by the inversion server to hand out request result buffers inside its HBM budget.
// Offsets and sizes are multiples of 256 bytes; freed ranges coalesce with their neighbours.
#pragma once
#include <cstddef>
#include <iterator>
#include <map>

namespace gn {

class Arena {
public:
    static constexpr size_t kAlign = 256;
    static size_t round(size_t b) { return (b + kAlign - 1) & ~(kAlign - 1); }

    explicit Arena(size_t capacity = 0) { reset(capacity); }

    void reset(size_t capacity) {
        cap_ = capacity & ~(kAlign - 1);
        used_ = 0;
        peak_ = 0;
        free_.clear();
        if (cap_) free_[0] = cap_;
    }

    // First fit; false when no free range is large enough.
    bool alloc(size_t bytes, size_t *off) {
        const size_t need = round(bytes ? bytes : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < need) continue;
            *off = it->first;
            const size_t rest = it->second - need, at = it->first + need;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used_ += need;
            if (used_ > peak_) peak_ = used_;
            return true;
        }
        return false;
    }

    // Returns a range obtained from alloc (same byte count).
    void release(size_t off, size_t bytes) {
        const size_t sz = round(bytes ? bytes : 1);
        used_ -= sz;
        auto next = free_.lower_bound(off);
        size_t start = off, len = sz;
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                start = prev->first;
                len += prev->second;
                free_.erase(prev);
            }
        }
        if (next != free_.end() && off + sz == next->first) {
            len += next->second;
            free_.erase(next);
        }
        free_[start] = len;
    }

    size_t capacity() const { return cap_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }
    size_t ranges() const { return free_.size(); }
    size_t largest_free() const {
        size_t m = 0;
        for (const auto &f : free_) m = f.second > m ? f.second : m;
        return m;
    }

private:
    size_t cap_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;  // offset -> length
};

}  // namespace gn
