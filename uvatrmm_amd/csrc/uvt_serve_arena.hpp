// uvt_serve_arena.hpp -- first-fit allocator of byte offsets in [0, size),
// 256-byte granules, with coalescing frees.  Host code only (the serving
// loop maps offsets onto one device allocation); kept free of HIP so the
// admission arithmetic can be exercised on a machine without a GPU.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <iterator>
#include <map>

namespace uvt {

class OffsetArena {
public:
    static size_t granule(size_t x) { return (x + 255) & ~(size_t)255; }

    explicit OffsetArena(size_t bytes = 0) { reset(bytes); }
    void reset(size_t bytes)
    {
        size_ = granule(bytes);
        used_ = 0;
        free_.clear();
        if (size_) free_[0] = size_;
    }
    // offset of a block of `bytes` (first fit), or SIZE_MAX when no free
    // block is large enough
    size_t alloc(size_t bytes)
    {
        bytes = granule(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                const size_t off = it->first, rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                used_ += bytes;
                return off;
            }
        }
        return SIZE_MAX;
    }
    // returns a block from alloc(bytes); merges it with free neighbours
    void free(size_t off, size_t bytes)
    {
        bytes = granule(bytes);
        used_ -= bytes;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + bytes == next->first) {
            bytes += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += bytes;
                return;
            }
        }
        free_[off] = bytes;
    }
    size_t used() const { return used_; }
    size_t size() const { return size_; }
    size_t free_blocks() const { return free_.size(); }
    size_t largest_free() const
    {
        size_t b = 0;
        for (const auto& f : free_) b = f.second > b ? f.second : b;
        return b;
    }

private:
    size_t size_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;  // offset -> bytes, disjoint, non-adjacent
};

}  // namespace uvt
