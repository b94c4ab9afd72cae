// First-fit device-memory carving shared by the serving loops: offsets into
// one device allocation, handed out and returned by the host.
#pragma once

#include <cstdint>
#include <map>
#include <stdexcept>

namespace hz {
namespace capi {

constexpr uint64_t POOL_ALIGN = 256;

inline uint64_t pool_align(uint64_t x) { return (x + POOL_ALIGN - 1) / POOL_ALIGN * POOL_ALIGN; }

// First-fit allocator over [0, size) with coalescing of free neighbours.
class Pool {
   public:
    explicit Pool(uint64_t size) : size_(size) {
        if (size) free_[0] = size;
    }
    static constexpr uint64_t NONE = ~0ull;

    uint64_t alloc(uint64_t bytes) {
        bytes = pool_align(bytes ? bytes : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                uint64_t off = it->first, len = it->second;
                free_.erase(it);
                if (len > bytes) free_[off + bytes] = len - bytes;
                used_ += bytes;
                return off;
            }
        }
        return NONE;
    }

    void free(uint64_t off, uint64_t bytes) {
        bytes = pool_align(bytes ? bytes : 1);
        if (off + bytes > size_) throw std::invalid_argument("pool: free outside the pool");
        auto next = free_.lower_bound(off);
        if (next != free_.end() && next->first < off + bytes) throw std::invalid_argument("pool: double free");
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second > off) throw std::invalid_argument("pool: double free");
        }
        used_ -= bytes;
        uint64_t start = off, len = bytes;
        if (next != free_.end() && next->first == off + bytes) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == start) {
                start = prev->first;
                len += prev->second;
                free_.erase(prev);
            }
        }
        free_[start] = len;
    }

    uint64_t largest_free() const {
        uint64_t best = 0;
        for (auto& kv : free_) best = kv.second > best ? kv.second : best;
        return best;
    }
    uint64_t used() const { return used_; }
    uint64_t size() const { return size_; }
    size_t fragments() const { return free_.size(); }

   private:
    uint64_t size_;
    uint64_t used_ = 0;
    std::map<uint64_t, uint64_t> free_;
};

}  // namespace capi
}  // namespace hz
