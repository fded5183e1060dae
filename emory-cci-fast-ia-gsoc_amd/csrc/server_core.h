// server_core.h -- host-only parts of the serving loop: the arena's first-fit
// free list and the admission of queued requests into a slot's set.  No HIP
// here: the same code is built by the host sanitizer tests (ASan/UBSan and
// TSan) with plain host memory standing in for the device block.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <deque>
#include <iterator>
#include <map>
#include <vector>

namespace nsx_srv {

constexpr size_t kAlign = 256;
inline size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over one block of `size` bytes; offsets are
// kAlign-aligned and the free list stays coalesced.  Not thread-safe: the
// server calls it under its mutex.
struct Arena {
    size_t size = 0, used = 0, peak = 0;
    std::map<size_t, size_t> free_blocks;   // offset -> length

    void reset(size_t bytes) {
        size = align_up(bytes);
        used = peak = 0;
        free_blocks.clear();
        if (size) free_blocks.emplace(0, size);
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = align_up(std::max<size_t>(bytes, 1));
        for (auto it = free_blocks.begin(); it != free_blocks.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t o = it->first, rest = it->second - bytes;
            free_blocks.erase(it);
            if (rest) free_blocks.emplace(o + bytes, rest);
            used += bytes;
            peak = std::max(peak, used);
            *off = o;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = align_up(std::max<size_t>(bytes, 1));
        used -= bytes;
        auto it = free_blocks.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_blocks.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_blocks.erase(nx);
        }
        if (it != free_blocks.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_blocks.erase(it);
            }
        }
    }
    // Free list sorted, aligned, inside the block, non-overlapping and
    // coalesced, and free + used == size.
    bool consistent() const {
        size_t free_sum = 0, end = 0;
        bool first = true;
        for (const auto& kv : free_blocks) {
            if (kv.second == 0 || kv.first % kAlign || kv.second % kAlign) return false;
            if (kv.first + kv.second > size) return false;
            if (!first && kv.first <= end) return false;   // overlapping or touching (not coalesced)
            end = kv.first + kv.second;
            first = false;
            free_sum += kv.second;
        }
        return free_sum + used == size;
    }
};

// Default: only requests of at most 2^20 pixels are served together, at most
// 2^22 pixels per set.  A small image's chain is bound by launch and round
// latency, which a shared set pays once, while a large image fills the device
// alone and keeps its own slot (so its transfers overlap other work).
struct SharingPolicy {
    int max_requests = 16;              // requests a slot takes at once
    int64_t max_request_px = 1 << 20;   // largest request that joins a set
    int64_t max_set_px = 1 << 22;       // pixels per set
};

// Item: a queued request with fields `bytes` (arena bytes), `off` (set on
// admission) and `px` (pixels).

// Admits the oldest queued request when its block fits (it is then still at
// the queue's front, with `off` set).  Later requests wait behind it, so a
// large request cannot be starved.
template <class Item>
bool admit_front(std::deque<Item*>& q, Arena& a) {
    if (q.empty()) return false;
    size_t off = 0;
    if (!a.alloc(q.front()->bytes, &off)) return false;
    q.front()->off = off;
    return true;
}

// After admit_front returned true: moves the admitted front request into
// `out` and, when it is small, the small requests behind it, in order, while
// they fit in the arena and the policy's limits.
template <class Item>
void take_set(std::deque<Item*>& q, Arena& a, const SharingPolicy& p, std::vector<Item*>& out) {
    int64_t px = 0;
    for (;;) {
        Item* r = q.front();
        q.pop_front();
        out.push_back(r);
        px += r->px;
        if (r->px > p.max_request_px || (int)out.size() >= p.max_requests || q.empty()) return;
        const Item* nx = q.front();
        if (nx->px > p.max_request_px || px + nx->px > p.max_set_px) return;
        if (!admit_front(q, a)) return;
    }
}

}  // namespace nsx_srv
