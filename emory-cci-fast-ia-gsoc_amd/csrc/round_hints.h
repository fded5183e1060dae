// round_hints.h -- host-only: the propagation-round predictions of a context.
//
// A reconstruction launches, without a host check, as many propagation rounds
// as the last problem of the same kind and size class needed; the host reads
// the convergence word once afterwards.  Size classes are quarter octaves of
// the tile count, so problems within ~19% of each other share a prediction.
// A prediction rises at once to what a run needed and decays by one round per
// run below it, so one unusually long input costs its successors at most a
// few empty rounds each.  No HIP here: the logic is also built by the host
// sanitizer tests.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <unordered_map>

namespace nsx {

struct RoundHints {
    static constexpr int kDefault = 4;       // rounds launched for an unseen key
    static constexpr int kMax = 4096;        // a prediction never exceeds this

    // Key of (sweep pairs per visit, tile count, mode, dtype, connectivity).
    static uint64_t key(int pairs, int ntiles, int mode, int dtype, int conn) {
        const uint64_t bucket = (uint64_t)std::lround(4.0 * std::log2((double)std::max(1, ntiles)));
        return ((uint64_t)(uint32_t)pairs << 40) ^ (bucket << 16) ^ ((uint64_t)(uint32_t)mode << 8) ^
               ((uint64_t)(uint32_t)dtype << 4) ^ (uint64_t)(uint32_t)conn;
    }
    // Rounds to launch for `k` (kDefault when unseen).
    int launch(uint64_t k) const {
        auto it = map.find(k);
        return it != map.end() ? it->second : kDefault;
    }
    // The stored prediction, 0 when unseen.
    int predicted(uint64_t k) const {
        auto it = map.find(k);
        return it != map.end() ? it->second : 0;
    }
    static int need_of(int64_t rounds_with_work) {
        return (int)std::min<int64_t>(kMax, std::max<int64_t>(0, rounds_with_work) + 1);
    }
    // Records that a run of `k` had work in `rounds_with_work` rounds.
    void note(uint64_t k, int64_t rounds_with_work) {
        const int need = need_of(rounds_with_work);
        auto it = map.find(k);
        if (it == map.end()) map.emplace(k, need);
        else it->second = std::max(need, it->second - 1);
    }
    // As note(), but only for a key already present (callers that must not
    // allocate); returns whether it was.
    bool note_existing(uint64_t k, int64_t rounds_with_work) {
        auto it = map.find(k);
        if (it == map.end()) return false;
        it->second = std::max(need_of(rounds_with_work), it->second - 1);
        return true;
    }
    bool has(uint64_t k) const { return map.count(k) != 0; }

    std::unordered_map<uint64_t, int> map;
};

}  // namespace nsx
