// Host sanitizer tests of the serving loop's host logic and the round
// predictions: the arena free list (against a brute-force occupancy map), the
// admission of queued requests into sets (policy limits, arrival order, no
// starvation of a large request) under many producer and worker threads, and
// the round-prediction table.  Built with -fsanitize=address,undefined and
// with -fsanitize=thread by the Makefile beside it; a plain host buffer stands
// in for the device block, and every served set writes and re-reads its
// blocks, so an overlapping or out-of-range block shows up as a corrupted
// pattern or a sanitizer report.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "round_hints.h"
#include "server_core.h"

#define CHECK(c)                                                             \
    do {                                                                     \
        if (!(c)) {                                                          \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
            std::exit(1);                                                    \
        }                                                                    \
    } while (0)

using nsx_srv::Arena;
using nsx_srv::kAlign;

// Lowest offset at which a maximal free run of at least `units` starts.
static long first_fit_model(const std::vector<int>& occ, size_t units) {
    size_t i = 0;
    while (i < occ.size()) {
        if (occ[i]) { ++i; continue; }
        size_t j = i;
        while (j < occ.size() && !occ[j]) ++j;
        if (j - i >= units) return (long)i;
        i = j;
    }
    return -1;
}

static void test_arena_model() {
    std::mt19937_64 rng(12345);
    for (int trial = 0; trial < 40; ++trial) {
        const size_t units = 64 + rng() % 512;
        Arena a;
        a.reset(units * kAlign - (rng() % 2 ? 0 : 17));   // also an unaligned budget
        CHECK(a.size == units * kAlign && a.consistent());
        std::vector<int> occ(units, 0);
        std::vector<std::pair<size_t, size_t>> live;       // (offset, bytes asked)
        for (int step = 0; step < 4000; ++step) {
            if (live.empty() || rng() % 100 < 55) {
                const size_t bytes = 1 + rng() % (kAlign * (1 + rng() % 24));
                const size_t need = (bytes + kAlign - 1) / kAlign;
                size_t off = 0;
                const long want = first_fit_model(occ, need);
                const bool ok = a.alloc(bytes, &off);
                CHECK(ok == (want >= 0));
                if (ok) {
                    CHECK(off % kAlign == 0 && (long)(off / kAlign) == want);
                    for (size_t u = 0; u < need; ++u) { CHECK(!occ[off / kAlign + u]); occ[off / kAlign + u] = 1; }
                    live.push_back({off, bytes});
                }
            } else {
                const size_t k = rng() % live.size();
                const size_t need = (live[k].second + kAlign - 1) / kAlign;
                for (size_t u = 0; u < need; ++u) occ[live[k].first / kAlign + u] = 0;
                a.release(live[k].first, live[k].second);
                live[k] = live.back();
                live.pop_back();
            }
            CHECK(a.consistent());
            CHECK(a.peak >= a.used);
        }
        for (auto& l : live) a.release(l.first, l.second);
        CHECK(a.consistent() && a.used == 0 && a.free_blocks.size() == 1);
    }
}

static void test_round_hints() {
    using nsx::RoundHints;
    RoundHints h;
    const uint64_t k = RoundHints::key(4, 4489, 0, 0, 8);
    CHECK(h.launch(k) == RoundHints::kDefault && h.predicted(k) == 0 && !h.has(k));
    CHECK(!h.note_existing(k, 10) && !h.has(k));            // never inserts
    h.note(k, 29);
    CHECK(h.launch(k) == 30 && h.predicted(k) == 30);
    h.note(k, 5);                                            // decays by one per run below it
    CHECK(h.launch(k) == 29);
    h.note(k, 60);                                           // rises at once
    CHECK(h.launch(k) == 61);
    CHECK(h.note_existing(k, 0) && h.launch(k) == 60);
    h.note(k, int64_t(1) << 40);                             // capped
    CHECK(h.launch(k) == RoundHints::kMax);
    h.note(k, -5);
    CHECK(h.launch(k) == RoundHints::kMax - 1);
    // quarter-octave classes: within ~19% share a key, fields never collide
    CHECK(RoundHints::key(4, 4489, 0, 0, 8) == RoundHints::key(4, 4600, 0, 0, 8));
    CHECK(RoundHints::key(4, 4489, 0, 0, 8) != RoundHints::key(4, 6000, 0, 0, 8));
    CHECK(RoundHints::key(4, 0, 0, 0, 8) == RoundHints::key(4, 1, 0, 0, 8));
    std::vector<uint64_t> keys;
    for (int pairs = 1; pairs <= 8; ++pairs)
        for (int mode = 0; mode < 4; ++mode)
            for (int dt = 0; dt < 3; ++dt)
                for (int conn : {4, 8})
                    for (int nt : {1, 7, 100, 4489, 1 << 20}) keys.push_back(RoundHints::key(pairs, nt, mode, dt, conn));
    std::sort(keys.begin(), keys.end());
    CHECK(std::adjacent_find(keys.begin(), keys.end()) == keys.end());
}

struct Item {
    uint64_t ticket = 0;
    size_t bytes = 0, off = 0;
    int64_t px = 0;
    uint64_t admit_seq = 0;
    std::atomic<int> done{0};
};

static void test_admission_threads(int producers, int workers, int per_producer) {
    const size_t arena_bytes = 64 * kAlign * 16;
    std::vector<unsigned char> mem(arena_bytes);        // stands in for the device block
    Arena arena;
    arena.reset(arena_bytes);
    nsx_srv::SharingPolicy pol;
    pol.max_requests = 5;
    pol.max_request_px = 5000;
    pol.max_set_px = 12000;
    std::mutex mu;
    std::condition_variable cv_work;
    std::deque<Item*> queue;
    bool stopping = false;
    uint64_t next_ticket = 1, next_seq = 1, last_admitted_ticket = 0;
    std::atomic<long> served{0};
    const int total = producers * per_producer;
    std::vector<Item*> all(total + 1, nullptr);

    auto worker = [&](int id) {
        std::vector<Item*> set;
        std::mt19937 rng(id * 7919 + 1);
        for (;;) {
            set.clear();
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return queue.empty() ? stopping : nsx_srv::admit_front(queue, arena); });
                if (queue.empty()) return;
                nsx_srv::take_set(queue, arena, pol, set);
                // arrival order: each set continues where the previous left off
                CHECK(!set.empty() && (int)set.size() <= pol.max_requests);
                int64_t px = 0;
                for (size_t i = 0; i < set.size(); ++i) {
                    CHECK(set[i]->ticket == last_admitted_ticket + 1);
                    last_admitted_ticket = set[i]->ticket;
                    set[i]->admit_seq = next_seq++;
                    px += set[i]->px;
                    if (set.size() > 1) CHECK(set[i]->px <= pol.max_request_px);
                }
                if (set.size() > 1) CHECK(px <= pol.max_set_px);
                CHECK(arena.consistent());
            }
            // "serve": fill each block with a pattern, yield, verify it
            for (Item* r : set) {
                CHECK(r->off + r->bytes <= mem.size());
                std::memset(mem.data() + r->off, (int)(r->ticket & 0xFF), r->bytes);
            }
            if (rng() % 4 == 0) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 200));
            else std::this_thread::yield();
            for (Item* r : set)
                for (size_t b = 0; b < r->bytes; ++b) CHECK(mem[r->off + b] == (unsigned char)(r->ticket & 0xFF));
            {
                std::lock_guard<std::mutex> lk(mu);
                for (Item* r : set) {
                    arena.release(r->off, r->bytes);
                    CHECK(r->done.fetch_add(1) == 0);
                    ++served;
                }
                CHECK(arena.consistent());
            }
            cv_work.notify_all();
        }
    };
    auto producer = [&](int id) {
        std::mt19937 rng(id * 104729 + 3);
        for (int i = 0; i < per_producer; ++i) {
            Item* r = new Item();
            // mostly small requests, a few that take a quarter to all of the arena
            const int kind = rng() % 20;
            r->bytes = kind == 0 ? arena_bytes / 4 + rng() % (arena_bytes * 3 / 4) : 1 + rng() % (8 * kAlign);
            r->px = kind == 0 ? 1 << 20 : 1 + rng() % 6000;
            {
                std::lock_guard<std::mutex> lk(mu);
                r->ticket = next_ticket++;
                all[r->ticket] = r;
                queue.push_back(r);
            }
            cv_work.notify_all();
            if (rng() % 8 == 0) std::this_thread::yield();
        }
    };
    std::vector<std::thread> ws, ps;
    for (int i = 0; i < workers; ++i) ws.emplace_back(worker, i);
    for (int i = 0; i < producers; ++i) ps.emplace_back(producer, i);
    for (auto& t : ps) t.join();
    {
        std::lock_guard<std::mutex> lk(mu);
        stopping = true;
    }
    cv_work.notify_all();
    for (auto& t : ws) t.join();
    CHECK(served.load() == total && queue.empty());
    CHECK(arena.used == 0 && arena.consistent() && arena.free_blocks.size() == 1);
    for (int t = 1; t <= total; ++t) {
        CHECK(all[t] && all[t]->done.load() == 1 && all[t]->admit_seq == (uint64_t)t);
        delete all[t];
    }
}

int main() {
    test_arena_model();
    std::printf("arena model: ok\n");
    test_round_hints();
    std::printf("round hints: ok\n");
    test_admission_threads(3, 4, 400);
    test_admission_threads(6, 8, 150);
    test_admission_threads(1, 1, 300);
    std::printf("admission threads: ok\n");
    return 0;
}
