// CPU model of the persistent drain's tile queue (switched off in the
// library), run with host threads standing in for GPU waves.
//
// The removed drain (one persistent kernel over a device FIFO of tile ids)
// hung the GPU on a 1 x 130 image (3 tiles, 3 worker waves).  Its pop read the
// queue's head and tail with relaxed atomic loads and, after losing the
// compare-and-swap on the head, went straight back to those loads:
//
//     h = load_relaxed(head); t = load_relaxed(tail);
//     if (h < t) { if (CAS(head, h, h + 1) == h) { ...take slot h... } continue; }
//
// Every other wait had a wall-clock bound; this one had none.  It ends only
// when a later relaxed load returns the head that the winning CAS wrote.  On
// MI355X each XCD has its own L2, and a plain or relaxed load can be served
// from a line that XCD cached earlier, while the CAS is performed at the
// memory side.  A waiter that keeps reading its XCD's stale copy of the head
// therefore retries the CAS forever, and the kernel never ends.
//
// This model gives each simulated XCD such a cache for relaxed loads (a value
// once read is returned again until the same XCD writes that word with a
// plain store; read-modify-writes go to memory and do not refresh it).  Two
// protocols run on it:
//   old: the removed pop, with its unbounded CAS-retry loop;
//   new: a ticket queue.  A worker claims a position with fetch_add on the
//        head (a read-modify-write, always current) and waits for that slot's
//        sequence number; a push claims a position on the tail and waits until
//        the slot's previous position (one lap back) was taken.  Every control
//        word a wait depends on is read through a read-modify-write, and every
//        wait is bounded: it ends the worker (error flag) when nothing moves
//        for too long.  An earlier version of this model let a push reuse a
//        slot whose item from any older lap was taken; a later position could
//        then overtake an earlier push, and the model caught the deadlock.
// The old protocol is shown to spin without progress in the 3-tile case; the
// new one drains random workloads to the right visit counts under the same
// caches: every activation of a tile is followed by a visit of that tile that
// starts after it, and nothing stays queued.  Exit status 0 when both results
// hold.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

namespace {

constexpr uint32_t RING_EMPTY = 0xFFFFFFFFu;

// Per-XCD cache for relaxed loads (stale until the XCD itself stores).
struct XcdCache {
    std::mutex mu;
    std::map<const void*, uint32_t> lines;
};

struct Mem {
    std::atomic<uint32_t> head{0}, tail{0}, pending{0}, error{0}, visits{0};
    std::vector<std::atomic<uint32_t>> ring;
    std::vector<std::atomic<uint64_t>> slots;   // new protocol: (sequence << 32) | tile
    std::vector<std::atomic<uint32_t>> tstate;
    XcdCache xcd[8];
    bool stale = true;
    explicit Mem(int ntiles, int cap) : ring(cap), slots(cap), tstate(ntiles) {
        for (auto& r : ring) r.store(RING_EMPTY);
        for (size_t i = 0; i < slots.size(); ++i)   // "position i - cap, taken"
            slots[i].store(((uint64_t)(uint32_t)((uint32_t)i - (uint32_t)cap) << 32) | RING_EMPTY);
        for (auto& t : tstate) t.store(0);
    }
    uint32_t load_relaxed(int x, const std::atomic<uint32_t>& a) {
        if (!stale) return a.load(std::memory_order_relaxed);
        std::lock_guard<std::mutex> lk(xcd[x].mu);
        auto it = xcd[x].lines.find(&a);
        if (it != xcd[x].lines.end()) return it->second;
        const uint32_t v = a.load(std::memory_order_relaxed);
        xcd[x].lines.emplace(&a, v);
        return v;
    }
};

constexpr uint32_t TQ_DIRTY = 1u, TQ_RUN = 2u;

// The workload: tile t's visits activate its neighbours (t - 1, t + 1) while
// the tile still has raises to make (a stand-in for halo pixels that rose).
struct Script {
    int ntiles;
    std::vector<std::atomic<int>> left;   // raises each tile still has to make
    explicit Script(int n, int each) : ntiles(n), left(n) { for (auto& l : left) l.store(each); }
    // one visit: returns the neighbours to activate
    std::vector<int> visit(int t) {
        std::vector<int> act;
        if (left[t].fetch_sub(1) > 0) {
            if (t > 0) act.push_back(t - 1);
            if (t + 1 < ntiles) act.push_back(t + 1);
        }
        return act;
    }
};

// ---- the removed protocol ------------------------------------------------------
struct Old {
    Mem& m;
    uint32_t mask;
    long spin_cap;   // model-only bound, to report the livelock instead of hanging
    std::atomic<long> max_cas_retries{0};

    void push(int, uint32_t t) {
        m.pending.fetch_add(1);
        const uint32_t pos = m.tail.fetch_add(1);
        uint32_t expect = RING_EMPTY;
        while (!m.ring[pos & mask].compare_exchange_weak(expect, t)) expect = RING_EMPTY;
    }
    void activate(int x, uint32_t t) {
        if (m.tstate[t].fetch_or(TQ_DIRTY) == 0u) push(x, t);
    }
    // returns a tile, -1 when drained, -2 when the CAS-retry loop exceeded spin_cap
    int pop(int x) {
        long retries = 0;
        while (true) {
            const uint32_t h = m.load_relaxed(x, m.head);
            const uint32_t t = m.load_relaxed(x, m.tail);
            if (h < t) {
                uint32_t e = h;
                if (m.head.compare_exchange_strong(e, h + 1)) {
                    while (true) {
                        const uint32_t v = m.ring[h & mask].exchange(RING_EMPTY);
                        if (v != RING_EMPTY) return (int)v;
                        std::this_thread::yield();
                    }
                }
                if (++retries > spin_cap) {
                    long prev = max_cas_retries.load();
                    while (retries > prev && !max_cas_retries.compare_exchange_weak(prev, retries)) {}
                    return -2;   // the real loop has no exit here
                }
                continue;       // back to the (stale) relaxed loads
            }
            if (m.pending.load() == 0u) return -1;
            std::this_thread::yield();
        }
    }
};

// ---- the bounded ticket queue ---------------------------------------------------
struct New {
    Mem& m;
    uint32_t cap;
    std::chrono::milliseconds stall{2000};

    static uint64_t pack(uint32_t seq, uint32_t t) { return ((uint64_t)seq << 32) | t; }
    uint32_t rmw_read(std::atomic<uint32_t>& a) { return a.fetch_add(0); }   // always the current value

    bool push(uint32_t t) {
        m.pending.fetch_add(1);
        const uint32_t pos = m.tail.fetch_add(1);
        auto& slot = m.slots[pos % cap];
        // the slot's previous position (pos - cap) must have been taken
        const auto t0 = std::chrono::steady_clock::now();
        while (true) {
            uint64_t cur = slot.fetch_add(0);
            const uint32_t seq = (uint32_t)(cur >> 32);
            // free for position pos: the item of position pos - cap was taken
            // (a slot starts as "position i - cap, taken"); an older taken
            // item is not enough, or pos could overtake pos - cap's push
            if ((uint32_t)cur == RING_EMPTY && seq == pos - cap) {
                if (slot.compare_exchange_strong(cur, pack(pos, t))) return true;
                continue;    // re-read through the returned value
            }
            if (std::chrono::steady_clock::now() - t0 > stall || rmw_read(m.error)) {
                m.error.fetch_or(1);
                return false;
            }
            std::this_thread::yield();
        }
    }
    std::vector<std::atomic<uint32_t>>* acts = nullptr;   // activations per tile (the check)
    void activate(uint32_t t) {
        if (acts) (*acts)[t].fetch_add(1);
        if (m.tstate[t].fetch_or(TQ_DIRTY) == 0u) push(t);
    }
    // a claimed position's tile, or -1 when everything is drained (or an error)
    int pop() {
        const uint32_t my = m.head.fetch_add(1);
        auto& slot = m.slots[my % cap];
        auto t0 = std::chrono::steady_clock::now();
        uint32_t last_visits = rmw_read(m.visits);
        while (true) {
            const uint64_t cur = slot.fetch_add(0);
            if ((uint32_t)(cur >> 32) == my && (uint32_t)cur != RING_EMPTY) {
                // take it: mark the slot free for position my + cap
                slot.store(pack(my, RING_EMPTY));
                return (int)(uint32_t)cur;
            }
            if (rmw_read(m.pending) == 0u || rmw_read(m.error)) return -1;
            const uint32_t v = rmw_read(m.visits);
            if (v != last_visits) { last_visits = v; t0 = std::chrono::steady_clock::now(); }
            else if (std::chrono::steady_clock::now() - t0 > stall) { m.error.fetch_or(2); return -1; }
            std::this_thread::yield();
        }
    }
};

// Runs the first pass (every tile once) and the drain with `workers` threads,
// worker w on XCD w % 8.  Returns the number of visits in the drain, -1 on a
// reported livelock (old) or error (new).
long run_old(int ntiles, int workers, int each, long spin_cap, long* worst_retries, bool stale = true) {
    int cap = 1;
    while (cap < ntiles) cap <<= 1;
    Mem m(ntiles, cap);
    m.stale = stale;
    Script sc(ntiles, each);
    Old q{m, (uint32_t)(cap - 1), spin_cap};
    for (int t = 0; t < ntiles; ++t)
        for (int a : sc.visit(t)) q.activate(0, (uint32_t)a);
    std::atomic<bool> livelock{false};
    std::vector<std::thread> th;
    for (int w = 0; w < workers; ++w)
        th.emplace_back([&, w] {
            const int x = w % 8;
            while (true) {
                const int t = q.pop(x);
                if (t == -2) { livelock = true; return; }
                if (t < 0) return;
                m.tstate[t].store(TQ_RUN);
                for (int a : sc.visit(t)) q.activate(x, (uint32_t)a);
                uint32_t e = TQ_RUN;
                if (!m.tstate[t].compare_exchange_strong(e, 0u)) {
                    m.tstate[t].store(TQ_DIRTY);
                    q.push(x, (uint32_t)t);
                }
                m.visits.fetch_add(1);
                m.pending.fetch_sub(1);
            }
        });
    for (auto& t : th) t.join();
    *worst_retries = q.max_cas_retries.load();
    return livelock ? -1 : (long)m.visits.load();
}

long run_new(int ntiles, int workers, int each, unsigned seed) {
    int cap = 1;
    while (cap < ntiles + workers + 1) cap <<= 1;
    Mem m(ntiles, cap);
    Script sc(ntiles, each);
    New q{m, (uint32_t)cap};
    // every activation must be followed by a visit that starts after it
    std::vector<std::atomic<uint32_t>> acts(ntiles), seen(ntiles);
    for (int t = 0; t < ntiles; ++t) { acts[t].store(0); seen[t].store(0); }
    q.acts = &acts;
    for (int t = 0; t < ntiles; ++t)
        for (int a : sc.visit(t)) q.activate((uint32_t)a);
    std::vector<std::thread> th;
    for (int w = 0; w < workers; ++w)
        th.emplace_back([&, w] {
            std::mt19937 rng(seed * 131 + w);
            while (true) {
                const int t = q.pop();
                if (t < 0) return;
                m.tstate[t].store(TQ_RUN);
                const uint32_t a = acts[t].load();
                uint32_t prev = seen[t].load();
                while (a > prev && !seen[t].compare_exchange_weak(prev, a)) {}
                if (rng() % 4 == 0) std::this_thread::yield();
                for (int a : sc.visit(t)) q.activate((uint32_t)a);
                uint32_t e = TQ_RUN;
                if (!m.tstate[t].compare_exchange_strong(e, 0u)) {
                    m.tstate[t].store(TQ_DIRTY);
                    q.push((uint32_t)t);
                }
                m.visits.fetch_add(1);
                m.pending.fetch_sub(1);
            }
        });
    for (auto& t : th) t.join();
    if (m.error.load()) {
        std::printf("    error bits %u (head %u tail %u pending %u)\n", m.error.load(), m.head.load(), m.tail.load(),
                    m.pending.load());
        return -1;
    }
    for (int t = 0; t < ntiles; ++t)
        if (seen[t].load() != acts[t].load() || m.tstate[t].load() != 0u) {   // an activation never visited
            std::printf("    tile %d activations %u, last seen by a visit %u, state %u\n", t, acts[t].load(),
                        seen[t].load(), m.tstate[t].load());
            return -1;
        }
    return (long)m.visits.load();
}

}  // namespace

int main() {
    // the hung case: a 1 x 130 image, 3 tiles of 62 columns, 3 worker waves
    long worst = 0;
    const long old_visits = run_old(3, 3, 40, 2000000, &worst);
    std::printf("old protocol, 3 tiles / 3 waves, stale relaxed loads: %s (CAS retries without progress: %ld)\n",
                old_visits < 0 ? "LIVELOCK" : "drained", worst);
    long worst_c = 0;
    const long old_coherent = run_old(3, 3, 40, 2000000, &worst_c, false);
    std::printf("old protocol, same case, coherent relaxed loads: %s (%ld drain visits, most CAS retries %ld)\n",
                old_coherent < 0 ? "LIVELOCK" : "drained", old_coherent, worst_c);
    bool ok = old_visits < 0 && old_coherent >= 0;
    for (unsigned seed = 1; seed <= 40 && ok; ++seed) {
        std::mt19937 rng(seed);
        const int ntiles = 1 + (int)(rng() % 40), workers = 1 + (int)(rng() % 12), each = (int)(rng() % 30);
        const long v = run_new(ntiles, workers, each, seed);
        std::printf("  seed %2u: %2d tiles, %2d workers, %2d raises each: %ld drain visits\n", seed, ntiles, workers,
                    each, v);
        if (v < 0) {
            std::printf("new protocol FAILED: seed %u tiles %d workers %d\n", seed, ntiles, workers);
            ok = false;
        }
    }
    const long v3 = run_new(3, 3, 40, 7);
    std::printf("new protocol, 3 tiles / 3 waves, same caches: %s (%ld drain visits); 40 random workloads: %s\n",
                v3 >= 0 ? "drained" : "FAILED", v3, ok ? "drained" : "FAILED");
    return ok && v3 >= 0 ? 0 : 1;
}
