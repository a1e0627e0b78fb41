// Drives the continuous-batching core (sprout_amd/csrc/pool_core.h) with a backend that only
// records what it is asked to do, from several submitting threads, one serving loop
// (step()) and waiters that drive rounds themselves (get(block)), and checks what the
// device backend relies on:
//   - every request completes, and drain() returns;
//   - each scene's requests run in arrival order, each as begin, then one launch per round
//     for exactly its substep count in consecutive rounds, then finish, never overlapping;
//   - a launch holds only resident scenes of one key, at most max_group, and a scene at
//     most once per round;
//   - no scene is evicted or restored while a request of it is active;
//   - after every restore the resident scenes fit the budget;
//   - a request whose begin fails completes with that error and the rest go on;
//   - the statistics count every request and every particle-step.
// Usage: pool_core_check <seed> <scenes> <requests per scene> <streams> <budget units>
// Prints "ok ..." and exits 0, or prints the broken rule and exits 1.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "../sprout_amd/csrc/pool_core.h"

namespace {

[[noreturn]] void fail(const char* what, long a = 0, long b = 0) {
    std::fprintf(stderr, "FAIL: %s (%ld, %ld)\n", what, a, b);
    std::fflush(stderr);
    std::_Exit(1);
}

constexpr int BAD = -3;  // the error a refused begin reports

struct MockBackend {
    struct Scene {
        uint64_t bytes = 1, key = 0;
        bool resident = true;
        uint32_t active = 0;  // substeps of the active request, 0: idle
        uint32_t ran = 0;     // its launches so far
        uint64_t last_round = 0;
        bool last_seen = false;      // its latest launch was flagged as its request's last
        std::vector<uint32_t> done;  // substep counts of its finished requests, in order
        int begins = 0;
        int refuse_at = -1;  // this begin (0-based) is refused
    };
    std::vector<Scene> sc;
    uint64_t budget = 0;
    uint32_t max_group = 0;
    int streams = 0;
    uint64_t round = 1;
    std::vector<char> in_round;
    std::atomic<int> driving{0};

    uint64_t resident_total() const {
        uint64_t b = 0;
        for (const Scene& s : sc)
            if (s.resident) b += s.bytes;
        return b;
    }
    void enter() {
        if (driving.fetch_add(1) != 0) fail("two threads drive at once");
    }
    void leave() { driving.fetch_sub(1); }

    uint64_t bytes(uint32_t k) { return sc[k].bytes; }
    bool resident(uint32_t k) { return sc[k].resident; }
    uint64_t key(uint32_t k) { return sc[k].key; }
    void evict(uint32_t k) {
        enter();
        if (!sc[k].resident) fail("evict of an evicted scene", k);
        if (sc[k].active) fail("evict of an active scene", k);
        sc[k].resident = false;
        leave();
    }
    void restore(uint32_t k) {
        enter();
        if (sc[k].resident) fail("restore of a resident scene", k);
        if (sc[k].active) fail("restore of an active scene", k);
        sc[k].resident = true;
        if (resident_total() > budget) fail("resident scenes over the budget after a restore", (long)resident_total(), (long)budget);
        leave();
    }
    int begin(uint32_t k, uint32_t substeps, float) {
        enter();
        Scene& s = sc[k];
        if (!s.resident) fail("begin of an evicted scene", k);
        if (s.active) fail("begin while a request of the scene is active", k);
        const int b = s.begins++;
        leave();
        if (b == s.refuse_at) return BAD;
        s.active = substeps;
        s.ran = 0;
        return 0;
    }
    void launch(int stream, const std::vector<uint32_t>& ks, const std::vector<char>& last) {
        enter();
        if (stream < 0 || stream >= streams) fail("stream out of range", stream);
        if (ks.empty() || ks.size() > max_group) fail("launch group size", (long)ks.size());
        if (last.size() != ks.size()) fail("last flags per member");
        for (size_t i = 0; i < ks.size(); i++) {
            const uint32_t k = ks[i];
            Scene& s = sc[k];
            if ((last[i] != 0) != (s.ran + 1 == s.active)) fail("last-substep flag wrong", k, (long)s.ran);
            s.last_seen = last[i] != 0;
            if (!s.resident) fail("launch of an evicted scene", k);
            if (!s.active) fail("launch of an idle scene", k);
            if (s.key != sc[ks[0]].key) fail("launch mixes keys", k);
            if (in_round[k]) fail("a scene twice in one round", k);
            if (s.ran > 0 && s.last_round + 1 != round) fail("a request's substeps in rounds that are not consecutive", k);
            in_round[k] = 1;
            s.ran++;
            s.last_round = round;
            if (s.ran > s.active) fail("more launches than the request's substeps", k);
        }
        leave();
    }
    void sync() {
        enter();
        for (uint32_t k = 0; k < sc.size(); k++)
            if (sc[k].active && !in_round[k]) fail("an active scene left out of a round", k);
        std::fill(in_round.begin(), in_round.end(), 0);
        round++;
        leave();
    }
    uint64_t finish(uint32_t k, uint32_t substeps) {
        enter();
        Scene& s = sc[k];
        if (s.active != substeps || s.ran != substeps) fail("finish before every substep ran", k, (long)s.ran);
        if (!s.last_seen) fail("finish without a launch flagged last", k);
        s.done.push_back(substeps);
        s.active = 0;
        leave();
        return 1000ull * substeps;
    }
};

}  // namespace

int main(int argc, char** argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s seed scenes requests streams budget_units\n", argv[0]);
        return 2;
    }
    const unsigned seed = (unsigned)std::atoi(argv[1]);
    const int n_scenes = std::atoi(argv[2]), per = std::atoi(argv[3]), streams = std::atoi(argv[4]);
    const uint64_t budget = (uint64_t)std::atoll(argv[5]);
    std::mt19937 rng(seed);
    MockBackend be;
    be.budget = budget;
    be.max_group = 3;
    be.streams = streams;
    be.sc.resize((size_t)n_scenes);
    be.in_round.assign((size_t)n_scenes, 0);
    pool::Core<MockBackend> core(be, budget, be.max_group, streams, -2);
    // scenes of 1..4 units (none above the budget), three keys; every fifth scene refuses
    // one of its requests
    std::vector<std::vector<uint32_t>> plan((size_t)n_scenes);
    for (int k = 0; k < n_scenes; k++) {
        uint32_t id = 0;
        const uint64_t bytes = 1 + rng() % std::min<uint64_t>(4, budget);
        const uint64_t key = rng() % 3;
        const int refuse = k % 5 == 4 ? (int)(rng() % (unsigned)per) : -1;
        bool ok = core.add(&id, [&](uint32_t i) {
            be.sc[i].bytes = bytes;
            be.sc[i].key = key;
            be.sc[i].refuse_at = refuse;
            return true;
        });
        if (!ok || id != (uint32_t)k) fail("add", k);
        for (int r = 0; r < per; r++) plan[(size_t)k].push_back(1 + rng() % 3);
    }
    // four submitters, each owning every fourth scene, interleaving at random
    std::mutex tmu;
    std::vector<std::pair<uint64_t, int>> tickets;  // ticket, scene
    std::atomic<int> feeding{4};
    std::vector<std::thread> feeders;
    for (int f = 0; f < 4; f++)
        feeders.emplace_back([&, f] {
            std::mt19937 r(seed * 7 + (unsigned)f);
            std::vector<int> pos((size_t)n_scenes, 0);
            std::vector<int> mine;
            for (int k = f; k < n_scenes; k += 4)
                for (int q = 0; q < per; q++) mine.push_back(k);
            std::shuffle(mine.begin(), mine.end(), r);
            for (int k : mine) {
                uint64_t t = 0;
                const uint32_t n = plan[(size_t)k][(size_t)pos[(size_t)k]++];
                if (core.submit((uint32_t)k, n, n / 240.0f, &t) != 0) fail("submit", k);
                {
                    std::lock_guard<std::mutex> g(tmu);
                    tickets.push_back({t, k});
                }
                if (r() % 4 == 0) std::this_thread::sleep_for(std::chrono::microseconds(r() % 300));
            }
            feeding--;
        });
    // the serving loop, and two waiters that drive rounds when the loop does not
    std::atomic<bool> loop_on{true};
    std::thread loop([&] {
        while (loop_on)
            if (core.step(0.2) < 0) fail("step failed");
    });
    std::vector<std::thread> waiters;
    for (int w = 0; w < 2; w++)
        waiters.emplace_back([&, w] {
            std::mt19937 r(seed * 13 + (unsigned)w);
            while (feeding > 0) {
                uint64_t t = 0;
                {
                    std::lock_guard<std::mutex> g(tmu);
                    if (!tickets.empty()) t = tickets[r() % tickets.size()].first;
                }
                pool::Req q;
                if (t && core.get(t, true, &q) && q.status == pool::REQ_PENDING) fail("wait returned a pending request", (long)t);
                std::this_thread::sleep_for(std::chrono::microseconds(50));
            }
        });
    for (auto& t : feeders) t.join();
    for (auto& t : waiters) t.join();
    // a stall shows as drain() not returning: a watchdog ends the check
    std::atomic<bool> drained{false};
    std::thread dog([&] {
        for (int i = 0; i < 600 && !drained; i++) std::this_thread::sleep_for(std::chrono::milliseconds(100));
        if (!drained) fail("drain did not return within 60 s: a stall");
    });
    if (!core.drain()) fail("pool failed");
    drained = true;
    dog.join();
    loop_on = false;
    loop.join();
    uint64_t refused = 0;
    for (const auto& tk : tickets) {
        pool::Req q;
        if (!core.get(tk.first, false, &q) || q.status == pool::REQ_PENDING) fail("request not done", (long)tk.first);
        if (q.status == BAD) {
            refused++;
            continue;
        }
        if (q.status != 0) fail("request failed", (long)tk.first, q.status);
        if (!(q.submit_ms <= q.start_ms && q.start_ms <= q.done_ms)) fail("request times out of order", (long)tk.first);
        if (q.tanks < 1 || q.tanks > be.max_group) fail("request's group size", (long)q.tanks);
    }
    pool::Stats st = core.stats();
    uint64_t steps = 0, expect_refused = 0;
    for (int k = 0; k < n_scenes; k++) {
        std::vector<uint32_t> want = plan[(size_t)k];
        if (be.sc[(size_t)k].refuse_at >= 0) {
            want.erase(want.begin() + be.sc[(size_t)k].refuse_at);
            expect_refused++;
        }
        if (be.sc[(size_t)k].done != want) fail("a scene's requests did not run in arrival order", k);
        for (uint32_t n : want) steps += 1000ull * n;
        if (be.sc[(size_t)k].active) fail("a scene left active", k);
    }
    if (refused != expect_refused) fail("refused requests", (long)refused, (long)expect_refused);
    if (st.submitted != (uint64_t)n_scenes * per || st.completed + st.failed != st.submitted || st.failed != refused)
        fail("request counts", (long)st.completed, (long)st.failed);
    if (st.particle_steps != steps) fail("particle-steps", (long)st.particle_steps, (long)steps);
    if (!(st.p50_ms <= st.p99_ms && st.p99_ms <= st.max_ms)) fail("latency percentiles");
    if (st.resident_bytes > budget) fail("resident bytes over the budget at the end", (long)st.resident_bytes);
    std::printf("ok seed=%u scenes=%d requests=%llu rounds=%llu launches=%llu tanks/launch=%.2f evictions=%llu restores=%llu refused=%llu\n",
                seed, n_scenes, (unsigned long long)st.submitted, (unsigned long long)st.rounds, (unsigned long long)st.launches,
                st.tanks_per_launch, (unsigned long long)st.evictions, (unsigned long long)st.restores, (unsigned long long)refused);
    return 0;
}
