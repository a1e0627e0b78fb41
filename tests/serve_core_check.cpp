// Drives the request server's scheduler (sprout_amd/csrc/serve_core.h) with a backend that
// only keeps time, from several submitting threads, and checks what the device backend
// relies on:
//   - every request completes (no stall), with drain() returning;
//   - each scene's requests run in arrival order, one launch at a time;
//   - a launch holds only resident scenes of one key, at most max_group of them;
//   - a scene is never evicted or restored while a launch holds it;
//   - after every restore the resident scenes fit the budget;
//   - the statistics count every request and every particle-step.
// Usage: serve_core_check <seed> <scenes> <requests per scene> <slots> <budget units>
// Prints "ok ..." and exits 0, or prints the broken rule and exits 1.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <random>
#include <stdexcept>
#include <thread>
#include <vector>

#include "../sprout_amd/csrc/serve_core.h"

namespace {

[[noreturn]] void fail(const char* what, long a = 0, long b = 0) {
    std::fprintf(stderr, "FAIL: %s (%ld, %ld)\n", what, a, b);
    std::fflush(stderr);
    std::_Exit(1);
}

struct MockBackend {
    struct Scene {
        uint64_t bytes = 1;
        uint64_t key = 0;
        bool resident = true;
        bool held = false;
        std::vector<uint32_t> ran;  // substep counts of its launches, in order
    };
    static constexpr int MAX = 4096;
    Scene scenes[MAX];
    uint64_t budget = 0;
    uint32_t max_group = 0;
    std::vector<std::chrono::steady_clock::time_point> ends;
    std::vector<std::vector<uint32_t>> members;
    std::mt19937 rng{1};
    std::atomic<uint32_t> n{0};  // scenes entered
    std::atomic<uint64_t> restores{0}, evictions{0};

    explicit MockBackend(int slots) : ends((size_t)slots), members((size_t)slots) {}

    uint64_t resident_total() const {
        uint64_t b = 0;
        for (uint32_t k = 0; k < n; k++)
            if (scenes[k].resident) b += scenes[k].bytes;
        return b;
    }
    uint64_t bytes(uint32_t k) { return scenes[k].bytes; }
    bool resident(uint32_t k) { return scenes[k].resident; }
    uint64_t key(uint32_t k) { return scenes[k].key; }
    void evict(uint32_t k) {
        if (!scenes[k].resident) fail("evict of an evicted scene", k);
        if (scenes[k].held) fail("evict of a scene a launch holds", k);
        scenes[k].resident = false;
        evictions++;
    }
    void restore(uint32_t k) {
        if (scenes[k].resident) fail("restore of a resident scene", k);
        if (scenes[k].held) fail("restore of a scene a launch holds", k);
        scenes[k].resident = true;
        restores++;
        if (resident_total() > budget) fail("resident scenes over the budget after a restore", (long)resident_total(), (long)budget);
    }
    void launch(int slot, const std::vector<uint32_t>& ms, uint32_t substeps, const std::vector<float>& targets) {
        if (ms.empty() || ms.size() > max_group) fail("launch group size", (long)ms.size());
        if (targets.size() != ms.size()) fail("targets per member");
        for (uint32_t m : ms) {
            if (!scenes[m].resident) fail("launch of an evicted scene", m);
            if (scenes[m].held) fail("a scene in two launches", m);
            if (scenes[m].key != scenes[ms[0]].key) fail("launch mixes keys", m);
            scenes[m].held = true;
            scenes[m].ran.push_back(substeps);
        }
        std::uniform_int_distribution<int> us(20, 400);
        ends[(size_t)slot] = std::chrono::steady_clock::now() + std::chrono::microseconds(us(rng) * (int)substeps);
        members[(size_t)slot] = ms;
    }
    bool done(int slot) { return std::chrono::steady_clock::now() >= ends[(size_t)slot]; }
    uint64_t finish(int slot, uint32_t scene, uint32_t substeps) {
        bool found = false;
        for (uint32_t m : members[(size_t)slot]) found = found || m == scene;
        if (!found || !scenes[scene].held) fail("finish of a scene the slot does not hold", scene);
        scenes[scene].held = false;
        return 1000ull * substeps;
    }
    void drained() {}
};

}  // namespace

int main(int argc, char** argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s seed scenes requests slots budget_units\n", argv[0]);
        return 2;
    }
    const unsigned seed = (unsigned)std::atoi(argv[1]);
    const int n_scenes = std::atoi(argv[2]), per = std::atoi(argv[3]), slots = std::atoi(argv[4]);
    const uint64_t budget = (uint64_t)std::atoll(argv[5]);
    std::mt19937 rng(seed);
    MockBackend be(slots);
    be.budget = budget;
    be.max_group = 4;
    be.rng.seed(seed + 1);
    serve::Core<MockBackend> core(be, budget, slots, be.max_group, -2);
    core.start();
    // scenes of 1..4 units (none above the budget), three keys
    std::vector<std::vector<uint32_t>> plan((size_t)n_scenes);
    for (int k = 0; k < n_scenes; k++) {
        uint32_t id = 0;
        const uint64_t bytes = 1 + rng() % std::min<uint64_t>(4, budget);
        const uint64_t key = rng() % 3;
        bool ok = core.add(&id, [&](uint32_t i) {
            be.scenes[i].bytes = bytes;
            be.scenes[i].key = key;
            be.scenes[i].resident = true;
            be.n = i + 1;
            return true;
        });
        if (!ok || id != (uint32_t)k) fail("add", k);
        for (int r = 0; r < per; r++) plan[(size_t)k].push_back(1 + rng() % 3);
    }
    // four submitters, each owning every fourth scene, interleaving at random
    std::vector<std::thread> feeders;
    std::mutex tmu;
    std::vector<uint64_t> tickets;
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
                    tickets.push_back(t);
                }
                if (r() % 4 == 0) std::this_thread::sleep_for(std::chrono::microseconds(r() % 300));
            }
        });
    for (auto& t : feeders) t.join();
    // a stall shows as drain() not returning: a watchdog ends the check
    std::atomic<bool> drained{false};
    std::thread dog([&] {
        for (int i = 0; i < 600 && !drained; i++) std::this_thread::sleep_for(std::chrono::milliseconds(100));
        if (!drained) fail("drain did not return within 60 s: a stall");
    });
    if (!core.drain()) fail("server failed");
    drained = true;
    dog.join();
    for (uint64_t t : tickets) {
        serve::Req q;
        if (!core.get(t, false, &q) || q.status != 0) fail("request not done", (long)t);
        if (!(q.submit_ms <= q.launch_ms && q.launch_ms <= q.done_ms)) fail("request times out of order", (long)t);
    }
    serve::Stats st = core.stats();
    core.stop();
    uint64_t steps = 0;
    for (int k = 0; k < n_scenes; k++) {
        if (be.scenes[k].ran != plan[(size_t)k]) fail("a scene's requests did not run in arrival order", k);
        for (uint32_t n : plan[(size_t)k]) steps += 1000ull * n;
    }
    if (st.completed != (uint64_t)n_scenes * per || st.submitted != st.completed || st.failed != 0)
        fail("request counts", (long)st.completed, (long)st.submitted);
    if (st.particle_steps != steps) fail("particle-steps", (long)st.particle_steps, (long)steps);
    if (!(st.p50_ms <= st.p99_ms && st.p99_ms <= st.max_ms)) fail("latency percentiles");
    std::printf("ok seed=%u scenes=%d requests=%llu launches=%llu tanks/launch=%.2f evictions=%llu restores=%llu p50=%.3fms p99=%.3fms\n",
                seed, n_scenes, (unsigned long long)st.completed, (unsigned long long)st.launches, st.tanks_per_launch,
                (unsigned long long)st.evictions, (unsigned long long)st.restores, st.p50_ms, st.p99_ms);
    return 0;
}
