// pool_core.h — continuous batching of step requests in rounds: requests, scene queues,
// admission under the memory budget, launch groups and completions, with no device code and
// no thread of its own. The caller's serving loop drives it (step(), or wait() / drain(),
// which drive it themselves); submit() may be called from any thread at any time. The
// device side is a Backend (pool.hip: HIP launches on the tanks); tests drive the same core
// with a backend that only records what it is asked to do.
//
// One round:
//   1. the requests that arrived join their scenes' queues (arrival order per scene);
//   2. resident scenes over the budget that no request holds are evicted, least recently
//      used first;
//   3. every scene with a queued request and no active one starts its oldest request
//      (oldest request first): it is made resident within the budget (evicting idle scenes,
//      those with nothing queued first, then the least recently used; a scene that does not
//      fit waits, and the scenes behind it may only start if they are resident already, so
//      a large scene is not passed over for memory), and the backend prepares the request;
//   4. every active scene advances one substep: scenes of one key (same grid, same solver
//      schedule) share launches, at most `max_group` per launch, launches spread over the
//      streams so that they overlap on the device;
//   5. the backend waits for the round's launches;
//   6. requests whose last substep ran complete: the backend reads the scene's state back,
//      and the scene is idle again.
// A request of n substeps is thus active for n consecutive rounds, and its scene's device
// work is exactly that of sp_record(n) + sp_wait of the scene alone; requests that arrive
// join at the next round (iteration-level continuous batching). Evictions and restores
// happen between rounds, with no launch in flight.
//
// Backend interface (called by the driving thread only, one round at a time):
//   uint64_t bytes(uint32_t scene); bool resident(uint32_t scene);
//   void evict(uint32_t scene); void restore(uint32_t scene);
//   uint64_t key(uint32_t scene);
//   int begin(uint32_t scene, uint32_t substeps, float target);  // 0, or the request fails with it
//   void launch(int stream, const std::vector<uint32_t>& scenes,    // one substep of each; `last`:
//               const std::vector<char>& last);                      //   its request's final one
//   void sync();                                                    // the round's launches ended
//   uint64_t finish(uint32_t scene, uint32_t substeps);             // read back; particle-steps
// A thrown std::exception fails every request not done, and the pool with it.
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <exception>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

namespace pool {

enum : int { REQ_PENDING = 1 };

// Done requests whose record is kept for get(); older ones are dropped (their tickets then
// read as unknown).
constexpr size_t KEEP_DONE = 1u << 20;

struct Req {
    uint64_t ticket = 0;
    uint32_t scene = 0, substeps = 0;
    float target = 0.0f;
    int status = REQ_PENDING;  // else 0 (done) or the error code
    uint32_t tanks = 0;        // the largest launch group it ran in
    double submit_ms = 0, start_ms = 0, done_ms = 0;
};

struct Stats {
    uint64_t submitted = 0, completed = 0, failed = 0, particle_steps = 0, launches = 0, rounds = 0;
    uint64_t evictions = 0, restores = 0, resident_bytes = 0;
    uint32_t scenes = 0, resident = 0;
    double p50_ms = 0, p99_ms = 0, mean_ms = 0, max_ms = 0, span_s = 0, tanks_per_launch = 0;
};

template <class Backend>
class Core {
public:
    Core(Backend& b, uint64_t budget, uint32_t max_group, int streams, int error_code)
        : be_(b), budget_(budget), max_group_(std::max<uint32_t>(max_group, 1)), streams_(std::max(streams, 1)),
          error_code_(error_code), t0_(std::chrono::steady_clock::now()) {}

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
    }
    uint64_t budget() const { return budget_; }
    uint32_t scenes_added() {
        std::lock_guard<std::mutex> g(mu_);
        return (uint32_t)scenes_.size();
    }
    std::string error() {
        std::lock_guard<std::mutex> g(mu_);
        return err_;
    }

    // A new scene: `enter(id)` hands it to the backend while no round runs. False once the
    // pool failed, or when `enter` refuses.
    template <class F>
    bool add(uint32_t* id, F&& enter) {
        std::lock_guard<std::mutex> d(drive_);
        std::lock_guard<std::mutex> g(mu_);
        if (failed_) return false;
        const uint32_t k = (uint32_t)scenes_.size();
        if (!enter(k)) return false;
        scenes_.push_back(Scene{});
        *id = k;
        return true;
    }

    // Queues a request; never waits for a round. 0 ok, -1 bad scene, -2 the pool failed.
    int submit(uint32_t scene, uint32_t substeps, float target, uint64_t* ticket) {
        {
            std::lock_guard<std::mutex> g(mu_);
            if (scene >= scenes_.size()) return -1;
            if (failed_) return -2;
            Req q;
            q.ticket = next_ticket_++;
            q.scene = scene;
            q.substeps = substeps;
            q.target = target;
            q.submit_ms = now_ms();
            if (first_submit_ < 0.0) first_submit_ = q.submit_ms;
            reqs_[q.ticket] = q;
            incoming_.push_back(q.ticket);
            st_.submitted++;
            outstanding_++;
            *ticket = q.ticket;
        }
        cv_work_.notify_all();
        return 0;
    }

    // One round. With nothing to run it first waits up to `wait_ms` for a request. Returns
    // the scenes the round advanced (0: nothing to run), or -2 if the pool has failed.
    int step(double wait_ms) {
        std::lock_guard<std::mutex> d(drive_);
        return round(wait_ms);
    }

    // The request's state; false for an unknown ticket. With `block`, drives rounds (or
    // waits for the thread that drives them) until it is done.
    bool get(uint64_t ticket, bool block, Req* out) {
        for (;;) {
            {
                std::unique_lock<std::mutex> g(mu_);
                auto it = reqs_.find(ticket);
                if (it == reqs_.end()) return false;
                if (!block || it->second.status != REQ_PENDING) {
                    *out = it->second;
                    return true;
                }
            }
            drive_or_wait([&] {
                auto it = reqs_.find(ticket);
                return it == reqs_.end() || it->second.status != REQ_PENDING;
            });
        }
    }

    // Drives rounds until every submitted request is done. False if the pool failed.
    bool drain() {
        for (;;) {
            {
                std::lock_guard<std::mutex> g(mu_);
                if (outstanding_ == 0) return !failed_;
            }
            drive_or_wait([&] { return outstanding_ == 0; });
        }
    }

    Stats stats() {
        std::lock_guard<std::mutex> g(mu_);
        Stats s = st_;
        s.scenes = (uint32_t)scenes_.size();
        s.resident_bytes = resident_bytes_pub_;
        s.resident = resident_pub_;
        s.tanks_per_launch = s.launches ? (double)launched_tanks_ / (double)s.launches : 0.0;
        if (!lat_.empty()) {
            std::vector<double> l = lat_;
            std::sort(l.begin(), l.end());
            auto pct = [&](double p) {  // nearest rank
                size_t r = (size_t)std::ceil(p / 100.0 * (double)l.size());
                return l[std::min(std::max<size_t>(r, 1), l.size()) - 1];
            };
            s.p50_ms = pct(50.0);
            s.p99_ms = pct(99.0);
            double sum = 0;
            for (double x : l) sum += x;
            s.mean_ms = sum / (double)l.size();
            s.max_ms = l.back();
            s.span_s = first_submit_ >= 0.0 ? (last_done_ - first_submit_) * 1e-3 : 0.0;
        }
        return s;
    }

    void reset_stats() {
        std::lock_guard<std::mutex> g(mu_);
        st_ = Stats{};
        launched_tanks_ = 0;
        lat_.clear();
        first_submit_ = -1.0;
        last_done_ = 0.0;
    }

private:
    struct Scene {
        std::deque<uint64_t> queue;  // tickets not started, arrival order
        uint64_t active = 0;         // the ticket running, 0: idle
        uint32_t left = 0;           // its substeps still to run
        uint64_t last_used = 0;
        uint64_t start_seq = 0;      // when its active request started
    };

    // Runs one round if no other thread drives; otherwise waits a little for it.
    template <class Done>
    void drive_or_wait(Done&& done) {
        std::unique_lock<std::mutex> d(drive_, std::try_to_lock);
        if (d.owns_lock()) {
            {
                std::lock_guard<std::mutex> g(mu_);
                if (done()) return;
            }
            if (round(0.0) == 0) {
                // nothing could run although a request is pending: cannot happen with
                // every scene within the budget, but never wait for it forever
                std::lock_guard<std::mutex> g(mu_);
                if (!done()) fail_locked("a pending request could not be started");
            }
            return;
        }
        std::unique_lock<std::mutex> g(mu_);
        cv_done_.wait_for(g, std::chrono::milliseconds(1), [&] { return done(); });
    }

    int round(double wait_ms) {
        std::vector<uint64_t> arrived;
        {
            std::unique_lock<std::mutex> g(mu_);
            if (failed_) return -2;
            if (incoming_.empty() && !any_work() && wait_ms > 0.0)
                cv_work_.wait_for(g, std::chrono::duration<double, std::milli>(wait_ms),
                                  [&] { return !incoming_.empty() || failed_; });
            if (failed_) return -2;
            arrived.swap(incoming_);
            for (uint64_t tk : arrived) scenes_[reqs_[tk].scene].queue.push_back(tk);
        }
        try {
            return run_round();
        } catch (const std::exception& e) {
            std::lock_guard<std::mutex> g(mu_);
            fail_locked(e.what());
            return -2;
        } catch (...) {
            std::lock_guard<std::mutex> g(mu_);
            fail_locked("unexpected error in the pool");
            return -2;
        }
    }

    // Scenes with something to run (under mu_, or from the driving thread).
    bool any_work() const {
        for (const Scene& sc : scenes_)
            if (sc.active || !sc.queue.empty()) return true;
        return false;
    }

    uint64_t resident_total() {
        uint64_t b = 0;
        for (uint32_t k = 0; k < scenes_.size(); k++)
            if (be_.resident(k)) b += be_.bytes(k);
        return b;
    }

    // Evicts idle resident scenes until `need` more bytes fit; false if they cannot.
    bool make_room(uint64_t need) {
        uint64_t total = resident_total();
        while (total + need > budget_) {
            int victim = -1;
            for (uint32_t k = 0; k < scenes_.size(); k++) {
                const Scene& sc = scenes_[k];
                if (sc.active || !be_.resident(k)) continue;
                if (victim < 0) {
                    victim = (int)k;
                    continue;
                }
                const Scene& v = scenes_[(size_t)victim];
                const bool a = sc.queue.empty(), b = v.queue.empty();
                if (a != b ? a : sc.last_used < v.last_used) victim = (int)k;
            }
            if (victim < 0) return false;
            total -= be_.bytes((uint32_t)victim);
            be_.evict((uint32_t)victim);
            bump(&Stats::evictions);
        }
        return true;
    }

    void bump(uint64_t Stats::*f) {
        std::lock_guard<std::mutex> g(mu_);
        st_.*f += 1;
    }

    int run_round() {
        const uint32_t n = (uint32_t)scenes_.size();
        (void)make_room(0);  // over the budget (scenes arrive resident): idle ones go first
        // start requests, oldest first
        std::vector<std::pair<double, uint32_t>> ready;
        {
            std::lock_guard<std::mutex> g(mu_);
            for (uint32_t k = 0; k < n; k++)
                if (!scenes_[k].active && !scenes_[k].queue.empty())
                    ready.push_back({reqs_[scenes_[k].queue.front()].submit_ms, k});
        }
        std::sort(ready.begin(), ready.end());
        bool blocked = false;  // an older scene is waiting for memory
        for (const auto& r : ready) {
            const uint32_t k = r.second;
            Scene& sc = scenes_[k];
            if (!be_.resident(k)) {
                if (blocked || !make_room(be_.bytes(k))) {
                    blocked = true;
                    continue;
                }
                be_.restore(k);
                bump(&Stats::restores);
            }
            uint64_t tk;
            Req q;
            {
                std::lock_guard<std::mutex> g(mu_);
                tk = sc.queue.front();
                sc.queue.pop_front();
                q = reqs_[tk];
            }
            const int rc = be_.begin(k, q.substeps, q.target);
            if (rc != 0) {
                std::lock_guard<std::mutex> g(mu_);
                complete_locked(tk, rc, 0);
                continue;
            }
            sc.active = tk;
            sc.left = q.substeps;
            sc.start_seq = ++seq_;
            std::lock_guard<std::mutex> g(mu_);
            reqs_[tk].start_ms = now_ms();
        }
        // the round's groups: active scenes by key, in the order their requests started
        std::vector<uint32_t> active;
        for (uint32_t k = 0; k < n; k++)
            if (scenes_[k].active) active.push_back(k);
        publish_residency();
        if (active.empty()) return 0;
        std::sort(active.begin(), active.end(),
                  [&](uint32_t a, uint32_t b) { return scenes_[a].start_seq < scenes_[b].start_seq; });
        std::map<uint64_t, std::vector<uint32_t>> by_key;
        std::vector<uint64_t> key_order;
        for (uint32_t k : active) {
            const uint64_t key = be_.key(k);
            auto it = by_key.find(key);
            if (it == by_key.end()) {
                key_order.push_back(key);
                it = by_key.emplace(key, std::vector<uint32_t>()).first;
            }
            it->second.push_back(k);
        }
        std::vector<std::vector<uint32_t>> groups;
        for (uint64_t key : key_order) {
            const std::vector<uint32_t>& ks = by_key[key];
            for (size_t a = 0; a < ks.size(); a += max_group_)
                groups.emplace_back(ks.begin() + (long)a, ks.begin() + (long)std::min(ks.size(), a + max_group_));
        }
        for (size_t gi = 0; gi < groups.size(); gi++) {
            std::vector<char> last(groups[gi].size());
            for (size_t i = 0; i < last.size(); i++) last[i] = scenes_[groups[gi][i]].left == 1 ? 1 : 0;
            be_.launch((int)(gi % (size_t)streams_), groups[gi], last);
        }
        be_.sync();
        {
            std::lock_guard<std::mutex> g(mu_);
            st_.rounds++;
            st_.launches += groups.size();
            for (const auto& grp : groups) {
                launched_tanks_ += grp.size();
                for (uint32_t k : grp) reqs_[scenes_[k].active].tanks = std::max(reqs_[scenes_[k].active].tanks, (uint32_t)grp.size());
            }
        }
        // completions
        for (uint32_t k : active) {
            Scene& sc = scenes_[k];
            sc.last_used = ++use_clock_;
            if (--sc.left > 0) continue;
            uint32_t substeps;
            {
                std::lock_guard<std::mutex> g(mu_);
                substeps = reqs_[sc.active].substeps;
            }
            const uint64_t steps = be_.finish(k, substeps);
            std::lock_guard<std::mutex> g(mu_);
            complete_locked(sc.active, 0, steps);
            sc.active = 0;
        }
        cv_done_.notify_all();
        return (int)active.size();
    }

    void complete_locked(uint64_t tk, int status, uint64_t steps) {
        Req& q = reqs_[tk];
        q.status = status;
        q.done_ms = now_ms();
        outstanding_--;
        note_done_locked(tk);
        if (status == 0) {
            st_.completed++;
            st_.particle_steps += steps;
            lat_.push_back(q.done_ms - q.submit_ms);
            last_done_ = q.done_ms;
        } else {
            st_.failed++;
        }
        cv_done_.notify_all();
    }

    // Fails every request not done; the pool takes no more.
    void fail_locked(const std::string& why) {
        failed_ = true;
        if (err_.empty()) err_ = why;
        std::vector<uint64_t> failed;
        for (auto& kv : reqs_)
            if (kv.second.status == REQ_PENDING) {
                kv.second.status = error_code_;
                kv.second.done_ms = now_ms();
                st_.failed++;
                outstanding_--;
                failed.push_back(kv.first);
            }
        std::sort(failed.begin(), failed.end());
        for (uint64_t tk : failed) note_done_locked(tk);
        for (Scene& sc : scenes_) {
            sc.queue.clear();
            sc.active = 0;
        }
        incoming_.clear();
        cv_done_.notify_all();
        cv_work_.notify_all();
    }

    void note_done_locked(uint64_t tk) {
        done_order_.push_back(tk);
        while (done_order_.size() > KEEP_DONE) {
            reqs_.erase(done_order_.front());
            done_order_.pop_front();
        }
    }

    void publish_residency() {
        uint64_t b = 0;
        uint32_t r = 0;
        for (uint32_t k = 0; k < scenes_.size(); k++)
            if (be_.resident(k)) {
                b += be_.bytes(k);
                r++;
            }
        std::lock_guard<std::mutex> g(mu_);
        resident_bytes_pub_ = b;
        resident_pub_ = r;
    }

    Backend& be_;
    const uint64_t budget_;
    const uint32_t max_group_;
    const int streams_;
    const int error_code_;
    const std::chrono::steady_clock::time_point t0_;

    std::mutex drive_;  // held by the thread running a round (and by add)
    std::mutex mu_;     // requests, queues, statistics
    std::condition_variable cv_work_, cv_done_;
    std::vector<Scene> scenes_;
    std::vector<uint64_t> incoming_;
    std::unordered_map<uint64_t, Req> reqs_;
    std::deque<uint64_t> done_order_;
    uint64_t next_ticket_ = 1;
    uint64_t outstanding_ = 0;
    uint64_t use_clock_ = 0, seq_ = 0;
    bool failed_ = false;
    std::string err_;
    Stats st_;
    uint64_t launched_tanks_ = 0;
    std::vector<double> lat_;
    double first_submit_ = -1.0, last_done_ = 0.0;
    uint64_t resident_bytes_pub_ = 0;
    uint32_t resident_pub_ = 0;
};

}  // namespace pool
