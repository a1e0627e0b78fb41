// serve_core.h — the request server's scheduler, with no device code: requests, scene
// queues, admission under the memory budget, launch groups and completions. The device side
// is a Backend the server supplies (serve.hip: HIP launches on the tanks); tests drive the
// same scheduler with a backend that only keeps time.
//
// One worker thread owns the scenes. Each turn of its loop it
//   1. takes the requests and scenes that arrived (under the lock; callers never wait on it
//      for long),
//   2. finishes the launches the backend reports done,
//   3. launches while a slot is free: the oldest waiting request whose scene is idle picks
//      the group (same key, same substep count), the other idle scenes with such a request
//      at the head of their queue join it (oldest first, at most `max_group`), every member
//      is made resident within the budget (evicting the least recently used idle scenes that
//      no launch holds), and the backend records the group on the slot,
//   4. sleeps until something arrives or, with launches in flight, for a few microseconds.
// A scene has at most one launch in flight, so its requests run in arrival order.
//
// Backend interface (all called from the worker thread only):
//   uint64_t bytes(uint32_t scene); bool resident(uint32_t scene);
//   void evict(uint32_t scene); void restore(uint32_t scene);
//   uint64_t key(uint32_t scene);
//   void launch(int slot, const std::vector<uint32_t>& scenes, uint32_t substeps,
//               const std::vector<float>& targets);
//   bool done(int slot);   // the slot's launch has finished on the device
//   uint64_t finish(int slot, uint32_t scene, uint32_t substeps);  // fold in; particle-steps
//   void drained();        // the worker is about to stop: wait for the device
// Errors are thrown (std::exception); the worker then fails every request not done.
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <exception>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace serve {

enum : int { REQ_PENDING = 1 };

struct Req {
    uint64_t ticket = 0;
    uint32_t scene = 0, substeps = 0;
    float target = 0.0f;
    int status = REQ_PENDING;  // else 0 (done) or the error code
    uint32_t tanks = 0;
    double submit_ms = 0, launch_ms = 0, done_ms = 0;
};

struct Stats {
    uint64_t submitted = 0, completed = 0, failed = 0, particle_steps = 0, launches = 0;
    uint64_t evictions = 0, restores = 0, resident_bytes = 0;
    uint32_t scenes = 0, resident = 0;
    double p50_ms = 0, p99_ms = 0, mean_ms = 0, max_ms = 0, span_s = 0, tanks_per_launch = 0;
};

template <class Backend>
class Core {
public:
    Core(Backend& b, uint64_t budget, int slots, uint32_t max_group, int error_code)
        : be_(b), budget_(budget), n_slots_(slots), max_group_(max_group), error_code_(error_code),
          t0_(std::chrono::steady_clock::now()), slots_((size_t)slots) {}

    ~Core() { stop(); }

    void start() { worker_ = std::thread([this] { loop(); }); }

    // Runs what was submitted, then stops the worker.
    void stop() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_work_.notify_all();
        if (worker_.joinable()) worker_.join();
    }

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
    }

    // A new scene: `enter(id)` hands the scene to the backend under the lock, before the
    // worker can see it. False once the server failed or stops, or when `enter` refuses.
    template <class F>
    bool add(uint32_t* id, F&& enter) {
        std::lock_guard<std::mutex> g(mu_);
        if (failed_ || stop_) return false;
        if (!enter(n_scenes_)) return false;
        *id = n_scenes_++;
        cv_work_.notify_all();
        return true;
    }

    // Queues a request. 0 ok, -1 bad scene, -2 failed or stopping.
    int submit(uint32_t scene, uint32_t substeps, float target, uint64_t* ticket) {
        std::lock_guard<std::mutex> g(mu_);
        if (scene >= n_scenes_) return -1;
        if (failed_ || stop_) return -2;
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
        cv_work_.notify_all();
        return 0;
    }

    // The request, waiting for it to finish when `block`. False for an unknown ticket.
    bool get(uint64_t ticket, bool block, Req* out) {
        std::unique_lock<std::mutex> lk(mu_);
        if (reqs_.find(ticket) == reqs_.end()) return false;
        if (block)
            cv_done_.wait(lk, [&] { return reqs_.find(ticket)->second.status != REQ_PENDING; });
        *out = reqs_.find(ticket)->second;
        return true;
    }

    // Waits until no request is outstanding; false if the server failed.
    bool drain() {
        std::unique_lock<std::mutex> lk(mu_);
        cv_done_.wait(lk, [&] { return outstanding_ == 0; });
        return !failed_;
    }

    std::string error() {
        std::lock_guard<std::mutex> g(mu_);
        return err_;
    }

    Stats stats() {
        std::lock_guard<std::mutex> g(mu_);
        Stats s = st_;
        s.evictions = evictions_.load();
        s.restores = restores_.load();
        s.resident_bytes = resident_bytes_pub_;
        s.resident = resident_pub_;
        s.scenes = n_scenes_;
        s.tanks_per_launch = st_.launches ? (double)launched_tanks_ / (double)st_.launches : 0.0;
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
        lat_.clear();
        st_ = Stats{};
        launched_tanks_ = 0;
        evictions_ = 0;
        restores_ = 0;
        first_submit_ = -1.0;
        last_done_ = 0.0;
    }

    uint64_t budget() const { return budget_; }

    uint32_t scenes_added() {
        std::lock_guard<std::mutex> g(mu_);
        return n_scenes_;
    }

private:
    struct Scene {
        std::deque<uint64_t> queue;  // tickets, arrival order
        bool busy = false;           // a launch holds it
        uint64_t last_used = 0;
    };
    struct Slot {
        bool busy = false;
        std::vector<uint32_t> scenes;
        std::vector<uint64_t> tickets;
        uint32_t substeps = 0;
    };

    // Evicts least recently used idle scenes outside `keep` until `need` more bytes fit.
    bool make_room(uint64_t need, const std::vector<uint32_t>& keep) {
        while (resident_bytes_ + need > budget_) {
            int victim = -1;
            for (size_t k = 0; k < scenes_.size(); k++) {
                const Scene& sc = scenes_[k];
                if (!be_.resident((uint32_t)k) || sc.busy) continue;
                if (std::find(keep.begin(), keep.end(), (uint32_t)k) != keep.end()) continue;
                if (victim < 0) {
                    victim = (int)k;
                    continue;
                }
                // idle scenes with nothing queued go first, then the least recently used
                const Scene& vb = scenes_[victim];
                const bool a_idle = sc.queue.empty(), b_idle = vb.queue.empty();
                if (a_idle != b_idle ? a_idle : sc.last_used < vb.last_used) victim = (int)k;
            }
            if (victim < 0) return false;
            be_.evict((uint32_t)victim);
            resident_bytes_ -= be_.bytes((uint32_t)victim);
            evictions_++;
        }
        return true;
    }

    bool admit(uint32_t k, const std::vector<uint32_t>& keep) {
        if (be_.resident(k)) return true;
        if (!make_room(be_.bytes(k), keep)) return false;
        be_.restore(k);
        resident_bytes_ += be_.bytes(k);
        restores_++;
        return true;
    }

    void finish_slot(int si) {
        Slot& sl = slots_[si];
        const double t = now_ms();
        std::vector<uint64_t> psteps(sl.scenes.size());
        for (size_t k = 0; k < sl.scenes.size(); k++) {
            psteps[k] = be_.finish(si, sl.scenes[k], sl.substeps);
            scenes_[sl.scenes[k]].busy = false;
        }
        {
            std::lock_guard<std::mutex> g(mu_);
            for (size_t k = 0; k < sl.tickets.size(); k++) {
                Req& q = reqs_.find(sl.tickets[k])->second;
                q.status = 0;
                q.done_ms = t;
                lat_.push_back(t - q.submit_ms);
                st_.completed++;
                st_.particle_steps += psteps[k];
                outstanding_--;
            }
            last_done_ = t;
        }
        cv_done_.notify_all();
        sl.busy = false;
        sl.scenes.clear();
        sl.tickets.clear();
    }

    void fail_all(const std::string& why) {
        {
            std::lock_guard<std::mutex> g(mu_);
            failed_ = true;
            err_ = why;
            for (auto& kv : reqs_)
                if (kv.second.status == REQ_PENDING) {
                    kv.second.status = error_code_;
                    kv.second.done_ms = now_ms();
                    st_.failed++;
                    outstanding_--;
                }
            incoming_.clear();
        }
        cv_done_.notify_all();
    }

    // The substep count of scene k's oldest request (its queue is not empty).
    uint32_t head_substeps(uint32_t k) {
        std::lock_guard<std::mutex> g(mu_);
        return reqs_.find(scenes_[k].queue.front())->second.substeps;
    }

    // Launches while a slot is free and an idle scene has a request.
    bool launch_ready() {
        bool any = false;
        for (;;) {
            int si = -1;
            for (int k = 0; k < n_slots_; k++)
                if (!slots_[k].busy) {
                    si = k;
                    break;
                }
            if (si < 0) return any;
            std::vector<std::pair<double, uint32_t>> ready;  // idle scenes with a request, oldest head first
            {
                std::lock_guard<std::mutex> g(mu_);
                for (uint32_t k = 0; k < scenes_.size(); k++) {
                    const Scene& sc = scenes_[k];
                    if (sc.busy || sc.queue.empty()) continue;
                    ready.push_back({reqs_.find(sc.queue.front())->second.submit_ms, k});
                }
            }
            if (ready.empty()) return any;
            std::sort(ready.begin(), ready.end());
            bool launched = false;
            std::vector<bool> tried(scenes_.size(), false);
            for (size_t a = 0; a < ready.size() && !launched; a++) {
                const uint32_t lead = ready[a].second;
                if (tried[lead]) continue;
                const uint32_t substeps = head_substeps(lead);
                const uint64_t key = be_.key(lead);
                std::vector<uint32_t> members;
                for (size_t b = a; b < ready.size() && members.size() < max_group_; b++) {
                    const uint32_t k = ready[b].second;
                    if (tried[k] || be_.key(k) != key || head_substeps(k) != substeps) continue;
                    tried[k] = true;
                    if (admit(k, members)) members.push_back(k);
                }
                if (members.empty()) continue;
                std::vector<uint64_t> tickets;
                std::vector<float> targets;
                const double t = now_ms();
                {
                    std::lock_guard<std::mutex> g(mu_);
                    for (uint32_t m : members) {
                        const uint64_t tk = scenes_[m].queue.front();
                        scenes_[m].queue.pop_front();
                        Req& q = reqs_.find(tk)->second;
                        q.launch_ms = t;
                        q.tanks = (uint32_t)members.size();
                        tickets.push_back(tk);
                        targets.push_back(q.target);
                    }
                    st_.launches++;
                    launched_tanks_ += members.size();
                }
                be_.launch(si, members, substeps, targets);
                Slot& sl = slots_[si];
                sl.busy = true;
                sl.scenes = members;
                sl.tickets = tickets;
                sl.substeps = substeps;
                for (uint32_t m : members) {
                    scenes_[m].busy = true;
                    scenes_[m].last_used = ++use_clock_;
                }
                launched = true;
                any = true;
            }
            if (!launched) return any;
        }
    }

    void publish() {
        uint32_t n = 0;
        for (uint32_t k = 0; k < scenes_.size(); k++) n += be_.resident(k) ? 1u : 0u;
        std::lock_guard<std::mutex> g(mu_);
        resident_bytes_pub_ = resident_bytes_;
        resident_pub_ = n;
    }

    void loop() {
        try {
            for (;;) {
                bool stopping;
                {
                    std::lock_guard<std::mutex> g(mu_);
                    while (scenes_.size() < n_scenes_) scenes_.push_back(Scene{});
                    for (uint64_t tk : incoming_) scenes_[reqs_.find(tk)->second.scene].queue.push_back(tk);
                    incoming_.clear();
                    stopping = stop_;
                }
                // scenes arrive resident: over the budget, the idle ones go first
                uint64_t bytes = 0;
                for (uint32_t k = 0; k < scenes_.size(); k++)
                    if (be_.resident(k)) bytes += be_.bytes(k);
                resident_bytes_ = bytes;
                if (bytes > budget_) (void)make_room(0, {});
                bool progressed = false;
                for (int k = 0; k < n_slots_; k++)
                    if (slots_[k].busy && be_.done(k)) {
                        finish_slot(k);
                        progressed = true;
                    }
                if (launch_ready()) progressed = true;
                publish();
                bool in_flight = false;
                for (int k = 0; k < n_slots_; k++) in_flight = in_flight || slots_[k].busy;
                if (progressed) continue;
                std::unique_lock<std::mutex> lk(mu_);
                if (stopping && outstanding_ == 0 && !in_flight && incoming_.empty()) break;
                if (!incoming_.empty() || scenes_.size() < n_scenes_) continue;
                if (in_flight) {
                    cv_work_.wait_for(lk, std::chrono::microseconds(20));
                } else {
                    // nothing runs: either nothing is queued, or a queued scene can be launched
                    // now (a slot is free and every resident scene is idle), so wait only
                    // when no scene holds a request
                    bool queued = false;
                    for (const Scene& sc : scenes_) queued = queued || !sc.queue.empty();
                    if (queued) continue;
                    cv_work_.wait(lk, [&] { return stop_ || !incoming_.empty() || scenes_.size() < n_scenes_; });
                }
            }
        } catch (const std::exception& e) {
            fail_all(e.what());
        } catch (...) {
            fail_all("unexpected error in the server");
        }
        try {
            be_.drained();
        } catch (...) {
        }
    }

    Backend& be_;
    const uint64_t budget_;
    const int n_slots_;
    const uint32_t max_group_;
    const int error_code_;
    const std::chrono::steady_clock::time_point t0_;

    // shared with callers, under mu_
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<uint64_t> incoming_;
    std::unordered_map<uint64_t, Req> reqs_;
    uint64_t next_ticket_ = 1;
    uint32_t n_scenes_ = 0;
    uint64_t outstanding_ = 0;
    bool stop_ = false;
    bool failed_ = false;
    std::string err_;
    Stats st_;
    std::vector<double> lat_;
    uint64_t launched_tanks_ = 0;
    double first_submit_ = -1.0, last_done_ = 0.0;
    uint64_t resident_bytes_pub_ = 0;
    uint32_t resident_pub_ = 0;
    std::atomic<uint64_t> evictions_{0}, restores_{0};

    // the worker's own
    std::vector<Slot> slots_;
    std::vector<Scene> scenes_;
    uint64_t resident_bytes_ = 0;
    uint64_t use_clock_ = 0;
    std::thread worker_;
};

}  // namespace serve
