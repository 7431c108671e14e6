// qpso_sched.h -- the serving loop's scheduler: continuous batching of pose
// solves under an HBM budget and a particle-slot limit, independent of the
// device.  No counterpart in the reference, which solves one cluster per
// process (README.md:26-29).
//
// Server<Dev> owns the request queue, admission, the lanes of resident
// solves, cancellation and time-outs, and the ticket bookkeeping; Dev does
// the device work.  The library instantiates it with its HIP device
// (qpso_api.cpp); the CPU tests instantiate it with a mock device (virtual
// clock, fake allocator, injected failures).  Dev provides:
//
//   kLanes                     lanes of resident solves (queues of the device
//                              that run side by side)
//   typename Solve, Job        a resident solve; what a request carries
//   double now()               seconds on the clock latencies are measured with
//   int  create(Job&, Solve*&) allocate and upload one solve (status code)
//   int  init(Solve*, lane)    enqueue its swarm initialisation on the lane
//   bool ready(lane)           a step can be enqueued without waiting
//   int  step(lane, solves, changed, total_particles)
//                              enqueue one iteration of every solve of the lane
//                              (`changed`: the lane's set differs from the last
//                              step; `total_particles`: over every lane)
//   int  mark(lane, Solve*)    after the solve's last step: note where its work ends
//   int  done(Solve*)          1 once that work has completed, 0 before, < 0: error
//   int  read(Solve*, qpso_result*)   the finished solve's result (after done)
//   int  sync(lane)            wait until the lane's queued work is done
//   void destroy(Solve*, lane) free it once the work queued on the lane (all
//                              the work that uses it) is done, without waiting
//   void idle_wait()           wait a little for the device (nothing else to do)
//   void on_thread_start(), on_thread_end()   hooks of the server thread
//
// Requests join lanes by size (choose_lane): like sizes share the lanes, far
// lighter ones keep to a lane of their own.  Each lane is stepped as soon as
// it can take a step, independently of the others, so light solves advance
// at their own step time rather than at that of the largest resident ones.
// A finished solve's result is posted once its own last step has completed;
// its memory is released at once, in the order of its lane's queue, and its
// charge with it.
//
// Invariants kept (and checked by the mock-device tests): requests are
// admitted in arrival order; the device bytes charged (solve arenas plus
// each random table once while any resident solve uses it) never exceed the
// budget; resident particles never exceed the slots except for
// one solve larger than the slots, which then runs alone; every ticket ends
// with exactly one result, read exactly once; nothing is read before the
// work it depends on completed, nor freed except in its lane's order.
#ifndef QPSO_SCHED_H_
#define QPSO_SCHED_H_

#include <stdint.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

#include "qpso.h"

namespace qpso_sched {

// A random table is shared by the solves with equal (seed, particles, stride).
using TableKey = std::tuple<uint32_t, int32_t, int32_t>;

template <class Dev>
class Server {
   public:
    using Solve = typename Dev::Solve;
    using Job = typename Dev::Job;
    static constexpr int kLanes = Dev::kLanes;
    // requests whose step work differs by more than this factor run in
    // different lanes where they can
    static constexpr int64_t kRatio = 64;

    struct Request {
        Job job;
        uint64_t bytes = 0;        // the solve's own device bytes
        TableKey table{};          // its random table
        uint64_t table_bytes = 0;  // charged once while any resident solve uses the table
        int64_t particles = 0;
        int64_t cost = 0;          // work of one step: particles x faces rendered
        int32_t iters = 0;
        uint64_t ticket = 0;
        double t_submit = 0.0;
        double deadline = 0.0;     // 0: none
    };

    Server(Dev& dev, uint64_t budget) : dev_(dev), budget_(budget) {}
    ~Server() { shutdown(); }
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    // Run the loop on a thread of its own (the library); tests may instead
    // call tick() themselves.
    void start() {
        th_ = std::thread([this] {
            dev_.on_thread_start();
            run();
            dev_.on_thread_end();
        });
    }

    // Ends every request that has no result yet with QPSO_E_CANCELLED (queued
    // ones at once, resident ones at the next step boundary), stops the loop,
    // and returns once no thread is blocked in wait().  Calls that start after
    // shutdown() began are the caller's error.
    void shutdown() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (down_) return;
            stop_ = true;
            for (auto& r : pending_) post_locked(r.ticket, result_with(QPSO_E_CANCELLED), kCancelled, r.t_submit);
            pending_.clear();
            for (uint64_t t : live_)
                if (!results_.count(t)) cancel_.insert(t);
        }
        cv_.notify_all();
        if (th_.joinable()) th_.join();
        else drain_inline();
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return waiters_ == 0; });
        down_ = true;
    }

    int set_slots(int64_t max_particles) {
        if (max_particles <= 0) return QPSO_E_INVAL;
        std::lock_guard<std::mutex> lk(mu_);
        max_particles_ = max_particles;
        cv_.notify_all();
        return QPSO_OK;
    }

    int set_timeout(double seconds) {
        if (!(seconds >= 0.0)) return QPSO_E_INVAL;
        std::lock_guard<std::mutex> lk(mu_);
        timeout_s_ = seconds;
        return QPSO_OK;
    }

    int submit(Request&& r, uint64_t* ticket) {
        std::lock_guard<std::mutex> lk(mu_);
        if (stop_) return QPSO_E_INVAL;
        // a request that could never be admitted is refused now
        if (r.bytes + r.table_bytes > budget_) {
            ++rejected_;
            return QPSO_E_BUDGET;
        }
        r.t_submit = dev_.now();
        r.deadline = timeout_s_ > 0.0 ? r.t_submit + timeout_s_ : 0.0;
        r.ticket = next_ticket_++;
        *ticket = r.ticket;
        if (!have_first_) {
            t_first_ = r.t_submit;
            have_first_ = true;
        }
        ++submitted_;
        live_.insert(r.ticket);
        pending_.push_back(std::move(r));
        cv_.notify_all();
        return QPSO_OK;
    }

    // QPSO_E_NOTFOUND: unknown, or already read; QPSO_E_PENDING: not finished.
    int poll(uint64_t ticket, qpso_result* out) {
        std::lock_guard<std::mutex> lk(mu_);
        if (!live_.count(ticket)) return QPSO_E_NOTFOUND;
        if (!results_.count(ticket)) return QPSO_E_PENDING;
        return consume_locked(ticket, out);
    }

    int wait(uint64_t ticket, qpso_result* out) {
        std::unique_lock<std::mutex> lk(mu_);
        if (!live_.count(ticket)) return QPSO_E_NOTFOUND;
        ++waiters_;
        done_cv_.wait(lk, [&] { return results_.count(ticket) != 0 || !live_.count(ticket); });
        --waiters_;
        const int rc = results_.count(ticket) ? consume_locked(ticket, out) : QPSO_E_NOTFOUND;
        done_cv_.notify_all();  // shutdown() waits for the last waiter
        return rc;
    }

    // QPSO_OK: the request ends with QPSO_E_CANCELLED (queued: now; resident:
    // at the next step boundary), or had already finished and keeps its
    // result.  QPSO_E_NOTFOUND: unknown or already read.
    int cancel(uint64_t ticket) {
        std::lock_guard<std::mutex> lk(mu_);
        if (!live_.count(ticket)) return QPSO_E_NOTFOUND;
        if (results_.count(ticket)) return QPSO_OK;
        for (auto it = pending_.begin(); it != pending_.end(); ++it) {
            if (it->ticket != ticket) continue;
            post_locked(ticket, result_with(QPSO_E_CANCELLED), kCancelled, it->t_submit);
            pending_.erase(it);
            return QPSO_OK;
        }
        cancel_.insert(ticket);
        cv_.notify_all();
        return QPSO_OK;
    }

    void stats(qpso_server_stats* st) {
        std::lock_guard<std::mutex> lk(mu_);
        std::memset(st, 0, sizeof(*st));
        st->submitted = submitted_;
        st->completed = completed_;
        st->rejected = rejected_;
        st->cancelled = cancelled_;
        st->timed_out = timed_out_;
        st->steps = steps_;
        st->mean_active = steps_ ? active_sum_ / (double)steps_ : 0.0;
        if (!latencies_ms_.empty()) {
            std::vector<double> v = latencies_ms_;
            std::sort(v.begin(), v.end());
            auto q = [&](double f) { return v[std::min(v.size() - 1, (size_t)(f * (double)(v.size() - 1) + 0.5))]; };
            st->p50_ms = q(0.50);
            st->p99_ms = q(0.99);
            double s = 0;
            for (double x : v) s += x;
            st->mean_ms = s / (double)v.size();
            const double span = t_last_ - t_first_;
            st->solves_per_s = span > 0 ? (double)completed_ / span : 0.0;
        }
    }

    // Accessors for tests.
    uint64_t used_bytes() { std::lock_guard<std::mutex> lk(mu_); return used_; }
    int64_t resident_particles() { std::lock_guard<std::mutex> lk(mu_); return resident_; }
    int waiters() { std::lock_guard<std::mutex> lk(mu_); return waiters_; }
    size_t resident_solves() { std::lock_guard<std::mutex> lk(mu_); return n_active_; }
    std::vector<uint64_t> admission_order() { std::lock_guard<std::mutex> lk(mu_); return admitted_; }

    // One pass of the loop: expire and admit queued requests, end cancelled
    // and expired resident solves, step every lane that can take a step,
    // post the results of solves whose work has completed (their memory is
    // released in their lane's order as they leave).  Returns false when
    // there is nothing to do.  Only one thread may call it (the server
    // thread, or a test).
    bool tick() {
        std::vector<Request> admit;
        std::unordered_set<uint64_t> cancels;
        bool progress = false;
        const double now = dev_.now();
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (auto it = pending_.begin(); it != pending_.end();) {
                if (it->deadline > 0.0 && now >= it->deadline) {
                    post_locked(it->ticket, result_with(QPSO_E_TIMEOUT), kTimedOut, it->t_submit);
                    it = pending_.erase(it);
                } else {
                    ++it;
                }
            }
            // arrival order while the budget and the slots allow; a solve
            // larger than the slots runs when nothing else is resident
            while (!pending_.empty() && !stop_) {
                const Request& r = pending_.front();
                auto tr = table_refs_.find(r.table);
                const uint64_t need = r.bytes + (tr == table_refs_.end() ? r.table_bytes : 0);
                const bool fits = used_ + need <= budget_ &&
                                  (resident_ + r.particles <= max_particles_ || (resident_ == 0 && admit.empty()));
                if (!fits) break;
                used_ += need;
                resident_ += r.particles;
                ++table_refs_[r.table];
                admitted_.push_back(r.ticket);
                admit.push_back(std::move(pending_.front()));
                pending_.pop_front();
            }
            cancels.swap(cancel_);
        }
        for (auto& r : admit) {
            progress = true;
            Solve* s = nullptr;
            int rc = dev_.create(r.job, s);
            const int lane = choose_lane(r);
            if (rc == QPSO_OK) {
                rc = dev_.init(s, lane);
                if (rc != QPSO_OK) dev_.destroy(s, lane);
            }
            if (rc != QPSO_OK) {
                post(r, result_with(rc), kRejected);
                release(r);
                continue;
            }
            Lane& l = lanes_[lane];
            l.particles += r.particles;
            l.cost += r.cost;
            l.active.push_back(Active{std::move(r), s, 0});
            l.changed = true;
        }
        // cancelled and expired solves leave at this step boundary
        for (int k = 0; k < kLanes; ++k) {
            Lane& l = lanes_[k];
            auto leaving = [&](const Active& a) {
                return cancels.count(a.req.ticket) != 0 || (a.req.deadline > 0.0 && now >= a.req.deadline);
            };
            for (auto* v : {&l.active, &l.finishing}) {
                std::vector<Active> keep;
                for (auto& a : *v) {
                    if (!leaving(a)) {
                        keep.push_back(std::move(a));
                        continue;
                    }
                    const bool cancelled = cancels.count(a.req.ticket) != 0;
                    post(a.req, result_with(cancelled ? QPSO_E_CANCELLED : QPSO_E_TIMEOUT),
                         cancelled ? kCancelled : kTimedOut);
                    leave_lane(k, a, v == &l.active);
                    progress = true;
                }
                if (keep.size() != v->size()) l.changed = true;
                v->swap(keep);
            }
        }
        // step every lane that can take a step now
        size_t stepped = 0;
        int64_t all_particles = 0;  // over every lane: what the device has to run
        for (const auto& l : lanes_) all_particles += l.particles;
        for (int k = 0; k < kLanes; ++k) {
            Lane& l = lanes_[k];
            if (l.active.empty() || !dev_.ready(k)) continue;
            std::vector<Solve*> ss;
            ss.reserve(l.active.size());
            for (auto& a : l.active) ss.push_back(a.s);
            if (const int rc = dev_.step(k, ss, l.changed, all_particles)) {
                fail_lane(k, rc);
                continue;
            }
            progress = true;
            stepped += l.active.size();
            l.changed = false;
            // solves that have run all their iterations wait for their last
            // step to complete
            std::vector<Active> keep;
            int rc = QPSO_OK;
            for (auto& a : l.active) {
                if (++a.iter < a.req.iters) {
                    keep.push_back(std::move(a));
                    continue;
                }
                if (rc == QPSO_OK) rc = dev_.mark(k, a.s);
                l.particles -= a.req.particles;
                l.cost -= a.req.cost;
                l.finishing.push_back(std::move(a));
                l.changed = true;
            }
            l.active.swap(keep);
            if (rc != QPSO_OK) fail_lane(k, rc);
        }
        if (stepped) {
            std::lock_guard<std::mutex> lk(mu_);
            ++steps_;
            active_sum_ += (double)resident_solves_loop();
        }
        // results of the solves whose last step has completed
        for (int k = 0; k < kLanes; ++k) {
            Lane& l = lanes_[k];
            std::vector<Active> keep;
            int err = QPSO_OK;
            for (auto& a : l.finishing) {
                const int d = err ? 0 : dev_.done(a.s);
                if (d < 0) err = d;
                if (d != 1) {
                    keep.push_back(std::move(a));
                    continue;
                }
                qpso_result res{};
                const int rc = dev_.read(a.s, &res);
                if (rc != QPSO_OK && res.status == QPSO_OK) res.status = rc;
                post(a.req, res, kCompleted);
                leave_lane(k, a, false);
                progress = true;
            }
            l.finishing.swap(keep);
            if (err) fail_lane(k, err);
        }
        const size_t resident = resident_solves_loop();
        publish_active(resident);
        if (!progress && resident > 0) dev_.idle_wait();
        return progress || resident > 0;
    }

   private:
    enum Kind { kCompleted, kRejected, kCancelled, kTimedOut };
    struct Active {
        Request req;
        Solve* s;
        int32_t iter;  // iterations enqueued so far
    };
    struct Lane {
        std::vector<Active> active;     // stepped
        std::vector<Active> finishing;  // all iterations enqueued, result not yet read
        int64_t particles = 0;          // of the active solves
        int64_t cost = 0;
        bool changed = true;
    };

    static qpso_result result_with(int status) {
        qpso_result r{};
        r.status = status;
        return r;
    }

    size_t resident_solves_loop() const {
        size_t n = 0;
        for (const auto& l : lanes_) n += l.active.size() + l.finishing.size();
        return n;
    }

    // A request joins a lane whose resident requests' step work is within a
    // factor kRatio of its own both ways (or an empty one), the least loaded
    // of those: requests of like size share the lanes, so both keep the
    // device busy, while a light request never waits on the steps of a far
    // heavier one.  When no lane qualifies, the lane whose heaviest resident
    // is closest to it in ratio.
    int choose_lane(const Request& r) const {
        const double c = (double)std::max<int64_t>(r.cost, 1);
        int best = -1, nearest = 0;
        double best_cost = 0.0, nearest_gap = 0.0;
        for (int k = 0; k < kLanes; ++k) {
            const Lane& l = lanes_[k];
            double lo = 0.0, hi = 0.0, total = 0.0;
            bool any = false;
            for (const auto* v : {&l.active, &l.finishing}) {
                for (const auto& a : *v) {
                    const double x = (double)std::max<int64_t>(a.req.cost, 1);
                    lo = any ? std::min(lo, x) : x;
                    hi = any ? std::max(hi, x) : x;
                    any = true;
                }
            }
            for (const auto& a : l.active) total += (double)a.req.cost;
            if (!any || (lo * (double)kRatio >= c && c * (double)kRatio >= hi)) {
                if (best < 0 || total < best_cost) {
                    best = k;
                    best_cost = total;
                }
            }
            const double gap = any ? std::max(hi / c, c / hi) : 0.0;
            if (k == 0 || gap < nearest_gap) {
                nearest = k;
                nearest_gap = gap;
            }
        }
        return best >= 0 ? best : nearest;
    }

    // A solve leaves its lane (its result is posted): its slots are free and
    // its memory is released in the lane's order.  `stepped`: it was still
    // among the lane's active solves.
    void leave_lane(int k, Active& a, bool stepped) {
        Lane& l = lanes_[k];
        if (stepped) {
            l.particles -= a.req.particles;
            l.cost -= a.req.cost;
        }
        dev_.destroy(a.s, k);
        release(a.req);
    }

    void run() {
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || !pending_.empty() || n_active_ > 0 || !cancel_.empty(); });
                if (stop_ && n_active_ == 0 && cancel_.empty()) break;
            }
            tick();
        }
    }

    // shutdown() without a loop thread: end what is resident.
    void drain_inline() {
        for (;;) {
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (n_active_ == 0 && cancel_.empty()) break;
            }
            tick();
        }
    }

    void publish_active(size_t n) {
        std::lock_guard<std::mutex> lk(mu_);
        n_active_ = n;
    }

    // Every solve of the lane ends with `code` (a failed launch, copy, sync
    // or completion query).
    void fail_lane(int k, int code) {
        Lane& l = lanes_[k];
        dev_.sync(k);
        for (auto* v : {&l.active, &l.finishing}) {
            for (auto& a : *v) {
                dev_.destroy(a.s, k);
                post(a.req, result_with(code), kRejected);
                release(a.req);
            }
            v->clear();
        }
        l.particles = 0;
        l.cost = 0;
        l.changed = true;
    }

    // Post an admitted request's result; its particle slots are free.
    void post(const Request& r, const qpso_result& res, Kind kind) {
        std::lock_guard<std::mutex> lk(mu_);
        resident_ -= r.particles;
        post_locked(r.ticket, res, kind, r.t_submit);
    }

    // Release an admitted request's device bytes (its memory is released).
    void release(const Request& r) {
        std::lock_guard<std::mutex> lk(mu_);
        used_ -= r.bytes;
        auto tr = table_refs_.find(r.table);
        if (tr != table_refs_.end() && --tr->second == 0) {
            used_ -= r.table_bytes;
            table_refs_.erase(tr);
        }
    }

    void post_locked(uint64_t ticket, const qpso_result& res, Kind kind, double t_submit) {
        const double now = dev_.now();
        qpso_result& r = results_[ticket];
        r = res;
        r.latency_ms = (now - t_submit) * 1e3;
        switch (kind) {
            case kCompleted: {
                ++completed_;
                latencies_ms_.push_back(r.latency_ms);
                t_last_ = now;
                break;
            }
            case kRejected: ++rejected_; break;
            case kCancelled: ++cancelled_; break;
            case kTimedOut: ++timed_out_; break;
        }
        done_cv_.notify_all();
    }

    int consume_locked(uint64_t ticket, qpso_result* out) {
        auto it = results_.find(ticket);
        if (out) *out = it->second;
        const int st = it->second.status;
        results_.erase(it);
        live_.erase(ticket);
        return st;
    }

    Dev& dev_;
    const uint64_t budget_;
    std::thread th_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    // under mu_
    bool stop_ = false, down_ = false;
    std::deque<Request> pending_;
    std::unordered_set<uint64_t> live_;     // submitted and not yet read
    std::unordered_map<uint64_t, qpso_result> results_;
    std::unordered_set<uint64_t> cancel_;   // cancellations of admitted requests
    std::map<TableKey, int> table_refs_;    // resident solves per random table
    std::vector<uint64_t> admitted_;
    uint64_t used_ = 0;
    int64_t resident_ = 0, max_particles_ = 8192;
    double timeout_s_ = 0.0;
    int waiters_ = 0;
    size_t n_active_ = 0;
    uint64_t next_ticket_ = 1, submitted_ = 0, completed_ = 0, rejected_ = 0, cancelled_ = 0, timed_out_ = 0;
    uint64_t steps_ = 0;
    double active_sum_ = 0.0;
    std::vector<double> latencies_ms_;
    double t_first_ = 0.0, t_last_ = 0.0;
    bool have_first_ = false;
    // loop thread only
    Lane lanes_[kLanes];
};

}  // namespace qpso_sched

#endif
