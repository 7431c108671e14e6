// sched_mock.cpp -- CPU tests of the serving scheduler (q-pso_amd/csrc/qpso_sched.h)
// driven by a mock device: a virtual clock that each step advances by a cost
// model, a fake allocator whose failures are injected per request, launch
// failures injected per step, and seeded Poisson arrivals.  Built and run by
// tests/test_scheduler.py (with AddressSanitizer/UBSan, and ThreadSanitizer
// for the threaded scenario).  Prints one line per scenario; exits non-zero
// on the first violated check.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "qpso_sched.h"

namespace {

int g_failures = 0;
#define CHECK(cond)                                                                 \
    do {                                                                            \
        if (!(cond)) {                                                              \
            std::printf("  check failed at line %d: %s\n", __LINE__, #cond);        \
            ++g_failures;                                                           \
            return false;                                                           \
        }                                                                           \
    } while (0)

struct MockDev {
    static constexpr int kLanes = 2;  // as the library
    struct Job {
        int id = 0;
        int64_t particles = 0;
        bool fail_create = false;  // the allocator refuses this request
        bool fail_init = false;    // its initialisation launch fails
    };
    struct Solve {
        Job job;
        int lane = -1;
        int steps = 0;
        double last_use = 0.0;  // completion time of the last queued work that reads it
        double done_at = -1.0;  // completion of its last step, once marked
    };
    // Each lane is a queue of its own: a step starts when the lane's previous
    // one has completed (and not before it is enqueued) and costs step_base
    // plus step_per_particle per particle of the lane.
    std::atomic<double> clock{0.0};  // virtual seconds
    double step_base = 1e-4, step_per_particle = 1e-7;
    int fail_step_at = -1;  // the n-th step call fails (a failed launch)
    int n_steps = 0;
    std::set<Solve*> alive;
    std::vector<double> queued[kLanes];  // completion times of the lanes' queued work
    double lane_end[kLanes] = {};
    int violations = 0;  // device-use errors seen by the mock
    std::vector<int> lane_of_job = std::vector<int>(1024, -1);

    double now() const { return clock.load(); }
    void prune(int lane) {
        auto& q = queued[lane];
        q.erase(std::remove_if(q.begin(), q.end(), [&](double t) { return t <= now(); }), q.end());
    }
    double enqueue(int lane, double cost) {
        lane_end[lane] = std::max(now(), lane_end[lane]) + cost;
        queued[lane].push_back(lane_end[lane]);
        return lane_end[lane];
    }
    int create(Job& j, Solve*& s) {
        if (j.fail_create) return QPSO_E_NOMEM;
        s = new Solve{j};
        alive.insert(s);
        return QPSO_OK;
    }
    int init(Solve* s, int lane) {
        if (s->job.fail_init) return QPSO_E_HIP;
        s->lane = lane;
        if (s->job.id >= 0 && s->job.id < (int)lane_of_job.size()) lane_of_job[s->job.id] = lane;
        s->last_use = enqueue(lane, step_base);
        return QPSO_OK;
    }
    bool ready(int lane) {
        prune(lane);
        return queued[lane].size() < 2;
    }
    int step(int lane, const std::vector<Solve*>& ss, bool, int64_t) {
        if (n_steps++ == fail_step_at) return QPSO_E_HIP;
        if (!ready(lane)) ++violations;  // the server steps only lanes that are ready
        int64_t particles = 0;
        for (Solve* s : ss) particles += s->job.particles;
        const double end = enqueue(lane, step_base + step_per_particle * (double)particles);
        for (Solve* s : ss) {
            if (!alive.count(s) || s->lane != lane || s->done_at >= 0.0) ++violations;
            ++s->steps;
            s->last_use = end;
        }
        return QPSO_OK;
    }
    int mark(int lane, Solve* s) {
        if (s->lane != lane) ++violations;
        s->done_at = lane_end[lane];
        return QPSO_OK;
    }
    int done(Solve* s) {
        if (s->done_at < 0.0) ++violations;
        return now() >= s->done_at ? 1 : 0;
    }
    int sync(int lane) {
        clock.store(std::max(now(), lane_end[lane]));
        queued[lane].clear();
        return QPSO_OK;
    }
    int read(Solve* s, qpso_result* r) {
        if (s->done_at < 0.0 || now() < s->done_at) ++violations;  // read only after its work completed
        r->status = QPSO_OK;
        r->fit = (float)s->steps;  // iterations the solve was stepped
        r->raster_px = (uint64_t)s->job.id;
        return QPSO_OK;
    }
    // released in the lane's order: the lane must be the one that ran it
    void destroy(Solve* s, int lane) {
        if (!alive.count(s)) ++violations;  // double free
        if (s->lane >= 0 && s->lane != lane) ++violations;  // not ordered after its work
        alive.erase(s);
        delete s;
    }
    // nothing to do until the device moves: advance to its next completion
    void idle_wait() {
        double next = -1.0;
        for (int k = 0; k < kLanes; ++k)
            for (double t : queued[k])
                if (t > now() && (next < 0.0 || t < next)) next = t;
        if (next > 0.0) clock.store(next);
    }
    void on_thread_start() {}
    void on_thread_end() {}
};

using Srv = qpso_sched::Server<MockDev>;

Srv::Request req(int id, uint64_t bytes, int64_t particles, int iters, uint32_t table_seed = 1,
                 uint64_t table_bytes = 0, int64_t cost = 0) {
    Srv::Request r;
    r.job.id = id;
    r.job.particles = particles;
    r.bytes = bytes;
    r.particles = particles;
    r.cost = cost > 0 ? cost : particles;
    r.iters = iters;
    r.table = qpso_sched::TableKey(table_seed, (int32_t)particles, 512);
    r.table_bytes = table_bytes;
    return r;
}

uint64_t submit(Srv& sv, Srv::Request r) {
    uint64_t t = 0;
    if (sv.submit(std::move(r), &t) != QPSO_OK) return 0;
    return t;
}

// Tick until nothing is resident or queued, checking the budget and slots
// after every pass.
bool drain(Srv& sv, MockDev& dev, uint64_t budget, int64_t slots, int max_ticks = 100000) {
    for (int i = 0; i < max_ticks; ++i) {
        const bool busy = sv.tick();
        CHECK(sv.used_bytes() <= budget);
        CHECK(sv.resident_particles() <= slots || sv.resident_solves() == 1);
        CHECK(dev.violations == 0);
        if (!busy && sv.resident_solves() == 0) return true;
    }
    CHECK(false && "did not drain");
    return false;
}

bool fifo_and_budget() {
    MockDev dev;
    Srv sv(dev, 1000);
    const uint64_t a = submit(sv, req(1, 600, 10, 3)), b = submit(sv, req(2, 600, 10, 3)),
                   c = submit(sv, req(3, 100, 10, 3));
    CHECK(a && b && c);
    sv.tick();
    // b does not fit beside a; c would, but admission keeps arrival order
    CHECK(sv.admission_order() == std::vector<uint64_t>{a});
    CHECK(sv.used_bytes() == 600);
    if (!drain(sv, dev, 1000, 8192)) return false;
    CHECK((sv.admission_order() == std::vector<uint64_t>{a, b, c}));
    qpso_result r{};
    for (uint64_t t : {a, b, c}) {
        CHECK(sv.poll(t, &r) == QPSO_OK);
        CHECK(r.fit == 3.0f);
    }
    CHECK(sv.used_bytes() == 0 && sv.resident_particles() == 0 && dev.alive.empty());
    return true;
}

bool budget_refusal() {
    MockDev dev;
    Srv sv(dev, 1000);
    uint64_t t = 0;
    CHECK(sv.submit(req(1, 1001, 10, 3), &t) == QPSO_E_BUDGET);
    CHECK(sv.submit(req(2, 800, 10, 3, 1, 300), &t) == QPSO_E_BUDGET);  // arena + its random table
    CHECK(sv.submit(req(3, 700, 10, 3, 1, 300), &t) == QPSO_OK);
    qpso_server_stats st;
    sv.stats(&st);
    CHECK(st.rejected == 2 && st.submitted == 1);
    return drain(sv, dev, 1000, 8192);
}

bool table_charged_once() {
    MockDev dev;
    Srv sv(dev, 1000);
    const uint64_t x = submit(sv, req(1, 300, 10, 4, 7, 300)), y = submit(sv, req(2, 300, 10, 4, 7, 300)),
                   z = submit(sv, req(3, 300, 20, 4, 9, 300));
    sv.tick();
    // x and y share one table: 300 + 300 + 300; z's own table would exceed 1000
    CHECK((sv.admission_order() == std::vector<uint64_t>{x, y}));
    CHECK(sv.used_bytes() == 900);
    if (!drain(sv, dev, 1000, 8192)) return false;
    CHECK((sv.admission_order() == std::vector<uint64_t>{x, y, z}));
    CHECK(sv.used_bytes() == 0);
    return true;
}

bool slots_and_oversize() {
    MockDev dev;
    Srv sv(dev, 1u << 30);
    sv.set_slots(100);
    const uint64_t a = submit(sv, req(1, 10, 60, 2)), b = submit(sv, req(2, 10, 60, 2)),
                   big = submit(sv, req(3, 10, 150, 2)), d = submit(sv, req(4, 10, 30, 2));
    sv.tick();
    CHECK(sv.admission_order() == std::vector<uint64_t>{a});
    int big_alone_ticks = 0;
    for (int i = 0; i < 1000; ++i) {
        const bool busy = sv.tick();
        if (sv.resident_particles() > 100) {
            CHECK(sv.resident_solves() == 1);  // the oversize solve runs alone
            ++big_alone_ticks;
        }
        if (!busy && sv.resident_solves() == 0) break;
    }
    CHECK(big_alone_ticks > 0);
    CHECK((sv.admission_order() == std::vector<uint64_t>{a, b, big, d}));
    CHECK(sv.resident_particles() == 0 && dev.alive.empty() && dev.violations == 0);
    return true;
}

bool iterations_exact() {
    MockDev dev;
    Srv sv(dev, 1u << 30);
    const int iters[] = {1, 5, 17, 2, 9, 33};
    std::vector<uint64_t> t;
    for (int i = 0; i < 6; ++i) t.push_back(submit(sv, req(i, 10, 32 * (i + 1), iters[i])));
    if (!drain(sv, dev, 1u << 30, 8192)) return false;
    for (int i = 0; i < 6; ++i) {
        qpso_result r{};
        CHECK(sv.poll(t[i], &r) == QPSO_OK);
        CHECK(r.fit == (float)iters[i] && r.raster_px == (uint64_t)i);
        CHECK(sv.poll(t[i], &r) == QPSO_E_NOTFOUND);  // read once
    }
    return true;
}

bool cancel_queued_and_resident() {
    MockDev dev;
    Srv sv(dev, 1000);
    const uint64_t a = submit(sv, req(1, 600, 10, 100)), b = submit(sv, req(2, 600, 10, 100)),
                   c = submit(sv, req(3, 100, 10, 1));
    sv.tick();
    CHECK(sv.admission_order() == std::vector<uint64_t>{a});
    qpso_result r{};
    CHECK(sv.cancel(b) == QPSO_OK);  // queued: ends at once
    CHECK(sv.poll(b, &r) == QPSO_E_CANCELLED && r.status == QPSO_E_CANCELLED);
    CHECK(sv.cancel(b) == QPSO_E_NOTFOUND);  // already read
    // c now fits beside a and runs its one iteration
    int rc_c = QPSO_E_PENDING;
    for (int i = 0; i < 100 && rc_c == QPSO_E_PENDING; ++i) {
        sv.tick();
        rc_c = sv.poll(c, &r);
    }
    CHECK(rc_c == QPSO_OK && r.fit == 1.0f);
    CHECK(sv.cancel(a) == QPSO_OK);  // resident: leaves at the next step boundary
    CHECK(sv.poll(a, &r) == QPSO_E_PENDING);
    sv.tick();
    CHECK(sv.poll(a, &r) == QPSO_E_CANCELLED);
    CHECK(sv.used_bytes() == 0 && sv.resident_particles() == 0 && dev.alive.empty() && dev.violations == 0);
    // a finished, unread request keeps its result
    const uint64_t d = submit(sv, req(4, 10, 10, 1));
    if (!drain(sv, dev, 1000, 8192)) return false;
    CHECK(sv.cancel(d) == QPSO_OK);
    CHECK(sv.poll(d, &r) == QPSO_OK && r.fit == 1.0f);
    qpso_server_stats st;
    sv.stats(&st);
    CHECK(st.cancelled == 2 && st.completed == 2);
    return true;
}

bool timeouts() {
    MockDev dev;
    Srv sv(dev, 1000);
    sv.set_timeout(0.01);  // 10 ms of virtual time; a step of 10 particles costs ~0.1 ms
    const uint64_t a = submit(sv, req(1, 600, 10, 1000)), b = submit(sv, req(2, 600, 10, 1));
    if (!drain(sv, dev, 1000, 8192)) return false;
    qpso_result r{};
    CHECK(sv.poll(a, &r) == QPSO_E_TIMEOUT);  // resident past its deadline
    CHECK(sv.poll(b, &r) == QPSO_E_TIMEOUT);  // queued behind a until its deadline
    const uint64_t c = submit(sv, req(3, 100, 10, 5));
    if (!drain(sv, dev, 1000, 8192)) return false;
    CHECK(sv.poll(c, &r) == QPSO_OK);         // five steps, well within its time
    CHECK(sv.used_bytes() == 0 && dev.alive.empty());
    qpso_server_stats st;
    sv.stats(&st);
    CHECK(st.timed_out == 2 && st.completed == 1);
    return true;
}

bool create_and_init_failures() {
    MockDev dev;
    Srv sv(dev, 1000);
    auto r1 = req(1, 400, 10, 3);
    r1.job.fail_create = true;  // the allocator refuses it
    auto r2 = req(2, 400, 10, 3);
    r2.job.fail_init = true;
    const uint64_t a = submit(sv, r1), b = submit(sv, r2), c = submit(sv, req(3, 900, 10, 3));
    if (!drain(sv, dev, 1000, 8192)) return false;
    qpso_result r{};
    CHECK(sv.poll(a, &r) == QPSO_E_NOMEM);
    CHECK(sv.poll(b, &r) == QPSO_E_HIP);
    CHECK(sv.poll(c, &r) == QPSO_OK && r.fit == 3.0f);  // needs the whole budget: both charges were released
    CHECK(sv.used_bytes() == 0 && dev.alive.empty());
    return true;
}

bool step_failure_ends_one_lane() {
    MockDev dev;
    Srv sv(dev, 1u << 30);
    dev.fail_step_at = 0;  // the first step call: lane 0's first step
    // like sizes: one in each lane
    const uint64_t a = submit(sv, req(1, 10, 64, 4)), b = submit(sv, req(2, 10, 64, 4));
    if (!drain(sv, dev, 1u << 30, 8192)) return false;
    qpso_result r{};
    CHECK(sv.poll(a, &r) == QPSO_E_HIP);  // lane 0
    CHECK(sv.poll(b, &r) == QPSO_OK && r.fit == 4.0f);  // lane 1 carries on
    CHECK(dev.lane_of_job[1] == 0 && dev.lane_of_job[2] == 1);
    CHECK(dev.alive.empty());
    return true;
}

bool poisson_stream() {
    MockDev dev;
    const uint64_t budget = 64000;
    const int64_t slots = 1024;
    Srv sv(dev, budget);
    sv.set_slots(slots);
    std::mt19937_64 rng(12345);
    std::exponential_distribution<double> gap(400.0);  // 400 arrivals per virtual second
    const int n = 400;
    std::vector<double> arrive(n);
    double t = 0.0;
    for (auto& x : arrive) x = (t += gap(rng));
    const int sizes[] = {64, 128, 256}, iters[] = {5, 10, 20};
    std::vector<int> want(n);
    std::vector<uint64_t> tickets;
    std::map<uint64_t, int> index;
    int next = 0, done = 0;
    for (int guard = 0; guard < 1000000 && done < n; ++guard) {
        while (next < n && arrive[next] <= dev.now()) {
            const int p = sizes[rng() % 3], it = iters[rng() % 3];
            want[next] = it;
            const uint64_t tk = submit(sv, req(next, 40 * (uint64_t)p, p, it, (uint32_t)(rng() % 4), 2000));
            CHECK(tk != 0);
            index[tk] = next;
            tickets.push_back(tk);
            ++next;
        }
        const bool busy = sv.tick();
        CHECK(sv.used_bytes() <= budget);
        CHECK(sv.resident_particles() <= slots || sv.resident_solves() == 1);
        CHECK(dev.violations == 0);
        for (auto it = index.begin(); it != index.end();) {
            qpso_result r{};
            const int rc = sv.poll(it->first, &r);
            if (rc == QPSO_E_PENDING) {
                ++it;
                continue;
            }
            CHECK(rc == QPSO_OK && r.fit == (float)want[it->second]);
            ++done;
            it = index.erase(it);
        }
        if (!busy && next < n) dev.clock.store(std::max(dev.now(), arrive[next]));  // idle: jump to the next arrival
    }
    CHECK(done == n);
    CHECK(sv.admission_order() == tickets);  // arrival order
    qpso_server_stats st;
    sv.stats(&st);
    CHECK(st.completed == (uint64_t)n && st.p50_ms > 0.0 && st.p99_ms >= st.p50_ms && st.solves_per_s > 0.0);
    CHECK(st.mean_active > 1.0);  // requests overlapped on the device
    std::printf("  poisson: %d solves, p50 %.3f ms, p99 %.3f ms, %.0f solves/s, mean resident %.2f (virtual)\n", n,
                st.p50_ms, st.p99_ms, st.solves_per_s, st.mean_active);
    CHECK(sv.used_bytes() == 0 && dev.alive.empty());
    return true;
}

// A light request resident beside a heavy one runs in the other lane at
// its own step time: it finishes long before one iteration of its own at
// the heavy solve's step time would.  A second heavy request then joins the
// heavy one's lane, not the light one's.
bool light_lane_latency() {
    MockDev dev;
    Srv sv(dev, 1u << 30);
    sv.set_slots(1 << 20);
    // heavy: 100,000 particles, a step of ~10 ms; light: 64 particles, a
    // step of ~0.1 ms, 20 iterations
    const uint64_t h1 = submit(sv, req(1, 1000, 100000, 30, 1, 0, 100000LL * 100000));
    const uint64_t lt = submit(sv, req(3, 100, 64, 20, 3, 0, 64LL * 1000));
    const uint64_t h2 = submit(sv, req(2, 1000, 100000, 30, 2, 0, 100000LL * 100000));
    CHECK(h1 && h2 && lt);
    if (!drain(sv, dev, 1u << 30, 1 << 20)) return false;
    CHECK(dev.lane_of_job[1] == 0 && dev.lane_of_job[3] == 1 && dev.lane_of_job[2] == 0);
    qpso_result r{};
    CHECK(sv.poll(lt, &r) == QPSO_OK && r.fit == 20.0f);
    const double light_ms = r.latency_ms;
    CHECK(sv.poll(h1, &r) == QPSO_OK && r.fit == 30.0f);
    const double heavy_ms = r.latency_ms;
    CHECK(sv.poll(h2, &r) == QPSO_OK && r.fit == 30.0f);
    CHECK(heavy_ms > 250.0);  // 30 steps of ~10 ms or more
    CHECK(light_ms < 10.0);   // under one heavy step
    std::printf("  light lane: light solve %.2f ms, heavy solve %.1f ms (virtual)\n", light_ms, heavy_ms);
    CHECK(sv.used_bytes() == 0 && dev.alive.empty() && dev.violations == 0);
    return true;
}

// Requests of like size are spread over both lanes by load.
bool like_sizes_balanced() {
    MockDev dev;
    Srv sv(dev, 1u << 30);
    for (int k = 0; k < 8; ++k) submit(sv, req(10 + k, 10, 100 + 10 * k, 5, 1, 0, 1000 + 100 * k));
    if (!drain(sv, dev, 1u << 30, 8192)) return false;
    int in0 = 0;
    for (int k = 0; k < 8; ++k) in0 += dev.lane_of_job[10 + k] == 0;
    CHECK(in0 == 4);
    return true;
}

// The stress mix in miniature: heavy and light requests interleaved and
// submitted at once; every light one finishes before the first heavy one.
bool interleaved_mix() {
    MockDev dev;
    Srv sv(dev, 1u << 30);
    sv.set_slots(1 << 24);
    std::vector<uint64_t> heavy, light;
    for (int k = 0; k < 8; ++k) {
        heavy.push_back(submit(sv, req(100 + k, 1000, 2048, 50, 10 + k, 0, 2048LL * 100000)));
        light.push_back(submit(sv, req(200 + k, 100, 64, 20 + 20 * k, 30 + k, 0, 64LL * 1016)));
    }
    if (!drain(sv, dev, 1u << 30, 1 << 24)) return false;
    double light_max = 0.0, heavy_min = 1e30;
    qpso_result r{};
    for (uint64_t t : light) {
        CHECK(sv.poll(t, &r) == QPSO_OK);
        light_max = std::max(light_max, r.latency_ms);
    }
    for (uint64_t t : heavy) {
        CHECK(sv.poll(t, &r) == QPSO_OK && r.fit == 50.0f);
        heavy_min = std::min(heavy_min, r.latency_ms);
    }
    for (int k = 0; k < 8; ++k) CHECK(dev.lane_of_job[200 + k] == 1 && dev.lane_of_job[100 + k] == 0);
    std::printf("  interleaved: light max %.1f ms, heavy min %.1f ms (virtual)\n", light_max, heavy_min);
    CHECK(light_max < heavy_min);
    CHECK(sv.used_bytes() == 0 && dev.alive.empty() && dev.violations == 0);
    return true;
}

// A finished solve's bytes are released with it: a queued request that needs
// them is admitted while other solves still run.
bool freed_bytes_admit() {
    MockDev dev;
    Srv sv(dev, 1000);
    const uint64_t lng = submit(sv, req(1, 300, 10, 5000)), a = submit(sv, req(2, 600, 10, 2)),
                   b = submit(sv, req(3, 600, 10, 2));
    CHECK(lng && a && b);
    qpso_result r{};
    int rc_b = QPSO_E_PENDING;
    bool b_beside_long = false;
    for (int i = 0; i < 100000 && rc_b == QPSO_E_PENDING; ++i) {
        sv.tick();
        CHECK(sv.used_bytes() <= 1000 && dev.violations == 0);
        const auto order = sv.admission_order();
        if (order.size() == 3 && sv.poll(lng, &r) == QPSO_E_PENDING) b_beside_long = true;
        rc_b = sv.poll(b, &r);
    }
    CHECK(rc_b == QPSO_OK && r.fit == 2.0f);
    CHECK(b_beside_long);  // a's bytes went to b while the long solve ran
    CHECK(sv.poll(a, &r) == QPSO_OK && r.fit == 2.0f);
    CHECK(sv.cancel(lng) == QPSO_OK);
    if (!drain(sv, dev, 1000, 8192)) return false;
    CHECK(sv.poll(lng, &r) == QPSO_E_CANCELLED);
    CHECK(sv.used_bytes() == 0 && dev.alive.empty());
    return true;
}

bool threaded_shutdown_releases_waiters() {
    MockDev dev;
    Srv sv(dev, 1000);
    // a finished request read once: a second wait returns at once
    const uint64_t quick = submit(sv, req(9, 100, 10, 1));
    std::vector<uint64_t> t;
    for (int i = 0; i < 4; ++i) t.push_back(submit(sv, req(i, 400, 10, 1 << 30)));  // two resident, two queued
    sv.start();
    qpso_result r{};
    CHECK(sv.wait(quick, &r) == QPSO_OK);
    CHECK(sv.wait(quick, &r) == QPSO_E_NOTFOUND);
    std::vector<int> rc(4, 1);
    std::vector<std::thread> th;
    for (int i = 0; i < 4; ++i) th.emplace_back([&, i] { rc[i] = sv.wait(t[i], nullptr); });
    const auto t0 = std::chrono::steady_clock::now();
    while (sv.waiters() < 4 && std::chrono::steady_clock::now() - t0 < std::chrono::seconds(20))
        std::this_thread::yield();
    CHECK(sv.waiters() == 4);
    sv.shutdown();  // returns only once every waiter has left wait()
    for (auto& x : th) x.join();
    for (int i = 0; i < 4; ++i) CHECK(rc[i] == QPSO_E_CANCELLED);
    CHECK(sv.waiters() == 0 && sv.used_bytes() == 0);
    uint64_t tk = 0;
    CHECK(sv.submit(req(10, 10, 10, 1), &tk) == QPSO_E_INVAL);  // after shutdown
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    struct Case {
        const char* name;
        bool (*fn)();
    };
    const Case cases[] = {
        {"fifo_and_budget", fifo_and_budget},
        {"budget_refusal", budget_refusal},
        {"table_charged_once", table_charged_once},
        {"slots_and_oversize", slots_and_oversize},
        {"iterations_exact", iterations_exact},
        {"cancel_queued_and_resident", cancel_queued_and_resident},
        {"timeouts", timeouts},
        {"create_and_init_failures", create_and_init_failures},
        {"step_failure_ends_one_lane", step_failure_ends_one_lane},
        {"poisson_stream", poisson_stream},
        {"light_lane_latency", light_lane_latency},
        {"like_sizes_balanced", like_sizes_balanced},
        {"interleaved_mix", interleaved_mix},
        {"freed_bytes_admit", freed_bytes_admit},
        {"threaded_shutdown_releases_waiters", threaded_shutdown_releases_waiters},
    };
    const std::string only = argc > 1 ? argv[1] : "";
    for (const Case& c : cases) {
        if (!only.empty() && only != c.name) continue;
        const bool ok = c.fn();
        std::printf("%s %s\n", ok ? "PASS" : "FAIL", c.name);
    }
    if (g_failures) return 1;
    std::printf("ALL PASS\n");
    return 0;
}
