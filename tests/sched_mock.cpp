// sched_mock.cpp -- CPU tests of the serving scheduler (q-pso_amd/csrc/qpso_sched.h)
// driven by a mock device: a virtual clock that each step advances by a cost
// model, a fake allocator whose failures are injected per request, launch
// failures injected per step, and seeded Poisson arrivals.  Built and run by
// tests/test_scheduler.py (with AddressSanitizer/UBSan, and ThreadSanitizer
// for the threaded scenario).  Prints one line per scenario; exits non-zero
// on the first violated check.
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
    struct Job {
        int id = 0;
        bool fail_create = false;  // the allocator refuses this request
        bool fail_init = false;    // its initialisation launch fails
    };
    struct Solve {
        Job job;
        int lane = -1;
        int steps = 0;
    };
    std::atomic<double> clock{0.0};  // virtual seconds
    double step_base = 1e-4, step_per_particle = 1e-7;
    int fail_step_at = -1;  // the n-th step call fails (a failed launch)
    int n_steps = 0;
    std::set<Solve*> alive;
    bool unsynced[2] = {false, false};
    int violations = 0;  // device-use errors seen by the mock

    double now() const { return clock.load(); }
    int create(Job& j, Solve*& s) {
        if (j.fail_create) return QPSO_E_NOMEM;
        s = new Solve{j};
        alive.insert(s);
        return QPSO_OK;
    }
    int init(Solve* s, int lane) {
        if (s->job.fail_init) return QPSO_E_HIP;
        s->lane = lane;
        unsynced[lane] = true;
        return QPSO_OK;
    }
    int step(int lane, const std::vector<Solve*>& ss, bool, int64_t total) {
        if (n_steps++ == fail_step_at) return QPSO_E_HIP;
        for (Solve* s : ss) {
            if (!alive.count(s) || s->lane != lane) ++violations;
            ++s->steps;
        }
        clock.store(clock.load() + step_base + step_per_particle * (double)total);
        unsynced[lane] = true;
        return QPSO_OK;
    }
    int sync(int lane) {
        unsynced[lane] = false;
        return QPSO_OK;
    }
    int read(int lane, Solve* s, qpso_result* r) {
        if (unsynced[lane] || s->lane != lane) ++violations;  // results are read only after a sync
        r->status = QPSO_OK;
        r->fit = (float)s->steps;  // iterations the solve was stepped
        r->raster_px = (uint64_t)s->job.id;
        return QPSO_OK;
    }
    void destroy(Solve* s) {
        if (!alive.count(s)) ++violations;  // double free
        if (s->lane >= 0 && unsynced[s->lane]) ++violations;  // freed under queued work
        alive.erase(s);
        delete s;
    }
    void on_thread_start() {}
    void on_thread_end() {}
};

using Srv = qpso_sched::Server<MockDev>;

Srv::Request req(int id, uint64_t bytes, int64_t particles, int iters, uint32_t table_seed = 1,
                 uint64_t table_bytes = 0) {
    Srv::Request r;
    r.job.id = id;
    r.bytes = bytes;
    r.particles = particles;
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
    sv.tick();  // c now fits beside a, runs its one iteration and finishes
    CHECK(sv.poll(c, &r) == QPSO_OK);
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
    const uint64_t a = submit(sv, req(1, 10, 64, 4)), b = submit(sv, req(2, 10, 64, 4));
    if (!drain(sv, dev, 1u << 30, 8192)) return false;
    qpso_result r{};
    CHECK(sv.poll(a, &r) == QPSO_E_HIP);  // lane 0
    CHECK(sv.poll(b, &r) == QPSO_OK && r.fit == 4.0f);  // lane 1 carries on
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
