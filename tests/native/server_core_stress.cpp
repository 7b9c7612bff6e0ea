// The job server's host core driven by a stub device backend, for runs under
// AddressSanitizer and ThreadSanitizer (no GPU). Every job's result is a
// function of the job alone, so each scenario checks results against that
// function whatever the batching, the lanes and the order of waits.
//
// Scenarios: closed loop; small jobs waited first while a large job of the
// same round still runs (the early-completion path); concurrent submitters
// and waiters; arrival replay; admission against the budget and overlapping
// rounds; failures at submission, planning and execution; shutdown with jobs
// nobody waited for.
#include <algorithm>
#include <atomic>
#include <cassert>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <list>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "mh_error.h"
#include "server_core.hpp"

using namespace mh;
using namespace mh::serve;

namespace {

std::atomic<int> g_failures{0};

#define CHECK(cond)                                                                    \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
            ++g_failures;                                                              \
        }                                                                              \
    } while (0)

// A circuit of n qubits and one parameterised op, held by the caller.
struct TestCircuit {
    std::vector<int32_t> opcodes{5};
    std::vector<uint32_t> qubit_offsets{0, 1}, qubits{0}, param_offsets{0, 1};
    std::vector<double> params{0.0};
    mh_circuit c{};
    TestCircuit(uint32_t n, double angle) {
        params[0] = angle;
        c.num_qubits = n;
        c.num_ops = 1;
        c.opcodes = opcodes.data();
        c.qubit_offsets = qubit_offsets.data();
        c.qubits = qubits.data();
        c.param_offsets = param_offsets.data();
        c.params = params.data();
    }
};

// The stub device: an arena of host memory; a job's "state" range is stamped
// with its ticket at the start and checked at the end of its execution, so
// overlapping ranges of concurrent rounds are caught, and the sanitizer sees
// every write outside the arena.
class StubBackend final : public ServerBackend {
public:
    explicit StubBackend(uint64_t capacity) : capacity_(capacity) {}

    uint64_t capacity_bytes() override { return capacity_; }
    void validate_circuit(const mh_circuit* c) override {
        if (c->num_ops == 0) fail(MH_ERR_INVALID, "empty circuit");
    }
    uint32_t min_qubits(int32_t) override { return 4; }
    uint32_t max_qubits() override { return 30; }
    uint64_t state_bytes(uint32_t n, int32_t precision) override {
        const uint64_t b = (uint64_t{1} << n) * (precision == MH_C128 ? 16 : 8);
        return (b + 255) / 256 * 256;
    }
    std::shared_ptr<void> lane_setup() override { return std::make_shared<int>(7); }
    void* segment_alloc(uint64_t bytes) override {
        std::lock_guard<std::mutex> lock(mu_);
        segments_.emplace_back(bytes, 0);
        ++resizes;
        return segments_.back().data();
    }
    void segment_free(void* p) noexcept override {
        std::lock_guard<std::mutex> lock(mu_);
        for (auto it = segments_.begin(); it != segments_.end(); ++it)
            if (it->data() == p) {
                segments_.erase(it);
                return;
            }
        ++bad_frees;
    }
    void plan(Job& j, int32_t, uint32_t, std::shared_ptr<const void>& plan) override {
        if (j.circuit.view.params[0] < -100.0) fail(MH_ERR_INVALID, "a plan the backend refuses");
        if (!plan) plan = std::make_shared<const uint32_t>(j.circuit.view.num_qubits);
        j.work = std::make_shared<double>(j.circuit.view.params[0]);
    }
    void execute(int32_t precision, const std::vector<Job*>& jobs, void* arena, uint32_t, PhaseSeconds& s) override {
        unsigned char* at = static_cast<unsigned char*>(arena);
        for (Job* j : jobs)
            if (j->circuit.view.params[0] > 100.0) fail(MH_ERR_INTERNAL, "a device failure of the stub");
        std::vector<unsigned char*> where;
        uint32_t biggest = 0;
        for (Job* j : jobs) {
            const uint64_t b = state_bytes(j->circuit.view.num_qubits, precision);
            std::memcpy(at, &j->ticket, 8);
            std::memcpy(at + b - 8, &j->ticket, 8);
            where.push_back(at);
            at += b;
            biggest = std::max(biggest, j->circuit.view.num_qubits);
        }
        std::this_thread::sleep_for(std::chrono::milliseconds(biggest >= 16 ? 40 : 2));
        at = static_cast<unsigned char*>(arena);
        for (std::size_t k = 0; k < jobs.size(); ++k) {
            Job& j = *jobs[k];
            const uint64_t b = state_bytes(j.circuit.view.num_qubits, precision);
            uint64_t t0 = 0, t1 = 0;
            std::memcpy(&t0, where[k], 8);
            std::memcpy(&t1, where[k] + b - 8, 8);
            CHECK(t0 == j.ticket && t1 == j.ticket);
            const double w = *static_cast<const double*>(j.work.get());
            j.raw_terms.resize(2 * j.num_terms());
            for (uint32_t t = 0; t < j.num_terms(); ++t) {
                j.raw_terms[2 * t] = w + static_cast<double>(j.x[t]);
                j.raw_terms[2 * t + 1] = static_cast<double>(j.z[t]) - w;
            }
            if (j.shots) j.counts[j.seed % 5] = j.shots;
            j.work.reset();
            (void)at;
        }
        s.pass += 1e-3;
    }
    std::atomic<int> resizes{0};
    std::atomic<int> bad_frees{0};

private:
    uint64_t capacity_;
    std::mutex mu_;
    std::list<std::vector<unsigned char>> segments_;
};

// One job and its result buffers.
struct TestJob {
    TestCircuit circuit;
    int32_t kind;
    std::vector<uint32_t> offs{0, 2};
    std::vector<uint64_t> x{1, 2}, z{3, 4};
    std::vector<double> coefs{1.0, 0.0, 0.5, -0.25};
    mh_job job{};
    double values[4]{}, raw[4]{};
    uint64_t outcomes[8]{}, counts[8]{};
    mh_job_result result{};

    TestJob(uint32_t n, double angle, int32_t k, int32_t precision = MH_C64) : circuit(n, angle), kind(k) {
        job.kind = k;
        job.precision = k == MH_JOB_SAMPLER ? MH_C128 : precision;
        job.circuit = &circuit.c;
        if (k == MH_JOB_ESTIMATOR) {
            job.num_observables = 1;
            job.observable_offsets = offs.data();
            job.x_masks = x.data();
            job.z_masks = z.data();
            job.coefficients = coefs.data();
        } else {
            job.shots = 100;
            job.seed = static_cast<uint64_t>(angle * 1000);
        }
        result.observable_values = values;
        result.raw_terms = raw;
        result.outcomes = outcomes;
        result.counts = counts;
    }

    void check_result() {
        CHECK(result.status == MH_OK);
        if (result.status != MH_OK) {
            std::fprintf(stderr, "  job failed: %s\n", result.error);
            return;
        }
        const double w = circuit.params[0];
        if (kind == MH_JOB_ESTIMATOR) {
            CHECK(raw[0] == w + 1 && raw[1] == 3 - w && raw[2] == w + 2 && raw[3] == 4 - w);
            Job j;
            j.observable_offsets = offs;
            j.coefficients = coefs;
            j.raw_terms.assign(raw, raw + 4);
            combine_observables(j);
            CHECK(std::memcmp(j.values.data(), values, 2 * sizeof(double)) == 0);
        } else {
            CHECK(result.num_outcomes == 1 && outcomes[0] == job.seed % 5 && counts[0] == 100);
        }
        CHECK(result.completed_seconds >= result.started_seconds);
        CHECK(result.started_seconds >= result.submitted_seconds);
    }
};

mh_server_options options(uint64_t budget, uint32_t lanes) {
    mh_server_options o{};
    o.budget_bytes = budget;
    o.lanes = lanes;
    o.host_threads = 2;
    o.fusion = MH_FUSE_CONTAINMENT;
    return o;
}

void closed_loop_and_early_completion() {
    // One 20-qubit job (16 MiB at complex128) and small ones in the same
    // round: the small ones complete first; they are waited for first.
    Server s(options(1ull << 30, 1), std::make_unique<StubBackend>(1ull << 32));
    std::vector<std::unique_ptr<TestJob>> jobs;
    jobs.push_back(std::make_unique<TestJob>(20, 0.5, MH_JOB_ESTIMATOR, MH_C128));
    for (int i = 0; i < 12; ++i) jobs.push_back(std::make_unique<TestJob>(4 + i % 5, 0.1 * i, i % 3 ? MH_JOB_ESTIMATOR : MH_JOB_SAMPLER));
    std::vector<uint64_t> tickets;
    for (auto& j : jobs) tickets.push_back(s.submit(&j->job));
    for (std::size_t i = jobs.size(); i-- > 0;) s.wait(tickets[i], &jobs[i]->result);
    for (auto& j : jobs) j->check_result();
    mh_server_stats st{};
    s.stats(&st);
    CHECK(st.jobs_completed == jobs.size() && st.jobs_failed == 0);
}

void concurrent_submitters_and_waiters() {
    Server s(options(64ull << 20, 3), std::make_unique<StubBackend>(1ull << 32));
    constexpr int kThreads = 4, kPer = 60;
    std::vector<std::thread> th;
    std::atomic<int> done{0};
    for (int t = 0; t < kThreads; ++t)
        th.emplace_back([&, t] {
            std::mt19937 g(t);
            std::vector<std::unique_ptr<TestJob>> mine;
            std::vector<uint64_t> tickets;
            for (int i = 0; i < kPer; ++i) {
                const uint32_t n = 4 + g() % 14;
                mine.push_back(std::make_unique<TestJob>(n, 0.01 * (t * kPer + i), g() % 4 ? MH_JOB_ESTIMATOR : MH_JOB_SAMPLER,
                                                         g() % 2 ? MH_C64 : MH_C128));
                tickets.push_back(s.submit(&mine.back()->job));
                if (g() % 3 == 0) std::this_thread::sleep_for(std::chrono::microseconds(g() % 3000));
            }
            std::vector<std::size_t> order(kPer);
            for (int i = 0; i < kPer; ++i) order[i] = i;
            std::shuffle(order.begin(), order.end(), g);
            for (std::size_t i : order) {
                s.wait(tickets[i], &mine[i]->result);
                mine[i]->check_result();
                ++done;
            }
        });
    for (auto& t : th) t.join();
    CHECK(done == kThreads * kPer);
    mh_server_stats st{};
    s.stats(&st);
    CHECK(st.jobs_completed == kThreads * kPer && st.jobs_failed == 0);
    CHECK(st.peak_held_bytes <= st.budget_bytes);
    CHECK(st.lanes == 3);
}

void overlap_and_admission() {
    // A long round holds one lane; jobs that arrive meanwhile run on the
    // others and complete before it.
    Server s(options(1ull << 30, 3), std::make_unique<StubBackend>(1ull << 32));
    TestJob big(22, 0.25, MH_JOB_SAMPLER);
    const uint64_t tb = s.submit(&big.job);
    std::this_thread::sleep_for(std::chrono::milliseconds(5));
    std::vector<std::unique_ptr<TestJob>> small;
    std::vector<uint64_t> ts;
    for (int i = 0; i < 6; ++i) {
        small.push_back(std::make_unique<TestJob>(6, 0.3 + i, MH_JOB_ESTIMATOR));
        ts.push_back(s.submit(&small.back()->job));
    }
    for (std::size_t i = 0; i < small.size(); ++i) {
        s.wait(ts[i], &small[i]->result);
        small[i]->check_result();
    }
    s.wait(tb, &big.result);
    big.check_result();
    for (auto& j : small) CHECK(j->result.completed_seconds < big.result.completed_seconds);
    mh_server_stats st{};
    s.stats(&st);
    CHECK(st.peak_concurrent_rounds >= 2);

    // Over the budget alone: refused with MH_ERR_RESOURCE, others go on.
    Server s2(options(1ull << 20, 2), std::make_unique<StubBackend>(1ull << 32));
    TestJob huge(20, 0.5, MH_JOB_ESTIMATOR, MH_C128);   // 16 MiB
    TestJob ok(8, 0.5, MH_JOB_ESTIMATOR);
    const uint64_t th = s2.submit(&huge.job), tok = s2.submit(&ok.job);
    s2.wait(th, &huge.result);
    s2.wait(tok, &ok.result);
    CHECK(huge.result.status == MH_ERR_RESOURCE && std::strstr(huge.result.error, "budget") != nullptr);
    ok.check_result();
}

void large_jobs_one_at_a_time() {
    // With large_job_bytes set, large jobs run in rounds of their own, one at
    // a time, while the small jobs behind them pass; every result is still
    // the job's own, and nothing is left in the queue.
    mh_server_options o = options(1ull << 30, 3);
    o.large_job_bytes = uint64_t{16} << 16;   // complex128 states of 16+ qubits
    Server s(o, std::make_unique<StubBackend>(1ull << 32));
    std::vector<std::unique_ptr<TestJob>> jobs;
    std::vector<uint64_t> ts;
    for (int i = 0; i < 12; ++i) {
        const bool large = i % 3 == 0;
        jobs.push_back(std::make_unique<TestJob>(large ? 18 : 6, 0.1 * i, large ? MH_JOB_SAMPLER : MH_JOB_ESTIMATOR,
                                                 MH_C128));
        ts.push_back(s.submit(&jobs.back()->job));
    }
    for (std::size_t i = 0; i < jobs.size(); ++i) {
        s.wait(ts[i], &jobs[i]->result);
        jobs[i]->check_result();
    }
    // Large rounds never overlapped: their [started, completed] intervals are
    // disjoint; each ran in a round of its own.
    std::vector<std::pair<double, double>> spans;
    for (std::size_t i = 0; i < jobs.size(); i += 3)
        spans.push_back({jobs[i]->result.started_seconds, jobs[i]->result.completed_seconds});
    std::sort(spans.begin(), spans.end());
    for (std::size_t k = 1; k < spans.size(); ++k) CHECK(spans[k].first >= spans[k - 1].second);
    for (std::size_t i = 0; i < jobs.size(); i += 3)
        for (std::size_t k = 0; k < jobs.size(); ++k)
            if (k != i) CHECK(jobs[k]->result.round != jobs[i]->result.round);
    // The small jobs did not wait for the large ones behind them.
    CHECK(jobs[1]->result.completed_seconds < jobs[9]->result.completed_seconds);
    mh_server_stats st{};
    s.stats(&st);
    CHECK(st.jobs_completed == jobs.size() && st.jobs_failed == 0);
}

void failures_and_tickets() {
    Server s(options(1ull << 30, 2), std::make_unique<StubBackend>(1ull << 32));
    TestJob few(2, 0.5, MH_JOB_ESTIMATOR);            // below min_qubits
    bool refused = false;
    try {
        s.submit(&few.job);
    } catch (const Error& e) {
        refused = e.status == MH_ERR_INVALID;
    }
    CHECK(refused);
    TestJob plan_fails(8, -200.0, MH_JOB_ESTIMATOR);
    TestJob device_fails(8, 200.0, MH_JOB_ESTIMATOR);
    TestJob fine(8, 0.75, MH_JOB_ESTIMATOR);
    const uint64_t a = s.submit(&plan_fails.job), b = s.submit(&device_fails.job), c = s.submit(&fine.job);
    s.wait(a, &plan_fails.result);
    s.wait(b, &device_fails.result);
    s.wait(c, &fine.result);
    CHECK(plan_fails.result.status == MH_ERR_INVALID);
    CHECK(device_fails.result.status == MH_ERR_INTERNAL);
    // A device failure fails its group; a job planned in another round is unaffected.
    if (fine.result.round != device_fails.result.round) fine.check_result();
    bool again = false;
    try {
        s.wait(c, &fine.result);
    } catch (const Error& e) {
        again = e.status == MH_ERR_INVALID;
    }
    CHECK(again);
    // Jobs nobody waits for are completed by the shutdown.
    std::vector<std::unique_ptr<TestJob>> left;
    for (int i = 0; i < 20; ++i) {
        left.push_back(std::make_unique<TestJob>(5 + i % 10, 0.2, MH_JOB_ESTIMATOR));
        s.submit(&left.back()->job);
    }
}

void replay() {
    Server s(options(1ull << 30, 3), std::make_unique<StubBackend>(1ull << 32));
    std::vector<std::unique_ptr<TestJob>> jobs;
    std::vector<mh_job> cj;
    std::vector<double> arrival;
    std::vector<mh_job_result> results;
    for (int i = 0; i < 80; ++i) {
        jobs.push_back(std::make_unique<TestJob>(4 + (i * 7) % 15, 0.05 * i, i % 4 ? MH_JOB_ESTIMATOR : MH_JOB_SAMPLER));
        cj.push_back(jobs.back()->job);
        arrival.push_back(0.002 * i);
        results.push_back(jobs.back()->result);
    }
    s.replay(static_cast<uint32_t>(cj.size()), cj.data(), arrival.data(), results.data());
    for (std::size_t i = 0; i < jobs.size(); ++i) {
        jobs[i]->result = results[i];
        jobs[i]->check_result();
        CHECK(results[i].submitted_seconds + 1e-3 >= arrival[i] - 1e-2);
    }
    std::vector<double> bad{0.2, 0.1};
    bool refused = false;
    try {
        s.replay(2, cj.data(), bad.data(), results.data());
    } catch (const Error& e) {
        refused = e.status == MH_ERR_INVALID;
    }
    CHECK(refused);
}

void arena_growth_and_fragmentation() {
    // No reserve: segments are added as rounds need them, within the budget;
    // rounds of mixed sizes share them through the free-range ledger.
    auto backend = std::make_unique<StubBackend>(1ull << 32);
    StubBackend* raw = backend.get();
    Server s(options(8ull << 20, 3), std::move(backend));
    std::vector<std::unique_ptr<TestJob>> jobs;
    std::vector<uint64_t> ts;
    std::mt19937 g(11);
    for (int i = 0; i < 150; ++i) {
        jobs.push_back(std::make_unique<TestJob>(4 + g() % 15, 0.001 * i, MH_JOB_ESTIMATOR, MH_C64));
        ts.push_back(s.submit(&jobs.back()->job));
        if (i % 10 == 0) std::this_thread::sleep_for(std::chrono::milliseconds(3));
    }
    for (std::size_t i = 0; i < jobs.size(); ++i) {
        s.wait(ts[i], &jobs[i]->result);
        jobs[i]->check_result();
    }
    mh_server_stats st{};
    s.stats(&st);
    CHECK(st.peak_held_bytes <= (8ull << 20));
    CHECK(raw->resizes >= 1 && raw->bad_frees == 0);
}

}  // namespace

int main() {
    closed_loop_and_early_completion();
    concurrent_submitters_and_waiters();
    overlap_and_admission();
    large_jobs_one_at_a_time();
    failures_and_tickets();
    replay();
    arena_growth_and_fragmentation();
    if (g_failures) {
        std::fprintf(stderr, "%d checks failed\n", g_failures.load());
        return 1;
    }
    std::printf("server core: all scenarios passed\n");
    return 0;
}
