// The job server's host core (server_core.hpp): queue, lanes, ledger,
// plan cache and results. No device code.
#include "server_core.hpp"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <exception>

#include "mh_error.h"

namespace mh::serve {
namespace {

using clock_type = std::chrono::steady_clock;

constexpr uint64_t kDefaultBudget = 256'000'000'000ULL;  // the ledger budget: 256 GB of the 288 GB
constexpr uint32_t kDefaultCacheCapacity = 256;
constexpr uint32_t kDefaultLanes = 3;
constexpr uint32_t kMaxLanes = 8;
constexpr uint64_t kMinSegment = 256ull << 20;

template <typename T>
void put(std::string& key, const T& v) {
    key.append(reinterpret_cast<const char*>(&v), sizeof(v));
}

// Runs body(i) for i in [0, count) on up to `threads` threads; the first
// exception is rethrown after all have stopped.
template <typename F>
void for_each_index(std::size_t count, uint32_t threads, F&& body) {
    std::atomic<std::size_t> next{0};
    std::exception_ptr error;
    std::mutex m;
    auto worker = [&] {
        try {
            for (std::size_t i; (i = next.fetch_add(1)) < count;) body(i);
        } catch (...) {
            std::lock_guard<std::mutex> lock(m);
            if (!error) error = std::current_exception();
        }
    };
    std::vector<std::thread> pool;
    const std::size_t t = std::max<std::size_t>(1, std::min<std::size_t>(threads, count));
    for (std::size_t k = 1; k < t; ++k) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
    if (error) std::rethrow_exception(error);
}

void status_of(const std::exception_ptr& e, int32_t& status, std::string& what) {
    try {
        std::rethrow_exception(e);
    } catch (const Error& x) {
        status = x.status;
        what = x.what();
    } catch (const std::bad_alloc&) {
        status = MH_ERR_RESOURCE;
        what = "host allocation failed";
    } catch (const std::exception& x) {
        status = MH_ERR_INTERNAL;
        what = x.what();
    }
}

}  // namespace

void OwnedCircuit::copy_from(const mh_circuit* c) {
    const uint32_t m = c->num_ops;
    opcodes.assign(c->opcodes, c->opcodes + m);
    qubit_offsets.assign(c->qubit_offsets, c->qubit_offsets + m + 1);
    param_offsets.assign(c->param_offsets, c->param_offsets + m + 1);
    qubits.assign(c->qubits, c->qubits + qubit_offsets[m]);
    params.assign(c->params, c->params + param_offsets[m]);
    view.num_qubits = c->num_qubits;
    view.num_ops = m;
    view.opcodes = opcodes.data();
    view.qubit_offsets = qubit_offsets.data();
    view.qubits = qubits.data();
    view.param_offsets = param_offsets.data();
    view.params = params.data();
}

void combine_observables(Job& j) {
    const uint32_t m = static_cast<uint32_t>(j.observable_offsets.size()) - 1;
    j.values.assign(2 * static_cast<std::size_t>(m), 0.0);
    for (uint32_t o = 0; o < m; ++o) {
        double tr = 0.0, ti = 0.0;
        for (uint32_t t = j.observable_offsets[o]; t < j.observable_offsets[o + 1]; ++t) {
            const double er = j.raw_terms[2 * t], ei = j.raw_terms[2 * t + 1];
            const double pr = 1.0 * er - 0.0 * ei;
            const double pi = 1.0 * ei + 0.0 * er;
            const double cr = j.coefficients[2 * t], ci = j.coefficients[2 * t + 1];
            const double xr = cr * pr - ci * pi;
            const double xi = cr * pi + ci * pr;
            tr = tr + xr;
            ti = ti + xi;
        }
        j.values[2 * o] = tr;
        j.values[2 * o + 1] = ti;
    }
}

std::string structure_key(const mh_circuit* c, int32_t fusion, uint32_t width) {
    std::string key;
    key.reserve(32 + 12 * static_cast<std::size_t>(c->num_ops));
    put(key, uint32_t{1});  // key format
    put(key, fusion);
    put(key, width);
    put(key, c->num_qubits);
    put(key, c->num_ops);
    for (uint32_t i = 0; i < c->num_ops; ++i) {
        const uint32_t q0 = c->qubit_offsets[i], q1 = c->qubit_offsets[i + 1];
        const uint32_t p0 = c->param_offsets[i], p1 = c->param_offsets[i + 1];
        put(key, c->opcodes[i]);
        put(key, q1 - q0);
        key.append(reinterpret_cast<const char*>(c->qubits + q0), sizeof(uint32_t) * (q1 - q0));
        put(key, p1 - p0);
        if (c->opcodes[i] == 21)  // a unitary: its entries steer fusion
            key.append(reinterpret_cast<const char*>(c->params + p0), sizeof(double) * (p1 - p0));
    }
    return key;
}

std::shared_ptr<const void> PlanCache::find(const std::string& key) {
    std::lock_guard<std::mutex> lock(mu_);
    const auto it = index_.find(key);
    if (it == index_.end()) {
        ++misses_;
        return nullptr;
    }
    lru_.splice(lru_.begin(), lru_, it->second);
    ++hits_;
    return it->second->second;
}

void PlanCache::insert(const std::string& key, std::shared_ptr<const void> plan) {
    std::lock_guard<std::mutex> lock(mu_);
    if (const auto it = index_.find(key); it != index_.end()) {
        lru_.erase(it->second);
        index_.erase(it);
    }
    lru_.emplace_front(key, std::move(plan));
    index_[key] = lru_.begin();
    while (lru_.size() > capacity_) {
        index_.erase(lru_.back().first);
        lru_.pop_back();
    }
}

uint64_t PlanCache::hits() const {
    std::lock_guard<std::mutex> lock(mu_);
    return hits_;
}

uint64_t PlanCache::misses() const {
    std::lock_guard<std::mutex> lock(mu_);
    return misses_;
}

Server::Server(const mh_server_options& o, std::unique_ptr<ServerBackend> backend)
    : options_(o),
      backend_(std::move(backend)),
      cache_(o.plan_cache_capacity ? o.plan_cache_capacity : kDefaultCacheCapacity),
      t0_(clock_type::now()) {
    budget_ = o.budget_bytes ? o.budget_bytes : kDefaultBudget;
    budget_ = std::min<uint64_t>(budget_, backend_->capacity_bytes());
    threads_ = o.host_threads ? o.host_threads : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    lanes_ = o.lanes ? std::min(o.lanes, kMaxLanes) : kDefaultLanes;
    if (o.reserve_bytes) add_segment_locked(std::min(o.reserve_bytes, budget_));
    try {
        for (uint32_t l = 0; l < lanes_; ++l) workers_.emplace_back([this, l] { lane_loop(l); });
    } catch (...) {
        {
            std::lock_guard<std::mutex> lock(mu_);
            stop_ = true;
        }
        work_cv_.notify_all();
        for (auto& w : workers_) w.join();
        throw;
    }
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lock(mu_);
        stop_ = true;
    }
    work_cv_.notify_all();
    for (auto& w : workers_) w.join();
    for (Segment& g : segments_) backend_->segment_free(g.base);
}

double Server::now() const { return std::chrono::duration<double>(clock_type::now() - t0_).count(); }

uint64_t Server::submit(const mh_job* job) {
    auto j = std::make_shared<Job>();
    validate_and_copy(job, *j);
    std::lock_guard<std::mutex> lock(mu_);
    if (stop_) fail(MH_ERR_INVALID, "the server is shutting down");
    j->ticket = ++last_ticket_;
    j->submitted = now();
    const uint64_t t = j->ticket;
    queue_.push_back(j);
    jobs_.emplace(t, std::move(j));
    work_cv_.notify_one();
    return t;
}

void Server::wait(uint64_t ticket, mh_job_result* out) {
    JobPtr j;
    {
        std::unique_lock<std::mutex> lock(mu_);
        const auto it = jobs_.find(ticket);
        if (it == jobs_.end() || it->second->waited) fail(MH_ERR_INVALID, "unknown or already collected ticket");
        j = it->second;
        j->waited = true;
        done_cv_.wait(lock, [&] { return j->done; });
        jobs_.erase(ticket);
    }
    write_result(*j, out);
}

void Server::stats(mh_server_stats* s) {
    std::lock_guard<std::mutex> lock(mu_);
    std::memset(s, 0, sizeof(*s));
    s->jobs_completed = completed_;
    s->jobs_failed = failed_;
    s->rounds = rounds_;
    s->plan_cache_hits = cache_.hits();
    s->plan_cache_misses = cache_.misses();
    s->peak_round_bytes = peak_round_bytes_;
    s->budget_bytes = budget_;
    s->busy_seconds = busy_seconds_;
    s->plan_seconds = plan_seconds_;
    s->arena_seconds = arena_seconds_;
    s->compile_seconds = phases_.compile;
    s->pass_seconds = phases_.pass;
    s->expectation_seconds = phases_.expectation;
    s->sample_seconds = phases_.sample;
    s->lanes = lanes_;
    s->peak_concurrent_rounds = peak_overlap_;
    s->peak_held_bytes = peak_held_;
}

void Server::replay(uint32_t count, const mh_job* jobs, const double* arrival_seconds, mh_job_result* results) {
    for (uint32_t i = 1; i < count; ++i)
        if (arrival_seconds[i] < arrival_seconds[i - 1]) fail(MH_ERR_INVALID, "arrival times must not decrease");
    std::vector<uint64_t> tickets(count, 0);
    std::vector<int32_t> submit_status(count, MH_OK);
    std::vector<std::string> submit_error(count);
    const auto start = clock_type::now();
    std::thread submitter([&] {
        for (uint32_t i = 0; i < count; ++i) {
            std::this_thread::sleep_until(start + std::chrono::duration_cast<clock_type::duration>(
                                                      std::chrono::duration<double>(arrival_seconds[i])));
            try {
                tickets[i] = submit(&jobs[i]);
            } catch (...) {
                status_of(std::current_exception(), submit_status[i], submit_error[i]);
            }
        }
    });
    submitter.join();
    for (uint32_t i = 0; i < count; ++i) {
        if (submit_status[i] != MH_OK) {
            mh_job_result& r = results[i];
            r.status = submit_status[i];
            std::memset(r.error, 0, sizeof(r.error));
            std::strncpy(r.error, submit_error[i].c_str(), sizeof(r.error) - 1);
            r.num_outcomes = 0;
            r.round = 0;
            r.lane = 0;
            r.plan_cache_hit = 0;
            r.submitted_seconds = r.started_seconds = r.completed_seconds = 0;
            continue;
        }
        wait(tickets[i], &results[i]);
    }
}

void Server::validate_and_copy(const mh_job* job, Job& j) {
    if (job == nullptr) fail(MH_ERR_INVALID, "null job");
    if (job->kind != MH_JOB_ESTIMATOR && job->kind != MH_JOB_SAMPLER) fail(MH_ERR_INVALID, "unknown job kind");
    if (job->precision != MH_C128 && job->precision != MH_C64)
        fail(MH_ERR_INVALID, "precision must be MH_C128 or MH_C64");
    if (job->circuit == nullptr) fail(MH_ERR_INVALID, "null circuit");
    backend_->validate_circuit(job->circuit);
    const uint32_t n = job->circuit->num_qubits;
    const uint32_t lo = backend_->min_qubits(job->precision), hi = backend_->max_qubits();
    if (n < lo || n > hi)
        fail(MH_ERR_INVALID, "a served circuit needs " + std::to_string(lo) + " to " + std::to_string(hi) + " qubits");
    j.kind = job->kind;
    j.precision = job->precision;
    j.global_phase = job->global_phase;
    j.circuit.copy_from(job->circuit);
    if (job->kind == MH_JOB_ESTIMATOR) {
        if (job->num_observables > 0 && job->observable_offsets == nullptr)
            fail(MH_ERR_INVALID, "observables without offsets");
        if (job->observable_offsets)
            j.observable_offsets.assign(job->observable_offsets, job->observable_offsets + job->num_observables + 1);
        if (j.observable_offsets.empty()) j.observable_offsets.push_back(0);
        if (j.observable_offsets[0] != 0) fail(MH_ERR_INVALID, "observable_offsets must start at 0");
        for (std::size_t o = 1; o < j.observable_offsets.size(); ++o)
            if (j.observable_offsets[o] < j.observable_offsets[o - 1])
                fail(MH_ERR_INVALID, "observable_offsets must not decrease");
        const uint32_t terms = j.observable_offsets.back();
        if (terms > 0 && (job->x_masks == nullptr || job->z_masks == nullptr || job->coefficients == nullptr))
            fail(MH_ERR_INVALID, "terms without masks or coefficients");
        const uint64_t lim = n >= 64 ? ~uint64_t{0} : (uint64_t{1} << n) - 1;
        for (uint32_t t = 0; t < terms; ++t)
            if ((job->x_masks[t] | job->z_masks[t]) & ~lim) fail(MH_ERR_INVALID, "a Pauli term acts outside the circuit");
        if (terms) {
            j.x.assign(job->x_masks, job->x_masks + terms);
            j.z.assign(job->z_masks, job->z_masks + terms);
            j.coefficients.assign(job->coefficients, job->coefficients + 2 * static_cast<std::size_t>(terms));
        }
    } else {
        if (job->precision != MH_C128) fail(MH_ERR_INVALID, "a sampler job runs at MH_C128");
        if (job->shots == 0) fail(MH_ERR_INVALID, "a sampler job needs shots");
        if (job->bitstring_width > 64) fail(MH_ERR_INVALID, "bitstring_width above 64");
        if (job->num_targets > 0 && (job->target_qubits == nullptr || job->target_clbits == nullptr))
            fail(MH_ERR_INVALID, "null measure targets");
        for (uint32_t t = 0; t < job->num_targets; ++t) {
            if (job->target_qubits[t] >= n || job->target_clbits[t] >= std::max<uint32_t>(job->bitstring_width, 1))
                fail(MH_ERR_INVALID, "measure target out of range");
            j.targets.emplace_back(job->target_qubits[t], job->target_clbits[t]);
        }
        j.shots = job->shots;
        j.seed = job->seed;
    }
}

void Server::write_result(const Job& j, mh_job_result* out) {
    out->status = j.status;
    std::memset(out->error, 0, sizeof(out->error));
    std::strncpy(out->error, j.error.c_str(), sizeof(out->error) - 1);
    out->round = j.round;
    out->lane = j.lane;
    out->plan_cache_hit = j.cache_hit ? 1 : 0;
    out->submitted_seconds = j.submitted;
    out->started_seconds = j.started;
    out->completed_seconds = j.completed;
    out->num_outcomes = 0;
    if (j.status != MH_OK) return;
    if (out->observable_values && !j.values.empty())
        std::memcpy(out->observable_values, j.values.data(), j.values.size() * sizeof(double));
    if (out->raw_terms && !j.raw_terms.empty())
        std::memcpy(out->raw_terms, j.raw_terms.data(), j.raw_terms.size() * sizeof(double));
    if (out->outcomes && out->counts) {
        uint64_t at = 0;
        for (const auto& [k, v] : j.counts) {
            out->outcomes[at] = k;
            out->counts[at] = v;
            ++at;
        }
        out->num_outcomes = at;
    }
}

void Server::finish_locked(Job& j, int32_t status, const std::string& error) {
    j.status = status;
    j.error = error;
    j.completed = now();
    j.done = true;
    if (status == MH_OK)
        ++completed_;
    else
        ++failed_;
}

// First fit over the segments' free ranges.
bool Server::take_range_locked(uint64_t bytes, unsigned char*& segment, uint64_t& offset) {
    for (Segment& g : segments_) {
        std::map<uint64_t, uint64_t>& free = g.free;
        for (auto it = free.begin(); it != free.end(); ++it) {
            if (it->second < bytes) continue;
            segment = g.base;
            offset = it->first;
            const uint64_t rest = it->second - bytes;
            free.erase(it);
            if (rest) free[offset + bytes] = rest;
            return true;
        }
    }
    return false;
}

void Server::free_range_locked(unsigned char* segment, uint64_t offset, uint64_t bytes) {
    auto seg = std::find_if(segments_.begin(), segments_.end(), [&](const Segment& g) { return g.base == segment; });
    std::map<uint64_t, uint64_t>& free = seg->free;
    auto it = free.emplace(offset, bytes).first;
    if (it != free.begin()) {
        auto prev = std::prev(it);
        if (prev->first + prev->second == it->first) {
            prev->second += it->second;
            free.erase(it);
            it = prev;
        }
    }
    auto next = std::next(it);
    if (next != free.end() && it->first + it->second == next->first) {
        it->second += next->second;
        free.erase(next);
    }
}

void Server::add_segment_locked(uint64_t bytes) {
    Segment g;
    g.base = static_cast<unsigned char*>(backend_->segment_alloc(bytes));
    g.bytes = bytes;
    g.free[0] = bytes;
    segments_.push_back(std::move(g));
    arena_bytes_ += bytes;
}

// Admits the longest arrival-order run of the queue whose states fit both
// the budget the other lanes leave and one free range of the arena. With
// large_job_bytes set, a large job is admitted only at the head of a round,
// as a round of its own, and only while no other large round runs; until
// then it keeps its place and the smaller jobs behind it may pass. When no
// range fits, a segment is added (at least as large as the arena so far, up to
// the budget in all), after releasing segments no round holds if the budget
// requires it.
bool Server::admit_locked(Batch& b) {
    std::vector<uint64_t> sizes;
    std::vector<std::size_t> picked;  // queue positions, ascending
    uint64_t total = 0;
    const uint64_t room = budget_ - std::min(budget_, held_);
    const uint64_t large = options_.large_job_bytes;
    for (std::size_t i = 0; i < queue_.size(); ++i) {
        const Job& j = *queue_[i];
        const uint64_t s = backend_->state_bytes(j.circuit.view.num_qubits, j.precision);
        const bool is_large = large != 0 && s >= large;
        if (is_large && large_running_ > 0) continue;
        if (is_large && !picked.empty()) break;
        if (s > budget_ || total + s > room) break;
        if (options_.max_round_jobs && sizes.size() >= options_.max_round_jobs) break;
        sizes.push_back(s);
        picked.push_back(i);
        total += s;
        if (is_large) {
            b.large = true;
            break;
        }
    }
    while (!sizes.empty()) {
        unsigned char* segment = nullptr;
        uint64_t offset = 0;
        if (take_range_locked(total, segment, offset)) {
            b.segment = segment;
            b.offset = offset;
            b.bytes = total;
            b.base = segment + offset;
            break;
        }
        if (arena_bytes_ + total > budget_) {
            for (std::size_t g = segments_.size(); g-- > 0;) {
                const Segment& seg = segments_[g];
                if (seg.free.size() == 1 && seg.free.begin()->second == seg.bytes) {
                    backend_->segment_free(seg.base);
                    arena_bytes_ -= seg.bytes;
                    segments_.erase(segments_.begin() + static_cast<std::ptrdiff_t>(g));
                }
            }
        }
        if (arena_bytes_ + total <= budget_) {
            const uint64_t want =
                std::max(total, std::min(budget_ - arena_bytes_, std::max(kMinSegment, arena_bytes_)));
            const double a0 = now();
            add_segment_locked(want);  // throws on failure, leaving the arena as it was
            arena_seconds_ += now() - a0;
            continue;
        }
        total -= sizes.back();
        sizes.pop_back();
        picked.pop_back();
        if (picked.empty()) b.large = false;
    }
    if (sizes.empty()) return false;
    for (std::size_t k : picked) b.jobs.push_back(std::move(queue_[k]));
    for (std::size_t k = picked.size(); k-- > 0;) queue_.erase(queue_.begin() + static_cast<std::ptrdiff_t>(picked[k]));
    if (b.large) ++large_running_;
    held_ += b.bytes;
    peak_held_ = std::max(peak_held_, held_);
    peak_round_bytes_ = std::max(peak_round_bytes_, b.bytes);
    return true;
}

void Server::lane_loop(uint32_t lane) {
    std::shared_ptr<void> setup;
    int32_t setup_status = MH_OK;
    std::string setup_error;
    try {
        setup = backend_->lane_setup();
    } catch (...) {
        status_of(std::current_exception(), setup_status, setup_error);
    }
    std::unique_lock<std::mutex> lock(mu_);
    for (;;) {
        Batch b;
        bool admitted = false;
        for (;;) {
            // A job whose state alone exceeds the budget fails as it reaches
            // the head (the reference's DeviceBuffer refuses before allocating).
            while (!queue_.empty()) {
                Job& h = *queue_.front();
                const uint64_t s = backend_->state_bytes(h.circuit.view.num_qubits, h.precision);
                if (s <= budget_) break;
                finish_locked(h, MH_ERR_RESOURCE, "the state needs " + std::to_string(s) +
                                                      " device bytes, budget is " + std::to_string(budget_));
                queue_.pop_front();
                done_cv_.notify_all();
            }
            if (!queue_.empty()) {
                try {
                    admitted = admit_locked(b);
                } catch (...) {
                    int32_t st = MH_ERR_INTERNAL;
                    std::string what;
                    status_of(std::current_exception(), st, what);
                    finish_locked(*queue_.front(), st, what);
                    queue_.pop_front();
                    done_cv_.notify_all();
                    continue;
                }
                if (admitted) break;
            }
            if (stop_ && queue_.empty()) break;
            work_cv_.wait(lock);
        }
        if (!admitted) break;
        b.number = ++rounds_;
        ++running_;
        peak_overlap_ = std::max(peak_overlap_, running_);
        lock.unlock();
        const double r0 = now();
        if (setup_status == MH_OK) {
            run_batch(lane, b);  // completes every job of the batch
        } else {
            lock.lock();
            for (JobPtr& j : b.jobs) finish_locked(*j, setup_status, setup_error);
            done_cv_.notify_all();
            lock.unlock();
        }
        const double r1 = now();
        lock.lock();
        --running_;
        if (b.large) --large_running_;
        busy_seconds_ += r1 - r0;
        held_ -= b.bytes;
        free_range_locked(b.segment, b.offset, b.bytes);
        done_cv_.notify_all();
        work_cv_.notify_all();  // room for the jobs waiting on memory
    }
    work_cv_.notify_all();
}

// Plans, executes and combines one batch. Every job of the batch is finished
// (under the lock) exactly once: the small part when it is done, the rest at
// the end; a lane touches no job's fields after finishing it.
void Server::run_batch(uint32_t lane, Batch& b) {
    const std::size_t n = b.jobs.size();
    const double started = now();
    std::vector<int32_t> status(n, MH_OK);
    std::vector<std::string> error(n);
    std::vector<char> finished(n, 0);
    double planned = started;
    PhaseSeconds phases;
    unsigned char* const arena = b.base;
    try {
        for_each_index(n, threads_, [&](std::size_t i) {
            Job& j = *b.jobs[i];
            j.round = b.number;
            j.lane = lane;
            j.started = started;
            try {
                const int32_t fusion =
                    j.precision == MH_C64 && options_.complex64_pairs ? int32_t{MH_FUSE_PAIRS} : options_.fusion;
                const std::string key = structure_key(&j.circuit.view, fusion, options_.fusion_width);
                std::shared_ptr<const void> plan = cache_.find(key);
                j.cache_hit = plan != nullptr;
                backend_->plan(j, fusion, options_.fusion_width, plan);
                if (!j.cache_hit && plan) cache_.insert(key, plan);
            } catch (...) {
                status_of(std::current_exception(), status[i], error[i]);
            }
        });
        planned = now();
        // What the parts need of each job, read before any job is finished.
        std::vector<int32_t> precision_of(n);
        std::vector<uint64_t> bytes_of(n);
        uint64_t largest = 0;
        for (std::size_t i = 0; i < n; ++i) {
            precision_of[i] = b.jobs[i]->precision;
            bytes_of[i] = backend_->state_bytes(b.jobs[i]->circuit.view.num_qubits, precision_of[i]);
            if (status[i] == MH_OK) largest = std::max(largest, bytes_of[i]);
        }
        uint64_t at = 0;
        for (int part = 0; part < 2; ++part) {
            std::vector<std::size_t> members;
            for (const int32_t precision : {int32_t{MH_C128}, int32_t{MH_C64}}) {
                std::vector<std::size_t> group;
                uint64_t group_bytes = 0;
                for (std::size_t i = 0; i < n; ++i) {
                    if (finished[i] || status[i] != MH_OK || precision_of[i] != precision) continue;
                    if ((bytes_of[i] * 8 <= largest) != (part == 0)) continue;
                    group.push_back(i);
                    group_bytes += bytes_of[i];
                }
                if (group.empty()) continue;
                members.insert(members.end(), group.begin(), group.end());
                std::vector<Job*> jobs;
                for (std::size_t i : group) jobs.push_back(b.jobs[i].get());
                try {
                    backend_->execute(precision, jobs, arena + at, threads_, phases);
                    for (Job* j : jobs)
                        if (j->kind == MH_JOB_ESTIMATOR) combine_observables(*j);
                } catch (...) {
                    int32_t st = MH_ERR_INTERNAL;
                    std::string what;
                    status_of(std::current_exception(), st, what);
                    for (std::size_t i : group) {
                        status[i] = st;
                        error[i] = what;
                    }
                }
                at += group_bytes;
            }
            if (part == 0 && !members.empty()) {
                std::lock_guard<std::mutex> lock(mu_);
                for (std::size_t i : members) {
                    finish_locked(*b.jobs[i], status[i], error[i]);
                    finished[i] = 1;
                }
                done_cv_.notify_all();
            }
        }
    } catch (...) {
        int32_t st = MH_ERR_INTERNAL;
        std::string what;
        status_of(std::current_exception(), st, what);
        for (std::size_t i = 0; i < n; ++i)
            if (!finished[i] && status[i] == MH_OK) {
                status[i] = st;
                error[i] = what;
            }
    }
    std::lock_guard<std::mutex> lock(mu_);
    plan_seconds_ += planned - started;
    phases_.compile += phases.compile;
    phases_.pass += phases.pass;
    phases_.expectation += phases.expectation;
    phases_.sample += phases.sample;
    for (std::size_t i = 0; i < n; ++i)
        if (!finished[i]) finish_locked(*b.jobs[i], status[i], error[i]);
    done_cv_.notify_all();
}

}  // namespace mh::serve
