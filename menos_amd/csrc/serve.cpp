// The serving loop (mh_server_*): a resident job server over the batched
// round executor.
//
// The reference serves a stream of estimator and sampler requests from its
// Python primitives, one full `run` per observable or pub
// (python/menos/qiskit/estimator.py:91-146, sampler.py:93-129), each run
// planning (with a structure-keyed plan cache, src/plan/plan_cache.cpp:233-276)
// and executing alone. Here a worker thread owns the device:
//
//   1. submit() validates and copies a job into the queue, stamping its time;
//   2. the worker takes queued jobs in arrival order while their states fit
//      the budget (a job larger than the budget alone fails with
//      MH_ERR_RESOURCE), up to max_round_jobs;
//   3. it plans them on host threads: the fusion plan comes from a cache keyed
//      by circuit structure (a hit is a parameter-only update of a circuit
//      seen before), block matrices are built from the job's own parameters;
//   4. one execute_round per precision runs every state of the round in the
//      server's arena: pass p of every circuit in one launch, all terms in two
//      launches, each sampler's draw;
//   5. raw terms are combined into observable values as the reference's
//      combine_expectation does, and the jobs complete.
//
// Jobs arriving while a round runs form the next round, so the batch size
// follows the load.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>

#include "batch.h"
#include "engine.hpp"
#include "mh_internal.h"
#include "window.h"

namespace mh {
namespace {

using clock_type = std::chrono::steady_clock;

constexpr uint64_t kDefaultBudget = 256'000'000'000ULL;  // the ledger budget: 256 GB of the 288 GB
constexpr uint32_t kDefaultCacheCapacity = 256;
constexpr uint32_t kMaxJobQubits = 30;

// A deep copy of a submitted circuit.
struct OwnedCircuit {
    std::vector<int32_t> opcodes;
    std::vector<uint32_t> qubit_offsets, qubits, param_offsets;
    std::vector<double> params;
    mh_circuit view{};

    void copy_from(const mh_circuit* c) {
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
};

template <typename T>
void put(std::string& key, const T& v) {
    key.append(reinterpret_cast<const char*>(&v), sizeof(v));
}

// What the fusion plan depends on (the reference's plan_cache_key,
// plan_cache.cpp:150-227, restricted to what this engine's planner reads):
// the qubit count and, per operation, its opcode, qubits and parameter count;
// a unitary's entries byte for byte, since its diagonality steers fusion.
std::string structure_key(const mh_circuit* c, int32_t fusion, uint32_t width) {
    std::string key;
    key.reserve(32 + 12 * static_cast<std::size_t>(c->num_ops));
    put(key, uint32_t{1});  // key format
    put(key, fusion);
    put(key, width);
    put(key, c->num_qubits);
    put(key, c->num_ops);
    const CircuitView cv{c};
    for (uint32_t i = 0; i < c->num_ops; ++i) {
        put(key, cv.opcode(i));
        put(key, cv.num_qubits_of(i));
        key.append(reinterpret_cast<const char*>(cv.qubits(i)), sizeof(uint32_t) * cv.num_qubits_of(i));
        put(key, cv.num_params(i));
        if (cv.opcode(i) == kOpUnitary)
            key.append(reinterpret_cast<const char*>(cv.params(i)), sizeof(double) * cv.num_params(i));
    }
    return key;
}

// Least-recently-used fusion plans by structure key.
class PlanCache {
public:
    explicit PlanCache(std::size_t capacity) : capacity_(capacity) {}

    std::shared_ptr<const FusionPlan> find(const std::string& key) {
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

    void insert(const std::string& key, std::shared_ptr<const FusionPlan> plan) {
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

    uint64_t hits() const { std::lock_guard<std::mutex> lock(mu_); return hits_; }
    uint64_t misses() const { std::lock_guard<std::mutex> lock(mu_); return misses_; }

private:
    using Entry = std::pair<std::string, std::shared_ptr<const FusionPlan>>;
    mutable std::mutex mu_;
    std::size_t capacity_;
    std::list<Entry> lru_;
    std::unordered_map<std::string, std::list<Entry>::iterator> index_;
    uint64_t hits_ = 0, misses_ = 0;
};

struct Job {
    uint64_t ticket = 0;
    int32_t kind = MH_JOB_ESTIMATOR;
    int32_t precision = MH_C128;
    OwnedCircuit circuit;
    double global_phase = 0;
    std::vector<uint32_t> observable_offsets;
    std::vector<uint64_t> x, z;
    std::vector<double> coefficients;
    uint64_t shots = 0, seed = 0;
    std::vector<std::pair<uint32_t, uint32_t>> targets;
    // Planning.
    std::vector<Block> blocks;
    // Results.
    bool done = false;
    bool waited = false;
    int32_t status = MH_OK;
    std::string error;
    std::vector<double> raw_terms;   // (re, im) per term
    std::vector<double> values;      // (re, im) per observable
    std::map<uint64_t, uint64_t> counts;
    uint32_t round = 0;
    bool cache_hit = false;
    double submitted = 0, started = 0, completed = 0;

    uint32_t num_terms() const { return static_cast<uint32_t>(x.size()); }
};

// The reference's combine_expectation for one region (combine_expectation.cpp:7-25):
// total starts at (0, 0); per term, product = (1, 0) * term, then
// total += coefficient * product, with the naive complex product
// (a*c - b*d, a*d + b*c) that std::complex uses for finite values.
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

class Server {
public:
    explicit Server(const mh_server_options& o)
        : options_(o), cache_(o.plan_cache_capacity ? o.plan_cache_capacity : kDefaultCacheCapacity),
          t0_(clock_type::now()) {
        budget_ = o.budget_bytes ? o.budget_bytes : kDefaultBudget;
        size_t free_b = 0, total_b = 0;
        MH_HIP(hipMemGetInfo(&free_b, &total_b));
        budget_ = std::min<uint64_t>(budget_, static_cast<uint64_t>(free_b) + pool_reusable_bytes());
        threads_ = o.host_threads ? o.host_threads : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        if (o.reserve_bytes) arena_ = DeviceBuffer(std::min(o.reserve_bytes, budget_));
        worker_ = std::thread([this] { loop(); });
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> lock(mu_);
            stop_ = true;
        }
        work_cv_.notify_all();
        worker_.join();
    }

    double now() const { return std::chrono::duration<double>(clock_type::now() - t0_).count(); }

    uint64_t submit(const mh_job* job) {
        auto j = std::make_unique<Job>();
        validate_and_copy(job, *j);
        std::lock_guard<std::mutex> lock(mu_);
        j->ticket = ++last_ticket_;
        j->submitted = now();
        const uint64_t t = j->ticket;
        queue_.push_back(j.get());
        jobs_.emplace(t, std::move(j));
        work_cv_.notify_one();
        return t;
    }

    void wait(uint64_t ticket, mh_job_result* out) {
        std::unique_ptr<Job> j;
        {
            std::unique_lock<std::mutex> lock(mu_);
            auto it = jobs_.find(ticket);
            if (it == jobs_.end() || it->second->waited) fail(MH_ERR_INVALID, "unknown or already collected ticket");
            Job* p = it->second.get();
            p->waited = true;
            done_cv_.wait(lock, [&] { return p->done; });
            it = jobs_.find(ticket);  // other submissions may have rehashed the map
            j = std::move(it->second);
            jobs_.erase(it);
        }
        write_result(*j, out);
    }

    void stats(mh_server_stats* s) {
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
        s->compile_seconds = phases_.compile_seconds;
        s->pass_seconds = phases_.pass_seconds;
        s->expectation_seconds = phases_.expectation_seconds;
        s->sample_seconds = phases_.sample_seconds;
    }

private:
    void validate_and_copy(const mh_job* job, Job& j) const {
        if (job == nullptr) fail(MH_ERR_INVALID, "null job");
        if (job->kind != MH_JOB_ESTIMATOR && job->kind != MH_JOB_SAMPLER) fail(MH_ERR_INVALID, "unknown job kind");
        amp_bytes(job->precision);
        validate_circuit(job->circuit);
        const uint32_t n = job->circuit->num_qubits;
        const uint32_t width = window_width(job->precision);
        if (n < width || n > kMaxJobQubits)
            fail(MH_ERR_INVALID, "a served circuit needs " + std::to_string(width) + " to " +
                                     std::to_string(kMaxJobQubits) + " qubits");
        j.kind = job->kind;
        j.precision = job->precision;
        j.global_phase = job->global_phase;
        j.circuit.copy_from(job->circuit);
        if (job->kind == MH_JOB_ESTIMATOR) {
            if (job->num_observables > 0 && job->observable_offsets == nullptr)
                fail(MH_ERR_INVALID, "observables without offsets");
            j.observable_offsets.assign(job->observable_offsets ? job->observable_offsets : nullptr,
                                        job->observable_offsets ? job->observable_offsets + job->num_observables + 1
                                                                : nullptr);
            if (j.observable_offsets.empty()) j.observable_offsets.push_back(0);
            if (j.observable_offsets[0] != 0) fail(MH_ERR_INVALID, "observable_offsets must start at 0");
            for (std::size_t o = 1; o < j.observable_offsets.size(); ++o)
                if (j.observable_offsets[o] < j.observable_offsets[o - 1])
                    fail(MH_ERR_INVALID, "observable_offsets must not decrease");
            const uint32_t terms = j.observable_offsets.back();
            if (terms > 0 && (job->x_masks == nullptr || job->z_masks == nullptr || job->coefficients == nullptr))
                fail(MH_ERR_INVALID, "terms without masks or coefficients");
            const uint64_t lim = n == 64 ? ~uint64_t{0} : (uint64_t{1} << n) - 1;
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

    static void write_result(const Job& j, mh_job_result* out) {
        out->status = j.status;
        std::memset(out->error, 0, sizeof(out->error));
        std::strncpy(out->error, j.error.c_str(), sizeof(out->error) - 1);
        out->round = j.round;
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

    void finish(Job* j, int32_t status, const std::string& error) {
        j->status = status;
        j->error = error;
        j->completed = now();
        j->done = true;
        if (status == MH_OK) ++completed_; else ++failed_;
    }

    void loop() {
        std::unique_lock<std::mutex> lock(mu_);
        for (;;) {
            work_cv_.wait(lock, [&] { return stop_ || !queue_.empty(); });
            if (queue_.empty()) break;  // stop_ with nothing left
            std::vector<Job*> round;
            uint64_t bytes = 0;
            while (!queue_.empty()) {
                Job* j = queue_.front();
                const uint64_t b = round_state_bytes(j->circuit.view.num_qubits, j->precision);
                if (b > budget_) {
                    queue_.pop_front();
                    finish(j, MH_ERR_RESOURCE, "the state needs " + std::to_string(b) + " device bytes, budget is " +
                                                   std::to_string(budget_));
                    continue;
                }
                if (!round.empty() && (bytes + b > budget_ ||
                                       (options_.max_round_jobs && round.size() >= options_.max_round_jobs)))
                    break;
                round.push_back(j);
                bytes += b;
                queue_.pop_front();
            }
            if (round.empty()) {
                done_cv_.notify_all();
                continue;
            }
            const uint32_t number = ++rounds_;
            lock.unlock();
            const double r0 = now();
            int32_t round_status = MH_OK;
            std::string round_error;
            try {
                run_round(round, bytes, number);
            } catch (const Error& e) {
                round_status = e.status;
                round_error = e.what();
            } catch (const std::bad_alloc&) {
                round_status = MH_ERR_RESOURCE;
                round_error = "host allocation failed";
            } catch (const std::exception& e) {
                round_status = MH_ERR_INTERNAL;
                round_error = e.what();
            }
            const double r1 = now();
            lock.lock();
            busy_seconds_ += r1 - r0;
            peak_round_bytes_ = std::max(peak_round_bytes_, bytes);
            for (Job* j : round) {
                if (j->done) continue;
                finish(j, round_status, round_error);
            }
            done_cv_.notify_all();
        }
        done_cv_.notify_all();
    }

    // Plans, executes and combines one round; a job that fails is finished
    // with its error (under the lock the caller retakes), the rest go on.
    void run_round(std::vector<Job*>& round, uint64_t bytes, uint32_t number) {
        const double started = now();
        std::vector<std::string> errors(round.size());
        std::vector<int32_t> statuses(round.size(), MH_OK);
        parallel_for(round.size(), threads_, [&](std::size_t i) {
            Job& j = *round[i];
            j.round = number;
            j.started = started;
            try {
                const CircuitView cv{&j.circuit.view};
                const int32_t fusion =
                    j.precision == MH_C64 && options_.complex64_pairs ? int32_t{MH_FUSE_PAIRS} : options_.fusion;
                const std::string key = structure_key(&j.circuit.view, fusion, options_.fusion_width);
                std::shared_ptr<const FusionPlan> fp = cache_.find(key);
                j.cache_hit = fp != nullptr;
                if (!fp) {
                    fp = std::make_shared<const FusionPlan>(plan_fusion(cv, fusion, options_.fusion_width));
                    cache_.insert(key, fp);
                }
                j.blocks = build_blocks(cv, *fp);
            } catch (const Error& e) {
                statuses[i] = e.status;
                errors[i] = e.what();
            } catch (const std::exception& e) {
                statuses[i] = MH_ERR_INTERNAL;
                errors[i] = e.what();
            }
        });
        const double planned = now();
        ensure_arena(bytes);
        const double arena_ready = now();
        RoundStats rs;
        // Parts: jobs whose state is at most an eighth of the round's largest
        // run first and complete at once; the rest follow. A round that holds
        // a large job then keeps its small jobs waiting for their own work only.
        uint64_t largest = 0;
        for (std::size_t i = 0; i < round.size(); ++i)
            if (statuses[i] == MH_OK)
                largest = std::max(largest, round_state_bytes(round[i]->circuit.view.num_qubits, round[i]->precision));
        for (int part = 0; part < 2; ++part) {
            uint64_t arena_at = 0;
            std::vector<Job*> done_here;
            for (const int32_t precision : {MH_C128, MH_C64}) {
                std::vector<Job*> group;
                for (std::size_t i = 0; i < round.size(); ++i) {
                    if (statuses[i] != MH_OK || round[i]->precision != precision) continue;
                    const bool small = round_state_bytes(round[i]->circuit.view.num_qubits, precision) * 8 <= largest;
                    if (small == (part == 0)) group.push_back(round[i]);
                }
                if (group.empty()) continue;
                std::vector<RoundItem> items(group.size());
                uint32_t terms = 0;
                uint64_t group_bytes = 0;
                for (std::size_t k = 0; k < group.size(); ++k) {
                    Job& j = *group[k];
                    RoundItem& it = items[k];
                    it.n = j.circuit.view.num_qubits;
                    it.blocks = &j.blocks;
                    it.global_phase = j.global_phase;
                    it.num_terms = j.num_terms();
                    it.x = j.x.data();
                    it.z = j.z.data();
                    it.term_out = terms;
                    terms += j.num_terms();
                    it.shots = j.shots;
                    it.seed = j.seed;
                    it.targets = &j.targets;
                    it.counts = &j.counts;
                    group_bytes += round_state_bytes(it.n, precision);
                }
                std::vector<double> raw(2 * static_cast<std::size_t>(terms));
                try {
                    execute_round(precision, items, threads_, raw.data(), &rs,
                                  static_cast<unsigned char*>(arena_.get()) + arena_at);
                } catch (const Error& e) {
                    for (Job* j : group) mark_failed(*j, round, statuses, errors, e.status, e.what());
                    continue;
                } catch (const std::exception& e) {
                    for (Job* j : group) mark_failed(*j, round, statuses, errors, MH_ERR_INTERNAL, e.what());
                    continue;
                }
                arena_at += group_bytes;
                for (std::size_t k = 0; k < group.size(); ++k) {
                    Job& j = *group[k];
                    const uint32_t at = items[k].term_out;
                    j.raw_terms.assign(raw.begin() + 2 * at, raw.begin() + 2 * (at + j.num_terms()));
                    if (j.kind == MH_JOB_ESTIMATOR) combine_observables(j);
                    j.blocks.clear();
                    j.blocks.shrink_to_fit();
                    done_here.push_back(&j);
                }
            }
            if (part == 0 && !done_here.empty()) {
                std::lock_guard<std::mutex> lock(mu_);
                for (Job* j : done_here) finish(j, MH_OK, "");
                done_cv_.notify_all();
            }
        }
        std::lock_guard<std::mutex> lock(mu_);
        plan_seconds_ += planned - started;
        arena_seconds_ += arena_ready - planned;
        phases_.compile_seconds += rs.compile_seconds;
        phases_.pass_seconds += rs.pass_seconds;
        phases_.expectation_seconds += rs.expectation_seconds;
        phases_.sample_seconds += rs.sample_seconds;
        for (std::size_t i = 0; i < round.size(); ++i)
            if (statuses[i] != MH_OK) finish(round[i], statuses[i], errors[i]);
    }

    static void mark_failed(const Job& j, const std::vector<Job*>& round, std::vector<int32_t>& statuses,
                            std::vector<std::string>& errors, int32_t status, const std::string& what) {
        for (std::size_t i = 0; i < round.size(); ++i)
            if (round[i] == &j) {
                statuses[i] = status;
                errors[i] = what;
            }
    }

    // The server's arena is kept between rounds; it grows at least fourfold
    // (up to the budget) so that a rising load reallocates it a few times only.
    void ensure_arena(uint64_t bytes) {
        if (arena_.bytes() >= bytes) return;
        const uint64_t grown = std::min(budget_, std::max(bytes, 4 * arena_.bytes()));
        arena_.reset();
        arena_ = DeviceBuffer(grown);
    }

    mh_server_options options_;
    PlanCache cache_;
    clock_type::time_point t0_;
    uint64_t budget_ = 0;
    uint32_t threads_ = 1;
    DeviceBuffer arena_;

    std::mutex mu_;
    std::condition_variable work_cv_, done_cv_;
    std::deque<Job*> queue_;
    std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs_;
    uint64_t last_ticket_ = 0;
    bool stop_ = false;
    uint64_t completed_ = 0, failed_ = 0;
    uint32_t rounds_ = 0;
    uint64_t peak_round_bytes_ = 0;
    double busy_seconds_ = 0;
    double plan_seconds_ = 0;
    double arena_seconds_ = 0;
    RoundStats phases_;
    std::thread worker_;
};

}  // namespace
}  // namespace mh

using namespace mh;

struct mh_server {
    std::unique_ptr<Server> impl;
};

extern "C" {

int mh_server_create(const mh_server_options* options, mh_server** out) {
    return guarded([&] {
        if (out == nullptr) fail(MH_ERR_INVALID, "null output");
        mh_server_options o{};
        if (options) o = *options;
        if (options == nullptr) o.fusion = MH_FUSE_CONTAINMENT;
        if (o.fusion != MH_FUSE_NONE && o.fusion != MH_FUSE_CONTAINMENT && o.fusion != MH_FUSE_WIDTH &&
            o.fusion != MH_FUSE_PAIRS)
            fail(MH_ERR_INVALID, "unknown fusion mode");
        context();
        auto s = std::make_unique<mh_server>();
        s->impl = std::make_unique<Server>(o);
        *out = s.release();
    });
}

int mh_server_submit(mh_server* server, const mh_job* job, uint64_t* ticket) {
    return guarded([&] {
        if (server == nullptr || ticket == nullptr) fail(MH_ERR_INVALID, "null server or ticket");
        *ticket = server->impl->submit(job);
    });
}

int mh_server_wait(mh_server* server, uint64_t ticket, mh_job_result* result) {
    return guarded([&] {
        if (server == nullptr || result == nullptr) fail(MH_ERR_INVALID, "null server or result");
        server->impl->wait(ticket, result);
    });
}

int mh_server_replay(mh_server* server, uint32_t count, const mh_job* jobs, const double* arrival_seconds,
                     mh_job_result* results) {
    return guarded([&] {
        if (server == nullptr || (count > 0 && (jobs == nullptr || arrival_seconds == nullptr || results == nullptr)))
            fail(MH_ERR_INVALID, "null server, jobs, arrivals or results");
        for (uint32_t i = 1; i < count; ++i)
            if (arrival_seconds[i] < arrival_seconds[i - 1]) fail(MH_ERR_INVALID, "arrival times must not decrease");
        Server& s = *server->impl;
        std::vector<uint64_t> tickets(count, 0);
        std::vector<int32_t> submit_status(count, MH_OK);
        std::vector<std::string> submit_error(count);
        const auto start = clock_type::now();
        std::thread submitter([&] {
            for (uint32_t i = 0; i < count; ++i) {
                std::this_thread::sleep_until(start + std::chrono::duration_cast<clock_type::duration>(
                                                          std::chrono::duration<double>(arrival_seconds[i])));
                try {
                    tickets[i] = s.submit(&jobs[i]);
                } catch (const Error& e) {
                    submit_status[i] = e.status;
                    submit_error[i] = e.what();
                } catch (const std::exception& e) {
                    submit_status[i] = MH_ERR_INTERNAL;
                    submit_error[i] = e.what();
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
                r.plan_cache_hit = 0;
                r.submitted_seconds = r.started_seconds = r.completed_seconds = 0;
                continue;
            }
            s.wait(tickets[i], &results[i]);
        }
    });
}

int mh_server_stats_get(mh_server* server, mh_server_stats* out) {
    return guarded([&] {
        if (server == nullptr || out == nullptr) fail(MH_ERR_INVALID, "null server or output");
        server->impl->stats(out);
    });
}

int mh_server_destroy(mh_server* server) {
    return guarded([&] { delete server; });
}

int mh_derive_seeds(uint64_t seed, uint32_t count, uint64_t* out) {
    return guarded([&] {
        if (count > 0 && out == nullptr) fail(MH_ERR_INVALID, "null output");
        uint64_t state = seed;
        for (uint32_t i = 0; i < count; ++i) {
            state += 0x9E3779B97F4A7C15ULL;
            uint64_t z = state;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
            out[i] = z ^ (z >> 31);
        }
    });
}

}  // extern "C"
