// The job server's host core: the queue, admission against the ledger
// budget, the lanes that run admitted batches, the plan cache and the result
// bookkeeping. It holds no device code: the device work is a ServerBackend's
// (serve.cpp runs it on the batched round executor), so the core also builds
// and runs on its own, with a stub backend, under AddressSanitizer and
// ThreadSanitizer.
//
// Serving here is continuous batching. Several lanes, each a thread with its
// own streams, take jobs from one arrival-ordered queue whenever they are
// idle: a lane admits the queued jobs whose states fit the arena space the
// other lanes leave free (the memory ledger: at most the budget in all), runs
// them as one batch and completes them, while jobs that arrive meanwhile are
// admitted by the next idle lane and overlap on the device with the batches
// already running. Within a batch, the jobs whose state is at most an eighth
// of its largest run first and complete before the rest start.
//
// Lifetime. The job table and every batch hold shared_ptr<Job>, so a waiter
// that collects a finished job never frees it under a lane; a lane writes a
// job's fields only before it finishes the job (under the lock), and keeps the
// jobs' completion in a batch-local table instead of reading the job back.
#pragma once

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <utility>
#include <vector>

#include "menos_hip.h"

namespace mh::serve {

// A deep copy of a submitted circuit.
struct OwnedCircuit {
    std::vector<int32_t> opcodes;
    std::vector<uint32_t> qubit_offsets, qubits, param_offsets;
    std::vector<double> params;
    mh_circuit view{};

    void copy_from(const mh_circuit* c);
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
    // The backend's per-job plan (fusion plan and block matrices), between
    // planning and execution.
    std::shared_ptr<void> work;
    // Results, written by the lane before it finishes the job.
    int32_t status = MH_OK;
    std::string error;
    std::vector<double> raw_terms;   // (re, im) per term
    std::vector<double> values;      // (re, im) per observable
    std::map<uint64_t, uint64_t> counts;
    uint32_t round = 0;
    uint32_t lane = 0;
    bool cache_hit = false;
    double submitted = 0, started = 0, completed = 0;
    // Under the server's lock.
    bool done = false;
    bool waited = false;

    uint32_t num_terms() const { return static_cast<uint32_t>(x.size()); }
};

// The reference's combine_expectation for one region
// (src/runtime/combine_expectation.cpp:7-25): total starts at (0, 0); per
// term, product = (1, 0) * term, then total += coefficient * product, with the
// naive complex product (a*c - b*d, a*d + b*c) that std::complex uses for
// finite values.
void combine_observables(Job& j);

// What a fusion plan depends on (the reference's plan_cache_key,
// src/plan/plan_cache.cpp:150-227, restricted to what this engine's planner
// reads): the qubit count and, per operation, its opcode, qubits and parameter
// count; a unitary's entries byte for byte, since its diagonality steers fusion.
std::string structure_key(const mh_circuit* c, int32_t fusion, uint32_t width);

// Least-recently-used fusion plans by structure key; the plans are the
// backend's, opaque here.
class PlanCache {
public:
    explicit PlanCache(std::size_t capacity) : capacity_(capacity) {}
    std::shared_ptr<const void> find(const std::string& key);
    void insert(const std::string& key, std::shared_ptr<const void> plan);
    uint64_t hits() const;
    uint64_t misses() const;

private:
    using Entry = std::pair<std::string, std::shared_ptr<const void>>;
    mutable std::mutex mu_;
    std::size_t capacity_;
    std::list<Entry> lru_;
    std::unordered_map<std::string, std::list<Entry>::iterator> index_;
    uint64_t hits_ = 0, misses_ = 0;
};

// Wall time of a lane's device phases (the backend adds to them).
struct PhaseSeconds {
    double compile = 0, pass = 0, expectation = 0, sample = 0;
};

// The device side of the server.
class ServerBackend {
public:
    virtual ~ServerBackend() = default;
    // Device bytes the server may hold (the budget is capped to them).
    virtual uint64_t capacity_bytes() = 0;
    // Throws Error(MH_ERR_INVALID) for a malformed circuit.
    virtual void validate_circuit(const mh_circuit* c) = 0;
    // The fewest and most qubits a served circuit of this precision may have.
    virtual uint32_t min_qubits(int32_t precision) = 0;
    virtual uint32_t max_qubits() = 0;
    // Device bytes one state takes in a batch.
    virtual uint64_t state_bytes(uint32_t n, int32_t precision) = 0;
    // Runs on each lane's thread before its first batch; what it returns lives
    // as long as the lane (its streams).
    virtual std::shared_ptr<void> lane_setup() = 0;
    // Device memory for one segment of the server's arena, and its release
    // (no lane holds any of it then).
    virtual void* segment_alloc(uint64_t bytes) = 0;
    virtual void segment_free(void* segment) noexcept = 0;
    // The job's fusion plan: `plan` comes from the cache or is null, in which
    // case the backend makes it (and the core caches it); fills job.work.
    virtual void plan(Job& job, int32_t fusion, uint32_t width, std::shared_ptr<const void>& plan) = 0;
    // Runs jobs of one precision with their states laid out from `arena`
    // (state_bytes each, in order): fills raw_terms and counts, clears work.
    virtual void execute(int32_t precision, const std::vector<Job*>& jobs, void* arena, uint32_t threads,
                         PhaseSeconds& seconds) = 0;
};

class Server {
public:
    Server(const mh_server_options& options, std::unique_ptr<ServerBackend> backend);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    double now() const;
    // Validates and copies the job, queues it; returns its ticket.
    uint64_t submit(const mh_job* job);
    // Blocks until the ticket's job is complete; writes its result.
    void wait(uint64_t ticket, mh_job_result* out);
    void stats(mh_server_stats* s);
    // Submits job i at arrival_seconds[i] from a submitter thread, then waits
    // for every job.
    void replay(uint32_t count, const mh_job* jobs, const double* arrival_seconds, mh_job_result* results);

private:
    using JobPtr = std::shared_ptr<Job>;
    struct Batch {
        std::vector<JobPtr> jobs;
        uint64_t bytes = 0;
        unsigned char* segment = nullptr;  // its segment's base
        uint64_t offset = 0;                // in the segment
        unsigned char* base = nullptr;      // segment + offset
        uint32_t number = 0;
        bool large = false;                 // holds a job of at least large_job_bytes
    };
    // A segment of the arena and its free ranges (offset -> bytes).
    struct Segment {
        unsigned char* base = nullptr;
        uint64_t bytes = 0;
        std::map<uint64_t, uint64_t> free;
    };

    void validate_and_copy(const mh_job* job, Job& j);
    static void write_result(const Job& j, mh_job_result* out);
    void finish_locked(Job& j, int32_t status, const std::string& error);
    void lane_loop(uint32_t lane);
    bool admit_locked(Batch& b);
    bool take_range_locked(uint64_t bytes, unsigned char*& segment, uint64_t& offset);
    void free_range_locked(unsigned char* segment, uint64_t offset, uint64_t bytes);
    void add_segment_locked(uint64_t bytes);
    void run_batch(uint32_t lane, Batch& b);

    mh_server_options options_;
    std::unique_ptr<ServerBackend> backend_;
    PlanCache cache_;
    std::chrono::steady_clock::time_point t0_;
    uint64_t budget_ = 0;
    uint32_t threads_ = 1;
    uint32_t large_running_ = 0;  // rounds holding a large job that are running
    uint32_t lanes_ = 1;

    std::mutex mu_;
    std::condition_variable work_cv_, done_cv_;
    std::deque<JobPtr> queue_;
    std::unordered_map<uint64_t, JobPtr> jobs_;
    uint64_t last_ticket_ = 0;
    bool stop_ = false;
    // The arena (segments added as the load needs them, at most the budget
    // in all) and the ledger over it: free ranges per segment, bytes held.
    std::vector<Segment> segments_;
    uint64_t arena_bytes_ = 0;
    uint64_t held_ = 0;
    uint32_t running_ = 0;       // batches in flight
    // Statistics.
    uint64_t completed_ = 0, failed_ = 0;
    uint32_t rounds_ = 0;
    uint32_t peak_overlap_ = 0;
    uint64_t peak_round_bytes_ = 0, peak_held_ = 0;
    double busy_seconds_ = 0, plan_seconds_ = 0, arena_seconds_ = 0;
    PhaseSeconds phases_;
    std::vector<std::thread> workers_;
};

}  // namespace mh::serve
