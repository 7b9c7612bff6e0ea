// HipDenseEngine: the MI355X dense statevector engine behind MENOS's
// SimulationEngine contract (reference include/menos/simulation_engine.h:322-367).
//
// It is the {Dense, Hip, InMemory} engine: estimate() never throws and reports
// an infeasible request with a reason; execute() takes the region's
// operations and output request and returns the statevector, counts (sole
// region) and the raw, coefficient-free expectation terms, charging the
// device bytes it holds to the calling thread's ledger. The request and result
// shapes mirror EngineExecutionRequest / EngineExecutionResult
// (simulation_engine.h); the C ABI (menos_hip.h) carries them across.
#pragma once

#include <cstdint>
#include <map>
#include <optional>
#include <string>
#include <vector>

#include "mh_internal.h"

namespace mh {

// The calling thread's device ledger (reference src/common/memory_ledger.h:37-117,
// Space::Device): live and peak bytes the engine holds on the device.
class DeviceLedger {
public:
    static DeviceLedger& current();
    void charge(uint64_t bytes);
    void release(uint64_t bytes);
    void reset();
    uint64_t peak() const { return peak_; }
    uint64_t live() const { return live_; }

private:
    uint64_t live_ = 0;
    uint64_t peak_ = 0;
};

// A device allocation charged to the ledger for as long as it lives.
class ChargedBuffer {
public:
    explicit ChargedBuffer(uint64_t bytes) : buf_(bytes), bytes_(bytes) { DeviceLedger::current().charge(bytes); }
    ~ChargedBuffer() { DeviceLedger::current().release(bytes_); }
    ChargedBuffer(const ChargedBuffer&) = delete;
    ChargedBuffer& operator=(const ChargedBuffer&) = delete;
    void* get() const { return buf_.get(); }
    // The memory, handed over: no longer charged to this call's ledger.
    DeviceBuffer release() {
        DeviceLedger::current().release(bytes_);
        bytes_ = 0;
        return std::move(buf_);
    }

private:
    DeviceBuffer buf_;
    uint64_t bytes_;
};

struct ExecutionRequest {
    const mh_circuit* circuit = nullptr;
    int32_t precision = MH_C128;
    int32_t fusion = MH_FUSE_CONTAINMENT;
    uint32_t fusion_width = 0;
    uint32_t window_qubits = 0;
    double global_phase = 0.0;
    bool want_statevector = false;
    uint64_t shots = 0;
    uint64_t seed = 0;
    std::vector<std::pair<uint32_t, uint32_t>> measure_targets;
    uint32_t bitstring_width = 0;
    std::vector<std::pair<uint64_t, uint64_t>> terms;  // (x, z)
    uint64_t budget_bytes = 0;
    bool record_pass_times = false;
    uint32_t max_pass_cost = 0;
    mh_monitor_fn monitor = nullptr;
    void* monitor_data = nullptr;
    bool want_region_sampler = false;
};

struct ExecutionResult {
    std::optional<std::vector<unsigned char>> statevector;  // raw amplitudes of the request's precision
    std::optional<std::map<uint64_t, uint64_t>> counts;
    std::optional<std::vector<cplx>> expectation_terms;
    uint64_t actual_bytes = 0;
    uint64_t peak_device_bytes = 0;
    uint32_t num_blocks = 0;
    uint32_t num_passes = 0;
    uint64_t fused_blocks = 0, gates_absorbed = 0;
    std::vector<float> pass_ms;
    double execute_seconds = 0, gate_seconds = 0, sample_seconds = 0, expectation_seconds = 0;
    int32_t fusion_used = MH_FUSE_CONTAINMENT;
    uint32_t width_used = 0;
    mh_region_sampler* region_sampler = nullptr;  // want_region_sampler: the caller's to destroy
};

struct CostEstimate {
    bool feasible = false;
    std::string infeasible_reason;
    uint64_t device_bytes = 0, host_bytes = 0, work_units = 0;
    uint32_t blocks = 0, passes = 0;
    uint64_t launches = 0, h2d_bytes = 0, d2h_bytes = 0;
    double seconds = 0;
    int32_t fusion_used = MH_FUSE_CONTAINMENT;
    uint32_t width_used = 0;
};

// Draws `shots` outcomes from a device state and folds them into counts keyed
// by the clbit string (reference src/engines/dense/sample.cpp:24-75):
// mt19937_64(seed) words as uniforms, sorted, then the device cumulative walk;
// outcome bit q lands on clbit c for every (q, c) target.
std::map<uint64_t, uint64_t> sample_counts(const void* state, uint32_t n, int32_t precision, uint64_t shots,
                                           uint64_t seed, const std::vector<std::pair<uint32_t, uint32_t>>& targets);

// The blocks a request runs: its fusion rule's, or the autotuner's choice
// (containment, then widths 2..6, priced by the time model).
struct Plan {
    FusionPlan fusion;
    std::vector<Block> blocks;
    int32_t fusion_used = MH_FUSE_CONTAINMENT;
    uint32_t width_used = 0;
};
Plan make_plan(const ExecutionRequest& request);

// The installed time-model profile (mh_set_calibration, mh_calibrate).
const mh_calibration& calibration();

// The sampler's uniforms: mt19937_64(seed) words, 53 bits each, sorted
// (reference src/engines/dense/sample.cpp:24-75).
std::vector<double> sorted_uniforms(uint64_t seed, uint64_t shots);
// Basis indices folded into counts keyed by the clbit string (bit q of an
// index lands on clbit c for every (q, c) target).
std::map<uint64_t, uint64_t> fold_counts(const std::vector<uint64_t>& basis,
                                         const std::vector<std::pair<uint32_t, uint32_t>>& targets);

// The C request as the engine's, and the engine's result into the caller's
// buffers (ownership of a region sampler passes to the caller).
ExecutionRequest to_request(const mh_execution_request* request);
void write_result(ExecutionResult& result, mh_execution_result* out);

class HipDenseEngine {
public:
    const char* name() const { return "dense_hip"; }
    CostEstimate estimate(const ExecutionRequest& request) const noexcept;
    ExecutionResult execute(const ExecutionRequest& request) const;
};

}  // namespace mh
