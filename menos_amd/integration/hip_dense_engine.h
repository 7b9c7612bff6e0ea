// The reference-side adapter: MENOS's SimulationEngine interface over the
// MI355X engine's C ABI (include/menos_hip.h). This is the code a maintainer
// adds to the reference tree as src/engines/dense_hip/ (beside the CUDA
// backend, src/engines/dense_cuda/cuda_dense_engine.{h,cpp}); it compiles
// against the reference's own headers and links libmenos_hip.so.
//
// It keeps the CUDA engine's contract (include/menos/simulation_engine.h:
// 322-367): estimate() never throws and reports an infeasible request with a
// reason; execute() gives the statevector, the sole region's counts or a
// non-sole region's sampler, and the raw expectation terms; the device's peak
// bytes are charged to the memory ledger installed on the calling thread
// (MemoryLedger::Space::Device); an ExecutionMonitor is called at the engine's
// safe boundaries (after each HBM pass) and abandons by throwing, the
// exception reaching the caller unchanged.
#pragma once

#include <cstdint>
#include <vector>

#include "menos/plan.h"
#include "menos/simulation_engine.h"
#include "menos_hip.h"

namespace menos::engines::dense_hip {

// The Backend value of this engine. include/menos/plan.h:86 declares
// enum class Backend { Cpu, Cuda }; the integration adds Hip = 2 there. Until
// that line exists the adapter names the value by its number, so nothing
// reuses Backend::Cuda (whose profile fields describe another device).
inline constexpr Backend kBackendHip = static_cast<Backend>(2);

// True when libmenos_hip.so finds a usable gfx950 device (the runtime gate
// execution_engines() applies, as cuda_device_available() does for CUDA).
bool hip_device_available();

class HipDenseEngine final : public SimulationEngine {
public:
    [[nodiscard]] std::string_view name() const override { return "dense_hip"; }
    [[nodiscard]] Engine engine_id() const override { return Engine::Dense; }
    [[nodiscard]] Backend backend() const override { return kBackendHip; }
    [[nodiscard]] EngineCapabilities capabilities() const override;
    [[nodiscard]] CostEstimate estimate(const EngineEstimateRequest& request) const override;
    [[nodiscard]] EngineExecutionResult execute(const EngineExecutionRequest& request) const override;

    // The calibrated device seconds mh_engine_estimate predicts for a request
    // (the planner's hip_seconds, beside cuda_seconds, src/plan/planner.cpp:415-433);
    // negative when the request is infeasible.
    [[nodiscard]] double device_seconds(const EngineEstimateRequest& request) const;
};

// The host-offloaded tier, (Dense, Hip, HostOffloaded): the reference's
// Tier::HostOffloaded (include/menos/plan.h:88-98), which its planner refuses
// as not built (src/plan/planner.cpp:386-390). The state lives in pinned host
// memory and the passes run on the device through two windows
// (mh_tier_estimate, mh_tier_execute) on the disk tier's pass plan. Like the
// disk tier's engine (src/engines/dense_tiered/disk_dense_engine.h:8-40) it
// returns no statevector, runs a job's sole region only, and takes the
// request's tier seams (tier_chunk_qubits, tier_window_qubits); its counts and
// raw terms are HipDenseEngine's. budget_bytes bounds the host state; the
// windows take the device budget given here (0: the device's available bytes).
class HipHostTierEngine final : public SimulationEngine {
public:
    explicit HipHostTierEngine(std::uint64_t device_budget_bytes = 0) : device_budget_(device_budget_bytes) {}
    [[nodiscard]] std::string_view name() const override { return "dense_hip_host_offloaded"; }
    [[nodiscard]] Engine engine_id() const override { return Engine::Dense; }
    [[nodiscard]] Backend backend() const override { return kBackendHip; }
    [[nodiscard]] Tier tier() const override { return Tier::HostOffloaded; }
    [[nodiscard]] EngineCapabilities capabilities() const override;
    [[nodiscard]] CostEstimate estimate(const EngineEstimateRequest& request) const override;
    [[nodiscard]] EngineExecutionResult execute(const EngineExecutionRequest& request) const override;

private:
    std::uint64_t device_budget_;
};

// The operations of a region in the C ABI's flat circuit layout (the
// reference's Circuit::append_batch layout, src/ir/circuit.cpp:34-67).
// Holds the arrays the mh_circuit view points into.
class FlatCircuit {
public:
    FlatCircuit(std::uint32_t num_qubits, std::span<const Operation> operations);
    [[nodiscard]] const mh_circuit& view() const { return view_; }

private:
    std::vector<std::int32_t> opcodes_;
    std::vector<std::uint32_t> qubit_offsets_{0}, qubits_, param_offsets_{0};
    std::vector<double> params_;
    mh_circuit view_{};
};

}  // namespace menos::engines::dense_hip
