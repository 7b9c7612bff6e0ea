// The HIP engine in MENOS's registry and in its dense backend choice.
//
// Registry: execution_engines() (reference src/plan/execution_engines.cpp:12-34)
// gains the HIP engine, runtime-gated on a present device exactly as the CUDA
// entry is; find_execution_engine (:37-45) resolves (Dense, Hip, InMemory) to
// it. with_hip_engine() extends a registry list that way, and find_engine()
// resolves a triple over such a list.
//
// Backend choice: the planner's two-level rule (src/plan/planner.cpp:1021-1155):
// the algorithm is chosen by work units first; then the dense candidate moves to
// the device only when the device's calibrated seconds are strictly below the
// CPU's (ties to the CPU), the device fits the state, both sides are
// calibrated, and the job did not force the CPU. A forced HIP backend is
// refused before costing when no HIP engine is built, no device is present or
// the circuit is noisy (UnsupportedFeature: never a silent fallback, planner.cpp:363-380),
// and made infeasible, with the reason, when its estimate is infeasible or the
// state does not fit the device. An observed overrun raises a backend's figure
// before the comparison (raise_to_observed_floor, Phase 11).
#pragma once

#include <cstdint>
#include <optional>
#include <string>
#include <vector>

#include "hip_dense_engine.h"
#include "menos/plan.h"
#include "menos/simulation_engine.h"

namespace menos::engines::dense_hip {

// The registry list with the HIP engine appended when a device is present.
std::vector<const SimulationEngine*> with_hip_engine(std::vector<const SimulationEngine*> engines);

const SimulationEngine* find_engine(const std::vector<const SimulationEngine*>& engines, Engine engine,
                                    Backend backend, Tier tier = Tier::InMemory);

// The HIP twin of HardwareProfile's cuda_* fields (src/plan/hardware_profile.h:52-77),
// filled from mh_calibrate / mh_device_available_bytes.
struct HipProfile {
    bool present = false;
    std::uint64_t device_available_bytes = 0;
    bool calibrated = false;       // mh_calibrate ran (or a profile was installed)
};

enum class BackendRequest { Any, Cpu, Hip };

struct DenseBackendInputs {
    BackendRequest requested = BackendRequest::Any;
    bool hip_engine_registered = false;
    std::optional<std::string> noise_reason;  // ir::first_noise_reason of the job
    HipProfile hip;
    bool cpu_calibrated = false;
    // The dense candidate as the CPU planned it.
    bool cpu_feasible = true;
    Tier cpu_tier = Tier::InMemory;
    double cpu_seconds = 0.0;
    // The HIP engine's estimate for the same request and its seconds
    // (host preparation priced by the CPU profile + HipDenseEngine::device_seconds).
    CostEstimate hip_estimate{false, 0, 0, "", {}};
    double hip_seconds = 0.0;
    // What abandoned attempts of this job measured on each backend (0: none).
    double observed_cpu_seconds = 0.0;
    double observed_hip_seconds = 0.0;
};

struct DenseBackendChoice {
    Backend backend = Backend::Cpu;
    bool feasible = true;
    std::string infeasible_reason;
    double seconds = 0.0;  // the chosen backend's figure after the observed floors
};

// Throws common::UnsupportedFeature for a forced HIP backend that cannot be
// honoured at all.
DenseBackendChoice choose_dense_backend(const DenseBackendInputs& in);

}  // namespace menos::engines::dense_hip
